// seldon_hip_driver.cpp -- Seldon-style C++ program on the GPU storages
// (standalone mode: include/seldon_amd/SeldonCore.hxx + SeldonHip.hxx).
//
// Same command line and output as oracle/seldon_ref_harness.cpp so a test can
// compare the two line by line:
//   seldon_hip_driver solve <cg|bicgstab|gmres> <row|sym|array> <id|jacobi>
//                     A b x0|- tol maxit restart out
//   seldon_hip_driver spmv <row|sym|array> <d|f> A x y alpha beta out
// Inputs and outputs use Seldon's binary layouts (CSR: int m, n, nz; ptr;
// ind; data -- vectors: int m; data).
#include <cstdio>
#include <fstream>

#include "seldon_amd/SeldonHip.hxx"

using namespace Seldon;

template <class T>
static void read_vec(const char* path, std::vector<T>& v)
{
  std::ifstream f(path, std::ios::binary);
  int m = 0;
  f.read((char*)&m, sizeof(int));
  v.resize(m);
  f.read((char*)v.data(), sizeof(T) * (size_t)m);
  if (!f) throw Error("read_vec", path);
}

template <class T>
static void write_vec(const char* path, const std::vector<T>& v)
{
  std::ofstream f(path, std::ios::binary);
  int m = (int)v.size();
  f.write((const char*)&m, sizeof(int));
  f.write((const char*)v.data(), sizeof(T) * (size_t)m);
}

template <class T>
struct Csr
{
  int m, n, nz;
  std::vector<int> ptr, ind;
  std::vector<T> val;
};

template <class T>
static Csr<T> read_csr(const char* path)
{
  Csr<T> A;
  std::ifstream f(path, std::ios::binary);
  f.read((char*)&A.m, sizeof(int));
  f.read((char*)&A.n, sizeof(int));
  f.read((char*)&A.nz, sizeof(int));
  A.ptr.resize(A.m + 1);
  A.ind.resize(A.nz);
  A.val.resize(A.nz);
  f.read((char*)A.ptr.data(), sizeof(int) * (A.m + 1));
  f.read((char*)A.ind.data(), sizeof(int) * (size_t)A.nz);
  f.read((char*)A.val.data(), sizeof(T) * (size_t)A.nz);
  if (!f) throw Error("read_csr", path);
  return A;
}

template <class Mat, class Prec>
static int solve_with(const std::string& method, Mat& A, Vector<double, VectHip, HipAlloc<double> >& x,
                      const Vector<double, VectHip, HipAlloc<double> >& b, Prec& M, Iteration<double>& it)
{
  if (method == "cg") return Cg(A, x, b, M, it);
  if (method == "bicgstab") return BiCgStab(A, x, b, M, it);
  return Gmres(A, x, b, M, it);
}

template <class Mat>
static int solve_prec(const std::string& method, const std::string& prec, Mat& A,
                      Vector<double, VectHip, HipAlloc<double> >& x, const Vector<double, VectHip, HipAlloc<double> >& b,
                      Iteration<double>& it)
{
  if (prec == "jacobi")
    {
      JacobiPreconditioner M;
      return solve_with(method, A, x, b, M, it);
    }
  Preconditioner_Base M;
  return solve_with(method, A, x, b, M, it);
}

static int run_solve(char** a)
{
  std::string method = a[0], st = a[1], prec = a[2];
  Csr<double> C = read_csr<double>(a[3]);
  std::vector<double> hb, hx;
  read_vec(a[4], hb);
  Iteration<double> it(atoi(a[7]), atof(a[6]));
  it.SetRestart(atoi(a[8]));
  it.ShowFullHistory();
  Vector<double, VectHip, HipAlloc<double> > b(hb.size()), x(hb.size());
  b.Upload(hb.data(), (int)hb.size());
  if (std::string(a[5]) != "-")
    {
      read_vec(a[5], hx);
      x.Upload(hx.data(), (int)hx.size());
      it.SetInitGuess(false);
    }
  else
    x.Fill(0.0);
  cout.precision(17);
  int code;
  if (st == "sym")
    {
      Matrix<double, Symmetric, RowSymSparseHip, HipAlloc<double> > A;
      A.SetData(C.m, C.n, C.nz, C.ptr.data(), C.ind.data(), C.val.data());
      code = solve_prec(method, prec, A, x, b, it);
    }
  else if (st == "array")
    {
      Matrix<double, General, ArrayRowSparseHip, HipAlloc<double> > A;
      A.SetData(C.m, C.n, C.nz, C.ptr.data(), C.ind.data(), C.val.data());
      code = solve_prec(method, prec, A, x, b, it);
    }
  else
    {
      Matrix<double, General, RowSparse, MallocAlloc<double> > H;
      H.SetData(C.m, C.n, C.val, C.ptr, C.ind);
      Matrix<double, General, RowSparseHip, HipAlloc<double> > A(H);
      code = solve_prec(method, prec, A, x, b, it);
    }
  cout << "ERRCODE " << code << " NBITER " << it.GetNumberIteration() << endl;
  hx.resize(x.GetM());
  x.Download(hx.data());
  write_vec(a[9], hx);
  return 0;
}

template <class T>
static int run_spmv(const std::string& st, char** a)
{
  Csr<T> C = read_csr<T>(a[0]);
  std::vector<T> hx, hy;
  read_vec(a[1], hx);
  read_vec(a[2], hy);
  Vector<T, VectHip, HipAlloc<T> > x(hx.size()), y(hy.size());
  x.Upload(hx.data(), (int)hx.size());
  y.Upload(hy.data(), (int)hy.size());
  T alpha = T(atof(a[3])), beta = T(atof(a[4]));
  if (st == "sym")
    {
      Matrix<T, Symmetric, RowSymSparseHip, HipAlloc<T> > A;
      A.SetData(C.m, C.n, C.nz, C.ptr.data(), C.ind.data(), C.val.data());
      MltAdd(alpha, A, x, beta, y);
    }
  else if (st == "array")
    {
      Matrix<T, General, ArrayRowSparseHip, HipAlloc<T> > A;
      A.SetData(C.m, C.n, C.nz, C.ptr.data(), C.ind.data(), C.val.data());
      MltAdd(alpha, A, x, beta, y);
    }
  else
    {
      Matrix<T, General, RowSparseHip, HipAlloc<T> > A;
      A.SetData(C.m, C.n, C.nz, C.ptr.data(), C.ind.data(), C.val.data());
      MltAdd(alpha, A, x, beta, y);
    }
  y.Download(hy.data());
  write_vec(a[5], hy);
  return 0;
}

int main(int argc, char** argv)
{
  try
    {
      if (argc >= 12 && std::string(argv[1]) == "solve")
        return run_solve(argv + 2);
      if (argc >= 10 && std::string(argv[1]) == "spmv")
        return std::string(argv[3]) == "f" ? run_spmv<float>(argv[2], argv + 4) : run_spmv<double>(argv[2], argv + 4);
    }
  catch (Error& e)
    {
      cout << e.What() << endl;
      return 2;
    }
  cout << "usage: see the header of seldon_hip_driver.cpp" << endl;
  return 1;
}
