// seldon_ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A driver around the *real* Seldon headers (read in place from the
// reference tree by oracle/Makefile, never copied).  It loads inputs in
// Seldon's own binary layouts (CSR: matrix_sparse/Matrix_Sparse.cxx:
// 1532-1556 / 1649-1682; vectors: vector/Vector.cxx:846-874 / 1060-1091),
// calls Seldon's MltAdd / Mlt / Cg / BiCgStab / Gmres, and prints what the
// golden fixtures and the oracle cross-check need:
//   - every "Residu at iteration number k  r" line that Iteration::Finished
//     prints at print level 6, at 17 significant digits;
//   - "ERRCODE e NBITER n" at the end of a solve;
// and writes output vectors in Seldon's binary layout.
//
// Usage:
//   seldon_ref spmv  <row|sym|array> <d|f> A x y alpha beta out
//   seldon_ref mlt   <row|sym|array> <d|f> A x out
//   seldon_ref spmvt <row|sym|array> <d|f> A x y alpha beta out   (SeldonTrans)
//   seldon_ref mltt  <row|sym|array> <d|f> A x out                 (SeldonTrans)
//   seldon_ref sor <row|array|sym> <d|f> A x b omega iter type_ssor out [wide]
//   seldon_ref coo2csr <d|f> rows cols vals index m_min n_min out
//                    (ConvertMatrix_from_Coordinates into a RowSparse matrix
//                     first sized m_min x n_min; writes it as a binary CSR)
//   seldon_ref hbwrite <row|sym|array> <d|f> A out   (WriteHarwellBoeing)
//   seldon_ref hbread <d|f> file out
//                    (ReadHarwellBoeing into a ColSparse matrix, then
//                     ConvertMatrix_to_Coordinates / _from_Coordinates into a
//                     RowSparse matrix of the same size; a binary CSR)
//   seldon_ref solve <cg|bicgstab|gmres> <row|sym|array> <id|jacobi>
//                    A b x0|- tol maxit restart out
//   seldon_ref solvef ...  (the same arguments; float matrix and vectors:
//                    Iteration<double> for Cg / BiCgStab, Iteration<float>
//                    for Gmres)
// The Jacobi preconditioner does not exist in the reference; it is defined
// here as the Preconditioner_Base-shaped class z_i = r_i / a_ii (the
// division form of SOR's /a_jj, matrix_sparse/Relaxation_MatVect.cxx:79).
#include "Seldon.hxx"
#include "SeldonSolver.hxx"
#include "matrix_sparse/IOMatrixMarket.cxx"

#include <cstdlib>
#include <cstring>
#include <iostream>

using namespace Seldon;

template <class T>
class JacobiPreconditioner : public Preconditioner_Base
{
public:
  Vector<T> diag;
  template <class Matrix1, class Vector1>
  void Solve(const Matrix1&, const Vector1& r, Vector1& z)
  {
    for (int i = 0; i < r.GetM(); i++)
      z(i) = r(i) / diag(i);
  }
  template <class Matrix1, class Vector1>
  void TransSolve(const Matrix1& A, const Vector1& r, Vector1& z)
  {
    Solve(A, r, z);
  }
};

template <class T>
static void csr_diag(const Matrix<T, General, RowSparse>& A, Vector<T>& d)
{
  d.Reallocate(A.GetM());
  d.Fill(T(0));
  for (int i = 0; i < A.GetM(); i++)
    for (int j = A.GetPtr()[i]; j < A.GetPtr()[i + 1]; j++)
      if (A.GetInd()[j] == i)
        d(i) = A.GetData()[j];
}

template <class T>
static void sym_diag(const Matrix<T, Symmetric, RowSymSparse>& A, Vector<T>& d)
{
  d.Reallocate(A.GetM());
  d.Fill(T(0));
  for (int i = 0; i < A.GetM(); i++)
    for (int j = A.GetPtr()[i]; j < A.GetPtr()[i + 1]; j++)
      if (A.GetInd()[j] == i)
        d(i) = A.GetData()[j];
}

template <class T>
static void to_array(const Matrix<T, General, RowSparse>& A,
                     Matrix<T, General, ArrayRowSparse>& B)
{
  B.Reallocate(A.GetM(), A.GetN());
  for (int i = 0; i < A.GetM(); i++)
    {
      int b = A.GetPtr()[i], e = A.GetPtr()[i + 1];
      B.ReallocateRow(i, e - b);
      for (int k = 0; k < e - b; k++)
        {
          B.Index(i, k) = A.GetInd()[b + k];
          B.Value(i, k) = A.GetData()[b + k];
        }
    }
}

// Y <- the product selected by (add, trans) on A: Mlt / MltAdd, plain or
// with SeldonTrans (Functions_MatVect.cxx:123-130, :729-758, :918-927;
// Functions_MatrixArray.cxx:591-632).
template <class Mat, class T>
static void product(bool add, bool trans, const Mat& A, const Vector<T>& x, T alpha, T beta, Vector<T>& y,
                    int out_len)
{
  if (!add)
    {
      y.Reallocate(out_len);
      y.Fill(T(0));
      if (trans) Mlt(SeldonTrans, A, x, y);
      else Mlt(A, x, y);
    }
  else if (trans)
    MltAdd(alpha, SeldonTrans, A, x, beta, y);
  else
    MltAdd(alpha, A, x, beta, y);
}

template <class T>
static int run_spmv(bool add, const string& st, char** a, bool trans = false)
{
  Vector<T> x, y;
  x.Read(a[1]);
  T alpha = 1, beta = 0;
  string out;
  if (add)
    {
      y.Read(a[2]);
      alpha = T(atof(a[3]));
      beta = T(atof(a[4]));
      out = a[5];
    }
  else
    {
      y.Reallocate(0);
      out = a[2];
    }
  if (st == "sym")
    {
      Matrix<T, Symmetric, RowSymSparse> A;
      A.Read(a[0]);
      product(add, trans, A, x, alpha, beta, y, A.GetM());
    }
  else
    {
      Matrix<T, General, RowSparse> A;
      A.Read(a[0]);
      const int len = trans ? A.GetN() : A.GetM();
      if (st == "array")
        {
          Matrix<T, General, ArrayRowSparse> B;
          to_array(A, B);
          product(add, trans, B, x, alpha, beta, y, len);
        }
      else
        product(add, trans, A, x, alpha, beta, y, len);
    }
  y.Write(out);
  return 0;
}

// Seldon's Gmres compiles only where the vector type and Iteration's type
// agree (its rotations mix them, Gmres.cxx:75-76, :145-157), and only with
// SELDON_WITH_BLAS (Solve(ColUpTriang), Blas_2.cxx:3225-3238).
template <class Mat, class Prec, class T, class Titer>
static int run_gmres(Mat& A, Vector<T>& x, const Vector<T>& b, Prec& M, Iteration<Titer>& it, T*, Titer*)
{
  cerr << "Gmres does not compile for these types" << endl;
  exit(3);
}

template <class Mat, class Prec, class T>
static int run_gmres(Mat& A, Vector<T>& x, const Vector<T>& b, Prec& M, Iteration<T>& it, T*, T*)
{
#ifdef SELDON_WITH_BLAS
  return Gmres(A, x, b, M, it);
#else
  cerr << "method not available without SELDON_WITH_BLAS: gmres" << endl;
  exit(3);
#endif
}

template <class Mat, class Prec, class T, class Titer>
static int solve_with(const string& method, Mat& A, Vector<T>& x,
                      const Vector<T>& b, Prec& M, Iteration<Titer>& it)
{
  if (method == "cg")
    return Cg(A, x, b, M, it);
  if (method == "bicgstab")
    return BiCgStab(A, x, b, M, it);
  if (method == "gmres")
    return run_gmres(A, x, b, M, it, (T*)0, (Titer*)0);
  cerr << "method not available: " << method << endl;
  exit(3);
}

template <class Mat, class T, class Titer>
static int solve_prec(const string& method, const string& prec, Mat& A,
                      const Vector<T>& d, Vector<T>& x,
                      const Vector<T>& b, Iteration<Titer>& it)
{
  if (prec == "jacobi")
    {
      JacobiPreconditioner<T> M;
      M.diag = d;
      return solve_with(method, A, x, b, M, it);
    }
  Preconditioner_Base M;
  return solve_with(method, A, x, b, M, it);
}

// T: the vector and matrix value type; Titer: Iteration's type.  Seldon's
// Cg and BiCgStab on float vectors run with Iteration<double>; its Gmres
// needs Iteration<float> for float vectors (rotations_cos is a
// Vector<Titer> that GenRot takes beside float values, Gmres.cxx:75-76).
template <class T, class Titer>
static int run_solve(char** a)
{
  string method = a[0], st = a[1], prec = a[2];
  Vector<T> b, x, d;
  b.Read(a[4]);
  Iteration<Titer> it(atoi(a[7]), Titer(atof(a[6])));
  it.SetRestart(atoi(a[8]));
  it.ShowFullHistory();
  if (string(a[5]) != "-")
    {
      x.Read(a[5]);
      it.SetInitGuess(false);
    }
  else
    {
      x.Reallocate(b.GetM());
      x.Fill(T(0));
    }
  cout.precision(17);
  int code;
  if (st == "sym")
    {
      Matrix<T, Symmetric, RowSymSparse> A;
      A.Read(a[3]);
      sym_diag(A, d);
      code = solve_prec(method, prec, A, d, x, b, it);
    }
  else
    {
      Matrix<T, General, RowSparse> A;
      A.Read(a[3]);
      csr_diag(A, d);
      if (st == "array")
        {
          Matrix<T, General, ArrayRowSparse> B;
          to_array(A, B);
          code = solve_prec(method, prec, B, d, x, b, it);
        }
      else
        code = solve_prec(method, prec, A, d, x, b, it);
    }
  cout << "ERRCODE " << code << " NBITER " << it.GetNumberIteration() << endl;
  x.Write(a[9]);
  return 0;
}

template <class T, class W>
static int run_sor(const string& st, char** a)
{
  Vector<T> x, b;
  x.Read(a[1]);
  b.Read(a[2]);
  const W omega = W(atof(a[3]));
  if (st == "sym")
    {
      Matrix<T, Symmetric, RowSymSparse> S;
      S.Read(a[0]);
      SOR(S, x, b, omega, atoi(a[4]), atoi(a[5]));
    }
  else
    {
      Matrix<T, General, RowSparse> A;
      A.Read(a[0]);
      if (st == "array")
        {
          Matrix<T, General, ArrayRowSparse> B;
          to_array(A, B);
          SOR(B, x, b, omega, atoi(a[4]), atoi(a[5]));
        }
      else
        SOR(A, x, b, omega, atoi(a[4]), atoi(a[5]));
    }
  x.Write(a[6]);
  return 0;
}

template <class T>
static int run_coo2csr(char** a)
{
  // The conversion Clear()s the index vectors, and Seldon's Vector::Clear
  // runs the destructor explicitly (vector/Vector.cxx:336): at -O2 the
  // vectors' own destructors would free the same memory again, so they are
  // heap objects that are not deleted.
  Vector<int>* row = new Vector<int>;
  Vector<int>* col = new Vector<int>;
  Vector<T>* val = new Vector<T>;
  row->Read(a[0]);
  col->Read(a[1]);
  val->Read(a[2]);
  Matrix<T, General, RowSparse> A(atoi(a[4]), atoi(a[5]));
  ConvertMatrix_from_Coordinates(*row, *col, *val, A, atoi(a[3]));
  A.Write(a[6]);
  return 0;
}

template <class T>
static int run_hbwrite(const string& st, char** a)
{
  if (st == "sym")
    {
      Matrix<T, Symmetric, RowSymSparse> S;
      S.Read(a[0]);
      WriteHarwellBoeing(S, a[1]);
      return 0;
    }
  Matrix<T, General, RowSparse> A;
  A.Read(a[0]);
  if (st == "array")
    {
      Matrix<T, General, ArrayRowSparse> B;
      to_array(A, B);
      WriteHarwellBoeing(B, a[1]);
    }
  else
    WriteHarwellBoeing(A, a[1]);
  return 0;
}

template <class T>
static int run_hbread(char** a)
{
  Matrix<T, General, ColSparse> C;
  ReadHarwellBoeing(a[0], C);
  // the conversion Clear()s the index vectors (see run_coo2csr): heap
  // objects that are not deleted
  Vector<int>* row = new Vector<int>;
  Vector<int>* col = new Vector<int>;
  Vector<T>* val = new Vector<T>;
  ConvertMatrix_to_Coordinates(C, *row, *col, *val, 0, false);
  Matrix<T, General, RowSparse> A(C.GetM(), C.GetN());
  ConvertMatrix_from_Coordinates(*row, *col, *val, A, 0);
  A.Write(a[1]);
  return 0;
}

int main(int argc, char** argv)
{
  try
    {
      if (argc >= 10 && string(argv[1]) == "spmv")
        return string(argv[3]) == "f" ? run_spmv<float>(true, argv[2], argv + 4)
                                      : run_spmv<double>(true, argv[2], argv + 4);
      if (argc >= 7 && string(argv[1]) == "mlt")
        return string(argv[3]) == "f" ? run_spmv<float>(false, argv[2], argv + 4)
                                      : run_spmv<double>(false, argv[2], argv + 4);
      if (argc >= 10 && string(argv[1]) == "spmvt")
        return string(argv[3]) == "f" ? run_spmv<float>(true, argv[2], argv + 4, true)
                                      : run_spmv<double>(true, argv[2], argv + 4, true);
      if (argc >= 7 && string(argv[1]) == "mltt")
        return string(argv[3]) == "f" ? run_spmv<float>(false, argv[2], argv + 4, true)
                                      : run_spmv<double>(false, argv[2], argv + 4, true);
      if (argc >= 11 && string(argv[1]) == "sor")
        {
          // a last argument "wide": omega is a double (float values only)
          if (string(argv[3]) != "f")
            return run_sor<double, double>(argv[2], argv + 4);
          if (argc >= 12 && string(argv[11]) == "wide")
            return run_sor<float, double>(argv[2], argv + 4);
          return run_sor<float, float>(argv[2], argv + 4);
        }
      if (argc >= 6 && string(argv[1]) == "hbwrite")
        return string(argv[3]) == "f" ? run_hbwrite<float>(argv[2], argv + 4)
                                      : run_hbwrite<double>(argv[2], argv + 4);
      if (argc >= 5 && string(argv[1]) == "hbread")
        return string(argv[2]) == "f" ? run_hbread<float>(argv + 3) : run_hbread<double>(argv + 3);
      if (argc >= 10 && string(argv[1]) == "coo2csr")
        return string(argv[2]) == "f" ? run_coo2csr<float>(argv + 3) : run_coo2csr<double>(argv + 3);
      if (argc >= 12 && string(argv[1]) == "solve")
        return run_solve<double, double>(argv + 2);
      if (argc >= 12 && string(argv[1]) == "solvef")
        {
          // float vectors and matrix; Iteration<float> for Gmres only
          if (string(argv[2]) == "gmres")
            return run_solve<float, float>(argv + 2);
          return run_solve<float, double>(argv + 2);
        }
    }
  catch (Error& e)
    {
      cerr << e.What() << endl;
      return 2;
    }
  cerr << "usage: see the header of seldon_ref_harness.cpp" << endl;
  return 1;
}
