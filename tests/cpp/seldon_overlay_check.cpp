// seldon_overlay_check.cpp -- the GPU storages dropped into the real Seldon.
//
// Compiled against the reference headers (oracle/Makefile, into oracle/_ref)
// and linked to libshk.so.  On one 2-D Laplacian system it runs
//   1. Seldon's own Cg on its CPU RowSparse matrix,
//   2. Cg on RowSparseHip / VectHip with Preconditioner_Base: the fused GPU
//      solver selected by overload resolution,
//   3. Cg on the same GPU storages with a user preconditioner type: Seldon's
//      generic Cg template, each primitive (Mlt, DotProdConj, Add, Norm2,
//      Copy, ...) reaching the GPU through the overload set,
// then, on a 2-D convection-diffusion system (nonsymmetric), the solvers that
// use the transposed product -- BiCg, Qmr, Cgne, Lsqr -- and Cgs, QCgs,
// TfQmr, Gcr, BiCgStabl, and on the Laplacian the symmetric ones -- MinRes,
// Symmlq, CoCg, BiCgcr, QmrSym, and Cg with Seldon's SorPreconditioner
// (SSOR) -- with Seldon's own templates, once on the CPU storages and once on
// the GPU storages (Mlt(SeldonTrans, ...), SOR, Swap and the rest through the
// overload set; BiCg, Qmr, Cgne, Lsqr, Cgs, QCgs, TfQmr, Gcr, BiCgStabl,
// CoCg, BiCgcr, QmrSym, MinRes and Symmlq with
// Preconditioner_Base reach the library's fused solvers, and once more,
// with a user preconditioner type, Seldon's templates: labels
// <solver>_gpu_generic),
// and prints each residual history (Iteration::Finished lines) after a
// "== <label>" marker, then "ERRCODE e NBITER n".
#include "Seldon.hxx"
#include "SeldonSolver.hxx"
#include "matrix_sparse/IOMatrixMarket.cxx"
#include "seldon_amd/SeldonHip.hxx"

#include <cstdlib>

using namespace Seldon;

// A preconditioner Seldon does not know: identity through Copy
class UserIdentity : public Preconditioner_Base
{
public:
  template <class Matrix1, class Vector1>
  void Solve(const Matrix1&, const Vector1& r, Vector1& z) { Copy(r, z); }
  template <class Matrix1, class Vector1>
  void TransSolve(const Matrix1& A, const Vector1& r, Vector1& z) { Solve(A, r, z); }
};

// upwind convection-diffusion on a g x g grid, natural order, sorted columns:
// -lap u + h (vx d/dx + vy d/dy) u, rows scaled by h^2
static void convdiff2d(int g, double vx, double vy, Matrix<double, General, RowSparse>& A)
{
  const int n = g * g;
  const double h = 1.0 / (g + 1);
  Vector<double> values;
  Vector<int> ptr(n + 1), ind;
  std::vector<double> vv;
  std::vector<int> vi;
  ptr(0) = 0;
  for (int iy = 0; iy < g; iy++)
    for (int ix = 0; ix < g; ix++)
      {
        int r = iy * g + ix;
        if (iy > 0) { vi.push_back(r - g); vv.push_back(-1 - h * vy); }
        if (ix > 0) { vi.push_back(r - 1); vv.push_back(-1 - h * vx); }
        vi.push_back(r); vv.push_back(4 + h * (vx + vy));
        if (ix < g - 1) { vi.push_back(r + 1); vv.push_back(-1); }
        if (iy < g - 1) { vi.push_back(r + g); vv.push_back(-1); }
        ptr(r + 1) = (int)vi.size();
      }
  values.Reallocate(vv.size());
  ind.Reallocate(vi.size());
  for (size_t k = 0; k < vv.size(); k++) { values(k) = vv[k]; ind(k) = vi[k]; }
  A.SetData(n, n, values, ptr, ind);
}

template <class Mat, class Vec, class Prec = Preconditioner_Base>
static void run_trans_solver(const string& which, const string& label, Mat& A, const Vec& b, int maxit, double tol)
{
  cout << "== " << which << "_" << label << endl;
  Vec x(b);
  x.Fill(0.0);
  Iteration<double> it(maxit, tol);
  it.ShowFullHistory();
  Prec M;
  int code = 0;
  if (which == "bicg") code = BiCg(A, x, b, M, it);
  else if (which == "qmr") code = Qmr(A, x, b, M, it);
  else if (which == "cgne") code = Cgne(A, x, b, M, it);
  else if (which == "lsqr") code = Lsqr(A, x, b, M, it);
  else if (which == "cgs") code = Cgs(A, x, b, M, it);
  else if (which == "qcgs") code = QCgs(A, x, b, M, it);
  else if (which == "tfqmr") code = TfQmr(A, x, b, M, it);
  else if (which == "gcr") code = Gcr(A, x, b, M, it);
  else if (which == "bicgstabl") code = BiCgStabl(A, x, b, M, it);
  else if (which == "minres") code = MinRes(A, x, b, M, it);
  else if (which == "symmlq") code = Symmlq(A, x, b, M, it);
  else if (which == "cocg") code = CoCg(A, x, b, M, it);
  else if (which == "bicgcr") code = BiCgcr(A, x, b, M, it);
  else code = QmrSym(A, x, b, M, it);
  cout << "ERRCODE " << code << " NBITER " << it.GetNumberIteration() << endl;
}

// Cg / BiCgStab with Seldon's own SorPreconditioner (SSOR, one sweep each
// way, omega 1.2), whose Solve calls Seldon::SOR: on the GPU storages that is
// the level-scheduled device sweep
template <class Mat, class Vec>
static void run_ssor_solver(const string& which, const string& label, Mat& A, const Vec& b, int maxit, double tol)
{
  cout << "== " << which << "_" << label << endl;
  Vec x(b);
  x.Fill(0.0);
  Iteration<double> it(maxit, tol);
  it.ShowFullHistory();
  SorPreconditioner<double> M;
  M.SetParameterRelaxation(1.2);
  M.SetNumberIterations(1);
  int code = which.compare(0, 7, "ssor_cg") == 0 ? Cg(A, x, b, M, it) : BiCgStab(A, x, b, M, it);
  cout << "ERRCODE " << code << " NBITER " << it.GetNumberIteration() << endl;
}

// The upper triangle of A as a RowSymSparse matrix.
static void upper_of(const Matrix<double, General, RowSparse>& A, Matrix<double, Symmetric, RowSymSparse>& S)
{
  const int n = A.GetM();
  const int* p = A.GetPtr();
  const int* c = A.GetInd();
  const double* v = A.GetData();
  std::vector<int> vi;
  std::vector<double> vv;
  Vector<int> ptr(n + 1);
  ptr(0) = 0;
  for (int r = 0; r < n; r++)
    {
      for (int k = p[r]; k < p[r + 1]; k++)
        if (c[k] >= r) { vi.push_back(c[k]); vv.push_back(v[k]); }
      ptr(r + 1) = (int)vi.size();
    }
  Vector<double> values(vv.size());
  Vector<int> ind(vi.size());
  for (size_t k = 0; k < vv.size(); k++) { values(k) = vv[k]; ind(k) = vi[k]; }
  S.SetData(n, n, values, ptr, ind);
}

// Usage: seldon_overlay_check [g] [gc] [dir] [cpu|gpu|both]
//   cpu   only Seldon's CPU storages (what tests/golden/make_golden.py
//         records as a fixture, in this container);
//   gpu   only the GPU storages, reading dir/b_cpu.bin (Seldon's vector
//         file of the Laplacian right-hand side) from the fixture;
//   both  (default) the two, interleaved.
int main(int argc, char** argv)
{
  const int g = argc > 1 ? atoi(argv[1]) : 64, n = g * g;
  const string dir = argc > 3 ? argv[3] : ".";
  const string mode = argc > 4 ? argv[4] : "both";
  const bool cpu = mode != "gpu", gpu = mode != "cpu";
  // 5-point Laplacian, natural order, sorted columns
  Vector<double> values, xt(n), bh(n);
  Vector<int> ptr(n + 1), ind;
  std::vector<double> vv;
  std::vector<int> vi;
  ptr(0) = 0;
  for (int iy = 0; iy < g; iy++)
    for (int ix = 0; ix < g; ix++)
      {
        int r = iy * g + ix;
        if (iy > 0) { vi.push_back(r - g); vv.push_back(-1); }
        if (ix > 0) { vi.push_back(r - 1); vv.push_back(-1); }
        vi.push_back(r); vv.push_back(4);
        if (ix < g - 1) { vi.push_back(r + 1); vv.push_back(-1); }
        if (iy < g - 1) { vi.push_back(r + g); vv.push_back(-1); }
        ptr(r + 1) = (int)vi.size();
      }
  values.Reallocate(vv.size());
  ind.Reallocate(vi.size());
  for (size_t k = 0; k < vv.size(); k++) { values(k) = vv[k]; ind(k) = vi[k]; }
  Matrix<double, General, RowSparse> A;
  A.SetData(n, n, values, ptr, ind);  // takes ownership (Matrix_Sparse.cxx:300-377)
  for (int i = 0; i < n; i++) xt(i) = std::sin(0.37 * i) + 0.1 * std::cos(1.3 * i);
  Mlt(A, xt, bh);
  cout.precision(17);

  if (cpu)
    {
      cout << "== cpu" << endl;
      Vector<double> x(n);
      x.Fill(0.0);
      Iteration<double> it(2000, 1e-10);
      it.ShowFullHistory();
      Preconditioner_Base M;
      int code = Cg(A, x, bh, M, it);
      cout << "ERRCODE " << code << " NBITER " << it.GetNumberIteration() << endl;
    }
  Matrix<double, General, RowSparseHip, HipAlloc<double> > Ag;
  Vector<double, VectHip, HipAlloc<double> > bg;
  if (gpu)
    {
      Ag.Copy(A);
      bg.Upload(bh);
      {
        cout << "== gpu_fused" << endl;
        Vector<double, VectHip, HipAlloc<double> > x(n);
        x.Fill(0.0);
        Iteration<double> it(2000, 1e-10);
        it.ShowFullHistory();
        Preconditioner_Base M;
        int code = Cg(Ag, x, bg, M, it);
        cout << "ERRCODE " << code << " NBITER " << it.GetNumberIteration() << endl;
      }
      {
        cout << "== gpu_generic" << endl;
        Vector<double, VectHip, HipAlloc<double> > x(n);
        x.Fill(0.0);
        Iteration<double> it(2000, 1e-10);
        it.ShowFullHistory();
        UserIdentity M;
        int code = Cg(Ag, x, bg, M, it);
        cout << "ERRCODE " << code << " NBITER " << it.GetNumberIteration() << endl;
      }
    }
  {
    const int gc = argc > 2 ? atoi(argv[2]) : 24, nc = gc * gc;
    Matrix<double, General, RowSparse> C;
    convdiff2d(gc, 20.0, 10.0, C);
    Vector<double> xc(nc), bc(nc);
    for (int i = 0; i < nc; i++) xc(i) = std::cos(0.21 * i) - 0.3 * std::sin(0.7 * i);
    Mlt(C, xc, bc);
    Matrix<double, General, RowSparseHip, HipAlloc<double> > Cg_;
    Vector<double, VectHip, HipAlloc<double> > bcg;
    if (gpu)
      {
        Cg_.Copy(C);
        bcg.Upload(bc);
      }
    const char* solvers[9] = {"bicg", "qmr", "cgne", "lsqr", "cgs", "qcgs", "tfqmr", "gcr", "bicgstabl"};
    for (int k = 0; k < 9; k++)
      {
        const int maxit = (k == 2 || k == 3) ? 4000 : 2000;
        if (cpu) run_trans_solver(solvers[k], "cpu", C, bc, maxit, 1e-8);
        if (gpu) run_trans_solver(solvers[k], "gpu", Cg_, bcg, maxit, 1e-8);
        if (gpu)
          run_trans_solver<Matrix<double, General, RowSparseHip, HipAlloc<double> >,
                           Vector<double, VectHip, HipAlloc<double> >, UserIdentity>(solvers[k], "gpu_generic",
                                                                                   Cg_, bcg, maxit, 1e-8);
      }
  }
  {
    // the symmetric methods on the SPD Laplacian above
    const char* solvers[5] = {"minres", "symmlq", "cocg", "bicgcr", "qmrsym"};
    for (int k = 0; k < 5; k++)
      {
        if (cpu) run_trans_solver(solvers[k], "cpu", A, bh, 2000, 1e-8);
        if (gpu) run_trans_solver(solvers[k], "gpu", Ag, bg, 2000, 1e-8);
        if (gpu)
          run_trans_solver<Matrix<double, General, RowSparseHip, HipAlloc<double> >,
                           Vector<double, VectHip, HipAlloc<double> >, UserIdentity>(solvers[k], "gpu_generic",
                                                                                   Ag, bg, 2000, 1e-8);
      }
    if (cpu) run_ssor_solver("ssor_cg", "cpu", A, bh, 2000, 1e-10);
    if (gpu) run_ssor_solver("ssor_cg", "gpu", Ag, bg, 2000, 1e-10);
    // the same on the symmetric storage (upper triangle)
    Matrix<double, Symmetric, RowSymSparse> S;
    upper_of(A, S);
    Matrix<double, Symmetric, RowSymSparseHip, HipAlloc<double> > Sg;
    if (gpu) Sg.Copy(S);
    if (cpu) run_ssor_solver("ssor_cg_sym", "cpu", S, bh, 2000, 1e-10);
    if (gpu) run_ssor_solver("ssor_cg_sym", "gpu", Sg, bg, 2000, 1e-10);
    // Harwell-Boeing: the device matrices write Seldon's files
    if (cpu)
      {
        WriteHarwellBoeing(A, dir + "/cpu.rua");
        WriteHarwellBoeing(S, dir + "/cpu_sym.rua");
        bh.Write(dir + "/b_cpu.bin");  // Seldon's binary vector file
      }
    if (gpu)
      {
        WriteHarwellBoeing(Ag, dir + "/gpu.rua");
        WriteHarwellBoeing(Sg, dir + "/gpu_sym.rua");
        // the same vector file from the device vector, and Seldon's file
        // read back into one
        bg.Write(dir + "/b_gpu.bin");
        Vector<double, VectHip, HipAlloc<double> > rb;
        rb.Read(dir + "/b_cpu.bin");
        rb.Write(dir + "/b_back.bin");
      }
    cout << "== hb" << endl;
  }
  return 0;
}
