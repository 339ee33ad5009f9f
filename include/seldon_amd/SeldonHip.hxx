// SeldonHip.hxx -- Seldon's sparse Krylov API on an MI355X, as a drop-in.
//
// Include after Seldon.hxx + SeldonSolver.hxx to extend the real Seldon, or
// on its own to get the standalone core (SeldonCore.hxx).  Unmodified Seldon
// code such as
//
//     Matrix<double, General, RowSparseHip> A(A_host);   // upload
//     Vector<double, VectHip> x(n), b(b_host);
//     Iteration<double> iter(1000, 1e-10);
//     Preconditioner_Base M;
//     Cg(A, x, b, M, iter);
//
// then runs on the GPU: the storages below hold device memory; the overloads
// of Mlt / MltAdd / DotProd / DotProdConj / Norm2 / Add / Copy / Mlt(alpha,X)
// are what Seldon's solver templates call unqualified
// (computation/solver/iterative/Cg.cxx:45-116, doc/guide/iterative.dox:62,223),
// so Seldon's own templates run primitive by primitive on the GPU; and the
// more specialised Cg / BiCgStab / Gmres / Cgs / BiCg / TfQmr / Qmr / Cgne / Lsqr / CoCg / BiCgcr /
// QmrSym / MinRes / Symmlq / QCgs / Gcr / BiCgStabl below replace the whole loop by the
// library's fused, device-driven solvers (same Iteration semantics, residual
// history replayed through Iteration::Finished so messages print as Seldon's)
// when the preconditioner is Preconditioner_Base (identity) or
// JacobiPreconditioner.  Every call goes through the C ABI of include/shk.h.
//
// The transposed products (MltAdd / Mlt with SeldonTrans) and Swap are
// overloaded too, so the transpose-using solvers (BiCg, Qmr, Cgne, Lsqr with
// any other preconditioner type) also run on the GPU storages through
// Seldon's own templates.
//
// Names: VectHip, RowSparseHip, RowSymSparseHip, ArrayRowSparseHip are the
// device counterparts of Seldon's VectFull, RowSparse, RowSymSparse and
// ArrayRowSparse storages (share/Storage.hxx:56,82,115,138).
#ifndef SELDON_AMD_SELDON_HIP_HXX
#define SELDON_AMD_SELDON_HIP_HXX

#ifndef SELDON_FILE_SELDON_HXX
#include "SeldonCore.hxx"
#endif

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../shk.h"

namespace Seldon
{
  class VectHip {};
  class RowSparseHip {};
  class RowSymSparseHip {};
  class ArrayRowSparseHip {};

  template <class T>
  class HipAlloc {};

  // ----------------------------------------------------------- context --
  // One library context (device + stream + workspace) per process, created
  // on first use; SetDevice before the first GPU object selects the device.
  class HipContext
  {
    shk_ctx ctx_;
    HipContext(int device) : ctx_(0)
    {
      int st = shk_ctx_create(device, &ctx_);
      if (st != SHK_OK)
        throw Error("HipContext", std::string("no usable gfx950 device: ") + shk_last_error());
    }
    static int& Device() { static int d = 0; return d; }
  public:
    ~HipContext() { if (ctx_) shk_ctx_destroy(ctx_); }
    static HipContext& Get()
    {
      static HipContext c(Device());
      return c;
    }
    static void SetDevice(int d) { Device() = d; }
    shk_ctx Handle() const { return ctx_; }
  };

  inline void HipCheck(int st, const char* fn)
  {
    if (st == SHK_OK)
      return;
    if (st == SHK_ERR_NOMEM)
      throw NoMemory(fn, shk_last_error());
    if (st == SHK_ERR_DIM)
      throw WrongDim(fn, shk_last_error());
    throw Error(fn, shk_last_error());
  }

  template <class T> struct HipType;
  template <> struct HipType<double> { static const shk_dtype value = SHK_F64; };
  template <> struct HipType<float> { static const shk_dtype value = SHK_F32; };

  // ------------------------------------------------------ device vector --
  template <class T, class Allocator>
  class Vector<T, VectHip, Allocator>
  {
  protected:
    int m_;
    T* data_;

    void alloc(int m)
    {
      m_ = m > 0 ? m : 0;
      data_ = 0;
      if (m_ > 0)
        {
          void* p = 0;
          HipCheck(shk_malloc(HipContext::Get().Handle(), sizeof(T) * (size_t)m_, &p),
                   "Vector<VectHip>::Reallocate");
          data_ = static_cast<T*>(p);
        }
    }

  public:
    typedef T value_type;
    typedef T* pointer;
    typedef const T* const_pointer;

    Vector() : m_(0), data_(0) {}
    explicit Vector(int m) { alloc(m); }
    Vector(const Vector& X) { alloc(X.m_); CopyFrom(X); }
    // upload a host vector (any storage exposing GetM() and GetData())
    template <class HostVector>
    explicit Vector(const HostVector& h,
                    typename std::enable_if<std::is_class<HostVector>::value &&
                                            !std::is_same<HostVector, Vector>::value>::type* = 0)
    {
      alloc(h.GetM());
      Upload(h);
    }
    ~Vector() { Clear(); }

    Vector& operator=(const Vector& X)
    {
      if (this != &X)
        {
          if (m_ != X.m_) Reallocate(X.m_);
          CopyFrom(X);
        }
      return *this;
    }

    void Clear()
    {
      if (data_)
        shk_free(HipContext::Get().Handle(), data_);
      data_ = 0;
      m_ = 0;
    }
    void Reallocate(int m)
    {
      if (m == m_) return;
      Clear();
      alloc(m);
    }
    int GetM() const { return m_; }
    int GetLength() const { return m_; }
    int GetSize() const { return m_; }
    int GetDataSize() const { return m_; }
    T* GetData() const { return data_; }

    // Vector::Zero / Fill(x) (vector/Vector.cxx:686-713)
    void Zero() { Fill(T(0)); }
    template <class T0>
    void Fill(const T0& x)
    {
      HipCheck(shk_fill(HipContext::Get().Handle(), HipType<T>::value, m_, double(T(x)), data_),
               "Vector<VectHip>::Fill");
    }
    // Vector::Fill() -- 0, 1, 2, ...
    void Fill()
    {
      std::vector<T> h(m_);
      for (int i = 0; i < m_; i++) h[i] = T(i);
      Upload(h.empty() ? 0 : &h[0], m_);
    }
    // Vector::Copy (vector/Vector.cxx:579-586)
    void Copy(const Vector& X)
    {
      Reallocate(X.m_);
      CopyFrom(X);
    }
    void CopyFrom(const Vector& X)
    {
      if (m_ == 0) return;
      HipCheck(shk_copy(HipContext::Get().Handle(), HipType<T>::value, m_, X.data_, data_),
               "Vector<VectHip>::Copy");
    }
    // operator*= (vector/Vector.cxx:607-615)
    template <class T0>
    Vector& operator*=(const T0& alpha)
    {
      HipCheck(shk_scal(HipContext::Get().Handle(), HipType<T>::value, m_, double(alpha), data_),
               "Vector<VectHip>::operator*=");
      return *this;
    }

    void Upload(const T* host, int m)
    {
      if (m != m_) Reallocate(m);
      if (m_ > 0)
        HipCheck(shk_memcpy_h2d(HipContext::Get().Handle(), data_, host, sizeof(T) * (size_t)m_),
                 "Vector<VectHip>::Upload");
    }
    template <class HostVector>
    void Upload(const HostVector& h) { Upload(h.GetData(), h.GetM()); }

    void Download(T* host) const
    {
      if (m_ > 0)
        HipCheck(shk_memcpy_d2h(HipContext::Get().Handle(), host, data_, sizeof(T) * (size_t)m_),
                 "Vector<VectHip>::Download");
    }
    template <class HostVector>
    void Download(HostVector& h) const
    {
      h.Reallocate(m_);
      Download(h.GetData());
    }
    // exchange storage with Y (Swap)
    void SwapStorage(Vector& Y)
    {
      std::swap(m_, Y.m_);
      std::swap(data_, Y.data_);
    }
    // element read by value, one device-to-host transfer (CoCg.cxx:53 reads
    // b(0) to form a zero of the value type)
    T operator()(int i) const { return GetElement(i); }
    // Seldon's binary vector file (vector/Vector.cxx:846-874 Write,
    // :1060-1091 Read): the length as an int when with_size, then the
    // elements; Read without the length fills the vector's current length
    void Write(std::string FileName, bool with_size = true) const
    {
      HipCheck(shk_vec_write(HipContext::Get().Handle(), HipType<T>::value, m_, data_, FileName.c_str(),
                             with_size ? 1 : 0),
               "Vector<VectHip>::Write");
    }
    void Read(std::string FileName, bool with_size = true)
    {
      int64_t n = m_;
      if (with_size)
        {
          HipCheck(shk_vec_read(HipContext::Get().Handle(), HipType<T>::value, FileName.c_str(), 1, 0, 0, &n),
                   "Vector<VectHip>::Read");
          Reallocate((int)n);
        }
      HipCheck(shk_vec_read(HipContext::Get().Handle(), HipType<T>::value, FileName.c_str(), with_size ? 1 : 0,
                            m_, data_, &n),
               "Vector<VectHip>::Read");
    }
    // element read for inspection (one transfer)
    T GetElement(int i) const
    {
      T v;
      HipCheck(shk_memcpy_d2h(HipContext::Get().Handle(), &v, data_ + i, sizeof(T)),
               "Vector<VectHip>::GetElement");
      return v;
    }
  };

  // ------------------------------------------------ device sparse matrix --
  template <class T, int STORAGE>
  class HipSparseBase
  {
  protected:
    shk_csr h_;
    int m_, n_, nz_;

    void release()
    {
      if (h_) shk_csr_destroy(h_);
      h_ = 0;
    }

  public:
    typedef T value_type;
    HipSparseBase() : h_(0), m_(0), n_(0), nz_(0) {}
    ~HipSparseBase() { release(); }

    // Upload CSR arrays as Seldon stores them (0-based, RowSymSparse: upper
    // part); the caller keeps its arrays (Matrix_Sparse::SetData copies here).
    void SetData(int m, int n, int nz, const int* ptr, const int* ind, const T* values)
    {
      release();
      HipCheck(shk_csr_create(HipContext::Get().Handle(), (shk_storage)STORAGE, HipType<T>::value, m, n, nz,
                              ptr, ind, values, 0, &h_),
               "Matrix<Hip>::SetData");
      m_ = m;
      n_ = n;
      nz_ = nz;
    }
    // from a host CSR matrix (Seldon's Matrix_Sparse / Matrix_SymSparse)
    template <class HostCsr>
    void Copy(const HostCsr& A)
    {
      SetData(A.GetM(), A.GetN(), A.GetDataSize(), A.GetPtr(), A.GetInd(), A.GetData());
    }
    // from a Seldon ArrayRowSparse matrix: rows laid end to end, as
    // ConvertToCSR does (matrix_sparse/Matrix_Conversions.cxx:3676-3705)
    template <class HostArray>
    void CopyArray(const HostArray& A)
    {
      std::vector<int> ptr(A.GetM() + 1, 0), ind;
      std::vector<T> val;
      for (int i = 0; i < A.GetM(); i++)
        {
          for (int k = 0; k < A.GetRowSize(i); k++)
            {
              ind.push_back(A.Index(i, k));
              val.push_back(A.Value(i, k));
            }
          ptr[i + 1] = (int)ind.size();
        }
      SetData(A.GetM(), A.GetN(), (int)ind.size(), &ptr[0], ind.empty() ? 0 : &ind[0],
              val.empty() ? 0 : &val[0]);
    }
    // takes over a matrix the library created (read from a file)
    void Adopt(shk_csr h)
    {
      release();
      h_ = h;
      int64_t nz = 0;
      HipCheck(shk_csr_info(h, &m_, &n_, &nz, 0, 0, 0), "Matrix<Hip>::Adopt");
      nz_ = (int)nz;
    }
    void SetValues(const T* values)
    {
      HipCheck(shk_csr_set_values(HipContext::Get().Handle(), h_, values, 0), "Matrix<Hip>::SetValues");
    }
    int GetM() const { return m_; }
    int GetN() const { return n_; }
    int GetNonZeros() const { return nz_; }
    int GetDataSize() const { return nz_; }
    shk_csr GetHandle() const { return h_; }
  };

  template <class T, class Prop, class Allocator>
  class Matrix<T, Prop, RowSparseHip, Allocator> : public HipSparseBase<T, SHK_ROWSPARSE>
  {
  public:
    Matrix() {}
    template <class HostCsr>
    explicit Matrix(const HostCsr& A) { this->Copy(A); }
  };

  template <class T, class Prop, class Allocator>
  class Matrix<T, Prop, RowSymSparseHip, Allocator> : public HipSparseBase<T, SHK_ROWSYMSPARSE>
  {
  public:
    Matrix() {}
    template <class HostCsr>
    explicit Matrix(const HostCsr& A) { this->Copy(A); }
  };

  template <class T, class Prop, class Allocator>
  class Matrix<T, Prop, ArrayRowSparseHip, Allocator> : public HipSparseBase<T, SHK_ARRAYROWSPARSE>
  {
  public:
    Matrix() {}
    template <class HostArray>
    explicit Matrix(const HostArray& A) { this->CopyArray(A); }
  };

  // WriteHarwellBoeing(A, file_name) (matrix_sparse/IOMatrixMarket.cxx:397-510):
  // the file Seldon writes for the same storage and arrays, byte for byte.
  // One overload per device storage, more specialised than Seldon's template.
#define SELDON_HIP_WRITE_HB(S)                                                                           \
  template <class T, class Prop, class Allocator>                                                        \
  void WriteHarwellBoeing(const Matrix<T, Prop, S, Allocator>& A, const std::string& file_name)          \
  {                                                                                                      \
    HipCheck(shk_csr_write_hb(HipContext::Get().Handle(), A.GetHandle(), file_name.c_str()),             \
             "WriteHarwellBoeing");                                                                      \
  }
  SELDON_HIP_WRITE_HB(RowSparseHip)
  SELDON_HIP_WRITE_HB(RowSymSparseHip)
  SELDON_HIP_WRITE_HB(ArrayRowSparseHip)
#undef SELDON_HIP_WRITE_HB

  // ReadHarwellBoeing(filename, A) (:80-360) into a RowSparse device matrix:
  // the matrix a Seldon program gets by reading into ColSparse and converting
  // through coordinates to RowSparse, its rows assembled on the device.
  template <class T, class Prop, class Allocator>
  void ReadHarwellBoeing(std::string filename, Matrix<T, Prop, RowSparseHip, Allocator>& A)
  {
    shk_csr h = 0;
    HipCheck(shk_csr_read_hb(HipContext::Get().Handle(), filename.c_str(), HipType<T>::value, SHK_LAYOUT_AUTO,
                             &h),
             "ReadHarwellBoeing");
    A.Adopt(h);
  }

  template <class S> struct IsHipSparse : std::false_type {};
  template <> struct IsHipSparse<RowSparseHip> : std::true_type {};
  template <> struct IsHipSparse<RowSymSparseHip> : std::true_type {};
  template <> struct IsHipSparse<ArrayRowSparseHip> : std::true_type {};

  // ------------------------------------------------- the overload set --
  // MltAdd(alpha, M, X, beta, Y): Functions_MatVect.cxx:398-431 (RowSparse),
  // :526-564 (RowSymSparse); Functions_MatrixArray.cxx:534-575 (ArrayRowSparse)
  template <class T0, class T1, class Prop1, class Storage1, class Allocator1, class T2, class Allocator2,
            class T3, class T4, class Allocator4>
  typename std::enable_if<IsHipSparse<Storage1>::value>::type
  MltAdd(const T0 alpha, const Matrix<T1, Prop1, Storage1, Allocator1>& M, const Vector<T2, VectHip, Allocator2>& X,
         const T3 beta, Vector<T4, VectHip, Allocator4>& Y)
  {
    if (X.GetM() != M.GetN() || Y.GetM() != M.GetM())
      throw WrongDim("MltAdd(alpha, M, X, beta, Y)", "sizes do not match");
    HipCheck(shk_spmv(HipContext::Get().Handle(), M.GetHandle(), double(alpha), X.GetData(), double(beta),
                      Y.GetData()),
             "MltAdd");
  }

  // Mlt(M, X, Y) = Y.Fill(0); MltAdd(1, M, X, 0, Y) (Functions_MatVect.cxx:72-81)
  template <class T0, class Prop0, class Storage0, class Allocator0, class T1, class Allocator1, class T2,
            class Allocator2>
  typename std::enable_if<IsHipSparse<Storage0>::value>::type
  Mlt(const Matrix<T0, Prop0, Storage0, Allocator0>& M, const Vector<T1, VectHip, Allocator1>& X,
      Vector<T2, VectHip, Allocator2>& Y)
  {
    if (X.GetM() != M.GetN() || Y.GetM() != M.GetM())
      throw WrongDim("Mlt(M, X, Y)", "sizes do not match");
    HipCheck(shk_mlt(HipContext::Get().Handle(), M.GetHandle(), X.GetData(), Y.GetData()), "Mlt");
  }

  // MltAdd(alpha, SeldonTrans, M, X, beta, Y): Y <- beta Y + alpha M^T X
  // (Functions_MatVect.cxx:729-758 RowSparse, :918-927 RowSymSparse;
  // Functions_MatrixArray.cxx:591-632 ArrayRowSparse).  SeldonConjTrans is
  // the same for real values; SeldonNoTrans is the plain product.
  template <class T0, class T1, class Prop1, class Storage1, class Allocator1, class T2, class Allocator2,
            class T3, class T4, class Allocator4>
  typename std::enable_if<IsHipSparse<Storage1>::value>::type
  MltAdd(const T0 alpha, const SeldonTranspose& Trans, const Matrix<T1, Prop1, Storage1, Allocator1>& M,
         const Vector<T2, VectHip, Allocator2>& X, const T3 beta, Vector<T4, VectHip, Allocator4>& Y)
  {
    if (Trans.NoTrans())
      {
        MltAdd(alpha, M, X, beta, Y);
        return;
      }
    if (X.GetM() != M.GetM() || Y.GetM() != M.GetN())
      throw WrongDim("MltAdd(alpha, SeldonTrans, M, X, beta, Y)", "sizes do not match");
    HipCheck(shk_spmv_trans(HipContext::Get().Handle(), M.GetHandle(), double(alpha), X.GetData(),
                            double(beta), Y.GetData()),
             "MltAdd(SeldonTrans)");
  }

  // Mlt(SeldonTrans, M, X, Y) (Functions_MatVect.cxx:123-130): Y = M^T X,
  // summed per column of M in increasing row order like Seldon's generic loop
  // (:1307-1313); SeldonNoTrans is Mlt(M, X, Y).
  template <class T0, class Prop0, class Storage0, class Allocator0, class T1, class Allocator1, class T2,
            class Allocator2>
  typename std::enable_if<IsHipSparse<Storage0>::value>::type
  Mlt(const SeldonTranspose& Trans, const Matrix<T0, Prop0, Storage0, Allocator0>& M,
      const Vector<T1, VectHip, Allocator1>& X, Vector<T2, VectHip, Allocator2>& Y)
  {
    if (Trans.NoTrans())
      {
        Mlt(M, X, Y);
        return;
      }
    if (X.GetM() != M.GetM() || Y.GetM() != M.GetN())
      throw WrongDim("Mlt(SeldonTrans, M, X, Y)", "sizes do not match");
    HipCheck(shk_mlt_trans(HipContext::Get().Handle(), M.GetHandle(), X.GetData(), Y.GetData()),
             "Mlt(SeldonTrans)");
  }

  // SOR(A, X, B, omega, iter, type_ssor) (matrix_sparse/Relaxation_MatVect.cxx:
  // 41-95 RowSparse, :106-167 ArrayRowSparse, :179-300 RowSymSparse),
  // level-scheduled on the GPU and
  // bitwise Seldon's sweeps; with it Seldon's own SorPreconditioner
  // (Precond_Ssor.cxx), which calls Seldon::SOR, runs on the GPU storages.
  template <class T0, class Prop0, class Storage0, class Allocator0, class T1, class Allocator1, class T2,
            class Allocator2, class T3>
  typename std::enable_if<IsHipSparse<Storage0>::value>::type
  SOR(const Matrix<T0, Prop0, Storage0, Allocator0>& A, Vector<T2, VectHip, Allocator2>& X,
      const Vector<T1, VectHip, Allocator1>& B, const T3& omega, int iter, int type_ssor = 2)
  {
    if (X.GetM() != A.GetM() || B.GetM() != A.GetM())
      throw WrongDim("SOR", "Matrix and vector dimensions are incompatible.");
    // the relaxation step runs in the type C++ gives (value, omega)
    typedef typename std::common_type<T2, T3>::type W;
    HipCheck(shk_sor(HipContext::Get().Handle(), A.GetHandle(), X.GetData(), B.GetData(), double(omega),
                     std::is_same<W, float>::value ? SHK_F32 : SHK_F64, iter, type_ssor),
             "SOR");
  }

#ifdef SELDON_FILE_PRECOND_SSOR_CXX
  // Seldon's SorPreconditioner calls Seldon::SOR qualified
  // (Precond_Ssor.cxx:67-90), which binds to the overloads declared before it;
  // its Solve / TransSolve are therefore specialised here for the GPU
  // storages, with the same steps: z.Zero() when init_guess_null, then SOR
  // with type 0 (symmetric) or 2 / 3 (forward / backward).
#define SELDON_HIP_SOR_PRECOND(T, P, S)                                                                  \
  template <>                                                                                           \
  template <>                                                                                           \
  inline void SorPreconditioner<T>::Solve(const Matrix<T, P, S, HipAlloc<T> >& A,                       \
                                          const Vector<T, VectHip, HipAlloc<T> >& r,                    \
                                          Vector<T, VectHip, HipAlloc<T> >& z, bool init_guess_null)    \
  {                                                                                                     \
    if (init_guess_null) z.Zero();                                                                      \
    SOR(A, z, r, omega, nb_iter, symmetric_precond ? 0 : 2);                                           \
  }                                                                                                     \
  template <>                                                                                           \
  template <>                                                                                           \
  inline void SorPreconditioner<T>::TransSolve(const Matrix<T, P, S, HipAlloc<T> >& A,                  \
                                               const Vector<T, VectHip, HipAlloc<T> >& r,               \
                                               Vector<T, VectHip, HipAlloc<T> >& z, bool init_guess_null) \
  {                                                                                                     \
    if (init_guess_null) z.Zero();                                                                      \
    SOR(A, z, r, omega, nb_iter, symmetric_precond ? 0 : 3);                                           \
  }
  SELDON_HIP_SOR_PRECOND(double, General, RowSparseHip)
  SELDON_HIP_SOR_PRECOND(double, General, ArrayRowSparseHip)
  SELDON_HIP_SOR_PRECOND(double, Symmetric, RowSymSparseHip)
  SELDON_HIP_SOR_PRECOND(float, General, RowSparseHip)
  SELDON_HIP_SOR_PRECOND(float, General, ArrayRowSparseHip)
  SELDON_HIP_SOR_PRECOND(float, Symmetric, RowSymSparseHip)
#undef SELDON_HIP_SOR_PRECOND
#endif

  // Swap(X, Y): exchanges the two vectors' storage (Functions_Vector.cxx:295-304)
  template <class T, class Allocator>
  void Swap(Vector<T, VectHip, Allocator>& X, Vector<T, VectHip, Allocator>& Y)
  {
    X.SwapStorage(Y);
  }

  // Mlt(alpha, X): X *= alpha (Functions_Vector.cxx:69-74)
  template <class T0, class T1, class Allocator1>
  void Mlt(const T0 alpha, Vector<T1, VectHip, Allocator1>& X)
  {
    X *= alpha;
  }

  // Add(alpha, X, Y): Y += alpha X, alpha == 0 a no-op (Functions_Vector.cxx:85-105)
  template <class T0, class T1, class Allocator1, class T2, class Allocator2>
  void Add(const T0 alpha, const Vector<T1, VectHip, Allocator1>& X, Vector<T2, VectHip, Allocator2>& Y)
  {
    if (X.GetM() != Y.GetM())
      throw WrongDim("Add(alpha, X, Y)", "sizes do not match");
    HipCheck(shk_axpy(HipContext::Get().Handle(), HipType<T1>::value, X.GetM(), double(alpha), X.GetData(),
                      Y.GetData()),
             "Add");
  }

  // Copy(X, Y) (Functions_Vector.cxx:240-246)
  template <class T1, class Allocator1, class T2, class Allocator2>
  void Copy(const Vector<T1, VectHip, Allocator1>& X, Vector<T2, VectHip, Allocator2>& Y)
  {
    Y.Copy(X);
  }

  // DotProd / DotProdConj (Functions_Vector.cxx:330-346, :391-397)
  template <class T1, class Allocator1, class T2, class Allocator2>
  T1 DotProd(const Vector<T1, VectHip, Allocator1>& X, const Vector<T2, VectHip, Allocator2>& Y)
  {
    if (X.GetM() != Y.GetM())
      throw WrongDim("DotProd(X, Y)", "sizes do not match");
    double r = 0;
    HipCheck(shk_dot(HipContext::Get().Handle(), HipType<T1>::value, X.GetM(), X.GetData(), Y.GetData(), &r),
             "DotProd");
    return T1(r);
  }

  template <class T1, class Allocator1, class T2, class Allocator2>
  T1 DotProdConj(const Vector<T1, VectHip, Allocator1>& X, const Vector<T2, VectHip, Allocator2>& Y)
  {
    return DotProd(X, Y);
  }

  // Norm2 (Functions_Vector.cxx:540-549)
  template <class T1, class Allocator1>
  T1 Norm2(const Vector<T1, VectHip, Allocator1>& X)
  {
    double r = 0;
    HipCheck(shk_nrm2(HipContext::Get().Handle(), HipType<T1>::value, X.GetM(), X.GetData(), &r), "Norm2");
    return T1(r);
  }

  // --------------------------------------------------- preconditioners --
  // Jacobi, z_i = r_i / a_ii (new: Seldon has none; the division form of
  // SOR's / a_jj, matrix_sparse/Relaxation_MatVect.cxx:79).
  class JacobiPreconditioner : public Preconditioner_Base
  {
  public:
    template <class Matrix1, class Vector1>
    void Solve(const Matrix1& A, const Vector1& r, Vector1& z)
    {
      HipCheck(shk_jacobi(HipContext::Get().Handle(), A.GetHandle(), r.GetData(), z.GetData()),
               "JacobiPreconditioner::Solve");
    }
    template <class Matrix1, class Vector1>
    void TransSolve(const Matrix1& A, const Vector1& r, Vector1& z)
    {
      Solve(A, r, z);
    }
  };

  template <class P> struct HipPrecKind { static const int value = -1; };
  template <> struct HipPrecKind<Preconditioner_Base> { static const int value = SHK_PREC_IDENTITY; };
  template <> struct HipPrecKind<JacobiPreconditioner> { static const int value = SHK_PREC_JACOBI; };

  // Iteration's state is protected; these member pointers are formed through
  // a derived class (the standard way to reach protected members) so the
  // fused solvers can read and write it for both the real Seldon Iteration
  // and the standalone one.
  template <class Titer>
  struct HipIterationAccess : public Iteration<Titer>
  {
    static Titer Iteration<Titer>::*Tol() { return &HipIterationAccess::tolerance; }
    static Titer Iteration<Titer>::*Fac() { return &HipIterationAccess::facteur_reste; }
    static int Iteration<Titer>::*MaxIter() { return &HipIterationAccess::max_iter; }
    static int Iteration<Titer>::*PrintLevel() { return &HipIterationAccess::print_level; }
  };

  // Run one of the library's fused solvers with Seldon's Iteration object and
  // mirror the outcome back into it: iteration count, failure code, and the
  // residual history replayed through Iteration::Finished (so the messages
  // are Seldon's own, at the print level the caller chose).
  template <class Titer>
  int HipRunSolver(int which, shk_csr A, void* x, const void* b, int prec, Iteration<Titer>& iter)
  {
    typedef HipIterationAccess<Titer> Acc;
    shk_iteration it;
    std::memset(&it, 0, sizeof(it));
    it.tolerance = double(iter.*Acc::Tol());
    it.max_iter = iter.*Acc::MaxIter();
    it.nb_iter = iter.GetNumberIteration();
    it.init_guess_null = iter.IsInitGuess_Null() ? 1 : 0;
    it.parameter_restart = iter.GetRestart();
    const int print_level = iter.*Acc::PrintLevel();
    std::vector<double> hist;
    std::vector<int> hist_it;
    if (print_level >= 1)
      {
        size_t cap = 2 * (size_t)std::max(it.max_iter, 0) + 64;
        hist.resize(cap);
        hist_it.resize(cap);
        it.hist_cap = (int)cap;
        it.hist = &hist[0];
        it.hist_it = &hist_it[0];
      }
    int code = 0;
    shk_ctx ctx = HipContext::Get().Handle();
    if (which == 0 || which == 9) HipCheck(shk_cg(ctx, A, x, b, (shk_precond)prec, &it, &code), "Cg");
    else if (which == 1) HipCheck(shk_bicgstab(ctx, A, x, b, (shk_precond)prec, &it, &code), "BiCgStab");
    else if (which == 2) HipCheck(shk_gmres(ctx, A, x, b, (shk_precond)prec, &it, &code), "Gmres");
    else if (which == 3) HipCheck(shk_cgs(ctx, A, x, b, (shk_precond)prec, &it, &code), "Cgs");
    else if (which == 4) HipCheck(shk_bicg(ctx, A, x, b, (shk_precond)prec, &it, &code), "BiCg");
    else if (which == 5) HipCheck(shk_tfqmr(ctx, A, x, b, (shk_precond)prec, &it, &code), "TfQmr");
    else if (which == 6) HipCheck(shk_qmr(ctx, A, x, b, (shk_precond)prec, &it, &code), "Qmr");
    else if (which == 7) HipCheck(shk_cgne(ctx, A, x, b, (shk_precond)prec, &it, &code), "Cgne");
    else if (which == 8) HipCheck(shk_lsqr(ctx, A, x, b, (shk_precond)prec, &it, &code), "Lsqr");
    else if (which == 10) HipCheck(shk_bicgcr(ctx, A, x, b, (shk_precond)prec, &it, &code), "BiCgcr");
    else if (which == 11) HipCheck(shk_qmrsym(ctx, A, x, b, (shk_precond)prec, &it, &code), "QmrSym");
    else if (which == 12) HipCheck(shk_minres(ctx, A, x, b, (shk_precond)prec, &it, &code), "MinRes");
    else if (which == 13) HipCheck(shk_symmlq(ctx, A, x, b, (shk_precond)prec, &it, &code), "Symmlq");
    else if (which == 14) HipCheck(shk_qcgs(ctx, A, x, b, (shk_precond)prec, &it, &code), "QCgs");
    else if (which == 15) HipCheck(shk_gcr(ctx, A, x, b, (shk_precond)prec, &it, &code), "Gcr");
    else HipCheck(shk_bicgstabl(ctx, A, x, b, (shk_precond)prec, &it, &code), "BiCgStabl");
    iter.*Acc::Fac() = Titer(it.facteur_reste);
    const int n = std::min(it.hist_len, it.hist_cap);
    for (int k = 0; k < n; k++)
      {
        iter.SetNumberIteration(hist_it[k]);
        if (it.facteur_reste != 0)
          iter.Finished(Titer(hist[k] / it.facteur_reste));
      }
    iter.SetNumberIteration(it.nb_iter);
    if (it.fail_convergence)
      {
        static const char* what[17][7] = {{"", "Cg breakdown #1", "Cg breakdown #2", "", "", "", ""},
                                         {"", "Bicgstab breakdown #1", "Bicgstab breakdown #2",
                                          "Bicgstab breakdown #3", "", "", ""},
                                         {"", "", "", "", "", "", ""},
                                         {"", "Cgs breakdown #1", "Cgs breakdown #2", "", "", "", ""},
                                         {"", "Bicg breakdown #1", "Bicg breakdown #2", "", "", "", ""},
                                         {"", "Tfqmr breakdown: sigma=0", "Tfqmr breakdown: alpha=0",
                                          "Tfqmr breakdown: tau=0", "Tfqmr breakdown: tau=0",
                                          "tfqmr breakdown: beta=0", ""},
                                         {"", "Qmr breakdown #1", "Qmr breakdown #2", "Qmr breakdown #3",
                                          "Qmr breakdown #4", "Qmr breakdown #5", "Qmr breakdown #6"},
                                         {"", "Cgne breakdown #1", "Cgne breakdown #2", "", "", "", ""},
                                         {"", "Lsqr breakdown #1", "Lsqr breakdown #2", "Lsqr breakdown #3",
                                          "", "", ""},
                                         {"", "Cocg breakdown #1", "Cocg breakdown #2", "", "", "", ""},
                                         {"", "Bicgcr breakdown #1", "Bicgcr breakdown #2", "", "", "", ""},
                                         {"", "Qmr breakdown #1", "", "Qmr breakdown #2", "Qmr breakdown #3",
                                          "Qmr breakdown #4", "Qmr breakdown #5"},
                                         {"", "Minres breakdown #1", "MinRes breakdown #2", "", "", "", ""},
                                         {"", "", "", "", "", "", ""},
                                         {"", "QCgs breakdown #1", "Qcgs breakdown #2", "Qcgs breakdown #3", "", "",
                                          ""},
                                         {"", "Gcr breakdown #1", "", "", "", "", ""},
                                         {"", "Bicgstabl breakdown #1", "Bicgstabl Breakdown #2", "", "", "",
                                          ""}};
        const int c = it.error_code;
        iter.Fail(c, (c >= 0 && c < 7) ? what[which][c] : "breakdown");
      }
    return code;
  }

  // ------------------------------------- fused solver specialisations --
  // More specialised than Seldon's Cg / BiCgStab / Gmres templates
  // (Cg.cxx:45-47, BiCgStab.cxx:43-45, Gmres.cxx:44-46): chosen for GPU
  // storages with the identity or Jacobi preconditioner.  Any other
  // preconditioner falls through to Seldon's own templates, which then run
  // on the GPU through the overload set above.
  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  Cg(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
     Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(0, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  BiCgStab(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
           Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(1, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  // Cgs.cxx:44-46, BiCg.cxx:45-47, TfQmr.cxx:44-46 and Qmr.cxx:41-43: the
  // same specialisation for the Lanczos-type solvers the library fuses
  // (BiCg's and Qmr's A^T products on the library's transpose, built on
  // first use)
  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  Cgs(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
      Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(3, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  BiCg(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
       Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(4, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  TfQmr(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
        Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(5, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  Qmr(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
      Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(6, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  // CoCg.cxx:43-45: for real values the loop is Cg's, operation for
  // operation (DotProd and DotProdConj agree on reals): the fused Cg runs it
  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  CoCg(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
       Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(9, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  // Gcr.cxx:44-46
  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  Gcr(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
      Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(15, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  // BiCgStabl.cxx:42-44
  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  BiCgStabl(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
            Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(16, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  // QCgs.cxx:45-47
  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  QCgs(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
       Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(14, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  // Symmlq.cxx:44-46
  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  Symmlq(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
         Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(13, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  // MinRes.cxx:46-48
  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  MinRes(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
         Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(12, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  // QmrSym.cxx:41-43
  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  QmrSym(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
         Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(11, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  // BiCgcr.cxx:44-46
  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  BiCgcr(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
         Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(10, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  // Cgne.cxx:40-42 and Lsqr.cxx:40-42 (Lsqr applies no preconditioner, in
  // the template as here)
  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  Cgne(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
       Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(7, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  Lsqr(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
       Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(8, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

  template <class Titer, class T, class Prop, class Storage, class Alloc, class AllocV, class Prec>
  typename std::enable_if<IsHipSparse<Storage>::value && (HipPrecKind<Prec>::value >= 0), int>::type
  Gmres(Matrix<T, Prop, Storage, Alloc>& A, Vector<T, VectHip, AllocV>& x, const Vector<T, VectHip, AllocV>& b,
        Prec& M, Iteration<Titer>& iter)
  {
    return HipRunSolver(2, A.GetHandle(), x.GetData(), b.GetData(), HipPrecKind<Prec>::value, iter);
  }

} // namespace Seldon

#endif
