// SeldonCore.hxx -- standalone Seldon-compatible core for builds without the
// Seldon headers (e.g. a GPU host that has only this library).
//
// Provides, in namespace Seldon and with Seldon's names and semantics, the
// pieces the sparse Krylov path needs around the GPU storages:
//   * the primary templates Vector<T, Storage, Allocator> and
//     Matrix<T, Prop, Storage, Allocator> (Seldon.hxx:25-40) and the tags
//     General / Symmetric, VectFull / RowSparse / RowSymSparse /
//     ArrayRowSparse (share/Properties.hxx:26,30; share/Storage.hxx:56-138);
//   * Iteration<Titer> with Seldon's stopping rule, messages and error codes
//     (computation/solver/iterative/Iterative.hxx:47-108, Iterative.cxx:81-354);
//   * Preconditioner_Base (Iterative.hxx:25-39);
//   * the exception classes of share/Errors.hxx used on this path;
//   * the transpose tags SeldonTrans / SeldonNoTrans / SeldonConjTrans
//     (share/MatrixFlag.hxx:32-80);
//   * a host dense vector Vector<T, VectFull> and host CSR
//     Matrix<T, General, RowSparse> / Matrix<T, Symmetric, RowSymSparse>
//     to build systems and move them to and from the GPU.
// When Seldon.hxx has been included first, this file is not used: the GPU
// storages then plug into the real Seldon types.
#ifndef SELDON_AMD_SELDON_CORE_HXX
#define SELDON_AMD_SELDON_CORE_HXX

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

namespace Seldon
{
  using std::cout;
  using std::endl;
  using std::string;

  template <class T, class Storage, class Allocator>
  class Vector
  {
  };

  template <class T, class Prop, class Storage, class Allocator>
  class Matrix
  {
  };

  class General {};
  class Symmetric {};
  class VectFull {};
  class RowSparse {};
  class RowSymSparse {};
  class ArrayRowSparse {};

  template <class T>
  class MallocAlloc {};

  // transpose tags (share/MatrixFlag.hxx:32-80): the global objects
  // SeldonTrans / SeldonNoTrans / SeldonConjTrans select the transposed
  // products; status 0: Trans, 1: NoTrans, 2: ConjTrans
  class SeldonTranspose
  {
  protected:
    int status_;
  public:
    explicit SeldonTranspose(int status) : status_(status) {}
    bool Trans() const { return status_ == 0; }
    bool NoTrans() const { return status_ == 1; }
    bool ConjTrans() const { return status_ == 2; }
  };
  class class_SeldonTrans : public SeldonTranspose
  {
  public:
    class_SeldonTrans() : SeldonTranspose(0) {}
  };
  class class_SeldonNoTrans : public SeldonTranspose
  {
  public:
    class_SeldonNoTrans() : SeldonTranspose(1) {}
  };
  class class_SeldonConjTrans : public SeldonTranspose
  {
  public:
    class_SeldonConjTrans() : SeldonTranspose(2) {}
  };
  static const class_SeldonTrans SeldonTrans;
  static const class_SeldonNoTrans SeldonNoTrans;
  static const class_SeldonConjTrans SeldonConjTrans;

  // ------------------------------------------------------------ errors --
  class Error
  {
  protected:
    string description_, function_, comment_;
  public:
    Error(string function = "", string comment = "")
      : description_("ERROR!"), function_(function), comment_(comment) {}
    Error(string description, string function, string comment)
      : description_(description), function_(function), comment_(comment) {}
    virtual ~Error() {}
    virtual string What()
    {
      string message(description_);
      if (!function_.empty())
        message += string(" Error in ") + function_ + ".";
      if (!comment_.empty())
        message += string(" ") + comment_;
      return message;
    }
    void CoutWhat() { cout << What() << endl; }
  };

  class WrongDim : public Error
  {
  public:
    WrongDim(string function = "", string comment = "")
      : Error("Wrong dimensions in function", function, comment) {}
  };

  class NoMemory : public Error
  {
  public:
    NoMemory(string function = "", string comment = "")
      : Error("Out of memory in function", function, comment) {}
  };

  class WrongArgument : public Error
  {
  public:
    WrongArgument(string function = "", string comment = "")
      : Error("Wrong argument given to function", function, comment) {}
  };

  // ------------------------------------------------------- Iteration --
  class Preconditioner_Base
  {
  public:
    Preconditioner_Base() {}
  };

  template <class Titer>
  class Iteration
  {
  protected:
    Titer tolerance;
    Titer facteur_reste;
    int max_iter;
    int nb_iter;
    int error_code;
    bool fail_convergence;
    int print_level;
    bool init_guess_null;
    int type_solver;
    int parameter_restart;
    int type_preconditioning;

  public:
    Iteration()
      : tolerance(1e-6), facteur_reste(0), max_iter(100), nb_iter(0), error_code(0),
        fail_convergence(false), print_level(1), init_guess_null(true), type_solver(0),
        parameter_restart(10), type_preconditioning(0) {}
    Iteration(int max_iteration, const Titer& tol)
      : tolerance(tol), facteur_reste(0), max_iter(max_iteration), nb_iter(0), error_code(0),
        fail_convergence(false), print_level(1), init_guess_null(true), type_solver(0),
        parameter_restart(10), type_preconditioning(0) {}
    // copy: as Seldon's, the iteration count restarts at 0
    Iteration(const Iteration<Titer>& outer)
      : tolerance(outer.tolerance), facteur_reste(outer.facteur_reste),
        max_iter(outer.max_iter), nb_iter(0), error_code(outer.error_code),
        fail_convergence(false), print_level(outer.print_level), init_guess_null(true),
        type_solver(outer.type_solver), parameter_restart(outer.parameter_restart),
        type_preconditioning(outer.type_preconditioning) {}

    int GetTypeSolver() const { return type_solver; }
    int GetRestart() const { return parameter_restart; }
    Titer GetFactor() const { return facteur_reste; }
    Titer GetTolerance() const { return tolerance; }
    int GetNumberIteration() const { return nb_iter; }

    void SetSolver(int type_resolution, int param_restart, int type_prec)
    {
      type_solver = type_resolution;
      parameter_restart = param_restart;
      type_preconditioning = type_prec;
    }
    void SetRestart(int m) { parameter_restart = m; }
    void SetTolerance(Titer stopping_criterion) { tolerance = stopping_criterion; }
    void SetMaxNumberIteration(int max_iteration) { max_iter = max_iteration; }
    void SetNumberIteration(int nb) { nb_iter = nb; }

    void ShowMessages() { print_level = 1; }
    void ShowFullHistory() { print_level = 6; }
    void HideMessages() { print_level = 0; }

    // 1/|r| for the relative residual; -1 when r == 0 (nb_iter untouched)
    template <class Vector1>
    int Init(const Vector1& r)
    {
      Titer norme_rhs = Titer(Norm2(r));
      if (norme_rhs == Titer(0))
        return -1;
      facteur_reste = Titer(1) / norme_rhs;
      nb_iter = 0;
      return 0;
    }

    bool First() const { return nb_iter == 0; }
    bool IsInitGuess_Null() const { return init_guess_null; }
    void SetInitGuess(bool type) { init_guess_null = type; }

    template <class Vector1>
    bool Finished(const Vector1& r) const
    {
      return Finished(Titer(Norm2(r)));
    }

    bool Finished(const Titer& r) const
    {
      Titer reste = facteur_reste * r;
      if ((print_level >= 1 && nb_iter % 100 == 0) || print_level >= 6)
        cout << "Residu at iteration number " << GetNumberIteration() << "  " << reste << endl;
      return reste < tolerance || nb_iter >= max_iter;
    }

    void Fail(int i, const string& s)
    {
      fail_convergence = true;
      error_code = i;
      if (print_level >= 1 && nb_iter % 100 == 0)
        cout << "Error during resolution : " << s << endl;
    }

    Iteration& operator++(void)
    {
      ++nb_iter;
      return *this;
    }

    int ErrorCode() const
    {
      if (nb_iter >= max_iter)
        return -2;
      if (fail_convergence)
        return error_code;
      return 0;
    }
  };

  // ------------------------------------------------ host dense vector --
  template <class T, class Allocator>
  class Vector<T, VectFull, Allocator>
  {
  protected:
    std::vector<T> v_;
  public:
    typedef T value_type;
    Vector() {}
    explicit Vector(int m) : v_(m > 0 ? m : 0) {}
    int GetM() const { return (int)v_.size(); }
    int GetLength() const { return (int)v_.size(); }
    int GetSize() const { return (int)v_.size(); }
    void Reallocate(int m) { v_.assign(m > 0 ? m : 0, T(0)); }
    void Resize(int m) { v_.resize(m > 0 ? m : 0, T(0)); }
    T* GetData() { return v_.empty() ? 0 : &v_[0]; }
    const T* GetData() const { return v_.empty() ? 0 : &v_[0]; }
    T& operator()(int i) { return v_[i]; }
    const T& operator()(int i) const { return v_[i]; }
    void Zero() { std::fill(v_.begin(), v_.end(), T(0)); }
    template <class T0>
    void Fill(const T0& x) { std::fill(v_.begin(), v_.end(), T(x)); }
  };

  // --------------------------------------------------- host CSR types --
  // 0-based CSR as Matrix_Sparse holds it (matrix_sparse/Matrix_Sparse.hxx:
  // 42-142); RowSymSparse keeps the upper triangle with the diagonal.
  template <class T>
  class HostCsr
  {
  protected:
    int m_, n_;
    std::vector<int> ptr_, ind_;
    std::vector<T> data_;
  public:
    typedef T value_type;
    HostCsr() : m_(0), n_(0), ptr_(1, 0) {}
    void SetData(int m, int n, const std::vector<T>& values, const std::vector<int>& ptr,
                 const std::vector<int>& ind)
    {
      if ((int)ptr.size() != m + 1 || ind.size() != values.size() || ptr[m] != (int)ind.size())
        throw WrongDim("Matrix::SetData", "inconsistent CSR arrays");
      m_ = m; n_ = n; ptr_ = ptr; ind_ = ind; data_ = values;
    }
    int GetM() const { return m_; }
    int GetN() const { return n_; }
    int GetNonZeros() const { return (int)ind_.size(); }
    int GetDataSize() const { return (int)ind_.size(); }
    const int* GetPtr() const { return &ptr_[0]; }
    const int* GetInd() const { return ind_.empty() ? 0 : &ind_[0]; }
    const T* GetData() const { return data_.empty() ? 0 : &data_[0]; }
  };

  template <class T, class Prop, class Allocator>
  class Matrix<T, Prop, RowSparse, Allocator> : public HostCsr<T> {};

  template <class T, class Prop, class Allocator>
  class Matrix<T, Prop, RowSymSparse, Allocator> : public HostCsr<T> {};

} // namespace Seldon

#endif
