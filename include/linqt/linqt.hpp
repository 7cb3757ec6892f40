// linqt.hpp -- C++ host API of LinQT's Chebyshev-moment path, implemented over
// the gfx950 C ABI (linqt_amd.h).
//
// Keeps the reference's class and function names, argument meanings and
// file formats so that code written against
//   include/sparse_matrix.hpp, include/chebyshev_moments.hpp,
//   include/chebyshev_solver.hpp, include/Kubo_solver_FFT.hpp,
//   include/quantum_states.hpp
// compiles and runs unchanged (per-name headers next to this one include
// it).  The three solvers (SpectralMoments, CorrelationExpansionMoments,
// Kubo_solver_FFT::compute) run as single fused device pipelines; the
// remaining members (Multiply, Vectors::IterateAll, ...) keep their host
// semantics with device matrix products.  Errors abort with the C ABI's
// message, as the reference's asserts do.
#ifndef LINQT_AMD_LINQT_HPP
#define LINQT_AMD_LINQT_HPP

#include <array>
#include <cassert>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "../linqt_amd.h"

namespace linqt_amd {
// Process-wide context on device $LINQT_DEVICE (default 0), created on first use.
lq_ctx default_context();
// Aborts with the C ABI's last error when rc != LQ_OK.
void check(int rc, const char *what);
// A moment file (text or binary) through lq_mom_file_read: header and the
// M0*M1 interleaved complex values.
lq_mom_header read_moments(const std::string &path, std::vector<double> &mu);
}  // namespace linqt_amd

// ---------------------------------------------------------------------------
// include/sparse_matrix.hpp
namespace Sparse {
bool OPERATOR_FromCSRFile(const std::string input, int &dim, std::vector<int> &columns, std::vector<int> &rowIndex,
                          std::vector<std::complex<double> > &values);
}

class SparseMatrixType_BASE {
 public:
  typedef std::complex<double> value_t;
  typedef std::vector<value_t> vector_t;
  int numRows() { return numRows_; }
  int numCols() { return numCols_; }
  int rank() { return numRows() > numCols() ? numCols() : numRows(); }
  void setDimensions(const int r, const int c) {
    numRows_ = r;
    numCols_ = c;
  }
  void SetID(std::string id) { id_ = id; }
  std::string ID() const { return id_; }
  bool isIdentity() { return ID() == "1"; }

 private:
  int numRows_ = 0, numCols_ = 0;
  std::string id_;
};

// A Hermitian (upper fill) operator resident on the GPU.
class SparseMatrixType : public SparseMatrixType_BASE {
 public:
  SparseMatrixType() = default;
  SparseMatrixType(const SparseMatrixType &) = delete;
  SparseMatrixType &operator=(const SparseMatrixType &) = delete;
  ~SparseMatrixType();

  std::string matrixType() const { return "CSR Matrix on the GPU (gfx950)."; }
  void Multiply(const value_t a, const value_t *x, const value_t b, value_t *y);
  void Multiply(const value_t a, const vector_t &x, const value_t b, vector_t &y);
  void Multiply(const value_t *x, value_t *y) { Multiply(value_t(1, 0), x, value_t(0, 0), y); }
  void Multiply(const vector_t &x, vector_t &y) { Multiply(value_t(1, 0), x, value_t(0, 0), y); }
  void Rescale(const value_t a, const value_t b);
  // x, y column-major n x batchSize, as mkl_sparse_z_mm with SPARSE_LAYOUT_COLUMN_MAJOR
  void BatchMultiply(const int batchSize, const value_t a, const value_t *x, const value_t b, value_t *y);
  void ConvertFromCOO(std::vector<int> &rows, std::vector<int> &cols, std::vector<std::complex<double> > &vals);
  void ConvertFromCSR(std::vector<int> &rowIndex, std::vector<int> &cols, std::vector<std::complex<double> > &vals);
  std::vector<int> *rows() { return &rows_; }
  std::vector<int> *cols() { return &cols_; }
  std::vector<std::complex<double> > *vals() { return &vals_; }

  lq_csr handle() const { return h_; }
  // The stored arrays as a general (non-completed) matrix, made on first use:
  // the matrix the reference's Eigen map of these arrays stands for.
  lq_csr general_handle();

 private:
  lq_csr gh_ = nullptr;
  void device_vectors(size_t bytes);
  std::vector<int> rows_, cols_;
  std::vector<std::complex<double> > vals_;
  lq_csr h_ = nullptr;
  void *dx_ = nullptr, *dy_ = nullptr;
  size_t dbytes_ = 0;
};

class SparseMatrixBuilder {
 public:
  void setSparseMatrix(SparseMatrixType *b) { _matrix_type = b; }
  void BuildOPFromCSRFile(const std::string input);
  SparseMatrixType *_matrix_type = nullptr;
};

// ---------------------------------------------------------------------------
// include/linear_algebra.hpp (host vector helpers)
namespace linalg {
void scal(const int dim, std::complex<double> a, std::complex<double> *x);
void scal(const std::complex<double> &a, std::vector<std::complex<double> > &x);
void axpy(const int dim, std::complex<double> a, const std::complex<double> *x, std::complex<double> *y);
void axpy(const std::complex<double> &a, const std::vector<std::complex<double> > &x,
          std::vector<std::complex<double> > &y);
void copy(const int dim, const std::complex<double> *x, std::complex<double> *y);
void copy(const std::vector<std::complex<double> > &x, std::vector<std::complex<double> > &y);
std::complex<double> vdot(const std::vector<std::complex<double> > &x, const std::vector<std::complex<double> > &y);
std::complex<double> vdot(const int dim, const std::complex<double> *x, std::complex<double> *y);
double nrm2(const std::vector<std::complex<double> > &x);
double nrm2(const int dim, const std::complex<double> *x);
void batch_vdot(const int dim, const int batchSize, const std::complex<double> *leftb,
                const std::complex<double> *rightb, std::complex<double> *output);
void extract_segment(const std::vector<std::complex<double> > &x, size_t size_x, size_t start_x,
                     std::vector<std::complex<double> > &y, size_t size_y);
void introduce_segment(const std::vector<std::complex<double> > &x, size_t size_x,
                       std::vector<std::complex<double> > &y, size_t size_y, size_t start_y);
}  // namespace linalg
class SparseMatrixType;
namespace linalg {
// src/mkl_linear_algebra.cpp:100-131: orthogonalized = S^-1 original by
// BiCGSTAB (tol 1e-4, 2000 iterations) on the GPU (lq_overlap_solve).
void orthogonalize(SparseMatrixType &S, std::vector<std::complex<double> > &orthogonalized,
                   std::vector<std::complex<double> > &original);
}  // namespace linalg

// ---------------------------------------------------------------------------
// include/vector_list.hpp
template <class T>
class VectorList {
 public:
  typedef std::vector<T> vector_t;
  typedef std::vector<vector_t> list_t;
  VectorList() {}
  VectorList(size_t list_count, size_t count, const T value = T(0))
      : _list_size(list_count), _vector_size(count), _data(list_t(list_count)) {
    for (auto &e : _data) e = vector_t(count, value);
  }
  void Resize(size_t list_count, size_t count, const T value = T(0)) {
    _list_size = list_count;
    _vector_size = count;
    _data.resize(list_count);
    for (auto &e : _data) e = vector_t(count, value);
  }
  size_t VectorSize() const { return _vector_size; }
  size_t ListSize() const { return _list_size; }
  list_t &List() { return _data; }
  vector_t &ListElem(const size_t i) { return _data[i]; }
  T &operator()(const size_t i, const size_t j) { return _data[i][j]; }

 private:
  size_t _list_size = 0, _vector_size = 0;
  list_t _data;
};

// ---------------------------------------------------------------------------
// include/quantum_states.hpp
enum StateType { LOCAL_STATE, USER_STATE, RANDOM_STATE };

namespace qstates {
typedef std::complex<double> scalar;

struct generator {
  generator() : label("default"), count(0), dim(0), num_states(1), kind(RANDOM_STATE) {}
  int SystemSize() const { return dim; }
  int NumberOfStates() const { return num_states; }
  void NumberOfStates(const int n) {
    num_states = n;
    spos.reserve(n);
  }
  void SystemSize(const int n) {
    if (SystemSize() == 0) {
      dim = n;
      out = std::vector<scalar>(dim);
      return;
    }
    assert(SystemSize() == n);
  }
  // The "random" kind is a state localised at site dim/2 (include/quantum_states.hpp:85-93).
  bool getQuantumState();
  std::string StateLabel() { return label; }
  std::vector<scalar> &State() { return out; }

  std::string label;
  int count;
  std::vector<scalar> out;
  std::vector<scalar> data;
  std::vector<int> spos;
  int dim;
  int num_states;
  StateType kind;
};

StateType String2State(std::string kind);
void readLocalState(std::ifstream &file, generator &data);
void readCustomState(std::ifstream &file, generator &data);
generator CreateStateSet(std::ifstream &file, std::string kind);
generator LoadStateFile(std::string filename);
}  // namespace qstates

// ---------------------------------------------------------------------------
// include/chebyshev_coefficients.hpp
inline double delta_chebF(double x, const double m) {
  const double f0 = std::sqrt(1.0 - x * x);
  return std::cos(m * std::acos(x)) / f0 / M_PI;
}
inline std::complex<double> greenR_chebF(double x, const double m) {
  const std::complex<double> I(0.0, 1.0);
  const double f0 = std::sqrt(1.0 - x * x);
  return -I * std::pow(x - I * f0, m) / f0;
}
inline std::complex<double> DgreenR_chebF(double x, const double m) {
  const std::complex<double> I(0.0, 1.0);
  const double f0 = std::sqrt(1.0 - x * x);
  return -I * std::pow(x - I * f0, m) * (x + I * m * f0) / f0 / f0 / f0;
}

// ---------------------------------------------------------------------------
// include/chebyshev_moments.hpp
namespace chebyshev {
const double CUTOFF = 1.00;

class Moments {
 public:
  typedef std::complex<double> value_t;
  typedef std::vector<value_t> vector_t;

  Moments() {}
  void getMomentsParams(Moments &mom) {
    SetHamiltonian(mom.Hamiltonian());
    SystemLabel(mom.SystemLabel());
    BandWidth(mom.BandWidth());
    BandCenter(mom.BandCenter());
  }
  size_t SystemSize() const { return system_size; }
  std::string SystemLabel() const { return system_label; }
  double BandWidth() const { return band_width; }
  double HalfWidth() const { return BandWidth() / 2.0; }
  double BandCenter() const { return band_center; }
  double ScaleFactor() const { return CUTOFF / HalfWidth(); }
  double ShiftFactor() const { return -BandCenter() / HalfWidth() / CUTOFF; }
  vector_t &MomentVector() { return mu; }
  value_t &MomentVector(const int i) { return mu[i]; }
  vector_t &Chebyshev0() { return ChebV0; }
  vector_t &Chebyshev1() { return ChebV1; }
  SparseMatrixType &Hamiltonian() const { return *_pNHAM; }
  bool HasHamiltonian() const { return _pNHAM != nullptr; }

  void SetHamiltonian(SparseMatrixType &NHAM) {
    if (SystemSize() == 0) SystemSize(NHAM.rank());
    assert((size_t)NHAM.rank() == SystemSize());
    _pNHAM = &NHAM;
  }
  int Rescale2ChebyshevDomain() {
    Hamiltonian().Rescale(ScaleFactor(), ShiftFactor());
    return 0;
  }
  void SetAndRescaleHamiltonian(SparseMatrixType &NHAM) {
    SetHamiltonian(NHAM);
    Rescale2ChebyshevDomain();
  }
  void SystemSize(const int dim) { system_size = dim; }
  void SystemLabel(std::string label) { system_label = label; }
  void BandWidth(const double x) { band_width = x; }
  void BandCenter(const double x) { band_center = x; }
  void MomentVector(const vector_t _mu) { mu = _mu; }
  void SetInitVectors(const vector_t &T0);
  void SetInitVectors(SparseMatrixType &OP, const vector_t &T0);
  double Rescale2ChebyshevDomain(const double energ) { return (energ - BandCenter()) / HalfWidth(); }
  int Iterate();
  int JacksonKernelMomCutOff(const double broad);
  double JacksonKernel(const double m, const double Mom);

 private:
  SparseMatrixType *_pNHAM = nullptr;
  vector_t ChebV0, ChebV1;
  std::string system_label;
  size_t system_size = 0;
  double band_width = 0, band_center = 0;
  vector_t mu;
};

class Moments1D : public Moments {
 public:
  Moments1D() : _numMoms(0) {}
  Moments1D(const size_t m0) : _numMoms(m0) { MomentVector(vector_t(_numMoms, 0.0)); }
  Moments1D(std::string momfilename);
  size_t MomentNumber() const { return _numMoms; }
  size_t HighestMomentNumber() const { return _numMoms; }
  value_t &operator()(const size_t m0) { return MomentVector((int)m0); }
  void MomentNumber(const size_t numMoms);
  void saveIn(std::string filename);
  void ApplyJacksonKernel(const double broad);
  void Print();

 private:
  size_t _numMoms;
};

// include/chebyshev_moments.hpp:194-218
class Moments1D_nonOrth : public Moments1D {
 public:
  Moments1D_nonOrth(const size_t m0) : Moments1D(m0), S_(nullptr) {}
  Moments1D_nonOrth(SparseMatrixType &S) : Moments1D(), S_(&S) {}
  Moments1D_nonOrth(const size_t m0, SparseMatrixType &S) : Moments1D(m0), S_(&S) {}
  Moments1D_nonOrth(std::string &momfilename, SparseMatrixType &S) : Moments1D(momfilename), S_(&S) {}
  void set_S(SparseMatrixType &S) { S_ = &S; }
  SparseMatrixType *S() const { return S_; }
  void SetInitVectors_nonOrthogonal(vector_t &T0);
  void SetInitVectors_nonOrthogonal(SparseMatrixType &OP, vector_t &T0);
  int Iterate_nonOrthogonal();

 private:
  SparseMatrixType *S_;
};

class Moments2D : public Moments {
 public:
  Moments2D() : numMoms({0, 0}) {}
  Moments2D(const size_t m0, const size_t m1) : numMoms({int(m0), int(m1)}) {
    MomentVector(vector_t((size_t)numMoms[1] * numMoms[0], 0.0));
  }
  Moments2D(std::string momfilename);
  Moments2D(Moments2D mom, const size_t m0, const size_t m1) {
    getMomentsParams(mom);
    numMoms = {int(m0), int(m1)};
    MomentVector(vector_t((size_t)numMoms[1] * numMoms[0], 0.0));
  }
  std::array<int, 2> MomentNumber() const { return numMoms; }
  int HighestMomentNumber(const int i) const { return numMoms[i]; }
  int HighestMomentNumber() const { return numMoms[1] > numMoms[0] ? numMoms[1] : numMoms[0]; }
  void MomentNumber(const int mom0, const int mom1);
  value_t &operator()(const int m0, const int m1) { return MomentVector(m0 * numMoms[1] + m1); }
  void ApplyJacksonKernel(const double b0, const double b1);
  void saveIn(std::string filename);
  void AddSubMatrix(Moments2D &sub, const int mL, const int mR) {
    for (int m0 = 0; m0 < sub.HighestMomentNumber(0); m0++)
      for (int m1 = 0; m1 < sub.HighestMomentNumber(1); m1++) (*this)(mL + m0, mR + m1) += sub(m0, m1);
  }
  void InsertSubMatrix(Moments2D &sub, const int mL, const int mR) {
    for (int m0 = 0; m0 < sub.HighestMomentNumber(0); m0++)
      for (int m1 = 0; m1 < sub.HighestMomentNumber(1); m1++) (*this)(mL + m0, mR + m1) = sub(m0, m1);
  }
  void Print();

 private:
  std::array<int, 2> numMoms;
};

// Time-dependent moments mu(m, n), m < numMoms, n < maxTimeStep (row m stride
// maxTimeStep), include/chebyshev_moments.hpp:287-365.
class MomentsTD : public Moments {
 public:
  const double HBAR = 0.6582119624;  // eV fs
  MomentsTD() : _numMoms(1), _maxTimeStep(1), _timeStep(0), _dt(0) {}
  MomentsTD(const size_t m, const size_t n) : _numMoms(m), _maxTimeStep(n), _timeStep(0), _dt(0) {
    MomentVector(vector_t(m * n, 0.0));
  }
  MomentsTD(std::string momfilename);
  size_t MomentNumber() const { return _numMoms; }
  size_t HighestMomentNumber() const { return _numMoms; }
  size_t CurrentTimeStep() const { return _timeStep; }
  size_t MaxTimeStep() const { return _maxTimeStep; }
  double TimeDiff() const { return _dt; }
  double ChebyshevFreq() const { return HalfWidth() / CUTOFF / HBAR; }
  double ChebyshevFreq_0() const { return BandCenter() / HBAR; }
  void MomentNumber(const size_t mom);
  void MaxTimeStep(const size_t maxTimeStep) { _maxTimeStep = maxTimeStep; }
  void IncreaseTimeStep() { _timeStep++; }
  void ResetTime() { _timeStep = 0; }
  void TimeDiff(const double dt) { _dt = dt; }
  value_t &operator()(const size_t m, const size_t n) { return MomentVector((int)(m * MaxTimeStep() + n)); }
  void ApplyJacksonKernel(const double broad);
  void saveIn(std::string filename);
  void Print();

 private:
  size_t _numMoms, _maxTimeStep, _timeStep;
  double _dt;
};

// Stored Chebyshev vectors (host lists; the fused solvers never materialise them).
class Vectors_sliced : public Moments {
 public:
  typedef VectorList<Moments::value_t> vectorList_t;
  Vectors_sliced() { SetNumSections(1); }
  Vectors_sliced(Vectors_sliced &vec) { *this = vec; }
  Vectors_sliced(const int numVecs, const size_t num_sections) : numVecs_(numVecs), num_sections_((int)num_sections) {}
  Vectors_sliced(Moments &mom, const int numVecs, const size_t num_sections)
      : numVecs_(numVecs), num_sections_((int)num_sections) {
    getMomentsParams(mom);
    SetNumSections((int)num_sections);
  }
  int NumberOfVectors() const { return numVecs_; }
  int NumberOfSections() const { return num_sections_; }
  vector_t &OPV() { return OPV_; }
  int SectionSize() const { return (int)section_size_; }
  int LastSectionSize() const { return (int)last_section_size_; }
  void SetNumberOfVectors(const int x) { numVecs_ = x; }
  int SetNumSections(const int x) {
    num_sections_ = x;
    section_size_ = SystemSize() / num_sections_;
    last_section_size_ = SystemSize() / num_sections_ + SystemSize() % num_sections_;
    return 0;
  }
  int CreateVectorSet();
  void getMomentsParams(Moments &mom) {
    SetHamiltonian(mom.Hamiltonian());
    SystemLabel(mom.SystemLabel());
    BandWidth(mom.BandWidth());
    BandCenter(mom.BandCenter());
  }
  size_t Size() const { return section_size_ * (size_t)NumberOfVectors(); }
  double SizeInGB() const { return sizeof(value_t) * section_size_ * (double)NumberOfVectors() * std::pow(2.0, -30.0); }
  size_t HighestMomentNumber() const { return Chebmu_.ListSize(); }
  vectorList_t &List() { return Chebmu_; }
  vector_t &Vector(const size_t m0) { return Chebmu_.ListElem(m0); }
  value_t &operator()(const size_t m0) { return Chebmu_(m0, 0); }
  void operator=(Vectors_sliced &vec) {
    if (vec.HasHamiltonian()) SetHamiltonian(vec.Hamiltonian());
    SystemLabel(vec.SystemLabel());
    BandWidth(vec.BandWidth());
    BandCenter(vec.BandCenter());
    num_sections_ = vec.NumberOfSections();
    numVecs_ = vec.NumberOfVectors();
    section_size_ = vec.SectionSize();
    last_section_size_ = vec.LastSectionSize();
  }
  int IterateAllSliced(int s);
  int MultiplySliced(SparseMatrixType &OP, int s);
  double MemoryConsumptionInGB() { return SizeInGB() + 2.0 * (double)SystemSize() * std::pow(2.0, -30.0); }

 private:
  vector_t OPV_;
  int numVecs_ = 0, num_sections_ = 1;
  size_t section_size_ = 0, last_section_size_ = 0;
  vectorList_t Chebmu_;
};

class Vectors : public Vectors_sliced {
 public:
  Vectors(Moments1D &mom) {
    getMomentsParams(mom);
    SetNumSections(1);
    SetNumberOfVectors((int)mom.HighestMomentNumber());
    CreateVectorSet();
  }
  Vectors(Moments2D &mom) {
    getMomentsParams(mom);
    SetNumSections(1);
    SetNumberOfVectors(mom.HighestMomentNumber());
    CreateVectorSet();
  }
  int IterateAll() { return IterateAllSliced(0); }
  int Multiply(SparseMatrixType &OP) { return MultiplySliced(OP, 0); }
};

// ---------------------------------------------------------------------------
// include/kpm_noneqop.hpp
bool GetBatchSize(int &batchSize);

// ---------------------------------------------------------------------------
// include/chebyshev_solver.hpp
typedef std::complex<double> value_t;
typedef std::vector<value_t> vector_t;
namespace utility {
std::array<double, 2> SpectralBounds(SparseMatrixType &HAM);
}
int ComputeMomTable(chebyshev::Vectors &chevVecL, chebyshev::Vectors &chevVecR, chebyshev::Moments2D &sub);
int CorrelationExpansionMoments(const vector_t &PhiR, const vector_t &PhiL, SparseMatrixType &OPL,
                                SparseMatrixType &OPR, chebyshev::Vectors &chevVecL, chebyshev::Vectors &chevVecR,
                                chebyshev::Moments2D &chebMoms);
int CorrelationExpansionMoments(SparseMatrixType &OPL, SparseMatrixType &OPR, chebyshev::Moments2D &chebMoms,
                                qstates::generator &gen);
int SpectralMoments(SparseMatrixType &OP, chebyshev::Moments1D &chebMoms, qstates::generator &gen);
// src/chebyshev_solver.cpp:203-229: one device solve (lq_spectral_moments_nonorth).
int SpectralMoments_nonOrth(SparseMatrixType &OP, chebyshev::Moments1D_nonOrth &chebMoms, qstates::generator &gen);
int TimeDependentCorrelations(SparseMatrixType &OPL, SparseMatrixType &OPR, chebyshev::MomentsTD &chebMoms,
                              qstates::generator &gen);

// ---------------------------------------------------------------------------
// include/Kubo_solver_FFT.hpp
enum formula { NIL = -1, KUBO_GREENWOOD = 0, KUBO_BASTIN = 1 };

class Kubo_solver_FFT {
 public:
  typedef double r_value_t;
  Kubo_solver_FFT(int M, int num_sections, int nump, formula sym_formula, chebyshev::Vectors_sliced &chebVec,
                  std::string &outputfilename)
      : M_(M), num_sections_(num_sections), nump_(nump), sym_formula_(sym_formula), chebVec_(chebVec),
        outputfilename_(outputfilename) {}
  // Runs the whole solve on the device and writes the output file(s) with
  // the mean over all states (the reference rewrites them after each state;
  // the final file is the same).
  void compute(SparseMatrixType &OPL, SparseMatrixType &OPR, qstates::generator &gen);
  int nump() { return nump_; }
  int num_sections() { return num_sections_; }
  int num_vectors() { return M_; }
  std::string &OutputFilename() { return outputfilename_; }
  formula SolverFormula() { return sym_formula_; }
  // final_data of the last compute: mean over states of r_data (2*nump)
  const std::vector<std::complex<double> > &FinalData() const { return final_data_; }

 private:
  int M_, num_sections_, nump_;
  formula sym_formula_;
  chebyshev::Vectors_sliced &chebVec_;
  std::string outputfilename_;
  std::vector<std::complex<double> > final_data_;
};

}  // namespace chebyshev

#endif
