// linqt_facade.cpp -- implementation of include/linqt/linqt.hpp over the C ABI.
//
// Host-only code (g++): every device operation goes through linqt_amd.h.
#include "../../include/linqt/linqt.hpp"

#include <cstdio>
#include <cstring>
#include <mutex>

namespace linqt_amd {

void check(int rc, const char *what) {
  if (rc != LQ_OK) {
    std::cerr << "linqt_amd: " << what << " failed (" << rc << "): " << lq_last_error() << std::endl;
    std::abort();
  }
}

// Moments file constructors: native reader (text or binary), as `>>` reads the text.
lq_mom_header read_moments(const std::string &path, std::vector<double> &mu) {
  lq_mom_header h{};
  check(lq_mom_file_read(path.c_str(), &h, nullptr), "lq_mom_file_read");
  mu.assign(2 * (size_t)h.M0 * (size_t)h.M1, 0.0);
  check(lq_mom_file_read(path.c_str(), &h, mu.data()), "lq_mom_file_read");
  return h;
}

lq_ctx default_context() {
  static lq_ctx ctx = nullptr;
  static std::once_flag once;
  std::call_once(once, [] {
    const char *d = getenv("LINQT_DEVICE");
    check(lq_ctx_create(d ? atoi(d) : 0, &ctx), "lq_ctx_create");
  });
  return ctx;
}

}  // namespace linqt_amd

using linqt_amd::check;
using linqt_amd::default_context;
typedef std::complex<double> cplx;

// ---- Sparse::OPERATOR_FromCSRFile (src/sparse_matrix.cpp:3-50) -------------
bool Sparse::OPERATOR_FromCSRFile(const std::string input, int &dim, std::vector<int> &columns,
                                  std::vector<int> &rowIndex, std::vector<cplx> &values) {
  std::cout << "\nReading the CSR file located at: " << input << std::endl;
  // native multi-threaded parse (text or binary CSR), values bitwise as `>>` reads them
  long long nnz = 0;
  if (lq_csr_file_read(input.c_str(), &dim, &nnz, nullptr, nullptr, nullptr) != LQ_OK) {
    std::cerr << "ERROR: Can not open the matrix file (" << lq_last_error() << ")";
    return false;
  }
  values.assign((size_t)nnz, cplx(0));
  columns.assign((size_t)nnz, 0);
  rowIndex.assign((size_t)dim + 1, 0);
  if (lq_csr_file_read(input.c_str(), &dim, &nnz, rowIndex.data(), columns.data(),
                       reinterpret_cast<double *>(values.data())) != LQ_OK) {
    std::cerr << "ERROR: " << lq_last_error() << std::endl;
    return false;
  }
  std::cout << "FINISH READING.Status: SUCCEED" << std::endl;
  return true;
}

// ---- SparseMatrixType ---------------------------------------------------------
SparseMatrixType::~SparseMatrixType() {
  if (dx_) lq_free(default_context(), dx_);
  if (dy_) lq_free(default_context(), dy_);
  if (h_) lq_csr_destroy(h_);
  if (gh_) lq_csr_destroy(gh_);
}

lq_csr SparseMatrixType::general_handle() {
  if (!gh_)
    check(lq_csr_create(default_context(), numRows(), (int)cols_.size(), rows_.data(), cols_.data(),
                        reinterpret_cast<const double *>(vals_.data()), LQ_FILL_GENERAL, &gh_),
          "lq_csr_create (general)");
  return gh_;
}

// ---- linalg::orthogonalize (src/mkl_linear_algebra.cpp:100-131) ----------------------
void linalg::orthogonalize(SparseMatrixType &S, std::vector<cplx> &orthogonalized, std::vector<cplx> &original) {
  const int DIM = S.numRows();
  lq_ctx ctx = default_context();
  void *db = nullptr, *dx = nullptr;
  check(lq_malloc(ctx, sizeof(cplx) * (size_t)DIM, &db), "lq_malloc");
  check(lq_malloc(ctx, sizeof(cplx) * (size_t)DIM, &dx), "lq_malloc");
  check(lq_memcpy_h2d(ctx, db, original.data(), sizeof(cplx) * (size_t)DIM), "lq_memcpy_h2d");
  int its = 0;
  double err = 0;
  check(lq_overlap_solve(S.general_handle(), (const double *)db, (double *)dx, 1e-4, 2000, &its, &err),
        "lq_overlap_solve");
  orthogonalized.resize((size_t)DIM);
  check(lq_memcpy_d2h(ctx, orthogonalized.data(), dx, sizeof(cplx) * (size_t)DIM), "lq_memcpy_d2h");
  lq_free(ctx, db);
  lq_free(ctx, dx);
  double nrm = 0;
  for (const auto &z : original) nrm += std::norm(z);
  std::cout << "#iterations:     " << its << std::endl;
  std::cout << "  max#iterations:" << 2000 << std::endl;
  std::cout << "estimated error: " << err << std::endl;
  std::cout << "  tolerance :    " << 1e-4 << std::endl;
  std::cout << "Vector norm :    " << std::sqrt(nrm) << std::endl << std::endl;
}

void SparseMatrixType::ConvertFromCSR(std::vector<int> &rowIndex, std::vector<int> &cols, std::vector<cplx> &vals) {
  rows_ = rowIndex;
  cols_ = cols;
  vals_ = vals;
  if (h_) lq_csr_destroy(h_);
  h_ = nullptr;
  check(lq_csr_create(default_context(), numRows(), (int)cols_.size(), rows_.data(), cols_.data(),
                      reinterpret_cast<const double *>(vals_.data()), LQ_FILL_HERMITIAN_UPPER, &h_),
        "lq_csr_create");
}

void SparseMatrixType::ConvertFromCOO(std::vector<int> &rows, std::vector<int> &cols, std::vector<cplx> &vals) {
  // COO -> CSR (rows sorted, stable within a row)
  const int n = numRows();
  std::vector<int> rp(n + 1, 0), order(rows.size());
  for (int r : rows) rp[r + 1]++;
  for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
  std::vector<int> pos(rp.begin(), rp.end() - 1), c(cols.size());
  std::vector<cplx> v(vals.size());
  for (size_t q = 0; q < rows.size(); ++q) {
    const int p = pos[rows[q]]++;
    c[p] = cols[q];
    v[p] = vals[q];
  }
  ConvertFromCSR(rp, c, v);
}

void SparseMatrixType::device_vectors(size_t bytes) {
  if (bytes <= dbytes_) return;
  lq_ctx ctx = default_context();
  if (dx_) lq_free(ctx, dx_);
  if (dy_) lq_free(ctx, dy_);
  check(lq_malloc(ctx, bytes, &dx_), "lq_malloc");
  check(lq_malloc(ctx, bytes, &dy_), "lq_malloc");
  dbytes_ = bytes;
}

// mkl_sparse_z_mv (src/mkl_sparse_matrix.cpp:56-69) on host arrays.
void SparseMatrixType::Multiply(const value_t a, const value_t *x, const value_t b, value_t *y) {
  assert(h_);
  lq_ctx ctx = default_context();
  const size_t bytes = sizeof(cplx) * (size_t)numRows();
  device_vectors(bytes);
  check(lq_memcpy_h2d(ctx, dx_, x, bytes), "lq_memcpy_h2d");
  if (b != value_t(0)) check(lq_memcpy_h2d(ctx, dy_, y, bytes), "lq_memcpy_h2d");
  check(lq_spmv(h_, a.real(), a.imag(), (const double *)dx_, b.real(), b.imag(), (double *)dy_, 1), "lq_spmv");
  check(lq_memcpy_d2h(ctx, y, dy_, bytes), "lq_memcpy_d2h");
}

void SparseMatrixType::Multiply(const value_t a, const vector_t &x, const value_t b, vector_t &y) {
  assert(x.size() == y.size());
  Multiply(a, x.data(), b, y.data());
}

// mkl_sparse_z_mm with SPARSE_LAYOUT_COLUMN_MAJOR (src/mkl_sparse_matrix.cpp:102-106).
void SparseMatrixType::BatchMultiply(const int batchSize, const value_t a, const value_t *x, const value_t b,
                                     value_t *y) {
  assert(h_ && batchSize > 0 && batchSize <= 64);
  const size_t n = numRows();
  std::vector<cplx> xi(n * batchSize), yi(n * batchSize);
  for (int v = 0; v < batchSize; ++v)
    for (size_t i = 0; i < n; ++i) {
      xi[i * batchSize + v] = x[v * n + i];
      yi[i * batchSize + v] = y[v * n + i];
    }
  lq_ctx ctx = default_context();
  const size_t bytes = sizeof(cplx) * n * batchSize;
  device_vectors(bytes);
  check(lq_memcpy_h2d(ctx, dx_, xi.data(), bytes), "lq_memcpy_h2d");
  check(lq_memcpy_h2d(ctx, dy_, yi.data(), bytes), "lq_memcpy_h2d");
  check(lq_spmv(h_, a.real(), a.imag(), (const double *)dx_, b.real(), b.imag(), (double *)dy_, batchSize), "lq_spmv");
  check(lq_memcpy_d2h(ctx, yi.data(), dy_, bytes), "lq_memcpy_d2h");
  for (int v = 0; v < batchSize; ++v)
    for (size_t i = 0; i < n; ++i) y[v * n + i] = yi[i * batchSize + v];
}

// src/mkl_sparse_matrix.cpp:73-97: the stored CSR arrays keep their original values.
void SparseMatrixType::Rescale(const value_t a, const value_t b) {
  assert(h_);
  check(lq_csr_rescale(h_, a.real(), a.imag(), b.real(), b.imag()), "lq_csr_rescale");
}

void SparseMatrixBuilder::BuildOPFromCSRFile(const std::string input) {
  std::vector<int> columns, rowIndex;
  std::vector<cplx> values;
  int dim = 0;
  Sparse::OPERATOR_FromCSRFile(input, dim, columns, rowIndex, values);
  _matrix_type->setDimensions(dim, dim);
  _matrix_type->ConvertFromCSR(rowIndex, columns, values);
  std::cout << "OPERATOR SUCCESSFULLY BUILT" << std::endl;
}

// ---- linalg (src/mkl_linear_algebra.cpp) ----------------------------------------
namespace linalg {
void scal(const int dim, cplx a, cplx *x) {
  for (int i = 0; i < dim; ++i) x[i] *= a;
}
void scal(const cplx &a, std::vector<cplx> &x) { scal((int)x.size(), a, x.data()); }
void axpy(const int dim, cplx a, const cplx *x, cplx *y) {
  for (int i = 0; i < dim; ++i) y[i] += a * x[i];
}
void axpy(const cplx &a, const std::vector<cplx> &x, std::vector<cplx> &y) {
  assert(x.size() == y.size());
  axpy((int)x.size(), a, x.data(), y.data());
}
void copy(const int dim, const cplx *x, cplx *y) { std::copy(x, x + dim, y); }
void copy(const std::vector<cplx> &x, std::vector<cplx> &y) {
  assert(x.size() == y.size());
  copy((int)x.size(), x.data(), y.data());
}
cplx vdot(const int dim, const cplx *x, cplx *y) {
  cplx s = 0;
  for (int i = 0; i < dim; ++i) s += std::conj(x[i]) * y[i];
  return s;
}
cplx vdot(const std::vector<cplx> &x, const std::vector<cplx> &y) {
  assert(x.size() == y.size());
  return vdot((int)x.size(), x.data(), const_cast<cplx *>(y.data()));
}
double nrm2(const int dim, const cplx *x) {
  double s = 0;
  for (int i = 0; i < dim; ++i) s += std::norm(x[i]);
  return std::sqrt(s);
}
double nrm2(const std::vector<cplx> &x) { return nrm2((int)x.size(), x.data()); }
// output(m0, m1) = <left_m0|right_m1>, row major (src/mkl_linear_algebra.cpp:68-78)
void batch_vdot(const int dim, const int batchSize, const cplx *leftb, const cplx *rightb, cplx *output) {
  for (int a = 0; a < batchSize; ++a)
    for (int b = 0; b < batchSize; ++b)
      output[a * batchSize + b] += vdot(dim, leftb + (size_t)a * dim, const_cast<cplx *>(rightb + (size_t)b * dim));
}
void extract_segment(const std::vector<cplx> &x, size_t, size_t start_x, std::vector<cplx> &y, size_t size_y) {
  for (size_t i = 0; i < size_y; i++) y[i] = x[start_x + i];
}
void introduce_segment(const std::vector<cplx> &x, size_t size_x, std::vector<cplx> &y, size_t, size_t start_y) {
  for (size_t i = 0; i < size_x; i++) y[start_y + i] = x[i];
}
}  // namespace linalg

// ---- qstates (include/quantum_states.hpp) --------------------------------------
namespace qstates {
bool generator::getQuantumState() {
  if (count >= NumberOfStates()) return false;
  switch (kind) {
    case LOCAL_STATE:
      std::fill(out.begin(), out.end(), 0.0);
      out[spos[count]] = 1.0;
      break;
    case USER_STATE:
      out.assign(data.begin() + spos[count], data.begin() + spos[count] + out.size());
      break;
    case RANDOM_STATE:
      std::fill(out.begin(), out.end(), 0.0);
      out[dim / 2] = 1.0;
      break;
  }
  count++;
  return true;
}
StateType String2State(std::string kind) {
  if (kind == "local") return LOCAL_STATE;
  if (kind == "vector") return USER_STATE;
  return RANDOM_STATE;
}
void readLocalState(std::ifstream &file, generator &data) {
  int n;
  file >> n;
  data.NumberOfStates(n);
  for (int i = 0; i < data.num_states; i++) {
    int s;
    file >> s;
    data.spos.push_back(s);
  }
}
void readCustomState(std::ifstream &file, generator &data) {
  int n;
  file >> n;
  data.NumberOfStates(n);
  file >> n;
  data.SystemSize(n);
  const int size = data.SystemSize(), num = data.NumberOfStates();
  data.data = std::vector<scalar>((size_t)size * num);
  for (int i = 0; i < num; i++) {
    for (int j = 0; j < size; j++) {
      double re, im;
      file >> re >> im;
      data.data[(size_t)i * size + j] = scalar(re, im);
    }
    data.spos.push_back(i * size);
  }
}
generator CreateStateSet(std::ifstream &file, std::string kind) {
  generator data;
  data.kind = String2State(kind);
  if (data.kind == LOCAL_STATE) readLocalState(file, data);
  if (data.kind == USER_STATE) readCustomState(file, data);
  return data;
}
generator LoadStateFile(std::string filename) {
  std::ifstream infile(filename);
  std::string kind("random_phase");
  if (infile.good())
    infile >> kind;
  else
    std::cout << "State file not found, using default" << std::endl;
  std::cout << "Building " << kind << " states" << std::endl;
  generator states = CreateStateSet(infile, kind);
  states.label = filename;
  return states;
}
}  // namespace qstates

namespace chebyshev {

// Draws every state of the generator (the loop of the reference's solvers).
static std::vector<cplx> collect_states(qstates::generator &gen, int dim, int *R) {
  gen.SystemSize(dim);
  std::vector<cplx> phi;
  int r = 0;
  while (gen.getQuantumState()) {
    phi.insert(phi.end(), gen.State().begin(), gen.State().end());
    ++r;
  }
  *R = r;
  return phi;
}

// ---- Moments (src/chebyshev_moments.cpp) -----------------------------------------
void Moments::SetInitVectors(const vector_t &T0) {
  const size_t dim = SystemSize();
  assert(T0.size() == dim);
  if (Chebyshev0().size() != dim) Chebyshev0() = vector_t(dim, 0.0);
  if (Chebyshev1().size() != dim) Chebyshev1() = vector_t(dim, 0.0);
  linalg::copy(T0, Chebyshev0());
  Hamiltonian().Multiply(Chebyshev0(), Chebyshev1());
}

void Moments::SetInitVectors(SparseMatrixType &OP, const vector_t &T0) {
  const size_t dim = SystemSize();
  assert((size_t)OP.rank() == dim && T0.size() == dim);
  if (Chebyshev0().size() != dim) Chebyshev0() = vector_t(dim, 0.0);
  if (Chebyshev1().size() != dim) Chebyshev1() = vector_t(dim, 0.0);
  linalg::copy(T0, Chebyshev1());
  OP.Multiply(Chebyshev1(), Chebyshev0());
  Hamiltonian().Multiply(Chebyshev0(), Chebyshev1());
}

int Moments::Iterate() {
  Hamiltonian().Multiply(2.0, Chebyshev1(), -1.0, Chebyshev0());
  Chebyshev0().swap(Chebyshev1());
  return 0;
}

int Moments::JacksonKernelMomCutOff(const double broad) {
  assert(broad > 0);
  const double eta = 2.0 * broad / 1000 / BandWidth();
  return (int)ceil(M_PI / eta);
}

double Moments::JacksonKernel(const double m, const double Mom) {
  const double phi_J = M_PI / (double)(Mom + 1.0);
  return ((Mom - m + 1) * cos(phi_J * m) + sin(phi_J * m) * cos(phi_J) / sin(phi_J)) * phi_J / M_PI;
}

// ---- Moments1D (src/chebyshev_moments1D.cpp) ---------------------------------------
Moments1D::Moments1D(std::string momfilename) {
  const size_t ext = momfilename.find(".chebmom1D");
  if (ext == std::string::npos) {
    std::cerr << "The first argument does not seem to be a valid .chebmom1D file" << std::endl;
    assert(false);
    std::abort();
  }
  SystemLabel(momfilename.substr(0, ext));
  std::vector<double> mu;
  const lq_mom_header h = linqt_amd::read_moments(momfilename, mu);
  SystemSize((int)h.n);
  BandWidth(h.W);
  BandCenter(h.Ec);
  _numMoms = (size_t)h.M0;
  MomentVector(vector_t(_numMoms, 0.0));
  for (size_t m = 0; m < _numMoms; m++) (*this)(m) = value_t(mu[2 * m], mu[2 * m + 1]);
}

void Moments1D::MomentNumber(const size_t numMoms) {
  vector_t nm(numMoms, 0.0);
  for (size_t m = 0; m < numMoms; m++) nm[m] = (*this)(m);
  _numMoms = numMoms;
  MomentVector(nm);
}

void Moments1D::ApplyJacksonKernel(const double broad) {
  int maxMom = 0;
  vector_t &mu = MomentVector();
  check(lq_jackson_1d(default_context(), reinterpret_cast<double *>(mu.data()), (int)_numMoms, BandWidth(), broad,
                      &maxMom),
        "lq_jackson_1d");
  std::cout << "Kernel reduced the number of moments to " << maxMom << " for a broadening of " << M_PI / maxMom
            << std::endl;
  MomentNumber(maxMom);
}

void Moments1D::saveIn(std::string filename) {
  std::ofstream out(filename.c_str());
  out.precision(std::numeric_limits<double>::digits10);
  out << SystemSize() << " " << BandWidth() << " " << BandCenter() << std::endl;
  out << HighestMomentNumber() << std::endl;
  for (auto m : MomentVector()) out << m.real() << " " << m.imag() << std::endl;
}

void Moments1D::Print() {
  std::cout << "\n\nCHEBYSHEV 1D MOMENTS INFO" << std::endl;
  std::cout << "\tSYSTEM:\t\t\t" << SystemLabel() << std::endl;
  if (SystemSize() > 0) std::cout << "\tSIZE:\t\t\t" << SystemSize() << std::endl;
  std::cout << "\tMOMENTS SIZE:\t\t" << HighestMomentNumber() << std::endl;
  std::cout << "\tSCALE FACTOR:\t\t" << ScaleFactor() << std::endl;
  std::cout << "\tSHIFT FACTOR:\t\t" << ShiftFactor() << std::endl;
  std::cout << "\tENERGY SPECTRUM:\t(" << -HalfWidth() + BandCenter() << " , " << HalfWidth() + BandCenter() << ")"
            << std::endl
            << std::endl;
}

// ---- Moments2D (src/chebyshev_moments2D.cpp) ---------------------------------------
Moments2D::Moments2D(std::string momfilename) {
  const size_t ext = momfilename.find(".chebmom2D");
  if (ext == std::string::npos) {
    std::cerr << "The first argument does not seem to be a valid .chebmom2D file" << std::endl;
    assert(false);
    std::abort();
  }
  SystemLabel(momfilename.substr(0, ext));
  std::vector<double> mu;
  const lq_mom_header h = linqt_amd::read_moments(momfilename, mu);
  SystemSize((int)h.n);
  BandWidth(h.W);
  BandCenter(h.Ec);
  numMoms[0] = h.M0;
  numMoms[1] = h.M1;
  MomentVector(vector_t((size_t)numMoms[0] * numMoms[1], 0.0));
  for (int m0 = 0; m0 < numMoms[0]; m0++)
    for (int m1 = 0; m1 < numMoms[1]; m1++) {
      const size_t q = (size_t)m0 * numMoms[1] + m1;
      (*this)(m0, m1) = value_t(mu[2 * q], mu[2 * q + 1]);
    }
}

void Moments2D::MomentNumber(const int mom0, const int mom1) {
  assert(mom0 <= numMoms[0] && mom1 <= numMoms[1]);
  vector_t nm((size_t)mom0 * mom1, 0.0);
  for (int m0 = 0; m0 < mom0; m0++)
    for (int m1 = 0; m1 < mom1; m1++) nm[(size_t)m0 * mom1 + m1] = (*this)(m0, m1);
  numMoms = {mom0, mom1};
  MomentVector(nm);
}

void Moments2D::ApplyJacksonKernel(const double b0, const double b1) {
  int o0 = 0, o1 = 0;
  vector_t &mu = MomentVector();
  check(lq_jackson_2d(default_context(), reinterpret_cast<double *>(mu.data()), numMoms[0], numMoms[1], BandWidth(),
                      b0, b1, &o0, &o1),
        "lq_jackson_2d");
  std::cout << "Kernel reduced the number of moments to " << o0 << " " << o1 << std::endl;
  mu.resize((size_t)o0 * o1);
  numMoms = {o0, o1};
}

void Moments2D::saveIn(std::string filename) {
  std::ofstream out(filename.c_str());
  out.precision(std::numeric_limits<double>::digits10);
  out << SystemSize() << " " << BandWidth() << "  " << BandCenter() << std::endl;
  for (auto x : numMoms) out << x << " ";
  out << std::endl;
  for (auto m : MomentVector()) out << m.real() << " " << m.imag() << std::endl;
}

void Moments2D::Print() {
  std::cout << "\n\nCHEBYSHEV 2D MOMENTS INFO" << std::endl;
  std::cout << "\tSYSTEM:\t\t\t" << SystemLabel() << std::endl;
  if (SystemSize() > 0) std::cout << "\tSIZE:\t\t\t" << SystemSize() << std::endl;
  std::cout << "\tMOMENTS SIZE:\t\t(" << HighestMomentNumber(0) << " x " << HighestMomentNumber(1) << ")" << std::endl;
  std::cout << "\tSCALE FACTOR:\t\t" << ScaleFactor() << std::endl;
  std::cout << "\tSHIFT FACTOR:\t\t" << ShiftFactor() << std::endl;
  std::cout << "\tENERGY SPECTRUM:\t(" << -HalfWidth() + BandCenter() << " , " << HalfWidth() + BandCenter() << ")"
            << std::endl
            << std::endl;
}

// ---- MomentsTD (src/chebyshev_momentsTD.cpp) --------------------------------------------
MomentsTD::MomentsTD(std::string momfilename) : _numMoms(1), _maxTimeStep(1), _timeStep(0), _dt(0) {
  const std::size_t ext = momfilename.find(".chebmomTD");
  if (ext == std::string::npos) {
    std::cerr << "The first argument does not seem to be a valid .chebmomTD file" << std::endl;
    assert(false);
    std::abort();
  }
  SystemLabel(momfilename.substr(0, ext));
  std::vector<double> mu;
  const lq_mom_header h = linqt_amd::read_moments(momfilename, mu);
  SystemSize((int)h.n);
  BandWidth(h.W);
  BandCenter(h.Ec);
  MaxTimeStep(h.M1);
  TimeDiff(h.dt);
  _numMoms = (size_t)h.M0;
  _maxTimeStep = (size_t)h.M1;
  MomentVector(vector_t(_numMoms * _maxTimeStep, 0.0));
  for (size_t m = 0; m < _numMoms; m++)
    for (size_t n = 0; n < _maxTimeStep; n++) {
      const size_t q = m * _maxTimeStep + n;
      (*this)(m, n) = value_t(mu[2 * q], mu[2 * q + 1]);
    }
}

void MomentsTD::MomentNumber(const size_t numMoms) {
  assert(numMoms <= HighestMomentNumber());
  vector_t nm(numMoms * _maxTimeStep, 0.0);
  for (size_t m = 0; m < numMoms; m++)
    for (size_t n = 0; n < _maxTimeStep; n++) nm[m * _maxTimeStep + n] = (*this)(m, n);
  _numMoms = numMoms;
  MomentVector(nm);
}

void MomentsTD::ApplyJacksonKernel(const double broad) {
  int mo = 0;
  vector_t &mu = MomentVector();
  check(lq_jackson_td(default_context(), reinterpret_cast<double *>(mu.data()), (int)_numMoms, (int)_maxTimeStep,
                      BandWidth(), broad, &mo),
        "lq_jackson_td");
  std::cout << "Kernel reduced the number of moments to " << mo << " for a broadening of " << M_PI / mo << std::endl;
  mu.resize((size_t)mo * _maxTimeStep);
  _numMoms = (size_t)mo;
}

void MomentsTD::saveIn(std::string filename) {
  std::ofstream out(filename.c_str());
  out.precision(std::numeric_limits<double>::digits10);
  out << SystemSize() << " " << BandWidth() << " " << BandCenter() << " " << MaxTimeStep() << " " << TimeDiff() << " "
      << std::endl;
  out << HighestMomentNumber() << " " << MaxTimeStep() << " " << std::endl;
  for (auto m : MomentVector()) out << m.real() << " " << m.imag() << std::endl;
}

void MomentsTD::Print() {
  std::cout << "\n\nCHEBYSHEV TD MOMENTS INFO" << std::endl;
  std::cout << "\tSYSTEM:\t\t\t" << SystemLabel() << std::endl;
  if (SystemSize() > 0) std::cout << "\tSIZE:\t\t\t" << SystemSize() << std::endl;
  std::cout << "\tMOMENTS SIZE:\t\t(" << HighestMomentNumber() << " x " << MaxTimeStep() << ")" << std::endl;
  std::cout << "\tSCALE FACTOR:\t\t" << ScaleFactor() << std::endl;
  std::cout << "\tSHIFT FACTOR:\t\t" << ShiftFactor() << std::endl;
  std::cout << "\tTIME STEP:\t\t" << MaxTimeStep() << std::endl;
  std::cout << "\tTIME DIFF:\t" << TimeDiff() << std::endl;
}

// ---- Vectors_sliced (src/chebyshev_vectors.cpp) --------------------------------------
int Vectors_sliced::CreateVectorSet() {
  SystemSize(Hamiltonian().rank());
  section_size_ = SystemSize() / num_sections_;
  last_section_size_ = SystemSize() / num_sections_ + SystemSize() % num_sections_;
  Chebmu_.Resize(NumberOfVectors(), last_section_size_);
  return 0;
}

int Vectors_sliced::IterateAllSliced(int s) {
  const size_t seg = (s == num_sections_ - 1) ? last_section_size_ : section_size_;
  const size_t start = (size_t)s * section_size_, DIM = SystemSize();
  linalg::extract_segment(Chebyshev0(), DIM, start, Vector(0), seg);
  for (int m = 1; m < NumberOfVectors(); m++) {
    linalg::extract_segment(Chebyshev1(), DIM, start, Vector(m), seg);
    Hamiltonian().Multiply(2.0, Chebyshev1(), -1.0, Chebyshev0());
    Chebyshev0().swap(Chebyshev1());
  }
  return 0;
}

int Vectors_sliced::MultiplySliced(SparseMatrixType &OP, int s) {
  const size_t seg = (s == num_sections_ - 1) ? last_section_size_ : section_size_;
  const size_t start = (size_t)s * section_size_, DIM = SystemSize();
  vector_t tmp2(DIM, 0.0);
  OPV_.assign(OP.rank(), 0.0);
  for (int m = 0; m < NumberOfVectors(); m++) {
    linalg::introduce_segment(Chebmu_.ListElem(m), seg, OPV_, DIM, start);
    OP.Multiply(OPV_, tmp2);
    linalg::extract_segment(tmp2, DIM, start, Chebmu_.ListElem(m), seg);
  }
  return 0;
}

// ---- kpm_noneqop.hpp:27-39 ---------------------------------------------------------
bool GetBatchSize(int &batchSize) {
  batchSize = 3;
  const char *e = getenv("BATCH_SIZE");
  if (!e || atoi(e) <= 0) {
    std::cout << "\nEnviroment variable BATCH_SIZE not set or invalid.\nSet to your custom N value throught the "
                 "command export BATCH_SIZE N. Using sequential"
              << std::endl;
    return false;
  }
  batchSize = atoi(e);
  return true;
}

// ---- chebyshev_solver.cpp ------------------------------------------------------------
std::array<double, 2> utility::SpectralBounds(SparseMatrixType &) {
  double highest = 0, lowest = 0;
  std::ifstream f("BOUNDS");
  if (f.is_open()) {
    f >> lowest >> highest;
  } else {
    lowest = -100;
    highest = 100;
  }
  return {lowest, highest};
}

// ComputeMomTable (src/chebyshev_solver.cpp:26-42) on the MFMA contraction.
int ComputeMomTable(Vectors &L, Vectors &R, Moments2D &sub) {
  const int ML = L.NumberOfVectors(), MR = R.NumberOfVectors();
  const long long K = (long long)L.Vector(0).size();
  lq_ctx ctx = default_context();
  void *dl = nullptr, *dr = nullptr, *ds = nullptr;
  check(lq_malloc(ctx, sizeof(cplx) * ML * K, &dl), "lq_malloc");
  check(lq_malloc(ctx, sizeof(cplx) * MR * K, &dr), "lq_malloc");
  check(lq_malloc(ctx, sizeof(cplx) * ML * MR, &ds), "lq_malloc");
  for (int m = 0; m < ML; ++m)
    check(lq_memcpy_h2d(ctx, (cplx *)dl + (size_t)m * K, L.Vector(m).data(), sizeof(cplx) * K), "lq_memcpy_h2d");
  for (int m = 0; m < MR; ++m)
    check(lq_memcpy_h2d(ctx, (cplx *)dr + (size_t)m * K, R.Vector(m).data(), sizeof(cplx) * K), "lq_memcpy_h2d");
  check(lq_compute_mom_table(ctx, (const double *)dl, K, ML, (const double *)dr, K, MR, K, (double *)ds, MR),
        "lq_compute_mom_table");
  std::vector<cplx> h((size_t)ML * MR);
  check(lq_memcpy_d2h(ctx, h.data(), ds, sizeof(cplx) * h.size()), "lq_memcpy_d2h");
  for (int m0 = 0; m0 < ML && m0 < sub.HighestMomentNumber(1); ++m0)
    for (int m1 = 0; m1 < MR && m1 < sub.HighestMomentNumber(0); ++m1) sub(m0, m1) = h[(size_t)m0 * MR + m1];
  lq_free(ctx, dl);
  lq_free(ctx, dr);
  lq_free(ctx, ds);
  return 0;
}

// Inner overload (src/chebyshev_solver.cpp:45-82): host recursion with device
// products; accumulates the raw (unscaled) table.
int CorrelationExpansionMoments(const vector_t &PhiR, const vector_t &PhiL, SparseMatrixType &OPL,
                                SparseMatrixType &OPR, Vectors &chevVecL, Vectors &chevVecR, Moments2D &chebMoms) {
  const size_t numRVecs = chevVecR.NumberOfVectors(), numLVecs = chevVecL.NumberOfVectors();
  const size_t NumMoms = chebMoms.HighestMomentNumber();
  Moments2D sub(numLVecs, numRVecs);
  chevVecL.SetInitVectors(OPL, PhiL);
  for (size_t mL = 0; mL < NumMoms; mL += numLVecs) {
    chevVecL.IterateAll();
    chevVecR.SetInitVectors(PhiR);
    for (size_t mR = 0; mR < NumMoms; mR += numRVecs) {
      chevVecR.IterateAll();
      chevVecR.Multiply(OPR);
      ComputeMomTable(chevVecL, chevVecR, sub);
      chebMoms.AddSubMatrix(sub, (int)mL, (int)mR);
    }
  }
  return 0;
}

// Generator overload (src/chebyshev_solver.cpp:86-172): one fused device solve.
int CorrelationExpansionMoments(SparseMatrixType &OPL, SparseMatrixType &OPR, Moments2D &chebMoms,
                                qstates::generator &gen) {
  int batchSize;
  if (!GetBatchSize(batchSize)) {
    batchSize = chebMoms.HighestMomentNumber(0);
    std::cout << "Using default BATCH_SIZE = " << batchSize << std::endl;
  } else {
    std::cout << "Using user's BATCH_SIZE = " << batchSize << std::endl;
  }
  const int M0 = chebMoms.HighestMomentNumber(0), M1 = chebMoms.HighestMomentNumber(1);
  int R = 0;
  std::vector<cplx> phi = collect_states(gen, (int)chebMoms.SystemSize(), &R);
  lq_csr opl = OPL.isIdentity() ? nullptr : OPL.handle();
  lq_csr opr = OPR.isIdentity() ? nullptr : OPR.handle();
  check(lq_correlation_moments(chebMoms.Hamiltonian().handle(), opl, 1, &opr,
                               reinterpret_cast<const double *>(phi.data()), R, M0, M1, batchSize, LQ_PREC_FP64,
                               reinterpret_cast<double *>(chebMoms.MomentVector().data())),
        "lq_correlation_moments");
  return 0;
}

// TimeDependentCorrelations (src/chebyshev_solver.cpp:231-275): one device solve over
// all states and time steps.
int TimeDependentCorrelations(SparseMatrixType &OPL, SparseMatrixType &OPR, MomentsTD &chebMoms,
                              qstates::generator &gen) {
  int R = 0;
  std::vector<cplx> phi = collect_states(gen, (int)chebMoms.SystemSize(), &R);
  if (R == 0) return 0;
  check(lq_time_correlations(chebMoms.Hamiltonian().handle(), OPL.isIdentity() ? nullptr : OPL.handle(),
                             OPR.isIdentity() ? nullptr : OPR.handle(), reinterpret_cast<const double *>(phi.data()),
                             R, (int)chebMoms.HighestMomentNumber(), (int)chebMoms.MaxTimeStep(), chebMoms.TimeDiff(),
                             chebMoms.BandWidth(), chebMoms.BandCenter(),
                             reinterpret_cast<double *>(chebMoms.MomentVector().data())),
        "lq_time_correlations");
  return 0;
}

// ---- Moments1D_nonOrth (src/chebyshev_moments.cpp:63-153): host-vector semantics ----
void Moments1D_nonOrth::SetInitVectors_nonOrthogonal(vector_t &T0) {
  const size_t dim = SystemSize();
  assert(T0.size() == dim && S_);
  Chebyshev0().assign(dim, value_t(0));
  Chebyshev1().assign(dim, value_t(0));
  vector_t tmp(dim, 0.0);
  linalg::orthogonalize(*S_, Chebyshev0(), T0);
  Hamiltonian().Multiply(Chebyshev0(), tmp);
  linalg::orthogonalize(*S_, Chebyshev1(), tmp);
}

void Moments1D_nonOrth::SetInitVectors_nonOrthogonal(SparseMatrixType &OP, vector_t &T0) {
  const size_t dim = SystemSize();
  assert(T0.size() == dim && S_);
  Chebyshev0().assign(dim, value_t(0));
  Chebyshev1().assign(dim, value_t(0));
  vector_t tmp(dim, 0.0);
  // T0 = OP S^-1 phi, T1 = S^-1 H T0 (the reference first forms OP phi and
  // H OP phi, which the later assignments overwrite)
  linalg::orthogonalize(*S_, tmp, T0);
  Chebyshev1() = tmp;
  OP.Multiply(Chebyshev1(), Chebyshev0());
  Hamiltonian().Multiply(Chebyshev0(), tmp);
  linalg::orthogonalize(*S_, Chebyshev1(), tmp);
}

int Moments1D_nonOrth::Iterate_nonOrthogonal() {
  const size_t dim = Chebyshev0().size();
  vector_t tmp(dim, 0.0), tmp2(dim, 0.0);
  Hamiltonian().Multiply(Chebyshev1(), tmp);
  linalg::orthogonalize(*S_, tmp2, tmp);
  for (size_t i = 0; i < dim; i++) Chebyshev0()[i] = 2.0 * tmp2[i] - Chebyshev0()[i];
  Chebyshev0().swap(Chebyshev1());
  return 0;
}

// SpectralMoments_nonOrth (src/chebyshev_solver.cpp:203-229): one device solve.
int SpectralMoments_nonOrth(SparseMatrixType &OP, Moments1D_nonOrth &chebMoms, qstates::generator &gen) {
  int R = 0;
  std::vector<cplx> phi = collect_states(gen, (int)chebMoms.SystemSize(), &R);
  if (R == 0) return 0;
  assert(chebMoms.S());
  long long its = 0;
  check(lq_spectral_moments_nonorth(chebMoms.Hamiltonian().handle(), OP.isIdentity() ? nullptr : OP.handle(),
                                    chebMoms.S()->general_handle(), reinterpret_cast<const double *>(phi.data()), R,
                                    (int)chebMoms.HighestMomentNumber(), 1e-4, 2000,
                                    reinterpret_cast<double *>(chebMoms.MomentVector().data()), &its),
        "lq_spectral_moments_nonorth");
  std::cout << "BiCGSTAB overlap solves: " << its << " iterations in total (tolerance 1e-4)" << std::endl;
  return 0;
}

// SpectralMoments (src/chebyshev_solver.cpp:176-200): one fused device solve.
int SpectralMoments(SparseMatrixType &OP, Moments1D &chebMoms, qstates::generator &gen) {
  int R = 0;
  std::vector<cplx> phi = collect_states(gen, (int)chebMoms.SystemSize(), &R);
  if (R == 0) return 0;
  check(lq_spectral_moments(chebMoms.Hamiltonian().handle(), OP.isIdentity() ? nullptr : OP.handle(),
                            reinterpret_cast<const double *>(phi.data()), R, (int)chebMoms.HighestMomentNumber(),
                            LQ_ALGO_DIRECT, reinterpret_cast<double *>(chebMoms.MomentVector().data())),
        "lq_spectral_moments");
  return 0;
}

// ---- Kubo_solver_FFT (src/Kubo_solver_FFT.cpp, Kubo_solver_FFT_postProcess.cpp) ------
void Kubo_solver_FFT::compute(SparseMatrixType &OPL, SparseMatrixType &OPR, qstates::generator &gen) {
  const int DIM = (int)chebVec_.SystemSize();
  int R = 0;
  std::vector<cplx> phi = collect_states(gen, DIM, &R);
  if (R == 0) return;
  final_data_.assign(2 * (size_t)nump_, cplx(0));
  lq_csr opl = OPL.isIdentity() ? nullptr : OPL.handle();
  lq_csr opr = OPR.isIdentity() ? nullptr : OPR.handle();
  check(lq_kubo_fft(chebVec_.Hamiltonian().handle(), opl, 1, &opr, reinterpret_cast<const double *>(phi.data()), R,
                    M_, num_sections_, nump_, (int)sym_formula_, LQ_PREC_FP64,
                    reinterpret_cast<double *>(final_data_.data())),
        "lq_kubo_fft");
  std::vector<double> E(nump_), sigma(nump_), integ(nump_);
  check(lq_kubo_post(default_context(), (int)sym_formula_, reinterpret_cast<const double *>(final_data_.data()), nump_,
                     DIM, chebVec_.BandWidth(), chebVec_.BandCenter(), E.data(), sigma.data(), integ.data()),
        "lq_kubo_post");
  std::ofstream dataP(outputfilename_);
  for (int e = 0; e < nump_; e++) dataP << E[e] << "  " << sigma[e] << std::endl;
  if (sym_formula_ == KUBO_BASTIN) {
    std::ofstream data2(outputfilename_ + "_integrand");
    for (int e = 0; e < nump_; e++) data2 << E[e] << "  " << integ[e] << std::endl;
  }
}

}  // namespace chebyshev
