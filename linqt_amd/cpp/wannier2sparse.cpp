// wannier2sparse.cpp -- include/linqt/wannier2sparse.hpp over the C ABI.
//
// Host code (g++).  Unit-cell work follows utilities/wannier2sparse/src
// (hopping_list.cpp, tbmodel.cpp, wannier_parser.cpp); the supercell wrap and
// the CSR run on the device through lq_csr_from_hoppings.  Values are formed
// with the same std::complex<double> operations as the reference, so the
// operators it writes are reproduced bitwise before the 6-digit text output.
#include "../../include/linqt/wannier2sparse.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace w2s {

using linqt_amd::check;
typedef std::complex<double> cplx;

namespace {
[[noreturn]] void fail(const std::string &msg) {
  std::cerr << "wannier2sparse: " << msg << std::endl;
  std::exit(1);
}

std::string tag_of(const hopping_list::cellID_t &c, const hopping_list::edge_t &e) {
  char buf[96];
  snprintf(buf, sizeof buf, "%d %d %d %d %d ", c[0], c[1], c[2], e[0], e[1]);
  return buf;
}
}  // namespace

// ---- hopping_list -----------------------------------------------------------
void hopping_list::AddHopping(const cellID_t &c, const value_t &v, const edge_t &e) {
  cell.push_back(c);
  edge.push_back(e);
  val.push_back(v);
}

std::vector<int> hopping_list::ordered() const {
  std::vector<std::string> tags(val.size());
  for (size_t h = 0; h < val.size(); ++h) tags[h] = tag_of(cell[h], edge[h]);
  std::vector<int> idx(val.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return tags[a] < tags[b]; });
  return idx;
}

hopping_list &hopping_list::add_from_wannier(const std::tuple<int, std::vector<std::string>> &data) {
  SetWannierBasisSize(std::get<0>(data));
  const auto &lines = std::get<1>(data);
  std::vector<cellID_t> cs;
  std::vector<edge_t> es;
  std::vector<value_t> vs;
  double avg = 0.0;
  size_t counted = 0;
  for (const auto &line : lines) {
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;  // blank line
    std::istringstream ss(line);
    cellID_t c;
    edge_t e;
    double re, im;
    if (!(ss >> c[0] >> c[1] >> c[2] >> e[0] >> e[1] >> re >> im)) fail("unreadable hopping line: " + line);
    e[0] -= 1;  // the file's orbitals are one based
    e[1] -= 1;
    if (e[0] < 0 || e[1] < 0 || e[0] >= num_wann || e[1] >= num_wann) fail("orbital out of range in: " + line);
    const value_t hop(re, im);
    avg += std::sqrt(std::norm(hop));
    ++counted;
    cs.push_back(c);
    es.push_back(e);
    vs.push_back(hop);
  }
  avg /= (double)counted;
  if (!(avg > 0)) fail("the hopping list is empty or zero");
  for (size_t h = 0; h < vs.size(); ++h)
    if (std::sqrt(std::norm(vs[h])) / avg > 1e-10) AddHopping(cs[h], vs[h], es[h]);
  return *this;
}

hopping_list hopping_list::wrap_in_supercell(const cellID_t &cellDim) const {
  if (cellSizes != cellID_t{1, 1, 1}) fail("the list is already wrapped in a supercell");
  for (int d : cellDim)
    if (d <= 0) fail("supercell dimensions must be positive");
  hopping_list sc = *this;
  sc.cellSizes = cellDim;
  return sc;
}

lq_csr hopping_list::BuildOperator(lq_ctx ctx) const {
  const int nh = (int)val.size();
  std::vector<int> c(3 * (size_t)std::max(nh, 1)), e(2 * (size_t)std::max(nh, 1));
  std::vector<double> v(2 * (size_t)std::max(nh, 1));
  for (int h = 0; h < nh; ++h) {
    for (int d = 0; d < 3; ++d) c[3 * h + d] = cell[h][d];
    e[2 * h] = edge[h][0];
    e[2 * h + 1] = edge[h][1];
    v[2 * h] = val[h].real();
    v[2 * h + 1] = val[h].imag();
  }
  lq_csr A = nullptr;
  check(lq_csr_from_hoppings(ctx, num_wann, nh, c.data(), e.data(), v.data(), cellSizes.data(), &A),
        "lq_csr_from_hoppings");
  return A;
}

void hopping_list::save_hopping_list_as_csr(const std::string &output_filename) const {
  lq_csr A = BuildOperator(linqt_amd::default_context());
  int n = 0;
  long long nnz = 0, stored = 0;
  check(lq_csr_info(A, &n, &nnz, &stored), "lq_csr_info");
  check(lq_csr_export(A, &nnz, nullptr, nullptr, nullptr), "lq_csr_export");
  std::vector<int> rowptr((size_t)n + 1), col((size_t)std::max(nnz, 1LL));
  std::vector<double> v(2 * (size_t)std::max(nnz, 1LL));
  check(lq_csr_export(A, &nnz, rowptr.data(), col.data(), v.data()), "lq_csr_export");
  lq_csr_destroy(A);
  check(lq_csr_file_write(output_filename.c_str(), LQ_FMT_TEXT_REF, n, nnz, rowptr.data(), col.data(), v.data()),
        "lq_csr_file_write");
}

// ---- file readers (src/wannier_parser.cpp) ------------------------------------
std::tuple<int, std::vector<std::string>> read_wannier_file(const std::string &fname) {
  std::ifstream f(fname.c_str());
  if (!f.is_open()) fail("cannot open " + fname);
  int num_wann = 0, nrpts = 0;
  std::vector<std::string> hops;
  std::string line;
  for (int counter = 0; std::getline(f, line); counter++) {
    if (counter == 0) continue;  // date / comment
    if (counter == 1) {
      num_wann = std::stoi(line);
      continue;
    }
    if (counter == 2) {
      nrpts = std::stoi(line);
      continue;
    }
    if (counter == 3) {  // Wigner-Seitz degeneracies: this line and nrpts/15 more
      for (int k = 0; k < nrpts / 15 && std::getline(f, line); ++k) counter++;
      continue;
    }
    hops.push_back(line);
  }
  if (num_wann <= 0 || nrpts <= 0) fail("bad header in " + fname);
  return {num_wann, hops};
}

std::vector<std::tuple<std::string, std::array<double, 3>>> read_xyz_file(const std::string &fname) {
  std::ifstream f(fname.c_str());
  if (!f.is_open()) fail("cannot open " + fname);
  int num_sites = 0;
  f >> num_sites;
  if (num_sites <= 0) fail("bad site count in " + fname);
  std::vector<std::tuple<std::string, std::array<double, 3>>> out;
  for (int i = 0; i < num_sites; ++i) {
    std::string label;
    std::array<double, 3> pos;
    if (!(f >> label >> pos[0] >> pos[1] >> pos[2])) fail("truncated " + fname);
    out.emplace_back(label, pos);
  }
  return out;
}

std::array<std::array<double, 3>, 3> read_unit_cell_file(const std::string &fname) {
  std::ifstream f(fname.c_str());
  if (!f.is_open()) fail("cannot open " + fname);
  std::array<std::array<double, 3>, 3> uc;
  for (auto &v : uc)
    for (auto &x : v)
      if (!(f >> x)) fail("truncated " + fname);
  return uc;
}

// ---- operator_utils ---------------------------------------------------------
namespace oputil {
const char *const AVAIL_OPS[18] = {"VX", "VY", "VZ", "VXSX", "VXSY", "VXSZ", "VYSX", "VYSY", "VYSZ",
                                   "VZSX", "VZSY", "VZSZ", "SX", "SY", "SZ", "TX", "TY", "TZ"};
static bool in(const std::string &op, int i0, int i1) {
  for (int i = i0; i < i1; ++i)
    if (op == AVAIL_OPS[i]) return true;
  return false;
}
bool is_velocity(const std::string &op) { return in(op, 0, 3); }
bool is_spinvelocity(const std::string &op) { return in(op, 3, 12); }
bool is_spin(const std::string &op) { return in(op, 12, 15); }
bool is_torque(const std::string &op) { return in(op, 15, 18); }
}  // namespace oputil

std::array<double, 3> tag2cartesian(const std::array<int, 3> &tag, const std::array<std::array<double, 3>, 3> &lat) {
  std::array<double, 3> v = {0, 0, 0};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) v[i] += tag[j] * lat[j][i];
  return v;
}

double volume(const std::array<std::array<double, 3>, 3> &uc) {
  // triple product a2 . (a0 x a1), as the reference groups it
  return std::fabs((uc[0][1] * uc[1][2] - uc[1][1] * uc[0][2]) * uc[2][0] +
                   (uc[0][2] * uc[1][0] - uc[0][0] * uc[1][2]) * uc[2][1] +
                   (uc[0][0] * uc[1][1] - uc[0][1] * uc[1][0]) * uc[2][2]);
}

// ---- tbmodel (src/tbmodel.cpp) ----------------------------------------------
void tbmodel::readUnitCell(const std::string &f) { lat_vecs = read_unit_cell_file(f); }
void tbmodel::readOrbitalPositions(const std::string &f) { orbPos_list = read_xyz_file(f); }
void tbmodel::readWannierModel(const std::string &f) { hl.add_from_wannier(read_wannier_file(f)); }

void tbmodel::readStaticDisorder(const std::string &fname) {
  std::ifstream f(fname.c_str());
  if (!f.is_open()) {
    std::cout << "Pristine calculation" << std::endl;
    return;
  }
  std::string line;
  while (std::getline(f, line))
    if (!line.empty())
      fail("static disorder files (" + fname +
           ") are not supported: the reference draws their values from std::random_device");
  std::cout << "Pristine calculation" << std::endl;
}

hopping_list tbmodel::createLocalHopping_list() {
  if (!(volume(lat_vecs) > 0) || (int)orbPos_list.size() != hl.WannierBasisSize())
    fail("unit cell or orbital positions do not match the model");
  hopping_list chl;
  chl.SetWannierBasisSize(hl.WannierBasisSize());
  for (int h : hl.ordered())
    if (hl.cell[h][0] == 0 && hl.cell[h][1] == 0 && hl.cell[h][2] == 0) chl.AddHopping(hl.cell[h], hl.val[h], hl.edge[h]);
  return chl;
}

hopping_list tbmodel::createHoppingDensity_list(oputil::op_matrix OP) {
  if (!(volume(lat_vecs) > 0) || (int)orbPos_list.size() != hl.WannierBasisSize() ||
      OP.size() != hl.WannierBasisSize())
    fail("unit cell, orbital positions or operator size do not match the model");
  hopping_list chl;
  chl.SetWannierBasisSize(hl.WannierBasisSize());
  const int n = (int)orbPos_list.size();
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) chl.AddHopping({0, 0, 0}, OP(i, j), {i, j});
  return chl;
}

hopping_list tbmodel::createHoppingCurrents_list(int dir) {
  if (dir < 0 || dir > 2) fail("velocity direction must be X, Y or Z");
  if (!(volume(lat_vecs) > 0)) fail("the unit cell has no volume");
  if ((int)orbPos_list.size() != hl.WannierBasisSize()) fail("orbital positions do not match the model");
  const double av_length = std::pow(volume(lat_vecs), 1.0 / 3.0);
  hopping_list curr = hl;
  for (int h : curr.ordered()) {
    const auto &p1 = std::get<1>(orbPos_list[curr.edge[h][1]]);
    const auto &p0 = std::get<1>(orbPos_list[curr.edge[h][0]]);
    std::array<double, 3> d = p1;
    for (int i = 0; i < 3; ++i) d[i] += -p0[i];
    const auto disp = tag2cartesian(curr.cell[h], lat_vecs);
    for (int i = 0; i < 3; ++i) d[i] += disp[i];
    if (std::fabs(d[dir]) / av_length > 1e-10)
      curr.val[h] *= cplx(0.0, -d[dir]);
    else
      curr.val[h] *= cplx(0.0, 0.0);
  }
  return curr;
}

oputil::op_matrix tbmodel::createSpinorialOp(const std::array<cplx, 4> &op) {
  const int tot = (int)orbPos_list.size(), orb = tot / 2;
  oputil::op_matrix S(tot);
  for (int io = 0; io < orb; io++)
    for (int si = 0; si < 2; si++)
      for (int sj = 0; sj < 2; sj++) S(si * orb + io, sj * orb + io) = op[si * 2 + sj];
  return S;
}

hopping_list tbmodel::createHoppingSpinorialDensity_list(const std::array<cplx, 4> &op) {
  return createHoppingDensity_list(createSpinorialOp(op));
}

hopping_list tbmodel::createHoppingSpinDensity_list(double theta, double phi) {
  const cplx I(0, 1);
  std::array<cplx, 4> op = {cos(theta), -sin(theta) * exp(-I * phi), sin(theta) * exp(I * phi), -cos(theta)};
  return createHoppingDensity_list(createSpinorialOp(op));
}

oputil::op_matrix tbmodel::createTorqueMatrix(double theta, double phi) {
  const cplx I(0, 1);
  const int tot = (int)orbPos_list.size(), orb = tot / 2;
  oputil::op_matrix H00(tot);
  hopping_list on = createLocalHopping_list();
  for (int h : on.ordered()) H00(on.edge[h][0], on.edge[h][1]) += on.val[h];
  oputil::op_matrix Tu(tot);
  double avg_Jex = 0.0;
  for (int io = 0; io < orb; io++) {
    const int i = io, j = orb + io;
    const cplx J00 = 0.5 * (H00(i, i) - H00(j, j));
    const cplx J01 = 0.5 * (H00(i, j) + std::conj(H00(j, i)));
    const double J[3] = {J01.real(), -J01.imag(), J00.real()};
    const double Jex = std::sqrt(J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
    avg_Jex += Jex;
    Tu(i, i) = 1.0 * sin(theta) * (J[1] * cos(phi) - J[0] * sin(phi));
    Tu(j, j) = -1.0 * sin(theta) * (J[1] * cos(phi) - J[0] * sin(phi));
    Tu(i, j) = (-I * J[0] - J[1]) * cos(theta) + J[2] * sin(theta) * (I * cos(phi) + sin(phi));
    Tu(j, i) = std::conj(Tu(i, j));
  }
  std::cout << "Average exchange Jex = " << avg_Jex / orb << " " << std::endl;
  return Tu;
}

hopping_list tbmodel::createHoppingTorqueDensity_list(double theta, double phi) {
  return createHoppingDensity_list(createTorqueMatrix(theta, phi));
}

hopping_list tbmodel::createHoppingSpinCurrents_list(int dir, double theta, double phi) {
  const cplx I(0.0, 1.0);
  hopping_list scurr;
  scurr.SetWannierBasisSize(hl.WannierBasisSize());
  const hopping_list curr = createHoppingCurrents_list(dir);
  const int UP = 0, DW = 1;
  for (int h : curr.ordered()) {
    const auto &c = curr.cell[h];
    const auto &e = curr.edge[h];
    const cplx v = curr.val[h];
    const int is0 = getSpinIndex(e[0]), is1 = getSpinIndex(e[1]);
    const double sz = 1.0 - 2.0 * is0;
    if (is0 == is1) {  // up-up or down-down
      const double sval_uu = sz * cos(theta);
      const cplx sval_ud = 0.5 * sin(theta) * exp(-I * phi);
      if (std::norm(sval_uu) > 1e-5) {
        cplx vu = v;
        vu *= cplx(sval_uu);
        scurr.AddHopping(c, vu, ChangeSpinIndex(e, {is0, is0}));
      }
      if (std::norm(sval_ud) > 1e-5) {
        cplx vud = v, vdu = v;
        vud *= sval_ud;
        vdu *= std::conj(sval_ud);
        scurr.AddHopping(c, vud, ChangeSpinIndex(e, {UP, DW}));
        scurr.AddHopping(c, vdu, ChangeSpinIndex(e, {DW, UP}));
      }
    } else {  // up-down (sz = 1) or down-up (sz = -1)
      const cplx sval_uu = 0.5 * sin(theta) * exp(sz * I * phi);
      if (std::norm(sval_uu) > 1e-5) {
        cplx vu = v;
        vu *= sval_uu;
        scurr.AddHopping(c, vu, ChangeSpinIndex(e, {UP, UP}));
        scurr.AddHopping(c, vu, ChangeSpinIndex(e, {DW, DW}));
      }
    }
  }
  return scurr;
}

static double spin_theta(char d) { return d == 'Z' ? 0.0 * M_PI : 0.5 * M_PI; }
static double spin_phi(char d) { return d == 'Y' ? 0.5 * M_PI : 0.0 * M_PI; }

hopping_list tbmodel::createHoppingSpinDensity_list(const std::string &op) {
  const char d = op.back();
  if (d != 'X' && d != 'Y' && d != 'Z') fail("incorrect spin direction in " + op);
  return createHoppingSpinDensity_list(spin_theta(d), spin_phi(d));
}

hopping_list tbmodel::createHoppingTorqueDensity_list(const std::string &op) {
  const char d = op.back();
  if (d != 'X' && d != 'Y' && d != 'Z') fail("incorrect spin direction in " + op);
  return createHoppingTorqueDensity_list(spin_theta(d), spin_phi(d));
}

hopping_list tbmodel::createHoppingCurrents_list(const std::string &op) {
  const char d = op[1];
  return createHoppingCurrents_list(d == 'Y' ? 1 : d == 'Z' ? 2 : 0);
}

hopping_list tbmodel::createHoppingSpinCurrents_list(const std::string &op) {
  const char d = op[1];
  const int dir = d == 'X' ? 0 : d == 'Y' ? 1 : d == 'Z' ? 2 : -1;
  return createHoppingSpinCurrents_list(dir, op.back());
}

hopping_list tbmodel::createHoppingSpinCurrents_list(int dir, char sdir) {
  if (sdir != 'X' && sdir != 'Y' && sdir != 'Z') fail(std::string("incorrect spin direction ") + sdir);
  return createHoppingSpinCurrents_list(dir, spin_theta(sdir), spin_phi(sdir));
}

hopping_list tbmodel::WannierOperator(const std::string &op_id) {
  if (oputil::is_velocity(op_id)) return createHoppingCurrents_list(op_id);
  if (oputil::is_spinvelocity(op_id)) return createHoppingSpinCurrents_list(op_id);
  if (oputil::is_spin(op_id)) return createHoppingSpinDensity_list(op_id);
  if (oputil::is_torque(op_id)) return createHoppingTorqueDensity_list(op_id);
  std::ifstream f(op_id + ".OP");
  if (f.good()) {
    std::string type;
    f >> type;
    std::array<cplx, 4> elems;
    for (auto &x : elems) {
      double re, im;
      if (!(f >> re >> im)) fail("truncated " + op_id + ".OP");
      x = {re, im};
    }
    std::cout << op_id << " is of the " << type << " type" << std::endl;
    return createHoppingSpinorialDensity_list(elems);
  }
  fail("incorrect wannier operator requested: " + op_id);
}

}  // namespace w2s
