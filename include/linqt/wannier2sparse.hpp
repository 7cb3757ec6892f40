// wannier2sparse.hpp -- the reference's wannier2sparse tool over the C ABI.
//
// Keeps the names of utilities/wannier2sparse/include/{hopping_list,tbmodel,
// wannier_parser,operator_utils}.hpp for the constant-hopping models the tool
// builds: a tight-binding model read from LABEL_hr.dat / LABEL.xyz / LABEL.uc
// and the operators derived from it (velocities, spin densities, spin
// velocities, torques, custom spinorial densities).  The unit-cell work
// (parsing, operator construction) is host code; wrapping a list into a
// supercell and building the CSR runs on the GPU (lq_csr_from_hoppings).
// Static disorder files (LABEL.stdis) are not supported: the reference draws
// their values from std::random_device, so no run reproduces another.
#ifndef LINQT_AMD_WANNIER2SPARSE_HPP
#define LINQT_AMD_WANNIER2SPARSE_HPP

#include <array>
#include <complex>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "linqt.hpp"

namespace w2s {

// utilities/wannier2sparse/include/hopping_list.hpp:28-229.  The hoppings are
// kept in insertion order; every traversal (and the generator) visits them in
// the order of the reference's multimap: by tag string "c0 c1 c2 e0 e1 ",
// equal tags in insertion order.
class hopping_list {
 public:
  typedef std::complex<double> value_t;
  typedef std::array<int, 2> edge_t;
  typedef std::array<int, 3> cellID_t;

  hopping_list() = default;
  int WannierBasisSize() const { return num_wann; }
  void SetWannierBasisSize(int n) { num_wann = n; }
  cellID_t Bounds() const { return cellSizes; }
  void AddHopping(const cellID_t &cellID, const value_t &value, const edge_t &edge);
  size_t size() const { return val.size(); }
  // Indices in multimap order (hopping_list::hoppings iteration order).
  std::vector<int> ordered() const;

  // hopping_list::add_from_wannier (src/hopping_list.cpp:114-154).
  hopping_list &add_from_wannier(const std::tuple<int, std::vector<std::string>> &wannier_data);
  // hopping_list::wrap_in_supercell (src/hopping_list.cpp:156-225): records the
  // supercell; the wrap itself happens on the device when the operator is built.
  hopping_list wrap_in_supercell(const cellID_t &cellDim) const;
  // The wrapped operator on the device (lq_csr_from_hoppings).
  lq_csr BuildOperator(lq_ctx ctx) const;
  // hopping_list::save_hopping_list_as_csr (src/hopping_list.cpp:227-285):
  // the file the reference writes, byte for byte up to the device-built values.
  void save_hopping_list_as_csr(const std::string &output_filename) const;

  std::vector<cellID_t> cell;
  std::vector<edge_t> edge;
  std::vector<value_t> val;

 private:
  int num_wann = 0;
  cellID_t cellSizes{1, 1, 1};
};

// Reference file readers (src/wannier_parser.cpp).
std::tuple<int, std::vector<std::string>> read_wannier_file(const std::string &wannier_filename);
std::vector<std::tuple<std::string, std::array<double, 3>>> read_xyz_file(const std::string &xyz_filename);
std::array<std::array<double, 3>, 3> read_unit_cell_file(const std::string &uc_filename);

namespace oputil {
// include/operator_utils.hpp:12-33
struct op_matrix {
  typedef std::complex<double> value_t;
  explicit op_matrix(int dim) : n(dim), data((size_t)dim * dim, value_t(0.0)) {}
  value_t &operator()(int i, int j) { return data[(size_t)i * n + j]; }
  int size() const { return n; }
  int n;
  std::vector<value_t> data;
};
extern const char *const AVAIL_OPS[18];
bool is_velocity(const std::string &op);
bool is_spinvelocity(const std::string &op);
bool is_spin(const std::string &op);
bool is_torque(const std::string &op);
}  // namespace oputil

std::array<double, 3> tag2cartesian(const std::array<int, 3> &tag, const std::array<std::array<double, 3>, 3> &lat_vecs);
double volume(const std::array<std::array<double, 3>, 3> &uc);

// utilities/wannier2sparse/include/tbmodel.hpp:38-229 (src/tbmodel.cpp).
class tbmodel {
 public:
  typedef std::tuple<std::string, std::array<double, 3>> orbPos_t;
  typedef std::array<std::array<double, 3>, 3> unitCell_t;

  void readUnitCell(const std::string &inputfile);
  void readOrbitalPositions(const std::string &inputfile);
  void readWannierModel(const std::string &inputfile);
  // Fails when the file exists and holds a disorder group (see the header note).
  void readStaticDisorder(const std::string &inputfile);

  hopping_list &Hopping_List() { return hl; }
  hopping_list createLocalHopping_list();
  hopping_list createHoppingDensity_list(oputil::op_matrix OP);
  hopping_list createHoppingCurrents_list(int dir);
  hopping_list createHoppingSpinDensity_list(double theta, double phi);
  hopping_list createHoppingTorqueDensity_list(double theta, double phi);
  hopping_list createHoppingSpinCurrents_list(int dir, double theta, double phi);
  hopping_list createHoppingCurrents_list(const std::string &op);
  hopping_list createHoppingSpinorialDensity_list(const std::array<std::complex<double>, 4> &op);
  hopping_list createHoppingSpinDensity_list(const std::string &op);
  hopping_list createHoppingTorqueDensity_list(const std::string &op);
  hopping_list createHoppingSpinCurrents_list(const std::string &op);
  hopping_list createHoppingSpinCurrents_list(int dir, char sdir);
  hopping_list WannierOperator(const std::string &op_id);

  unitCell_t lat_vecs{};
  std::vector<orbPos_t> orbPos_list;
  hopping_list hl;

 private:
  int NumberOfSites() const { return (int)orbPos_list.size() / 2; }
  int getSpinIndex(int n) const { return n / NumberOfSites(); }
  hopping_list::edge_t ChangeSpinIndex(const hopping_list::edge_t &n, const hopping_list::edge_t &s) const {
    return {n[0] % NumberOfSites() + s[0] * NumberOfSites(), n[1] % NumberOfSites() + s[1] * NumberOfSites()};
  }
  oputil::op_matrix createSpinorialOp(const std::array<std::complex<double>, 4> &op);
  oputil::op_matrix createTorqueMatrix(double theta, double phi);
};

}  // namespace w2s

#endif
