// wannier2sparse LABEL Dim0 Dim1 Dim2 [Operators | all]
//
// The reference's tool (utilities/wannier2sparse/src/main.cpp,
// include/w2sp_arguments.hpp): reads LABEL.xyz, LABEL.uc, LABEL_hr.dat (and
// LABEL.stdis when present), wraps the model in a Dim0 x Dim1 x Dim2
// supercell and writes LABEL.HAM.CSR plus LABEL.<OP>.CSR per operator.
// The supercell operators are generated on the GPU.
#include <deque>
#include <iostream>
#include <string>

#include "linqt/wannier2sparse.hpp"

int main(int argc, char *argv[]) {
  std::deque<std::string> args(argv + 1, argv + argc);
  const std::string usage =
      "wannier2sparse LABEL Dim0 Dim1 Dim2 [Operators]\n"
      "LABEL names the LABEL_hr.dat, LABEL.uc, LABEL.xyz and LABEL.stdis files;\n"
      "Dim0 Dim1 Dim2 (> 0) are the supercell dimensions;\n"
      "[Operators]: none (the Hamiltonian only), 'all', or a list such as VX VY SZ VXSZ TX.";
  if (args.empty() || args.front() == "help") {
    std::cout << usage << std::endl;
    return args.empty() ? 1 : 0;
  }
  const std::string label = args.front();
  args.pop_front();
  w2s::hopping_list::cellID_t dims;
  for (int i = 0; i < 3; ++i) {
    if (args.empty()) {
      std::cerr << "ERROR: the program should be called with arguments (LABEL, Dim0, Dim1, Dim2).\n" << usage << std::endl;
      return 1;
    }
    dims[i] = std::stoi(args.front());
    args.pop_front();
  }
  std::deque<std::string> ops;
  if (!args.empty() && args.front() == "all") {
    for (const char *op : w2s::oputil::AVAIL_OPS) ops.push_back(op);
  } else {
    ops = args;
  }
  std::cout << "Using " << label << " as the system's identification label" << std::endl;
  w2s::tbmodel model;
  model.readOrbitalPositions(label + ".xyz");
  model.readUnitCell(label + ".uc");
  model.readWannierModel(label + "_hr.dat");
  model.readStaticDisorder(label + ".stdis");
  std::cout << "Creating Hamiltonian" << std::endl;
  model.Hopping_List().wrap_in_supercell(dims).save_hopping_list_as_csr(label + ".HAM.CSR");
  for (const auto &op : ops) {
    std::cout << "Creating " << op << std::endl;
    model.WannierOperator(op).wrap_in_supercell(dims).save_hopping_list_as_csr(label + "." + op + ".CSR");
  }
  std::cout << "The program finished" << std::endl;
  return 0;
}
