// inline_compute-kpm-nonEqOp LABEL OPR OPL numMom [statefile]
// Same contract as src/inline_compute-kpm-nonEqOp.cpp:23-85 (the first
// operator argument goes on the left of the table, as there); writes
// NonEqOp<OPR>-<OPL><LABEL>KPM_M<M>x<M>_state<label>.chebmom2D.  BATCH_SIZE
// from the environment as in include/kpm_noneqop.hpp:27-39.
#include <iostream>
#include <string>

#include "chebyshev_solver.hpp"
#include "kpm_noneqop.hpp"

int main(int argc, char *argv[]) {
  if (!(argc == 5 || argc == 6)) {
    std::cout << "The program should be called with the following options: Label Op1 Op2 numMom num_states (default 1)"
              << std::endl
              << std::endl;
    return 0;
  }
  const std::string LABEL = argv[1], S_OPR = argv[2], S_OPL = argv[3], S_NUM_MOM = argv[4];
  const int numMoms = atoi(argv[4]);
  chebyshev::Moments2D chebMoms(numMoms, numMoms);
  SparseMatrixType OP[3];
  OP[0].SetID("HAM");
  OP[1].SetID(S_OPR);
  OP[2].SetID(S_OPL);
  SparseMatrixBuilder builder;
  std::array<double, 2> bounds{-100, 100};
  for (int i = 0; i < 3; i++) {
    builder.setSparseMatrix(&OP[i]);
    builder.BuildOPFromCSRFile("operators/" + LABEL + "." + OP[i].ID() + ".CSR");
    if (i == 0) bounds = chebyshev::utility::SpectralBounds(OP[0]);
  }
  chebMoms.SystemLabel(LABEL);
  chebMoms.BandWidth((bounds[1] - bounds[0]) * 1.0);
  chebMoms.BandCenter((bounds[1] + bounds[0]) * 0.5);
  chebMoms.SetAndRescaleHamiltonian(OP[0]);
  chebMoms.Print();
  qstates::generator gen;
  if (argc == 6) gen = qstates::LoadStateFile(argv[5]);
  chebyshev::CorrelationExpansionMoments(OP[1], OP[2], chebMoms, gen);
  const std::string out = "NonEqOp" + S_OPR + "-" + S_OPL + LABEL + "KPM_M" + S_NUM_MOM + "x" + S_NUM_MOM + "_state" +
                          gen.StateLabel() + ".chebmom2D";
  chebMoms.saveIn(out);
  std::cout << "End of program" << std::endl;
  return 0;
}
