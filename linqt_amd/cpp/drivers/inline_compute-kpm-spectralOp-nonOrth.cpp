// inline_compute-kpm-spectralOp-nonOrth LABEL OP numMom [statefile]
// Same contract as src/inline_compute-kpm-spectralOp-nonOrth.cpp:22-96: reads
// operators/LABEL.HAM.CSR, LABEL.OP.CSR (unless OP is "1") and the overlap
// LABEL.S.CSR, ./BOUNDS, writes
// SpectralOp-nonOrth-<OP><LABEL>KPM_M<numMom>_state<label>.chebmom1D-nonOrth
// (Spectral-nonOrth-OP1... for the identity).
#include <iostream>
#include <string>

#include "chebyshev_solver.hpp"

int main(int argc, char *argv[]) {
  if (!(argc == 4 || argc == 5)) {
    std::cout << "The program should be called with the following options: Label Op numMom " << std::endl << std::endl;
    std::cout << "Label will be used to look for Label.HAM, Label.Op and the overlap Label.S" << std::endl;
    return 0;
  }
  const std::string LABEL = argv[1], S_OP = argv[2], S_NMOM = argv[3];
  const int numMoms = atoi(S_NMOM.c_str());
  chebyshev::Moments1D_nonOrth chebMoms(numMoms);
  SparseMatrixType OP[3];
  OP[0].SetID("HAM");
  OP[1].SetID(S_OP);
  OP[2].SetID("S");
  SparseMatrixBuilder builder;
  std::array<double, 2> bounds{-100, 100};
  for (int i = 0; i < 3; i++) {
    if (OP[i].isIdentity()) {
      std::cout << "The operator " << OP[i].ID() << " is treated as the identity" << std::endl;
      continue;
    }
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
  if (argc == 5) gen = qstates::LoadStateFile(argv[4]);
  chebMoms.set_S(OP[2]);
  chebyshev::SpectralMoments_nonOrth(OP[1], chebMoms, gen);
  std::string prefix = "SpectralOp-nonOrth-" + OP[1].ID();
  if (OP[1].isIdentity()) prefix = "Spectral-nonOrth-OP" + OP[1].ID();
  const std::string out = prefix + LABEL + "KPM_M" + S_NMOM + "_state" + gen.StateLabel() + ".chebmom1D-nonOrth";
  std::cout << "Saving the moments in  " << out << std::endl;
  chebMoms.saveIn(out);
  std::cout << "End of program" << std::endl;
  return 0;
}
