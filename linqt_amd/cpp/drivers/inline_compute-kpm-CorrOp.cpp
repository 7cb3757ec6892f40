// inline_compute-kpm-CorrOp LABEL OPL OPR numMom numTimeSteps TimeMax [num_states]
// Same contract as src/inline_compute-kpm-CorrOp.cpp:22-90: operators from
// operators/<LABEL>.<ID>.CSR, bounds from ./BOUNDS, time step TimeMax /
// (numTimeSteps - 1) fs; writes TimeCorr<OPL>-<OPR><LABEL>KPM_M<M>_state<label>.chebmomTD.
// As there, the states are the default generator's (the state-file branch of
// the source tests argc == 5, which the argument check excludes) and
// num_states is read but not used.
#include <iostream>
#include <string>

#include "chebyshev_solver.hpp"
#include "kpm_noneqop.hpp"

int main(int argc, char *argv[]) {
  if (!(argc == 7 || argc == 8)) {
    std::cout << "The program should be called with the following options: Label Op1 Op2 numMom numTimeSteps "
                 "TimeStep BandWidth BandCenter (optional) num_states"
              << std::endl
              << std::endl;
    return 0;
  }
  const std::string LABEL = argv[1], S_OPL = argv[2], S_OPR = argv[3], S_NMOM = argv[4], S_NTIME = argv[5],
                    S_TMAX = argv[6];
  const int numMoms = atoi(S_NMOM.c_str());
  const int numTimes = atoi(S_NTIME.c_str());
  const double tmax = std::stod(S_TMAX);
  chebyshev::MomentsTD chebMoms(numMoms, numTimes);
  SparseMatrixType OP[3];
  OP[0].SetID("HAM");
  OP[1].SetID(S_OPL);
  OP[2].SetID(S_OPR);
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
  chebMoms.TimeDiff(tmax / (numTimes - 1));
  chebMoms.SetAndRescaleHamiltonian(OP[0]);
  chebMoms.Print();
  qstates::generator gen;
  chebyshev::TimeDependentCorrelations(OP[1], OP[2], chebMoms, gen);
  const std::string out =
      "TimeCorr" + S_OPL + "-" + S_OPR + LABEL + "KPM_M" + S_NMOM + "_state" + gen.StateLabel() + ".chebmomTD";
  std::cout << "Saving convergence data in " << out << std::endl;
  chebMoms.saveIn(out);
  std::cout << "End of program" << std::endl;
  return 0;
}
