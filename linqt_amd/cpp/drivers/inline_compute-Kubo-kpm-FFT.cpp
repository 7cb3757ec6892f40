// inline_compute-Kubo-kpm-FFT LABEL OPR OPL numMom [statefile]
// Same contract as src/inline_compute-Kubo-kpm-FFT.cpp:24-94 (Greenwood,
// one section, nump = numMom); LQ_KUBO_FORMULA=bastin and LQ_KUBO_NUMP=<n>
// select the Bastin formula and another energy count, which the reference
// reaches only through its class API.
#include <cstring>
#include <iostream>
#include <string>

#include "Kubo_solver_FFT.hpp"

int main(int argc, char *argv[]) {
  if (!(argc == 5 || argc == 6)) {
    std::cout << "The program should be called with the following options: Label Op1 Op2 numMom num_states (default 1)"
              << std::endl
              << std::endl;
    return 0;
  }
  const std::string LABEL = argv[1], S_OPR = argv[2], S_OPL = argv[3], S_NUM_MOM = argv[4];
  const int numMoms = atoi(argv[4]);
  int num_sections = 1, nump = numMoms;
  chebyshev::formula sym_formula = chebyshev::KUBO_GREENWOOD;
  const char *f = getenv("LQ_KUBO_FORMULA");
  if (f && strcmp(f, "bastin") == 0) sym_formula = chebyshev::KUBO_BASTIN;
  const char *np = getenv("LQ_KUBO_NUMP");
  if (np && atoi(np) > 0) nump = atoi(np);
  chebyshev::Vectors_sliced chebVec(numMoms, num_sections);
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
  chebVec.SystemLabel(LABEL);
  chebVec.BandWidth((bounds[1] - bounds[0]) * 1.0);
  chebVec.BandCenter((bounds[1] + bounds[0]) * 0.5);
  chebVec.SetAndRescaleHamiltonian(OP[0]);
  qstates::generator gen;
  if (argc == 6) gen = qstates::LoadStateFile(argv[5]);
  std::string out = std::string(sym_formula == chebyshev::KUBO_BASTIN ? "Bastin_FFT" : "Greenwood_FFT") + S_OPR + "-" +
                    S_OPL + LABEL + "KPM_M" + S_NUM_MOM + "x" + S_NUM_MOM + "_state" + gen.StateLabel() +
                    ".conductivity";
  chebyshev::Kubo_solver_FFT solver(numMoms, num_sections, nump, sym_formula, chebVec, out);
  solver.compute(OP[1], OP[2], gen);
  std::cout << "End of program" << std::endl;
  return 0;
}
