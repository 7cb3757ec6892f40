// timeCorrelationsFromChebmom file.chebmomTD broadening_meV fermi_energy
// Same contract as src/timeCorrelationsFromChebmom.cpp:19-75: Jackson kernel on
// the moment axis, rows "n*dt value" with value = N/a sum_m delta_m(x_F) Re mu(m, n),
// written to mean<label>EF<fermi_energy>JACKSON.dat.
#include <iostream>
#include <string>

#include "chebyshev_moments.hpp"

int main(int argc, char *argv[]) {
  if (argc != 4) {
    std::cout << "The program should be called with the following options: moments_filename broadening(meV) "
                 "maximum_time (fs)"
              << std::endl
              << std::endl;
    return 0;
  }
  const double broadening = std::stod(argv[2]);
  const double fermiEner = std::stod(argv[3]);
  chebyshev::MomentsTD mu(argv[1]);
  mu.ApplyJacksonKernel(broadening);
  mu.Print();
  const int nT = (int)mu.MaxTimeStep();
  std::vector<double> t(nT), out(nT);
  linqt_amd::check(lq_timecorr_post(linqt_amd::default_context(), reinterpret_cast<const double *>(mu.MomentVector().data()),
                                    (int)mu.HighestMomentNumber(), nT, (int)mu.SystemSize(), mu.BandWidth(),
                                    mu.BandCenter(), mu.TimeDiff(), fermiEner, t.data(), out.data()),
                   "lq_timecorr_post");
  const std::string name = "mean" + mu.SystemLabel() + "EF" + std::to_string(fermiEner) + "JACKSON.dat";
  std::ofstream f(name.c_str());
  for (int n = 0; n < nT; n++) f << t[n] << " " << out[n] << std::endl;
  std::cout << "The program finished succesfully." << std::endl;
  return 0;
}
