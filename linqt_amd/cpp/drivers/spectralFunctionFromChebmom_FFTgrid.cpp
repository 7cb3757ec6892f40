// spectralFunctionFromChebmom_FFTgrid file.chebmom1D broadening_meV
// Same contract as src/spectralFunctionFromChebmom_FFTgrid.cpp:17-79: writes
// mean<label>JACKSON_FFTgrid.dat with "E rho" on num_div = #moments energies.
#include <iostream>
#include <string>

#include "chebyshev_moments.hpp"

int main(int argc, char *argv[]) {
  if (argc != 3) {
    std::cout << "The program should be called with the following options: moments_filename broadening(meV)"
              << std::endl
              << std::endl;
    return 0;
  }
  chebyshev::Moments1D mu(argv[1]);
  mu.ApplyJacksonKernel(std::stod(argv[2]));
  const int num_div = (int)mu.HighestMomentNumber();
  std::vector<double> E(num_div), rho(num_div);
  linqt_amd::check(lq_dos_fftgrid(linqt_amd::default_context(), reinterpret_cast<const double *>(mu.MomentVector().data()),
                                  num_div, (int)mu.SystemSize(), mu.BandWidth(), mu.BandCenter(), num_div, E.data(),
                                  rho.data()),
                   "lq_dos_fftgrid");
  const std::string out = "mean" + mu.SystemLabel() + "JACKSON_FFTgrid.dat";
  std::cout << "Saving the data in " << out << std::endl;
  std::ofstream f(out.c_str());
  for (int i = 0; i < num_div; i++) f << E[i] << " " << rho[i] << std::endl;
  std::cout << "The program finished succesfully." << std::endl;
  return 0;
}
