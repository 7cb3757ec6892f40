// kuboBastinFromChebmom_FFTgrid file.chebmom2D broadening_meV
// Same contract as src/kuboBastinFromChebmom_FFTgrid.cpp: Jackson on both
// axes, num_div = max(M0, M1) energies, writes num_div-1 rows of "E value"
// to KuboBastin_<label>JACKSON_FFTgrid.dat.
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
  const double b = std::stod(argv[2]);
  chebyshev::Moments2D mu(argv[1]);
  mu.ApplyJacksonKernel(b, b);
  const int num_div = mu.HighestMomentNumber();
  std::vector<double> E(num_div), out(num_div);
  linqt_amd::check(lq_bastin_fftgrid(linqt_amd::default_context(), reinterpret_cast<const double *>(mu.MomentVector().data()),
                                      mu.HighestMomentNumber(0), mu.HighestMomentNumber(1), (int)mu.SystemSize(),
                                      mu.BandWidth(), mu.BandCenter(), num_div, E.data(), out.data()),
                   "lq_bastin_fftgrid");
  const std::string name = "KuboBastin_" + mu.SystemLabel() + "JACKSON_FFTgrid.dat";
  std::cout << "Saving the data in " << name << std::endl;
  std::ofstream f(name.c_str());
  for (int i = 0; i < num_div - 1; i++) f << E[i] << " " << out[i] << std::endl;
  std::cout << "The program finished succesfully." << std::endl;
  return 0;
}
