// spectralFunctionFromChebmom file.chebmom1D broadening_meV
// Same contract as src/spectralFunctionFromChebmom.cpp:17-75: Jackson kernel, num_div = 30 * (highest
// moment number) energies on the tool's linear grid, "E value" rows written to
// mean<label>JACKSON.dat.
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
  chebyshev::Moments1D mu(argv[1]);
  mu.ApplyJacksonKernel(b);
  const int num_div = 30 * (int)mu.HighestMomentNumber();
  std::vector<double> E(num_div), out(num_div);
  linqt_amd::check(lq_linear_grid(linqt_amd::default_context(), LQ_LIN_SPECTRAL,
                                  reinterpret_cast<const double *>(mu.MomentVector().data()), (int)mu.HighestMomentNumber(), 1,
                                  (int)mu.SystemSize(), mu.BandWidth(), mu.BandCenter(), num_div, E.data(), out.data()),
                   "lq_linear_grid");
  const std::string file = "mean" + mu.SystemLabel() + "JACKSON.dat";
  std::cout << "Saving the data in " << file << std::endl;
  std::ofstream f(file.c_str());
  for (int i = 0; i < num_div; i++) f << E[i] << " " << out[i] << std::endl;
  std::cout << "The program finished succesfully." << std::endl;
  return 0;
}
