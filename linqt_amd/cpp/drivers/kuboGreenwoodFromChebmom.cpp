// kuboGreenwoodFromChebmom file.chebmom2D broadening_meV
// Same contract as src/kuboGreenwoodFromChebmom.cpp:17-80: Jackson kernel, num_div = 30 * (highest
// moment number) energies on the tool's linear grid, "E value" rows written to
// KuboGreenWood_<label>JACKSON.dat.
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
  const int num_div = 30 * (int)mu.HighestMomentNumber();
  std::vector<double> E(num_div), out(num_div);
  linqt_amd::check(lq_linear_grid(linqt_amd::default_context(), LQ_LIN_GREENWOOD,
                                  reinterpret_cast<const double *>(mu.MomentVector().data()), mu.HighestMomentNumber(0), mu.HighestMomentNumber(1),
                                  (int)mu.SystemSize(), mu.BandWidth(), mu.BandCenter(), num_div, E.data(), out.data()),
                   "lq_linear_grid");
  const std::string file = "KuboGreenWood_" + mu.SystemLabel() + "JACKSON.dat";
  std::cout << "Saving the data in " << file << std::endl;
  std::ofstream f(file.c_str());
  for (int i = 0; i < num_div - 1; i++) f << E[i] << " " << out[i] << std::endl;
  std::cout << "The program finished succesfully." << std::endl;
  return 0;
}
