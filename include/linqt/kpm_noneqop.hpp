// kpm_noneqop.hpp -- same name as the reference header (include/kpm_noneqop.hpp).
#pragma once
#include <iostream>

#include "linqt.hpp"

namespace chebyshev {
inline void printHelpMessage() {
  std::cout << "The program should be called with the following options: Label Op1 Op2 numMom BandWidth "
               "BandCenter (optional) num_states"
            << std::endl
            << std::endl;
}
inline void printWelcomeMessage() {
  std::cout << "WELCOME: This program will compute a table needed for expanding the correlation function in "
               "Chebyshev polynomialms"
            << std::endl;
}
}  // namespace chebyshev
