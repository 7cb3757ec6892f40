// chebyshev_coefficients.hpp -- same name as the reference header; the declarations live in linqt.hpp.
#pragma once
#include "linqt.hpp"
