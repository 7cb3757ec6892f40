// sparse_matrix.hpp -- same name as the reference header; the declarations live in linqt.hpp.
#pragma once
#include "linqt.hpp"
