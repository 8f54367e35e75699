#pragma once
#include <vector>

#include "common.h"

namespace cx {

// n-point Rys rule at x: s_i = t_i^2 (ascending) and weights, long double.
void rys_rule_ld(int n, long double x, long double* s, long double* w);

// Chebyshev table (see rys_table.cpp) and asymptotic constants.
void rys_table_build(std::vector<double>& coef, std::vector<double>& large,
                     int off[kMaxRoots + 1]);

}  // namespace cx
