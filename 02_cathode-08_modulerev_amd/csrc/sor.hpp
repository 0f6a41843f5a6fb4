// sor.hpp -- point SOR of include/solverFuncs.hpp:169-216 on the GPU.
#pragma once

#include "common.hpp"

namespace cathode {

int sor_solve_host(const double* aP, const double* aE, const double* aW, const double* aN,
                   const double* aS, const double* b, const int* boxes, int n_bl, int ni2, int nj2,
                   int ld, double alpha, int maxITR, double tol, double* x, hipStream_t s,
                   cathode_status* st);
void sor_release();

}  // namespace cathode
