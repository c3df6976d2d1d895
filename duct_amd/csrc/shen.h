// Shen-Legendre Galerkin solve: the batched symmetric tridiagonal systems along z
// (solV / solevenodd, solveshen.f90:335-513), one thread per (mode, y-eigenmode, parity).
#pragma once
#include "common.h"

namespace duct {

struct ShenTriArgs {
    d2v* V;                // in: h = (P^T W f W^T)(p, q) at p + q m per mode; out: v
    double* gam;           // scratch, m * m per mode (the forward sweep's factors)
    const double* md;      // mass diagonal, n = m - 2
    const double* mu;      // mass (k, k+2) entries
    const double* sd;      // stiffness diagonal
    const double* lam;     // y-eigenvalues (entry 0 unused for Neumann)
    double alambda, alp;
    int m, nmodes, mode0, neumann;
};
hipError_t launch_shen_tri(const ShenTriArgs& a, hipStream_t st);

}  // namespace duct
