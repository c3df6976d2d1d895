// Launchers of the bandwidth-bound helper kernels (aux.hip).
#pragma once
#include "common.h"

namespace duct {

hipError_t launch_transpose_f64(const double* A, double* C, int Nx, int Ny, int nbatch, double alpha, hipStream_t st);
hipError_t launch_transpose_c128(const d2v* A, d2v* C, int Nx, int Ny, int nbatch, double alpha, hipStream_t st);
hipError_t launch_invdiag(const double* F, const double* ED2, double* C, long n, double xlam, hipStream_t st);

}  // namespace duct
