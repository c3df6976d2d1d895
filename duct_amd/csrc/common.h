// Shared declarations for the duct cross-plane solve library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace duct {

typedef double d2v __attribute__((ext_vector_type(2)));
typedef double d4v __attribute__((ext_vector_type(4)));

// Status codes returned by every extern "C" entry point (0 = success).
enum Status : int {
    OK = 0,
    ERR_ARG = 1,         // bad argument (size, pointer, bound character)
    ERR_NOT_SETUP = 2,   // duct_setup_ has not been called for this bound / size
    ERR_HIP = 3,         // a HIP runtime call or kernel launch failed
    ERR_ALLOC = 4,       // device allocation failed
    ERR_UNSUPPORTED = 5, // size outside what the kernels are built for
};

constexpr int kMaxFields = 8;   // fields per batched launch
constexpr int kMinM = 4;        // smallest CGL grid (n = m-2 >= 2)
constexpr int kMaxM = 514;      // n = m-2 <= 512

// Padded interior size for the MFMA tiles: the interior n = m-2 is rounded up
// to one of the sizes the projection kernels are instantiated for.
inline int padded_n(int n) {
    if (n <= 32) return 32;
    if (n <= 64) return 64;
    if (n <= 128) return 128;
    if (n <= 256) return 256;
    if (n <= 512) return 512;
    return -1;
}

// One group of fields that share a basis (all Dirichlet or all Neumann) and
// are processed by one launch of each projection pass.
struct FieldGroup {
    const d2v* in[kMaxFields];   // complex(8) rhs (m, m, nmodes), interior read only
    d2v* out[kMaxFields];        // complex(8) sol (m, m, nmodes), interior written
    double alambda[kMaxFields];  // per-field alambda (helm3d.f90:67)
    int nf;                      // number of fields in the group
    int nmodes;                  // complex modes per field
    int mode0;                   // Fourier index of the first mode (ib)
    int m;                       // CGL points per direction
    int n;                       // interior points (m-2)
    double alp;                  // streamwise wavenumber unit (xalp(i)=alp*i)
};

// xlambda = alambda + xalp(imode)**2 with xalp(i) = alp*i (helm3d.f90:67,
// fourier.F:201-203), rounded after every operation on host and device alike
// so that the Neumann gauge test |xlambda| <= 1e-10 agrees between the two.
__host__ __device__ inline double xlambda_of(double alambda, double alp, int imode)
{
#pragma clang fp contract(off)
    const double xa = alp * (double)imode;
    return alambda + xa * xa;
}

}  // namespace duct
