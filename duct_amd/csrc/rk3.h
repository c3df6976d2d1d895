// Elementwise kernels of the Runge-Kutta substep (main1.f90:384-655) on complex(8)
// fields (m, m, nmodes), y fastest; the derivative products along y and z and the
// solves are the existing collocation and projection kernels.
#pragma once
#include "common.h"

namespace duct {

// The x-parts of steps (i)-(vii) of one substep, fused, for the three components c:
//   h_c  = s (-rho adv_c + (-gam) advnew_c + (clap (-(alp k)^2) + rdt) u_c
//             [c = 0: + cp (i alp k) p,  k = 0: + cdp on the real part])
//   adv_c = advnew_c                               (kept for the next substep)
// use_old = 0 drops the rho term without reading adv (the first substep, rkkc(1) = 0).
struct Rk3RhsArgs {
    const d2v* u[3];
    d2v* adv[3];           // in: adv of the previous substep; out: advnew
    const d2v* advnew[3];
    const d2v* p;
    d2v* h[3];
    long mm;               // points per mode plane
    int nmodes, mode0, use_old;
    double alp, s, rho, gam, clap, rdt, cp, cdp;
};
hipError_t launch_rk3_rhs(const Rk3RhsArgs& a, hipStream_t st);

// out = ((beta out + scal (i alp k)^ider in) + add0) + add1, ider = 1 or 2 (derivx,
// fourier.F:117-188); add0 / add1 may be null (the y and z terms of the same sum, computed beforehand)
hipError_t launch_xderiv(const d2v* in, d2v* out, long mm, int nmodes, int mode0, int ider, double alp, double scal,
                         double beta, const d2v* add0, const d2v* add1, hipStream_t st);

// update_uvwp's elementwise part (operators.F:564-577): p += phi - rhsphi, u += su (i alp k) phi
hipError_t launch_rk3_update(d2v* p, d2v* u, const d2v* phi, const d2v* rhsphi, long mm, int nmodes, int mode0,
                             double alp, double su, hipStream_t st);

}  // namespace duct
