// Elementwise kernels of the Runge-Kutta substep (main1.f90:384-655): one pass over
// the fields for everything that is diagonal in Fourier-physical-physical space.
// HBM-bound; grid-stride loops over complex points, 16 B loads and stores.
#include "rk3.h"

namespace duct {

namespace {
constexpr int kThreads = 256;
constexpr long kMaxBlocks = 8192;

unsigned blocks_for(long total)
{
    long b = (total + kThreads - 1) / kThreads;
    return (unsigned)(b < 1 ? 1 : (b > kMaxBlocks ? kMaxBlocks : b));
}
}  // namespace

__global__ __launch_bounds__(kThreads) void k_rk3_rhs(Rk3RhsArgs a)
{
    const long total = a.mm * a.nmodes;
    for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < total; e += (long)gridDim.x * kThreads) {
        const int mode = (int)(e / a.mm);
        const int k = a.mode0 + mode;
        const double xa = a.alp * (double)k;
        const double cu = a.clap * (-xa * xa) + a.rdt;     // laplacian3d's derivx(ider=2) + u/dt
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const d2v an = a.advnew[c][e];
            d2v h = -a.gam * an + cu * a.u[c][e];
            if (a.use_old) h -= a.rho * a.adv[c][e];
            if (c == 0) {
                const d2v pv = a.p[e];
                h += a.cp * xa * d2v{-pv[1], pv[0]};       // grad3d's derivx: (i alp k) p
                if (k == 0) h[0] += a.cdp;                  // the driving pressure gradient, real part
            }
            a.h[c][e] = a.s * h;
            a.adv[c][e] = an;
        }
    }
}

hipError_t launch_rk3_rhs(const Rk3RhsArgs& a, hipStream_t st)
{
    const long total = a.mm * a.nmodes;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rk3_rhs, dim3(blocks_for(total)), dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

__global__ __launch_bounds__(kThreads) void k_xderiv(const d2v* __restrict__ in, d2v* out, long mm, int nmodes,
                                                     int mode0, int ider, double alp, double scal, double beta,
                                                     const d2v* __restrict__ add0, const d2v* __restrict__ add1)
{
    const long total = mm * nmodes;
    for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < total; e += (long)gridDim.x * kThreads) {
        const double xa = alp * (double)(mode0 + (int)(e / mm));
        const d2v u = in[e];
        d2v v = ider == 1 ? scal * xa * d2v{-u[1], u[0]} : (-scal * xa * xa) * u;
        if (beta != 0.0) v += beta * out[e];
        if (add0) v += add0[e];
        if (add1) v += add1[e];
        out[e] = v;
    }
}

hipError_t launch_xderiv(const d2v* in, d2v* out, long mm, int nmodes, int mode0, int ider, double alp, double scal,
                         double beta, const d2v* add0, const d2v* add1, hipStream_t st)
{
    const long total = mm * nmodes;
    if (total <= 0) return hipSuccess;
    if (ider != 1 && ider != 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_xderiv, dim3(blocks_for(total)), dim3(kThreads), 0, st, in, out, mm, nmodes, mode0, ider,
                       alp, scal, beta, add0, add1);
    return hipGetLastError();
}

__global__ __launch_bounds__(kThreads) void k_rk3_update(d2v* p, d2v* u, const d2v* __restrict__ phi,
                                                         const d2v* __restrict__ rhsphi, long mm, int nmodes,
                                                         int mode0, double alp, double su)
{
    const long total = mm * nmodes;
    for (long e = (long)blockIdx.x * kThreads + threadIdx.x; e < total; e += (long)gridDim.x * kThreads) {
        const double xa = alp * (double)(mode0 + (int)(e / mm));
        const d2v f = phi[e];
        p[e] = p[e] + (f - rhsphi[e]);
        u[e] = u[e] + su * xa * d2v{-f[1], f[0]};
    }
}

hipError_t launch_rk3_update(d2v* p, d2v* u, const d2v* phi, const d2v* rhsphi, long mm, int nmodes, int mode0,
                             double alp, double su, hipStream_t st)
{
    const long total = mm * nmodes;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rk3_update, dim3(blocks_for(total)), dim3(kThreads), 0, st, p, u, phi, rhsphi, mm, nmodes,
                       mode0, alp, su);
    return hipGetLastError();
}

}  // namespace duct
