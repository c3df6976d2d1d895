// Wall-data correction and edges of the mixed-boundary solve (solvediag_temp.f90:2-352).
// Both kernels touch O(m) values per line and are bandwidth- or latency-bound.
#include "temp.h"

namespace duct {

namespace {
// The effective wall values (g1, g2) of one line: the interior correction of
// mod_term_FT* is -D2(i, 1) g1 - D2(i, m) g2 for every wall-type pair
// (solvediag_temp.f90:136-247), with D0 = D(1, :), Dm = D(m, :).
__device__ __forceinline__ void wall_values(int b1, int b2, const double* D0, const double* Dm, int m, double v1,
                                            double v2, double& g1, double& g2)
{
    const double d00 = D0[0], d0m = D0[m - 1], dm0 = Dm[0], dmm = Dm[m - 1];
    if (!b1 && !b2) { g1 = v1; g2 = v2; }
    else if (!b1 && b2) { g1 = v1; g2 = (v2 - dm0 * v1) / dmm; }
    else if (b1 && !b2) { g1 = (v1 - d0m * v2) / d00; g2 = v2; }
    else {
        const double det = d00 * dmm - d0m * dm0;
        g1 = (dmm * v1 - d0m * v2) / det;
        g2 = (d00 * v2 - dm0 * v1) / det;
    }
}

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
}  // namespace

__global__ __launch_bounds__(256) void k_temp_rhs(TempArgs a)
{
    const int my = a.my, mz = a.mz, ny = my - 2, nz = mz - 2;
    const long total = (long)ny * nz * a.nmodes;
    const double* D2y0 = a.vec;
    const double* D2ym = a.vec + my;
    const double* Dy0 = a.vec + 2 * my;
    const double* Dym = a.vec + 3 * my;
    const double* D2z0 = a.vec + 4 * my;
    const double* D2zm = D2z0 + mz;
    const double* Dz0 = D2z0 + 2 * mz;
    const double* Dzm = D2z0 + 3 * mz;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const int i = 1 + (int)(e % ny);              // y
        const int j = 1 + (int)((e / ny) % nz);       // z
        const int mode = (int)(e / ((long)ny * nz));
        const double* by = a.bcy + (size_t)mode * 2 * mz;
        const double* bz = a.bcz + (size_t)mode * 2 * my;
        double gy1, gy2, gz1, gz2;
        wall_values(a.by1, a.by2, Dy0, Dym, my, by[j], by[mz + j], gy1, gy2);
        wall_values(a.bz1, a.bz2, Dz0, Dzm, mz, bz[i], bz[my + i], gz1, gz2);
        const double c = (-D2y0[i] * gy1 - D2ym[i] * gy2) + (-D2z0[j] * gz1 - D2zm[j] * gz2);
        const size_t idx = (size_t)mode * my * mz + i + (size_t)j * my;
        const d2v r = a.rhs[idx];
        a.out[idx] = d2v{r[0] + c, r[1] + c};
    }
}

hipError_t launch_temp_rhs(const TempArgs& a, hipStream_t st)
{
    const long total = (long)(a.my - 2) * (a.mz - 2) * a.nmodes;
    if (total <= 0) return hipSuccess;
    long blocks = (total + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_temp_rhs, dim3((unsigned)blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

// blockIdx.x = mode; blockIdx.y = 0: the y = 1 and y = my edges of every interior z line (sums
// along the contiguous y), 1: the z = 1 and z = mz edges of every interior y line; one wave per line.
__global__ __launch_bounds__(256) void k_temp_edges(TempArgs a)
{
    const int my = a.my, mz = a.mz, mode = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool zdir = blockIdx.y == 1;
    const bool k0 = a.mode0 + mode == 0;
    d2v* U = a.out + (size_t)mode * my * mz;
    const int mp = zdir ? mz : my;                   // points of a line
    const int ml = zdir ? my : mz;                   // lines (the other direction), walls included
    const int b1 = zdir ? a.bz1 : a.by1, b2 = zdir ? a.bz2 : a.by2;
    const double* D0 = zdir ? a.vec + 4 * my + 2 * mz : a.vec + 2 * my;
    const double* Dm = D0 + mp;
    const double* bc = (zdir ? a.bcz : a.bcy) + (size_t)mode * 2 * ml;
    const double d00 = D0[0], d0m = D0[mp - 1], dm0 = Dm[0], dmm = Dm[mp - 1];
    const size_t sp = zdir ? (size_t)my : 1;         // stride between the points of a line
    const size_t sl = zdir ? 1 : (size_t)my;         // stride between lines
    for (int line = 1 + wave; line < ml - 1; line += 4) {
        d2v sa = {0.0, 0.0}, sb = {0.0, 0.0};
        for (int p = 1 + lane; p < mp - 1; p += 64) {
            const d2v u = U[(size_t)line * sl + (size_t)p * sp];
            sa += D0[p] * u;
            sb += Dm[p] * u;
        }
        sa = d2v{wsum(sa[0]), wsum(sa[1])};
        sb = d2v{wsum(sb[0]), wsum(sb[1])};
        if (lane == 0) {
            const double v1 = bc[line], v2 = bc[ml + line];
            d2v e0, e1;
            if (!b1 && !b2) {
                e0 = d2v{v1, v1}; e1 = d2v{v2, v2};
            } else if (!b1 && b2) {
                e0 = d2v{v1, v1};
                e1 = (d2v{v2, v2} - dm0 * e0 - sb) / dmm;
            } else if (b1 && !b2) {
                e1 = d2v{v2, v2};
                e0 = (d2v{v1, v1} - d0m * e1 - sa) / d00;
            } else {
                const double det = d00 * dmm - d0m * dm0;
                const d2v ra = d2v{v1, v1} - sa, rb = d2v{v2, v2} - sb;
                e0 = (dmm * ra - d0m * rb) / det;
                e1 = (-dm0 * ra + d00 * rb) / det;
            }
            if (k0) { e0[1] = 0.0; e1[1] = 0.0; }
            U[(size_t)line * sl] = e0;
            U[(size_t)line * sl + (size_t)(mp - 1) * sp] = e1;
        }
    }
    if (!zdir && threadIdx.x < 4) {   // corners keep the right-hand side's values
        const size_t c = threadIdx.x == 0 ? 0 : threadIdx.x == 1 ? (size_t)(my - 1)
                         : threadIdx.x == 2 ? (size_t)(mz - 1) * my : (size_t)(mz - 1) * my + my - 1;
        d2v v = a.rhs[(size_t)mode * my * mz + c];
        if (k0) v[1] = 0.0;
        U[c] = v;
    }
}

hipError_t launch_temp_edges(const TempArgs& a, hipStream_t st)
{
    if (a.nmodes <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_temp_edges, dim3((unsigned)a.nmodes, 2), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace duct
