// Small per-plane kernels around the projection passes:
//   k_gauge      Neumann k=0 gauge fix          solvediag.f90:161-206
//   k_boundary   Dirichlet zero / keep, Neumann edges + corners
//                                               helm3d.f90:48-60, solvediag.f90:251-322
//   k_bc_correct Dirichlet RHS correction       solvediag.f90:37-45
// All three are bandwidth- or latency-bound and touch O(n) or O(n^2) values
// per plane; re and im of a complex plane are processed together as double2.
#include "common.h"
#include "proj.h"

namespace duct {

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ d2v wave_sum2(d2v v)
{
    return d2v{wave_sum(v[0]), wave_sum(v[1])};
}

// ---------------------------------------------------------------------------
// Gauge.  For a plane G = P^-1 F P^-T ./ den with G(a,a) = 0, set
//   G(a,a) = (wanted - sum_l P(a,l) sum_k P(a,k) G(l,k)) / P(a,a)^2
// so that (P G P^T)(a,a) = wanted.  One workgroup per listed plane.
__global__ __launch_bounds__(256) void k_gauge(GaugeArgs a, GaugeList list)
{
    const int plane = list.plane[blockIdx.x];
    const int NP = a.NP, n = a.g.n, io = a.gauge;
    double* G = a.G + (size_t)plane * NP * NP;
    const int tid = threadIdx.x;
    double part = 0.0;
    for (int l = tid; l < n; l += 256) {
        double val = 0.0;
        for (int k = 0; k < n; ++k) val += a.prow[k] * G[l + (size_t)k * NP];
        part += a.prow[l] * val;
    }
    __shared__ double red[4];
    part = wave_sum(part);
    if ((tid & 63) == 0) red[tid >> 6] = part;
    __syncthreads();
    if (tid == 0) {
        const double valore = (red[0] + red[1]) + (red[2] + red[3]);
        G[io + (size_t)io * NP] = (a.wanted - valore) / (a.pdiag * a.pdiag);
    }
}

hipError_t launch_gauge(const GaugeArgs& a, const GaugeList& list, hipStream_t st)
{
    if (list.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gauge, dim3(list.count), dim3(256), 0, st, a, list);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Boundary of each complex plane after pass 4.  One workgroup per (field, mode).
__global__ __launch_bounds__(256) void k_boundary(BoundaryArgs a)
{
    const FieldGroup& g = a.g;
    const int fm = blockIdx.x;
    const int f = fm / g.nmodes, mode = fm % g.nmodes;
    const int m = g.m;
    const size_t mm = (size_t)m * m;
    d2v* out = g.out[f] + (size_t)mode * mm;
    d2v* in = const_cast<d2v*>(g.in[f]) + (size_t)mode * mm;
    const int tid = threadIdx.x;

    if (a.mode == BND_DIR_ZERO || a.mode == BND_DIR_KEEP) {
        const bool zero = a.mode == BND_DIR_ZERO;
        for (int e = tid; e < 4 * m; e += 256) {
            const int side = e / m, t = e % m;
            size_t idx;
            if (side == 0) idx = (size_t)t;                         // z = 1 column
            else if (side == 1) idx = (size_t)t + (size_t)(m - 1) * m; // z = m
            else if (side == 2) idx = (size_t)t * m;                // y = 1 row
            else idx = (size_t)(m - 1) + (size_t)t * m;             // y = m row
            d2v v = {0.0, 0.0};
            if (!zero) v = in[idx];
            if (zero) in[idx] = v;
            if (g.mode0 + mode == 0) v[1] = 0.0;   // int2d3d.f90:43: Im not solved at k=0
            out[idx] = v;
        }
        return;
    }

    // Neumann: U(i,1), U(i,m) from the two DC rows along z, U(1,j), U(m,j) along y.
    __shared__ d2v eb[kMaxM], et[kMaxM], el[kMaxM], er[kMaxM];
    const double* DC = a.DC;
    const double DC00 = DC[0], DCmm = DC[(m - 1) + (size_t)(m - 1) * m];
    const double DC0m = DC[(size_t)(m - 1) * m], DCm0 = DC[m - 1];
    const double det = DC00 * DCmm - DC0m * DCm0;

    // bottom/top lines (z = 1, z = m): sums along z, one thread per y
    for (int i = 1 + tid; i < m - 1; i += 256) {
        d2v sb = {0.0, 0.0}, stp = {0.0, 0.0};
        for (int j = 1; j < m - 1; ++j) {
            const d2v u = out[(size_t)i + (size_t)j * m];
            sb += DC[(size_t)j * m] * u;
            stp += DC[(m - 1) + (size_t)j * m] * u;
        }
        const d2v vb = (-DCmm * sb + DC0m * stp) / det;
        const d2v vt = (DCm0 * sb - DC00 * stp) / det;
        eb[i] = vb;
        et[i] = vt;
        out[(size_t)i] = vb;
        out[(size_t)i + (size_t)(m - 1) * m] = vt;
    }
    // left/right lines (y = 1, y = m): sums along y, one wave per z column
    const int lane = tid & 63, wave = tid >> 6;
    for (int j = 1 + wave; j < m - 1; j += 4) {
        d2v sl = {0.0, 0.0}, sr = {0.0, 0.0};
        for (int i = 1 + lane; i < m - 1; i += 64) {
            const d2v u = out[(size_t)i + (size_t)j * m];
            sl += DC[(size_t)i * m] * u;
            sr += DC[(m - 1) + (size_t)i * m] * u;
        }
        sl = wave_sum2(sl);
        sr = wave_sum2(sr);
        if (lane == 0) {
            const d2v vl = (-DCmm * sl + DC0m * sr) / det;
            const d2v vr = (DCm0 * sl - DC00 * sr) / det;
            el[j] = vl;
            er[j] = vr;
            out[(size_t)j * m] = vl;
            out[(m - 1) + (size_t)j * m] = vr;
        }
    }
    __syncthreads();
    // corners from the four summed Neumann equations: wave 0
    if (wave == 0) {
        d2v f1 = {0, 0}, f2 = {0, 0}, f3 = {0, 0}, f4 = {0, 0};
        for (int i = 1 + lane; i < m - 1; i += 64) {
            const double d0 = DC[(size_t)i * m], dm = DC[(m - 1) + (size_t)i * m];
            f1 -= d0 * (eb[i] + el[i]);
            f2 += -d0 * er[i] + dm * eb[i];
            f3 -= dm * (et[i] + er[i]);
            f4 += -d0 * et[i] + dm * el[i];
        }
        f1 = wave_sum2(f1);
        f2 = wave_sum2(f2);
        f3 = wave_sum2(f3);
        f4 = wave_sum2(f4);
        if (lane < 4) {
            const double* A = a.AM1;  // 4x4 column-major, row `lane`
            const d2v v = A[lane] * f1 + A[lane + 4] * f2 + A[lane + 8] * f3 + A[lane + 12] * f4;
            size_t idx;
            if (lane == 0) idx = 0;                                        // U(1,1)
            else if (lane == 1) idx = (size_t)(m - 1);                     // U(m,1)
            else if (lane == 2) idx = (size_t)(m - 1) + (size_t)(m - 1) * m; // U(m,m)
            else idx = (size_t)(m - 1) * m;                                // U(1,m)
            out[idx] = v;
        }
    }
}

hipError_t launch_boundary(const BoundaryArgs& a, hipStream_t st)
{
    const int planes = a.g.nf * a.g.nmodes;
    if (planes <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_boundary, dim3(planes), dim3(256), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// F(i,j) = U(i,j) - [D2(i,1) U(1,j) + D2(i,m) U(m,j)] - [D2(j,1) U(i,1) + D2(j,m) U(i,m)]
// on the interior, in place (the boundary values it reads are not written).
__global__ __launch_bounds__(256) void k_bc_correct(BcCorrectArgs a)
{
    const FieldGroup& g = a.g;
    const int fm = blockIdx.y;
    const int f = fm / g.nmodes, mode = fm % g.nmodes;
    const int m = g.m, n = g.n;
    d2v* U = const_cast<d2v*>(g.in[f]) + (size_t)mode * m * m;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)n * n) return;
    const int i = 1 + (int)(e % n), j = 1 + (int)(e / n);
    const double* D2 = a.D2b;
    const d2v bcontv = D2[i] * U[(size_t)j * m] + D2[i + m] * U[(m - 1) + (size_t)j * m];
    const d2v bconth = D2[j] * U[(size_t)i] + D2[j + m] * U[(size_t)i + (size_t)(m - 1) * m];
    const size_t idx = (size_t)i + (size_t)j * m;
    U[idx] = U[idx] - bcontv - bconth;
}

hipError_t launch_bc_correct(const BcCorrectArgs& a, hipStream_t st)
{
    const int planes = a.g.nf * a.g.nmodes;
    const long elems = (long)a.g.n * a.g.n;
    if (planes <= 0 || elems <= 0) return hipSuccess;
    dim3 grid((unsigned)((elems + 255) / 256), planes);
    hipLaunchKernelGGL(k_bc_correct, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace duct
