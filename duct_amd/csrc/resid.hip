// FP64 collocation-operator products on complex(8) fields, v_mfma_f64_16x16x4f64.
//
//   k_colloc<CL_RESID>   r = f - (D2B_y u + u D2B_z^T - xlambda u)   interior (n x n)
//   launch_colloc(CL_DY / CL_DZ): out = beta out + scal (D u | u D^T) on full planes (m x m),
//                        served by the slab kernel k_cline of deriv.hip
//
// The residual of the mixed solve uses D2B, the interior operator the eigenbasis
// diagonalises (pseudomat.f90:232-240 Dirichlet, :114-128 Neumann), so r is the
// residual of exactly the system solvediag solves.  The derivative products
// apply the collocation derivative matrix DC (pseudomat.f90:90-107, CHQZ
// (2.4.31), cheb_collocation.F:19-75) or its transpose: d/dy and d/dz of the
// nodal interpolant, which is what derivyz/transyz compute in coefficient
// space (chebyshev.f90:1-295, 458-540).
//
// Every product accumulates into a register tile whose MFMA row index is z and
// column index is y, so lanes run along the contiguous y of the field:
//   z term:  sum_k D[z][k] u(y, k)    A = matrix fragments,  B = u tile [y][k]
//   y term:  sum_k u(k, z) D[y][k]    A = u tile [z][k],     B = matrix fragments
// The matrix's fragment array (build_frag layout) is the A operand of one and
// the B operand of the other.  A 256-thread workgroup owns a 64 z x 64 y block
// of one complex plane (re and im); the u slabs it streams are staged through
// LDS, double buffered.  Workgroups are numbered so that all blocks of a plane
// land on the same XCD (and its L2).
#include "common.h"
#include "proj.h"
#include "proj_lp.h"

namespace duct {

namespace {
constexpr int RB = 64;          // block edge (z and y)
constexpr int RKT = 16;         // k per LDS stage
constexpr int RLD = RKT + 1;    // LDS stride (doubles)
constexpr int kXcd = 8;         // XCDs per MI355X

__device__ __forceinline__ void atomic_max_pos(unsigned long long* p, double v)
{
    // non-negative doubles order like their bit patterns
    atomicMax(p, (unsigned long long)__builtin_bit_cast(unsigned long long, v));
}
}  // namespace

template <int MODE>
__global__ __launch_bounds__(256, 2) void k_colloc(CollocArgs a, int bpd, long nplanes)
{
    constexpr bool RESID = MODE == CL_RESID;
    constexpr bool YT = MODE != CL_DZ;   // y term present
    constexpr bool ZT = MODE != CL_DY;   // z term present
    const FieldGroup& g = a.g;
    const int NP = a.NP, m = g.m;
    const int off = RESID ? 1 : 0;            // interior or full plane
    const int sz = RESID ? g.n : m;
    // XCD-aware numbering: 8 consecutive planes share a group of 8*bpd^2 ids,
    // plane = group*8 + id%8, block = (id/8) % bpd^2.
    const long bid = blockIdx.x;
    const int tpp = bpd * bpd;
    const long grp = bid / ((long)kXcd * tpp);
    const long plane = grp * kXcd + bid % kXcd;
    const int blk = (int)((bid / kXcd) % tpp);
    if (plane >= nplanes) return;
    const int f = (int)(plane / g.nmodes), mode = (int)(plane % g.nmodes);
    const int z0 = (blk / bpd) * RB, y0 = (blk % bpd) * RB;
    // RESID: g.out = current solution u, g.in = right-hand side f.
    // DY/DZ: g.in = operand, g.out = result.
    const d2v* U = (RESID ? (const d2v*)g.out[f] : g.in[f]) + (size_t)mode * m * m;

    __shared__ double Uz[ZT ? 2 : 1][2][ZT ? RB * RLD : 1];   // [buf][re/im][y][k]: u(y0+y, k0+k)
    __shared__ double Uy[YT ? 2 : 1][2][YT ? RB * RLD : 1];   // [buf][re/im][z][k]: u(k0+k, z0+z)

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int wz = wave & 1, wy = wave >> 1;           // wave tile: 32 z x 32 y

    // loaders: 4 complex values of each slab per thread per stage
    //   Uz: thread -> k = tid/16, y = (tid%16)*4 .. +3   (y contiguous in memory)
    //   Uy: thread -> z = tid/4,  k = (tid%4)*4 .. +3    (k = y' contiguous)
    const int uz_k = tid >> 4, uz_y = (tid & 15) * 4;
    const int uy_z = tid >> 2, uy_k = (tid & 3) * 4;
    d2v rz[4], ry[4];
    auto load = [&](int ks) {
        const int k0 = ks * RKT;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if constexpr (ZT) {
                const int y = y0 + uz_y + e, kz = k0 + uz_k;
                rz[e] = (y < sz && kz < sz) ? U[(size_t)(y + off) + (size_t)(kz + off) * m] : d2v{0.0, 0.0};
            }
            if constexpr (YT) {
                const int z = z0 + uy_z, ky = k0 + uy_k + e;
                ry[e] = (z < sz && ky < sz) ? U[(size_t)(ky + off) + (size_t)(z + off) * m] : d2v{0.0, 0.0};
            }
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if constexpr (ZT) {
                Uz[buf][0][(uz_y + e) * RLD + uz_k] = rz[e][0];
                Uz[buf][1][(uz_y + e) * RLD + uz_k] = rz[e][1];
            }
            if constexpr (YT) {
                Uy[buf][0][uy_z * RLD + uy_k + e] = ry[e][0];
                Uy[buf][1][uy_z * RLD + uy_k + e] = ry[e][1];
            }
        }
    };

    // matrix fragments: row tile t (16 rows), k-pair kp: 64 lanes x 2 doubles
    const d2v* Df = (const d2v*)a.Dfrag;
    auto dfrag = [&](int rowtile, int kp) -> d2v {
        if (rowtile * 16 >= NP) return d2v{0.0, 0.0};
        return Df[((size_t)rowtile * (NP / 8) + kp) * 64 + lane];
    };

    d4v acc[2][2][2];   // [zt][yt][re/im]
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int p = 0; p < 2; ++p) acc[i][j][p] = d4v{0.0, 0.0, 0.0, 0.0};

    const int zt0 = (z0 + wz * 32) / 16, yt0 = (y0 + wy * 32) / 16;   // global 16-row tiles
    const int NK = (sz + RKT - 1) / RKT;
    load(0);
    store(0);
    __syncthreads();
    for (int ks = 0; ks < NK; ++ks) {
        const int buf = ks & 1;
        if (ks + 1 < NK) load(ks + 1);
        d2v az[2][2], by[2][2];   // matrix fragments for the z rows (A) and y rows (B), 2 k-pairs
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if constexpr (ZT) az[t][h] = dfrag(zt0 + t, ks * 2 + h);
                if constexpr (YT) by[t][h] = dfrag(yt0 + t, ks * 2 + h);
            }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int kl = kk * 4 + l4;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                double bu[2], au[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    if constexpr (ZT) bu[t] = Uz[buf][p][(wy * 32 + t * 16 + l16) * RLD + kl];   // B[k][y] = u(y,k)
                    if constexpr (YT) au[t] = Uy[buf][p][(wz * 32 + t * 16 + l16) * RLD + kl];   // A[z][k] = u(k,z)
                }
#pragma unroll
                for (int zt = 0; zt < 2; ++zt)
#pragma unroll
                    for (int yt = 0; yt < 2; ++yt) {
                        if constexpr (ZT)
                            acc[zt][yt][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                az[zt][kk >> 1][kk & 1], bu[yt], acc[zt][yt][p], 0, 0, 0);
                        if constexpr (YT)
                            acc[zt][yt][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                au[zt], by[yt][kk >> 1][kk & 1], acc[zt][yt][p], 0, 0, 0);
                    }
            }
        }
        if (ks + 1 < NK) store(buf ^ 1);
        __syncthreads();
    }

    // epilogue: D[i=z][j=y] in lane j + 16 (i%4), element i/4
    if constexpr (RESID) {
        const d2v* F = g.in[f] + (size_t)mode * m * m;
        const double xlam = xlambda_of(g.alambda[f], g.alp, g.mode0 + mode);
        const bool zim = (g.mode0 + mode) == 0;
        const bool singular = a.neumann && fabs(xlam) <= 1.0e-10;
        d2v* R = a.r[f] + (size_t)mode * m * m;
        double rmx = 0.0, fmx = 0.0;
#pragma unroll
        for (int zt = 0; zt < 2; ++zt)
#pragma unroll
            for (int yt = 0; yt < 2; ++yt)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int z = z0 + wz * 32 + zt * 16 + 4 * e + l4;
                    const int y = y0 + wy * 32 + yt * 16 + l16;
                    if (z < sz && y < sz) {
                        const size_t idx = (size_t)(y + 1) + (size_t)(z + 1) * m;
                        const d2v u = U[idx], fv = F[idx];
                        d2v lu = {acc[zt][yt][0][e], acc[zt][yt][1][e]};
                        d2v rv = fv - (lu - xlam * u);
                        if (zim) rv[1] = 0.0;
                        R[idx] = rv;
                        if (!singular) rmx = fmax(rmx, fmax(fabs(rv[0]), fabs(rv[1])));
                        fmx = fmax(fmx, fmax(fabs(fv[0]), zim ? 0.0 : fabs(fv[1])));
                    }
                }
        if (a.stats) {
#pragma unroll
            for (int sh = 32; sh > 0; sh >>= 1) {
                rmx = fmax(rmx, __shfl_xor(rmx, sh, 64));
                fmx = fmax(fmx, __shfl_xor(fmx, sh, 64));
            }
            if (lane == 0) {
                atomic_max_pos(a.stats + 2 * a.slot[f], rmx);
                atomic_max_pos(a.stats + 2 * a.slot[f] + 1, fmx);
            }
        }
    } else {
        d2v* O = g.out[f] + (size_t)mode * m * m;
#pragma unroll
        for (int zt = 0; zt < 2; ++zt)
#pragma unroll
            for (int yt = 0; yt < 2; ++yt)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int z = z0 + wz * 32 + zt * 16 + 4 * e + l4;
                    const int y = y0 + wy * 32 + yt * 16 + l16;
                    if (z < sz && y < sz) {
                        const size_t idx = (size_t)y + (size_t)z * m;
                        d2v v = {a.scal * acc[zt][yt][0][e], a.scal * acc[zt][yt][1][e]};
                        if (a.beta != 0.0) v += a.beta * O[idx];
                        O[idx] = v;
                    }
                }
    }
}

hipError_t launch_colloc(int mode, const CollocArgs& a, hipStream_t st)
{
    // the single-direction products on whole planes take the slab kernel (deriv.hip)
    if (mode == CL_DY || mode == CL_DZ) return launch_cline(mode == CL_DZ, a, st);
    const int sz = mode == CL_RESID ? a.g.n : a.g.m;
    const int bpd = (sz + RB - 1) / RB;
    const long nplanes = (long)a.g.nf * a.g.nmodes;
    const long groups = (nplanes + kXcd - 1) / kXcd;
    const long blocks = groups * kXcd * bpd * bpd;
    if (blocks <= 0) return hipSuccess;
    switch (mode) {
        case CL_RESID: hipLaunchKernelGGL(k_colloc<CL_RESID>, dim3((unsigned)blocks), dim3(256), 0, st, a, bpd, nplanes); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_resid(const ResidArgs& a, hipStream_t st) { return launch_colloc(CL_RESID, a, st); }

// out = beta out + scal * (i alp k) in, k = mode0 + mode  (derivx, fourier.F:117-188, ider=1)
__global__ __launch_bounds__(256) void k_derivx(const d2v* __restrict__ in, d2v* out, long mm, int nmodes, int mode0,
                                                double alp, double scal, double beta)
{
    const long total = mm * nmodes;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const int mode = (int)(e / mm);
        const double xa = alp * (double)(mode0 + mode);
        const d2v u = in[e];
        d2v v = {-scal * xa * u[1], scal * xa * u[0]};
        if (beta != 0.0) v += beta * out[e];
        out[e] = v;
    }
}

hipError_t launch_derivx(const d2v* in, d2v* out, int m, int nmodes, int mode0, double alp, double scal, double beta,
                         hipStream_t st)
{
    const long total = (long)m * m * nmodes;
    if (total <= 0) return hipSuccess;
    long blocks = (total + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_derivx, dim3((unsigned)blocks), dim3(256), 0, st, in, out, (long)m * m, nmodes, mode0, alp,
                       scal, beta);
    return hipGetLastError();
}

}  // namespace duct
