// Folded (even/odd) mixed-precision refined solve.
//
// The refinement of the reference system L u = f (solvediag.f90; L = D2B_y +
// D2B_z - xlambda) runs in the parity-folded basis of proj_fold.hip:
//   F  = fold(f)                       FP32 quarters                 k_fold32
//   U0 = S~(F)                         fused low-precision solve      k_lpq
//   R  = f - S U0                      FP64 residual, folded          k_resid_fold
//   D1 = S~(R)                         fused low-precision solve      k_lpq
//   u  = unfold(U0 + D1)               FP64                           k_unfold_sum
// S is the centrosymmetric part (D2B + J D2B J)/2 of the interior operator: its
// folded blocks are exactly block-diagonal, so the residual costs two n^3/2
// products per plane.  D2B itself is centrosymmetric up to rounding (~2e-15 of
// its norm); the converged solutions of S and of D2B lie within 2e-11 of each
// other and both within 2e-11 of the FP64 oracle on the Poisson planes (the
// rounding level of those ill-conditioned planes), so the refinement targets S.
// U0 and D1 are FP32: an FP32-rounded U0 is an exact FP64 iterate, and D1 is a
// correction whose own relative accuracy only needs to match the solve's.
//
// Plane order in every folded array: p = 2 (mode nf + f) + part, so the first K
// modes of all fields are a prefix (a second refinement of the low Poisson
// modes reuses the same arrays).
#include <cmath>
#include <cstring>

#include "mixed_fold.h"

namespace duct {

constexpr int kSplitExp = 12;   // the fp16 hi/lo split works on values scaled to [2^11, 2^12)

typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef _Float16 h4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

// Maximum over the 64 lanes of a wave of non-negative values: DPP row moves (quad_perm [1,0,3,2],
// [2,3,0,1], row_half_mirror, row_mirror) leave each row's maximum in all its lanes, the four
// rows are read with readlane.  A chain of __shfl_xor would take six LDS-crossbar round trips.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF,
                                                                 false));
}

__device__ __forceinline__ float wave_max_nonneg(float v)
{
    v = fmaxf(v, dpp_f<0xB1>(v));
    v = fmaxf(v, dpp_f<0x4E>(v));
    v = fmaxf(v, dpp_f<0x141>(v));
    v = fmaxf(v, dpp_f<0x140>(v));
    const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0));
    const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16));
    const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
    const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48));
    return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

__device__ __forceinline__ void plane_decode(const FieldGroup& g, long fmp, int& f, int& mode)
{
    mode = (int)(fmp / g.nf);
    f = (int)(fmp % g.nf);
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fold32(FieldGroup g, int NH, float* X, float* Xlo)
{
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    const long per = (long)NH * NH;
    if (idx >= (long)g.nf * g.nmodes * per) return;
    const long fmp = idx / per;
    const int zq = (int)((idx % per) / NH), yq = (int)(idx % NH);
    int f, mode;
    plane_decode(g, fmp, f, mode);
    const int n = g.n, m = g.m, he = (n + 1) / 2;
    d2v a = {0.0, 0.0}, b = a, c = a, d = a;
    if (yq < he && zq < he) {
        const d2v* base = g.in[f] + (size_t)mode * m * m;
        const int y0 = yq + 1, y1 = n - yq, z0 = zq + 1, z1 = n - zq;
        a = base[y0 + (size_t)z0 * m];
        b = base[y1 + (size_t)z0 * m];
        c = base[y0 + (size_t)z1 * m];
        d = base[y1 + (size_t)z1 * m];
    }
    const d2v s0 = a + b, d0 = a - b, s1 = c + d, d1 = c - d;
    const d2v q[4] = {0.25 * (s0 + s1), 0.25 * (s0 - s1), 0.25 * (d0 + d1), 0.25 * (d0 - d1)};
    const bool zim = (g.mode0 + mode) == 0;
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        const long p = 2 * fmp + part;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const size_t o = ((size_t)(p * 4 + k) * NH + zq) * NH + yq;
            const double v = (part && zim) ? 0.0 : q[k][part];
            const float hi = (float)v;
            X[o] = hi;
            if (Xlo) Xlo[o] = (float)(v - (double)hi);
        }
    }
}

hipError_t launch_fold32(const FieldGroup& g, int NH, float* X, hipStream_t st, float* Xlo)
{
    const long total = (long)g.nf * g.nmodes * NH * NH;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fold32, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, g, NH, X, Xlo);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
struct UnfoldArgs {
    FieldGroup g;
    const float* src[kMaxFoldSources];
    int nsrc, NH, accumulate;
    const unsigned char* udig;   // DIG: the first term in launch_lpq's digit form (Ucol, gexp)
    const int* gexp;
    int zero_edges;              // also write 0 to the boundary of sol and rhs (Dirichlet, helm3d.f90:48-60)
};

template <bool DIG>
__global__ __launch_bounds__(256) void k_unfold_sum(UnfoldArgs a)
{
    const FieldGroup& g = a.g;
    const int NH = a.NH, n = g.n, m = g.m, he = (n + 1) / 2;
    const long idx = (long)blockIdx.x * 256 + threadIdx.x;
    const long per = (long)he * he;
    if (idx >= (long)g.nf * g.nmodes * per) return;
    const long fmp = idx / per;
    const int zq = (int)((idx % per) / he), yq = (int)(idx % he);
    int f, mode;
    plane_decode(g, fmp, f, mode);
    const size_t NN = (size_t)NH * NH;
    d2v V[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        V[k] = d2v{0.0, 0.0};
#pragma unroll
        for (int part = 0; part < 2; ++part) {
            const size_t qi = (size_t)(2 * fmp + part) * 4 + k;
            const size_t o = qi * NN + (size_t)zq * NH + yq;
            double s = 0.0;
            // every input is read once and the output is not read back here: nontemporal
            // accesses keep them from displacing cached data (~8% of this kernel's time)
            if (DIG) {   // U = 2^(g - 8 kUDig) (c1 256^(kUDig-1) + ... + c_kUDig), exact
                const unsigned char* ud = a.udig + qi * kUDig * NN + (size_t)zq * NH + yq;
                int Z = 0;
#pragma unroll
                for (int d = 0; d < kUDig; ++d) Z = Z * 256 + (int)(signed char)__builtin_nontemporal_load(ud + d * NN);
                s = ldexp((double)Z, a.gexp[qi] - 8 * kUDig);
            }
            for (int t = 0; t < a.nsrc; ++t) s += (double)__builtin_nontemporal_load(&a.src[t][o]);
            V[k][part] = s;
        }
    }
    const bool zim = (g.mode0 + mode) == 0;
    // V[2 r + c]: r = parity of y, c = parity of z; the mirrored y (z) flips the odd-r (odd-c) terms
    const d2v e0 = V[0] + V[1], e1 = V[0] - V[1], o0 = V[2] + V[3], o1 = V[2] - V[3];
    const int y0 = yq + 1, y1 = n - yq, z0 = zq + 1, z1 = n - zq;
    d2v* out = g.out[f] + (size_t)mode * m * m;
    auto put = [&](int y, int z, d2v v) {
        if (zim) v[1] = 0.0;
        d2v* p = out + y + (size_t)z * m;
        if (a.accumulate) v += *p;
        __builtin_nontemporal_store(v, p);
    };
    put(y0, z0, e0 + o0);
    if (y1 != y0) put(y1, z0, e0 - o0);
    if (z1 != z0) {
        put(y0, z1, e1 + o1);
        if (y1 != y0) put(y1, z1, e1 - o1);
    }
    if (a.zero_edges) {
        // helm3d.f90:48-60: the boundary of rhs and sol set to zero; each thread takes the walls
        // next to its points (y0, y1 cover 1..n, as z0, z1), the first thread the corners
        d2v* in = const_cast<d2v*>(g.in[f]) + (size_t)mode * m * m;
        auto zero = [&](size_t i) {
            out[i] = d2v{0.0, 0.0};
            in[i] = d2v{0.0, 0.0};
        };
        if (yq == 0) {
            zero((size_t)z0 * m);
            zero((m - 1) + (size_t)z0 * m);
            zero((size_t)z1 * m);
            zero((m - 1) + (size_t)z1 * m);
        }
        if (zq == 0) {
            zero(y0);
            zero(y1);
            zero(y0 + (size_t)(m - 1) * m);
            zero(y1 + (size_t)(m - 1) * m);
        }
        if (yq == 0 && zq == 0) {
            zero(0);
            zero(m - 1);
            zero((size_t)(m - 1) * m);
            zero((m - 1) + (size_t)(m - 1) * m);
        }
    }
}

hipError_t launch_unfold_sum(const FieldGroup& g, int NH, const float* const* src, int nsrc, int accumulate,
                             hipStream_t st, const unsigned char* udig, const int* gexp, int zero_edges)
{
    if (nsrc < 0 || nsrc > kMaxFoldSources || (nsrc == 0 && !udig)) return hipErrorInvalidValue;
    if (udig && !gexp) return hipErrorInvalidValue;
    UnfoldArgs a{};
    a.g = g;
    for (int t = 0; t < nsrc; ++t) a.src[t] = src[t];
    a.nsrc = nsrc;
    a.NH = NH;
    a.accumulate = accumulate;
    a.udig = udig;
    a.gexp = gexp;
    a.zero_edges = zero_edges;
    const long he = (g.n + 1) / 2;
    const long total = (long)g.nf * g.nmodes * he * he;
    if (total <= 0) return hipSuccess;
    const dim3 grid((unsigned)((total + 255) / 256));
    if (udig) hipLaunchKernelGGL(k_unfold_sum<true>, grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_unfold_sum<false>, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fused low-precision solve of one quarter per workgroup.
//
// Each pass computes Y = (B X)^T for the pass's basis B with the data X as the
// MFMA A operand (A[j][k] = X[k][j]) and B as the B operand (Bop[k][i] = B[i][k]),
// so D[j][i] = Y[j][i] lands with four consecutive j per lane: stored column by
// column (column i, rows j) it is the next pass's X without a transpose.  X lives
// in LDS as two fp16 terms, hi = fp16(x s) and lo = fp16(x s - hi), with s a power
// of two per pass that puts max |x s| in [2^11, 2^12): hi then carries 11 bits, lo
// the next 11 or more, and a lo below fp16's normal range is below 2^-25 of the
// maximum.  The products hi*hi, hi*lo and lo*hi go to one FP32 accumulator (the
// bases are split the same way).  Lane layout of v_mfma_f32_16x16x32_f16:
// A[i][k] lane i + 16 (k/8) element k%8, B[k][j] lane j + 16 (k/8) element k%8,
// D[i][j] lane j + 16 (i/4) element i%4.
// Passes: 0 B = P_r^-1, 1 B = P_c^-1 then ./ (lam_r + lam_c - xlambda),
// 2 B = P_r, 3 B = P_c -> U(y', z').
// Work split: NW waves in a 2 x NW/2 grid over the quarter (j x i).  With NW = 4 a wave owns 64 x 64
// and two workgroups share a CU, so one workgroup's epilogue (VALU, LDS writes, barriers) runs while
// the other's products keep the matrix cores busy; the B fragments stream one k chunk ahead.
template <int NH, int NW>
__global__ __launch_bounds__(64 * NW, 8 / NW) void k_lpq(LpqArgs a)
{
    constexpr int NTH = 64 * NW;
    constexpr int LDX = NH + 8;          // halves per LDS column
    constexpr int WI = NW / 2;           // waves along i (2 along j)
    constexpr int TJ = NH / 32;          // MFMA tiles per wave along j
    constexpr int TI = NH / (16 * WI);   // MFMA tiles per wave along i
    constexpr int KC = NH / 32;          // k chunks per pass
    constexpr int NT = NH / 16;          // 16-row tiles of a basis
    constexpr int PER = NH * NH / (4 * NTH);   // float4 loads per thread
    __shared__ __attribute__((aligned(16))) _Float16 Xs[2][NH * LDX];
    __shared__ float red[NW];
    __shared__ float lam_s[2 * NH];      // [lam_e | lam_o] for the divide of pass 1

    const FieldGroup& g = a.g;
    const long plane = blockIdx.x >> 2;
    const int q = blockIdx.x & 3, r = q >> 1, c = q & 1;
    const long fmp = plane >> 1;
    const int part = (int)(plane & 1);
    int f, mode;
    plane_decode(g, fmp, f, mode);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int n = g.n;
    const int hh[2] = {(n + 1) / 2, n / 2};
    const size_t qoff = ((size_t)plane * 4 + q) * NH * NH;
    const float* X = a.X + qoff;
    float* Y = a.Y + qoff;

    if (part == 1 && g.mode0 + mode == 0) {   // Im of mode 0 is not solved (int2d3d.f90:43)
        if (a.Ucol) {
            for (int e = tid; e < kUDig * NH * NH / 16; e += NTH) {
                ((uint4*)(a.Ucol + qoff * kUDig))[e] = uint4{0u, 0u, 0u, 0u};
                ((uint4*)(a.Urow + qoff * kUDig))[e] = uint4{0u, 0u, 0u, 0u};
            }
            if (tid == 0) a.gexp[plane * 4 + q] = 0;
            return;
        }
        for (int e = tid; e < NH * NH / 4; e += NTH) ((f4v*)Y)[e] = f4v{0.f, 0.f, 0.f, 0.f};
        return;
    }
    const double xl64 = xlambda_of(g.alambda[f], g.alp, g.mode0 + mode);
    const float xlam = (float)xl64;

    auto block_max = [&](float v) -> float {   // v >= 0
        v = wave_max_nonneg(v);
        if (lane == 0) red[wave] = v;
        __syncthreads();
        float mx = red[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) mx = fmaxf(mx, red[w]);
        return mx;
    };
    auto scale_of = [](float mx) -> float {   // power of two s with mx * s in [2^11, 2^12)
        if (!(mx > 0.f) || !(mx < 3.0e38f)) return 1.f;
        int e;
        (void)frexpf(mx, &e);
        return ldexpf(1.f, kSplitExp - e);
    };
    auto put4 = [&](int col, int row, f4v v, float s) {   // 4 consecutive rows of a column
        h4v hi, lo;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float x = v[e] * s;
            const _Float16 h = (_Float16)x;
            hi[e] = h;
            lo[e] = (_Float16)(x - (float)h);
        }
        *(h4v*)&Xs[0][col * LDX + row] = hi;
        *(h4v*)&Xs[1][col * LDX + row] = lo;
    };

    // ---- load X (column z', rows y'), scale and split
    for (int e = tid; e < 2 * NH; e += NTH) lam_s[e] = a.lam[e];   // visible after the barriers below
    float sx;
    {
        f4v xv[PER];
        float mx = 0.f;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            xv[u] = ((const f4v*)X)[tid + NTH * u];
#pragma unroll
            for (int e = 0; e < 4; ++e) mx = fmaxf(mx, fabsf(xv[u][e]));
        }
        sx = scale_of(block_max(mx));
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int idx = (tid + NTH * u) * 4;
            put4(idx / NH, idx % NH, xv[u], sx);
        }
        __syncthreads();
    }

    const int wj = wave & 1, wi = wave >> 1;   // 2 x WI waves
    const int j0 = wj * (NH / 2), i0 = wi * (NH / WI);
    // B fragments of one k chunk (2 terms x TI tiles), the next chunk's loaded while this one's
    // products run; the next pass's first chunk is issued before this pass's epilogue
    h8v bq[2][2][TI];
    auto load_b = [&](int pass, int kc, h8v (&dst)[2][TI]) {
        const int par = (pass == 0 || pass == 2) ? r : c;
        const h8v* Bf = (const h8v*)(pass < 2 ? a.fwd[par] : a.bwd[par]);
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int y = 0; y < TI; ++y) dst[t][y] = Bf[((size_t)(t * NT + i0 / 16 + y) * KC + kc) * 64 + lane];
    };
    load_b(0, 0, bq[0]);
    f4v am[TJ][TI];
    float inv = 1.f, mx = 0.f;
    // passes 0-2 end in the split into LDS; pass 3's epilogue (the stores) follows the loop, so its
    // addresses are not loop invariants held through all four passes
    for (int pass = 0; pass < 4; ++pass) {
        const int par = (pass == 0 || pass == 2) ? r : c;
        const float sb = pass < 2 ? a.sfwd[par] : a.sbwd[par];
#pragma unroll
        for (int x = 0; x < TJ; ++x)
#pragma unroll
            for (int y = 0; y < TI; ++y) am[x][y] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
            h8v (&bc)[2][TI] = bq[kc & 1];
            // the next chunk, unconditionally (after pass 3's last chunk a harmless reload of pass
            // 0's first): a load under a branch would make every outstanding load be waited for
            load_b(kc + 1 < KC ? pass : (pass + 1) & 3, (kc + 1) % KC, bq[(kc + 1) & 1]);
            // keep the chunk's products below these loads and above the next chunk's, so that the
            // loads have a whole chunk of products to land behind
            __builtin_amdgcn_sched_barrier(0);
            // one j tile's A fragments at a time (the registers go to the accumulator); the three
            // products into am[x][y] are TI launches apart, not back to back
#pragma unroll
            for (int x = 0; x < TJ; ++x) {
                const int j = j0 + x * 16 + l16;
                const int k = kc * 32 + l4 * 8;
                const h8v ah = *(const h8v*)&Xs[0][j * LDX + k];
                const h8v al = *(const h8v*)&Xs[1][j * LDX + k];
#pragma unroll
                for (int y = 0; y < TI; ++y)
                    am[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bc[0][y], am[x][y], 0, 0, 0);
#pragma unroll
                for (int y = 0; y < TI; ++y)
                    am[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bc[1][y], am[x][y], 0, 0, 0);
#pragma unroll
                for (int y = 0; y < TI; ++y)
                    am[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bc[0][y], am[x][y], 0, 0, 0);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        // KC is even: the next pass's first chunk sits in bq[0]
        static_assert(KC % 2 == 0, "the B chunk double buffer assumes an even chunk count");
        // true values: am inv, inv = 1 / (sx sb) a power of two that is applied
        // together with the next pass's scale; D[j][i]: lane i + 16 (j/4), element j%4
        inv = 1.f / (sx * sb);
        mx = 0.f;
        if (pass == 1) {   // G = ... ./ (lam_r(j) + lam_c(i) - xlambda), solvediag.f90:61-66, :151-159
            // one uniform branch for the tile; lam_r - xlambda per row and lam_c per column read
            // once, +inf past the parity's size so that rcp gives 0 on the padding
            float lr[TJ][4], lc[TI];
#pragma unroll
            for (int x = 0; x < TJ; ++x)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int j = j0 + x * 16 + l4 * 4 + e;
                    lr[x][e] = j < hh[r] ? lam_s[r * NH + j] - xlam : __builtin_huge_valf();
                }
#pragma unroll
            for (int y = 0; y < TI; ++y) {
                const int i = i0 + y * 16 + l16;
                lc[y] = i < hh[c] ? lam_s[c * NH + i] : __builtin_huge_valf();
            }
#pragma unroll
            for (int x = 0; x < TJ; ++x)
#pragma unroll
                for (int y = 0; y < TI; ++y)
#pragma unroll
                    for (int e = 0; e < 4; ++e)   // rcp: ~1 ulp, inside the split's 2^-22
                        am[x][y][e] *= __builtin_amdgcn_rcpf(lr[x][e] + lc[y]);
        }
#pragma unroll
        for (int x = 0; x < TJ; ++x)
#pragma unroll
            for (int y = 0; y < TI; ++y)
#pragma unroll
                for (int e = 0; e < 4; ++e) mx = fmaxf(mx, fabsf(am[x][y][e]));
        mx *= inv;
        if (pass == 3) break;
        sx = scale_of(block_max(mx));   // the barrier inside: every wave is done reading Xs
        const float sput = inv * sx;     // a power of two
#pragma unroll
        for (int x = 0; x < TJ; ++x)
#pragma unroll
            for (int y = 0; y < TI; ++y) {
                const int j = j0 + x * 16 + l4 * 4, i = i0 + y * 16 + l16;
                put4(i, j, am[x][y], sput);
            }
        __syncthreads();
    }
    // ---- pass 3's epilogue: U(y' = j, z' = i) -> column i, rows j
#pragma unroll
    for (int x = 0; x < TJ; ++x)
#pragma unroll
        for (int y = 0; y < TI; ++y) am[x][y] *= inv;
    {
            if (!a.Ucol) {
                // through LDS like the digit planes below: columns of NH floats + 4 of padding,
                // stored as whole columns, 16 bytes per lane
                constexpr int SLF = NH + 4;
                float* stf = (float*)&Xs[0][0];
                static_assert((size_t)NH * SLF * sizeof(float) <= sizeof(Xs), "FP32 staging exceeds the LDS tile");
                __syncthreads();   // every wave is done reading Xs
#pragma unroll
                for (int x = 0; x < TJ; ++x)
#pragma unroll
                    for (int y = 0; y < TI; ++y) {
                        const int j = j0 + x * 16 + l4 * 4, i = i0 + y * 16 + l16;
                        *(f4v*)&stf[(size_t)i * SLF + j] = am[x][y];
                    }
                __syncthreads();
                for (int e = tid; e < NH * NH / 4; e += NTH) {
                    const int col = e / (NH / 4), ch = e % (NH / 4);
                    *(f4v*)&Y[(size_t)col * NH + ch * 4] = *(const f4v*)&stf[(size_t)col * SLF + ch * 4];
                }
                return;
            }
            // digit form: Z = rint(U 2^(8 kUDig - g)), |Z| < 0.49 256^kUDig, as signed base-256 digits
            const float gmx = block_max(mx);   // its barrier: every wave is done reading Xs
            int ge = 0;
            if (gmx > 0.f && gmx < 3.0e38f) (void)frexpf(gmx / 0.49f, &ge);
            const float sc = ldexpf(1.f, 8 * kUDig - ge);
            // The digit planes go out through LDS (Xs is free now): each lane's words are written
            // to a staging image of the quarter's three planes (rows of NH bytes, 16 bytes of
            // padding against bank conflicts), then stored as whole rows, 16 bytes per lane, so
            // that every store instruction writes one contiguous kilobyte.
            constexpr int SLD = NH + 16;
            unsigned char* stg = (unsigned char*)&Xs[0][0];
            static_assert((size_t)kUDig * NH * SLD <= sizeof(Xs), "digit staging exceeds the LDS tile");
            auto flush = [&](unsigned char* dst) {   // staging -> dst[d][row][col], coalesced
                __syncthreads();
                constexpr int C16 = NH / 16;
                for (int e = tid; e < kUDig * NH * C16; e += NTH) {
                    const int d = e / (NH * C16), row = (e / C16) % NH, ch = e % C16;
                    *(uint4*)&dst[((size_t)d * NH + row) * NH + ch * 16] =
                        *(const uint4*)&stg[((size_t)d * NH + row) * SLD + ch * 16];
                }
            };
            unsigned w[TJ][TI][kUDig];
#pragma unroll
            for (int x = 0; x < TJ; ++x)
#pragma unroll
                for (int y = 0; y < TI; ++y) {
                    const int j = j0 + x * 16 + l4 * 4, i = i0 + y * 16 + l16;
#pragma unroll
                    for (int d = 0; d < kUDig; ++d) w[x][y][d] = 0u;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        int Z = __float2int_rn(am[x][y][e] * sc);
                        int cd[kUDig];   // cd[0] most significant
#pragma unroll
                        for (int d = kUDig - 1; d >= 1; --d) {
                            cd[d] = ((Z + 128) & 255) - 128;
                            Z = (Z - cd[d]) >> 8;
                        }
                        cd[0] = Z;
#pragma unroll
                        for (int d = 0; d < kUDig; ++d) w[x][y][d] |= (unsigned)(cd[d] & 255) << (8 * e);
                    }
                    // Ucol: column i, rows j .. j + 3
#pragma unroll
                    for (int d = 0; d < kUDig; ++d) *(unsigned*)&stg[((size_t)d * NH + i) * SLD + j] = w[x][y][d];
                }
            flush(a.Ucol + qoff * kUDig);
            __syncthreads();   // the staging image is read out before it is rewritten
            // the 4 x 4 byte transpose within a lane quad (lanes l16 = 4 a .. 4 a + 3 hold columns
            // i .. i + 3 of rows j .. j + 3): quad_perm [1,0,3,2] and a byte select, then [2,3,0,1]
            // and a half-word select leave row j + (l16 % 4), columns i .. i + 3 in each lane
            const int qv = l16 & 3;
            const unsigned selA = (qv & 1) ? 0x03070105u : 0x06020400u;
            const unsigned selB = (qv & 2) ? 0x03020706u : 0x05040100u;
#pragma unroll
            for (int x = 0; x < TJ; ++x)
#pragma unroll
                for (int y = 0; y < TI; ++y) {
                    const int j = j0 + x * 16 + l4 * 4, i = i0 + y * 16 + l16;
#pragma unroll
                    for (int d = 0; d < kUDig; ++d) {
                        const unsigned wd = w[x][y][d];
                        const unsigned wx = (unsigned)__builtin_amdgcn_update_dpp(0, (int)wd, 0xB1, 0xF, 0xF, false);
                        const unsigned t1 = __builtin_amdgcn_perm(wx, wd, selA);
                        const unsigned tx = (unsigned)__builtin_amdgcn_update_dpp(0, (int)t1, 0x4E, 0xF, 0xF, false);
                        const unsigned t2 = __builtin_amdgcn_perm(tx, t1, selB);
                        // Urow: row j + qv, columns i - qv .. i - qv + 3
                        *(unsigned*)&stg[((size_t)d * NH + j + qv) * SLD + (i - qv)] = t2;
                    }
                }
            flush(a.Urow + qoff * kUDig);
            if (tid == 0) a.gexp[plane * 4 + q] = ge;
    }
}

hipError_t launch_lpq(int NH, const LpqArgs& a, hipStream_t st)
{
    const long planes = (long)a.g.nf * a.g.nmodes * 2;
    if (planes <= 0) return hipSuccess;
    const dim3 grid((unsigned)(planes * 4));
    switch (NH) {
        case 64: hipLaunchKernelGGL((k_lpq<64, 4>), grid, dim3(256), 0, st, a); break;
        case 128: hipLaunchKernelGGL((k_lpq<128, 4>), grid, dim3(256), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

float build_lpq_basis(const double* B, int ld, int NH, std::vector<unsigned short>& frag)
{
    // frag[((t NT + it) KC + kc) 64 + lane][e] = term t of B(it 16 + lane%16, kc 32 + (lane/16) 8 + e)
    double mx = 0.0;
    for (int k = 0; k < NH; ++k)
        for (int i = 0; i < NH; ++i) mx = std::fmax(mx, std::fabs(B[i + (size_t)k * ld]));
    int ex = 0;
    if (mx > 0.0) (void)std::frexp(mx, &ex);
    const float s = std::ldexp(1.f, kSplitExp - ex);
    const int NT = NH / 16, KC = NH / 32;
    frag.assign((size_t)2 * NT * KC * 64 * 8, 0);
    for (int it = 0; it < NT; ++it)
        for (int kc = 0; kc < KC; ++kc)
            for (int lane = 0; lane < 64; ++lane)
                for (int e = 0; e < 8; ++e) {
                    const int i = it * 16 + (lane & 15), k = kc * 32 + (lane >> 4) * 8 + e;
                    const float x = (float)B[i + (size_t)k * ld] * s;
                    const _Float16 h = (_Float16)x;
                    const _Float16 l = (_Float16)(x - (float)h);
                    unsigned short hb, lb;
                    std::memcpy(&hb, &h, 2);
                    std::memcpy(&lb, &l, 2);
                    frag[((((size_t)0 * NT + it) * KC + kc) * 64 + lane) * 8 + e] = hb;
                    frag[((((size_t)1 * NT + it) * KC + kc) * 64 + lane) * 8 + e] = lb;
                }
    return s;
}

// ---------------------------------------------------------------------------
// Folded FP64 residual, one 64 x 64 block of one quarter of one complex plane
// (re and im) per workgroup; the k_colloc scheme of resid.hip with the folded
// blocks S_r (y term, B operand) and S_c (z term, A operand), K = NH.
namespace {
constexpr int RB = 64, RKT = 16, RLD = RKT + 1, kXcd = 8;
__device__ __forceinline__ void atomic_max_pos(unsigned long long* p, double v)
{
    atomicMax(p, (unsigned long long)__builtin_bit_cast(unsigned long long, v));
}
}  // namespace

// NSRC, the number of FP32 sources summed into u, is a template parameter so that the loads of
// a stage are all issued before any of them is waited for: they land while the MFMAs of the
// current stage run, and are summed (in FP64, in source order) only as they go to LDS.
template <int NSRC>
__global__ __launch_bounds__(256, 2) void k_resid_fold(ResidFoldArgs a, long ncplx)
{
    const FieldGroup& g = a.g;
    const int NH = a.NH, m = g.m, n = g.n;
    const int bpd = NH / RB, tpp = 4 * bpd * bpd;   // blocks per complex plane
    const long bid = blockIdx.x;
    const long grp = bid / ((long)kXcd * tpp);
    const long fmp = grp * kXcd + bid % kXcd;       // the blocks of one plane share an XCD
    const int blk = (int)((bid / kXcd) % tpp);
    if (fmp >= ncplx) return;
    int f, mode;
    plane_decode(g, fmp, f, mode);
    const int q = blk / (bpd * bpd), r = q >> 1, c = q & 1;
    const int b2 = blk % (bpd * bpd);
    const int z0 = (b2 / bpd) * RB, y0 = (b2 % bpd) * RB;
    const int hh[2] = {(n + 1) / 2, n / 2};

    __shared__ double Uz[2][2][RB * RLD];   // [buf][part][y][k]: u(y0 + y, k0 + k)
    __shared__ double Uy[2][2][RB * RLD];   // [buf][part][z][k]: u(k0 + k, z0 + z)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int wz = wave & 1, wy = wave >> 1;

    size_t qbase[2];
#pragma unroll
    for (int part = 0; part < 2; ++part) qbase[part] = ((size_t)(2 * fmp + part) * 4 + q) * NH * NH;
    // Uz loader: k = tid/16 (z' = k0 + k), y = (tid%16)*4..+3;  Uy loader: z = tid/4, k = (tid%4)*4..+3 (y')
    const int uz_k = tid >> 4, uz_y = (tid & 15) * 4;
    const int uy_z = tid >> 2, uy_k = (tid & 3) * 4;
    f4v vz[NSRC][2], vy[NSRC][2];           // this thread's FP32 source values of the next stage
    auto load = [&](int ks) {
        const int k0 = ks * RKT;
#pragma unroll
        for (int t = 0; t < NSRC; ++t)
#pragma unroll
            for (int part = 0; part < 2; ++part) {
                vz[t][part] = ((const f4v*)(a.src[t] + qbase[part] + (size_t)(k0 + uz_k) * NH + y0 + uz_y))[0];
                vy[t][part] = ((const f4v*)(a.src[t] + qbase[part] + (size_t)(z0 + uy_z) * NH + k0 + uy_k))[0];
            }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int part = 0; part < 2; ++part)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                double sz = 0.0, sy = 0.0;       // sums in FP64 of the FP32 sources
#pragma unroll
                for (int t = 0; t < NSRC; ++t) {
                    sz += (double)vz[t][part][e];
                    sy += (double)vy[t][part][e];
                }
                Uz[buf][part][(uz_y + e) * RLD + uz_k] = sz;
                Uy[buf][part][uy_z * RLD + uy_k + e] = sy;
            }
    };
    const d2v* Sr = (const d2v*)a.Sfrag[r];
    const d2v* Sc = (const d2v*)a.Sfrag[c];
    auto dfrag = [&](const d2v* D, int rowtile, int kp) -> d2v {
        return D[((size_t)rowtile * (NH / 8) + kp) * 64 + lane];
    };

    d4v acc[2][2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int p = 0; p < 2; ++p) acc[i][j][p] = d4v{0.0, 0.0, 0.0, 0.0};
    const int zt0 = (z0 + wz * 32) / 16, yt0 = (y0 + wy * 32) / 16;
    const int NK = NH / RKT;
    d2v az[2][2], by[2][2], azn[2][2], byn[2][2];   // matrix fragments, this and the next stage
    auto load_frags = [&](int ks, d2v (&za)[2][2], d2v (&yb)[2][2]) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                za[t][h] = dfrag(Sc, zt0 + t, ks * 2 + h);
                yb[t][h] = dfrag(Sr, yt0 + t, ks * 2 + h);
            }
    };
    load(0);
    load_frags(0, az, by);
    store(0);
    __syncthreads();
    for (int ks = 0; ks < NK; ++ks) {
        const int buf = ks & 1;
        if (ks + 1 < NK) {
            load(ks + 1);
            load_frags(ks + 1, azn, byn);
        }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int kl = kk * 4 + l4;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                double bu[2], au[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    bu[t] = Uz[buf][p][(wy * 32 + t * 16 + l16) * RLD + kl];   // B[k][y] = u(y, k)
                    au[t] = Uy[buf][p][(wz * 32 + t * 16 + l16) * RLD + kl];   // A[z][k] = u(k, z)
                }
#pragma unroll
                for (int zt = 0; zt < 2; ++zt)
#pragma unroll
                    for (int yt = 0; yt < 2; ++yt) {
                        acc[zt][yt][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(az[zt][kk >> 1][kk & 1], bu[yt],
                                                                              acc[zt][yt][p], 0, 0, 0);
                        acc[zt][yt][p] = __builtin_amdgcn_mfma_f64_16x16x4f64(au[zt], by[yt][kk >> 1][kk & 1],
                                                                              acc[zt][yt][p], 0, 0, 0);
                    }
            }
        }
        if (ks + 1 < NK) {
            store(buf ^ 1);
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    az[t][h] = azn[t][h];
                    by[t][h] = byn[t][h];
                }
        }
        __syncthreads();
    }

    // epilogue: D[i = z][j = y] in lane j + 16 (i%4), element i/4
    const double xlam = xlambda_of(g.alambda[f], g.alp, g.mode0 + mode);
    const bool zim = (g.mode0 + mode) == 0;
    const d2v* F = g.in[f] + (size_t)mode * m * m;
    const double sr = r ? -1.0 : 1.0, sc = c ? -1.0 : 1.0;
    double rmx = 0.0, fmx = 0.0;
    // One 16 x 16 output tile at a time: its f and u operands are all loaded before its first store
    // (R and the sources may be the same allocation for the compiler, which would otherwise keep
    // every load behind the previous store). Points outside the quarter read its first point.
#pragma unroll
    for (int zt = 0; zt < 2; ++zt)
#pragma unroll
        for (int yt = 0; yt < 2; ++yt) {
            d2v fq[4][4];
            float us[4][NSRC][2];
            bool in[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int z = z0 + wz * 32 + zt * 16 + 4 * e + l4;
                const int y = y0 + wy * 32 + yt * 16 + l16;
                in[e] = y < hh[r] && z < hh[c];
                const int yc = in[e] ? y : 0, zc = in[e] ? z : 0;
                const int ya = yc + 1, yb = n - yc, za = zc + 1, zb = n - zc;
                fq[e][0] = F[ya + (size_t)za * m];
                fq[e][1] = F[yb + (size_t)za * m];
                fq[e][2] = F[ya + (size_t)zb * m];
                fq[e][3] = F[yb + (size_t)zb * m];
#pragma unroll
                for (int t = 0; t < NSRC; ++t)
#pragma unroll
                    for (int p = 0; p < 2; ++p) us[e][t][p] = a.src[t][qbase[p] + (size_t)zc * NH + yc];
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int z = z0 + wz * 32 + zt * 16 + 4 * e + l4;
                const int y = y0 + wy * 32 + yt * 16 + l16;
                d2v rv = {0.0, 0.0};
                if (in[e]) {
                    const d2v fv = 0.25 * ((fq[e][0] + sr * fq[e][1]) + sc * (fq[e][2] + sr * fq[e][3]));
                    d2v u;
#pragma unroll
                    for (int p = 0; p < 2; ++p) {
                        double s = 0.0;
#pragma unroll
                        for (int t = 0; t < NSRC; ++t) s += (double)us[e][t][p];
                        u[p] = s;
                    }
                    const d2v lu = {acc[zt][yt][0][e], acc[zt][yt][1][e]};
                    rv = fv - (lu - xlam * u);
                    if (zim) rv[1] = 0.0;
                    rmx = fmax(rmx, fmax(fabs(rv[0]), fabs(rv[1])));
                    fmx = fmax(fmx, fmax(fabs(fv[0]), zim ? 0.0 : fabs(fv[1])));
                }
#pragma unroll
                for (int p = 0; p < 2; ++p) a.R[qbase[p] + (size_t)z * NH + y] = (float)rv[p];
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
}

hipError_t launch_resid_fold(const ResidFoldArgs& a, hipStream_t st)
{
    if (a.NH % RB || a.nsrc < 1 || a.nsrc > kMaxFoldSources) return hipErrorInvalidValue;
    const long ncplx = (long)a.g.nf * a.g.nmodes;
    if (ncplx <= 0) return hipSuccess;
    const int bpd = a.NH / RB, tpp = 4 * bpd * bpd;
    const long groups = (ncplx + kXcd - 1) / kXcd;
    const dim3 grid((unsigned)(groups * kXcd * tpp)), block(256);
    switch (a.nsrc) {
        case 1: hipLaunchKernelGGL(k_resid_fold<1>, grid, block, 0, st, a, ncplx); break;
        case 2: hipLaunchKernelGGL(k_resid_fold<2>, grid, block, 0, st, a, ncplx); break;
        case 3: hipLaunchKernelGGL(k_resid_fold<3>, grid, block, 0, st, a, ncplx); break;
        case 4: hipLaunchKernelGGL(k_resid_fold<4>, grid, block, 0, st, a, ncplx); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace duct
