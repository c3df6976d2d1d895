// Folded residual of the refined solve on the int8 matrix cores.
//
// The refinement (mixed_fold.hip) needs r = f - (S_r U + U S_c^T - xlambda U) on every
// quarter to FP64 accuracy: r is ~1e-6 of the terms it is the difference of, and the
// products' rounding error is what limits the refined solution.  Here both products are
// formed exactly in integer arithmetic from byte digits:
//   S(y, k) = 2^e(y) sum_{t=1..7} d_t(y, k) 2^-8t      (host, build_sdig; exact to 2^-57 e)
//   U(k, z) = 2^g    sum_{u=1..kUDig} c_u(k, z) 2^-8u  (launch_lpq's digit output; exact)
//   (S U)(y, z) = 2^(e(y)+g) sum_w 2^-8w sum_{t+u=w} (d_t c_u)(y, z)
// with every digit product (d_t c_u) an exact int32 v_mfma_i32_16x16x64_i8 sum over k.  The
// pairs with t + u > kDigT = 8 are left out: they are below 2^-64 of the row's terms (the
// int8 products stand in for FP64 ones whose rounding is 2^-53 of the same terms).  The pairs
// of one level w = t + u share a scale and one int32 accumulator; the seven levels are exact
// integers, converted to FP64 and summed from the smallest, which rounds at 2^-53 of the terms
// as an FP64 product does.
//
// Work decomposition: a persistent workgroup owns one 64 x 64 block of one quarter type
// (parities r, c) of one part (Re or Im) and walks a stride of complex planes; the S digits of
// its rows (S_r) and columns (S_c) stay in LDS for its lifetime.  Its 4 waves own 32 x 32
// each.  The block configurations of one plane stream sit on one XCD, so a plane's U digits
// and f are read from HBM once.
#include <cmath>
#include <cstring>
#include <vector>

#include "mixed_fold.h"

namespace duct {

typedef float f4v __attribute__((ext_vector_type(4)));

namespace {
constexpr int kXcds = 8;

__device__ __forceinline__ void plane_decode_i8(const FieldGroup& g, long fmp, int& f, int& mode)
{
    mode = (int)(fmp / g.nf);
    f = (int)(fmp % g.nf);
}

__device__ __forceinline__ void atomic_max_pos_i8(unsigned long long* p, double v)
{
    atomicMax(p, (unsigned long long)__builtin_bit_cast(unsigned long long, v));
}

// One direction of the residual for the wave's 2 x 2 output tiles D[y][z] (D[at][bt], at the A
// operand's 16-row tile, bt the B operand's), in units of 2^(e + g):
// v = sum_w 2^-8w L_w with L_w = sum_{t+u=w} (d_t c_u) exact in int32 (|L_w| < 3 2^21), summed in
// FP64 from the smallest level up.  The S digit t is on one side, the U digit u on the other:
//   SA = true : A = S digits, B = U digits (y-product: A[y][k] = S_r[y][k], B[k][z] = U[k][z])
//   SA = false: A = U digits, B = S digits (z-product: A[y][k] = U[y][k], B[k][z] = S_c[z][k])
template <int KC, bool SA>
__device__ __forceinline__ void digit_products(const i4v (&uf)[2][kUDig][KC], const i4v* sfr, int lane,
                                               double (&v)[2][2][4])
{
    constexpr int NL = kDigT - 1;   // levels w = 2 .. kDigT
    i4v Lv[NL][2][2];
#pragma unroll
    for (int w = 0; w < NL; ++w)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) Lv[w][a][b] = i4v{0, 0, 0, 0};
    // S digit t of the wave's two S tiles, read from LDS one digit ahead of its products
    i4v sfb[2][2][KC];
    auto read_s = [&](int t, i4v (&dst)[2][KC]) {
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) dst[s][kc] = sfr[((s * kSDig + (t - 1)) * KC + kc) * 64 + lane];
    };
    read_s(1, sfb[0]);
#pragma unroll
    for (int t = 1; t <= kSDig; ++t) {
        i4v (&sf)[2][KC] = sfb[(t - 1) & 1];
        if (t < kSDig) read_s(t + 1, sfb[t & 1]);
        // consecutive products go to different accumulators (the two k chunks of one tile apart)
#pragma unroll
        for (int kc = 0; kc < KC; ++kc)
#pragma unroll
            for (int u = 1; u <= kUDig; ++u) {
                if (t + u > kDigT) continue;
#pragma unroll
                for (int at = 0; at < 2; ++at)
#pragma unroll
                    for (int bt = 0; bt < 2; ++bt) {
                        i4v& L = Lv[t + u - 2][at][bt];
                        if (SA)
                            L = __builtin_amdgcn_mfma_i32_16x16x64_i8(sf[at][kc], uf[bt][u - 1][kc], L, 0, 0, 0);
                        else
                            L = __builtin_amdgcn_mfma_i32_16x16x64_i8(uf[at][u - 1][kc], sf[bt][kc], L, 0, 0, 0);
                    }
            }
    }
    // adjacent levels joined exactly in int32 first (|L_w 256 + L_(w+1)| < 3 2^29 + 3 2^21 < 2^31)
    static_assert(kDigT == 8, "level grouping below assumes levels 2..8");
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int j23 = Lv[0][a][b][e] * 256 + Lv[1][a][b][e];
                const int j45 = Lv[2][a][b][e] * 256 + Lv[3][a][b][e];
                const int j67 = Lv[4][a][b][e] * 256 + Lv[5][a][b][e];
                double x = (double)Lv[6][a][b][e] * 0x1p-64;
                x = fma((double)j67, 0x1p-56, x);
                x = fma((double)j45, 0x1p-40, x);
                v[a][b][e] = fma((double)j23, 0x1p-24, x);
            }
}
}  // namespace

// Operands of one plane for one wave, loaded a plane ahead: the U digits of the first product
// and, for the epilogue, f as its FP32 pair and the digits of U at the wave's output points.
// Output tile D[y][z] of 16 x 16 holds 4 consecutive y per lane (lane z%16 + 16 (y%16)/4,
// element y%4), so f is one 16-byte load per tile from the column-major quarters and each digit
// of U one 4-byte load.
template <int KC>
struct PlaneOps {
    i4v ub[2][kUDig][KC];   // B operand of the y-product: columns z of Ucol (k = y')
    f4v fh[2][2], fl[2][2];
    unsigned uw[2][2][kUDig];
    int g;
};

template <int NH>
__device__ __forceinline__ void load_plane(const ResidI8Args& a, size_t qp, int y0, int z0, int l16, int l4,
                                           PlaneOps<NH / 64>& o)
{
    constexpr int KC = NH / 64;
    constexpr size_t NN = (size_t)NH * NH;
    o.g = a.gexp[qp];
    const unsigned char* ub0 = a.Ucol + qp * kUDig * NN + (size_t)(z0 + l16) * NH + l4 * 16;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int d = 0; d < kUDig; ++d)
#pragma unroll
            for (int kc = 0; kc < KC; ++kc)
                o.ub[t][d][kc] = *(const i4v*)(ub0 + d * NN + t * 16 * NH + kc * 64);
    // output tile (yt, zt): z = z0 + zt 16 + l16, y = y0 + yt 16 + 4 l4 + e
    const size_t pt = qp * NN + (size_t)(z0 + l16) * NH + y0 + 4 * l4;
#pragma unroll
    for (int yt = 0; yt < 2; ++yt)
#pragma unroll
        for (int zt = 0; zt < 2; ++zt) {
            const size_t off = pt + (size_t)zt * 16 * NH + yt * 16;
            o.fh[yt][zt] = *(const f4v*)(a.Fhi + off);
            o.fl[yt][zt] = *(const f4v*)(a.Flo + off);
            const size_t ud = qp * kUDig * NN + (size_t)(z0 + l16 + zt * 16) * NH + y0 + 4 * l4 + yt * 16;
#pragma unroll
            for (int d = 0; d < kUDig; ++d) o.uw[yt][zt][d] = *(const unsigned*)(a.Ucol + ud + d * NN);
        }
}

// One wave per SIMD: the products' operands of the next plane are loaded while this plane's
// products run (the registers for that, ~400 per wave, allow no second wave).
template <int NH, bool STATS>
__global__ __launch_bounds__(256, 1) void k_resid_i8(ResidI8Args a, long ncplx, int ncfg2, int spx)
{
    constexpr int KC = NH / 64;
    constexpr int BPD = NH / 64;
    constexpr int FR = 4 * kSDig * KC;   // fragments of one 64-row block of S
    __shared__ i4v Ssh[2][FR * 64];      // [0]: S_r rows of the y block, [1]: S_c rows of the z block

    const FieldGroup& g = a.g;
    const int xcd = blockIdx.x % kXcds, k8 = blockIdx.x / kXcds;
    const int cfg2 = k8 % ncfg2, stream = k8 / ncfg2;
    const int part = cfg2 & 1, cfg = cfg2 >> 1;
    const long stride = (long)kXcds * spx;
    const long p0 = (long)xcd * spx + stream;
    const int q = cfg / (BPD * BPD), r = q >> 1, c = q & 1;
    const int yb = (cfg / BPD) % BPD, zb = cfg % BPD;
    {
        const i4v* gr = a.Sdig[r] + (size_t)yb * FR * 64;
        const i4v* gc = a.Sdig[c] + (size_t)zb * FR * 64;
        for (int e = threadIdx.x; e < FR * 64; e += 256) {
            Ssh[0][e] = gr[e];
            Ssh[1][e] = gc[e];
        }
    }
    __syncthreads();

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int wy = wave >> 1, wz = wave & 1;
    const int y0 = yb * 64 + wy * 32, z0 = zb * 64 + wz * 32;
    const int n = g.n;
    const int hh[2] = {(n + 1) / 2, n / 2};
    const i4v* sy = &Ssh[0][(size_t)(wy * 2) * kSDig * KC * 64];   // the wave's two S_r row tiles
    const i4v* sz = &Ssh[1][(size_t)(wz * 2) * kSDig * KC * 64];   // the wave's two S_c row tiles
    int ey[2][4], ez[2];
#pragma unroll
    for (int yt = 0; yt < 2; ++yt)
#pragma unroll
        for (int e = 0; e < 4; ++e) ey[yt][e] = a.Sexp[r][y0 + yt * 16 + 4 * l4 + e];
#pragma unroll
    for (int zt = 0; zt < 2; ++zt) ez[zt] = a.Sexp[c][z0 + zt * 16 + l16];

    // the planes this workgroup visits (Im of mode 0 is not solved: skipped)
    auto next_plane = [&](long fmp) -> long {
        for (; fmp < ncplx; fmp += stride) {
            int f, mode;
            plane_decode_i8(g, fmp, f, mode);
            if (!(part == 1 && g.mode0 + mode == 0)) return fmp;
        }
        return ncplx;
    };
    long fmp = next_plane(p0);
    if (fmp >= ncplx) return;
    PlaneOps<KC> cur, nxt;
    load_plane<NH>(a, (size_t)(2 * fmp + part) * 4 + q, y0, z0, l16, l4, cur);
    while (fmp < ncplx) {
        const size_t qp = (size_t)(2 * fmp + part) * 4 + q;
        // A operand of the z-product (rows y of Urow, k = z'), issued before the next plane's
        // loads so that waiting for it does not wait for them
        i4v ua[2][kUDig][KC];
        {
            const unsigned char* ua0 = a.Urow + qp * kUDig * NH * NH + (size_t)(y0 + l16) * NH + l4 * 16;
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int d = 0; d < kUDig; ++d)
#pragma unroll
                    for (int kc = 0; kc < KC; ++kc)
                        ua[t][d][kc] = *(const i4v*)(ua0 + (size_t)d * NH * NH + t * 16 * NH + kc * 64);
        }
        // the scheduler keeps these loads after ua's: the z-product then waits for ua alone
        __builtin_amdgcn_sched_barrier(0);
        const long fnext = next_plane(fmp + stride);
        if (fnext < ncplx) load_plane<NH>(a, (size_t)(2 * fnext + part) * 4 + q, y0, z0, l16, l4, nxt);
        __builtin_amdgcn_sched_barrier(0);
        int f, mode;
        plane_decode_i8(g, fmp, f, mode);
        const int gq = cur.g;
        double ly[2][2][4], lzv[2][2][4];
        digit_products<KC, true>(cur.ub, sy, lane, ly);     // S_r U: A = S_r digits, B = U digits
        __builtin_amdgcn_sched_barrier(0);                  // one product's levels live at a time
        digit_products<KC, false>(ua, sz, lane, lzv);       // U S_c^T: A = U digits, B = S_c digits
        // r = f - ((S_r U + U S_c^T) - xlambda U), the order of the FP64 residual
        const double xlam = xlambda_of(g.alambda[f], g.alp, g.mode0 + mode);
        double rmx = 0.0, fmx = 0.0;
        float* R = a.R + qp * NH * NH + (size_t)(z0 + l16) * NH + y0 + 4 * l4;
#pragma unroll
        for (int yt = 0; yt < 2; ++yt)
#pragma unroll
            for (int zt = 0; zt < 2; ++zt) {
                f4v rout;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int y = y0 + yt * 16 + 4 * l4 + e, z = z0 + zt * 16 + l16;
                    const double lyv = ldexp(ly[yt][zt][e], ey[yt][e] + gq);
                    const double lz = ldexp(lzv[yt][zt][e], ez[zt] + gq);
                    // U = 2^(g - 8 kUDig) (c1 256^(kUDig-1) + ... + c_kUDig), the digits
                    // sign-extended from byte e
                    int Z = 0;
#pragma unroll
                    for (int d = 0; d < kUDig; ++d) Z = Z * 256 + ((int)(cur.uw[yt][zt][d] << (24 - 8 * e)) >> 24);
                    const double uv = ldexp((double)Z, gq - 8 * kUDig);
                    const double fv = (double)cur.fh[yt][zt][e] + (double)cur.fl[yt][zt][e];
                    double rv = fv - ((lyv + lz) - xlam * uv);
                    const bool in = y < hh[r] && z < hh[c];
                    if (!in) rv = 0.0;
                    if (STATS) {
                        if (in) rmx = fmax(rmx, fabs(rv));
                        if (in) fmx = fmax(fmx, fabs(fv));
                    }
                    rout[e] = (float)rv;
                }
                *(f4v*)(R + (size_t)zt * 16 * NH + yt * 16) = rout;
            }
        if (STATS) {
#pragma unroll
            for (int sh = 32; sh > 0; sh >>= 1) {
                rmx = fmax(rmx, __shfl_xor(rmx, sh, 64));
                fmx = fmax(fmx, __shfl_xor(fmx, sh, 64));
            }
            if (lane == 0) {
                atomic_max_pos_i8(a.stats + 2 * a.slot[f], rmx);
                atomic_max_pos_i8(a.stats + 2 * a.slot[f] + 1, fmx);
            }
        }
        cur = nxt;
        fmp = fnext;
    }
}

hipError_t launch_resid_i8(const ResidI8Args& a, hipStream_t st)
{
    const long ncplx = (long)a.g.nf * a.g.nmodes;
    if (ncplx <= 0) return hipSuccess;
    const int bpd = a.NH / 64, ncfg2 = 2 * 4 * bpd * bpd;   // (part, quarter, 64-block)
    const int spx = ncfg2 >= 32 ? 1 : 32 / ncfg2;          // plane streams per XCD: 256 workgroups
    const dim3 grid((unsigned)(kXcds * ncfg2 * spx)), block(256);
    const bool stats = a.stats != nullptr;
    switch (a.NH) {
        case 64:
            if (stats) hipLaunchKernelGGL((k_resid_i8<64, true>), grid, block, 0, st, a, ncplx, ncfg2, spx);
            else hipLaunchKernelGGL((k_resid_i8<64, false>), grid, block, 0, st, a, ncplx, ncfg2, spx);
            break;
        case 128:
            if (stats) hipLaunchKernelGGL((k_resid_i8<128, true>), grid, block, 0, st, a, ncplx, ncfg2, spx);
            else hipLaunchKernelGGL((k_resid_i8<128, false>), grid, block, 0, st, a, ncplx, ncfg2, spx);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

void build_sdig(const double* B, int NH, std::vector<i4v>& frag, std::vector<int>& exps)
{
    const int KC = NH / 64, RT = NH / 16;
    exps.assign(NH, 0);
    std::vector<signed char> dig((size_t)kSDig * NH * NH, 0);   // [t][row][k]
    for (int i = 0; i < NH; ++i) {
        double mx = 0.0;
        for (int k = 0; k < NH; ++k) mx = std::fmax(mx, std::fabs(B[i + (size_t)k * NH]));
        if (!(mx > 0.0)) continue;
        int E = 0;
        (void)std::frexp(mx / 0.49, &E);        // mx 2^-E < 0.49
        exps[i] = E;
        for (int k = 0; k < NH; ++k) {
            // Z = rint(x 2^56), |Z| < 0.49 2^56, as signed base-256 digits, most significant first
            long long Z = std::llrint(std::ldexp(B[i + (size_t)k * NH], 8 * kSDig - E));
            for (int t = kSDig - 1; t >= 1; --t) {
                const long long d = ((Z + 128) & 255) - 128;
                dig[((size_t)t * NH + i) * NH + k] = (signed char)d;
                Z = (Z - d) / 256;
            }
            dig[(size_t)i * NH + k] = (signed char)Z;   // |Z| <= 126
        }
    }
    frag.assign((size_t)RT * kSDig * KC * 64, i4v{0, 0, 0, 0});
    for (int rt = 0; rt < RT; ++rt)
        for (int t = 0; t < kSDig; ++t)
            for (int kc = 0; kc < KC; ++kc)
                for (int lane = 0; lane < 64; ++lane) {
                    signed char b[16];
                    for (int e = 0; e < 16; ++e)
                        b[e] = dig[((size_t)t * NH + rt * 16 + (lane & 15)) * NH + kc * 64 + (lane >> 4) * 16 + e];
                    std::memcpy(&frag[(((size_t)rt * kSDig + t) * KC + kc) * 64 + lane], b, 16);
                }
}

}  // namespace duct
