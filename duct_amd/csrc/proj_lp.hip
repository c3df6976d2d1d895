// Low-precision eigenbasis projection passes for the mixed-precision solve.
//
// Same four-pass chain as proj_f64.hip (Y_p = (A X_p)^T per plane; complex
// split fused into passes 1 and 4, divide fused into pass 2), with the products
// on the fast matrix cores and FP32 planar intermediates:
//   LP_F32     v_mfma_f32_16x16x4_f32                     (variant A)
//   LP_BF16X3  v_mfma_f32_16x16x32_bf16, operands split into 2 bf16 terms,
//              products hi*hi + hi*lo + lo*hi              (variant B, three terms)
//   LP_BF16X6  3 bf16 terms, the 6 products with ta+tx <= 2 (FP32-level accuracy)
//   LP_BF16X1  plain bf16 (1 product)
// Accumulation is FP32.  The eigenvalue divide of pass 2 is done in FP64 on
// the FP32 product and rounded back.  Pass 4 either writes the FP64 complex
// field or adds to it (the correction step of iterative refinement).
//
// Lane layouts (gfx950): f32 16x16x4: A[i][k] lane i+16k, B[k][j] lane j+16k;
// bf16 16x16x32: A[i][k] lane i+16(k/8) element k%8, B[k][j] lane j+16(k/8)
// element k%8; both write D[i][j] to lane j+16(i/4), element i%4.
#include "common.h"
#include "proj.h"
#include "proj_lp.h"

namespace duct {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef unsigned short u16;
typedef u16 u8v __attribute__((ext_vector_type(8)));

template <int S> struct LpT;
template <> struct LpT<LP_F32> { static constexpr int T = 1, MAXO = 0; static constexpr bool BF = false; };
template <> struct LpT<LP_BF16X1> { static constexpr int T = 1, MAXO = 0; static constexpr bool BF = true; };
template <> struct LpT<LP_BF16X3> { static constexpr int T = 2, MAXO = 1; static constexpr bool BF = true; };
template <> struct LpT<LP_BF16X6> { static constexpr int T = 3, MAXO = 2; static constexpr bool BF = true; };

__host__ __device__ inline u16 to_bf16_bits(float x)
{
    // round to nearest even (finite inputs; NaN stays NaN through the carry-free path)
    unsigned u = __builtin_bit_cast(unsigned, x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (u16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__host__ __device__ inline float bf16_bits_to_float(u16 h)
{
    return __builtin_bit_cast(float, (unsigned)h << 16);
}

// Split x into T bf16 terms with x ~= sum of terms (each remainder exact in FP64).
template <int T>
__device__ __forceinline__ void split_terms(double x, u16 (&t)[T])
{
    double r = x;
#pragma unroll
    for (int i = 0; i < T; ++i) {
        t[i] = to_bf16_bits((float)r);
        r -= (double)bf16_bits_to_float(t[i]);
    }
}

// Two workgroups per CU, except the 3-term bf16 split at the smallest tiles (NP <= 64: its operand
// terms do not fit two waves per SIMD).
template <int NP, int MODE, int S>
__global__ __launch_bounds__(256, (NP <= 64 && S == LP_BF16X6) ? 1 : 2) void k_proj_lp(ProjLpArgs a)
{
    using Sh = ProjShape<NP>;
    using Tr = LpT<S>;
    constexpr int NT = Sh::NT, TM = Sh::TM;
    constexpr int KT = 32;
    constexpr int NK = NP / KT;
    constexpr int T = Tr::T;
    constexpr bool BF = Tr::BF;
    constexpr bool CIN = MODE == PROJ_C2P;
    constexpr bool COUT = MODE == PROJ_P2C;
    constexpr int LDF = KT + 1;   // f32 LDS column stride (floats)
    constexpr int LDH = KT + 8;   // bf16 LDS column stride (elements, 80 bytes)

    // LDS: f32 tiles [2][NT][LDF] or bf16 tiles [2][T][NT][LDH]
    constexpr int LDS_BYTES = BF ? 2 * T * NT * LDH * 2 : 2 * NT * LDF * 4;
    __shared__ __attribute__((aligned(16))) unsigned char smem[LDS_BYTES];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave % Sh::WM, wn = wave / Sh::WM;
    const int l16 = lane & 15, l4 = lane >> 4;
    const long tile = blockIdx.x;
    const int rowblk = blockIdx.y;
    const FieldGroup& g = a.g;
    const int n = g.n, m = g.m;

    // ---------------- B-tile loader geometry ----------------
    constexpr int DPT = NT * KT / 256;   // planar floats per thread per stage
    constexpr int TPC = KT / DPT;
    constexpr int CPT = NT * KT / 512;   // complex values per thread per stage
    constexpr int TPQ = KT / CPT;

    const float* srcp = nullptr;
    const d2v* srcc = nullptr;
    int v_dst = 0, v_dst_im = 0, kk0 = 0;
    bool src_ok = false, zero_im = false;
    if constexpr (MODE == PROJ_P2P || MODE == PROJ_P2P_DIV) {
        const int v = tid / TPC;
        kk0 = (tid % TPC) * DPT;
        const long c = tile * NT + v;
        v_dst = v;
        src_ok = c < a.ncols;
        if (src_ok) srcp = (const float*)a.X + (size_t)c * NP;
    } else if constexpr (COUT) {
        const int v = tid / TPC;
        kk0 = (tid % TPC) * DPT;
        v_dst = v;
        const int t = (v & 63) >> 4;
        const int ql = (v >> 6) * 32 + (t >> 1) * 16 + (v & 15);
        const int part = t & 1;
        const long q = tile * (NT / 2) + ql;
        src_ok = q < a.ncols;
        if (src_ok) {
            const long fm = q / NP;
            const int j = (int)(q % NP);
            srcp = (const float*)a.X + ((size_t)(fm * 2 + part) * NP + j) * NP;
        }
    } else {
        const int ql = tid / TPQ;
        kk0 = (tid % TPQ) * CPT;
        v_dst = (ql >> 5) * 64 + ((ql >> 4) & 1) * 32 + (ql & 15);
        v_dst_im = v_dst + 16;
        const long q = tile * (NT / 2) + ql;
        if (q < a.ncols) {
            const long fm = q / NP;
            const int j = (int)(q % NP);
            const int f = (int)(fm / g.nmodes), mode = (int)(fm % g.nmodes);
            src_ok = j < n;
            zero_im = (g.mode0 + mode) == 0;
            if (src_ok) srcc = g.in[f] + (size_t)mode * m * m + (size_t)(j + 1) * m + 1;
        }
    }

    float breg[CIN ? 1 : DPT];
    d2v creg[CIN ? CPT : 1];
    auto load_b = [&](int ks) {
        const int k0 = ks * KT + kk0;
        if constexpr (CIN) {
#pragma unroll
            for (int e = 0; e < CPT; ++e) {
                d2v v = {0.0, 0.0};
                if (src_ok && k0 + e < n) v = srcc[k0 + e];
                creg[e] = v;
            }
        } else {
#pragma unroll
            for (int e = 0; e < DPT; e += 4) {
                f4v v = {0.f, 0.f, 0.f, 0.f};
                if (src_ok) v = *(const f4v*)(srcp + k0 + e);
                breg[e] = v[0];
                breg[e + 1] = v[1];
                breg[e + 2] = v[2];
                breg[e + 3] = v[3];
            }
        }
    };
    auto put = [&](int buf, int v, int k, double x) {
        if constexpr (BF) {
            u16 t[T];
            split_terms<T>(x, t);
            u16* B = (u16*)smem + (size_t)buf * T * NT * LDH;
#pragma unroll
            for (int i = 0; i < T; ++i) B[(size_t)i * NT * LDH + v * LDH + k] = t[i];
        } else {
            float* B = (float*)smem + (size_t)buf * NT * LDF;
            B[v * LDF + k] = (float)x;
        }
    };
    auto store_b = [&](int buf) {
        if constexpr (CIN) {
#pragma unroll
            for (int e = 0; e < CPT; ++e) {
                put(buf, v_dst, kk0 + e, creg[e][0]);
                put(buf, v_dst_im, kk0 + e, zero_im ? 0.0 : creg[e][1]);
            }
        } else {
#pragma unroll
            for (int e = 0; e < DPT; ++e) put(buf, v_dst, kk0 + e, (double)breg[e]);
        }
    };

    // ---------------- A fragments ----------------
    // f32:  [mt][stage][lane][8]  = A[mt*16 + l%16][stage*32 + e*4 + l/16]
    // bf16: [term][mt][stage][lane][8] = term(A[mt*16 + l%16][stage*32 + 8*(l/16) + e])
    const int mt0 = (rowblk * Sh::MT + wm * Sh::WR) / 16;
    constexpr int AV = BF ? T : 2;   // 16-byte vectors per (mt, stage)
    typedef f4v AVec;
    AVec acur[TM][AV], anext[TM][AV];
    auto load_a = [&](int ks, AVec (&dst)[TM][AV]) {
#pragma unroll
        for (int mt = 0; mt < TM; ++mt) {
            if constexpr (BF) {
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    const size_t idx = ((((size_t)t * (NP / 16) + mt0 + mt) * NK + ks) * 64 + lane);
                    dst[mt][t] = ((const AVec*)a.Afrag)[idx];
                }
            } else {
                const size_t idx = (((size_t)(mt0 + mt) * NK + ks) * 64 + lane) * 2;
                dst[mt][0] = ((const AVec*)a.Afrag)[idx];
                dst[mt][1] = ((const AVec*)a.Afrag)[idx + 1];
            }
        }
    };

    f4v acc[TM][4];
#pragma unroll
    for (int mt = 0; mt < TM; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f4v{0.f, 0.f, 0.f, 0.f};

    load_b(0);
    load_a(0, acur);
    store_b(0);
    __syncthreads();

    for (int ks = 0; ks < NK; ++ks) {
        const int buf = ks & 1;
        if (ks + 1 < NK) {
            load_b(ks + 1);
            load_a(ks + 1, anext);
        }
        if constexpr (BF) {
            const u16* B = (const u16*)smem + (size_t)buf * T * NT * LDH;
            u8v bx[4][T];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int t = 0; t < T; ++t)
                    bx[nt][t] = *(const u8v*)(B + (size_t)t * NT * LDH + (wn * 64 + nt * 16 + l16) * LDH + 8 * l4);
#pragma unroll
            for (int mt = 0; mt < TM; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                    for (int o = Tr::MAXO; o >= 0; --o)        // small products first
#pragma unroll
                        for (int ta = 0; ta < T; ++ta) {
                            const int tx = o - ta;
                            if (tx < 0 || tx >= T) continue;
                            acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                __builtin_bit_cast(bf8v, acur[mt][ta]), __builtin_bit_cast(bf8v, bx[nt][tx]),
                                acc[mt][nt], 0, 0, 0);
                        }
        } else {
            const float* B = (const float*)smem + (size_t)buf * NT * LDF;
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) {
                float bf[4];
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) bf[nt] = B[(wn * 64 + nt * 16 + l16) * LDF + kk * 4 + l4];
#pragma unroll
                for (int mt = 0; mt < TM; ++mt) {
                    const float av = acur[mt][kk >> 2][kk & 3];
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt)
                        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bf[nt], acc[mt][nt], 0, 0, 0);
                }
            }
        }
        if (ks + 1 < NK) {
            store_b(buf ^ 1);
#pragma unroll
            for (int mt = 0; mt < TM; ++mt)
#pragma unroll
                for (int t = 0; t < AV; ++t) acur[mt][t] = anext[mt][t];
        }
        __syncthreads();
    }

    // ---------------- epilogue ----------------
    const int rbase = rowblk * Sh::MT + wm * Sh::WR + l4 * 4;
    if constexpr (COUT) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int ql = wn * 32 + s * 16 + l16;
            const long q = tile * (NT / 2) + ql;
            if (q >= a.ncols) continue;
            const long fm = q / NP;
            const int j = (int)(q % NP);
            if (j >= n) continue;
            const int f = (int)(fm / g.nmodes), mode = (int)(fm % g.nmodes);
            const bool zim = (g.mode0 + mode) == 0;
            d2v* out = g.out[f] + (size_t)mode * m * m + (j + 1);
#pragma unroll
            for (int mt = 0; mt < TM; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rbase + mt * 16 + r;
                    if (row < n) {
                        d2v v = {(double)acc[mt][2 * s][r], zim ? 0.0 : (double)acc[mt][2 * s + 1][r]};
                        d2v* p = out + (size_t)(row + 1) * m;
                        if (a.accumulate) v += *p;
                        *p = v;
                    }
                }
        }
    } else {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            long plane;
            int j;
            bool ok;
            if constexpr (CIN) {
                const int ql = wn * 32 + (nt >> 1) * 16 + l16;
                const long q = tile * (NT / 2) + ql;
                ok = q < a.ncols;
                plane = (q / NP) * 2 + (nt & 1);
                j = (int)(q % NP);
            } else {
                const long c = tile * NT + wn * 64 + nt * 16 + l16;
                ok = c < a.ncols;
                plane = c / NP;
                j = (int)(c % NP);
            }
            if (!ok) continue;
            float* Y = a.Y + (size_t)plane * NP * NP + j;
            double xlam = 0.0;
            bool gauge_plane = false;
            if constexpr (MODE == PROJ_P2P_DIV) {
                const long fm = plane >> 1;
                const int f = (int)(fm / g.nmodes), mode = (int)(fm % g.nmodes);
                xlam = xlambda_of(g.alambda[f], g.alp, g.mode0 + mode);
                gauge_plane = a.gauge >= 0 && fabs(xlam) <= 1.0e-10;
            }
#pragma unroll
            for (int mt = 0; mt < TM; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rbase + mt * 16 + r;
                    float v = acc[mt][nt][r];
                    if constexpr (MODE == PROJ_P2P_DIV) {
                        if (j < n && row < n && !(gauge_plane && j == a.gauge && row == a.gauge))
                            v = (float)((double)v / (a.lam[j] + a.lam[row] - xlam));
                        else
                            v = 0.f;
                    }
                    Y[(size_t)row * NP] = v;
                }
        }
    }
}

template <int NP, int S>
static hipError_t launch_np_s(int mode, const ProjLpArgs& a, hipStream_t st)
{
    using Sh = ProjShape<NP>;
    const bool cplx = mode == PROJ_C2P || mode == PROJ_P2C;
    const long per = cplx ? Sh::NT / 2 : Sh::NT;
    const long tiles = (a.ncols + per - 1) / per;
    if (tiles <= 0) return hipSuccess;
    dim3 grid((unsigned)tiles, Sh::ROWBLOCKS), block(256);
    switch (mode) {
        case PROJ_C2P: hipLaunchKernelGGL((k_proj_lp<NP, PROJ_C2P, S>), grid, block, 0, st, a); break;
        case PROJ_P2P: hipLaunchKernelGGL((k_proj_lp<NP, PROJ_P2P, S>), grid, block, 0, st, a); break;
        case PROJ_P2P_DIV: hipLaunchKernelGGL((k_proj_lp<NP, PROJ_P2P_DIV, S>), grid, block, 0, st, a); break;
        case PROJ_P2C: hipLaunchKernelGGL((k_proj_lp<NP, PROJ_P2C, S>), grid, block, 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int NP>
static hipError_t launch_np(int scheme, int mode, const ProjLpArgs& a, hipStream_t st)
{
    switch (scheme) {
        case LP_F32: return launch_np_s<NP, LP_F32>(mode, a, st);
        case LP_BF16X1: return launch_np_s<NP, LP_BF16X1>(mode, a, st);
        case LP_BF16X3: return launch_np_s<NP, LP_BF16X3>(mode, a, st);
        case LP_BF16X6: return launch_np_s<NP, LP_BF16X6>(mode, a, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_proj_lp(int NP, int scheme, int mode, const ProjLpArgs& a, hipStream_t st)
{
    switch (NP) {
        case 32: return launch_np<32>(scheme, mode, a, st);
        case 64: return launch_np<64>(scheme, mode, a, st);
        case 128: return launch_np<128>(scheme, mode, a, st);
        case 256: return launch_np<256>(scheme, mode, a, st);
        case 512: return launch_np<512>(scheme, mode, a, st);
        default: return hipErrorInvalidValue;
    }
}

size_t lp_frag_bytes(int NP, int scheme)
{
    if (scheme == LP_F32) return (size_t)NP * NP * sizeof(float);
    const int T = scheme == LP_BF16X6 ? 3 : scheme == LP_BF16X3 ? 2 : 1;
    return (size_t)T * NP * NP * sizeof(u16);
}

void build_frag_lp(const double* A, int n, int lda, int NP, int scheme, void* out)
{
    const int NK = NP / 32;
    auto at = [&](int i, int k) -> double { return (i < n && k < n) ? A[(size_t)i + (size_t)k * lda] : 0.0; };
    if (scheme == LP_F32) {
        float* f = (float*)out;
        for (int mt = 0; mt < NP / 16; ++mt)
            for (int s = 0; s < NK; ++s)
                for (int lane = 0; lane < 64; ++lane)
                    for (int e = 0; e < 8; ++e)
                        f[(((size_t)mt * NK + s) * 64 + lane) * 8 + e] =
                            (float)at(mt * 16 + (lane & 15), s * 32 + e * 4 + (lane >> 4));
        return;
    }
    const int T = scheme == LP_BF16X6 ? 3 : scheme == LP_BF16X3 ? 2 : 1;
    u16* h = (u16*)out;
    for (int mt = 0; mt < NP / 16; ++mt)
        for (int s = 0; s < NK; ++s)
            for (int lane = 0; lane < 64; ++lane)
                for (int e = 0; e < 8; ++e) {
                    double r = at(mt * 16 + (lane & 15), s * 32 + 8 * (lane >> 4) + e);
                    for (int t = 0; t < T; ++t) {
                        const u16 b = to_bf16_bits((float)r);
                        r -= (double)bf16_bits_to_float(b);
                        h[((((size_t)t * (NP / 16) + mt) * NK + s) * 64 + lane) * 8 + e] = b;
                    }
                }
}

}  // namespace duct
