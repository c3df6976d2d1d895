// Batched FP64 eigenbasis projection passes on v_mfma_f64_16x16x4f64.
//
// The matrix-diagonalisation solve of solvediag.f90 (Dirichlet :51-80,
// Neumann :140-219) is U = P * ((P^-1 F P^-T) ./ (lam_i + lam_j - xlambda)) * P^T
// for every real plane F.  The reference runs it per plane as dgemm('t','n'),
// an explicit transpose, dgemm, divide, dgemm, transpose, dgemm.  Here every
// one of the four products is one launch over ALL planes of a field group:
//
//     Y_p = (A * X_p)^T        for every plane p,   A = NP x NP basis
//
// Storing the product transposed is what makes the chain close:
//   pass 1 (C2P)      Y1 = (P^-1 F)^T                 complex field  -> planar
//   pass 2 (P2P)      Y2 = (P^-1 Y1)^T                planar -> planar
//   pass 3 (P2P_LDIV) Y3 = (P (Y2 ./ den))^T          planar -> planar, divide on the way into LDS
//   pass 4 (P2C)      Y4 = (P Y3)^T = P G P^T = U     planar -> complex field
// The complex<->real split of int2d3d.f90 (com2re/re2com) is fused into the
// loads of pass 1 and the stores of pass 4; the divide of :61-66 / :151-159 is
// applied to pass 3's input as it is staged in LDS (the same rounding as dividing in
// pass 2's epilogue, where it would stall every wave at once).  P2P_DIV, the
// divide in pass 2's epilogue, serves the adjoint and the mixed solve, whose
// gauge terms act on the divided G.
//
// Work decomposition.  A workgroup of 4 waves owns MT rows of A (all NP of them
// for NP <= 256) and NT columns of the concatenated planes.  A (<= 2 MB) is
// L2-resident and each wave streams its own rows of A straight from global
// memory in a fragment-major layout (one 16-byte load per lane covers two
// k-steps).  The data tile (16 k x NT columns) is staged through LDS, double
// buffered, and shared by the waves of the workgroup.  Each wave accumulates a
// (TM*16) x 64 tile in 4*TM v4f64 accumulators.
//
// MFMA lane layout (v_mfma_f64_16x16x4f64, checked by tests/test_gpu_solve.py):
//   A[i][k] from lane i + 16 k, B[k][j] from lane j + 16 k,
//   D[i][j] in lane j + 16 (i % 4), element i / 4 (rows interleaved by 4).
#include "common.h"
#include "proj.h"

namespace duct {


// Virtual column v of a complex tile -> (complex column within tile, part).
// Wave wn owns virtual columns [64 wn, 64 wn + 64) = 4 tiles of 16: tiles 0,1 are
// (re, im) of complex columns 32 wn + [0,16), tiles 2,3 of 32 wn + [16,32).
__device__ __forceinline__ int vcol_re(int ql) {
    return (ql >> 5) * 64 + ((ql >> 4) & 1) * 32 + (ql & 15);
}

template <int NP, int MODE>
__global__ __launch_bounds__(256, 2) void k_proj(ProjArgs a)
{
    using S = ProjShape<NP>;
    constexpr int NT = S::NT, TM = S::TM, LDB = S::LDB, NK = S::NK;
    constexpr bool CIN = MODE == PROJ_C2P;
    constexpr bool COUT = MODE == PROJ_P2C;
    constexpr bool CPLX = CIN || COUT;
    constexpr bool LDIV = MODE == PROJ_P2P_LDIV;

    __shared__ double Bs[2][NT * LDB];
    __shared__ double lam_s[LDIV ? NP : 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave % S::WM, wn = wave / S::WM;
    const int l16 = lane & 15, l4 = lane >> 4;
    const long tile = blockIdx.x;
    const int rowblk = blockIdx.y;
    const FieldGroup& g = a.g;
    const int n = g.n, m = g.m;
    // valid extents: K and output rows kn, columns cn; complex fields (ldf, ., .) of fpl points
    const int kn = a.kn ? a.kn : n, cn = a.cn ? a.cn : n, ldf = a.ldf ? a.ldf : m;
    const size_t fpl = a.ldf ? (size_t)a.fplane : (size_t)m * m;

    // ---------------- B-tile loader geometry ----------------
    // planar source (P2P, P2C): each thread loads DPT consecutive k of one column
    constexpr int DPT = NT / 16;           // doubles per thread per stage
    constexpr int TPC = 16 / DPT;          // threads per column
    // complex source (C2P): each thread loads CPT consecutive k of one complex column
    constexpr int CPT = NT / 32;
    constexpr int TPQ = 16 / CPT;

    // The B loads are unconditional: lanes with no column read a valid address and store zeros.
    // A load under a branch makes the compiler wait for every outstanding load, the next
    // stage's HBM reads included, before the first instruction that follows the branch.
    const double* srcp = a.X;              // planar column base (k = 0)
    const d2v* srcc = a.g.in[0] + 1 + ldf;     // complex column base (k = 0 -> y = 1)
    int v_dst = 0, v_dst_im = 0, kk0 = 0;
    bool src_ok = false, zero_im = false;

    // LDIV: column j of plane `plane` is divided by lam_j + lam_k - xlambda as it is staged in LDS
    int ld_j = 0;
    double ld_xlam = 0.0, ld_lamj = 0.0;
    bool ld_gauge = false;
    if constexpr (MODE == PROJ_P2P || MODE == PROJ_P2P_DIV || LDIV) {
        const int v = tid / TPC;
        kk0 = (tid % TPC) * DPT;
        const long c = tile * NT + v;
        v_dst = v;
        src_ok = c < a.ncols;
        if (src_ok) srcp = a.X + (size_t)c * NP;
        if constexpr (LDIV) {
            const long plane = c / NP;
            ld_j = (int)(c % NP);
            const long fm = plane >> 1;
            const int f = (int)(fm / g.nmodes), mode = (int)(fm % g.nmodes);
            if (src_ok) ld_xlam = xlambda_of(g.alambda[f], g.alp, g.mode0 + mode);
            if (src_ok && ld_j < cn) ld_lamj = (a.lam2 ? a.lam2 : a.lam)[ld_j];
            ld_gauge = a.gauge >= 0 && fabs(ld_xlam) <= 1.0e-10 && ld_j == a.gauge;
        }
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
            srcp = a.X + ((size_t)(fm * 2 + part) * NP + j) * NP;
        }
    } else {  // C2P
        const int ql = tid / TPQ;
        kk0 = (tid % TPQ) * CPT;
        v_dst = vcol_re(ql);
        v_dst_im = v_dst + 16;
        const long q = tile * (NT / 2) + ql;
        if (q < a.ncols) {
            const long fm = q / NP;
            const int j = (int)(q % NP);
            const int f = (int)(fm / g.nmodes), mode = (int)(fm % g.nmodes);
            src_ok = j < cn;
            zero_im = (g.mode0 + mode) == 0;   // int2d3d.f90:43 skips Im at k=0
            if (src_ok) srcc = g.in[f] + (size_t)mode * fpl + (size_t)(j + 1) * ldf + 1;
        }
    }

    if constexpr (LDIV) {
        for (int i = tid; i < NP; i += 256) lam_s[i] = a.lam[i];
        __syncthreads();
    }

    double breg[CPLX && CIN ? 1 : DPT];
    d2v creg[CIN ? CPT : 1];
    bool okc[CIN ? CPT : 1];

    auto load_b = [&](int ks) {
        const int k0 = ks * S::KT + kk0;
        if constexpr (CIN) {
#pragma unroll
            for (int e = 0; e < CPT; ++e) {
                okc[e] = src_ok && k0 + e < kn;
                creg[e] = srcc[okc[e] ? k0 + e : 0];
            }
        } else {
#pragma unroll
            for (int e = 0; e < DPT; e += 2) {
                const d2v v = *(const d2v*)(srcp + k0 + e);
                breg[e] = v[0];
                breg[e + 1] = v[1];
            }
        }
    };
    auto store_b = [&](int buf, int ks) {
        double* B = Bs[buf];
        if constexpr (CIN) {
#pragma unroll
            for (int e = 0; e < CPT; ++e) {
                B[v_dst * LDB + kk0 + e] = okc[e] ? creg[e][0] : 0.0;
                B[v_dst_im * LDB + kk0 + e] = (okc[e] && !zero_im) ? creg[e][1] : 0.0;
            }
        } else {
#pragma unroll
            for (int e = 0; e < DPT; ++e) {
                double v = src_ok ? breg[e] : 0.0;
                if constexpr (LDIV) {   // solvediag.f90:61-66 / :151-159, the gauge entry kept as set
                    const int k = ks * S::KT + kk0 + e;
                    if (ld_j < cn && k < kn && !(ld_gauge && k == a.gauge)) v = v / (ld_lamj + lam_s[k] - ld_xlam);
                }
                B[v_dst * LDB + kk0 + e] = v;
            }
        }
    };

    // ---------------- A fragments (global, fragment-major) ----------------
    const int mt0 = (rowblk * S::MT + wm * S::WR) / 16;
    const d2v* Ab = (const d2v*)a.Afrag + (size_t)mt0 * (NP / 8) * 64 + lane;
    d2v acur[TM][2], anext[TM][2];
    auto load_a = [&](int ks, d2v (&dst)[TM][2]) {
#pragma unroll
        for (int mt = 0; mt < TM; ++mt)
#pragma unroll
            for (int h = 0; h < 2; ++h) dst[mt][h] = Ab[((size_t)mt * (NP / 8) + ks * 2 + h) * 64];
    };

    d4v acc[TM][4];
#pragma unroll
    for (int mt = 0; mt < TM; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = d4v{0.0, 0.0, 0.0, 0.0};

    load_b(0);
    load_a(0, acur);
    store_b(0, 0);
    __syncthreads();

    for (int ks = 0; ks < NK; ++ks) {
        const int buf = ks & 1;
        if (ks + 1 < NK) {
            load_b(ks + 1);
            load_a(ks + 1, anext);
        }
        const double* B = Bs[buf];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            double bf[4];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) bf[nt] = B[(wn * 64 + nt * 16 + l16) * LDB + kk * 4 + l4];
#pragma unroll
            for (int mt = 0; mt < TM; ++mt) {
                const double av = acur[mt][kk >> 1][kk & 1];
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bf[nt], acc[mt][nt], 0, 0, 0);
            }
        }
        if (ks + 1 < NK) {
            store_b(buf ^ 1, ks + 1);
#pragma unroll
            for (int mt = 0; mt < TM; ++mt) {
                acur[mt][0] = anext[mt][0];
                acur[mt][1] = anext[mt][1];
            }
        }
        __syncthreads();
    }

    // ---------------- epilogue ----------------
    const int rbase = rowblk * S::MT + wm * S::WR + l4;
    if constexpr (COUT) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int ql = wn * 32 + s * 16 + l16;
            const long q = tile * (NT / 2) + ql;
            if (q >= a.ncols) continue;
            const long fm = q / NP;
            const int j = (int)(q % NP);
            if (j >= cn) continue;
            const int f = (int)(fm / g.nmodes), mode = (int)(fm % g.nmodes);
            const bool zim = (g.mode0 + mode) == 0;
            d2v* out = g.out[f] + (size_t)mode * fpl + (j + 1);
#pragma unroll
            for (int mt = 0; mt < TM; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rbase + mt * 16 + 4 * r;
                    if (row < kn) {
                        d2v v = {acc[mt][2 * s][r], zim ? 0.0 : acc[mt][2 * s + 1][r]};
                        out[(size_t)(row + 1) * ldf] = v;
                    }
                }
        }
    } else {
        // divide epilogue: the row eigenvalues are loaded once, before any store
        double lrow[TM][4];
        if constexpr (MODE == PROJ_P2P_DIV) {
            const double* lr = a.lam2 ? a.lam2 : a.lam;
#pragma unroll
            for (int mt = 0; mt < TM; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rbase + mt * 16 + 4 * r;
                    lrow[mt][r] = row < kn ? lr[row] : 0.0;
                }
        }
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
            double* Y = a.Y + (size_t)plane * NP * NP + j;
            double xlam = 0.0, lj = 0.0;
            bool gauge_plane = false;
            if constexpr (MODE == PROJ_P2P_DIV) {
                const long fm = plane >> 1;
                const int f = (int)(fm / g.nmodes), mode = (int)(fm % g.nmodes);
                xlam = xlambda_of(g.alambda[f], g.alp, g.mode0 + mode);
                gauge_plane = a.gauge >= 0 && fabs(xlam) <= 1.0e-10;
                lj = j < cn ? a.lam[j] : 0.0;
            }
#pragma unroll
            for (int mt = 0; mt < TM; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rbase + mt * 16 + 4 * r;
                    double v = acc[mt][nt][r];
                    if constexpr (MODE == PROJ_P2P_DIV) {
                        if (j < cn && row < kn && !(gauge_plane && j == a.gauge && row == a.gauge))
                            v = v / (lj + lrow[mt][r] - xlam);
                        else
                            v = 0.0;
                    }
                    Y[(size_t)row * NP] = v;
                }
        }
    }
}

template <int NP>
static hipError_t launch_np(int mode, const ProjArgs& a, hipStream_t st)
{
    using S = ProjShape<NP>;
    const bool cplx = mode == PROJ_C2P || mode == PROJ_P2C;
    const long per = cplx ? S::NT / 2 : S::NT;
    const long tiles = (a.ncols + per - 1) / per;
    if (tiles <= 0) return hipSuccess;
    dim3 grid((unsigned)tiles, S::ROWBLOCKS), block(256);
    switch (mode) {
        case PROJ_C2P: hipLaunchKernelGGL((k_proj<NP, PROJ_C2P>), grid, block, 0, st, a); break;
        case PROJ_P2P: hipLaunchKernelGGL((k_proj<NP, PROJ_P2P>), grid, block, 0, st, a); break;
        case PROJ_P2P_DIV: hipLaunchKernelGGL((k_proj<NP, PROJ_P2P_DIV>), grid, block, 0, st, a); break;
        case PROJ_P2C: hipLaunchKernelGGL((k_proj<NP, PROJ_P2C>), grid, block, 0, st, a); break;
        case PROJ_P2P_LDIV: hipLaunchKernelGGL((k_proj<NP, PROJ_P2P_LDIV>), grid, block, 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_proj(int NP, int mode, const ProjArgs& a, hipStream_t st)
{
    switch (NP) {
        case 32: return launch_np<32>(mode, a, st);
        case 64: return launch_np<64>(mode, a, st);
        case 128: return launch_np<128>(mode, a, st);
        case 256: return launch_np<256>(mode, a, st);
        case 512: return launch_np<512>(mode, a, st);
        default: return hipErrorInvalidValue;
    }
}

// Host-side builder of the fragment-major basis layout consumed above.
void build_frag(const double* A, int n, int lda, int NP, double* frag)
{
    // frag[((mt*(NP/8) + kp)*64 + lane)*2 + e] = Apad[mt*16 + lane%16][kp*8 + e*4 + lane/16]
    for (int mt = 0; mt < NP / 16; ++mt)
        for (int kp = 0; kp < NP / 8; ++kp)
            for (int lane = 0; lane < 64; ++lane)
                for (int e = 0; e < 2; ++e) {
                    const int i = mt * 16 + (lane & 15);
                    const int k = kp * 8 + e * 4 + (lane >> 4);
                    const double v = (i < n && k < n) ? A[(size_t)i + (size_t)k * lda] : 0.0;
                    frag[(((size_t)mt * (NP / 8) + kp) * 64 + lane) * 2 + e] = v;
                }
}

}  // namespace duct
