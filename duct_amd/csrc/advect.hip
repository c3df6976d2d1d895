// Products of the advection term in divergence form (operators.F:9-276).
//
// advection3d takes u, v, w (Fourier in x, physical in y and z), transforms
// each to physical x with the 3/2 rule (fourx with fill, fourier.F:1-115),
// forms the six products uu, vv, ww, uv, uw, vw (mult.F) and transforms them
// back, keeping modes 0..M-1 scaled by 1/mgalx (pack_fftw).  The reference
// does this with one FFTW call per (y, z) line and keeps the physical fields
// in memory.  Here one kernel does all of it per block of 16 lines without
// writing a physical field: the x-transforms are dense real DFT matrices
// applied on the FP64 matrix cores (v_mfma_f64_16x16x4f64),
//     phys = C  Einv^T      C    = [Re c_0..c_{M-1} | Im c_0..c_{M-1}]  (16 x 2M)
//     P    = p  Fwd^T       Fwd  = 2M x mgalx, rows cos / -sin over mgalx
// with Einv(x, k) = w_k cos(2 pi x k / mgalx), Einv(x, M+k) = -w_k sin(...),
// w_0 = 1 (Im c_0 ignored, as FFTW's c2r does), w_k = 2.  The products stay in
// registers between the two transforms.  At 192 x 64 modes a dense DFT costs
// about 7x the FLOPs of an FFT, but it runs on the matrix cores at a high
// fraction of their peak, where a strided FFT would be latency-bound; the
// whole product stage is ~8 GFLOP at 129^2 x 64 modes.
//
// Layout: fields are complex(8) (m, m, M), y fastest.  A workgroup owns 16
// consecutive y at one z: for a fixed mode those 16 values are contiguous.
#include <algorithm>
#include <cmath>

#include "advect.h"
#include "proj.h"

namespace duct {

namespace {
constexpr int kLines = 16;     // lines (consecutive y) per workgroup
constexpr int kLdp = kLines + 1;
constexpr int kSmem = (3 * 2 * kXprodMaxM * kLines > 3 * kXprodMaxX * kLdp) ? 3 * 2 * kXprodMaxM * kLines
                                                                           : 3 * kXprodMaxX * kLdp;
}  // namespace

__global__ __launch_bounds__(256, 2) void k_xprod(XprodArgs a)
{
    const int m = a.m, M = a.M, Q = 2 * M, X = a.X;
    const int yblocks = (m + kLines - 1) / kLines;
    const int z = blockIdx.x / yblocks, y0 = (blockIdx.x % yblocks) * kLines;
    const size_t mm = (size_t)m * m;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: scalar branches only
    const int l16 = lane & 15, l4 = lane >> 4;

    // LDS: the coefficient tiles [3][Q][16] (step B), then product tiles [3][X][kLdp] (step D)
    __shared__ double smem[kSmem];

    // Every phase issues its loads ahead of their use: the coefficients all at once, the DFT
    // fragments one k-step ahead. A loop that loads and consumes in the same iteration waits out
    // the full L2 / HBM latency on every step, which made this kernel latency-bound.
    // ---- A: coefficients of u, v, w for 16 lines -> coef[f][q][line]
    constexpr int kAIt = kXprodMaxM * kLines / 256;   // loads per thread and field
    {
        d2v cv[3][kAIt];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const d2v* src = a.in[f] + (size_t)z * m + y0;
#pragma unroll
            for (int it = 0; it < kAIt; ++it) {
                const int e = tid + 256 * it, k = e / kLines, l = e % kLines;
                const bool ok = e < M * kLines && y0 + l < m;
                cv[f][it] = src[ok ? (size_t)k * mm + l : 0];
                if (!ok) cv[f][it] = d2v{0.0, 0.0};
            }
        }
#pragma unroll
        for (int f = 0; f < 3; ++f)
#pragma unroll
            for (int it = 0; it < kAIt; ++it) {
                const int e = tid + 256 * it, k = e / kLines, l = e % kLines;
                if (e < M * kLines) {
                    smem[((size_t)f * Q + k) * kLines + l] = cv[f][it][0];
                    smem[((size_t)f * Q + M + k) * kLines + l] = cv[f][it][1];
                }
            }
    }
    __syncthreads();

    // ---- B: phys_f (16 x X) = coef_f (16 x Q) * Einv^T;  wave w owns x-tiles w, w+4, ...
    const int XT = X / 16;
    constexpr int kMaxXT = 3;                 // X <= 192: at most 3 x-tiles per wave
    d4v acc[3][kMaxXT];
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
        for (int s = 0; s < kMaxXT; ++s) acc[f][s] = d4v{0.0, 0.0, 0.0, 0.0};
    const double* Ei = a.Einv;                // frag: [(xt * (Q/4) + ks) * 64 + lane]
    const int KB = Q / 4;
    const double* Eit[kMaxXT];                // this wave's x-tiles (past XT: the last tile, unused)
#pragma unroll
    for (int s = 0; s < kMaxXT; ++s) {
        const int xt = wave + 4 * s;
        Eit[s] = Ei + (size_t)(xt < XT ? xt : XT - 1) * KB * 64 + lane;
    }
    double bcur[kMaxXT], bnext[kMaxXT], av[3], avn[3];
    // One k-step: prefetch the next step's operands into (xb, xa), then the products with (cb, ca).
    // Called twice per iteration with the buffers swapped (KB = M / 2 is even), so the prefetched
    // operands are never copied: a copy at the bottom of the loop would wait for the loads it rotates
    // in, and the compiler does not unroll loops holding MFMAs when that would need a remainder.
    auto stepB = [&](int ks, double (&cb)[kMaxXT], double (&ca)[3], double (&xb)[kMaxXT], double (&xa)[3]) {
        const int kn = ks + 1 < KB ? ks + 1 : ks;
#pragma unroll
        for (int s = 0; s < kMaxXT; ++s) xb[s] = Eit[s][(size_t)kn * 64];
#pragma unroll
        for (int f = 0; f < 3; ++f) xa[f] = smem[((size_t)f * Q + kn * 4 + l4) * kLines + l16];
#pragma unroll
        for (int s = 0; s < kMaxXT; ++s) {
            if (wave + 4 * s >= XT) continue;
#pragma unroll
            for (int f = 0; f < 3; ++f)
                acc[f][s] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[f], cb[s], acc[f][s], 0, 0, 0);
        }
    };
#pragma unroll
    for (int s = 0; s < kMaxXT; ++s) bcur[s] = Eit[s][0];
#pragma unroll
    for (int f = 0; f < 3; ++f) av[f] = smem[((size_t)f * Q + l4) * kLines + l16];
    for (int ks = 0; ks < KB; ks += 2) {
        stepB(ks, bcur, av, bnext, avn);
        stepB(ks + 1, bnext, avn, bcur, av);
    }
    __syncthreads();   // coefficient tiles are dead from here on

    // ---- C + D, two rounds of three products: write p (line, x) to LDS, then P = p Fwd^T
    // D layout: D[i][j] in lane j + 16 (i % 4), element i / 4 -> line 4 r + l4, x = 16 xt + l16
    const int QT2 = M / 16;                   // Re q-tiles (the Im tile of q-tile t is t + QT2)
    const double* Fw = a.Fwd;                 // frag: [(qt * (X/4) + ks) * 64 + lane]
    for (int round = 0; round < 2; ++round) {
#pragma unroll
        for (int s = 0; s < kMaxXT; ++s) {
            const int xt = wave + 4 * s;
            if (xt >= XT) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double u = acc[0][s][r], v = acc[1][s][r], w = acc[2][s][r];
                double p0, p1, p2;
                if (round == 0) { p0 = u * u; p1 = v * v; p2 = w * w; }
                else { p0 = u * v; p1 = u * w; p2 = v * w; }
                const size_t o = (size_t)(xt * 16 + l16) * kLdp + 4 * r + l4;
                smem[o] = p0;
                smem[(size_t)X * kLdp + o] = p1;
                smem[(size_t)2 * X * kLdp + o] = p2;
            }
        }
        __syncthreads();
        for (int t = wave; t < QT2; t += 4) {
            d4v o[3][2];
#pragma unroll
            for (int p = 0; p < 3; ++p) o[p][0] = o[p][1] = d4v{0.0, 0.0, 0.0, 0.0};
            const int KD = X / 4;
            const double* Fre = Fw + (size_t)t * KD * 64 + lane;
            const double* Fim = Fw + (size_t)(t + QT2) * KD * 64 + lane;
            double bv[2], bn[2], pv[3], pn[3];   // (re, im) fragments of Fwd; products of the lines
            auto stepD = [&](int ks, double (&cb)[2], double (&cp)[3], double (&xb)[2], double (&xp)[3]) {
                const int kn = ks + 1 < KD ? ks + 1 : ks;
                xb[0] = Fre[(size_t)kn * 64];
                xb[1] = Fim[(size_t)kn * 64];
#pragma unroll
                for (int p = 0; p < 3; ++p) xp[p] = smem[(size_t)p * X * kLdp + (size_t)(kn * 4 + l4) * kLdp + l16];
#pragma unroll
                for (int p = 0; p < 3; ++p) {
                    o[p][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(cp[p], cb[0], o[p][0], 0, 0, 0);
                    o[p][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(cp[p], cb[1], o[p][1], 0, 0, 0);
                }
            };
            bv[0] = Fre[0];
            bv[1] = Fim[0];
#pragma unroll
            for (int p = 0; p < 3; ++p) pv[p] = smem[(size_t)p * X * kLdp + (size_t)l4 * kLdp + l16];
            for (int ks = 0; ks < KD; ks += 2) {   // KD = X / 4 is even
                stepD(ks, bv, pv, bn, pn);
                stepD(ks + 1, bn, pn, bv, pv);
            }
            const int k = t * 16 + l16;
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                d2v* dst = a.out[round * 3 + p] + (size_t)k * mm + (size_t)z * m + y0;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int l = 4 * r + l4;
                    if (y0 + l < m) dst[l] = d2v{o[p][0][r], o[p][1][r]};
                }
            }
        }
        __syncthreads();
    }
}

hipError_t launch_xprod(const XprodArgs& a, hipStream_t st)
{
    if (a.M % 16 || a.M > kXprodMaxM || a.X % 16 || a.X > kXprodMaxX || a.m <= 0) return hipErrorInvalidValue;
    const int yblocks = (a.m + kLines - 1) / kLines;
    hipLaunchKernelGGL(k_xprod, dim3((unsigned)(yblocks * a.m)), dim3(256), 0, st, a);
    return hipGetLastError();
}

// Fragment layout of a B operand of v_mfma_f64_16x16x4f64 (B[k][j] from lane j + 16 k):
// frag[(jt * (K/4) + ks) * 64 + lane] = Bmat[ks*4 + lane/16][jt*16 + lane%16], Bmat(k, j) = src[j*ld + k]
static void frag_b(const std::vector<double>& src, int J, int K, std::vector<double>& frag)
{
    frag.assign((size_t)J * K, 0.0);
    for (int jt = 0; jt < J / 16; ++jt)
        for (int ks = 0; ks < K / 4; ++ks)
            for (int lane = 0; lane < 64; ++lane) {
                const int j = jt * 16 + (lane & 15), k = ks * 4 + (lane >> 4);
                frag[((size_t)jt * (K / 4) + ks) * 64 + lane] = src[(size_t)j * K + k];
            }
}

// ---------------------------------------------------------------------------
// fourx (fourier.F:1-115) on the reference's z-cut layouts, 16 (y, z) lines per
// workgroup, the same dense x-DFTs on the matrix cores:
//   fou(0:M-1, my, nz) complex, mode fastest;  phys(mgalx+2, my, nz) real, x fastest
//   (the two padding values of a physical line are written as zero).
__global__ __launch_bounds__(256, 2) void k_fourx(FourxArgs a)
{
    const int M = a.M, Q = 2 * M, X = a.X, XP = a.X + 2;
    const long lines = a.lines, L0 = (long)blockIdx.x * kLines;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    __shared__ double smem[kSmem];
    if (a.inverse) {
        for (int e = tid; e < M * kLines; e += 256) {          // coef[q][line], a line's modes contiguous
            const int l = e / M, k = e % M;
            d2v c = {0.0, 0.0};
            if (L0 + l < lines) c = a.fou[(L0 + l) * M + k];
            smem[(size_t)k * kLines + l] = c[0];
            smem[(size_t)(M + k) * kLines + l] = c[1];
        }
        __syncthreads();
        const int XT = X / 16;
        for (int xt = wave; xt < XT; xt += 4) {
            d4v acc = {0.0, 0.0, 0.0, 0.0};
            for (int ks = 0; ks < Q / 4; ++ks)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(smem[(size_t)(ks * 4 + l4) * kLines + l16],
                                                          a.Einv[((size_t)xt * (Q / 4) + ks) * 64 + lane], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long L = L0 + 4 * r + l4;
                if (L < lines) a.phys[L * XP + xt * 16 + l16] = acc[r];
            }
        }
        for (int l = tid; l < kLines * 2; l += 256) {            // the padding of each line
            const long L = L0 + l / 2;
            if (L < lines) a.phys[L * XP + X + (l & 1)] = 0.0;
        }
    } else {
        for (int e = tid; e < X * kLines; e += 256) {           // p[x][line], a line's x contiguous
            const int l = e / X, x = e % X;
            smem[(size_t)x * kLdp + l] = (L0 + l < lines) ? a.phys[(L0 + l) * XP + x] : 0.0;
        }
        __syncthreads();
        const int QT2 = M / 16;
        for (int t = wave; t < QT2; t += 4) {
            d4v ore = {0.0, 0.0, 0.0, 0.0}, oim = ore;
            for (int ks = 0; ks < X / 4; ++ks) {
                const double av = smem[(size_t)(ks * 4 + l4) * kLdp + l16];
                ore = __builtin_amdgcn_mfma_f64_16x16x4f64(av, a.Fwd[((size_t)t * (X / 4) + ks) * 64 + lane], ore, 0, 0, 0);
                oim = __builtin_amdgcn_mfma_f64_16x16x4f64(av, a.Fwd[((size_t)(t + QT2) * (X / 4) + ks) * 64 + lane],
                                                          oim, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long L = L0 + 4 * r + l4;
                if (L < lines) a.fou[L * M + t * 16 + l16] = d2v{ore[r], oim[r]};
            }
        }
    }
}

hipError_t launch_fourx(const FourxArgs& a, hipStream_t st)
{
    if (a.M % 16 || a.M > kXprodMaxM || a.X % 16 || a.X > kXprodMaxX || a.lines <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fourx, dim3((unsigned)((a.lines + kLines - 1) / kLines)), dim3(256), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// derivyz (chebyshev.f90:1-295): first or second derivative of Chebyshev
// coefficients along y or z, one line per thread.  From the top coefficient
// down, c_k = c_{k+2} + 2 k u_k, du_k = -c_{k+1}, du_0 = -c_1 / 2 (the minus
// sign: the coefficients belong to cos(pi j / N) = -x_j).  The second
// derivative repeats the recurrence on the first (through aux).
__device__ __forceinline__ void cheb_recurrence(const d2v* in, d2v* out, long stride, int m)
{
    d2v aari = {0.0, 0.0}, bri = aari, tri = aari;
    for (int kk = m - 1; kk >= 1; --kk) {
        const d2v cri = bri + (2.0 * (double)kk) * in[kk * stride];
        out[kk * stride] = -tri;
        bri = aari;
        tri = cri;
        aari = cri;
    }
    out[0] = -0.5 * tri;
}

__global__ __launch_bounds__(256) void k_derivyz(DerivyzArgs a)
{
    const long line = (long)blockIdx.x * 256 + threadIdx.x;
    const int m = a.m;
    if (line >= (long)a.nmodes * m) return;
    const long mode = line / m, t = line % m;
    const size_t base = (size_t)mode * m * m + (a.dirz ? (size_t)t : (size_t)t * m);
    const long stride = a.dirz ? m : 1;
    const d2v* u = a.u + base;
    d2v* w = a.aux + base;
    cheb_recurrence(u, w, stride, m);
    if (a.ider == 2) cheb_recurrence(w, w, stride, m);   // in place: each entry is read before it is written
    d2v* du = a.du + base;
    for (int k = 0; k < m; ++k) {
        const d2v v = a.scal * w[k * stride];
        du[k * stride] = a.accumulate ? du[k * stride] + v : v;
    }
}

hipError_t launch_derivyz(const DerivyzArgs& a, hipStream_t st)
{
    const long lines = (long)a.nmodes * a.m;
    if (lines <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_derivyz, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

void build_cheb_transforms(int m, std::vector<double>& fwd, std::vector<double>& inv)
{
    // tcheb_fftw (tchebw.f:62-150, nnyc = 0): c_k = s_k sum_j w_j cos(pi j k / N) u_j with
    // w_j = 1 at the ends, 2 inside, s_k = 1/N inside, 1/(2N) at the ends; u_j = sum_k c_k cos(pi j k / N).
    const int N = m - 1;
    fwd.assign((size_t)m * m, 0.0);
    inv.assign((size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j)
        for (int k = 0; k < m; ++k) {
            const long ph = ((long)j * k) % (2L * N);            // exact reduction of the angle
            const double c = std::cos(3.14159265358979323846 * (double)ph / (double)N);
            const double w = (j == 0 || j == N) ? 1.0 : 2.0;
            const double sk = (k == 0 || k == N) ? 0.5 / N : 1.0 / N;
            fwd[k + (size_t)j * m] = sk * w * c;                 // column-major: row k, column j
            inv[j + (size_t)k * m] = c;                          // row j, column k
        }
}

namespace {
// D and D2 = D D of the m CGL nodes in extended precision (column-major, D[i + k m]).
void colloc_ld(int m, std::vector<long double>& D, std::vector<long double>& D2)
{
    // D_ik = (c_i / c_k) (-1)^(i+k) / (x_i - x_k), x_j = -cos(pi j / N) (CHQZ (2.4.31), the matrix of
    // pseudomat.f90:90-107), evaluated in extended precision: x_i - x_k from the product
    // 2 sin(pi (i+k) / 2N) sin(pi (i-k) / 2N), which has no cancellation, the diagonal from the
    // zero row sums, D2 = D D accumulated in extended precision.
    const int N = m - 1;
    typedef long double X;
    const X pi = 3.141592653589793238462643383279502884L;
    D.assign((size_t)m * m, 0.0L);
    for (int i = 0; i < m; ++i) {
        X rs = 0.0L;
        for (int k = 0; k < m; ++k) {
            if (k == i) continue;
            const X ci = (i == 0 || i == N) ? 2.0L : 1.0L, ck = (k == 0 || k == N) ? 2.0L : 1.0L;
            const X dx = 2.0L * sinl(pi * (X)(i + k) / (2.0L * N)) * sinl(pi * (X)(i - k) / (2.0L * N));
            const X v = (ci / ck) * (((i + k) & 1) ? -1.0L : 1.0L) / dx;
            D[i + (size_t)k * m] = v;
            rs += v;
        }
        D[i + (size_t)i * m] = -rs;
    }
    D2.assign((size_t)m * m, 0.0L);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
            X acc = 0.0L;
            for (int k = 0; k < m; ++k) acc += D[i + (size_t)k * m] * D[k + (size_t)j * m];
            D2[i + (size_t)j * m] = acc;
        }
}
}  // namespace

void build_colloc_matrices(int m, std::vector<double>& d1, std::vector<double>& d2)
{
    // Both rounded once to double. The product of the double-rounded textbook entries loses ~1e-5
    // relative in second derivatives at m = 129; these matrices reproduce the interpolant's
    // derivatives to ~1e-10 of the rows' scale.
    std::vector<long double> D, D2;
    colloc_ld(m, D, D2);
    d1.assign((size_t)m * m, 0.0);
    d2.assign((size_t)m * m, 0.0);
    for (size_t e = 0; e < (size_t)m * m; ++e) {
        d1[e] = (double)D[e];
        d2[e] = (double)D2[e];
    }
}

namespace {
void fold_rows(int m, const std::vector<long double>& A, std::vector<double>& frag)
{
    // Rows y < he of D (or D2) with the contraction folded about the centre:
    //   Dp(y, k) = D(y, k) + D(y, m-1-k)  (k < ho),  Dp(y, c) = D(y, c)  (odd m, c = ho)
    //   Dm(y, k) = D(y, k) - D(y, m-1-k)  (k < ho)
    // each sum formed in extended precision and rounded once; zero-padded to RT 16 rows and KH
    // columns, in the fragment layout of build_frag_rect, Dp then Dm.
    const int he = (m + 1) / 2, ho = m / 2;
    const int KH = (he + 15) / 16 * 16, R = KH;
    std::vector<double> P((size_t)R * KH, 0.0), Mn((size_t)R * KH, 0.0);
    for (int y = 0; y < he; ++y)
        for (int k = 0; k < he; ++k) {
            const long double a = A[y + (size_t)k * m], b = A[y + (size_t)(m - 1 - k) * m];
            if (k < ho) {
                P[y + (size_t)k * R] = (double)(a + b);
                Mn[y + (size_t)k * R] = (double)(a - b);
            } else {
                P[y + (size_t)k * R] = (double)a;   // the centre column of odd m
            }
        }
    frag.assign((size_t)2 * R * KH, 0.0);
    build_frag_rect(P.data(), R, KH, R, frag.data());
    build_frag_rect(Mn.data(), R, KH, R, frag.data() + (size_t)R * KH);
}
}  // namespace

void build_colloc_folded(int m, std::vector<double>& fold1, std::vector<double>& fold2)
{
    std::vector<long double> D, D2;
    colloc_ld(m, D, D2);
    fold_rows(m, D, fold1);
    fold_rows(m, D2, fold2);
}

void build_xdft(int M, int X, std::vector<double>& einv, std::vector<double>& fwd)
{
    const int Q = 2 * M;
    const double tp = 2.0 * 3.14159265358979323846 / (double)X;
    std::vector<double> E((size_t)X * Q), F((size_t)Q * X);   // E[x][q] (j = x), F[q][x] (j = q)
    for (int x = 0; x < X; ++x)
        for (int k = 0; k < M; ++k) {
            const double ang = tp * (double)(((long)x * k) % X);   // reduce first: exact angle
            const double c = std::cos(ang), s = std::sin(ang);
            const double w = k == 0 ? 1.0 : 2.0;
            E[(size_t)x * Q + k] = w * c;
            E[(size_t)x * Q + M + k] = k == 0 ? 0.0 : -w * s;
            F[(size_t)k * X + x] = c / (double)X;
            F[(size_t)(M + k) * X + x] = k == 0 ? 0.0 : -s / (double)X;
        }
    frag_b(E, X, Q, einv);   // B[k = q][j = x]
    frag_b(F, Q, X, fwd);    // B[k = x][j = q]
}

}  // namespace duct
