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
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;

    // LDS: the coefficient tiles [3][Q][16] (step B), then product tiles [3][X][kLdp] (step D)
    __shared__ double smem[kSmem];

    // ---- A: coefficients of u, v, w for 16 lines -> coef[f][q][line]
    for (int f = 0; f < 3; ++f) {
        const d2v* src = a.in[f] + (size_t)z * m + y0;
        for (int e = tid; e < M * kLines; e += 256) {
            const int k = e / kLines, l = e % kLines;
            d2v c = {0.0, 0.0};
            if (y0 + l < m) c = src[(size_t)k * mm + l];
            smem[((size_t)f * Q + k) * kLines + l] = c[0];
            smem[((size_t)f * Q + M + k) * kLines + l] = c[1];
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
    for (int ks = 0; ks < Q / 4; ++ks) {
        double av[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) av[f] = smem[((size_t)f * Q + ks * 4 + l4) * kLines + l16];
#pragma unroll
        for (int s = 0; s < kMaxXT; ++s) {
            const int xt = wave + 4 * s;
            if (xt >= XT) continue;
            const double bv = Ei[((size_t)xt * (Q / 4) + ks) * 64 + lane];
#pragma unroll
            for (int f = 0; f < 3; ++f)
                acc[f][s] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[f], bv, acc[f][s], 0, 0, 0);
        }
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
            for (int ks = 0; ks < X / 4; ++ks) {
                const double b_re = Fw[((size_t)t * (X / 4) + ks) * 64 + lane];
                const double b_im = Fw[((size_t)(t + QT2) * (X / 4) + ks) * 64 + lane];
#pragma unroll
                for (int p = 0; p < 3; ++p) {
                    const double av = smem[(size_t)p * X * kLdp + (size_t)(ks * 4 + l4) * kLdp + l16];
                    o[p][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b_re, o[p][0], 0, 0, 0);
                    o[p][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b_im, o[p][1], 0, 0, 0);
                }
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
