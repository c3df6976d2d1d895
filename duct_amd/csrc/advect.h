// Advection-term products on the matrix cores (declarations).
#pragma once
#include <vector>

#include "common.h"

namespace duct {

constexpr int kXprodMaxM = 64;    // the dense kernels' largest complex x-mode count (2/3 rule: mx1 + 1)
constexpr int kXprodMaxX = 192;   // and physical x point count (mgalx); rocFFT past them

struct XprodArgs {
    const d2v* in[3];      // u, v, w: complex(8) (m, m, M)
    d2v* out[6];           // uu, vv, ww, uv, uw, vw: complex(8) (m, m, M)
    const double* Einv;    // inverse x-DFT, B-operand fragment layout (2M x X)
    const double* Fwd;     // forward x-DFT with 1/X, B-operand fragment layout (X x 2M)
    int m, M, X;
};
// M % 16 == 0, M <= 64, X % 16 == 0, X <= 192.
hipError_t launch_xprod(const XprodArgs& a, hipStream_t st);
// The same products for X = 192 and M <= 64 with the x-transforms as 64-point FFTs across the
// lanes of DPP rows (xlane.hip; Einv and Fwd unused); tab = build_xlane_table on the device.
hipError_t launch_xprod_lane(const XprodArgs& a, const d2v* tab, hipStream_t st);
void build_xlane_table(std::vector<double>& tab);

struct FourxArgs {
    d2v* fou;              // fou(0:M-1, lines) complex, mode fastest
    double* phys;          // phys(X+2, lines) real, x fastest
    const double* Einv;    // as XprodArgs
    const double* Fwd;
    long lines;            // my * nz
    int M, X, inverse;     // inverse: fou -> phys (fourx iopt >= 0)
};
hipError_t launch_fourx(const FourxArgs& a, hipStream_t st);

struct DerivyzArgs {
    const d2v* u;          // Chebyshev coefficients, complex(8) (m, m, nmodes)
    d2v* du;               // result (= or +=)
    d2v* aux;              // scratch of the same size (may be du when not accumulating)
    int m, nmodes, dirz, ider, accumulate;
    double scal;           // scal / aspect^ider for z
};
hipError_t launch_derivyz(const DerivyzArgs& a, hipStream_t st);
// Chebyshev transform matrices (column-major m x m): fwd = phys -> coefficients, inv = the inverse.
void build_cheb_transforms(int m, std::vector<double>& fwd, std::vector<double>& inv);
// The collocation first and second derivative matrices of the m CGL nodes, column-major m x m,
// correctly rounded from an extended-precision evaluation (the derivative products' matrices).
void build_colloc_matrices(int m, std::vector<double>& d1, std::vector<double>& d2);
// The rows y < ceil(m/2) of D (fold1) and of D2 (fold2) with the contraction folded about the
// centre node: [Dp | Dm] in fragment layout, each (16 ceil(ceil(m/2)/16))^2 (the derivative
// products of k_cline_fold on the even and odd parts of their input).
void build_colloc_folded(int m, std::vector<double>& fold1, std::vector<double>& fold2);
void build_xdft(int M, int X, std::vector<double>& einv, std::vector<double>& fwd);

// Sizes beyond the dense x-DFT kernels: batched rocFFT real transforms (capi.hip) between these
// element-wise stages (xfft.hip).  H = X/2 + 1 Hermitian modes of an X-point real line.
struct XpadArgs {
    const d2v* in[3];      // mode-major (m, m, M) fields, or one line-major fou[line][M]
    d2v* out;              // mode-major: [f][h][line]; line-major: [line][h]; h < H
    long lines;
    int nf, M, H, X;
};
// mode h < M copied, h >= M zero, Im of the zero mode (and of the Nyquist mode) zero
hipError_t launch_xpad(const XpadArgs& a, bool line_major, hipStream_t st);
// p[0..5] (each n doubles) = uu, vv, ww, uv, uw, vw
hipError_t launch_xmul6(const double* u, const double* v, const double* w, double* p, long n, hipStream_t st);
struct XpackArgs {
    const d2v* in;         // mode-major: [f][h][line] (nf fields); line-major: [line][h]
    d2v* out[6];           // mode-major: (m, m, M) fields; line-major: fou[line][M]
    long lines;
    int nf, M, H, X, line_major;
};
// out = in / X for the modes h < M
hipError_t launch_xpack(const XpackArgs& a, hipStream_t st);
// phys(X+2, lines): the two values past X of every line set to 0
hipError_t launch_xzero_tail(double* phys, int X, long lines, hipStream_t st);

}  // namespace duct
