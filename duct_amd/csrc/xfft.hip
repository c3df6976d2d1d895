// x-transforms of the pseudo-spectral step beyond the dense kernels' sizes (more than 64 modes or
// 192 physical points): batched real FFTs of rocFFT along x, with the fill of fourier.F:1-115
// (modes M..mgalx/2 zero, Im of mode 0 ignored as FFTW's c2r does), the products of mult.F in
// physical space and the pack of pack_fftw (modes 0..M-1, 1/mgalx).  These kernels are the
// element-wise stages around the transforms; all are HBM streams.
#include "advect.h"

namespace duct {

namespace {
constexpr int kXT = 256;
long xgrid(long total)
{
    const long b = (total + kXT - 1) / kXT;
    return b < 65536 ? b : 65536;
}
}  // namespace

// Mode-major fields [mode][line] (the solver's (m, m, M) layout, line = z m + y): W[f][h][line] =
// field f's mode h for h < M, 0 for M <= h < H; Im of mode 0 (and of the Nyquist mode) is 0.
__global__ __launch_bounds__(kXT) void k_xpad_modes(XpadArgs a)
{
    const long per = (long)a.H * a.lines, total = per * a.nf;
    for (long e = (long)blockIdx.x * kXT + threadIdx.x; e < total; e += (long)gridDim.x * kXT) {
        const int f = (int)(e / per);
        const long r = e % per;
        const int h = (int)(r / a.lines);
        const long line = r % a.lines;
        d2v v = {0.0, 0.0};
        if (h < a.M) v = a.in[f][(size_t)h * a.lines + line];
        if (h == 0 || 2 * h == a.X) v[1] = 0.0;
        a.out[(size_t)e] = v;
    }
}

// Line-major coefficients fou[line][mode] (fourx's fou(0:M-1, lines)) -> W[line][h], h < H.
__global__ __launch_bounds__(kXT) void k_xpad_lines(XpadArgs a)
{
    const long total = (long)a.H * a.lines;
    for (long e = (long)blockIdx.x * kXT + threadIdx.x; e < total; e += (long)gridDim.x * kXT) {
        const int h = (int)(e % a.H);
        const long line = e / a.H;
        d2v v = {0.0, 0.0};
        if (h < a.M) v = a.in[0][(size_t)line * a.M + h];
        if (h == 0 || 2 * h == a.X) v[1] = 0.0;
        a.out[(size_t)e] = v;
    }
}

// The six products uu, vv, ww, uv, uw, vw (mult.F) of three physical fields [line][x].
__global__ __launch_bounds__(kXT) void k_xmul6(const double* u, const double* v, const double* w, double* p,
                                                long n)
{
    for (long e = (long)blockIdx.x * kXT + threadIdx.x; e < n; e += (long)gridDim.x * kXT) {
        const double a = u[e], b = v[e], c = w[e];
        p[e] = a * a;
        p[n + e] = b * b;
        p[2 * n + e] = c * c;
        p[3 * n + e] = a * b;
        p[4 * n + e] = a * c;
        p[5 * n + e] = b * c;
    }
}

// pack_fftw: out[f][mode][line] = W[f][mode][line] / X for mode < M (mode-major), or
// out[line][mode] = W[line][mode] / X (line-major, one field).
__global__ __launch_bounds__(kXT) void k_xpack(XpackArgs a)
{
    const long per = (long)a.M * a.lines, total = per * a.nf;
    const double s = 1.0 / (double)a.X;
    for (long e = (long)blockIdx.x * kXT + threadIdx.x; e < total; e += (long)gridDim.x * kXT) {
        const int f = (int)(e / per);
        const long r = e % per;
        size_t src;
        if (a.line_major) {
            const long line = r / a.M;
            const int h = (int)(r % a.M);
            src = (size_t)line * a.H + h;
        } else {
            const int h = (int)(r / a.lines);
            const long line = r % a.lines;
            src = (size_t)f * a.H * a.lines + (size_t)h * a.lines + line;
        }
        const d2v v = a.in[src];
        a.out[f][(size_t)r] = d2v{v[0] * s, v[1] * s};
    }
}

// fourx's phys(X+2, lines): the two values past X of every line set to 0.
__global__ __launch_bounds__(kXT) void k_xzero_tail(double* phys, int X, long lines)
{
    for (long e = (long)blockIdx.x * kXT + threadIdx.x; e < 2 * lines; e += (long)gridDim.x * kXT)
        phys[(size_t)(e >> 1) * (X + 2) + X + (e & 1)] = 0.0;
}

hipError_t launch_xpad(const XpadArgs& a, bool line_major, hipStream_t st)
{
    const long total = (long)a.H * a.lines * (line_major ? 1 : a.nf);
    if (total <= 0) return hipSuccess;
    if (line_major) hipLaunchKernelGGL(k_xpad_lines, dim3((unsigned)xgrid(total)), dim3(kXT), 0, st, a);
    else hipLaunchKernelGGL(k_xpad_modes, dim3((unsigned)xgrid(total)), dim3(kXT), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_xmul6(const double* u, const double* v, const double* w, double* p, long n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_xmul6, dim3((unsigned)xgrid(n)), dim3(kXT), 0, st, u, v, w, p, n);
    return hipGetLastError();
}

hipError_t launch_xpack(const XpackArgs& a, hipStream_t st)
{
    const long total = (long)a.M * a.lines * a.nf;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_xpack, dim3((unsigned)xgrid(total)), dim3(kXT), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_xzero_tail(double* phys, int X, long lines, hipStream_t st)
{
    if (lines <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_xzero_tail, dim3((unsigned)xgrid(2 * lines)), dim3(kXT), 0, st, phys, X, lines);
    return hipGetLastError();
}

}  // namespace duct
