// Bandwidth-bound helpers that mirror the reference's GPU glue:
//   k_transpose   C = alpha * A^T, batched, FP64 or complex(8)
//                 (transpose_gpu_ / transposeCoalesced, mytranspose_cuda.cu:69-88,132-162;
//                  cxyz2yzx / cyzx2xyz, parallel.F:476-529)
//   k_invdiag     C = F ./ (ED2 - xlam)   (mat_invdiag_gpu_, mytranspose_cuda.cu:45-53,181-196)
//
// The reference kernels need both sizes to be multiples of 32 and drop the
// remainder silently (mytranspose_cuda.cu:141); these take any size, mask the
// edge tiles, and pad the LDS tile by one element against bank conflicts.
#include "aux.h"
#include "common.h"

namespace duct {

// Tile edge: 64 doubles or 32 complex(8) values (33 / 17 KB of LDS), read
// and written by a 256-thread block in kTile/kRows passes.
template <class T> struct TileOf { static constexpr int v = 64; };
template <> struct TileOf<d2v> { static constexpr int v = 32; };

template <class T>
__device__ __forceinline__ T scale(T v, double a) { return v * a; }

// A is Nx x Ny column-major (x fastest), C = A^T is Ny x Nx column-major.
template <class T>
__global__ __launch_bounds__(256) void k_transpose(const T* __restrict__ A, T* __restrict__ C,
                                                   int Nx, int Ny, long batch_stride, double alpha)
{
    constexpr int kTile = TileOf<T>::v, kRows = 256 / kTile;
    __shared__ T tile[kTile][kTile + 1];
    const size_t off = (size_t)blockIdx.z * batch_stride;
    A += off;
    C += off;
    const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    for (int r = ty; r < kTile; r += kRows) {
        const int x = x0 + tx, y = y0 + r;
        if (x < Nx && y < Ny) tile[r][tx] = A[(size_t)x + (size_t)y * Nx];
    }
    __syncthreads();
    for (int r = ty; r < kTile; r += kRows) {
        const int y = y0 + tx, x = x0 + r;  // C(y, x) = A(x, y)
        if (x < Nx && y < Ny) C[(size_t)y + (size_t)x * Ny] = scale(tile[tx][r], alpha);
    }
}

template <class T>
static hipError_t launch_t(const T* A, T* C, int Nx, int Ny, int nbatch, double alpha, hipStream_t st)
{
    constexpr int kTile = TileOf<T>::v;
    if (Nx <= 0 || Ny <= 0 || nbatch <= 0) return hipSuccess;
    dim3 grid((Nx + kTile - 1) / kTile, (Ny + kTile - 1) / kTile, nbatch);
    hipLaunchKernelGGL(k_transpose<T>, grid, dim3(256), 0, st, A, C, Nx, Ny, (long)Nx * Ny, alpha);
    return hipGetLastError();
}

hipError_t launch_transpose_f64(const double* A, double* C, int Nx, int Ny, int nbatch, double alpha, hipStream_t st)
{
    return launch_t<double>(A, C, Nx, Ny, nbatch, alpha, st);
}

hipError_t launch_transpose_c128(const d2v* A, d2v* C, int Nx, int Ny, int nbatch, double alpha, hipStream_t st)
{
    return launch_t<d2v>(A, C, Nx, Ny, nbatch, alpha, st);
}

// C(i,j) = F(i,j) / (ED2(i,j) - xlam) over n = Nx*Ny elements.
__global__ __launch_bounds__(256) void k_invdiag(const double* __restrict__ F, const double* __restrict__ ED2,
                                                 double* __restrict__ C, long n, double xlam)
{
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256)
        C[e] = F[e] / (ED2[e] - xlam);
}

hipError_t launch_invdiag(const double* F, const double* ED2, double* C, long n, double xlam, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_invdiag, dim3((unsigned)blocks), dim3(256), 0, st, F, ED2, C, n, xlam);
    return hipGetLastError();
}

}  // namespace duct
