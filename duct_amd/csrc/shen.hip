// Batched symmetric tridiagonal solves of the Shen-Legendre Galerkin method.
//
// After the y-diagonalisation, row p of the coefficient plane obeys
//     (lam_p (xlambda M + S) + M) v = h        (solV, solveshen.f90:335-471)
// and M couples only z indices k and k+2, so each row splits into an even and an
// odd tridiagonal system.  The Neumann y-constant row (b_0 = 0) obeys
// (S + xlambda M) v = h; at xlambda = 0 its z-constant coefficient is pinned to 0.
// One thread solves one (mode, p, parity) system for Re and Im together with the
// Thomas recurrence of symtri (solveshen.f90:515-546); consecutive threads take
// consecutive p, so every step of the sweep is a coalesced row access.
#include "shen.h"

namespace duct {

__global__ __launch_bounds__(256) void k_shen_tri(ShenTriArgs a)
{
    const int m = a.m, n = m - 2;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)a.nmodes * 2 * n) return;
    const int p = (int)(t % n), par = (int)((t / n) & 1), mode = (int)(t / (2L * n));
    const double alpha = xlambda_of(a.alambda, a.alp, a.mode0 + mode);
    const bool yconst = a.neumann && p == 0;
    const bool pin = yconst && par == 0 && alpha <= 1.0e-10;
    const bool zim = a.mode0 + mode == 0;                  // int2d3d: Im not solved at k = 0
    const double rp = a.lam[p];
    const double ce = yconst ? alpha : rp * alpha + 1.0;    // scale of the off-diagonal M_{k,k+2}
    d2v* V = a.V + (size_t)mode * m * m + p;
    double* G = a.gam + (size_t)mode * m * m + p;
    auto diag = [&](int k) {
        return yconst ? a.sd[k] + alpha * a.md[k] : rp * (alpha * a.md[k] + a.sd[k]) + a.md[k];
    };
    int k = par;
    if (k >= n) return;
    double bet = pin ? 1.0 : diag(k);
    d2v r = pin ? d2v{0.0, 0.0} : V[(size_t)k * m];
    if (zim) r[1] = 0.0;
    d2v x = r / bet;
    V[(size_t)k * m] = x;
    int last = k;
    for (k = par + 2; k < n; k += 2) {
        const double e = ce * a.mu[k - 2];
        const double g = e / bet;
        G[(size_t)k * m] = g;
        bet = diag(k) - e * g;
        d2v rk = V[(size_t)k * m];
        if (zim) rk[1] = 0.0;
        x = (rk - e * x) / bet;
        V[(size_t)k * m] = x;
        last = k;
    }
    d2v xn = x;
    for (k = last - 2; k >= par; k -= 2) {
        d2v xk = V[(size_t)k * m];
        xk -= G[(size_t)(k + 2) * m] * xn;
        V[(size_t)k * m] = xk;
        xn = xk;
    }
}

hipError_t launch_shen_tri(const ShenTriArgs& a, hipStream_t st)
{
    const long total = (long)a.nmodes * 2 * (a.m - 2);
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_shen_tri, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace duct
