// FP64 projection passes in the even/odd (folded) eigenbasis.
//
// The interior operator D2B of the Dirichlet problem (pseudomat.f90:232-240)
// commutes with the reflection y -> -y of the CGL nodes, so its eigenvectors
// P(:,k) (solvediag.f90:51-80 uses PD and PDM1) are even or odd about the
// centre node.  With h_e = ceil(n/2) even and h_o = floor(n/2) odd columns,
// the half-sum / half-difference fold of a line x,
//     x_e(i) = (x(i) + x(n-1-i)) / 2,   x_o(i) = (x(i) - x(n-1-i)) / 2,   i < h_e,
// splits every product P^-1 x or P x into two independent products of size
// h_e and h_o.  On a plane both directions fold, so one n x n plane becomes
// four quarters (ee, eo, oe, oo) and each of the four eigenbasis products
// costs 2 (n/2)^3 * 2 instead of 2 n^3: half the FLOPs of the unfolded chain.
//
// Folded bases (built on the host from the caller's P, P^-1, see capi.hip):
//     Pe(i,k)   = (P(i,k) + P(n-1-i,k)) / 2          k even, i < h_e
//     Po(i,k)   = (P(i,k) - P(n-1-i,k)) / 2          k odd,  i < h_o
//     Pie(k,j)  = Pinv(k,j) + Pinv(k,n-1-j)  (j < h_o), Pinv(k,c) at the centre c
//     Pio(k,j)  = Pinv(k,j) - Pinv(k,n-1-j)
// stacked as one (2 NH) x NH matrix [A_e; A_o] (NH = padded h_e).
//
// Planar layout: plane p (Re or Im of a mode) holds four NH x NH column-major
// quarters, quarter (rp, cp) at ((p*4 + 2 rp + cp) NH) NH, rp = parity of the
// row (contraction) index, cp = parity of the column index.  A pass maps
//     Y(cp, rp) = (A_rp X(rp, cp))^T
// so the chain of solvediag (P^-1 F P^-T ./ den, then P G P^T) closes exactly
// as in the unfolded kernel (proj_f64.hip).  The 2-D fold of the complex
// right-hand side is the load stage of pass 1 (each thread reads the four
// mirror points of a quarter position), the 2-D unfold is the store stage of
// pass 4: one lane holds all four quarter values of its output position.
//
// Work decomposition (ProjShape<2 NH>): a wave owns 32 rows of each parity
// (2 x 2 M-tiles), so the unfold needs no exchange between waves; the data
// tile is staged in LDS per parity (2 x 16 k x NT, double buffered).  A
// group of 64 virtual columns = 16 folded columns x (cp, part) 4 N-tiles.
#include <algorithm>
#include <vector>

#include "common.h"
#include "proj.h"

namespace duct {

template <int NP2, int MODE>
__global__ __launch_bounds__(256, 2) void k_projf(ProjFoldArgs a)
{
    using S = ProjShape<NP2>;
    constexpr int NH = NP2 / 2;
    constexpr int NT = S::NT, TM = S::TM, LDB = S::LDB;
    constexpr int TH = TM / 2;                // M-tiles per parity per wave
    constexpr int NK = NH / S::KT;            // LDS stages over the folded contraction
    constexpr bool CIN = MODE == PROJ_C2P;
    constexpr bool COUT = MODE == PROJ_P2C;
    constexpr bool LDIV = MODE == PROJ_P2P_LDIV;
    static_assert(TH >= 1 && NH % 64 == 0, "folded kernel needs NH >= 64");

    __shared__ double Bs[2][2][NT * LDB];     // [buffer][row parity]
    __shared__ double lam_s[LDIV ? 2 * NH : 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave % S::WM, wn = wave / S::WM;
    const int l16 = lane & 15, l4 = lane >> 4;
    const long tile = blockIdx.x;
    const int rowblk = blockIdx.y;
    const FieldGroup& g = a.g;
    const int n = g.n, m = g.m;
    const int he = (n + 1) >> 1;
    const long qbase = tile * (NT / 4);       // first folded complex column of the tile

    // ---------------- B-tile loader ----------------
    constexpr int DPT = NT / 16;              // planar: doubles per thread per stage per parity
    constexpr int TPC = 16 / DPT;
    constexpr int PPT = NT / 64;              // C2P: (column, k) pairs per thread per stage

    // The B loads are unconditional (lanes with no column read a valid address and store zeros):
    // a load under a branch makes the compiler wait for every outstanding load, the next stage's
    // HBM reads included, before the first MFMA that follows the branch.
    const double* srcp[2] = {a.X, a.X};
    int v_dst = 0, kk0 = 0;
    bool src_ok = false;
    // LDIV: column l (parity cp) of the undivided pass-2 output, row k of parity rp,
    // is divided by lam_rp(k) + lam_cp(l) - xlambda as it is staged in LDS
    int ld_l = 0, ld_cp = 0;
    double ld_xlam = 0.0;

    const d2v* colA[PPT];                     // C2P: column j' and its mirror n-1-j'
    const d2v* colB[PPT];
    int vq[PPT], kq[PPT];
    bool okq[PPT], zimq[PPT];

    if constexpr (CIN) {
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
            const int pi = tid + 256 * p;
            const int ql = pi >> 4;
            kq[p] = pi & 15;
            vq[p] = (ql >> 4) * 64 + (ql & 15);
            const long q = qbase + ql;
            okq[p] = false;
            zimq[p] = false;
            colA[p] = colB[p] = g.in[0] + 1 + m;   // a valid address for lanes with no column
            if (q < a.ncols) {
                const long fm = q / NH;
                const int j = (int)(q % NH);
                const int f = (int)(fm / g.nmodes), mode = (int)(fm % g.nmodes);
                okq[p] = j < he;
                zimq[p] = (g.mode0 + mode) == 0;     // int2d3d.f90:43 skips Im at k=0
                const d2v* base = g.in[f] + (size_t)mode * m * m + 1;
                colA[p] = base + (size_t)(j + 1) * m;
                colB[p] = base + (size_t)(n - j) * m;
            }
        }
    } else {
        const int v = tid / TPC;
        kk0 = (tid % TPC) * DPT;
        v_dst = v;
        const int w = v & 63, nt = w >> 4;
        const long q = qbase + (v >> 6) * 16 + (w & 15);
        src_ok = q < a.ncols;
        if (src_ok) {
            const long fm = q / NH;
            const int l = (int)(q % NH);
            const long plane = fm * 2 + (nt & 1);
            const int cp = nt >> 1;
            srcp[0] = a.X + ((size_t)(plane * 4 + cp) * NH + l) * NH;
            srcp[1] = a.X + ((size_t)(plane * 4 + 2 + cp) * NH + l) * NH;
            if constexpr (LDIV) {
                const int f = (int)(fm / g.nmodes), mode = (int)(fm % g.nmodes);
                ld_l = l;
                ld_cp = cp;
                ld_xlam = xlambda_of(g.alambda[f], g.alp, g.mode0 + mode);
            }
        }
    }

    double ld_lamc = 0.0;
    if constexpr (LDIV) {
        for (int i = threadIdx.x; i < 2 * NH; i += 256) lam_s[i] = a.lam[i];
        if (src_ok) ld_lamc = a.lam[ld_cp * NH + ld_l] - ld_xlam;
        __syncthreads();
    }
    double breg[CIN ? 1 : 2][CIN ? 1 : DPT];
    d2v creg[PPT][4];
    bool okc[PPT];                            // C2P: the loaded column is real (else stored as zero)

    auto load_b = [&](int ks) {
        if constexpr (CIN) {
#pragma unroll
            for (int p = 0; p < PPT; ++p) {
                const int i = ks * S::KT + kq[p];
                const bool ok = okq[p] && i < he;
                const int ii = ok ? i : 0;
                creg[p][0] = colA[p][ii];
                creg[p][1] = colA[p][n - 1 - ii];
                creg[p][2] = colB[p][ii];
                creg[p][3] = colB[p][n - 1 - ii];
                okc[p] = ok;
            }
        } else {
            const int k0 = ks * S::KT + kk0;
#pragma unroll
            for (int rp = 0; rp < 2; ++rp)
#pragma unroll
                for (int e = 0; e < DPT; e += 2) {
                    const d2v v = *(const d2v*)(srcp[rp] + k0 + e);
                    breg[rp][e] = v[0];
                    breg[rp][e + 1] = v[1];
                }
        }
    };
    int st_k0 = 0;                            // first contraction index of the stage being stored
    auto store_b = [&](int buf) {
        if constexpr (CIN) {
#pragma unroll
            for (int p = 0; p < PPT; ++p) {
                const d2v z = {0.0, 0.0};
                const bool ok = okc[p];
                const d2v x00 = ok ? creg[p][0] : z, x10 = ok ? creg[p][1] : z;
                const d2v x01 = ok ? creg[p][2] : z, x11 = ok ? creg[p][3] : z;
                const d2v s0 = x00 + x10, d0 = x00 - x10, s1 = x01 + x11, d1 = x01 - x11;
                // (row parity, column parity): ee, eo, oe, oo, each /4
                const d2v q00 = 0.25 * (s0 + s1), q01 = 0.25 * (s0 - s1);
                const d2v q10 = 0.25 * (d0 + d1), q11 = 0.25 * (d0 - d1);
                const int k = kq[p];
                double* Be = Bs[buf][0];
                double* Bo = Bs[buf][1];
                const int v = vq[p];
                const bool zi = zimq[p];
                Be[(v + 0) * LDB + k] = q00[0];
                Be[(v + 16) * LDB + k] = zi ? 0.0 : q00[1];
                Be[(v + 32) * LDB + k] = q01[0];
                Be[(v + 48) * LDB + k] = zi ? 0.0 : q01[1];
                Bo[(v + 0) * LDB + k] = q10[0];
                Bo[(v + 16) * LDB + k] = zi ? 0.0 : q10[1];
                Bo[(v + 32) * LDB + k] = q11[0];
                Bo[(v + 48) * LDB + k] = zi ? 0.0 : q11[1];
            }
        } else {
#pragma unroll
            for (int rp = 0; rp < 2; ++rp)
#pragma unroll
                for (int e = 0; e < DPT; ++e) {
                    double v = src_ok ? breg[rp][e] : 0.0;
                    if constexpr (LDIV) {   // solvediag.f90:61-66
                        const int k = st_k0 + kk0 + e;
                        const int hh = rp ? (n >> 1) : he;
                        const int hc = ld_cp ? (n >> 1) : he;
                        if (ld_l < hc && k < hh)
                            v = div_eig(v, lam_s[rp * NH + k] + ld_lamc);
                    }
                    Bs[buf][rp][v_dst * LDB + kk0 + e] = v;
                }
        }
    };

    // ---------------- A fragments (global, fragment-major, (2 NH) x NH) ----------------
    const int idx0 = rowblk * (S::MT / 2) + wm * (S::WR / 2);   // row index within a parity
    const d2v* Ab[TM];
#pragma unroll
    for (int mt = 0; mt < TM; ++mt) {
        const int rp = mt / TH;
        const int mtg = (rp * NH + idx0 + (mt % TH) * 16) / 16;
        Ab[mt] = (const d2v*)a.Afrag + (size_t)mtg * (NH / 8) * 64 + lane;
    }
    d2v acur[TM][2];
    auto load_a = [&](int ks, d2v (&dst)[TM][2]) {
#pragma unroll
        for (int mt = 0; mt < TM; ++mt)
#pragma unroll
            for (int h = 0; h < 2; ++h) dst[mt][h] = Ab[mt][(size_t)(ks * 2 + h) * 64];
    };

    d4v acc[TM][4];
#pragma unroll
    for (int mt = 0; mt < TM; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = d4v{0.0, 0.0, 0.0, 0.0};

    load_b(0);
    load_a(0, acur);
    st_k0 = 0;
    store_b(0);
    __syncthreads();

#pragma unroll 2
    for (int ks = 0; ks < NK; ++ks) {
        const int buf = ks & 1;
        if (ks + 1 < NK) {
            load_b(ks + 1);
        }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            double bf[2][4];
#pragma unroll
            for (int rp = 0; rp < 2; ++rp)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    bf[rp][nt] = Bs[buf][rp][(wn * 64 + nt * 16 + l16) * LDB + kk * 4 + l4];
#pragma unroll
            for (int mt = 0; mt < TM; ++mt) {
                const double av = acur[mt][kk >> 1][kk & 1];
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bf[mt / TH][nt], acc[mt][nt], 0, 0, 0);
            }
            // This half of the A fragments is dead: fetch the next stage's. The planar passes fetch
            // unconditionally (after the last stage from the kFoldFragSlack padding), so the loop has
            // no branch around a load and the compiler's wait counts stay exact.
            if ((kk & 1) && (!(CIN || COUT) || ks + 1 < NK)) {
#pragma unroll
                for (int mt = 0; mt < TM; ++mt) acur[mt][kk >> 1] = Ab[mt][(size_t)((ks + 1) * 2 + (kk >> 1)) * 64];
            }
        }
        if (ks + 1 < NK) {
            st_k0 = (ks + 1) * S::KT;
            store_b(buf ^ 1);
        }
        __syncthreads();
    }

    // ---------------- epilogue ----------------
    const long q = qbase + wn * 16 + l16;     // this lane's folded column
    if (q >= a.ncols) return;
    const long fm = q / NH;
    const int l = (int)(q % NH);
    const int rbase = idx0 + l4;
    if constexpr (COUT) {
        // unfold: V[cp][rp] at (i' = l, j' = row); U(i,j) = sum, mirrors flip the odd terms
        if (l >= he) return;
        const int f = (int)(fm / g.nmodes), mode = (int)(fm % g.nmodes);
        const bool zim = (g.mode0 + mode) == 0;
        d2v* out = g.out[f] + (size_t)mode * m * m;
        const int i0 = l + 1, i1 = n - l;    // 1-based grid rows of y' and its mirror
#pragma unroll
        for (int mi = 0; mi < TH; ++mi)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int jr = rbase + mi * 16 + 4 * r;
                if (jr >= he) continue;
                d2v v[2][2];   // [cp][rp]
#pragma unroll
                for (int cp = 0; cp < 2; ++cp)
#pragma unroll
                    for (int rp = 0; rp < 2; ++rp) {
                        const int mt = rp * TH + mi;
                        v[cp][rp] = d2v{acc[mt][cp * 2][r], zim ? 0.0 : acc[mt][cp * 2 + 1][r]};
                    }
                const d2v a0 = v[0][0] + v[0][1], a1 = v[0][0] - v[0][1];
                const d2v b0 = v[1][0] + v[1][1], b1 = v[1][0] - v[1][1];
                const int j0 = jr + 1, j1 = n - jr;
                d2v* c0 = out + (size_t)j0 * m;
                d2v* c1 = out + (size_t)j1 * m;
                if (a.accumulate) {
                    c0[i0] += a0 + b0;
                    if (i1 != i0) c0[i1] += a0 - b0;
                    if (j1 != j0) {
                        c1[i0] += a1 + b1;
                        if (i1 != i0) c1[i1] += a1 - b1;
                    }
                } else {
                    c0[i0] = a0 + b0;
                    if (i1 != i0) c0[i1] = a0 - b0;
                    if (j1 != j0) {
                        c1[i0] = a1 + b1;
                        if (i1 != i0) c1[i1] = a1 - b1;
                    }
                }
            }
    } else {
        double xlam = 0.0;
        if constexpr (MODE == PROJ_P2P_DIV) {
            const int f = (int)((fm) / g.nmodes), mode = (int)(fm % g.nmodes);
            xlam = xlambda_of(g.alambda[f], g.alp, g.mode0 + mode);
        }
        const int hh[2] = {he, n >> 1};
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int cp = nt >> 1;
            const long plane = fm * 2 + (nt & 1);
#pragma unroll
            for (int mt = 0; mt < TM; ++mt) {
                const int rp = mt / TH;
                double* Y = a.Y + ((size_t)(plane * 4 + cp * 2 + rp) * NH) * NH + l;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rbase + (mt % TH) * 16 + 4 * r;
                    double v = acc[mt][nt][r];
                    if constexpr (MODE == PROJ_P2P_DIV) {
                        if (l < hh[cp] && row < hh[rp])
                            v = v / (a.lam[cp * NH + l] + a.lam[rp * NH + row] - xlam);
                        else
                            v = 0.0;
                    }
                    Y[(size_t)row * NH] = v;
                }
            }
        }
    }
}

template <int NP2>
static hipError_t launch_f(int mode, const ProjFoldArgs& a, hipStream_t st)
{
    using S = ProjShape<NP2>;
    const long tiles = (a.ncols * 4 + S::NT - 1) / S::NT;
    if (tiles <= 0) return hipSuccess;
    dim3 grid((unsigned)tiles, S::ROWBLOCKS), block(256);
    switch (mode) {
        case PROJ_C2P: hipLaunchKernelGGL((k_projf<NP2, PROJ_C2P>), grid, block, 0, st, a); break;
        case PROJ_P2P: hipLaunchKernelGGL((k_projf<NP2, PROJ_P2P>), grid, block, 0, st, a); break;
        case PROJ_P2P_DIV: hipLaunchKernelGGL((k_projf<NP2, PROJ_P2P_DIV>), grid, block, 0, st, a); break;
        case PROJ_P2C: hipLaunchKernelGGL((k_projf<NP2, PROJ_P2C>), grid, block, 0, st, a); break;
        case PROJ_P2P_LDIV: hipLaunchKernelGGL((k_projf<NP2, PROJ_P2P_LDIV>), grid, block, 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_proj_fold(int NH, int mode, const ProjFoldArgs& a, hipStream_t st)
{
    switch (NH) {
        case 64: return launch_f<128>(mode, a, st);
        case 128: return launch_f<256>(mode, a, st);
        case 256: return launch_f<512>(mode, a, st);
        default: return hipErrorInvalidValue;
    }
}

int fold_size(int n)
{
    const int he = (n + 1) / 2;
    if (he <= 32) return -1;          // a wave's 32 rows per parity must fit in NH >= 64
    if (he <= 64) return 64;
    if (he <= 128) return 128;
    if (he <= 256) return 256;
    return -1;
}

namespace {
// In-place LU with partial pivoting of an h x h column-major matrix; false if a pivot is 0.
bool lu_factor(std::vector<double>& A, int h, std::vector<int>& piv)
{
    piv.resize(h);
    for (int k = 0; k < h; ++k) {
        int p = k;
        for (int i = k + 1; i < h; ++i)
            if (std::fabs(A[i + (size_t)k * h]) > std::fabs(A[p + (size_t)k * h])) p = i;
        piv[k] = p;
        if (A[p + (size_t)k * h] == 0.0) return false;
        if (p != k)
            for (int j = 0; j < h; ++j) std::swap(A[k + (size_t)j * h], A[p + (size_t)j * h]);
        const double d = A[k + (size_t)k * h];
        for (int i = k + 1; i < h; ++i) A[i + (size_t)k * h] /= d;
        for (int j = k + 1; j < h; ++j) {
            const double t = A[k + (size_t)j * h];
            if (t != 0.0)
                for (int i = k + 1; i < h; ++i) A[i + (size_t)j * h] -= A[i + (size_t)k * h] * t;
        }
    }
    return true;
}

void lu_solve(const std::vector<double>& A, int h, const std::vector<int>& piv, double* b)
{
    for (int k = 0; k < h; ++k) std::swap(b[k], b[piv[k]]);
    for (int k = 0; k < h; ++k)
        for (int i = k + 1; i < h; ++i) b[i] -= A[i + (size_t)k * h] * b[k];
    for (int k = h - 1; k >= 0; --k) {
        b[k] /= A[k + (size_t)k * h];
        for (int i = 0; i < k; ++i) b[i] -= A[i + (size_t)k * h] * b[k];
    }
}
}  // namespace

int fold_matrices(const double* P, const double* Pinv, const double* lam, int n, int NH, double* Af, double* Ab,
                  double* lamf, int* pos, const double* op)
{
    const int he = (n + 1) / 2, ho = n / 2;
    // parity of every eigenvector: its even energy fraction; the he most even ones are the even
    // ones.  A pair of eigenvalues of the two parities that nearly coincide may come back from
    // the eigensolver as two mixtures of the pair's even and odd vectors: each then has a large
    // share of both, and this split gives the pair one of each.
    std::vector<double> fe(n), fmin(n);
    for (int k = 0; k < n; ++k) {
        double se = 0.0, so = 0.0;
        for (int i = 0; i < n; ++i) {
            const double a = P[i + (size_t)k * n], b = P[(n - 1 - i) + (size_t)k * n];
            se += (a + b) * (a + b);
            so += (a - b) * (a - b);
        }
        const double t = se + so;
        if (!(t > 0.0)) return -1;
        fe[k] = se / t;
        fmin[k] = std::min(se, so) / t;
    }
    std::vector<int> order(n);
    for (int k = 0; k < n; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return fe[x] > fe[y]; });
    std::vector<char> even(n, 0);
    for (int r = 0; r < he; ++r) even[order[r]] = 1;
    std::vector<int> ev, od;
    bool mixed = false;
    for (int k = 0; k < n; ++k) {
        (even[k] ? ev : od).push_back(k);
        if ((even[k] && fe[k] < 0.5) || (!even[k] && fe[k] > 0.5)) mixed = true;   // a balanced split
        if (fmin[k] > 1e-12) mixed = true;
    }
    if (pos) {
        for (int r = 0; r < he; ++r) pos[ev[r]] = r;
        for (int r = 0; r < ho; ++r) pos[od[r]] = -1 - r;
    }
    const int R = 2 * NH;
    std::fill(Af, Af + (size_t)R * NH, 0.0);
    std::fill(Ab, Ab + (size_t)R * NH, 0.0);
    for (int c = 0; c < 2; ++c) {
        const std::vector<int>& ks = c ? od : ev;
        const int h = c ? ho : he;
        const double s = c ? -1.0 : 1.0;
        for (int r = 0; r < h; ++r) {
            const int k = ks[r];
            for (int j = 0; j < h; ++j) {
                const int jm = n - 1 - j;
                // P^-1 rows: Pinv(k,j) +/- Pinv(k,n-1-j), the centre counted once
                const double vi = (j == jm) ? Pinv[k + (size_t)j * n] : Pinv[k + (size_t)j * n] + s * Pinv[k + (size_t)jm * n];
                Af[(c * NH + r) + (size_t)j * R] = vi;
                // P columns as rows of the left factor: Pfold(j, k) = (P(j,k) +/- P(n-1-j,k)) / 2
                Ab[(c * NH + j) + (size_t)r * R] = 0.5 * (P[j + (size_t)k * n] + s * P[jm + (size_t)k * n]);
            }
            lamf[c * NH + r] = lam[k];
        }
        for (int r = h; r < NH; ++r) lamf[c * NH + r] = 0.0;
    }
    if (!mixed) return 0;
    // Mixed vectors: their parity parts span the pair's even and odd eigenvectors, so the
    // projected columns Ab are still eigenvectors of the folded blocks, but the caller's P^-1
    // rows no longer match them.  The folded inverse is then recomputed from Ab, and the
    // eigenvalue of each mixed vector is its Rayleigh quotient on the folded block of the
    // centrosymmetric part (op + J op J) / 2 of the interior operator.
    if (!op) return -1;
    for (int c = 0; c < 2; ++c) {
        const std::vector<int>& ks = c ? od : ev;
        const int h = c ? ho : he;
        const double sg = c ? -1.0 : 1.0;
        std::vector<double> V((size_t)h * h), Sb((size_t)h * h);
        for (int r = 0; r < h; ++r)
            for (int j = 0; j < h; ++j) V[j + (size_t)r * h] = Ab[(c * NH + j) + (size_t)r * R];
        auto Sv = [&](int i, int j) { return 0.5 * (op[i + (size_t)j * n] + op[(n - 1 - i) + (size_t)(n - 1 - j) * n]); };
        for (int j = 0; j < h; ++j)
            for (int i = 0; i < h; ++i) {
                const int jm = n - 1 - j;
                Sb[i + (size_t)j * h] = (j == jm) ? Sv(i, j) : Sv(i, j) + sg * Sv(i, jm);
            }
        for (int r = 0; r < h; ++r) {
            const int k = ks[r];
            if (fmin[k] <= 1e-12 && ((c == 0) == (fe[k] > 0.5))) continue;
            double num = 0.0, den = 0.0;
            for (int i = 0; i < h; ++i) {
                double sx = 0.0;
                for (int j = 0; j < h; ++j) sx += Sb[i + (size_t)j * h] * V[j + (size_t)r * h];
                num += V[i + (size_t)r * h] * sx;
                den += V[i + (size_t)r * h] * V[i + (size_t)r * h];
            }
            if (!(den > 0.0)) return -1;
            lamf[c * NH + r] = num / den;
        }
        std::vector<double> LU = V;
        std::vector<int> piv;
        if (!lu_factor(LU, h, piv)) return -1;
        std::vector<double> col(h);
        for (int j = 0; j < h; ++j) {       // column j of V^-1
            std::fill(col.begin(), col.end(), 0.0);
            col[j] = 1.0;
            lu_solve(LU, h, piv, col.data());
            for (int r = 0; r < h; ++r) Af[(c * NH + r) + (size_t)j * R] = col[r];
        }
    }
    return 0;
}

int build_fold_bases(const double* P, const double* Pinv, const double* lam, int n, int NH,
                     double* fwd, double* bwd, double* lamf, int* pos, const double* op)
{
    const int R = 2 * NH;
    std::vector<double> Af((size_t)R * NH), Ab((size_t)R * NH);
    if (fold_matrices(P, Pinv, lam, n, NH, Af.data(), Ab.data(), lamf, pos, op)) return -1;
    build_frag_rect(Af.data(), R, NH, R, fwd);
    build_frag_rect(Ab.data(), R, NH, R, bwd);
    return 0;
}

// frag[((mt*(K/8) + kp)*64 + lane)*2 + e] = A[mt*16 + lane%16][kp*8 + e*4 + lane/16]
void build_frag_rect(const double* A, int R, int K, int lda, double* frag)
{
    for (int mt = 0; mt < R / 16; ++mt)
        for (int kp = 0; kp < K / 8; ++kp)
            for (int lane = 0; lane < 64; ++lane)
                for (int e = 0; e < 2; ++e) {
                    const int i = mt * 16 + (lane & 15);
                    const int k = kp * 8 + e * 4 + (lane >> 4);
                    frag[(((size_t)mt * (K / 8) + kp) * 64 + lane) * 2 + e] = A[(size_t)i + (size_t)k * lda];
                }
}

}  // namespace duct
