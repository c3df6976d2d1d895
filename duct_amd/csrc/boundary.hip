// Small per-plane kernels around the projection passes:
//   k_gauge      Neumann k=0 gauge fix          solvediag.f90:161-206
//   k_boundary   Dirichlet zero / keep          helm3d.f90:48-60
//   k_bnd_cols, k_bnd_rows  Neumann edges + corners  solvediag.f90:251-322
//   k_bc_correct Dirichlet RHS correction       solvediag.f90:37-45
// All are bandwidth- or latency-bound and touch O(n) or O(n^2) values
// per plane; re and im of a complex plane are processed together as double2.
#include "common.h"
#include "proj.h"

namespace duct {

// Sum over the 64 lanes of a wave (all lanes active), the same value returned to every lane:
// DPP moves within rows of 16 lanes (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror, row_mirror)
// leave each row's sum in all its lanes; the four row sums are read from lanes 0, 16, 32, 48.
// These are VALU moves, where a shuffle (__shfl_xor) goes through the LDS crossbar and a chain
// of six of them per value made the reductions the bound of the Neumann edge kernels.
template <int CTRL>
__device__ __forceinline__ double dpp_row(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double read_lane(double v, int lane)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_row<0xB1>(v);    // quad_perm [1,0,3,2]: pairs
    v += dpp_row<0x4E>(v);    // quad_perm [2,3,0,1]: quads
    v += dpp_row<0x141>(v);   // row_half_mirror: 8 lanes
    v += dpp_row<0x140>(v);   // row_mirror: 16 lanes
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

__device__ __forceinline__ d2v wave_sum2(d2v v)
{
    return d2v{wave_sum(v[0]), wave_sum(v[1])};
}

// ---------------------------------------------------------------------------
// Gauge.  For a plane G = P^-1 F P^-T ./ den with G(a,a) = 0, set
//   G(a,a) = (wanted - sum_l P(a,l) sum_k P(a,k) G(l,k)) / P(a,a)^2
// so that (P G P^T)(a,a) = wanted.  One 1024-thread workgroup per listed
// plane: lanes run down the contiguous l index, the 16 waves split k.
constexpr int kAuxThreads = 1024;
constexpr int kAuxWaves = kAuxThreads / 64;

// The sums of the gauge kernels: lanes run down the contiguous l (LI steps of 64), the waves split
// k, and each wave takes its k values in chunks of KCH whose loads are all issued before any is
// summed; a loop that loaded and summed one term at a time left every load's latency exposed.
template <int LI>
struct GaugeShape {
    static constexpr int KCH = LI <= 2 ? 8 : (16 / LI >= 2 ? 16 / LI : 2);
};

template <int LI, bool F32>
__global__ __launch_bounds__(kAuxThreads) void k_gauge(GaugeArgs a, GaugeList list)
{
    constexpr int KCH = GaugeShape<LI>::KCH;
    const int plane = list.plane[blockIdx.x];
    const int NP = a.NP, n = a.g.n, io = a.gauge;
    const size_t base = (size_t)plane * NP * NP;
    const double* Gd = (const double*)a.G + base;
    const float* Gf = (const float*)a.G + base;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool divide = a.lam != nullptr;   // undivided G: divide each term as pass 2 would have
    double xlam = 0.0;
    if (divide) {
        const int fm = plane >> 1;
        xlam = xlambda_of(a.g.alambda[fm / a.g.nmodes], a.g.alp, a.g.mode0 + fm % a.g.nmodes);
    }
    double pl[LI], ll[LI];
#pragma unroll
    for (int h = 0; h < LI; ++h) {
        const int l = lane + 64 * h, lc = l < NP ? l : 0;
        pl[h] = a.prow[lc];
        ll[h] = divide ? a.lam[lc] : 0.0;
    }
    double part = 0.0;
    for (int k0 = wave; k0 < n; k0 += kAuxWaves * KCH) {
        double gv[KCH][LI], pk[KCH], lk[KCH];
#pragma unroll
        for (int t = 0; t < KCH; ++t) {
            const int k = k0 + kAuxWaves * t, kc = k < n ? k : 0;
            pk[t] = a.prow[kc];
            lk[t] = divide ? a.lam[kc] : 0.0;
#pragma unroll
            for (int h = 0; h < LI; ++h) {
                const int l = lane + 64 * h, lc = l < NP ? l : 0;
                gv[t][h] = F32 ? (double)Gf[(size_t)kc * NP + lc] : Gd[(size_t)kc * NP + lc];
            }
        }
#pragma unroll
        for (int t = 0; t < KCH; ++t) {
            const int k = k0 + kAuxWaves * t;
#pragma unroll
            for (int h = 0; h < LI; ++h) {
                const int l = lane + 64 * h;
                if (k < n && l < n) {
                    if (divide) {   // the (a,a) entry counts 0
                        if (!(l == io && k == io)) part += pl[h] * (gv[t][h] / (ll[h] + lk[t] - xlam)) * pk[t];
                    } else {
                        part += pl[h] * gv[t][h] * pk[t];
                    }
                }
            }
        }
    }
    __shared__ double red[kAuxWaves];
    part = wave_sum(part);
    if (lane == 0) red[wave] = part;
    __syncthreads();
    if (tid == 0) {
        double valore = 0.0;
        for (int w = 0; w < kAuxWaves; ++w) valore += red[w];
        double target = a.wanted;
        if (a.relative) {
            const int fm = plane >> 1, part = plane & 1;
            const int f = fm / a.g.nmodes, mode = fm % a.g.nmodes, m = a.g.m;
            const d2v u = a.g.out[f][(size_t)mode * m * m + (size_t)(io + 1) + (size_t)(io + 1) * m];
            target -= u[part];
        }
        const double v = (target - valore) / (a.pdiag * a.pdiag);
        if (F32) ((float*)a.G)[base + io + (size_t)io * NP] = (float)v;
        else ((double*)a.G)[base + io + (size_t)io * NP] = v;
    }
}

template <int LI>
static void launch_gauge_li(const GaugeArgs& a, const GaugeList& list, hipStream_t st)
{
    if (a.f32) hipLaunchKernelGGL((k_gauge<LI, true>), dim3(list.count), dim3(kAuxThreads), 0, st, a, list);
    else hipLaunchKernelGGL((k_gauge<LI, false>), dim3(list.count), dim3(kAuxThreads), 0, st, a, list);
}

hipError_t launch_gauge(const GaugeArgs& a, const GaugeList& list, hipStream_t st)
{
    if (list.count <= 0) return hipSuccess;
    switch (a.NP <= 64 ? 1 : a.NP / 64) {   // lanes cover the NP (<= 512) rows of a column
        case 1: launch_gauge_li<1>(a, list, st); break;
        case 2: launch_gauge_li<2>(a, list, st); break;
        case 4: launch_gauge_li<4>(a, list, st); break;
        case 8: launch_gauge_li<8>(a, list, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// r -= s p p^T, s = q^T r q, on one part (re or im) of each listed plane.
__global__ __launch_bounds__(kAuxThreads) void k_nullproj(NullProjArgs a, GaugeList list)
{
    const int plane = list.plane[blockIdx.x];
    const int fm = plane >> 1, part = plane & 1;
    const int f = fm / a.nmodes, mode = fm % a.nmodes, m = a.m, n = a.n;
    double* R = (double*)(a.r[f] + (size_t)mode * m * m) + part;   // stride 2 doubles
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double s = 0.0;
    for (int k = wave; k < n; k += kAuxWaves) {
        const double qk = a.q[k];
        for (int l = lane; l < n; l += 64)
            s += a.q[l] * R[2 * ((size_t)(l + 1) + (size_t)(k + 1) * m)] * qk;
    }
    __shared__ double red[kAuxWaves];
    __shared__ double tot;
    s = wave_sum(s);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    if (tid == 0) {
        double v = 0.0;
        for (int w = 0; w < kAuxWaves; ++w) v += red[w];
        tot = v;
    }
    __syncthreads();
    const double st = tot;
    double rmx = 0.0;
    for (int k = wave; k < n; k += kAuxWaves) {
        const double pk = a.p[k];
        for (int l = lane; l < n; l += 64) {
            double* x = R + 2 * ((size_t)(l + 1) + (size_t)(k + 1) * m);
            const double v = *x - st * a.p[l] * pk;
            *x = v;
            rmx = fmax(rmx, fabs(v));
        }
    }
    if (a.stats) {
        for (int off = 32; off > 0; off >>= 1) rmx = fmax(rmx, __shfl_xor(rmx, off, 64));
        if (lane == 0)
            atomicMax(a.stats + 2 * a.slot[f], (unsigned long long)__builtin_bit_cast(unsigned long long, rmx));
    }
}

hipError_t launch_nullproj(const NullProjArgs& a, const GaugeList& list, hipStream_t st)
{
    if (list.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_nullproj, dim3(list.count), dim3(kAuxThreads), 0, st, a, list);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Boundary of each complex plane after pass 4.  One 1024-thread workgroup per
// (field, mode).
//
// Neumann (solvediag.f90:251-322): with det = DC(1,1)DC(m,m) - DC(1,m)DC(m,1),
//   U(i,1) = (-DC(m,m) sb + DC(1,m) st)/det, U(i,m) = (DC(m,1) sb - DC(1,1) st)/det,
//   sb = sum_j DC(1,j) U(i,j), st = sum_j DC(m,j) U(i,j) over interior j,
// and the same along y for U(1,j), U(m,j).  The four corner right-hand sides
// ff(1..4) are linear in those edge values, so every thread that produces an
// edge value adds its share to ff directly and the block reduces ff at the
// end; the corners are AM1 * ff.
__global__ __launch_bounds__(kAuxThreads) void k_boundary(BoundaryArgs a)
{
    const FieldGroup& g = a.g;
    const int fm = blockIdx.x;
    const int f = fm / g.nmodes, mode = fm % g.nmodes;
    const int m = g.m;
    const size_t mm = (size_t)m * m;
    d2v* out = g.out[f] + (size_t)mode * mm;
    d2v* in = const_cast<d2v*>(g.in[f]) + (size_t)mode * mm;
    const int tid = threadIdx.x;
    // BND_DIR_ZERO, BND_DIR_KEEP, BND_OUT_ZERO (BND_NEUMANN runs k_bnd_cols + k_bnd_rows)
    const bool zero = a.mode != BND_DIR_KEEP;
    if (a.mode == BND_OUT_ZERO) in = out;   // leave the input field alone
    for (int e = tid; e < 4 * m; e += kAuxThreads) {
        const int side = e / m, t = e % m;
        size_t idx;
        if (side == 0) idx = (size_t)t;                            // z = 1 column
        else if (side == 1) idx = (size_t)t + (size_t)(m - 1) * m; // z = m
        else if (side == 2) idx = (size_t)t * m;                   // y = 1 row
        else idx = (size_t)(m - 1) + (size_t)t * m;                // y = m row
        d2v v = {0.0, 0.0};
        if (!zero) v = in[idx];
        if (zero) in[idx] = v;
        if (g.mode0 + mode == 0) v[1] = 0.0;   // int2d3d.f90:43: Im not solved at k=0
        out[idx] = v;
    }
}

// Neumann edges and corners (solvediag.f90:251-322) in two launches:
//  k_bnd_cols: kBndSplit workgroups per plane, each owning a range of interior columns j. They read
//    the plane's interior once between them. Each computes the edges U(1, j), U(m, j) of its columns
//    from the sums along y (sl_j = sum_i DC(1,i) U(i,j), sr_j with DC(m,:)), and its share of the sums
//    along z of every interior row (sb_i = sum_j DC(1,j) U(i,j), st_i), which go to a.part.
//  k_bnd_rows: one workgroup per plane adds the shares, writes the edges U(i, 1), U(i, m), and
//    sets the corners from AM1 and the sums ff over all four edges.
// A wave owns whole columns (its lanes run down all the rows), so each column needs one reduction
// across the lanes; the row shares of the four waves are added in LDS at the end. The loads of
// kBndCols columns are issued before any of them is summed.
namespace {
constexpr int kBndThreads = 256;
constexpr int kBndWaves = kBndThreads / 64;
constexpr int kBndColsMax = (kMaxM + kBndSplit - 1) / kBndSplit;   // columns per workgroup
constexpr int kBndCols = 2;                                        // columns per wave per batch of loads
}  // namespace

template <int RW>   // rows per lane: ceil((m - 2) / 64)
__global__ __launch_bounds__(kBndThreads) void k_bnd_cols(BoundaryArgs a)
{
    const FieldGroup& g = a.g;
    const int fm = blockIdx.x / kBndSplit, sp = blockIdx.x % kBndSplit;
    const int f = fm / g.nmodes, mode = fm % g.nmodes;
    const int m = g.m, nc = m - 2;
    d2v* out = g.out[f] + (size_t)mode * m * m;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int J = (nc + kBndSplit - 1) / kBndSplit;
    const int j0 = 1 + sp * J, j1 = min(m - 1, j0 + J);

    __shared__ double dc0[kMaxM], dcm[kMaxM];   // DC(1, :), DC(m, :)
    __shared__ d2v edge[2][kBndColsMax];        // U(1, j), U(m, j) of this workgroup's columns
    __shared__ d2v racc[2][kMaxM];              // the waves' row shares, added in turn
    const double* DC = a.DC;
    for (int t = tid; t < m; t += kBndThreads) {
        dc0[t] = DC[(size_t)t * m];
        dcm[t] = DC[(m - 1) + (size_t)t * m];
    }
    __syncthreads();
    const double DC00 = DC[0], DCmm = DC[(m - 1) + (size_t)(m - 1) * m];
    const double DC0m = DC[(size_t)(m - 1) * m], DCm0 = DC[m - 1];
    const double det = DC00 * DCmm - DC0m * DCm0;

    // rows of this lane: i = 1 + lane + 64 r
    double w0[RW], wm[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        const int i = 1 + lane + 64 * r;
        w0[r] = i < m - 1 ? dc0[i] : 0.0;   // zero weights past the interior
        wm[r] = i < m - 1 ? dcm[i] : 0.0;
    }
    d2v rb[RW], rt[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) rb[r] = rt[r] = d2v{0.0, 0.0};
    for (int jb = j0 + wave * kBndCols; jb < j1; jb += kBndWaves * kBndCols) {
        d2v u[kBndCols][RW];
#pragma unroll
        for (int c = 0; c < kBndCols; ++c) {
            const int j = jb + c < j1 ? jb + c : j0;
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                const int i = 1 + lane + 64 * r;
                u[c][r] = i < m - 1 ? out[(size_t)i + (size_t)j * m] : d2v{0.0, 0.0};
            }
        }
#pragma unroll
        for (int c = 0; c < kBndCols; ++c) {
            const int j = jb + c;
            if (j >= j1) break;   // uniform over the wave
            const double d0j = dc0[j], dmj = dcm[j];
            d2v sl = {0.0, 0.0}, sr = {0.0, 0.0};
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                sl += w0[r] * u[c][r];
                sr += wm[r] * u[c][r];
                rb[r] += d0j * u[c][r];
                rt[r] += dmj * u[c][r];
            }
            sl = wave_sum2(sl);
            sr = wave_sum2(sr);
            if (lane == 0) {
                edge[0][j - j0] = (-DCmm * sl + DC0m * sr) / det;
                edge[1][j - j0] = (DCm0 * sl - DC00 * sr) / det;
            }
        }
    }
    // the row shares of the four waves, added in LDS in wave order
    for (int w = 0; w < kBndWaves; ++w) {
        if (wave == w) {
#pragma unroll
            for (int r = 0; r < RW; ++r) {
                const int i = 1 + lane + 64 * r;
                if (i < m - 1) {
                    racc[0][i] = w ? racc[0][i] + rb[r] : rb[r];
                    racc[1][i] = w ? racc[1][i] + rt[r] : rt[r];
                }
            }
        }
        __syncthreads();
    }
    d2v* P = a.part + (size_t)(fm * kBndSplit + sp) * 2 * m;
    for (int i = 1 + tid; i < m - 1; i += kBndThreads) {
        P[i] = racc[0][i];
        P[m + i] = racc[1][i];
    }
    for (int j = j0 + tid; j < j1; j += kBndThreads) {
        out[(size_t)j * m] = edge[0][j - j0];
        out[(m - 1) + (size_t)j * m] = edge[1][j - j0];
    }
}

__global__ __launch_bounds__(kBndThreads) void k_bnd_rows(BoundaryArgs a)
{
    const FieldGroup& g = a.g;
    const int fm = blockIdx.x;
    const int f = fm / g.nmodes, mode = fm % g.nmodes;
    const int m = g.m;
    d2v* out = g.out[f] + (size_t)mode * m * m;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double* DC = a.DC;
    const double DC00 = DC[0], DCmm = DC[(m - 1) + (size_t)(m - 1) * m];
    const double DC0m = DC[(size_t)(m - 1) * m], DCm0 = DC[m - 1];
    const double det = DC00 * DCmm - DC0m * DCm0;
    const d2v* P = a.part + (size_t)fm * kBndSplit * 2 * m;
    d2v f1 = {0, 0}, f2 = {0, 0}, f3 = {0, 0}, f4 = {0, 0};
    for (int t = 1 + tid; t < m - 1; t += kBndThreads) {
        d2v pb[kBndSplit], pt[kBndSplit], el, er;
#pragma unroll
        for (int s = 0; s < kBndSplit; ++s) {
            pb[s] = P[(size_t)s * 2 * m + t];
            pt[s] = P[(size_t)s * 2 * m + m + t];
        }
        el = out[(size_t)t * m];            // U(1, t), U(m, t): the edges of k_bnd_cols
        er = out[(m - 1) + (size_t)t * m];
        const double d0 = DC[(size_t)t * m], dm = DC[(m - 1) + (size_t)t * m];
        d2v sb = {0.0, 0.0}, st = {0.0, 0.0};
#pragma unroll
        for (int s = 0; s < kBndSplit; ++s) {
            sb += pb[s];
            st += pt[s];
        }
        const d2v vb = (-DCmm * sb + DC0m * st) / det;
        const d2v vt = (DCm0 * sb - DC00 * st) / det;
        out[(size_t)t] = vb;                // U(t, 1)
        out[(size_t)t + (size_t)(m - 1) * m] = vt;   // U(t, m)
        f1 -= d0 * vb;
        f2 += dm * vb;
        f3 -= dm * vt;
        f4 -= d0 * vt;
        f1 -= d0 * el;
        f2 -= d0 * er;
        f3 -= dm * er;
        f4 += dm * el;
    }
    // corners: reduce ff over the block, U_corner = AM1 * ff
    __shared__ d2v fred[kBndWaves][4];
    f1 = wave_sum2(f1);
    f2 = wave_sum2(f2);
    f3 = wave_sum2(f3);
    f4 = wave_sum2(f4);
    if (lane == 0) {
        fred[wave][0] = f1;
        fred[wave][1] = f2;
        fred[wave][2] = f3;
        fred[wave][3] = f4;
    }
    __syncthreads();
    if (tid < 4) {
        d2v ff[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
        for (int w = 0; w < kBndWaves; ++w)
            for (int c = 0; c < 4; ++c) ff[c] += fred[w][c];
        const double* A = a.AM1;  // 4x4 column-major, row `tid`
        const d2v v = A[tid] * ff[0] + A[tid + 4] * ff[1] + A[tid + 8] * ff[2] + A[tid + 12] * ff[3];
        size_t idx;
        if (tid == 0) idx = 0;                                          // U(1,1)
        else if (tid == 1) idx = (size_t)(m - 1);                       // U(m,1)
        else if (tid == 2) idx = (size_t)(m - 1) + (size_t)(m - 1) * m; // U(m,m)
        else idx = (size_t)(m - 1) * m;                                 // U(1,m)
        out[idx] = v;
    }
}

size_t boundary_scratch_bytes(const FieldGroup& g)
{
    return (size_t)g.nf * g.nmodes * kBndSplit * 2 * g.m * sizeof(d2v);
}

hipError_t launch_boundary(const BoundaryArgs& a, hipStream_t st)
{
    const int planes = a.g.nf * a.g.nmodes;
    if (planes <= 0) return hipSuccess;
    if (a.mode != BND_NEUMANN) {
        hipLaunchKernelGGL(k_boundary, dim3(planes), dim3(kAuxThreads), 0, st, a);
        return hipGetLastError();
    }
    if (!a.part || a.g.m < 3 || a.g.m > kMaxM) return hipErrorInvalidValue;
    const dim3 grid((unsigned)planes * kBndSplit), block(kBndThreads);
    const int rw = (a.g.m - 2 + 63) / 64;   // rows per lane, <= 8 (m <= kMaxM)
    if (rw <= 1) hipLaunchKernelGGL(k_bnd_cols<1>, grid, block, 0, st, a);
    else if (rw <= 2) hipLaunchKernelGGL(k_bnd_cols<2>, grid, block, 0, st, a);
    else if (rw <= 4) hipLaunchKernelGGL(k_bnd_cols<4>, grid, block, 0, st, a);
    else hipLaunchKernelGGL(k_bnd_cols<8>, grid, block, 0, st, a);
    hipLaunchKernelGGL(k_bnd_rows, dim3(planes), dim3(kBndThreads), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// F(i,j) = U(i,j) - [D2(i,1) U(1,j) + D2(i,m) U(m,j)] - [D2(j,1) U(i,1) + D2(j,m) U(i,m)]
// on the interior, in place (the boundary values it reads are not written).
__global__ __launch_bounds__(256) void k_bc_correct(BcCorrectArgs a)
{
    const FieldGroup& g = a.g;
    const int fm = blockIdx.y;
    const int f = fm / g.nmodes, mode = fm % g.nmodes;
    const int m = g.m, n = g.n;
    d2v* U = const_cast<d2v*>(g.in[f]) + (size_t)mode * m * m;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)n * n) return;
    const int i = 1 + (int)(e % n), j = 1 + (int)(e / n);
    const double* D2 = a.D2b;
    const d2v bcontv = D2[i] * U[(size_t)j * m] + D2[i + m] * U[(m - 1) + (size_t)j * m];
    const d2v bconth = D2[j] * U[(size_t)i] + D2[j + m] * U[(size_t)i + (size_t)(m - 1) * m];
    const size_t idx = (size_t)i + (size_t)j * m;
    U[idx] = U[idx] - bcontv - bconth;
}

hipError_t launch_bc_correct(const BcCorrectArgs& a, hipStream_t st)
{
    const int planes = a.g.nf * a.g.nmodes;
    const long elems = (long)a.g.n * a.g.n;
    if (planes <= 0 || elems <= 0) return hipSuccess;
    dim3 grid((unsigned)((elems + 255) / 256), planes);
    hipLaunchKernelGGL(k_bc_correct, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace duct

namespace duct {

// ---------------------------------------------------------------------------
// out_int = in_int + E^T in_boundary for one complex plane per workgroup.
// With d0_i = DC(1,i), dm_i = DC(m,i) and the corner sums ff of the forward
// (solvediag.f90:280-302), U_corner = AM1 ff gives ffbar = AM1^T Ubar_corner;
// each edge value collects its own Ubar plus its share of ffbar; the edges are
// U(i,1) = (-DC(m,m) sb_i + DC(1,m) st_i)/det, U(i,m) = (DC(m,1) sb_i - DC(1,1) st_i)/det
// with sb_i = sum_j DC(1,j) U(i,j), st_i = sum_j DC(m,j) U(i,j) (and the same
// along y), so
//   out(i,j) = in(i,j) + DC(1,j) sbbar_i + DC(m,j) stbar_i + DC(1,i) slbar_j + DC(m,i) srbar_j.
__global__ __launch_bounds__(kAuxThreads) void k_boundary_adj(BoundaryAdjArgs a)
{
    const FieldGroup& g = a.g;
    const int fm = blockIdx.x;
    const int f = fm / g.nmodes, mode = fm % g.nmodes;
    const int m = g.m;
    const size_t mm = (size_t)m * m;
    const d2v* in = g.in[f] + (size_t)mode * mm;
    d2v* out = g.out[f] + (size_t)mode * mm;
    const int tid = threadIdx.x;
    const double* DC = a.DC;
    const double DC00 = DC[0], DCmm = DC[(m - 1) + (size_t)(m - 1) * m];
    const double DC0m = DC[(size_t)(m - 1) * m], DCm0 = DC[m - 1];
    const double det = DC00 * DCmm - DC0m * DCm0;

    __shared__ d2v ffb[4];
    __shared__ d2v sb[kMaxM], stt[kMaxM], sl[kMaxM], sr[kMaxM];
    if (tid < 4) {   // ffbar(c) = sum_r AM1(r,c) Ubar_corner(r)
        const size_t cidx[4] = {0, (size_t)(m - 1), (size_t)(m - 1) + (size_t)(m - 1) * m, (size_t)(m - 1) * m};
        d2v v = {0.0, 0.0};
        for (int r = 0; r < 4; ++r) v += a.AM1[r + 4 * tid] * in[cidx[r]];
        ffb[tid] = v;
    }
    __syncthreads();
    for (int i = 1 + tid; i < m - 1; i += kAuxThreads) {
        const double d0 = DC[(size_t)i * m], dm = DC[(m - 1) + (size_t)i * m];
        const d2v eb = in[(size_t)i] - d0 * ffb[0] + dm * ffb[1];
        const d2v et = in[(size_t)i + (size_t)(m - 1) * m] - dm * ffb[2] - d0 * ffb[3];
        const d2v el = in[(size_t)i * m] - d0 * ffb[0] + dm * ffb[3];
        const d2v er = in[(m - 1) + (size_t)i * m] - d0 * ffb[1] - dm * ffb[2];
        sb[i] = (-DCmm * eb + DCm0 * et) / det;
        stt[i] = (DC0m * eb - DC00 * et) / det;
        sl[i] = (-DCmm * el + DCm0 * er) / det;
        sr[i] = (DC0m * el - DC00 * er) / det;
    }
    __syncthreads();
    const int n = m - 2;
    for (long e = tid; e < (long)n * n; e += kAuxThreads) {
        const int i = 1 + (int)(e % n), j = 1 + (int)(e / n);   // i = y, j = z
        const size_t idx = (size_t)i + (size_t)j * m;
        out[idx] = in[idx] + DC[(size_t)j * m] * sb[i] + DC[(m - 1) + (size_t)j * m] * stt[i] +
                   DC[(size_t)i * m] * sl[j] + DC[(m - 1) + (size_t)i * m] * sr[j];
    }
}

hipError_t launch_boundary_adj(const BoundaryAdjArgs& a, hipStream_t st)
{
    const int planes = a.g.nf * a.g.nmodes;
    if (planes <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_boundary_adj, dim3(planes), dim3(kAuxThreads), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kAuxThreads) void k_gauge_adj(GaugeAdjArgs a, GaugeList list)
{
    const int plane = list.plane[blockIdx.x];
    const int fm = plane >> 1, part = plane & 1;
    const FieldGroup& g = a.g;
    const int f = fm / g.nmodes, mode = fm % g.nmodes, m = g.m, n = g.n, NP = a.NP, io = a.gauge;
    const double* U = (const double*)(g.in[f] + (size_t)mode * m * m) + part;   // stride 2 doubles
    double* G = a.G + (size_t)plane * NP * NP;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double s = 0.0;
    for (int k = wave; k < n; k += kAuxWaves) {
        const double pk = a.pcol[k];
        for (int l = lane; l < n; l += 64) s += a.pcol[l] * U[2 * ((size_t)(l + 1) + (size_t)(k + 1) * m)] * pk;
    }
    __shared__ double red[kAuxWaves];
    __shared__ double cc;
    s = wave_sum(s);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    if (tid == 0) {
        double v = 0.0;
        for (int w = 0; w < kAuxWaves; ++w) v += red[w];
        cc = v / (a.pdiag * a.pdiag);
    }
    __syncthreads();
    const double c = cc;
    const double xlam = xlambda_of(g.alambda[f], g.alp, g.mode0 + mode);
    for (int k = wave; k < n; k += kAuxWaves)
        for (int l = lane; l < n; l += 64) {
            if (l == io && k == io) continue;
            G[l + (size_t)k * NP] -= c * a.prow[l] * a.prow[k] / (a.lam[l] + a.lam[k] - xlam);
        }
}

hipError_t launch_gauge_adj(const GaugeAdjArgs& a, const GaugeList& list, hipStream_t st)
{
    if (list.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gauge_adj, dim3(list.count), dim3(kAuxThreads), 0, st, a, list);
    return hipGetLastError();
}

}  // namespace duct
