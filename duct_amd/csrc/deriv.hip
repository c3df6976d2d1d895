// Collocation derivative products on whole complex(8) planes:
//   y:  out(y, z) = beta out + scal sum_k D(y, k) u(k, z)      (D u, along the contiguous y)
//   z:  out(y, z) = beta out + scal sum_k D(z, k) u(y, k)      (u D^T, along z)
// for every plane of one or more fields (derivyz / transyz applied along y or z,
// chebyshev.f90:1-540, as a dense product with the matrices of the m CGL nodes).
//
// Work decomposition: a workgroup owns one 16-line slab of one plane of one op, the lines being
// the direction that is not contracted (16 z for y, 16 y for z), and stages that slab of u, all
// K = 16 ceil(m/16) contraction points, re and im, in LDS once. Its three waves then cover the
// NT = ceil(m/16) output tiles along the contracted direction, TPW tiles each, streaming the
// matrix fragments from L2. At m = 129 the padding costs (144/129)^3 = 1.4x the useful MFMA work,
// against 2.5x for 64 x 64 output blocks, and a slab is 37 KB of LDS, so four workgroups share a
// CU (static LDS sized for NP <= 48 TPW).
//
// One launch runs a list of independent ops (its own field, matrix, direction, scale and beta
// each): a single field's planes give about two waves per SIMD at m = 129, too few to hide the
// L2 latency of the fragments or to balance the CUs, so callers group the products that do not
// depend on each other.
#include "common.h"
#include "proj_lp.h"

namespace duct {

namespace {
constexpr int kWaves = 3;
constexpr int kThreads = 64 * kWaves;
}  // namespace

template <bool DIRZ, int TPW>
__device__ __forceinline__ void cline_slab(const d2v* __restrict__ U, d2v* O, const double* Dfrag, int m, int NP,
                                           int lb, double scal, double beta, double* slab)
{
    const int NT = NP / 16, K = NP, LDK = K + 1;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: scalar branches only
    const int l16 = lane & 15, l4 = lane >> 4;
    const int L0 = lb * 16;

    // stage the slab: y: u(k, z = L0 + l) is contiguous in k;  z: u(y = L0 + l, k) has y contiguous.
    // All loads of a thread are issued before its LDS stores, so their latencies overlap; they are
    // unconditional (points outside the plane read u(0, 0) and are stored as zeros).
    double* Sre = slab;
    double* Sim = slab + 16 * LDK;
    constexpr int IT = (16 * 16 * kWaves * TPW + kThreads - 1) / kThreads;
    const float invK = 1.0f / (float)K;       // e / K below is exact: e < 2^13, K <= 288
    d2v v[IT];
    bool ok[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = tid + it * kThreads;
        int l, k;
        if constexpr (DIRZ) {
            l = e & 15;
            k = e >> 4;
        } else {
            l = (int)(((float)e + 0.5f) * invK);
            k = e - l * K;
        }
        ok[it] = e < 16 * K && L0 + l < m && k < m;
        const size_t idx = DIRZ ? (size_t)(L0 + l) + (size_t)k * m : (size_t)k + (size_t)(L0 + l) * m;
        v[it] = U[ok[it] ? idx : 0];
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = tid + it * kThreads;
        if (e < 16 * K) {
            int l, k;
            if constexpr (DIRZ) {
                l = e & 15;
                k = e >> 4;
            } else {
                l = (int)(((float)e + 0.5f) * invK);
                k = e - l * K;
            }
            Sre[l * LDK + k] = ok[it] ? v[it][0] : 0.0;
            Sim[l * LDK + k] = ok[it] ? v[it][1] : 0.0;
        }
    }
    __syncthreads();

    // tiles of this wave: t = wave + kWaves s
    const d2v* Df = (const d2v*)Dfrag;
    d4v acc[TPW][2];
#pragma unroll
    for (int s = 0; s < TPW; ++s) acc[s][0] = acc[s][1] = d4v{0.0, 0.0, 0.0, 0.0};
    const int KP = K / 8;                         // fragment pairs of k-steps
    // Tiles past NT (when 3 TPW > NT) recompute the last tile and are not stored: the loop has
    // no branches, so the prefetch of the next pair's fragments overlaps this pair's products.
    const d2v* Dt[TPW];
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
        const int t = wave + kWaves * s;
        Dt[s] = Df + (size_t)(t < NT ? t : NT - 1) * KP * 64 + lane;
    }
    d2v df[TPW], dn[TPW];                         // this pair's fragments, the next pair's (prefetched)
    auto fetch = [&](int kp, d2v (&dst)[TPW]) {
#pragma unroll
        for (int s = 0; s < TPW; ++s) dst[s] = Dt[s][(size_t)kp * 64];
    };
    // One fragment pair: prefetch the next pair's fragments into `nxt`, then the products with `cur`.
    // The loop below calls it twice per iteration with the buffers swapped (KP = NP / 8 is even), so
    // the prefetched fragments are never copied: a copy at the bottom of the loop would wait for the
    // very loads it rotates in (and the compiler does not unroll loops holding MFMAs when that would
    // need a remainder).
    auto step = [&](int kp, d2v (&cur)[TPW], d2v (&nxt)[TPW]) {
        fetch(kp + 1 < KP ? kp + 1 : kp, nxt);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = (kp * 2 + h) * 4 + l4;
            const double ure = Sre[l16 * LDK + k], uim = Sim[l16 * LDK + k];
#pragma unroll
            for (int s = 0; s < TPW; ++s) {
                if constexpr (!DIRZ) {
                    // acc[i = z][j = y] = sum_k u(k, z) D(y, k): A = slab [z][k], B = D fragment [k][y]
                    acc[s][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ure, cur[s][h], acc[s][0], 0, 0, 0);
                    acc[s][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(uim, cur[s][h], acc[s][1], 0, 0, 0);
                } else {
                    // acc[i = z][j = y] = sum_k D(z, k) u(y, k): A = D fragment [z][k], B = slab [k][y]
                    acc[s][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur[s][h], ure, acc[s][0], 0, 0, 0);
                    acc[s][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur[s][h], uim, acc[s][1], 0, 0, 0);
                }
            }
        }
    };
    fetch(0, df);
    for (int kp = 0; kp < KP; kp += 2) {
        step(kp, df, dn);
        step(kp + 1, dn, df);
    }

    // epilogue: D[i][j] in lane j + 16 (i % 4), element i / 4; y runs along lanes (coalesced).
    // The accumulated outputs are all loaded before the first store: out and u may be the same
    // allocation for the compiler, which would otherwise serialise every load behind a store.
    d2v ov[TPW][4];
    if (beta != 0.0) {
#pragma unroll
        for (int s = 0; s < TPW; ++s) {
            const int t = wave + kWaves * s;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = 4 * e + l4;
                const int z = DIRZ ? 16 * t + i : L0 + i;
                const int y = DIRZ ? L0 + l16 : 16 * t + l16;
                ov[s][e] = (t < NT && z < m && y < m) ? O[(size_t)y + (size_t)z * m] : d2v{0.0, 0.0};
            }
        }
    }
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
        const int t = wave + kWaves * s;
        if (t >= NT) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = 4 * e + l4;
            const int z = DIRZ ? 16 * t + i : L0 + i;
            const int y = DIRZ ? L0 + l16 : 16 * t + l16;
            if (z < m && y < m) {
                d2v r = {scal * acc[s][0][e], scal * acc[s][1][e]};
                if (beta != 0.0) r += beta * ov[s][e];
                O[(size_t)y + (size_t)z * m] = r;
            }
        }
    }
}

// The same product for a matrix with J D J = s D (D: s = -1, D2: s = +1; J reverses the m CGL
// nodes), in half the MFMA work. With u split along the contraction into its even and odd parts,
//   ue(k) = (u(k) + u(m-1-k)) / 2,  uo(k) = (u(k) - u(m-1-k)) / 2   (k < floor(m/2)),  ue(c) = u(c),
// the output rows y < ceil(m/2) are Pe(y) + Po(y) with Pe = Dp ue and Po = Dm uo over the folded
// contraction (Dp, Dm of build_colloc_folded), and the mirrored rows are
//   out(m-1-y) = s (Pe(y) - Po(y)),
// since D ue has parity s and D uo parity -s. The slab holds ue and uo (re, im) for its 16 lines;
// a wave owns whole tiles of output rows with both parities, so it writes y and m-1-y itself.
template <bool DIRZ, int TPW>
__device__ __forceinline__ void cline_slab_fold(const d2v* __restrict__ U, d2v* O, const double* Dfrag, int m, int KH,
                                                int lb, double scal, double beta, double sgn, double* slab)
{
    const int RT = KH / 16, LDK = KH + 1;
    const int he = (m + 1) / 2, ho = m / 2;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int l16 = lane & 15, l4 = lane >> 4;
    const int L0 = lb * 16;

    // stage ue, uo: slab[(part * 2 + parity) * 16 * LDK + line * LDK + k]
    constexpr int IT = (16 * 16 * kWaves * TPW + kThreads - 1) / kThreads;
    const float invK = 1.0f / (float)KH;      // e / KH below is exact: e < 2^13, KH <= 144
    d2v va[IT], vb[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = tid + it * kThreads;
        int l, k;
        if constexpr (DIRZ) {
            l = e & 15;
            k = e >> 4;
        } else {
            l = (int)(((float)e + 0.5f) * invK);
            k = e - l * KH;
        }
        const bool line = e < 16 * KH && L0 + l < m;
        const bool oka = line && k < he, okb = line && k < ho;
        const int kb = m - 1 - k;
        const size_t ia = DIRZ ? (size_t)(L0 + l) + (size_t)k * m : (size_t)k + (size_t)(L0 + l) * m;
        const size_t ib = DIRZ ? (size_t)(L0 + l) + (size_t)kb * m : (size_t)kb + (size_t)(L0 + l) * m;
        va[it] = U[oka ? ia : 0];
        vb[it] = U[okb ? ib : 0];
        if (!oka) va[it] = d2v{0.0, 0.0};
        if (!okb) vb[it] = d2v{0.0, 0.0};
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = tid + it * kThreads;
        if (e < 16 * KH) {
            int l, k;
            if constexpr (DIRZ) {
                l = e & 15;
                k = e >> 4;
            } else {
                l = (int)(((float)e + 0.5f) * invK);
                k = e - l * KH;
            }
            d2v ue, uo;
            if (k < ho) {
                ue = 0.5 * (va[it] + vb[it]);
                uo = 0.5 * (va[it] - vb[it]);
            } else {
                ue = va[it];                       // the centre node of odd m, or zero padding
                uo = d2v{0.0, 0.0};
            }
            slab[(0 * 16 + l) * LDK + k] = ue[0];
            slab[(1 * 16 + l) * LDK + k] = uo[0];
            slab[(2 * 16 + l) * LDK + k] = ue[1];
            slab[(3 * 16 + l) * LDK + k] = uo[1];
        }
    }
    __syncthreads();

    const d2v* Dp = (const d2v*)Dfrag;
    const d2v* Dm = Dp + (size_t)KH * KH / 2;
    const int KP = KH / 8;
    d4v acc[TPW][4];                              // Pe re, Pe im, Po re, Po im
#pragma unroll
    for (int s = 0; s < TPW; ++s)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[s][q] = d4v{0.0, 0.0, 0.0, 0.0};
    const d2v* Pt[TPW];
    const d2v* Mt[TPW];
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
        const int t = wave + kWaves * s, tc = t < RT ? t : RT - 1;
        Pt[s] = Dp + (size_t)tc * KP * 64 + lane;
        Mt[s] = Dm + (size_t)tc * KP * 64 + lane;
    }
    d2v fp[TPW], fm[TPW], np[TPW], nm[TPW];
    // as in cline_slab: one fragment pair per call, called twice per iteration with the buffers
    // swapped (KP = KH / 8 is even)
    auto step = [&](int kp, d2v (&cp)[TPW], d2v (&cm)[TPW], d2v (&xp)[TPW], d2v (&xm)[TPW]) {
        const int kn = kp + 1 < KP ? kp + 1 : kp;
#pragma unroll
        for (int s = 0; s < TPW; ++s) {
            xp[s] = Pt[s][(size_t)kn * 64];
            xm[s] = Mt[s][(size_t)kn * 64];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = (kp * 2 + h) * 4 + l4;
            const double er = slab[(0 * 16 + l16) * LDK + k], orr = slab[(1 * 16 + l16) * LDK + k];
            const double ei = slab[(2 * 16 + l16) * LDK + k], oi = slab[(3 * 16 + l16) * LDK + k];
#pragma unroll
            for (int s = 0; s < TPW; ++s) {
                if constexpr (!DIRZ) {
                    acc[s][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(er, cp[s][h], acc[s][0], 0, 0, 0);
                    acc[s][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ei, cp[s][h], acc[s][1], 0, 0, 0);
                    acc[s][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(orr, cm[s][h], acc[s][2], 0, 0, 0);
                    acc[s][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(oi, cm[s][h], acc[s][3], 0, 0, 0);
                } else {
                    acc[s][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(cp[s][h], er, acc[s][0], 0, 0, 0);
                    acc[s][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(cp[s][h], ei, acc[s][1], 0, 0, 0);
                    acc[s][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(cm[s][h], orr, acc[s][2], 0, 0, 0);
                    acc[s][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(cm[s][h], oi, acc[s][3], 0, 0, 0);
                }
            }
        }
    };
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
        fp[s] = Pt[s][0];
        fm[s] = Mt[s][0];
    }
    for (int kp = 0; kp < KP; kp += 2) {
        step(kp, fp, fm, np, nm);
        step(kp + 1, np, nm, fp, fm);
    }

    // epilogue: D[i][j] in lane j + 16 (i % 4), element i / 4. Row r of the folded output (r < he)
    // is y = r (!DIRZ) or z = r (DIRZ); its mirror m-1-r exists for r < ho. All loads first.
    d2v o1[TPW][4], o2[TPW][4];
    if (beta != 0.0) {
#pragma unroll
        for (int s = 0; s < TPW; ++s) {
            const int t = wave + kWaves * s;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = 4 * e + l4;
                const int r = DIRZ ? 16 * t + i : 16 * t + l16;       // folded output row
                const int c = DIRZ ? L0 + l16 : L0 + i;               // line (the other direction)
                const bool v1 = t < RT && r < he && c < m, v2 = t < RT && r < ho && c < m;
                const size_t a1 = DIRZ ? (size_t)c + (size_t)r * m : (size_t)r + (size_t)c * m;
                const size_t a2 = DIRZ ? (size_t)c + (size_t)(m - 1 - r) * m : (size_t)(m - 1 - r) + (size_t)c * m;
                o1[s][e] = v1 ? O[a1] : d2v{0.0, 0.0};
                o2[s][e] = v2 ? O[a2] : d2v{0.0, 0.0};
            }
        }
    }
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
        const int t = wave + kWaves * s;
        if (t >= RT) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int i = 4 * e + l4;
            const int r = DIRZ ? 16 * t + i : 16 * t + l16;
            const int c = DIRZ ? L0 + l16 : L0 + i;
            if (c >= m) continue;
            const d2v pe = {acc[s][0][e], acc[s][1][e]}, po = {acc[s][2][e], acc[s][3][e]};
            if (r < he) {
                d2v v = scal * (pe + po);
                if (beta != 0.0) v += beta * o1[s][e];
                O[DIRZ ? (size_t)c + (size_t)r * m : (size_t)r + (size_t)c * m] = v;
            }
            if (r < ho) {
                d2v v = (scal * sgn) * (pe - po);
                if (beta != 0.0) v += beta * o2[s][e];
                O[DIRZ ? (size_t)c + (size_t)(m - 1 - r) * m : (size_t)(m - 1 - r) + (size_t)c * m] = v;
            }
        }
    }
}

template <int TPW>
__global__ __launch_bounds__(kThreads, 2) void k_cline_fold(ClineArgs a)
{
    __shared__ double slab[4 * 16 * (16 * kWaves * TPW + 1)];   // [part][parity][line][k], stride KH + 1
    const int NT = a.NP / 16, m = a.m;
    const int KH = ((m + 1) / 2 + 15) / 16 * 16;
    const long per_op = (long)a.nmodes * NT;
    const int o = (int)(blockIdx.x / per_op);
    const long r = blockIdx.x % per_op;
    const int mode = (int)(r / NT), lb = (int)(r % NT);
    const ClineOp& op = a.op[o];
    const d2v* U = op.in + (size_t)mode * m * m;
    d2v* O = op.out + (size_t)mode * m * m;
    const double sgn = op.fold < 0 ? -1.0 : 1.0;
    if (op.dirz)
        cline_slab_fold<true, TPW>(U, O, op.Dfrag, m, KH, lb, op.scal, op.beta, sgn, slab);
    else
        cline_slab_fold<false, TPW>(U, O, op.Dfrag, m, KH, lb, op.scal, op.beta, sgn, slab);
}

template <int TPW>
__global__ __launch_bounds__(kThreads, 2) void k_cline(ClineArgs a)
{
    __shared__ double slab[2 * 16 * (16 * kWaves * TPW + 1)];   // [part][line 0..15][k 0..K-1], stride LDK
    const int NT = a.NP / 16, m = a.m;
    const long per_op = (long)a.nmodes * NT;
    const int o = (int)(blockIdx.x / per_op);
    const long r = blockIdx.x % per_op;
    const int mode = (int)(r / NT), lb = (int)(r % NT);
    const ClineOp& op = a.op[o];
    const d2v* U = op.in + (size_t)mode * m * m;
    d2v* O = op.out + (size_t)mode * m * m;
    if (op.dirz)
        cline_slab<true, TPW>(U, O, op.Dfrag, m, a.NP, lb, op.scal, op.beta, slab);
    else
        cline_slab<false, TPW>(U, O, op.Dfrag, m, a.NP, lb, op.scal, op.beta, slab);
}

hipError_t launch_cline_ops(const ClineArgs& a, hipStream_t st)
{
    // the matrix fragments cover NP = 16 ceil(m/16) >= m; at most 6 tiles per wave (m <= 288)
    if (a.NP % 16 || a.NP < a.m || a.NP > 16 * kWaves * 6 || a.nops < 0 || a.nops > kMaxClineOps)
        return hipErrorInvalidValue;
    for (int i = 0; i < a.nops; ++i) {
        if (!a.op[i].in || !a.op[i].out || !a.op[i].Dfrag) return hipErrorInvalidValue;
        for (int j = 0; j < a.nops; ++j)   // every output is written by one op and read by no other
            if ((j != i && a.op[j].out == a.op[i].out) || a.op[j].in == a.op[i].out) return hipErrorInvalidValue;
    }
    for (int i = 1; i < a.nops; ++i)
        if ((a.op[i].fold != 0) != (a.op[0].fold != 0)) return hipErrorInvalidValue;
    if (a.nops == 0 || a.nmodes <= 0) return hipSuccess;
    const int NT = a.NP / 16;
    const long blocks = (long)a.nops * a.nmodes * NT;
    if (a.op[0].fold != 0) {   // folded rows: RT = ceil(ceil(m/2) / 16) tiles, TPW = ceil(RT / 3) <= 3
        const int RT = ((a.m + 1) / 2 + 15) / 16;
        switch ((RT + kWaves - 1) / kWaves) {
            case 1: hipLaunchKernelGGL(k_cline_fold<1>, dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
            case 2: hipLaunchKernelGGL(k_cline_fold<2>, dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
            case 3: hipLaunchKernelGGL(k_cline_fold<3>, dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch ((NT + kWaves - 1) / kWaves) {
        case 1: hipLaunchKernelGGL(k_cline<1>, dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
        case 2: hipLaunchKernelGGL(k_cline<2>, dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
        case 3: hipLaunchKernelGGL(k_cline<3>, dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
        case 4: hipLaunchKernelGGL(k_cline<4>, dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
        case 5: hipLaunchKernelGGL(k_cline<5>, dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
        case 6: hipLaunchKernelGGL(k_cline<6>, dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_cline(int dirz, const CollocArgs& a, hipStream_t st)
{
    if (a.g.nf < 1 || a.g.nf > kMaxClineOps) return hipErrorInvalidValue;
    ClineArgs c{};
    c.nops = a.g.nf;
    c.nmodes = a.g.nmodes;
    c.m = a.g.m;
    c.NP = a.NP;
    for (int f = 0; f < a.g.nf; ++f) c.op[f] = ClineOp{a.g.in[f], a.g.out[f], a.Dfrag, a.scal, a.beta, dirz, 0};
    return launch_cline_ops(c, st);
}

}  // namespace duct
