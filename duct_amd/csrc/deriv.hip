// Collocation derivative products on whole complex(8) planes, one direction per launch:
//   y:  out(y, z) = beta out + scal sum_k D(y, k) u(k, z)      (D u, along the contiguous y)
//   z:  out(y, z) = beta out + scal sum_k D(z, k) u(y, k)      (u D^T, along z)
// for every plane of every field of a FieldGroup (derivyz / transyz applied along y or z,
// chebyshev.f90:1-540, as a dense product with the matrices of the m CGL nodes).
//
// Work decomposition: a workgroup owns one 16-line slab of one plane, the lines being the
// direction that is not contracted (16 z for y, 16 y for z), and stages that slab of u,
// all K = 16 ceil(m/16) contraction points, re and im, in LDS once. Its three waves then
// cover the NT = ceil(m/16) output tiles along the contracted direction, TPW tiles each,
// streaming the matrix fragments from L2. At m = 129 the padding costs (144/129)^3 = 1.4x
// the useful MFMA work, against 2.5x for 64 x 64 output blocks, and a slab is 37 KB of LDS,
// so four workgroups share a CU (static LDS sized for NP <= 48 TPW).
#include "common.h"
#include "proj_lp.h"

namespace duct {

namespace {
constexpr int kWaves = 3;
constexpr int kThreads = 64 * kWaves;
}  // namespace

template <bool DIRZ, int TPW>
__global__ __launch_bounds__(kThreads, 2) void k_cline(CollocArgs a)
{
    __shared__ double slab[2 * 16 * (16 * kWaves * TPW + 1)];   // [part][line 0..15][k 0..K-1], stride LDK
    const FieldGroup& g = a.g;
    const int m = g.m, NT = a.NP / 16, K = a.NP, LDK = K + 1;
    const long plane = blockIdx.x / NT;
    const int lb = (int)(blockIdx.x % NT);        // 16-line block along the uncontracted direction
    const int f = (int)(plane / g.nmodes), mode = (int)(plane % g.nmodes);
    const d2v* U = g.in[f] + (size_t)mode * m * m;
    d2v* O = g.out[f] + (size_t)mode * m * m;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: scalar branches only
    const int l16 = lane & 15, l4 = lane >> 4;
    const int L0 = lb * 16;

    // stage the slab: y: u(k, z = L0 + l) is contiguous in k;  z: u(y = L0 + l, k) has y contiguous.
    // All loads of a thread are issued before its LDS stores, so their latencies overlap.
    double* Sre = slab;
    double* Sim = slab + 16 * LDK;
    constexpr int IT = (16 * 16 * kWaves * TPW + kThreads - 1) / kThreads;
    d2v v[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = tid + it * kThreads;
        v[it] = d2v{0.0, 0.0};
        if (e < 16 * K) {
            const int l = DIRZ ? e % 16 : e / K, k = DIRZ ? e / 16 : e % K;
            if (L0 + l < m && k < m)
                v[it] = DIRZ ? U[(size_t)(L0 + l) + (size_t)k * m] : U[(size_t)k + (size_t)(L0 + l) * m];
        }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int e = tid + it * kThreads;
        if (e < 16 * K) {
            const int l = DIRZ ? e % 16 : e / K, k = DIRZ ? e / 16 : e % K;
            Sre[l * LDK + k] = v[it][0];
            Sim[l * LDK + k] = v[it][1];
        }
    }
    __syncthreads();

    // tiles of this wave: t = wave + kWaves s
    const d2v* Df = (const d2v*)a.Dfrag;
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
    fetch(0, df);
    for (int kp = 0; kp < KP; ++kp) {
        if (kp + 1 < KP) fetch(kp + 1, dn);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = (kp * 2 + h) * 4 + l4;
            const double ure = Sre[l16 * LDK + k], uim = Sim[l16 * LDK + k];
#pragma unroll
            for (int s = 0; s < TPW; ++s) {
                if constexpr (!DIRZ) {
                    // acc[i = z][j = y] = sum_k u(k, z) D(y, k): A = slab [z][k], B = D fragment [k][y]
                    acc[s][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ure, df[s][h], acc[s][0], 0, 0, 0);
                    acc[s][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(uim, df[s][h], acc[s][1], 0, 0, 0);
                } else {
                    // acc[i = z][j = y] = sum_k D(z, k) u(y, k): A = D fragment [z][k], B = slab [k][y]
                    acc[s][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(df[s][h], ure, acc[s][0], 0, 0, 0);
                    acc[s][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(df[s][h], uim, acc[s][1], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int s = 0; s < TPW; ++s) df[s] = dn[s];
    }

    // epilogue: D[i][j] in lane j + 16 (i % 4), element i / 4; y runs along lanes (coalesced).
    // The accumulated outputs are all loaded before the first store: out and u may be the same
    // allocation for the compiler, which would otherwise serialise every load behind a store.
    d2v ov[TPW][4];
    if (a.beta != 0.0) {
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
                d2v v = {a.scal * acc[s][0][e], a.scal * acc[s][1][e]};
                if (a.beta != 0.0) v += a.beta * ov[s][e];
                O[(size_t)y + (size_t)z * m] = v;
            }
        }
    }
}

template <bool DIRZ>
static hipError_t launch_dir(const CollocArgs& a, hipStream_t st)
{
    const int NT = a.NP / 16;
    const long blocks = (long)a.g.nf * a.g.nmodes * NT;
    switch ((NT + kWaves - 1) / kWaves) {
        case 1: hipLaunchKernelGGL((k_cline<DIRZ, 1>), dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_cline<DIRZ, 2>), dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
        case 3: hipLaunchKernelGGL((k_cline<DIRZ, 3>), dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
        case 4: hipLaunchKernelGGL((k_cline<DIRZ, 4>), dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
        case 5: hipLaunchKernelGGL((k_cline<DIRZ, 5>), dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
        case 6: hipLaunchKernelGGL((k_cline<DIRZ, 6>), dim3((unsigned)blocks), dim3(kThreads), 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_cline(int dirz, const CollocArgs& a, hipStream_t st)
{
    // the matrix fragments cover NP = 16 ceil(m/16) >= m; at most 6 tiles per wave (m <= 288)
    if (a.NP % 16 || a.NP < a.g.m || a.NP > 16 * kWaves * 6 || a.g.nf < 1) return hipErrorInvalidValue;
    if ((long)a.g.nf * a.g.nmodes <= 0) return hipSuccess;
    return dirz ? launch_dir<true>(a, st) : launch_dir<false>(a, st);
}

}  // namespace duct
