// Passes 2 and 3 of the folded FP64 solve on one LDS-resident quarter.
//
// The folded chain (proj_fold.hip) maps each NH x NH quarter of a plane to a
// quarter of the next pass's input, Y(cp, rp) = (A_rp X(rp, cp))^T, so a
// quarter stays within its own slot pair through the middle of the solve:
//     pass 2:  G = P_rp^-1 X          (X = pass-1 output in slot 2 rp + cp)
//     divide:  G(i, c) /= lam_cp(c) + (lam_rp(i) - xlambda)      (solvediag.f90:61-66, :151-159)
//     pass 3:  V = P_cp G^T           (written back to slot 2 rp + cp for pass 4)
// At NH = 128 one FP64 quarter is 128 KB and fits the 160 KB LDS of a gfx950
// CU.  One workgroup therefore owns one quarter: pass 2 streams it from HBM
// into LDS stage by stage under its own MFMAs, its output replaces it in LDS
// (divided on the way), pass 3 reads it from there and writes the slot once.
// Against two k_projf launches this removes one HBM round trip of the planar
// intermediate and the cold start and epilogue of a launch; the arithmetic
// (the same MFMA sequence over k, the same divide) is that of k_projf's P2P
// and P2P_LDIV passes, so the results are identical to the bit.
//
// Work decomposition: 8 waves (two per SIMD), wave w owns output rows
// 32 (w & 3) .. +32 and columns 64 (w >> 2) .. +64: 2 x 4 tiles of
// v_mfma_f64_16x16x4f64.  The basis rows come from L2 in the fragment-major
// layout of build_frag_rect, two stages ahead; B fragments are read from LDS
// one k-step ahead (column-major quarter, column stride NH + 1 doubles: the 16
// lanes of a row that the LDS serves together hit distinct banks).
#include <algorithm>

#include "common.h"
#include "proj.h"

namespace duct {

namespace {
constexpr int kQNH = 128;                  // the folded size this kernel is built for
constexpr int kQLD = kQNH + 1;             // LDS column stride (doubles)
constexpr int kQThreads = 512;
constexpr int kQStages = kQNH / 16;        // 16-k stages per pass
}  // namespace

__global__ __launch_bounds__(kQThreads, 1) void k_projq(ProjQuarterArgs a)
{
    constexpr int NH = kQNH, LDX = kQLD, NS = kQStages;
    __shared__ __attribute__((aligned(16))) double Xs[NH * LDX];
    __shared__ double lam_s[2 * NH];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, l4 = lane >> 4;
    const int wr = wave & 3, wc = wave >> 2;
    const FieldGroup& g = a.g;
    const long nq = a.nplanes * 4;
    const int n = g.n;
    const int hh[2] = {(n + 1) >> 1, n >> 1};

    // quarter q = plane * 4 + slot; skipped: Im of mode 0 (int2d3d.f90:43, not used by pass 4)
    auto xlam_of = [&](long q) {
        const long fm = q >> 3;
        return xlambda_of(g.alambda[(int)(fm / g.nmodes)], g.alp, g.mode0 + (int)(fm % g.nmodes));
    };
    auto valid = [&](long q) {
        const long fm = q >> 3;
        return !(((q >> 2) & 1) && g.mode0 + (int)(fm % g.nmodes) == 0);
    };
    auto next_valid = [&](long q) {   // first valid quarter of this workgroup's sequence at or after q
        while (q < nq && !valid(q)) q += gridDim.x;
        return q;
    };

    long q = next_valid((long)blockIdx.x);
    if (q >= nq) return;
    const double lamv = a.lam[tid & (2 * NH - 1)];

    // ---- quarter stages: stage s = k in [16 s, 16 s + 16) of every column, 2 x 16 B per thread ----
    auto load_q = [&](const double* Qs, int s, d2v (&r)[2]) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int idx = tid + kQThreads * h;
            r[h] = *(const d2v*)(Qs + (size_t)(idx >> 3) * NH + s * 16 + 2 * (idx & 7));
        }
    };
    auto store_q = [&](int s, const d2v (&r)[2]) {   // odd column stride: 8-byte stores
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int idx = tid + kQThreads * h;
            double* d = &Xs[(idx >> 3) * LDX + s * 16 + 2 * (idx & 7)];
            d[0] = r[h][0];
            d[1] = r[h][1];
        }
    };

    // ---- basis fragments of parity p for this wave's 32 rows: P^-1 (pass 2) or P (pass 3) ----
    auto frag = [&](const double* base, int p, int mt) {
        return (const d2v*)base + (size_t)((p * NH + wr * 32 + mt * 16) / 16) * (NH / 8) * 64 + lane;
    };
    auto load_a = [&](const d2v* const (&A)[2], int s, d2v (&r)[2][2]) {
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int h = 0; h < 2; ++h) r[mt][h] = A[mt][(size_t)(2 * s + h) * 64];
    };

    d4v acc[2][4];
    auto zero_acc = [&]() {
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = d4v{0.0, 0.0, 0.0, 0.0};
    };
    // B fragments of k-step ks (k = 4 ks + l4) for this wave's 64 columns (the scheduling barrier
    // keeps the reads ahead of the MFMAs that follow them: the scheduler would otherwise sink them
    // to their first use and expose the LDS latency)
    auto read_b = [&](int ks, double (&b)[4]) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) b[nt] = Xs[(wc * 64 + nt * 16 + l16) * LDX + ks * 4 + l4];
        __builtin_amdgcn_sched_barrier(0);
    };
    // the 8 MFMAs of k-step kk of a stage (A fragments of the stage in af)
    auto mma = [&](int kk, const d2v (&af)[2][2], const double (&b)[4]) {
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            const double av = af[mt][kk >> 1][kk & 1];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b[nt], acc[mt][nt], 0, 0, 0);
        }
    };

    d2v areg[3][2][2];
    d2v qreg[3][2];
    double bcur[4], bnxt[4];

    // the first quarter's first stages; later quarters' come in under the previous pass 3
    {
        const double* Q0 = a.X + (size_t)q * NH * NH;
        const d2v* const A0[2] = {frag(a.fwd, (int)(q & 3) >> 1, 0), frag(a.fwd, (int)(q & 3) >> 1, 1)};
        load_q(Q0, 0, qreg[0]);
        load_a(A0, 0, areg[0]);
        load_q(Q0, 1, qreg[1]);
        load_a(A0, 1, areg[1]);
        load_q(Q0, 2, qreg[2]);
    }

    while (true) {
        const int slot = (int)(q & 3);
        const int rp = slot >> 1, cp = slot & 1;
        const double xlam = xlam_of(q);
        double* Q = a.X + (size_t)q * NH * NH;
        const d2v* const Af[2] = {frag(a.fwd, rp, 0), frag(a.fwd, rp, 1)};
        const d2v* const Ab[2] = {frag(a.bwd, cp, 0), frag(a.bwd, cp, 1)};

        // ------------ pass 2: the quarter streams into LDS under its own MFMAs ------------
        // Iteration s stores stage s + 1 (loaded two iterations earlier) and computes stage s.  Its
        // barrier sits after the second k-step, when the B reads issued before it are long complete,
        // so the wait for them at the barrier costs nothing and the first B fragments of stage
        // s + 1 are read under the last k-step's MFMAs.
        store_q(0, qreg[0]);
        zero_acc();
        __syncthreads();
        read_b(0, bcur);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (s + 3 < NS) load_q(Q, s + 3, qreg[(s + 3) % 3]);
            if (s + 2 < NS) load_a(Af, s + 2, areg[(s + 2) % 3]);
            else load_a(Ab, s + 2 - NS, areg[(s + 2) % 3]);   // pass 3's first stages
            if (s + 1 < NS) store_q(s + 1, qreg[(s + 1) % 3]);
            if (s == 0 && tid < 2 * NH) lam_s[tid] = lamv;   // read by the divide, after the last barrier
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                if (kk == 2 && s + 1 < NS) {   // (pinned: the MFMAs before it stay before it)
                    __builtin_amdgcn_sched_barrier(0);
                    __syncthreads();
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (4 * s + kk + 1 < 4 * NS) read_b(4 * s + kk + 1, bnxt);
                mma(kk, areg[s % 3], bcur);
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) bcur[nt] = bnxt[nt];
            }
        }
        __syncthreads();

        // ------------ divide, transpose into LDS: X3(c, i) = G(i, c) / den ------------
        // den = lam_cp(c) + (lam_rp(i) - xlambda), the order of k_projf's divide on load; the
        // eigenvalues are read before the first store so that the reads are not held behind it
        {
            double lc[4], li[2][4];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) lc[nt] = lam_s[cp * NH + wc * 64 + nt * 16 + l16];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r) li[mt][r] = lam_s[rp * NH + wr * 32 + mt * 16 + 4 * r + l4] - xlam;
            double out[2][4][4];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int c = wc * 64 + nt * 16 + l16, i = wr * 32 + mt * 16 + 4 * r + l4;
                        const double v = acc[mt][nt][r];
                        const double qv = div_eig(v, lc[nt] + li[mt][r]);
                        out[mt][nt][r] = (i < hh[rp] && c < hh[cp]) ? qv : v;
                    }
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        Xs[(wr * 32 + mt * 16 + 4 * r + l4) * LDX + wc * 64 + nt * 16 + l16] = out[mt][nt][r];
        }
        zero_acc();
        __syncthreads();

        // ------------ pass 3: V = P_cp X3, written to the slot (row-major i', column i) ------------
        // The next quarter's first stages are fetched under it (a valid address when there is none).
        const long qn = next_valid(q + gridDim.x);
        const long qf = qn < nq ? qn : q;
        const double* Qn = a.X + (size_t)qf * NH * NH;
        const d2v* const An[2] = {frag(a.fwd, (int)(qf & 3) >> 1, 0), frag(a.fwd, (int)(qf & 3) >> 1, 1)};
        d2v an[2][2][2];
        read_b(0, bcur);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            if (s + 2 < NS) load_a(Ab, s + 2, areg[(s + 2 + NS) % 3]);
            if (s == 1) load_q(Qn, 0, qreg[0]);
            if (s == 2) load_q(Qn, 1, qreg[1]);
            if (s == 3) load_q(Qn, 2, qreg[2]);
            if (s == 4) load_a(An, 0, an[0]);
            if (s == 5) load_a(An, 1, an[1]);
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                if (4 * s + kk + 1 < 4 * NS) read_b(4 * s + kk + 1, bnxt);
                mma(kk, areg[(s + NS) % 3], bcur);
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) bcur[nt] = bnxt[nt];
            }
        }
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const int c = wc * 64 + nt * 16 + l16;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = wr * 32 + mt * 16 + 4 * r + l4;
                    Q[(size_t)i * NH + c] = acc[mt][nt][r];
                }
            }
        if (qn >= nq) break;
        q = qn;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                areg[0][mt][h] = an[0][mt][h];
                areg[1][mt][h] = an[1][mt][h];
            }
        __syncthreads();   // pass 3's LDS reads are done before the next quarter's first stage lands
    }
}

bool proj_quarter_supported(int NH) { return NH == kQNH; }

hipError_t launch_proj_quarter(int NH, const ProjQuarterArgs& a, hipStream_t st)
{
    if (NH != kQNH) return hipErrorInvalidValue;
    if (a.nplanes <= 0) return hipSuccess;
    // one workgroup per CU (the quarter fills the LDS), each walking its quarters in turn
    int dev = 0, ncu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const long nq = a.nplanes * 4;
    const unsigned grid = (unsigned)std::min<long>(nq, std::max(ncu, 1));
    hipLaunchKernelGGL(k_projq, dim3(grid), dim3(kQThreads), 0, st, a);
    return hipGetLastError();
}

}  // namespace duct
