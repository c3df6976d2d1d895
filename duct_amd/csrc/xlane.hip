// The advection products of k_xprod (advect.hip) for mgalx = 192 points and up to 64 modes, with
// the x-transforms as FFTs on the vector units instead of dense DFT matrices on the matrix cores.
//
// One x-line (a (y, z) point) is transformed by 16 lanes, one DPP row.  With x = 3 j + r
// (j < 64, r < 3) the 192-point transforms split into 64-point ones:
//   inverse  phys(3 j + r) = Re sum_{k<64} a_k e^(2 pi i k r / 192) e^(2 pi i k j / 64),
//            a_k = w_k c_k (w_0 = 1 with Im c_0 dropped, w_k = 2, c_k = 0 for k >= M)
//   forward  P(k) = (1/192) sum_r e^(-2 pi i k r / 192) sum_j p(3 j + r) e^(-2 pi i k j / 64)
// (the definitions of build_xdft's matrices).  A 64-point transform with k = 16 k1 + k2 and
// j = j1 + 4 j2 is a 4-point DFT over k1 inside each lane (lane = k2), a twiddle
// e^(2 pi i k2 j1 / 64), and a 16-point DFT over the 16 lanes for each j1, done as four radix-2
// stages whose partner values come from DPP lane moves (xor 8, 4, 2, 1 within the row).  The
// inverse runs decimation in frequency (natural lane order in, bit-reversed out: lane l holds
// j2 = bitrev4(l)); the forward runs decimation in time on that bit-reversed order and ends in
// natural order (lane l holds k2 = l), so the pointwise products need no exchange.  The six
// real products go through three complex forward transforms, two products per transform
// (z = p_a + i p_b), separated with the mirrored modes Z(-k) of lane (16 - l) mod 16.
//
// Per line: 9 inverse and 9 forward 64-point transforms of ~200 vector operations per lane, so
// the kernel is bound by HBM (3 x M modes in, 6 x M out per line) rather than by arithmetic;
// the dense matrix form costs ~20x the operations of this one.
//
// Layout: fields complex(8) (m, m, M), y fastest.  A workgroup takes 16 consecutive lines of
// the flattened (y, z) plane (z m + y): for a fixed mode their 16 values are contiguous, so
// the coefficients go in, and the products out, through LDS tiles [field][mode][line].
#include <cmath>
#include <type_traits>
#include <vector>

#include "advect.h"

namespace duct {

namespace {
constexpr int kLL = 16;               // lines per workgroup (four per wave)
constexpr int kLLs = kLL + 1;         // LDS stride of a mode's lines (complex values)
constexpr int kTabLanes = 16;

// lane l receives the value of lane l ^ H of its row
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    // mov_dpp: every lane's source is inside its row, so no "old" value has to be initialised
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ d2v dpp_mov2(d2v v)
{
    return d2v{dpp_mov<CTRL>(v[0]), dpp_mov<CTRL>(v[1])};
}
template <int H>
__device__ __forceinline__ d2v partner(d2v v)
{
    if constexpr (H == 1) return dpp_mov2<0xB1>(v);                      // quad_perm [1,0,3,2]
    else if constexpr (H == 2) return dpp_mov2<0x4E>(v);                 // quad_perm [2,3,0,1]
    else if constexpr (H == 4) return dpp_mov2<0x1B>(dpp_mov2<0x141>(v)); // row_half_mirror, then [3,2,1,0]
    else return dpp_mov2<0x128>(v);                                      // row_ror:8
}
// lane l receives the value of lane (16 - l) mod 16 of its row: row_mirror, then row_ror:1
__device__ __forceinline__ d2v mirror16(d2v v)
{
    return dpp_mov2<0x121>(dpp_mov2<0x140>(v));
}

__device__ __forceinline__ d2v cmul(d2v a, d2v b) { return d2v{a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0]}; }
__device__ __forceinline__ d2v cmulc(d2v a, d2v b) { return d2v{a[0] * b[0] + a[1] * b[1], a[1] * b[0] - a[0] * b[1]}; }
__device__ __forceinline__ d2v ipl(d2v a) { return d2v{-a[1], a[0]}; }   // i a

// Twiddles of one lane l (k2 = l): e^(2 pi i l j1 / 64) for j1 = 1..3, e^(2 pi i l r / 192) for
// r = 1, 2, the radix-2 stage factors (DIF sign, 1 on the stage's lower lanes) for the stages of
// half-size 8, 4, 2, 1, and the lane-independent e^(2 pi i k1 r / 12) for k1 = 0..3, r = 1, 2.
struct LaneTw {
    d2v t64[3], tr[2], ts[3];   // ts: the stages of half-size 8, 4, 2 (that of 1 is 1)
};
}  // namespace

// Table layout (d2v): [0, 48) t64[j1-1][l], [48, 80) tr[r-1][l], [80, 144) ts[s][l], [144, 152) c12[r-1][k1].
constexpr int kXlTab = 152;

__global__ __launch_bounds__(256, 2) void k_xprod_lane(XprodArgs a, const d2v* __restrict__ tab)
{
    const int mm = a.m, M = a.M;
    const long lines = (long)mm * mm;
    const long L0 = (long)blockIdx.x * kLL;
    const size_t plane = (size_t)mm * mm;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l = lane & 15, g = lane >> 4;   // lane of the row (k2 or bitrev(j2)), line of the wave
    const int line = wave * 4 + g;
    __shared__ d2v sm[3 * 64 * kLLs];

    // ---- coefficients of u, v, w for the 16 lines -> sm[f][k][line] (modes >= M zero)
    {
        constexpr int kIt = 64 * kLL / 256;
        d2v cv[3][kIt];
#pragma unroll
        for (int f = 0; f < 3; ++f)
#pragma unroll
            for (int it = 0; it < kIt; ++it) {
                const int e = tid + 256 * it, k = e / kLL, i = e % kLL;
                const bool ok = k < M && L0 + i < lines;
                const d2v* src = a.in[f] + (ok ? (size_t)k * plane + L0 + i : 0);
                cv[f][it] = *src;
                if (!ok) cv[f][it] = d2v{0.0, 0.0};
            }
#pragma unroll
        for (int f = 0; f < 3; ++f)
#pragma unroll
            for (int it = 0; it < kIt; ++it) {
                const int e = tid + 256 * it, k = e / kLL, i = e % kLL;
                sm[(f * 64 + k) * kLLs + i] = cv[f][it];
            }
    }
    LaneTw tw;
#pragma unroll
    for (int j = 0; j < 3; ++j) tw.t64[j] = tab[j * kTabLanes + l];
#pragma unroll
    for (int r = 0; r < 2; ++r) tw.tr[r] = tab[48 + r * kTabLanes + l];
#pragma unroll
    for (int s = 0; s < 3; ++s) tw.ts[s] = tab[80 + s * kTabLanes + l];
    __syncthreads();

    // the sign of the lane's own value in the radix-2 stage of half-size H: -1 on the upper lanes
    // (l & H), applied by an fma (no bit manipulation of vector elements: clang's
    // __builtin_bit_cast of an ext_vector element reads element 0 whatever the index)
    const double sg8 = (l & 8) ? -1.0 : 1.0, sg4 = (l & 4) ? -1.0 : 1.0;
    const double sg2 = (l & 2) ? -1.0 : 1.0, sg1 = (l & 1) ? -1.0 : 1.0;
    auto flip_add = [&](int H, d2v v, d2v q) -> d2v {
        const double s = H == 8 ? sg8 : H == 4 ? sg4 : H == 2 ? sg2 : sg1;
        return d2v{s * v[0] + q[0], s * v[1] + q[1]};
    };
    // decimation-in-frequency stage of half-size H (factor index S, none for H = 1)
    auto dif = [&](auto hc, d2v (&v)[4]) {
        constexpr int H = decltype(hc)::value;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const d2v q = partner<H>(v[j]);
            const d2v t = flip_add(H, v[j], q);
            if constexpr (H == 1) v[j] = t;
            else v[j] = cmul(t, tw.ts[H == 8 ? 0 : H == 4 ? 1 : 2]);
        }
    };
    // decimation-in-time stage (forward sign): the upper lane premultiplies by the conjugate factor
    auto dit = [&](auto hc, d2v (&v)[4]) {
        constexpr int H = decltype(hc)::value;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            d2v vp = v[j];
            if constexpr (H != 1) vp = cmulc(vp, tw.ts[H == 8 ? 0 : H == 4 ? 1 : 2]);
            const d2v q = partner<H>(vp);
            v[j] = flip_add(H, vp, q);
        }
    };
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I4 = std::integral_constant<int, 4>;
    using I8 = std::integral_constant<int, 8>;

    // the r twiddle of mode k = 16 k1 + l: e^(2 pi i k1 r / 12) e^(2 pi i l r / 192)
    auto twr = [&](int r, int k1) -> d2v {
        return r == 0 ? d2v{1.0, 0.0} : cmul(tab[144 + (r - 1) * 4 + k1], tw.tr[r - 1]);
    };

    // ---- inverse transforms: phys[f][r][j1] at x = 3 (j1 + 4 bitrev4(l)) + r
    double phys[3][3][4];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            d2v v[4];
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) {
                d2v c = sm[(f * 64 + 16 * k1 + l) * kLLs + line];
                const bool k0 = k1 == 0 && l == 0;
                c = k0 ? d2v{c[0], 0.0} : d2v{2.0 * c[0], 2.0 * c[1]};
                v[k1] = r == 0 ? c : cmul(c, twr(r, k1));
            }
            // 4-point inverse DFT over k1 (e^(2 pi i / 4) = i), then e^(2 pi i l j1 / 64)
            const d2v s0 = v[0] + v[2], d0 = v[0] - v[2], s1 = v[1] + v[3], d1 = v[1] - v[3];
            v[0] = s0 + s1;
            v[2] = cmul(s0 - s1, tw.t64[1]);
            v[1] = cmul(d0 + ipl(d1), tw.t64[0]);
            v[3] = cmul(d0 - ipl(d1), tw.t64[2]);
            dif(I8{}, v);
            dif(I4{}, v);
            dif(I2{}, v);
            dif(I1{}, v);
#pragma unroll
            for (int j = 0; j < 4; ++j) phys[f][r][j] = v[j][0];
        }

    // ---- products, two per forward transform, pair pp = (uu, vv), (ww, uv), (uw, vw) = outputs
    // (2 pp, 2 pp + 1) in the order uu vv ww uv uw vw; each pair goes out through the LDS tile
    const double scale = 0.5 / 192.0;   // the separation's 1/2 and the forward transform's 1/X
#pragma unroll
    for (int pp = 0; pp < 3; ++pp) {
        d2v acc[2][4];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) acc[0][k1] = acc[1][k1] = d2v{0.0, 0.0};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            d2v z[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double u = phys[0][r][j], v = phys[1][r][j], w = phys[2][r][j];
                if (pp == 0) z[j] = d2v{u * u, v * v};
                else if (pp == 1) z[j] = d2v{w * w, u * v};
                else z[j] = d2v{u * w, v * w};
            }
            dit(I1{}, z);
            dit(I2{}, z);
            dit(I4{}, z);
            dit(I8{}, z);
            // e^(-2 pi i l j1 / 64), then the 4-point forward DFT over j1 (e^(-2 pi i / 4) = -i)
            z[1] = cmulc(z[1], tw.t64[0]);
            z[2] = cmulc(z[2], tw.t64[1]);
            z[3] = cmulc(z[3], tw.t64[2]);
            const d2v s0 = z[0] + z[2], d0 = z[0] - z[2], s1 = z[1] + z[3], d1 = z[1] - z[3];
            const d2v Z[4] = {s0 + s1, d0 - ipl(d1), s0 - s1, d0 + ipl(d1)};   // Z[k1], k = 16 k1 + l
            // the mirrored modes -k: lane (16 - l) mod 16, element 3 - k1 (l > 0) or -k1 mod 4 (l = 0)
            d2v G[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) G[e] = mirror16(Z[e]);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) {
                const d2v zm = l == 0 ? G[(4 - k1) & 3] : G[3 - k1];
                // P_a = (Z + conj zm) / 2, P_b = -i (Z - conj zm) / 2 (the 1/2 in scale)
                const d2v pa = d2v{Z[k1][0] + zm[0], Z[k1][1] - zm[1]};
                const d2v dm = d2v{Z[k1][0] - zm[0], Z[k1][1] + zm[1]};
                const d2v pb = d2v{dm[1], -dm[0]};
                if (r == 0) {
                    acc[0][k1] += pa;
                    acc[1][k1] += pb;
                } else {
                    const d2v t = twr(r, k1);
                    acc[0][k1] += cmulc(pa, t);
                    acc[1][k1] += cmulc(pb, t);
                }
            }
        }
        __syncthreads();   // the tile's previous contents (coefficients, the last pair) have been read
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) {
                const d2v v = acc[q][k1];
                sm[(q * 64 + 16 * k1 + l) * kLLs + line] = d2v{v[0] * scale, v[1] * scale};
            }
        __syncthreads();
        constexpr int kIt = 64 * kLL / 256;
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int it = 0; it < kIt; ++it) {
                const int e = tid + 256 * it, k = e / kLL, i = e % kLL;
                if (k < M && L0 + i < lines)
                    a.out[2 * pp + q][(size_t)k * plane + L0 + i] = sm[(q * 64 + k) * kLLs + i];
            }
    }
}

hipError_t launch_xprod_lane(const XprodArgs& a, const d2v* tab, hipStream_t st)
{
    if (a.X != 192 || a.M < 1 || a.M > 64 || a.m <= 0) return hipErrorInvalidValue;
    const long lines = (long)a.m * a.m;
    hipLaunchKernelGGL(k_xprod_lane, dim3((unsigned)((lines + kLL - 1) / kLL)), dim3(256), 0, st, a, tab);
    return hipGetLastError();
}

void build_xlane_table(std::vector<double>& tab)
{
    tab.assign((size_t)2 * kXlTab, 0.0);
    const long double tp = 6.283185307179586476925286766559005768L;
    auto put = [&](int idx, long double ang) {
        tab[2 * idx] = (double)cosl(ang);
        tab[2 * idx + 1] = (double)sinl(ang);
    };
    for (int l = 0; l < kTabLanes; ++l) {
        for (int j = 1; j <= 3; ++j) put((j - 1) * kTabLanes + l, tp * (long double)((l * j) % 64) / 64.0L);
        for (int r = 1; r <= 2; ++r) put(48 + (r - 1) * kTabLanes + l, tp * (long double)(l * r) / 192.0L);
        for (int s = 0; s < 4; ++s) {
            const int h = 8 >> s, pos = l % (2 * h);
            if (pos >= h) put(80 + s * kTabLanes + l, tp * (long double)(pos - h) / (long double)(2 * h));
            else put(80 + s * kTabLanes + l, 0.0L);
        }
    }
    for (int r = 1; r <= 2; ++r)
        for (int k1 = 0; k1 < 4; ++k1) put(144 + (r - 1) * 4 + k1, tp * (long double)((k1 * r) % 12) / 12.0L);
}

}  // namespace duct
