// Folded (even/odd) mixed-precision refined solve: declarations.
//
// Intermediates are FP32 "folded planar" arrays: plane p = 2 (f nmodes + mode)
// + part holds four NH x NH column-major quarters, quarter q = 2 r + c (r the
// parity of y, c of z), element (y', z') at ((p 4 + q) NH + z') NH + y'.
#pragma once
#include <vector>

#include "common.h"
#include "proj.h"

namespace duct {

constexpr int kMaxFoldSources = 4;

// fold: FP32 quarters of the interior of complex fields g.in (half-sum/half-difference
// in both directions, /4), zero padding, Im of mode 0 zero.  With Xlo non-null it also
// receives the FP32 remainder (q - X) of every FP64 quarter value q, so that X + Xlo
// carries q to ~2^-48.
hipError_t launch_fold32(const FieldGroup& g, int NH, float* X, hipStream_t st, float* Xlo = nullptr);

// unfold: g.out interior (= or +=) the unfolded sum of nsrc folded FP32 arrays (summed in FP64),
// preceded, with udig non-null, by the digit-form array of launch_lpq (udig = Ucol, gexp);
// zero_edges: the boundary of g.out and of g.in set to 0 as well (helm3d.f90:48-60).
hipError_t launch_unfold_sum(const FieldGroup& g, int NH, const float* const* src, int nsrc, int accumulate,
                             hipStream_t st, const unsigned char* udig = nullptr, const int* gexp = nullptr,
                             int zero_edges = 0);

// The fused low-precision solve of every quarter: U = P_r ((P_r^-1 X P_c^-T) ./ den) P_c^T
// with fp16 hi/lo splits (3 products, FP32 accumulation), per-pass power-of-two scaling.
struct LpqArgs {
    const void* fwd[2];     // folded P^-1 of parity e / o, fp16 hi/lo fragments (build_lpq_basis)
    const void* bwd[2];     // folded P
    float sfwd[2], sbwd[2]; // the bases' power-of-two scales (stored = basis * scale)
    const float* lam;       // [lam_e | lam_o], NH each
    const float* X;         // folded FP32 input
    float* Y;               // folded FP32 output
    // Digit output (Ucol non-null, Y unused): the solution U of each quarter is rounded to the
    // grid 2^(g - 8 kUDig) Z with 2^g > max|U| / 0.49 (g per quarter, in gexp[quarter]); Ucol /
    // Urow receive the kUDig signed base-256 digits of Z = c1 256^(kUDig-1) + ... + c_kUDig,
    // |c| <= 127, as bytes: Ucol[quarter][d][z'][y'] and the transposed Urow[quarter][d][y'][z']
    // (d = 0 for c1).
    unsigned char* Ucol;
    unsigned char* Urow;
    int* gexp;
    FieldGroup g;           // planes, xlambda per plane (no singular plane: |den| > 0)
};
hipError_t launch_lpq(int NH, const LpqArgs& a, hipStream_t st);
// Digits of the quantised first solution U0 (24 bits of each quarter's maximum).  Its rounding
// is white noise, which L amplifies into the residual by up to its largest eigenvalue, and the
// correction solve's error grows with |r| cond(L), not with the correction: with 16 bits (two
// digits) the refined Helmholtz field at m = 200 ended 1.1e-10 from the oracle.
constexpr int kUDig = 3;
// fp16 hi/lo fragments of an NH x NH block B (row i, column k, column-major with leading dim ld);
// returns the scale s (B * s has max magnitude in [2^11, 2^12); hi = fp16(B s), lo = fp16(B s - hi)).
float build_lpq_basis(const double* B, int ld, int NH, std::vector<unsigned short>& frag);

// Folded FP64 residual r = f - (S_r u + u S_c^T - xlambda u) on every quarter, u = sum of
// nsrc folded FP32 arrays (FP64 sum), f folded from the complex fields g.in on load.
struct ResidFoldArgs {
    FieldGroup g;                       // g.in = f (complex fields)
    const float* src[kMaxFoldSources];  // u terms
    int nsrc;
    float* R;                           // folded FP32 residual
    const double* Sfrag[2];             // folded S_e, S_o, FP64 fragment layout (build_frag, NP = NH)
    unsigned long long* stats;          // max|r| (bit patterns) at stats[2 slot[f]], max|f| at +1, or null
    int slot[kMaxFields];
    int NH;
};
hipError_t launch_resid_fold(const ResidFoldArgs& a, hipStream_t st);

// The same folded residual with the two products S_r U and U S_c^T on the int8 matrix cores
// (resid_i8.hip): S = 2^e(row) sum_t d_t 2^-8t with kSDig base-256 digits per row (exact to
// 2^-57 of the row's maximum), U = the digit form that launch_lpq wrote (exact), and every
// digit product with t + u <= kDigT formed exactly in int32, the levels summed in FP64, so that
// the products come out with FP64-level error.  f is read as the FP32 pair Fhi + Flo of
// launch_fold32.
constexpr int kSDig = 7;
constexpr int kDigT = 8;
typedef int i4v __attribute__((ext_vector_type(4)));
struct ResidI8Args {
    FieldGroup g;                       // planes and xlambda
    const float* Fhi;                   // folded f (launch_fold32 X)
    const float* Flo;                   // its remainder (launch_fold32 Xlo)
    const unsigned char* Ucol;          // digits of U (launch_lpq)
    const unsigned char* Urow;
    const int* gexp;
    float* R;                           // folded FP32 residual
    const i4v* Sdig[2];                 // S_e, S_o digits, layout of build_sdig
    const int* Sexp[2];                 // their row exponents, NH each
    unsigned long long* stats;          // as ResidFoldArgs
    int slot[kMaxFields];
    int NH;
};
hipError_t launch_resid_i8(const ResidI8Args& a, hipStream_t st);
// Digits of an NH x NH block B (h x h valid, column-major, leading dimension NH) in the fragment
// layout of the int8 products: frag[((rt kSDig + t) KC + kc) 64 + lane] holds the 16 bytes
// d_(t+1)(rt 16 + lane%16, kc 64 + (lane/16) 16 + b), KC = NH/64; exps[row] = e.
void build_sdig(const double* B, int NH, std::vector<i4v>& frag, std::vector<int>& exps);

}  // namespace duct
