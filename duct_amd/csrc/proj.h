// Launch interfaces of the projection and helper kernels.
#pragma once
#include "common.h"

namespace duct {

enum ProjMode : int {
    PROJ_C2P = 0,      // complex field interior -> planar  (pass 1)
    PROJ_P2P = 1,      // planar -> planar                  (pass 3)
    PROJ_P2P_DIV = 2,  // planar -> planar with the eigenvalue divide (pass 2)
    PROJ_P2C = 3,      // planar -> complex field interior  (pass 4)
    PROJ_P2P_LDIV = 4, // planar -> planar, the eigenvalue divide applied to the
                       // input as it is staged in LDS (pass 3 after an undivided pass 2)
};

// Workgroup geometry of the projection kernels for a padded interior NP.
template <int NP>
struct ProjShape {
    static constexpr int WR = NP < 64 ? NP : 64;             // rows per wave
    static constexpr int WM = (NP / WR) > 4 ? 4 : NP / WR;   // waves along rows
    static constexpr int MT = WR * WM;                       // rows per workgroup
    static constexpr int WN = 4 / WM;                        // waves along columns
    static constexpr int NT = 64 * WN;                       // (virtual) columns per workgroup
    static constexpr int TM = WR / 16;                       // 16-row tiles per wave
    static constexpr int KT = 16;                            // k per LDS stage
    static constexpr int LDB = KT + 1;                       // LDS column stride (doubles)
    static constexpr int NK = NP / KT;                       // LDS stages
    static constexpr int ROWBLOCKS = NP / MT;
};

struct ProjArgs {
    const double* Afrag;  // NP x NP basis in fragment-major layout (build_frag)
    const double* X;      // planar input, planes of NP x NP column-major
    double* Y;            // planar output
    const double* lam;    // eigenvalues padded with zeros to NP (divide epilogue)
    const double* lam2;   // when set, the eigenvalues of the other direction: den = lam[i] + lam2[j] - xlambda
                          // with i the y-eigen index, j the z-eigen index (separate y and z bases)
    long ncols;           // planar: planes*NP columns; complex: nf*nmodes*NP complex columns
    int gauge;            // 0-based eigen-index of the Neumann null mode, -1 if none
    FieldGroup g;
    // Rectangular planes (separate y and z bases of sizes my, mz): the pass's valid K (= output
    // row) count kn, its valid column count cn, the complex fields' y extent ldf and plane size
    // fplane (my mz).  All 0: the square plane, kn = cn = g.n, ldf = g.m.
    int kn, cn, ldf;
    long fplane;
};

hipError_t launch_proj(int NP, int mode, const ProjArgs& a, hipStream_t st);
void build_frag(const double* A, int n, int lda, int NP, double* frag);
// Fragment layout of a general R x K matrix (R, K multiples of 16 and 8).
void build_frag_rect(const double* A, int R, int K, int lda, double* frag);

// Even/odd (folded) projection passes: planes of four NH x NH quarters,
// basis = [A_e; A_o] as a (2 NH) x NH fragment-layout matrix.
// Zero doubles appended to each folded basis (fragment layout): the planar passes of k_projf
// fetch one stage of A fragments past the end after their last stage.
constexpr int kFoldFragSlack = 256;

struct ProjFoldArgs {
    const double* Afrag;  // folded basis, fragment layout
    const double* X;      // planar input (quarters)
    double* Y;            // planar output (quarters)
    const double* lam;    // [lam_e | lam_o], NH each, zero padded
    long ncols;           // folded complex columns: nf * nmodes * NH
    int accumulate;       // pass 4: out += U instead of out = U
    FieldGroup g;         // no singular plane (the folded solves never take one): |den| > 0
};
hipError_t launch_proj_fold(int NH, int mode, const ProjFoldArgs& a, hipStream_t st);

// The eigenvalue divide of the folded solve, a / d for a finite d of moderate magnitude (an
// eigenvalue sum, |d| well inside the normal range; no scaling for overflow, underflow or special
// values): v_rcp_f64 refined by one Newton step, then the quotient corrected by its residual
// (Markstein): 6 FP64 operations against 11 for the fully scaled IEEE sequence.  On gfx950 it
// returned the IEEE quotient in all of 4.2e6 random pairs (|a| in 1e-6 .. 1e9, d in -2e9 .. -1e-6),
// as did the same with two Newton steps.  The folded passes that divide (k_projf's divide on load,
// k_projq) share it, so their results agree to the bit.
__device__ inline double div_eig(double a, double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
    const double q = a * r;
    return __builtin_fma(r, __builtin_fma(-d, q, a), q);
}

// Passes 2 and 3 of the folded solve (undivided product, eigenvalue divide, product) on
// LDS-resident quarters, one workgroup per quarter, in place on the planar buffer.  Same
// results as PROJ_P2P followed by PROJ_P2P_LDIV.  The Im planes of mode 0 are left untouched.
struct ProjQuarterArgs {
    const double* fwd;    // folded P^-1 [P_e^-1; P_o^-1], fragment layout
    const double* bwd;    // folded P [P_e; P_o], fragment layout
    const double* lam;    // [lam_e | lam_o], NH each, zero padded
    double* X;            // planar quarters: pass-1 output in, pass-3 output out
    long nplanes;         // real planes: nf * nmodes * 2
    FieldGroup g;
};
bool proj_quarter_supported(int NH);
hipError_t launch_proj_quarter(int NH, const ProjQuarterArgs& a, hipStream_t st);

// Padded folded size NH for interior n, or -1 when the folded kernels do not apply.
int fold_size(int n);
// Classify P's columns by parity (the ceil(n/2) most even ones are even) and build the folded
// P^-1 (fwd) and P (bwd) fragments, each (2 NH) x NH, and the eigenvalues per parity.  A column
// that mixes the parities (an eigensolver's rotation of a nearly degenerate even/odd pair) is
// projected onto its parity; the folded inverse is then recomputed and its eigenvalue taken
// from the folded centrosymmetric part of op (the n x n interior operator, column-major;
// required only then).  Returns -1 if the folded blocks cannot be built.
// The folded matrices themselves: Af = [P_e^-1; P_o^-1], Ab = [P_e; P_o], each
// (2 NH) x NH column-major (rows of parity p start at p NH).
int fold_matrices(const double* P, const double* Pinv, const double* lam, int n, int NH, double* Af, double* Ab,
                  double* lamf, int* pos, const double* op);
// pos (may be null) receives, for every eigenvector k, its index within its
// parity (even: 0..h_e-1) or -1-index for odd ones.
int build_fold_bases(const double* P, const double* Pinv, const double* lam, int n, int NH,
                     double* fwd, double* bwd, double* lamf, int* pos, const double* op);

// Boundary treatment after pass 4 (one workgroup per complex plane).
enum BoundaryMode : int {
    BND_DIR_ZERO = 0,  // helm3d.f90:48-60: rhs and sol boundary set to zero
    BND_DIR_KEEP = 1,  // solvediag.f90 'd': boundary of U left as given (copied in -> out)
    BND_NEUMANN = 2,   // solvediag.f90:251-322: edges from DC rows, corners from AM1
    BND_OUT_ZERO = 3,  // adjoint solve: sol boundary set to zero, rhs untouched
};

struct BoundaryArgs {
    FieldGroup g;
    const double* DC;   // m x m column-major
    const double* AM1;  // 4 x 4 column-major
    int mode;           // BoundaryMode
    d2v* part;          // BND_NEUMANN: scratch of boundary_scratch_bytes(g) for the row sums
};
constexpr int kBndSplit = 8;   // workgroups per plane of the Neumann edge sums
size_t boundary_scratch_bytes(const FieldGroup& g);
hipError_t launch_boundary(const BoundaryArgs& a, hipStream_t st);

// Neumann k=0 gauge (solvediag.f90:161-206) on planar G after pass 2.
struct GaugeArgs {
    void* G;            // planar buffer (pass-2 output), double or float
    int f32;            // 1 if G holds float
    const double* prow; // P(iofzero, 1:n), padded to NP
    double pdiag;       // P(iofzero, iofzero)
    double wanted;
    int relative;       // 1: target wanted - u(iofzero+1, iofzero+1) of the field in g.out
                        //    (a refinement correction pins the corrected sum, not itself)
    const double* lam;  // non-null: G is undivided (pass 2 without the divide); each
                        //    term is divided by lam_l + lam_k - xlambda as it is summed
    int gauge;          // 0-based iofzero-1
    int NP;
    FieldGroup g;
};
constexpr int kMaxGaugePlanes = 64;
struct GaugeList {
    int count;
    int plane[kMaxGaugePlanes];  // planar plane indices (2*(f*nmodes+mode)+part)
};
hipError_t launch_gauge(const GaugeArgs& a, const GaugeList& list, hipStream_t st);

// Null-space projection of the residual on the singular (Neumann, xlambda = 0)
// planes of a refinement step: r -= (q^T r q) p p^T with q the null row of P^-1
// and p the null column of P, i.e. the component the gauge solve discards
// (solvediag.f90:161-206) is removed in FP64 before the low-precision correction.
struct NullProjArgs {
    d2v* r[kMaxFields];            // residual fields (interior)
    const double* q;               // P^-1(iofzero, :), n
    const double* p;               // P(:, iofzero), n
    unsigned long long* stats;     // max|r| after projection at stats[2*slot[f]], or null
    int slot[kMaxFields];
    int nmodes, m, n;
};
hipError_t launch_nullproj(const NullProjArgs& a, const GaugeList& list, hipStream_t st);

// Adjoint of the Neumann boundary reconstruction (solvediag.f90:251-322):
// out_interior = in_interior + E^T in_boundary, E mapping the interior to the
// edges (DC rows) and corners (AM1).  One workgroup per complex plane.
struct BoundaryAdjArgs {
    FieldGroup g;       // g.in = adjoint field (full plane), g.out = interior result
    const double* DC;   // m x m
    const double* AM1;  // 4 x 4
};
hipError_t launch_boundary_adj(const BoundaryAdjArgs& a, hipStream_t st);

// Adjoint of the k=0 gauge on planar pass-2 output G = (P^T Ubar P) ./ den:
//   G(l,k) -= c P(a,l) P(a,k) / den(l,k) for (l,k) != (a,a),  c = P(:,a)^T Ubar P(:,a) / P(a,a)^2
// (forward: G(a,a) = (wanted - sum_{(l,k) != (a,a)} P(a,l) P(a,k) G(l,k)) / P(a,a)^2).
struct GaugeAdjArgs {
    FieldGroup g;        // g.in = adjoint input field Ubar (interior read)
    double* G;           // planar pass-2 output
    const double* prow;  // P(a, :)
    const double* pcol;  // P(:, a)
    const double* lam;   // eigenvalues
    double pdiag;
    int gauge, NP;
};
hipError_t launch_gauge_adj(const GaugeAdjArgs& a, const GaugeList& list, hipStream_t st);

// Dirichlet boundary-condition correction of the right-hand side
// (solvediag.f90:37-45), applied in place to the interior of a complex field.
struct BcCorrectArgs {
    FieldGroup g;
    const double* D2b;  // m x 2 column-major: D2(:,1), D2(:,m)
};
hipError_t launch_bc_correct(const BcCorrectArgs& a, hipStream_t st);

}  // namespace duct
