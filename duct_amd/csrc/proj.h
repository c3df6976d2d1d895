// Launch interfaces of the projection and helper kernels.
#pragma once
#include "common.h"

namespace duct {

enum ProjMode : int {
    PROJ_C2P = 0,      // complex field interior -> planar  (pass 1)
    PROJ_P2P = 1,      // planar -> planar                  (pass 3)
    PROJ_P2P_DIV = 2,  // planar -> planar with the eigenvalue divide (pass 2)
    PROJ_P2C = 3,      // planar -> complex field interior  (pass 4)
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
    long ncols;           // planar: planes*NP columns; complex: nf*nmodes*NP complex columns
    int gauge;            // 0-based eigen-index of the Neumann null mode, -1 if none
    FieldGroup g;
};

hipError_t launch_proj(int NP, int mode, const ProjArgs& a, hipStream_t st);
void build_frag(const double* A, int n, int lda, int NP, double* frag);

// Boundary treatment after pass 4 (one workgroup per complex plane).
enum BoundaryMode : int {
    BND_DIR_ZERO = 0,  // helm3d.f90:48-60: rhs and sol boundary set to zero
    BND_DIR_KEEP = 1,  // solvediag.f90 'd': boundary of U left as given (copied in -> out)
    BND_NEUMANN = 2,   // solvediag.f90:251-322: edges from DC rows, corners from AM1
};

struct BoundaryArgs {
    FieldGroup g;
    const double* DC;   // m x m column-major
    const double* AM1;  // 4 x 4 column-major
    int mode;           // BoundaryMode
};
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

// Dirichlet boundary-condition correction of the right-hand side
// (solvediag.f90:37-45), applied in place to the interior of a complex field.
struct BcCorrectArgs {
    FieldGroup g;
    const double* D2b;  // m x 2 column-major: D2(:,1), D2(:,m)
};
hipError_t launch_bc_correct(const BcCorrectArgs& a, hipStream_t st);

}  // namespace duct
