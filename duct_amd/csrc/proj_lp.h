// Low-precision projection passes and the FP64 residual of the mixed solve.
#pragma once
#include "common.h"
#include "proj.h"

namespace duct {

// Precision of the eigenbasis products (duct_helm3d_mixed_batch_dev_ scheme codes).
enum LpScheme : int {
    LP_F64 = 0,     // the plain FP64 MFMA passes of proj_f64.hip
    LP_F32 = 1,     // FP32 MFMA
    LP_BF16X3 = 2,  // bf16, 2-term split, 3 products
    LP_BF16X6 = 3,  // bf16, 3-term split, 6 products
    LP_BF16X1 = 4,  // plain bf16
    LP_FOLD16 = 5,  // even/odd folded, fused per-quarter solve, fp16 hi/lo split (3 products)
    LP_FOLD16Q = 6, // LP_FOLD16 with the FP64-accurate residual formed from int8 digit products
};

struct ProjLpArgs {
    const void* Afrag;    // basis in the scheme's fragment layout (build_frag_lp)
    const float* X;       // planar float input
    float* Y;             // planar float output
    const double* lam;    // eigenvalues padded to NP
    long ncols;
    int gauge;            // 0-based Neumann null-mode index, -1 if none
    int accumulate;       // pass 4: add to the FP64 field instead of overwriting it
    FieldGroup g;
};

hipError_t launch_proj_lp(int NP, int scheme, int mode, const ProjLpArgs& a, hipStream_t st);
size_t lp_frag_bytes(int NP, int scheme);
void build_frag_lp(const double* A, int n, int lda, int NP, int scheme, void* out);

// r = f - (D2B u + u D2B^T - xlambda u) on the interior, FP64.
// FP64 collocation-operator products (resid.hip).
enum CollocMode : int {
    CL_RESID = 0,   // r = f - (D2B_y u + u D2B_z^T - xlambda u), interior
    CL_DY = 1,      // out = beta out + scal (D u) along y, full plane
    CL_DZ = 2,      // out = beta out + scal (u D^T) along z, full plane
};

struct CollocArgs {
    FieldGroup g;                 // RESID: g.in = f, g.out = u.  DY/DZ: g.in = operand, g.out = result
    d2v* r[kMaxFields];           // RESID: residual fields (interior written)
    const double* Dfrag;          // the matrix in the FP64 fragment layout of build_frag, NP padded
    unsigned long long* stats;    // RESID: bit patterns of max|r|, max|f| at stats[2*slot[f]], or null
    int neumann;                  // RESID: singular planes (|xlambda| <= 1e-10) leave max|r| to k_nullproj
    int slot[kMaxFields];         // RESID: stats slot of each field
    int NP;                       // padded matrix size (multiple of 16)
    double scal, beta;            // DY/DZ
};
typedef CollocArgs ResidArgs;
hipError_t launch_colloc(int mode, const CollocArgs& a, hipStream_t st);
hipError_t launch_resid(const ResidArgs& a, hipStream_t st);
// Whole-plane derivative product along y (dirz = 0) or z (dirz = 1) for every plane of a.g:
// out = beta out + scal (D u | u D^T), one 16-line slab per workgroup (deriv.hip); NP <= 288.
hipError_t launch_cline(int dirz, const CollocArgs& a, hipStream_t st);

// Several independent derivative products in one launch (deriv.hip): op k computes
//   out_k = beta_k out_k + scal_k (D_k u_k | u_k D_k^T)     (dirz_k = 0 | 1)
// for every plane of u_k. The outputs must be distinct and no output may be another op's input.
constexpr int kMaxClineOps = 8;
struct ClineOp {
    const d2v* in;
    d2v* out;
    const double* Dfrag;          // fold 0: the matrix in the fragment layout of build_frag, NP padded;
                                  // fold +-1: [Dp | Dm] of build_colloc_folded
    double scal, beta;
    int dirz;
    int fold;                     // 0, or the parity s of a matrix with J D J = s D (J the reflection
                                  // of the CGL nodes): -1 for D, +1 for D2; all ops of a launch alike
};
struct ClineArgs {
    ClineOp op[kMaxClineOps];
    int nops, nmodes, m, NP;
};
hipError_t launch_cline_ops(const ClineArgs& a, hipStream_t st);
hipError_t launch_derivx(const d2v* in, d2v* out, int m, int nmodes, int mode0, double alp, double scal, double beta,
                         hipStream_t st);

}  // namespace duct
