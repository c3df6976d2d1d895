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
struct ResidArgs {
    FieldGroup g;                 // g.in = f (right-hand sides), g.out = u (current solutions)
    d2v* r[kMaxFields];           // residual fields (interior written)
    const double* Dfrag;          // D2B (n x n) in the FP64 fragment layout of build_frag, NP padded
    unsigned long long* stats;    // bit patterns of max|r|, max|f| (atomicMax) at stats[2*slot[f]], or null
    int neumann;                  // singular planes (|xlambda| <= 1e-10) leave max|r| to k_nullproj
    int slot[kMaxFields];         // stats slot of each field
    int NP;
};
hipError_t launch_resid(const ResidArgs& a, hipStream_t st);

}  // namespace duct
