// extern "C" boundary of libduct_hip.so (declared in include/duct_hip.h).
//
// Library state: one device, one stream, the device copies of the Dirichlet
// and Neumann eigenbases (uploaded once by duct_setup_), and two planar
// scratch buffers of nplanes * NP^2 doubles that the projection passes
// ping-pong through.  Host-array entry points (duct_helm3d_, duct_int2d3d_,
// duct_solvediag_) additionally keep device copies of their fields.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <rocfft/rocfft.h>

#include "../../include/duct_hip.h"
#include "advect.h"
#include "aux.h"
#include "common.h"
#include "h5read.h"
#include "mixed_fold.h"
#include "proj.h"
#include "proj_lp.h"
#include "rk3.h"
#include "temp.h"
#include "shen.h"

using namespace duct;

// Scratch of the allocation-free entry duct_helm3d_batch_noalloc_dev_: the two planar buffers of the
// projection passes and the Neumann edge sums of one boundary group, in a fixed device array.  Every
// GPU the library runs on holds its own copy until the process ends.  Host code never passes the
// array's name to a kernel or a copy (that is a host-side stand-in, not a device address): it uses
// the per-device address that hipGetSymbolAddress returns (Lib::scratch_ptr).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

struct DevBasis {
    bool ready = false;
    int m = 0, n = 0, NP = 0;
    double* Afwd = nullptr;  // P^-1, fragment layout (passes 1, 2)
    double* Abwd = nullptr;  // P,    fragment layout (passes 3, 4)
    double* lam = nullptr;   // NP, zero padded
    double* DC = nullptr;    // m x m
    double* AM1 = nullptr;   // 4 x 4
    double* D2b = nullptr;   // m x 2: D2(:,1), D2(:,m)
    double* prow = nullptr;  // P(iofzero, :), NP
    void* lpf_fwd = nullptr; // P^-1, FP32 fragment layout (mixed solve)
    void* lpf_bwd = nullptr; // P
    void* lph_fwd = nullptr; // P^-1, 3 bf16 terms
    void* lph_bwd = nullptr; // P
    double* opfrag = nullptr;// interior operator D2B, FP64 fragment layout (residual)
    double* qrow = nullptr;  // P^-1(iofzero, :), Neumann null projection
    double* pcol = nullptr;  // P(:, iofzero)
    double* AfwdT = nullptr; // P^T,  fragment layout (adjoint passes 1, 2)
    double* AbwdT = nullptr; // P^-T, fragment layout (adjoint passes 3, 4)
    int NH = 0;              // folded (even/odd) size, 0 if the basis does not fold
    double* ffwd = nullptr;  // folded P^-1, (2 NH) x NH fragment layout
    double* fbwd = nullptr;  // folded P
    double* flam = nullptr;  // eigenvalues per parity, 2 NH
    // folded mixed solve (NH = 64 or 128): fp16 hi/lo bases, float eigenvalues, folded S blocks
    bool lq_ready = false;
    void* lq_fwd[2] = {nullptr, nullptr};
    void* lq_bwd[2] = {nullptr, nullptr};
    float lq_sfwd[2] = {1.f, 1.f}, lq_sbwd[2] = {1.f, 1.f};
    float* lq_lam = nullptr;
    double* sfold[2] = {nullptr, nullptr};   // centrosymmetric part of D2B, folded e / o, FP64 fragments
    i4v* sdig[2] = {nullptr, nullptr};       // the same blocks as int8 digits (build_sdig)
    int* sexp[2] = {nullptr, nullptr};       // and their row exponents
    double pdiag = 0.0, wanted = 0.0;
    int gauge = -1;          // 0-based eigen-index of the null mode (Neumann)
};

// Mixed-boundary (temperature) bases: separate y and z eigenbases (pseudomat_temp.f90)
struct TempDev {
    bool ready = false;
    int my = 0, mz = 0, NP = 0;
    int b[4] = {0, 0, 0, 0};                 // y=-1, y=+1, z=-1, z=+1: 0 = 'd', 1 = 'n'
    double* Ayi = nullptr;                   // PTYM1, fragment layout (pass 1)
    double* Azi = nullptr;                   // PTZM1 (pass 2)
    double* Ay = nullptr;                    // PTY (pass 3)
    double* Az = nullptr;                    // PTZ (pass 4)
    double* ly = nullptr;                    // ETY, NP (zero padded)
    double* lz = nullptr;                    // ETZ
    double* vec = nullptr;                   // the 4 my + 4 mz wall vectors of TempArgs
};

// Shen-Legendre Galerkin matrices of one boundary type (preshen.f90)
struct ShenDev {
    bool ready = false;
    int m = 0;
    double* frag[4] = {nullptr, nullptr, nullptr, nullptr};   // A_y, A_z, B_y, B_z (m x m padded), k_cline layout
    double* vec = nullptr;                                     // md, mu, sd, lam (n each)
};

struct ProfRec {
    int kind;
    hipEvent_t a, b;
};

enum ProfKind { PK_C2P = 0, PK_P2P_DIV = 1, PK_GAUGE = 2, PK_P2P = 3, PK_P2C = 4, PK_BOUNDARY = 5, PK_BCCORR = 6,
                PK_TRANSPOSE = 7, PK_INVDIAG = 8, PK_LP_C2P = 9, PK_LP_P2P_DIV = 10, PK_LP_P2P = 11,
                PK_LP_P2C = 12, PK_RESID = 13, PK_DERIV = 14, PK_ADJ = 15, PK_F_C2P = 16,
                PK_F_P2P_DIV = 17, PK_F_P2P = 18, PK_F_P2C = 19, PK_XPROD = 20, PK_TRIDIAG = 21,
                PK_F_MID = 22 };

struct Lib {
    bool init = false;
    int device = 0;
    hipStream_t stream = nullptr;
    DevBasis basis[2];  // 0 = 'd', 1 = 'n'
    double* plan[2] = {nullptr, nullptr};
    size_t plan_bytes = 0;
    d2v* hfield[2] = {nullptr, nullptr};  // device fields for host-array entries
    size_t hfield_bytes = 0;
    d2v* rfield = nullptr;                // residual fields of the mixed solve
    size_t rfield_bytes = 0;
    unsigned long long* stats = nullptr;  // residual maxima of the mixed solve
    double* dcfrag[2] = {nullptr, nullptr}; // DC and DC^T, FP64 fragment layout (derivative products)
    double* tfrag[2] = {nullptr, nullptr};  // Chebyshev transforms phys->cheb, cheb->phys (transyz)
    double* d2frag = nullptr;             // D2 = DC DC, FP64 fragment layout (laplacian3d)
    double* dfold[2] = {nullptr, nullptr};  // D and D2 with the contraction folded (build_colloc_folded)
    d2v* rk = nullptr;                    // Runge-Kutta substep work fields: h1..h3, advnew1..3, phi
    size_t rk_bytes = 0;
    ShenDev shen[2];                      // Shen-Legendre solve, 'd' and 'n'
    double* sgam = nullptr;               // its tridiagonal factors
    size_t sgam_bytes = 0;
    TempDev temp;                         // mixed-boundary solve
    d2v* tfield = nullptr;                // its corrected right-hand side
    size_t tfield_bytes = 0;
    d2v* bscr = nullptr;                  // row sums of the Neumann edge kernels
    size_t bscr_bytes = 0;
    int dc_m = 0, dc_np = 0;
    int fold = 0;                         // 1: Dirichlet groups use the folded passes, 2: Neumann too
    int fuse_quarters = 1;                // folded passes 2 and 3 on LDS-resident quarters where built
    double* xdft[2] = {nullptr, nullptr}; // advection x-DFTs (inverse, forward), fragment layout
    d2v* xlane = nullptr;                  // twiddles of the 192-point lane FFTs (build_xlane_table)
    int xdft_M = 0, xdft_X = 0;
    d2v* prods = nullptr;                 // the six product fields of the advection term
    size_t prods_bytes = 0;
    // x-transforms past the dense kernels' sizes: rocFFT plans (kept per layout), their work
    // buffer and execution info, the Hermitian and the physical lines
    struct XfftPlan {
        int forward;
        size_t X, batch, is, id, os, od;
        rocfft_plan plan;
        size_t work;
    };
    std::vector<XfftPlan> xplans;
    bool rocfft_on = false;
    rocfft_execution_info xinfo = nullptr;
    void* xwork = nullptr;
    size_t xwork_bytes = 0;
    d2v* xw = nullptr;
    size_t xw_bytes = 0;
    double* xp = nullptr;
    size_t xp_bytes = 0;
    hipEvent_t ev_switch = nullptr;       // orders a newly set stream after the previous one
    bool noalloc = false;                 // inside duct_helm3d_batch_noalloc_dev_: nothing may allocate
    void* scratch_ptr[kNoAllocDevices] = {};  // per device: the address of the static scratch array
    float* mfold = nullptr;               // folded FP32 arrays of the folded mixed solve
    size_t mfold_bytes = 0;
    unsigned char* mdig = nullptr;        // its digit-form iterate (Ucol, Urow) and exponents
    size_t mdig_bytes = 0;
    size_t stats_bytes = 0;
    bool prof = false;
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> pool;
    std::string err;
    // a second stream for the Neumann group of a batched solve (default priority, non-blocking), its
    // planar buffers, and the fork / join events
    hipStream_t side = nullptr;
    double* plan2[2] = {nullptr, nullptr};
    size_t plan2_bytes = 0;
    float* mfold2 = nullptr;              // and its folded mixed-solve arrays, digits, residual fields
    size_t mfold2_bytes = 0;
    unsigned char* mdig2 = nullptr;
    size_t mdig2_bytes = 0;
    d2v* rfield2 = nullptr;
    size_t rfield2_bytes = 0;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int overlap = 1;
};

Lib L;

template <class P>
void free_ptr(P*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

void free_basis(DevBasis& B)
{
    free_ptr(B.Afwd); free_ptr(B.Abwd); free_ptr(B.lam); free_ptr(B.DC); free_ptr(B.AM1);
    free_ptr(B.D2b); free_ptr(B.prow); free_ptr(B.lpf_fwd); free_ptr(B.lpf_bwd); free_ptr(B.lph_fwd);
    free_ptr(B.lph_bwd); free_ptr(B.opfrag); free_ptr(B.qrow); free_ptr(B.pcol);
    free_ptr(B.AfwdT); free_ptr(B.AbwdT); free_ptr(B.ffwd); free_ptr(B.fbwd); free_ptr(B.flam);
    for (int p = 0; p < 2; ++p) {
        free_ptr(B.lq_fwd[p]); free_ptr(B.lq_bwd[p]); free_ptr(B.sfold[p]);
        free_ptr(B.sdig[p]); free_ptr(B.sexp[p]);
    }
    free_ptr(B.lq_lam); 
    B = DevBasis{};
}

void free_temp()
{
    TempDev& T = L.temp;
    free_ptr(T.Ayi); free_ptr(T.Azi); free_ptr(T.Ay); free_ptr(T.Az);
    free_ptr(T.ly); free_ptr(T.lz); free_ptr(T.vec);
    T.ready = false;
}

void free_shen(ShenDev& S)
{
    for (auto& f : S.frag) free_ptr(f);
    free_ptr(S.vec);
    S.ready = false;
}

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    L.err = buf;
    return code;
}

#define HIPCHK(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                        __FILE__, __LINE__);                                              \
    } while (0)

// The device address of the static scratch array on `device` (current), looked up once per device.
int noalloc_scratch_setup(int device)
{
    if (device < 0 || device >= kNoAllocDevices) return OK;   // the entry then reports it
    if (L.scratch_ptr[device]) return OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    L.scratch_ptr[device] = ptr;
    return OK;
}

int ensure_init()
{
    if (L.init) return OK;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    L.device = dev;
    L.init = true;
    return OK;
}

template <class T>
int ensure_buf(T*& p, size_t& cap, size_t bytes)
{
    if (bytes <= cap && p) return OK;
    if (L.noalloc)
        return fail(ERR_ARG, "the allocation-free entry needs %zu bytes of a buffer that holds %zu", bytes, cap);
    if (p) {
        HIPCHK(hipStreamSynchronize(L.stream));
        HIPCHK(hipFree(p));
        p = nullptr;
        cap = 0;
    }
    hipError_t e = hipMalloc((void**)&p, bytes);
    if (e != hipSuccess) {
        p = nullptr;
        return fail(ERR_ALLOC, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    cap = bytes;
    return OK;
}

int ensure_planar(size_t bytes)
{
    if (bytes <= L.plan_bytes && L.plan[0]) return OK;
    if (L.noalloc)
        return fail(ERR_ARG, "the allocation-free entry needs %zu planar bytes, its scratch holds %zu", bytes,
                    L.plan_bytes);
    size_t c0 = L.plan_bytes, c1 = L.plan_bytes;
    int st = ensure_buf(L.plan[0], c0, bytes);
    if (st) return st;
    st = ensure_buf(L.plan[1], c1, bytes);
    if (st) return st;
    L.plan_bytes = bytes;
    return OK;
}

// k_boundary and, for the Neumann edges, its two kernels with their scratch of row sums.
hipError_t run_boundary(BoundaryArgs& ba)
{
    if (ba.mode == BND_NEUMANN) {
        if (ensure_buf(L.bscr, L.bscr_bytes, boundary_scratch_bytes(ba.g)) != OK) return hipErrorOutOfMemory;
        ba.part = L.bscr;
    }
    return launch_boundary(ba, L.stream);
}

hipEvent_t take_event()
{
    if (!L.pool.empty()) {
        hipEvent_t e = L.pool.back();
        L.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Bracket one launch with events when timing is enabled.
struct Timed {
    int kind;
    hipEvent_t a = nullptr, b = nullptr;
    explicit Timed(int k) : kind(k)
    {
        if (!L.prof || L.noalloc) return;   // events come from a pool that may create them
        a = take_event();
        b = take_event();
        if (a) (void)hipEventRecord(a, L.stream);
    }
    ~Timed()
    {
        if (!L.prof || !a || !b) return;
        (void)hipEventRecord(b, L.stream);
        L.recs.push_back({kind, a, b});
    }
};

int bound_index(char c)
{
    if (c == 'd') return 0;
    if (c == 'n') return 1;
    return -1;
}

// The folded solves use the even/odd eigenbases of the centrosymmetric part S = (D2B + J D2B J)/2
// of the interior operator.  D2B is centrosymmetric only up to rounding (~2e-15 of its norm), and
// on the ill-conditioned low Poisson modes (xlambda < kLowXlam) S^-1 f and D2B^-1 f differ by up
// to 4e-11 of the field at m = 257 (1e-12 from xlambda ~ 25 up, 2e-13 from 64 up).  Those
// leading modes of a Neumann group, the singular k=0 plane among them, are therefore solved
// with the unfolded FP64 passes, solvediag's own arithmetic; the folded solves take the rest.
constexpr double kLowXlam = 64.0;

// Number of leading modes of g with xlambda < kLowXlam in some field (xlambda grows with the
// mode index, so these modes are a prefix).
int low_mode_prefix(const FieldGroup& g)
{
    int k = 0;
    for (; k < g.nmodes; ++k) {
        bool low = false;
        for (int f = 0; f < g.nf; ++f) low = low || xlambda_of(g.alambda[f], g.alp, g.mode0 + k) < kLowXlam;
        if (!low) break;
    }
    return k;
}

// Modes [k0, k0 + nk) of g as a group of its own.
FieldGroup mode_range(const FieldGroup& g, int k0, int nk)
{
    FieldGroup s = g;
    s.nmodes = nk;
    s.mode0 = g.mode0 + k0;
    const size_t off = (size_t)k0 * g.m * g.m;
    for (int f = 0; f < g.nf; ++f) {
        s.in[f] = g.in[f] + off;
        s.out[f] = g.out[f] + off;
    }
    return s;
}

// Folded passes 2 and 3 as two planar launches: X (pass-1 output) -> G (undivided) in W -> X
// (pass-3 output, divided on load).
int run_fold_middle(const FieldGroup& g, const DevBasis& B, double* X, double* W)
{
    const int NH = B.NH;
    ProjFoldArgs a{};
    a.g = g;
    a.lam = B.flam;
    a.ncols = (long)g.nf * g.nmodes * NH;
    {   // pass 2 undivided: the divide is applied by pass 3 as it loads G
        Timed t(PK_F_P2P_DIV);
        a.Afrag = B.ffwd; a.X = X; a.Y = W;
        HIPCHK(launch_proj_fold(NH, PROJ_P2P, a, L.stream));
    }
    {
        Timed t(PK_F_P2P);
        a.Afrag = B.fbwd; a.X = W; a.Y = X;
        HIPCHK(launch_proj_fold(NH, PROJ_P2P_LDIV, a, L.stream));
    }
    return OK;
}

// The four passes in the even/odd folded basis, then the boundary.  bmode is a BoundaryMode.  g holds
// no singular plane (run_group gives the low Poisson modes to the unfolded passes).
int run_group_fold(const FieldGroup& g, int bidx, int bmode)
{
    DevBasis& B = L.basis[bidx];
    const int NH = B.NH;
    const long cols = (long)g.nf * g.nmodes * NH;
    const long planes = (long)g.nf * g.nmodes * 2;
    int st = ensure_planar((size_t)planes * 4 * NH * NH * sizeof(double));
    if (st) return st;
    ProjFoldArgs a{};
    a.g = g;
    a.lam = B.flam;
    a.ncols = cols;
    {
        Timed t(PK_F_C2P);
        a.Afrag = B.ffwd; a.X = nullptr; a.Y = L.plan[0];
        HIPCHK(launch_proj_fold(NH, PROJ_C2P, a, L.stream));
    }
    if (L.fuse_quarters && proj_quarter_supported(NH)) {
        {   // passes 2 and 3 on LDS-resident quarters, in place on plan[0]
            Timed t(PK_F_MID);
            ProjQuarterArgs q{};
            q.fwd = B.ffwd; q.bwd = B.fbwd; q.lam = B.flam; q.X = L.plan[0];
            q.nplanes = planes; q.g = g;
            HIPCHK(launch_proj_quarter(NH, q, L.stream));
        }
    } else {
        int st1 = run_fold_middle(g, B, L.plan[0], L.plan[1]);
        if (st1) return st1;
    }
    {
        Timed t(PK_F_P2C);
        a.Afrag = B.fbwd; a.X = L.plan[0]; a.Y = nullptr;
        HIPCHK(launch_proj_fold(NH, PROJ_P2C, a, L.stream));
    }
    {
        Timed t(PK_BOUNDARY);
        BoundaryArgs ba{};
        ba.g = g; ba.DC = B.DC; ba.AM1 = B.AM1;
        ba.mode = bidx == 1 ? BND_NEUMANN : bmode;
        HIPCHK(run_boundary(ba));
    }
    return OK;
}

int run_group_unfolded(const FieldGroup& g, int bidx, int bmode);

int run_group(const FieldGroup& g, int bidx, int bmode)
{
    DevBasis& B = L.basis[bidx];
    if (!B.ready) return fail(ERR_NOT_SETUP, "duct_setup_ not called for bound '%c'", bidx ? 'n' : 'd');
    if (g.m != B.m) return fail(ERR_ARG, "m=%d but the '%c' basis was set up for m=%d", g.m, bidx ? 'n' : 'd', B.m);
    if (!(B.NH > 0 && L.fold > bidx)) return run_group_unfolded(g, bidx, bmode);
    // folded, except the low Poisson modes of a Neumann group (kLowXlam)
    const int klow = bidx == 1 ? low_mode_prefix(g) : 0;
    int st;
    if (klow > 0 && (st = run_group_unfolded(mode_range(g, 0, klow), bidx, bmode))) return st;
    if (klow < g.nmodes) return run_group_fold(mode_range(g, klow, g.nmodes - klow), bidx, bmode);
    return OK;
}

// The four FP64 passes of solvediag in the unfolded eigenbasis, the gauge, the boundary.
int run_group_unfolded(const FieldGroup& g, int bidx, int bmode)
{
    DevBasis& B = L.basis[bidx];
    const int NP = B.NP;
    const long cplx_cols = (long)g.nf * g.nmodes * NP;
    const long planes = (long)g.nf * g.nmodes * 2;
    int st = ensure_planar((size_t)planes * NP * NP * sizeof(double));
    if (st) return st;

    ProjArgs a{};
    a.g = g;
    a.lam = B.lam;
    a.gauge = bidx == 1 ? B.gauge : -1;

    {   // pass 1: T = (P^-1 F)^T
        Timed t(PK_C2P);
        a.Afrag = B.Afwd; a.X = nullptr; a.Y = L.plan[0]; a.ncols = cplx_cols;
        HIPCHK(launch_proj(NP, PROJ_C2P, a, L.stream));
    }
    {   // pass 2: G = (P^-1 T)^T, undivided; pass 3 applies ./ (lam_i + lam_j - xlambda) as it loads G
        Timed t(PK_P2P_DIV);
        a.X = L.plan[0]; a.Y = L.plan[1]; a.ncols = planes * NP;
        HIPCHK(launch_proj(NP, PROJ_P2P, a, L.stream));
    }
    if (bidx == 1) {  // Neumann gauge planes: |xlambda| <= 1e-10 (solvediag.f90:151)
        GaugeArgs ga{};
        ga.G = L.plan[1]; ga.f32 = 0; ga.prow = B.prow; ga.pdiag = B.pdiag; ga.wanted = B.wanted;
        ga.lam = B.lam;
        ga.gauge = B.gauge; ga.NP = NP; ga.g = g;
        GaugeList gl{};
        for (int f = 0; f < g.nf; ++f)
            for (int k = 0; k < g.nmodes; ++k) {
                const double xl = xlambda_of(g.alambda[f], g.alp, g.mode0 + k);
                if (!(std::fabs(xl) <= 1.0e-10)) continue;
                const int fm = f * g.nmodes + k;
                for (int part = 0; part < 2; ++part) {
                    if (part == 1 && g.mode0 + k == 0) continue;  // Im skipped at k=0
                    gl.plane[gl.count++] = 2 * fm + part;
                    if (gl.count == kMaxGaugePlanes) {
                        Timed t(PK_GAUGE);
                        HIPCHK(launch_gauge(ga, gl, L.stream));
                        gl.count = 0;
                    }
                }
            }
        if (gl.count) {
            Timed t(PK_GAUGE);
            HIPCHK(launch_gauge(ga, gl, L.stream));
        }
    }
    {   // pass 3: T = (P (G ./ den))^T
        Timed t(PK_P2P);
        a.Afrag = B.Abwd; a.X = L.plan[1]; a.Y = L.plan[0]; a.ncols = planes * NP;
        HIPCHK(launch_proj(NP, PROJ_P2P_LDIV, a, L.stream));
    }
    {   // pass 4: U = (P T)^T = P G P^T into the field interior
        Timed t(PK_P2C);
        a.X = L.plan[0]; a.Y = nullptr; a.ncols = cplx_cols;
        HIPCHK(launch_proj(NP, PROJ_P2C, a, L.stream));
    }
    {
        Timed t(PK_BOUNDARY);
        BoundaryArgs ba{};
        ba.g = g; ba.DC = B.DC; ba.AM1 = B.AM1;
        ba.mode = bidx == 1 ? BND_NEUMANN : bmode;
        HIPCHK(run_boundary(ba));
    }
    return OK;
}

int check_common(int m, int nmodes);

// Adjoint of run_group (the transposed solve): out = S^T in, boundary of out = 0.
// For Neumann the boundary of `in` is first folded into its interior (E^T) in
// the residual scratch, and the gauge adjoint corrects the singular planes.
int run_group_adj(const FieldGroup& g0, int bidx)
{
    DevBasis& B = L.basis[bidx];
    if (!B.ready) return fail(ERR_NOT_SETUP, "duct_setup_ not called for bound '%c'", bidx ? 'n' : 'd');
    if (g0.m != B.m) return fail(ERR_ARG, "m=%d but the basis was set up for m=%d", g0.m, B.m);
    const int NP = B.NP, m = g0.m;
    FieldGroup g = g0;
    int st;
    if (bidx == 1) {
        const size_t field_bytes = (size_t)g.nmodes * m * m * sizeof(d2v);
        if ((st = ensure_buf(L.rfield, L.rfield_bytes, field_bytes * g.nf))) return st;
        BoundaryAdjArgs ba{};
        ba.g = g0;
        ba.DC = B.DC;
        ba.AM1 = B.AM1;
        for (int i = 0; i < g.nf; ++i) ba.g.out[i] = L.rfield + (size_t)i * g.nmodes * m * m;
        {
            Timed t(PK_ADJ);
            HIPCHK(launch_boundary_adj(ba, L.stream));
        }
        for (int i = 0; i < g.nf; ++i) g.in[i] = ba.g.out[i];
    }
    const long cplx_cols = (long)g.nf * g.nmodes * NP;
    const long planes = (long)g.nf * g.nmodes * 2;
    if ((st = ensure_planar((size_t)planes * NP * NP * sizeof(double)))) return st;
    ProjArgs a{};
    a.g = g;
    a.lam = B.lam;
    a.gauge = bidx == 1 ? B.gauge : -1;
    {
        Timed t(PK_C2P);
        a.Afrag = B.AfwdT; a.X = nullptr; a.Y = L.plan[0]; a.ncols = cplx_cols;
        HIPCHK(launch_proj(NP, PROJ_C2P, a, L.stream));
    }
    {
        Timed t(PK_P2P_DIV);
        a.X = L.plan[0]; a.Y = L.plan[1]; a.ncols = planes * NP;
        HIPCHK(launch_proj(NP, PROJ_P2P_DIV, a, L.stream));
    }
    if (bidx == 1) {
        GaugeAdjArgs ga{};
        ga.g = g; ga.G = L.plan[1]; ga.prow = B.prow; ga.pcol = B.pcol; ga.lam = B.lam;
        ga.pdiag = B.pdiag; ga.gauge = B.gauge; ga.NP = NP;
        GaugeList gl{};
        for (int f = 0; f < g.nf; ++f)
            for (int k = 0; k < g.nmodes; ++k) {
                if (!(std::fabs(xlambda_of(g.alambda[f], g.alp, g.mode0 + k)) <= 1.0e-10)) continue;
                for (int part = 0; part < 2; ++part) {
                    if (part == 1 && g.mode0 + k == 0) continue;
                    gl.plane[gl.count++] = 2 * (f * g.nmodes + k) + part;
                    if (gl.count == kMaxGaugePlanes) {
                        Timed t(PK_GAUGE);
                        HIPCHK(launch_gauge_adj(ga, gl, L.stream));
                        gl.count = 0;
                    }
                }
            }
        if (gl.count) {
            Timed t(PK_GAUGE);
            HIPCHK(launch_gauge_adj(ga, gl, L.stream));
        }
    }
    {
        Timed t(PK_P2P);
        a.Afrag = B.AbwdT; a.X = L.plan[1]; a.Y = L.plan[0]; a.ncols = planes * NP;
        HIPCHK(launch_proj(NP, PROJ_P2P, a, L.stream));
    }
    {
        Timed t(PK_P2C);
        a.X = L.plan[0]; a.Y = nullptr; a.ncols = cplx_cols;
        HIPCHK(launch_proj(NP, PROJ_P2C, a, L.stream));
    }
    {
        Timed t(PK_BOUNDARY);
        BoundaryArgs ba{};
        ba.g = g0; ba.DC = B.DC; ba.AM1 = B.AM1;
        ba.mode = BND_OUT_ZERO;
        HIPCHK(run_boundary(ba));
    }
    return OK;
}


// Low-precision pass chain for one field group: sol interior = (or +=) S(in).
int run_group_lp(const FieldGroup& g, int bidx, int scheme, int accumulate, double wanted)
{
    DevBasis& B = L.basis[bidx];
    const int NP = B.NP;
    const long cplx_cols = (long)g.nf * g.nmodes * NP;
    const long planes = (long)g.nf * g.nmodes * 2;
    int st = ensure_planar((size_t)planes * NP * NP * sizeof(float));
    if (st) return st;
    const bool f32 = scheme == LP_F32;
    ProjLpArgs a{};
    a.g = g;
    a.lam = B.lam;
    a.gauge = bidx == 1 ? B.gauge : -1;
    a.accumulate = accumulate;
    float* p0 = (float*)L.plan[0];
    float* p1 = (float*)L.plan[1];
    {
        Timed t(PK_LP_C2P);
        a.Afrag = f32 ? B.lpf_fwd : B.lph_fwd; a.X = nullptr; a.Y = p0; a.ncols = cplx_cols;
        HIPCHK(launch_proj_lp(NP, scheme, PROJ_C2P, a, L.stream));
    }
    {
        Timed t(PK_LP_P2P_DIV);
        a.X = p0; a.Y = p1; a.ncols = planes * NP;
        HIPCHK(launch_proj_lp(NP, scheme, PROJ_P2P_DIV, a, L.stream));
    }
    if (bidx == 1) {
        GaugeArgs ga{};
        ga.G = p1; ga.f32 = 1; ga.prow = B.prow; ga.pdiag = B.pdiag; ga.wanted = wanted;
        ga.relative = accumulate;
        ga.gauge = B.gauge; ga.NP = NP; ga.g = g;
        GaugeList gl{};
        for (int f = 0; f < g.nf; ++f)
            for (int k = 0; k < g.nmodes; ++k) {
                const double xl = xlambda_of(g.alambda[f], g.alp, g.mode0 + k);
                if (!(std::fabs(xl) <= 1.0e-10)) continue;
                const int fm = f * g.nmodes + k;
                for (int part = 0; part < 2; ++part) {
                    if (part == 1 && g.mode0 + k == 0) continue;
                    gl.plane[gl.count++] = 2 * fm + part;
                    if (gl.count == kMaxGaugePlanes) {
                        Timed t(PK_GAUGE);
                        HIPCHK(launch_gauge(ga, gl, L.stream));
                        gl.count = 0;
                    }
                }
            }
        if (gl.count) {
            Timed t(PK_GAUGE);
            HIPCHK(launch_gauge(ga, gl, L.stream));
        }
    }
    {
        Timed t(PK_LP_P2P);
        a.Afrag = f32 ? B.lpf_bwd : B.lph_bwd; a.X = p1; a.Y = p0; a.ncols = planes * NP;
        HIPCHK(launch_proj_lp(NP, scheme, PROJ_P2P, a, L.stream));
    }
    {
        Timed t(PK_LP_P2C);
        a.X = p0; a.Y = nullptr; a.ncols = cplx_cols;
        HIPCHK(launch_proj_lp(NP, scheme, PROJ_P2C, a, L.stream));
    }
    return OK;
}

// Mixed-precision refined solve of a batch:
//   u = S0(f);  repeat niter times: r = f - L u (FP64), u += S1(r);  boundary (FP64).
// stats (host, may be null): [(niter+1) x nf x 2] = max|r_it|, max|f| per field,
// including the residual after the last correction.
// Folded mixed solve of one boundary group (mixed_fold.hip): fold, fused low-precision
// solve, niter x (folded FP64 residual, fused solve), unfold-and-sum, boundary.  The group
// holds no low Poisson mode (run_mixed_fold_split gives those to the unfolded FP64 passes),
// so no plane of it is singular.  stats as in solve_fields_mixed (the last entry is the
// residual after the final correction, computed only when asked for).
int run_group_mixed_fold(const FieldGroup& g, int bidx, int niter, unsigned long long* stats, const int* fidx,
                         int nf_all, bool i8)
{
    DevBasis& B = L.basis[bidx];
    const int NH = B.NH;
    const size_t arr = (size_t)g.nf * g.nmodes * 2 * 4 * NH * NH;   // floats per folded array
    // F, R, U0 (FP64-residual form) or Flo (int8 form: U0 is in digits), D1..Dniter
    const int narr = 2 + 1 + niter;
    int st = ensure_buf(L.mfold, L.mfold_bytes, arr * narr * sizeof(float));
    if (st) return st;
    float* F = L.mfold;
    float* R = L.mfold + arr;
    float* D[kMaxFoldSources];
    for (int t = 0; t < niter + 1; ++t) D[t] = L.mfold + (size_t)(2 + t) * arr;   // D[0] = U0
    float* Flo = i8 ? D[0] : nullptr;
    unsigned char* Ucol = nullptr;
    unsigned char* Urow = nullptr;
    int* gexp = nullptr;
    if (i8) {
        if ((st = ensure_buf(L.mdig, L.mdig_bytes, arr * 2 * kUDig + (arr / ((size_t)NH * NH)) * sizeof(int))))
            return st;
        Ucol = L.mdig;
        Urow = L.mdig + arr * kUDig;
        gexp = (int*)(L.mdig + arr * 2 * kUDig);
    }
    {
        Timed t(PK_LP_C2P);
        HIPCHK(launch_fold32(g, NH, F, L.stream, Flo));
    }
    LpqArgs la{};
    for (int p = 0; p < 2; ++p) {
        la.fwd[p] = B.lq_fwd[p]; la.bwd[p] = B.lq_bwd[p];
        la.sfwd[p] = B.lq_sfwd[p]; la.sbwd[p] = B.lq_sbwd[p];
    }
    la.lam = B.lq_lam;
    auto lpq = [&](const FieldGroup& gg, const float* in, float* out) -> int {
        Timed t(PK_LP_P2P);
        la.g = gg; la.X = in; la.Y = out;
        la.Ucol = nullptr; la.Urow = nullptr; la.gexp = nullptr;
        HIPCHK(launch_lpq(NH, la, L.stream));
        return OK;
    };
    if (i8) {
        // U0 in digit form, the residual on the int8 matrix cores, one correction (niter == 1)
        {
            Timed t(PK_LP_P2P);
            la.g = g; la.X = F; la.Y = nullptr;
            la.Ucol = Ucol; la.Urow = Urow; la.gexp = gexp;
            HIPCHK(launch_lpq(NH, la, L.stream));
        }
        ResidI8Args ra{};
        ra.g = g;
        ra.Fhi = F; ra.Flo = Flo;
        ra.Ucol = Ucol; ra.Urow = Urow; ra.gexp = gexp;
        ra.R = R;
        ra.Sdig[0] = B.sdig[0]; ra.Sdig[1] = B.sdig[1];
        ra.Sexp[0] = B.sexp[0]; ra.Sexp[1] = B.sexp[1];
        ra.stats = stats;
        for (int i = 0; i < g.nf; ++i) ra.slot[i] = fidx[i];
        ra.NH = NH;
        {
            Timed t(PK_RESID);
            HIPCHK(launch_resid_i8(ra, L.stream));
        }
        if ((st = lpq(g, R, D[1]))) return st;
        Timed t(PK_LP_P2C);
        const float* src[kMaxFoldSources] = {D[1], nullptr, nullptr, nullptr};
        HIPCHK(launch_unfold_sum(g, NH, src, 1, 0, L.stream, Ucol, gexp, bidx == 0));
    } else {
        if ((st = lpq(g, F, D[0]))) return st;
        for (int it = 0; it < niter; ++it) {
            ResidFoldArgs ra{};
            ra.g = g;
            for (int t = 0; t <= it; ++t) ra.src[t] = D[t];
            ra.nsrc = it + 1;
            ra.R = R;
            ra.Sfrag[0] = B.sfold[0]; ra.Sfrag[1] = B.sfold[1];
            ra.stats = stats ? stats + (size_t)2 * nf_all * it : nullptr;
            for (int i = 0; i < g.nf; ++i) ra.slot[i] = fidx[i];
            ra.NH = NH;
            {
                Timed t(PK_RESID);
                HIPCHK(launch_resid_fold(ra, L.stream));
            }
            if ((st = lpq(g, R, D[it + 1]))) return st;
        }
        Timed t(PK_LP_P2C);
        const float* src[kMaxFoldSources] = {D[0], niter >= 1 ? D[1] : nullptr, niter >= 2 ? D[2] : nullptr, nullptr};
        HIPCHK(launch_unfold_sum(g, NH, src, niter + 1, 0, L.stream, nullptr, nullptr, bidx == 0));
    }
    if (stats) {   // the residual of the returned solution, by the unfolded FP64 residual kernel
        const size_t fb = (size_t)g.nmodes * g.m * g.m;
        if ((st = ensure_buf(L.rfield, L.rfield_bytes, fb * g.nf * sizeof(d2v)))) return st;
        ResidArgs ra{};
        ra.g = g;
        ra.Dfrag = B.opfrag;
        ra.NP = B.NP;
        ra.stats = stats + (size_t)2 * nf_all * niter;
        ra.neumann = bidx == 1;
        for (int i = 0; i < g.nf; ++i) { ra.r[i] = L.rfield + (size_t)i * fb; ra.slot[i] = fidx[i]; }
        Timed t(PK_RESID);
        HIPCHK(launch_resid(ra, L.stream));
    }
    if (bidx == 1) {   // (the Dirichlet boundary was zeroed by the unfold)
        Timed t(PK_BOUNDARY);
        BoundaryArgs ba{};
        ba.g = g; ba.DC = B.DC; ba.AM1 = B.AM1;
        ba.mode = BND_NEUMANN;
        HIPCHK(run_boundary(ba));
    }
    return OK;
}

// While alive, the library's stream and work buffers (planar, folded, digit and residual arrays) are
// the side stream's (the Neumann group of a batched solve runs there, concurrently with the
// Dirichlet group on the caller's stream).
struct SideLane {
    SideLane() { swap(); }
    ~SideLane() { swap(); }
    static void swap()
    {
        std::swap(L.stream, L.side);
        std::swap(L.plan[0], L.plan2[0]);
        std::swap(L.plan[1], L.plan2[1]);
        std::swap(L.plan_bytes, L.plan2_bytes);
        std::swap(L.mfold, L.mfold2);
        std::swap(L.mfold_bytes, L.mfold2_bytes);
        std::swap(L.mdig, L.mdig2);
        std::swap(L.mdig_bytes, L.mdig2_bytes);
        std::swap(L.rfield, L.rfield2);
        std::swap(L.rfield_bytes, L.rfield2_bytes);
    }
};

int solve_group_mixed(const FieldGroup& g, int bidx, bool folded, bool i8, int scheme0, int scheme1, int niter,
                      int nstat, int nf, const int* fidx);

// The folded refined solve of one boundary group; the low Poisson modes of a Neumann group
// (kLowXlam), the singular k=0 plane among them, take the unfolded FP64 passes instead.
int run_mixed_fold_split(const FieldGroup& g, int bidx, int niter, unsigned long long* stats, const int* fidx,
                         int nf_all, bool i8)
{
    const int klow = bidx == 1 ? low_mode_prefix(g) : 0;
    int st;
    if (klow > 0 && (st = run_group_unfolded(mode_range(g, 0, klow), bidx, BND_NEUMANN))) return st;
    if (klow < g.nmodes)
        return run_group_mixed_fold(mode_range(g, klow, g.nmodes - klow), bidx, niter, stats, fidx, nf_all, i8);
    return OK;
}

int solve_fields_mixed(int nf, const duct_devptr_t* rhs, const duct_devptr_t* sol, const char* bounds,
                       const double* alambda, int m, int nmodes, int ib, double alp, int scheme0, int scheme1,
                       int niter, double* stats)
{
    if (nf < 1 || nf > kMaxFields) return fail(ERR_ARG, "nfields=%d outside [1, %d]", nf, kMaxFields);
    int st = check_common(m, nmodes);
    if (st) return st;
    const bool folded = scheme0 == LP_FOLD16 || scheme1 == LP_FOLD16 || scheme0 == LP_FOLD16Q ||
                        scheme1 == LP_FOLD16Q;
    const bool i8 = scheme0 == LP_FOLD16Q;
    if (folded && scheme0 != scheme1)
        return fail(ERR_ARG, "schemes 5 and 6 (folded) are used for both solves or neither");
    for (int sc : {scheme0, scheme1})
        if (sc < LP_F32 || sc > LP_FOLD16Q) return fail(ERR_ARG, "precision scheme %d is not one of 1..6", sc);
    if (niter < 0 || niter > 16) return fail(ERR_ARG, "niter=%d outside [0, 16]", niter);
    if (folded && niter > 2) return fail(ERR_ARG, "the folded solve takes niter <= 2 (got %d)", niter);
    if (i8 && niter != 1) return fail(ERR_ARG, "scheme 6 takes exactly one refinement (got niter=%d)", niter);
    const size_t field_bytes = (size_t)nmodes * m * m * sizeof(d2v);
    if ((st = ensure_buf(L.rfield, L.rfield_bytes, field_bytes * nf))) return st;
    const int nstat = stats ? (niter + 1) : 0;
    if (nstat) {
        if ((st = ensure_buf(L.stats, L.stats_bytes, sizeof(unsigned long long) * 2 * nf * nstat))) return st;
        HIPCHK(hipMemsetAsync(L.stats, 0, sizeof(unsigned long long) * 2 * nf * nstat, L.stream));
    }
    FieldGroup gs[2]{};
    int fidxs[2][kMaxFields];
    for (int bidx = 0; bidx < 2; ++bidx) {
        FieldGroup& g = gs[bidx];
        g.nmodes = nmodes; g.mode0 = ib; g.m = m; g.n = m - 2; g.alp = alp;
        for (int f = 0; f < nf; ++f) {
            const int bi = bound_index(bounds[f]);
            if (bi < 0) return fail(ERR_ARG, "bound '%c' of field %d is neither 'd' nor 'n'", bounds[f], f);
            if (bi != bidx) continue;
            if (!rhs[f] || !sol[f]) return fail(ERR_ARG, "null device pointer for field %d", f);
            fidxs[bidx][g.nf] = f;
            g.in[g.nf] = (const d2v*)rhs[f];
            g.out[g.nf] = (d2v*)sol[f];
            g.alambda[g.nf] = alambda[f];
            ++g.nf;
        }
        if (g.nf == 0) continue;
        DevBasis& B = L.basis[bidx];
        if (!B.ready) return fail(ERR_NOT_SETUP, "duct_setup_ not called for bound '%c'", bidx ? 'n' : 'd');
        if (g.m != B.m) return fail(ERR_ARG, "m=%d but the basis was set up for m=%d", g.m, B.m);
        if (folded && !B.lq_ready)
            return fail(ERR_UNSUPPORTED, "scheme 5 needs a folding basis with ceil((m-2)/2) in (32, 128] "
                        "(m=%d, bound '%c')", m, bidx ? 'n' : 'd');
    }
    // The folded solve with both groups: the Neumann group on the side stream with its own work
    // arrays, so that its latency-bound kernels (the low modes' FP64 passes, gauge, edges) fill the
    // gaps of the Dirichlet group's (not while launches are timed).
    if (folded && gs[0].nf && gs[1].nf && L.overlap && !L.prof) {
        if (!L.side) {
            HIPCHK(hipStreamCreateWithFlags(&L.side, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&L.ev_fork, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&L.ev_join, hipEventDisableTiming));
        }
        HIPCHK(hipEventRecord(L.ev_fork, L.stream));
        HIPCHK(hipStreamWaitEvent(L.side, L.ev_fork, 0));
        st = run_group_mixed_fold(gs[0], 0, niter, nstat ? L.stats : nullptr, fidxs[0], nf, i8);
        int st_side;
        {
            SideLane lane;
            st_side = run_mixed_fold_split(gs[1], 1, niter, nstat ? L.stats : nullptr, fidxs[1], nf, i8);
        }
        HIPCHK(hipEventRecord(L.ev_join, L.side));
        HIPCHK(hipStreamWaitEvent(L.stream, L.ev_join, 0));
        if (st || st_side) return st ? st : st_side;
    } else {
        for (int bidx = 0; bidx < 2; ++bidx) {
            if (gs[bidx].nf == 0) continue;
            if ((st = solve_group_mixed(gs[bidx], bidx, folded, i8, scheme0, scheme1, niter, nstat, nf,
                                        fidxs[bidx])))
                return st;
        }
    }
    if (nstat) {
        std::vector<unsigned long long> h((size_t)2 * nf * nstat);
        HIPCHK(hipMemcpyAsync(h.data(), L.stats, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              L.stream));
        HIPCHK(hipStreamSynchronize(L.stream));
        for (size_t e = 0; e < h.size(); ++e) stats[e] = __builtin_bit_cast(double, h[e]);
    }
    return OK;
}

// One boundary group of solve_fields_mixed, on the library's current stream.
int solve_group_mixed(const FieldGroup& g, int bidx, bool folded, bool i8, int scheme0, int scheme1, int niter,
                      int nstat, int nf, const int* fidx)
{
    DevBasis& B = L.basis[bidx];
    const int m = g.m, nmodes = g.nmodes;
    int st;
    {
        if (folded) return run_mixed_fold_split(g, bidx, niter, nstat ? L.stats : nullptr, fidx, nf, i8);
        if ((st = run_group_lp(g, bidx, scheme0, 0, B.wanted))) return st;
        ResidArgs ra{};
        ra.g = g;
        ra.Dfrag = B.opfrag;
        ra.NP = B.NP;
        for (int i = 0; i < g.nf; ++i) ra.r[i] = L.rfield + (size_t)i * nmodes * m * m;
        FieldGroup gc = g;   // correction: right-hand side = residual, accumulated into sol
        for (int i = 0; i < g.nf; ++i) gc.in[i] = ra.r[i];
        for (int it = 0; it <= niter; ++it) {
            if (it == niter && !nstat) break;   // the last residual is only for the report
            {
                Timed t(PK_RESID);
                ra.stats = nstat ? L.stats + (size_t)2 * nf * it : nullptr;
                for (int i = 0; i < g.nf; ++i) ra.slot[i] = fidx[i];
                ra.neumann = bidx == 1;
                HIPCHK(launch_resid(ra, L.stream));
            }
            if (bidx == 1) {   // drop the null-space part of the singular planes' residual
                NullProjArgs np{};
                np.q = B.qrow; np.p = B.pcol; np.stats = ra.stats;
                np.nmodes = g.nmodes; np.m = g.m; np.n = g.n;
                for (int i = 0; i < g.nf; ++i) { np.r[i] = ra.r[i]; np.slot[i] = fidx[i]; }
                GaugeList gl{};
                for (int f = 0; f < g.nf; ++f)
                    for (int k = 0; k < g.nmodes; ++k) {
                        if (!(std::fabs(xlambda_of(g.alambda[f], g.alp, g.mode0 + k)) <= 1.0e-10)) continue;
                        for (int part = 0; part < 2; ++part) {
                            if (part == 1 && g.mode0 + k == 0) continue;
                            gl.plane[gl.count++] = 2 * (f * g.nmodes + k) + part;
                            if (gl.count == kMaxGaugePlanes) {
                                Timed t(PK_GAUGE);
                                HIPCHK(launch_nullproj(np, gl, L.stream));
                                gl.count = 0;
                            }
                        }
                    }
                if (gl.count) {
                    Timed t(PK_GAUGE);
                    HIPCHK(launch_nullproj(np, gl, L.stream));
                }
            }
            if (it == niter) break;
            if ((st = run_group_lp(gc, bidx, scheme1, 1, B.wanted))) return st;
        }
        {
            Timed t(PK_BOUNDARY);
            BoundaryArgs ba{};
            ba.g = g; ba.DC = B.DC; ba.AM1 = B.AM1;
            ba.mode = bidx == 1 ? BND_NEUMANN : BND_DIR_ZERO;
            HIPCHK(run_boundary(ba));
        }
    }
    return OK;
}

int check_common(int m, int nmodes)
{
    if (m < kMinM || m > kMaxM) return fail(ERR_UNSUPPORTED, "m=%d outside [%d, %d]", m, kMinM, kMaxM);
    if (nmodes < 1) return fail(ERR_ARG, "nmodes=%d < 1", nmodes);
    return OK;
}

// Solve a batch of device fields: group by boundary type, one run per group.
int solve_fields(int nf, const duct_devptr_t* rhs, const duct_devptr_t* sol, const char* bounds,
                 const double* alambda, int m, int nmodes, int ib, double alp, int dir_bmode)
{
    if (nf < 1 || nf > kMaxFields) return fail(ERR_ARG, "nfields=%d outside [1, %d]", nf, kMaxFields);
    int st = check_common(m, nmodes);
    if (st) return st;
    FieldGroup g[2]{};
    for (int bidx = 0; bidx < 2; ++bidx) {
        g[bidx].nmodes = nmodes; g[bidx].mode0 = ib; g[bidx].m = m; g[bidx].n = m - 2; g[bidx].alp = alp;
        for (int f = 0; f < nf; ++f) {
            const int bi = bound_index(bounds[f]);
            if (bi < 0) return fail(ERR_ARG, "bound '%c' of field %d is neither 'd' nor 'n'", bounds[f], f);
            if (bi != bidx) continue;
            if (!rhs[f] || !sol[f]) return fail(ERR_ARG, "null device pointer for field %d", f);
            FieldGroup& G = g[bidx];
            G.in[G.nf] = (const d2v*)rhs[f];
            G.out[G.nf] = (d2v*)sol[f];
            G.alambda[G.nf] = alambda[f];
            ++G.nf;
        }
    }
    // Both groups present: the Neumann group runs on the side stream, so its latency-bound kernels
    // (k=0 gauge, edges) overlap the Dirichlet group's passes. Not while launches are timed: the
    // per-launch hipEvents would then measure kernels that share the GPU.
    if (g[0].nf && g[1].nf && L.overlap && !L.prof) {
        if (!L.side) {
            HIPCHK(hipStreamCreateWithFlags(&L.side, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&L.ev_fork, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&L.ev_join, hipEventDisableTiming));
        }
        HIPCHK(hipEventRecord(L.ev_fork, L.stream));
        HIPCHK(hipStreamWaitEvent(L.side, L.ev_fork, 0));
        st = run_group(g[0], 0, dir_bmode);
        int st_side;
        {
            SideLane lane;
            st_side = run_group(g[1], 1, dir_bmode);
        }
        HIPCHK(hipEventRecord(L.ev_join, L.side));
        HIPCHK(hipStreamWaitEvent(L.stream, L.ev_join, 0));
        return st ? st : st_side;
    }
    for (int bidx = 0; bidx < 2; ++bidx) {
        if (g[bidx].nf == 0) continue;
        st = run_group(g[bidx], bidx, dir_bmode);
        if (st) return st;
    }
    return OK;
}

// The advection x-DFT tables for (M, X), rebuilt when either changes.
int ensure_xdft(int M, int X)
{
    if (L.xdft_M == M && L.xdft_X == X) return OK;
    std::vector<double> ei, fw;
    build_xdft(M, X, ei, fw);
    HIPCHK(hipStreamSynchronize(L.stream));
    free_ptr(L.xdft[0]);
    free_ptr(L.xdft[1]);
    L.xdft_M = L.xdft_X = 0;
    HIPCHK(hipMalloc((void**)&L.xdft[0], ei.size() * sizeof(double)));
    HIPCHK(hipMemcpy(L.xdft[0], ei.data(), ei.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc((void**)&L.xdft[1], fw.size() * sizeof(double)));
    HIPCHK(hipMemcpy(L.xdft[1], fw.data(), fw.size() * sizeof(double), hipMemcpyHostToDevice));
    L.xdft_M = M;
    L.xdft_X = X;
    return OK;
}

// Independent derivative products (y or z, each its own field, matrix, scale and beta) in one
// launch on the library's stream, with the derivative matrices of duct_setup_ (padded to L.dc_np).
int cline_run(int mm, int M, std::initializer_list<ClineOp> ops)
{
    ClineArgs c{};
    c.nmodes = M;
    c.m = mm;
    c.NP = L.dc_np;
    for (const ClineOp& o : ops) {
        if (c.nops == kMaxClineOps) return fail(ERR_ARG, "more than %d derivative products in one launch", kMaxClineOps);
        c.op[c.nops++] = o;
    }
    Timed t(PK_DERIV);
    HIPCHK(launch_cline_ops(c, L.stream));
    return OK;
}

// advection3d (operators.F:9-276) of device fields: the x-transforms and the six products in
// k_xprod, then the nine derivative terms in the order of operators.F:91-272 per output
// (adv1: uu_x, uv_y, uw_z; adv2: vv_y, uv_x, vw_z; adv3: ww_z, uw_x, vw_y); the caller has
// checked the sizes.
// The dense x-DFT kernels (k_xprod, k_fourx) cover M % 16 == 0, M <= 64, X % 16 == 0, X <= 192;
// other sizes go through batched rocFFT real transforms.
bool dense_x(int M, int X)
{
    return M % 16 == 0 && M <= kXprodMaxM && X % 16 == 0 && X <= kXprodMaxX;
}

// One batched real transform of X-point lines on the library's stream: forward (r2c: real in,
// X/2+1 Hermitian out) or inverse (c2r; its input may be overwritten).  Strides and distances in
// elements (doubles for the real side, complex values for the Hermitian side).  Plans are kept.
int xfft_run(bool forward, void* in, void* out, size_t X, size_t batch, size_t is, size_t id, size_t os, size_t od)
{
    if (!L.rocfft_on) {
        if (rocfft_setup() != rocfft_status_success) return fail(ERR_HIP, "rocfft_setup failed");
        L.rocfft_on = true;
    }
    Lib::XfftPlan* pl = nullptr;
    for (auto& p : L.xplans)
        if (p.forward == (int)forward && p.X == X && p.batch == batch && p.is == is && p.id == id && p.os == os &&
            p.od == od)
            pl = &p;
    if (!pl) {
        rocfft_plan_description desc = nullptr;
        if (rocfft_plan_description_create(&desc) != rocfft_status_success)
            return fail(ERR_HIP, "rocfft_plan_description_create failed");
        const rocfft_array_type ti = forward ? rocfft_array_type_real : rocfft_array_type_hermitian_interleaved;
        const rocfft_array_type to = forward ? rocfft_array_type_hermitian_interleaved : rocfft_array_type_real;
        rocfft_status rs = rocfft_plan_description_set_data_layout(desc, ti, to, nullptr, nullptr, 1, &is, id, 1, &os, od);
        rocfft_plan plan = nullptr;
        if (rs == rocfft_status_success)
            rs = rocfft_plan_create(&plan, rocfft_placement_notinplace,
                                    forward ? rocfft_transform_type_real_forward : rocfft_transform_type_real_inverse,
                                    rocfft_precision_double, 1, &X, batch, desc);
        (void)rocfft_plan_description_destroy(desc);
        if (rs != rocfft_status_success) return fail(ERR_HIP, "rocfft_plan_create failed (%d)", (int)rs);
        size_t work = 0;
        if (rocfft_plan_get_work_buffer_size(plan, &work) != rocfft_status_success) {
            (void)rocfft_plan_destroy(plan);
            return fail(ERR_HIP, "rocfft_plan_get_work_buffer_size failed");
        }
        L.xplans.push_back(Lib::XfftPlan{(int)forward, X, batch, is, id, os, od, plan, work});
        pl = &L.xplans.back();
    }
    if (!L.xinfo && rocfft_execution_info_create(&L.xinfo) != rocfft_status_success)
        return fail(ERR_HIP, "rocfft_execution_info_create failed");
    if (rocfft_execution_info_set_stream(L.xinfo, L.stream) != rocfft_status_success)
        return fail(ERR_HIP, "rocfft_execution_info_set_stream failed");
    if (pl->work) {
        int st = ensure_buf(L.xwork, L.xwork_bytes, pl->work);
        if (st) return st;
        if (rocfft_execution_info_set_work_buffer(L.xinfo, L.xwork, pl->work) != rocfft_status_success)
            return fail(ERR_HIP, "rocfft_execution_info_set_work_buffer failed");
    }
    void* ib[1] = {in};
    void* ob[1] = {out};
    if (rocfft_execute(pl->plan, ib, ob, L.xinfo) != rocfft_status_success) return fail(ERR_HIP, "rocfft_execute failed");
    HIPCHK(hipGetLastError());
    return OK;
}

// The six products of the advection term (k_xprod's output, prods (m, m, M) each) through rocFFT:
// fill (modes M..X/2 zero), c2r of u, v, w, the products in physical space, r2c, pack with 1/X.
int xprod_fft(const d2v* const in[3], d2v* prods, int mm, int M, int X)
{
    const long lines = (long)mm * mm;
    const int H = X / 2 + 1;
    int st;
    if ((st = ensure_buf(L.xw, L.xw_bytes, (size_t)6 * H * lines * sizeof(d2v)))) return st;
    if ((st = ensure_buf(L.xp, L.xp_bytes, (size_t)9 * X * lines * sizeof(double)))) return st;
    Timed t(PK_XPROD);
    XpadArgs pa{};
    for (int f = 0; f < 3; ++f) pa.in[f] = in[f];
    pa.out = L.xw;
    pa.lines = lines; pa.nf = 3; pa.M = M; pa.H = H; pa.X = X;
    HIPCHK(launch_xpad(pa, false, L.stream));
    double* phys = L.xp;                          // u, v, w: [line][x]
    double* prodp = L.xp + (size_t)3 * X * lines; // uu, vv, ww, uv, uw, vw: [line][x]
    for (int f = 0; f < 3; ++f)
        if ((st = xfft_run(false, L.xw + (size_t)f * H * lines, phys + (size_t)f * X * lines, X, lines, lines, 1, 1, X)))
            return st;
    HIPCHK(launch_xmul6(phys, phys + (size_t)X * lines, phys + (size_t)2 * X * lines, prodp, (long)X * lines,
                        L.stream));
    for (int p = 0; p < 6; ++p)
        if ((st = xfft_run(true, prodp + (size_t)p * X * lines, L.xw + (size_t)p * H * lines, X, lines, 1, X, lines, 1)))
            return st;
    XpackArgs ka{};
    ka.in = L.xw;
    for (int p = 0; p < 6; ++p) ka.out[p] = prods + (size_t)p * M * lines;
    ka.lines = lines; ka.nf = 6; ka.M = M; ka.H = H; ka.X = X; ka.line_major = 0;
    HIPCHK(launch_xpack(ka, L.stream));
    return OK;
}

int advection_run(const d2v* const in[3], d2v* const out[3], int mm, int M, int X, double alp, double aspect)
{
    int st;
    const size_t field = (size_t)M * mm * mm;
    if ((st = ensure_buf(L.prods, L.prods_bytes, 6 * field * sizeof(d2v)))) return st;
    if (X == 192 && M <= 64) {   // the x-transforms as 64-point FFTs across DPP rows
        if (!L.xlane) {
            std::vector<double> tab;
            build_xlane_table(tab);
            HIPCHK(hipMalloc((void**)&L.xlane, tab.size() * sizeof(double)));
            HIPCHK(hipMemcpy(L.xlane, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        XprodArgs xa{};
        for (int i = 0; i < 3; ++i) xa.in[i] = in[i];
        for (int p = 0; p < 6; ++p) xa.out[p] = L.prods + p * field;
        xa.m = mm; xa.M = M; xa.X = X;
        Timed t(PK_XPROD);
        HIPCHK(launch_xprod_lane(xa, L.xlane, L.stream));
    } else if (!dense_x(M, X)) {
        if ((st = xprod_fft(in, L.prods, mm, M, X))) return st;
    } else {
        if ((st = ensure_xdft(M, X))) return st;
        XprodArgs xa{};
        for (int i = 0; i < 3; ++i) xa.in[i] = in[i];
        for (int p = 0; p < 6; ++p) xa.out[p] = L.prods + p * field;
        xa.Einv = L.xdft[0];
        xa.Fwd = L.xdft[1];
        xa.m = mm; xa.M = M; xa.X = X;
        Timed t(PK_XPROD);
        HIPCHK(launch_xprod(xa, L.stream));
    }
    // Each output keeps that order of its three terms; the y and z terms run as two grouped
    // launches of three independent products.
    const d2v* P[6];
    for (int p = 0; p < 6; ++p) P[p] = L.prods + p * field;
    enum { UU, VV, WW, UV, UW, VW };
    const double* D = L.dfold[0];   // D folded (J D J = -D)
    const double sz = 1.0 / aspect;
    {
        Timed tm(PK_DERIV);
        HIPCHK(launch_derivx(P[UU], out[0], mm, M, 0, alp, 1.0, 0.0, L.stream));
    }
    if ((st = cline_run(mm, M, {{P[UV], out[0], D, 1.0, 1.0, 0, -1}, {P[VV], out[1], D, 1.0, 0.0, 0, -1},
                                {P[WW], out[2], D, sz, 0.0, 1, -1}})))
        return st;
    {
        Timed tm(PK_DERIV);
        HIPCHK(launch_derivx(P[UV], out[1], mm, M, 0, alp, 1.0, 1.0, L.stream));
        HIPCHK(launch_derivx(P[UW], out[2], mm, M, 0, alp, 1.0, 1.0, L.stream));
    }
    return cline_run(mm, M, {{P[UW], out[0], D, sz, 1.0, 1, -1}, {P[VW], out[1], D, sz, 1.0, 1, -1},
                             {P[VW], out[2], D, 1.0, 1.0, 0, -1}});
}

}  // namespace

extern "C" {

int duct_init_(const int* device)
{
    if (!device) return fail(ERR_ARG, "null device");
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    if (*device < 0 || *device >= count) return fail(ERR_ARG, "device %d not in [0, %d)", *device, count);
    HIPCHK(hipSetDevice(*device));
    L.device = *device;
    L.init = true;
    return noalloc_scratch_setup(*device);
}

int duct_shutdown_(void)
{
    if (!L.init) return OK;
    (void)hipStreamSynchronize(L.stream);
    if (L.side) (void)hipStreamSynchronize(L.side);
    for (auto& B : L.basis) free_basis(B);
    for (auto& p : L.plan) free_ptr(p);
    for (auto& p : L.plan2) free_ptr(p);
    L.plan2_bytes = 0;
    free_ptr(L.mfold2);
    free_ptr(L.mdig2);
    free_ptr(L.rfield2);
    L.mfold2_bytes = L.mdig2_bytes = L.rfield2_bytes = 0;
    if (L.ev_fork) (void)hipEventDestroy(L.ev_fork);
    if (L.ev_join) (void)hipEventDestroy(L.ev_join);
    if (L.ev_switch) (void)hipEventDestroy(L.ev_switch);
    if (L.side) (void)hipStreamDestroy(L.side);
    L.ev_fork = L.ev_join = L.ev_switch = nullptr;
    L.stream = nullptr;   // the caller's stream may not outlive the library state
    L.side = nullptr;
    for (auto& p : L.hfield) free_ptr(p);
    free_ptr(L.rfield);
    free_ptr(L.stats);
    free_ptr(L.dcfrag[0]);
    free_ptr(L.dcfrag[1]);
    free_ptr(L.tfrag[0]);
    free_ptr(L.tfrag[1]);
    free_ptr(L.d2frag);
    free_ptr(L.dfold[0]);
    free_ptr(L.dfold[1]);
    L.dc_m = L.dc_np = 0;
    free_ptr(L.rk);
    L.rk_bytes = 0;
    free_temp();
    for (auto& S : L.shen) free_shen(S);
    free_ptr(L.sgam);
    L.sgam_bytes = 0;
    free_ptr(L.tfield);
    L.tfield_bytes = 0;
    free_ptr(L.bscr);
    L.bscr_bytes = 0;
    free_ptr(L.xlane);
    free_ptr(L.xdft[0]);
    free_ptr(L.xdft[1]);
    L.xdft_M = L.xdft_X = 0;
    free_ptr(L.prods);
    L.prods_bytes = 0;
    for (auto& xp : L.xplans) (void)rocfft_plan_destroy(xp.plan);
    L.xplans.clear();
    if (L.xinfo) (void)rocfft_execution_info_destroy(L.xinfo);
    L.xinfo = nullptr;
    if (L.rocfft_on) (void)rocfft_cleanup();
    L.rocfft_on = false;
    free_ptr(L.xwork);
    free_ptr(L.xw);
    free_ptr(L.xp);
    L.xwork_bytes = L.xw_bytes = L.xp_bytes = 0;
    free_ptr(L.mfold);
    L.mfold_bytes = 0;
    free_ptr(L.mdig);
    L.mdig_bytes = 0;
    L.plan_bytes = L.hfield_bytes = L.rfield_bytes = L.stats_bytes = 0;
    for (auto& r : L.recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto& e : L.pool) (void)hipEventDestroy(e);
    L.recs.clear();
    L.pool.clear();
    L.prof = false;
    L.stream = nullptr;
    L.init = false;
    return OK;
}

int duct_set_overlap_(const int* on)
{
    if (!on) return fail(ERR_ARG, "null argument to duct_set_overlap_");
    int st = ensure_init();
    if (st) return st;
    L.overlap = *on ? 1 : 0;
    return OK;
}

int duct_set_fused_quarters_(const int* on)
{
    if (!on) return fail(ERR_ARG, "null argument to duct_set_fused_quarters_");
    int st = ensure_init();
    if (st) return st;
    L.fuse_quarters = *on ? 1 : 0;
    return OK;
}

int duct_set_stream_(const duct_devptr_t* stream)
{
    if (!stream) return fail(ERR_ARG, "null stream pointer");
    int st = ensure_init();
    if (st) return st;
    const hipStream_t ns = (hipStream_t)(*stream);
    if (ns == L.stream) return OK;
    // Every call shares the library's scratch buffers (planar intermediates, folded arrays, residual
    // fields), so work still queued on the old stream may yet read or write them: the new stream
    // waits for it.  Under graph capture of either stream the caller orders the streams itself.
    hipStreamCaptureStatus cold = hipStreamCaptureStatusNone, cnew = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(L.stream, &cold));
    HIPCHK(hipStreamIsCapturing(ns, &cnew));
    if (cold == hipStreamCaptureStatusNone && cnew == hipStreamCaptureStatusNone) {
        if (!L.ev_switch) HIPCHK(hipEventCreateWithFlags(&L.ev_switch, hipEventDisableTiming));
        HIPCHK(hipEventRecord(L.ev_switch, L.stream));
        HIPCHK(hipStreamWaitEvent(ns, L.ev_switch, 0));
    }
    L.stream = ns;
    return OK;
}

int duct_synchronize_(void)
{
    int st = ensure_init();
    if (st) return st;
    HIPCHK(hipStreamSynchronize(L.stream));
    return OK;
}

int duct_last_error_(char* buf, const int* len)
{
    if (!buf || !len || *len <= 0) return ERR_ARG;
    snprintf(buf, (size_t)*len, "%s", L.err.c_str());
    return OK;
}

int duct_setup_(const char* bound, const int* m_, const double* P, const double* Pinv, const double* lam,
                const double* DC, const double* AM1, const int* iofzero, const double* wanted,
                size_t bound_len)
{
    if (!bound || !m_ || !P || !Pinv || !lam || !DC) return fail(ERR_ARG, "null argument to duct_setup_");
    (void)bound_len;
    const int bidx = bound_index(bound[0]);
    if (bidx < 0) return fail(ERR_ARG, "in pseudomat error: bound '%c'", bound[0]);
    const int m = *m_;
    if (m < kMinM || m > kMaxM) return fail(ERR_UNSUPPORTED, "m=%d outside [%d, %d]", m, kMinM, kMaxM);
    if (bidx == 1 && (!AM1 || !iofzero || !wanted)) return fail(ERR_ARG, "Neumann setup needs AM1, iofzero, wanted");
    const int n = m - 2, NP = padded_n(n);
    if (bidx == 1 && (*iofzero < 1 || *iofzero > n)) return fail(ERR_ARG, "iofzero=%d outside [1, %d]", *iofzero, n);
    int st = ensure_init();
    if (st) return st;

    DevBasis& B = L.basis[bidx];
    HIPCHK(hipStreamSynchronize(L.stream));
    free_basis(B);

    std::vector<double> fwd((size_t)NP * NP), bwd((size_t)NP * NP), lp(NP, 0.0), d2b((size_t)m * 2, 0.0),
        prow(NP, 0.0), am1(16, 0.0);
    build_frag(Pinv, n, n, NP, fwd.data());
    build_frag(P, n, n, NP, bwd.data());
    for (int i = 0; i < n; ++i) lp[i] = lam[i];
    for (int i = 0; i < m; ++i) {  // D2(:,1) and D2(:,m) of D2 = DC*DC (pseudomat.f90:110, :225-230)
        double s0 = 0.0, s1 = 0.0;
        for (int k = 0; k < m; ++k) {
            s0 += DC[i + (size_t)k * m] * DC[k];
            s1 += DC[i + (size_t)k * m] * DC[k + (size_t)(m - 1) * m];
        }
        d2b[i] = s0;
        d2b[i + m] = s1;
    }
    if (bidx == 1) {
        const int io = *iofzero - 1;
        for (int k = 0; k < n; ++k) prow[k] = P[io + (size_t)k * n];
        B.pdiag = P[io + (size_t)io * n];
        B.wanted = *wanted;
        B.gauge = io;
        for (int e = 0; e < 16; ++e) am1[e] = AM1[e];
    }
    auto up = [&](double*& dst, const double* src, size_t count) -> int {
        HIPCHK(hipMalloc((void**)&dst, count * sizeof(double)));
        HIPCHK(hipMemcpy(dst, src, count * sizeof(double), hipMemcpyHostToDevice));
        return OK;
    };
    if ((st = up(B.Afwd, fwd.data(), fwd.size()))) return st;
    if ((st = up(B.Abwd, bwd.data(), bwd.size()))) return st;
    {   // adjoint bases: P^T and P^-T
        std::vector<double> T((size_t)n * n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) T[i + (size_t)j * n] = P[j + (size_t)i * n];
        build_frag(T.data(), n, n, NP, fwd.data());
        if ((st = up(B.AfwdT, fwd.data(), fwd.size()))) return st;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) T[i + (size_t)j * n] = Pinv[j + (size_t)i * n];
        build_frag(T.data(), n, n, NP, bwd.data());
        if ((st = up(B.AbwdT, bwd.data(), bwd.size()))) return st;
    }
    {   // collocation derivative matrices of the derivative products (DC, DC^T, D2), padded to a
        // multiple of 16: built here from m in extended precision, not from the caller's DC
        const int MP = (m + 15) / 16 * 16;
        std::vector<double> f((size_t)MP * MP), T((size_t)m * m), d1, d2;
        build_colloc_matrices(m, d1, d2);
        free_ptr(L.dcfrag[0]);
        free_ptr(L.dcfrag[1]);
        build_frag(d1.data(), m, m, MP, f.data());
        if ((st = up(L.dcfrag[0], f.data(), f.size()))) return st;
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) T[i + (size_t)j * m] = d1[j + (size_t)i * m];
        build_frag(T.data(), m, m, MP, f.data());
        if ((st = up(L.dcfrag[1], f.data(), f.size()))) return st;
        std::vector<double> tf, ti;
        build_cheb_transforms(m, tf, ti);
        free_ptr(L.tfrag[0]);
        free_ptr(L.tfrag[1]);
        build_frag(tf.data(), m, m, MP, f.data());
        if ((st = up(L.tfrag[0], f.data(), f.size()))) return st;
        build_frag(ti.data(), m, m, MP, f.data());
        if ((st = up(L.tfrag[1], f.data(), f.size()))) return st;
        // D2 = DC DC (pseudomat.f90:110), the full second-derivative matrix of laplacian3d
        free_ptr(L.d2frag);
        build_frag(d2.data(), m, m, MP, f.data());
        if ((st = up(L.d2frag, f.data(), f.size()))) return st;
        {   // the same products in half the work (k_cline_fold)
            std::vector<double> f1, f2;
            build_colloc_folded(m, f1, f2);
            free_ptr(L.dfold[0]);
            free_ptr(L.dfold[1]);
            if ((st = up(L.dfold[0], f1.data(), f1.size()))) return st;
            if ((st = up(L.dfold[1], f2.data(), f2.size()))) return st;
        }
        L.dc_m = m;
        L.dc_np = MP;
    }
    // the interior operator D2B of the solve (pseudomat.f90:110-128, :225-240), column-major
    std::vector<double> op((size_t)n * n, 0.0);
    {
        std::vector<double> D2((size_t)m * m, 0.0);
        for (int j = 0; j < m; ++j)          // D2 = DC * DC (pseudomat.f90:110)
            for (int k = 0; k < m; ++k) {
                const double b = DC[k + (size_t)j * m];
                for (int i = 0; i < m; ++i) D2[i + (size_t)j * m] += DC[i + (size_t)k * m] * b;
            }
        const int N = m - 1;
        const double Det = DC[0] * DC[N + (size_t)N * m] - DC[(size_t)N * m] * DC[N];
        for (int jj = 0; jj < n; ++jj)
            for (int ii = 0; ii < n; ++ii) {
                const int i = ii + 1, j = jj + 1;
                double v = D2[i + (size_t)j * m];               // Dirichlet: pseudomat.f90:232-240
                if (bidx == 1) {                                 // Neumann: :114-128
                    const double ter1 = D2[i] * (DC[(size_t)N * m] * DC[N + (size_t)j * m] -
                                                 DC[N + (size_t)N * m] * DC[(size_t)j * m]) / Det;
                    const double ter3 = D2[i + (size_t)N * m] * (DC[N] * DC[(size_t)j * m] -
                                                                 DC[0] * DC[N + (size_t)j * m]) / Det;
                    v = ter1 + v + ter3;
                }
                op[ii + (size_t)jj * n] = v;
            }
    }
    {   // even/odd folded bases (used when duct_set_fold_ enables them)
        const int NH = fold_size(n);
        if (NH > 0) {
            std::vector<double> ff((size_t)2 * NH * NH), fb((size_t)2 * NH * NH), fl((size_t)2 * NH);
            std::vector<int> pos(n);
            bool ok = build_fold_bases(P, Pinv, lam, n, NH, ff.data(), fb.data(), fl.data(), pos.data(),
                                       op.data()) == 0;
            if (ok) {
                ff.resize(ff.size() + kFoldFragSlack, 0.0);
                fb.resize(fb.size() + kFoldFragSlack, 0.0);
                if ((st = up(B.ffwd, ff.data(), ff.size()))) return st;
                if ((st = up(B.fbwd, fb.data(), fb.size()))) return st;
                if ((st = up(B.flam, fl.data(), fl.size()))) return st;
                B.NH = NH;
            }
        }
    }
    if ((st = up(B.lam, lp.data(), lp.size()))) return st;
    if ((st = up(B.DC, DC, (size_t)m * m))) return st;
    if ((st = up(B.AM1, am1.data(), 16))) return st;
    if ((st = up(B.D2b, d2b.data(), d2b.size()))) return st;
    if ((st = up(B.prow, prow.data(), prow.size()))) return st;

    // mixed-precision solve: low-precision bases and the FP64 interior operator
    {
        std::vector<double> opf((size_t)NP * NP);
        build_frag(op.data(), n, n, NP, opf.data());
        if ((st = up(B.opfrag, opf.data(), opf.size()))) return st;
        if (B.NH == 64 || B.NH == 128) {   // folded mixed solve (LP_FOLD16)
            const int NH = B.NH, he = (n + 1) / 2, ho = n / 2;
            std::vector<double> Af((size_t)2 * NH * NH), Ab((size_t)2 * NH * NH), fl((size_t)2 * NH);
            std::vector<int> pos(n);
            if (fold_matrices(P, Pinv, lam, n, NH, Af.data(), Ab.data(), fl.data(), pos.data(), op.data()) == 0) {
                auto upraw2 = [&](void*& dst, const void* src, size_t bytes) -> int {
                    HIPCHK(hipMalloc(&dst, bytes));
                    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
                    return OK;
                };
                std::vector<unsigned short> fr;
                for (int par = 0; par < 2; ++par) {
                    B.lq_sfwd[par] = build_lpq_basis(Af.data() + (size_t)par * NH, 2 * NH, NH, fr);
                    if ((st = upraw2(B.lq_fwd[par], fr.data(), fr.size() * 2))) return st;
                    B.lq_sbwd[par] = build_lpq_basis(Ab.data() + (size_t)par * NH, 2 * NH, NH, fr);
                    if ((st = upraw2(B.lq_bwd[par], fr.data(), fr.size() * 2))) return st;
                }
                std::vector<float> lf(fl.begin(), fl.end());
                void* lv = nullptr;
                if ((st = upraw2(lv, lf.data(), lf.size() * sizeof(float)))) return st;
                B.lq_lam = (float*)lv;
                // S = (op + J op J)/2 folded: S_e(i,j) = S(i,j) + S(i,n-1-j) (the centre once), S_o with -
                auto Sv = [&](int i, int j) {
                    return 0.5 * (op[i + (size_t)j * n] + op[(n - 1 - i) + (size_t)(n - 1 - j) * n]);
                };
                std::vector<double> Sb((size_t)NH * NH), sfr((size_t)NH * NH);
                for (int par = 0; par < 2; ++par) {
                    const int h = par ? ho : he;
                    const double sg = par ? -1.0 : 1.0;
                    std::fill(Sb.begin(), Sb.end(), 0.0);
                    for (int j = 0; j < h; ++j)
                        for (int i = 0; i < h; ++i) {
                            const int jm = n - 1 - j;
                            Sb[i + (size_t)j * NH] = (j == jm) ? Sv(i, j) : Sv(i, j) + sg * Sv(i, jm);
                        }
                    build_frag(Sb.data(), h, NH, NH, sfr.data());
                    if ((st = up(B.sfold[par], sfr.data(), sfr.size()))) return st;
                    std::vector<i4v> sd;
                    std::vector<int> se;
                    build_sdig(Sb.data(), NH, sd, se);
                    void* dv = nullptr;
                    if ((st = upraw2(dv, sd.data(), sd.size() * sizeof(i4v)))) return st;
                    B.sdig[par] = (i4v*)dv;
                    dv = nullptr;
                    if ((st = upraw2(dv, se.data(), se.size() * sizeof(int)))) return st;
                    B.sexp[par] = (int*)dv;
                }
                B.lq_ready = true;
            }
        }
        if (bidx == 1) {
            const int io = *iofzero - 1;
            std::vector<double> q(n), pc(n);
            for (int k = 0; k < n; ++k) {
                q[k] = Pinv[io + (size_t)k * n];
                pc[k] = P[k + (size_t)io * n];
            }
            if ((st = up(B.qrow, q.data(), q.size()))) return st;
            if ((st = up(B.pcol, pc.data(), pc.size()))) return st;
        }
        auto upraw = [&](void*& dst, const void* src, size_t bytes) -> int {
            HIPCHK(hipMalloc(&dst, bytes));
            HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
            return OK;
        };
        std::vector<unsigned char> buf(lp_frag_bytes(NP, LP_BF16X6));
        build_frag_lp(Pinv, n, n, NP, LP_F32, buf.data());
        if ((st = upraw(B.lpf_fwd, buf.data(), lp_frag_bytes(NP, LP_F32)))) return st;
        build_frag_lp(P, n, n, NP, LP_F32, buf.data());
        if ((st = upraw(B.lpf_bwd, buf.data(), lp_frag_bytes(NP, LP_F32)))) return st;
        build_frag_lp(Pinv, n, n, NP, LP_BF16X6, buf.data());
        if ((st = upraw(B.lph_fwd, buf.data(), buf.size()))) return st;
        build_frag_lp(P, n, n, NP, LP_BF16X6, buf.data());
        if ((st = upraw(B.lph_bwd, buf.data(), buf.size()))) return st;
    }
    B.m = m;
    B.n = n;
    B.NP = NP;
    B.ready = true;
    return OK;
}

int duct_helm3d_dev_(const duct_devptr_t* rhs, const duct_devptr_t* sol, const int* m, const int* nmodes,
                     const int* ib, const double* alp, const double* alambda, const char* bound,
                     size_t bound_len)
{
    if (!rhs || !sol || !m || !nmodes || !ib || !alp || !alambda || !bound)
        return fail(ERR_ARG, "null argument to duct_helm3d_dev_");
    (void)bound_len;
    int st = ensure_init();
    if (st) return st;
    return solve_fields(1, rhs, sol, bound, alambda, *m, *nmodes, *ib, *alp, BND_DIR_ZERO);
}

int duct_helm3d_batch_dev_(const int* nfields, const duct_devptr_t* rhs, const duct_devptr_t* sol,
                           const char* bounds, const double* alambda, const int* m, const int* nmodes,
                           const int* ib, const double* alp, size_t bounds_len)
{
    if (!nfields || !rhs || !sol || !bounds || !alambda || !m || !nmodes || !ib || !alp)
        return fail(ERR_ARG, "null argument to duct_helm3d_batch_dev_");
    if (bounds_len && (size_t)*nfields > bounds_len)
        return fail(ERR_ARG, "bounds has %zu characters for %d fields", bounds_len, *nfields);
    int st = ensure_init();
    if (st) return st;
    return solve_fields(*nfields, rhs, sol, bounds, alambda, *m, *nmodes, *ib, *alp, BND_DIR_ZERO);
}

// duct_helm3d_batch_dev_ with the planar intermediates and the Neumann edge sums in the static scratch
// array: no allocation of any kind.  The unfolded FP64 passes (the reference's products), the two
// boundary groups one after the other on the caller's stream, no per-launch events.  Every kernel
// writes the scratch it reads, so nothing in it needs clearing; calls that overlap on one GPU race.
int duct_helm3d_batch_noalloc_dev_(const int* nfields, const duct_devptr_t* rhs, const duct_devptr_t* sol,
                                   const char* bounds, const double* alambda, const int* m, const int* nmodes,
                                   const int* ib, const double* alp, size_t bounds_len)
{
    if (!nfields || !rhs || !sol || !bounds || !alambda || !m || !nmodes || !ib || !alp)
        return fail(ERR_ARG, "null argument to duct_helm3d_batch_noalloc_dev_");
    if (bounds_len && (size_t)*nfields > bounds_len)
        return fail(ERR_ARG, "bounds has %zu characters for %d fields", bounds_len, *nfields);
    int st = ensure_init();
    if (st) return st;
    if (L.device < 0 || L.device >= kNoAllocDevices)
        return fail(ERR_UNSUPPORTED, "device %d: the allocation-free entry serves devices 0..%d", L.device,
                    kNoAllocDevices - 1);
    if ((st = noalloc_scratch_setup(L.device))) return st;
    if (*nfields < 1 || *nfields > kMaxFields) return fail(ERR_ARG, "nfields=%d outside [1, %d]", *nfields, kMaxFields);
    if ((st = check_common(*m, *nmodes))) return st;
    // the largest boundary group sets the scratch: 2 planar buffers + the Neumann edge sums
    size_t need_plan = 0, need_bnd = 0;
    for (int bidx = 0; bidx < 2; ++bidx) {
        FieldGroup g{};
        g.nmodes = *nmodes; g.mode0 = *ib; g.m = *m; g.n = *m - 2; g.alp = *alp;
        for (int f = 0; f < *nfields; ++f) {
            const int bi = bound_index(bounds[f]);
            if (bi < 0) return fail(ERR_ARG, "bound '%c' of field %d is neither 'd' nor 'n'", bounds[f], f);
            if (bi == bidx) ++g.nf;
        }
        if (!g.nf) continue;
        const DevBasis& B = L.basis[bidx];
        if (!B.ready) return fail(ERR_NOT_SETUP, "duct_setup_ not called for bound '%c'", bidx ? 'n' : 'd');
        const size_t planes = (size_t)g.nf * g.nmodes * 2;
        need_plan = std::max(need_plan, planes * B.NP * B.NP * sizeof(double));
        if (bidx == 1) need_bnd = boundary_scratch_bytes(g);
    }
    const size_t plan_al = (need_plan + 255) & ~(size_t)255;
    if (2 * plan_al + need_bnd > kNoAllocScratchBytes)
        return fail(ERR_ARG, "the allocation-free entry holds %zu bytes of scratch; this batch (%d fields x %d "
                    "modes, m=%d) needs %zu", kNoAllocScratchBytes, *nfields, *nmodes, *m, 2 * plan_al + need_bnd);
    unsigned char* base = (unsigned char*)L.scratch_ptr[L.device];
    // the library's own buffers step aside for the call
    double* plan_keep[2] = {L.plan[0], L.plan[1]};
    const size_t plan_bytes_keep = L.plan_bytes;
    d2v* bscr_keep = L.bscr;
    const size_t bscr_bytes_keep = L.bscr_bytes;
    const int fold_keep = L.fold;
    L.plan[0] = (double*)base;
    L.plan[1] = (double*)(base + plan_al);
    L.plan_bytes = plan_al;
    L.bscr = need_bnd ? (d2v*)(base + 2 * plan_al) : nullptr;
    L.bscr_bytes = need_bnd;
    L.fold = 0;
    L.noalloc = true;
    const int overlap_keep = L.overlap;
    L.overlap = 0;
    st = solve_fields(*nfields, rhs, sol, bounds, alambda, *m, *nmodes, *ib, *alp, BND_DIR_ZERO);
    L.overlap = overlap_keep;
    L.noalloc = false;
    L.fold = fold_keep;
    L.plan[0] = plan_keep[0];
    L.plan[1] = plan_keep[1];
    L.plan_bytes = plan_bytes_keep;
    L.bscr = bscr_keep;
    L.bscr_bytes = bscr_bytes_keep;
    return st;
}

int duct_helm3d_(double* rhs, double* sol, const int* m, const int* nmodes, const int* ib, const double* alp,
                 const double* alambda, const char* bound, size_t bound_len)
{
    if (!rhs || !sol || !m || !nmodes || !ib || !alp || !alambda || !bound)
        return fail(ERR_ARG, "null argument to duct_helm3d_");
    (void)bound_len;
    int st = ensure_init();
    if (st) return st;
    if ((st = check_common(*m, *nmodes))) return st;
    const size_t bytes = (size_t)(*nmodes) * (*m) * (*m) * 2 * sizeof(double);
    if (bytes > L.hfield_bytes || !L.hfield[0]) {
        size_t c0 = L.hfield_bytes, c1 = L.hfield_bytes;
        if ((st = ensure_buf(L.hfield[0], c0, bytes))) return st;
        if ((st = ensure_buf(L.hfield[1], c1, bytes))) return st;
        L.hfield_bytes = bytes;
    }
    HIPCHK(hipMemcpyAsync(L.hfield[0], rhs, bytes, hipMemcpyHostToDevice, L.stream));
    duct_devptr_t r = (duct_devptr_t)L.hfield[0], s = (duct_devptr_t)L.hfield[1];
    if ((st = solve_fields(1, &r, &s, bound, alambda, *m, *nmodes, *ib, *alp, BND_DIR_ZERO))) return st;
    HIPCHK(hipMemcpyAsync(sol, L.hfield[1], bytes, hipMemcpyDeviceToHost, L.stream));
    if (bound[0] == 'd') HIPCHK(hipMemcpyAsync(rhs, L.hfield[0], bytes, hipMemcpyDeviceToHost, L.stream));
    HIPCHK(hipStreamSynchronize(L.stream));
    return OK;
}

// Shared by duct_int2d3d_ and duct_solvediag_: one complex plane held in
// hfield[0] (in) and solved into hfield[1] (out).
static int solve_plane(const char* bound, double xlambda, int imode)
{
    const int bidx = bound_index(bound[0]);
    if (bidx < 0) return fail(ERR_ARG, "in solvediag error in BC: bound '%c'", bound[0]);
    DevBasis& B = L.basis[bidx];
    if (!B.ready) return fail(ERR_NOT_SETUP, "duct_setup_ not called for bound '%c'", bound[0]);
    FieldGroup g{};
    g.nf = 1; g.nmodes = 1; g.mode0 = imode; g.m = B.m; g.n = B.n; g.alp = 0.0;
    g.in[0] = L.hfield[0];
    g.out[0] = L.hfield[1];
    g.alambda[0] = xlambda;
    if (bidx == 0) {
        Timed t(PK_BCCORR);
        BcCorrectArgs bc{};
        bc.g = g;
        bc.D2b = B.D2b;
        HIPCHK(launch_bc_correct(bc, L.stream));
    }
    return run_group(g, bidx, BND_DIR_KEEP);
}

static int plane_buffers(int m)
{
    const size_t bytes = (size_t)m * m * 2 * sizeof(double);
    if (bytes > L.hfield_bytes || !L.hfield[0]) {
        size_t c0 = L.hfield_bytes, c1 = L.hfield_bytes;
        int st;
        if ((st = ensure_buf(L.hfield[0], c0, bytes))) return st;
        if ((st = ensure_buf(L.hfield[1], c1, bytes))) return st;
        L.hfield_bytes = bytes;
    }
    return OK;
}

int duct_int2d3d_(double* sol, const double* rhs, const char* bound, const double* xlambda, const int* imode,
                  size_t bound_len)
{
    if (!sol || !rhs || !bound || !xlambda || !imode) return fail(ERR_ARG, "null argument to duct_int2d3d_");
    (void)bound_len;
    int st = ensure_init();
    if (st) return st;
    const int bidx = bound_index(bound[0]);
    if (bidx < 0) return fail(ERR_ARG, "mistake in bound '%c'", bound[0]);
    const int m = L.basis[bidx].m;
    if (!L.basis[bidx].ready) return fail(ERR_NOT_SETUP, "duct_setup_ not called for bound '%c'", bound[0]);
    if ((st = plane_buffers(m))) return st;
    const size_t bytes = (size_t)m * m * 2 * sizeof(double);
    HIPCHK(hipMemcpyAsync(L.hfield[0], rhs, bytes, hipMemcpyHostToDevice, L.stream));
    if ((st = solve_plane(bound, *xlambda, *imode))) return st;
    HIPCHK(hipMemcpyAsync(sol, L.hfield[1], bytes, hipMemcpyDeviceToHost, L.stream));
    HIPCHK(hipStreamSynchronize(L.stream));
    return OK;
}

int duct_solvediag_(double* U, const char* bound, const double* xlambda, size_t bound_len)
{
    if (!U || !bound || !xlambda) return fail(ERR_ARG, "null argument to duct_solvediag_");
    (void)bound_len;
    int st = ensure_init();
    if (st) return st;
    const int bidx = bound_index(bound[0]);
    if (bidx < 0) return fail(ERR_ARG, "in solvediag error in BC: bound '%c'", bound[0]);
    if (!L.basis[bidx].ready) return fail(ERR_NOT_SETUP, "duct_setup_ not called for bound '%c'", bound[0]);
    const int m = L.basis[bidx].m;
    if ((st = plane_buffers(m))) return st;
    std::vector<double> c((size_t)m * m * 2, 0.0);
    for (size_t e = 0; e < (size_t)m * m; ++e) c[2 * e] = U[e];
    const size_t bytes = c.size() * sizeof(double);
    HIPCHK(hipMemcpyAsync(L.hfield[0], c.data(), bytes, hipMemcpyHostToDevice, L.stream));
    // a nonzero Fourier index keeps the (all-zero) imaginary plane ordinary
    if ((st = solve_plane(bound, *xlambda, 1))) return st;
    HIPCHK(hipMemcpyAsync(c.data(), L.hfield[1], bytes, hipMemcpyDeviceToHost, L.stream));
    HIPCHK(hipStreamSynchronize(L.stream));
    for (size_t e = 0; e < (size_t)m * m; ++e) U[e] = c[2 * e];
    return OK;
}

int duct_set_fold_(const int* mode)
{
    if (!mode) return fail(ERR_ARG, "null argument");
    if (*mode < 0 || *mode > 2) return fail(ERR_ARG, "fold mode %d is not 0, 1 or 2", *mode);
    int st = ensure_init();
    if (st) return st;
    L.fold = *mode;
    return OK;
}

int duct_fold_size_(const char* bound, int* NH, size_t bound_len)
{
    (void)bound_len;
    if (!bound || !NH) return fail(ERR_ARG, "null argument");
    const int b = bound_index(*bound);
    if (b < 0) return fail(ERR_ARG, "bound '%c' is neither 'd' nor 'n'", *bound);
    *NH = L.basis[b].ready ? L.basis[b].NH : 0;
    return OK;
}

int duct_prof_enable_(const int* on)
{
    if (!on) return fail(ERR_ARG, "null argument");
    int st = ensure_init();
    if (st) return st;
    HIPCHK(hipStreamSynchronize(L.stream));
    for (auto& r : L.recs) {
        L.pool.push_back(r.a);
        L.pool.push_back(r.b);
    }
    L.recs.clear();
    L.prof = *on != 0;
    return OK;
}

int duct_prof_read_(const int* nclass, double* ms, int* count)
{
    if (!nclass || !ms || !count) return fail(ERR_ARG, "null argument");
    int st = ensure_init();
    if (st) return st;
    HIPCHK(hipStreamSynchronize(L.stream));
    for (int k = 0; k < *nclass; ++k) { ms[k] = 0.0; count[k] = 0; }
    for (auto& r : L.recs) {
        if (r.kind >= *nclass) continue;
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
        ms[r.kind] += t;
        count[r.kind] += 1;
    }
    return OK;
}

int duct_transpose_dev_(const int* Nx, const int* Ny, const int* nbatch, const double* alpha, const duct_devptr_t* A,
                        const duct_devptr_t* C)
{
    if (!Nx || !Ny || !nbatch || !alpha || !A || !C) return fail(ERR_ARG, "null argument to duct_transpose_dev_");
    if (*Nx < 0 || *Ny < 0 || *nbatch < 0) return fail(ERR_ARG, "negative size");
    if (!*A || !*C) return fail(ERR_ARG, "null device pointer");
    if (*A == *C) return fail(ERR_ARG, "in-place transpose is not supported (A == C)");
    int st = ensure_init();
    if (st) return st;
    Timed t(PK_TRANSPOSE);
    HIPCHK(launch_transpose_f64((const double*)*A, (double*)*C, *Nx, *Ny, *nbatch, *alpha, L.stream));
    return OK;
}

static int pencil(int nrow, int ncol, duct_devptr_t src, duct_devptr_t dst)
{
    if (nrow < 0 || ncol < 0) return fail(ERR_ARG, "negative size");
    if (!src || !dst) return fail(ERR_ARG, "null device pointer");
    if (src == dst) return fail(ERR_ARG, "in-place pencil transpose is not supported");
    int st = ensure_init();
    if (st) return st;
    Timed t(PK_TRANSPOSE);
    HIPCHK(launch_transpose_c128((const d2v*)src, (d2v*)dst, nrow, ncol, 1, 1.0, L.stream));
    return OK;
}

int duct_cxyz2yzx_dev_(const int* nx, const int* my, const int* nz, const duct_devptr_t* phys, const duct_devptr_t* phys2)
{
    if (!nx || !my || !nz || !phys || !phys2) return fail(ERR_ARG, "null argument to duct_cxyz2yzx_dev_");
    // phys(0:mx1, my, kb:ke) -> phys2(my, kb:ke, 0:mx1): the (nx) x (my*nz) complex transpose
    return pencil(*nx, (*my) * (*nz), *phys, *phys2);
}

int duct_cyzx2xyz_dev_(const int* nx, const int* my, const int* nz, const duct_devptr_t* phys2, const duct_devptr_t* phys)
{
    if (!nx || !my || !nz || !phys || !phys2) return fail(ERR_ARG, "null argument to duct_cyzx2xyz_dev_");
    return pencil((*my) * (*nz), *nx, *phys2, *phys);
}

int duct_mat_invdiag_dev_(const int* Nx, const int* Ny, const double* xlam, const duct_devptr_t* F,
                          const duct_devptr_t* ED2, const duct_devptr_t* C)
{
    if (!Nx || !Ny || !xlam || !F || !ED2 || !C) return fail(ERR_ARG, "null argument to duct_mat_invdiag_dev_");
    if (*Nx < 0 || *Ny < 0) return fail(ERR_ARG, "negative size");
    if (!*F || !*ED2 || !*C) return fail(ERR_ARG, "null device pointer");
    int st = ensure_init();
    if (st) return st;
    Timed t(PK_INVDIAG);
    HIPCHK(launch_invdiag((const double*)*F, (const double*)*ED2, (double*)*C, (long)(*Nx) * (*Ny), *xlam, L.stream));
    return OK;
}

int duct_helm3d_mixed_batch_dev_(const int* nfields, const duct_devptr_t* rhs, const duct_devptr_t* sol,
                                 const char* bounds, const double* alambda, const int* m, const int* nmodes,
                                 const int* ib, const double* alp, const int* scheme0, const int* scheme1,
                                 const int* niter, double* stats, size_t bounds_len)
{
    if (!nfields || !rhs || !sol || !bounds || !alambda || !m || !nmodes || !ib || !alp || !scheme0 || !scheme1 ||
        !niter)
        return fail(ERR_ARG, "null argument to duct_helm3d_mixed_batch_dev_");
    if (bounds_len && (size_t)*nfields > bounds_len)
        return fail(ERR_ARG, "bounds has %zu characters for %d fields", bounds_len, *nfields);
    int st = ensure_init();
    if (st) return st;
    return solve_fields_mixed(*nfields, rhs, sol, bounds, alambda, *m, *nmodes, *ib, *alp, *scheme0, *scheme1,
                              *niter, stats);
}

int duct_helm3d_adjoint_batch_dev_(const int* nfields, const duct_devptr_t* solbar, const duct_devptr_t* rhsbar,
                                   const char* bounds, const double* alambda, const int* m, const int* nmodes,
                                   const int* ib, const double* alp, size_t bounds_len)
{
    if (!nfields || !solbar || !rhsbar || !bounds || !alambda || !m || !nmodes || !ib || !alp)
        return fail(ERR_ARG, "null argument to duct_helm3d_adjoint_batch_dev_");
    if (bounds_len && (size_t)*nfields > bounds_len)
        return fail(ERR_ARG, "bounds has %zu characters for %d fields", bounds_len, *nfields);
    const int nf = *nfields;
    if (nf < 1 || nf > kMaxFields) return fail(ERR_ARG, "nfields=%d outside [1, %d]", nf, kMaxFields);
    int st = ensure_init();
    if (st) return st;
    if ((st = check_common(*m, *nmodes))) return st;
    for (int bidx = 0; bidx < 2; ++bidx) {
        FieldGroup g{};
        g.nmodes = *nmodes; g.mode0 = *ib; g.m = *m; g.n = *m - 2; g.alp = *alp;
        for (int f = 0; f < nf; ++f) {
            const int bi = bound_index(bounds[f]);
            if (bi < 0) return fail(ERR_ARG, "bound '%c' of field %d is neither 'd' nor 'n'", bounds[f], f);
            if (bi != bidx) continue;
            if (!solbar[f] || !rhsbar[f]) return fail(ERR_ARG, "null device pointer for field %d", f);
            if (solbar[f] == rhsbar[f]) return fail(ERR_ARG, "adjoint input and output of field %d alias", f);
            g.in[g.nf] = (const d2v*)solbar[f];
            g.out[g.nf] = (d2v*)rhsbar[f];
            g.alambda[g.nf] = alambda[f];
            ++g.nf;
        }
        if (g.nf == 0) continue;
        if ((st = run_group_adj(g, bidx))) return st;
    }
    return OK;
}

int duct_fourx_dev_(const duct_devptr_t* fou, const duct_devptr_t* phys, const int* my, const int* nz,
                    const int* nmodes, const int* mgalx, const int* iopt)
{
    if (!fou || !phys || !my || !nz || !nmodes || !mgalx || !iopt) return fail(ERR_ARG, "null argument to duct_fourx_dev_");
    if (!*fou || !*phys) return fail(ERR_ARG, "null device pointer");
    if (*my < 1 || *nz < 1) return fail(ERR_ARG, "my=%d, nz=%d", *my, *nz);
    const int M = *nmodes, X = *mgalx;
    if (M < 1 || X < 2 || X % 2 || 2 * M > X)
        return fail(ERR_ARG, "fourx needs nmodes >= 1 and an even mgalx >= 2 nmodes (got %d, %d)", M, X);
    int st = ensure_init();
    if (st) return st;
    const long lines = (long)*my * *nz;
    if (!dense_x(M, X)) {   // batched rocFFT along x: fou(0:M-1, lines) <-> phys(X+2, lines)
        const int H = X / 2 + 1;
        if ((st = ensure_buf(L.xw, L.xw_bytes, (size_t)H * lines * sizeof(d2v)))) return st;
        Timed t(PK_XPROD);
        if (*iopt >= 0) {
            XpadArgs pa{};
            pa.in[0] = (const d2v*)*fou;
            pa.out = L.xw;
            pa.lines = lines; pa.nf = 1; pa.M = M; pa.H = H; pa.X = X;
            HIPCHK(launch_xpad(pa, true, L.stream));
            if ((st = xfft_run(false, L.xw, (void*)*phys, X, lines, 1, H, 1, X + 2))) return st;
            HIPCHK(launch_xzero_tail((double*)*phys, X, lines, L.stream));
        } else {
            if ((st = xfft_run(true, (void*)*phys, L.xw, X, lines, 1, X + 2, 1, H))) return st;
            XpackArgs ka{};
            ka.in = L.xw;
            ka.out[0] = (d2v*)*fou;
            ka.lines = lines; ka.nf = 1; ka.M = M; ka.H = H; ka.X = X; ka.line_major = 1;
            HIPCHK(launch_xpack(ka, L.stream));
        }
        return OK;
    }
    if ((st = ensure_xdft(M, X))) return st;
    FourxArgs a{};
    a.fou = (d2v*)*fou;
    a.phys = (double*)*phys;
    a.Einv = L.xdft[0];
    a.Fwd = L.xdft[1];
    a.lines = lines;
    a.M = M; a.X = X; a.inverse = *iopt >= 0;
    Timed t(PK_XPROD);
    HIPCHK(launch_fourx(a, L.stream));
    return OK;
}

int duct_advection_dev_(const duct_devptr_t* u, const duct_devptr_t* v, const duct_devptr_t* w,
                        const duct_devptr_t* adv1, const duct_devptr_t* adv2, const duct_devptr_t* adv3,
                        const int* m, const int* nmodes, const int* mgalx, const double* alp, const double* aspect)
{
    if (!u || !v || !w || !adv1 || !adv2 || !adv3 || !m || !nmodes || !mgalx || !alp || !aspect)
        return fail(ERR_ARG, "null argument to duct_advection_dev_");
    const duct_devptr_t in[3] = {*u, *v, *w}, out[3] = {*adv1, *adv2, *adv3};
    for (int i = 0; i < 3; ++i) {
        if (!in[i] || !out[i]) return fail(ERR_ARG, "null device pointer");
        for (int j = 0; j < 3; ++j)
            if (out[i] == in[j]) return fail(ERR_ARG, "adv%d aliases an input velocity", i + 1);
    }
    int st = ensure_init();
    if (st) return st;
    if ((st = check_common(*m, *nmodes))) return st;
    const int M = *nmodes, X = *mgalx, mm = *m;
    if (X % 2 || X < 3 * M)
        return fail(ERR_ARG, "advection needs an even mgalx >= 3 nmodes (the 3/2 rule; got %d, %d)", M, X);
    if (!(*aspect > 0.0)) return fail(ERR_ARG, "aspect must be positive");
    if (!L.dcfrag[0] || !L.dfold[0] || L.dc_m != mm)
        return fail(ERR_NOT_SETUP, "no collocation matrix for m=%d (duct_setup_ builds it)", mm);
    const d2v* vin[3] = {(const d2v*)in[0], (const d2v*)in[1], (const d2v*)in[2]};
    d2v* vout[3] = {(d2v*)out[0], (d2v*)out[1], (d2v*)out[2]};
    return advection_run(vin, vout, mm, M, X, *alp, *aspect);
}

int duct_transyz_dev_(const duct_devptr_t* in, const duct_devptr_t* out, const int* m, const int* nmodes,
                      const char* cdir, const int* iopt, size_t cdir_len)
{
    (void)cdir_len;
    if (!in || !out || !m || !nmodes || !cdir || !iopt) return fail(ERR_ARG, "null argument to duct_transyz_dev_");
    if (*cdir != 'y' && *cdir != 'z') return fail(ERR_ARG, "err in transyz: cdir '%c'", *cdir);
    if (!*in || !*out) return fail(ERR_ARG, "null device pointer");
    if (*in == *out) return fail(ERR_ARG, "in and out must not alias");
    int st = ensure_init();
    if (st) return st;
    if ((st = check_common(*m, *nmodes))) return st;
    if (!L.tfrag[0] || L.dc_m != *m)
        return fail(ERR_NOT_SETUP, "no Chebyshev transforms for m=%d (duct_setup_ builds them)", *m);
    CollocArgs a{};
    a.g.nf = 1; a.g.nmodes = *nmodes; a.g.m = *m; a.g.n = *m - 2;
    a.g.in[0] = (const d2v*)*in;
    a.g.out[0] = (d2v*)*out;
    a.Dfrag = L.tfrag[*iopt < 0 ? 0 : 1];
    a.NP = L.dc_np;
    a.scal = 1.0;
    a.beta = 0.0;
    Timed t(PK_DERIV);
    HIPCHK(launch_colloc(*cdir == 'y' ? CL_DY : CL_DZ, a, L.stream));
    return OK;
}

int duct_derivyz_dev_(const duct_devptr_t* u, const duct_devptr_t* du, const int* m, const int* nmodes,
                      const char* cdir, const int* ider, const double* scal, const char* what, const double* aspect,
                      size_t cdir_len, size_t what_len)
{
    (void)cdir_len;
    (void)what_len;
    if (!u || !du || !m || !nmodes || !cdir || !ider || !scal || !what || !aspect)
        return fail(ERR_ARG, "null argument to duct_derivyz_dev_");
    if (*cdir != 'y' && *cdir != 'z') return fail(ERR_ARG, "cdir '%c' is neither 'y' nor 'z'", *cdir);
    if (*ider != 1 && *ider != 2) return fail(ERR_ARG, "this order of deriv not implemented: %d", *ider);
    if (!*u || !*du) return fail(ERR_ARG, "null device pointer");
    if (*u == *du) return fail(ERR_ARG, "u and du can not have the same address");
    if (!(*aspect > 0.0)) return fail(ERR_ARG, "aspect must be positive");
    int st = ensure_init();
    if (st) return st;
    if ((st = check_common(*m, *nmodes))) return st;
    DerivyzArgs a{};
    a.u = (const d2v*)*u;
    a.du = (d2v*)*du;
    a.m = *m; a.nmodes = *nmodes; a.dirz = *cdir == 'z'; a.ider = *ider;
    a.accumulate = *what != 'n';
    a.scal = *scal;
    if (a.dirz) a.scal = *ider == 1 ? *scal / *aspect : *scal / (*aspect * *aspect);
    if (a.accumulate) {
        const size_t bytes = (size_t)*nmodes * *m * *m * sizeof(d2v);
        if ((st = ensure_buf(L.rfield, L.rfield_bytes, bytes))) return st;
        a.aux = L.rfield;
    } else {
        a.aux = a.du;
    }
    Timed t(PK_DERIV);
    HIPCHK(launch_derivyz(a, L.stream));
    return OK;
}

int duct_colloc_deriv_dev_(const duct_devptr_t* in, const duct_devptr_t* out, const int* m, const int* nmodes,
                           const char* cdir, const int* trans, const double* scal, const double* beta,
                           size_t cdir_len)
{
    if (!in || !out || !m || !nmodes || !cdir || !trans || !scal || !beta)
        return fail(ERR_ARG, "null argument to duct_colloc_deriv_dev_");
    (void)cdir_len;
    if (*cdir != 'y' && *cdir != 'z') return fail(ERR_ARG, "cdir '%c' is neither 'y' nor 'z'", *cdir);
    if (!*in || !*out) return fail(ERR_ARG, "null device pointer");
    if (*in == *out) return fail(ERR_ARG, "in and out must not alias");
    int st = ensure_init();
    if (st) return st;
    if ((st = check_common(*m, *nmodes))) return st;
    if (!L.dcfrag[0] || L.dc_m != *m)
        return fail(ERR_NOT_SETUP, "no collocation matrix for m=%d (duct_setup_ builds it)", *m);
    CollocArgs a{};
    a.g.nf = 1; a.g.nmodes = *nmodes; a.g.m = *m; a.g.n = *m - 2;
    a.g.in[0] = (const d2v*)*in;
    a.g.out[0] = (d2v*)*out;
    a.Dfrag = L.dcfrag[*trans ? 1 : 0];
    a.NP = L.dc_np;
    a.scal = *scal;
    a.beta = *beta;
    Timed t(PK_DERIV);
    HIPCHK(launch_colloc(*cdir == 'y' ? CL_DY : CL_DZ, a, L.stream));
    return OK;
}

int duct_derivx_dev_(const duct_devptr_t* in, const duct_devptr_t* out, const int* m, const int* nmodes,
                     const int* ib, const double* alp, const double* scal, const double* beta)
{
    if (!in || !out || !m || !nmodes || !ib || !alp || !scal || !beta)
        return fail(ERR_ARG, "null argument to duct_derivx_dev_");
    if (!*in || !*out) return fail(ERR_ARG, "null device pointer");
    int st = ensure_init();
    if (st) return st;
    if ((st = check_common(*m, *nmodes))) return st;
    Timed t(PK_DERIV);
    HIPCHK(launch_derivx((const d2v*)*in, (d2v*)*out, *m, *nmodes, *ib, *alp, *scal, *beta, L.stream));
    return OK;
}

// One Runge-Kutta substep of the velocity and pressure, main1.f90:384-655 (items i-xi),
// with the low-storage coefficients of mod.f90:51-58.
int duct_rk3_substep_dev_(const duct_devptr_t* u, const duct_devptr_t* v, const duct_devptr_t* w,
                          const duct_devptr_t* p, const duct_devptr_t* adv1, const duct_devptr_t* adv2,
                          const duct_devptr_t* adv3, const int* m, const int* nmodes, const int* mgalx,
                          const double* alp, const double* aspect, const double* fnu, const double* dt, const int* krk,
                          const double* dpdx)
{
    if (!u || !v || !w || !p || !adv1 || !adv2 || !adv3 || !m || !nmodes || !mgalx || !alp || !aspect || !fnu || !dt ||
        !krk || !dpdx)
        return fail(ERR_ARG, "null argument to duct_rk3_substep_dev_");
    const duct_devptr_t f[7] = {*u, *v, *w, *p, *adv1, *adv2, *adv3};
    for (int i = 0; i < 7; ++i) {
        if (!f[i]) return fail(ERR_ARG, "null device pointer (argument %d)", i + 1);
        for (int j = 0; j < i; ++j)
            if (f[i] == f[j]) return fail(ERR_ARG, "arguments %d and %d alias", j + 1, i + 1);
    }
    if (*krk < 1 || *krk > 3) return fail(ERR_ARG, "krk=%d outside [1, 3]", *krk);
    if (!(*fnu > 0.0) || !(*dt > 0.0) || !(*aspect > 0.0)) return fail(ERR_ARG, "fnu, dt and aspect must be positive");
    int st = ensure_init();
    if (st) return st;
    const int mm = *m, M = *nmodes, X = *mgalx;
    if ((st = check_common(mm, M))) return st;
    if (X % 2 || X < 3 * M)
        return fail(ERR_ARG, "the substep's advection needs an even mgalx >= 3 nmodes (got %d, %d)", M, X);
    if (!L.dfold[0] || !L.dfold[1] || L.dc_m != mm)
        return fail(ERR_NOT_SETUP, "no collocation matrices for m=%d (duct_setup_ builds them)", mm);
    for (int b = 0; b < 2; ++b)
        if (!L.basis[b].ready || L.basis[b].m != mm)
            return fail(ERR_NOT_SETUP, "bound '%c' not set up for m=%d", b ? 'n' : 'd', mm);
    // Spalart-Moser-Rogers coefficients (mod.f90:51-58)
    static const double akkc[3] = {29.0 / 96.0, -3.0 / 40.0, 1.0 / 6.0};
    static const double bkkc[3] = {37.0 / 160.0, 5.0 / 24.0, 1.0 / 6.0};
    static const double gkkc[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
    static const double rkkc[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
    const int k = *krk - 1;
    const double a = akkc[k], b = bkkc[k], nu = *fnu, ab = a + b, asp = *aspect;
    const size_t field = (size_t)M * mm * mm;
    if ((st = ensure_buf(L.rk, L.rk_bytes, 9 * field * sizeof(d2v)))) return st;
    d2v* h[3] = {L.rk, L.rk + field, L.rk + 2 * field};
    d2v* advnew[3] = {L.rk + 3 * field, L.rk + 4 * field, L.rk + 5 * field};
    d2v* phi = L.rk + 6 * field;
    d2v* ty = L.rk + 7 * field;   // the y and z terms of a sum, computed in one launch before the x term
    d2v* tz = L.rk + 8 * field;
    d2v* U[3] = {(d2v*)*u, (d2v*)*v, (d2v*)*w};
    d2v* P = (d2v*)*p;
    d2v* ADV[3] = {(d2v*)*adv1, (d2v*)*adv2, (d2v*)*adv3};
    const double* D1 = L.dfold[0];   // D and D2 folded: J D J = -D, J D2 J = D2
    const double* D2 = L.dfold[1];
    // (ii) adv(u^{k-1}) into advnew
    const d2v* vin[3] = {U[0], U[1], U[2]};
    if ((st = advection_run(vin, advnew, mm, M, X, *alp, asp))) return st;
    // (i), (iii)-(vii): h = -1/(nu b) [-rho adv_old - gam adv + nu a lap(u) + u/dt - (a+b) grad p - (a+b) dpdx]
    const double s = -1.0 / (nu * b);
    {
        Rk3RhsArgs r{};
        for (int c = 0; c < 3; ++c) {
            r.u[c] = U[c];
            r.adv[c] = ADV[c];
            r.advnew[c] = advnew[c];
            r.h[c] = h[c];
        }
        r.p = P;
        r.mm = (long)mm * mm;
        r.nmodes = M;
        r.mode0 = 0;
        r.use_old = k > 0;
        r.alp = *alp;
        r.s = s;
        r.rho = rkkc[k];
        r.gam = gkkc[k];
        r.clap = nu * a;
        r.rdt = 1.0 / *dt;
        r.cp = -ab;
        r.cdp = -ab * *dpdx;
        Timed t(PK_DERIV);
        HIPCHK(launch_rk3_rhs(r, L.stream));
    }
    // laplacian3d's y parts, then its z parts, scaled as in (vii); then grad3d's y and z parts of p
    const double sy2 = s * nu * a, sz2 = s * nu * a / (asp * asp);
    if ((st = cline_run(mm, M, {{U[0], h[0], D2, sy2, 1.0, 0, 1}, {U[1], h[1], D2, sy2, 1.0, 0, 1},
                                {U[2], h[2], D2, sy2, 1.0, 0, 1}})))
        return st;
    if ((st = cline_run(mm, M, {{U[0], h[0], D2, sz2, 1.0, 1, 1}, {U[1], h[1], D2, sz2, 1.0, 1, 1},
                                {U[2], h[2], D2, sz2, 1.0, 1, 1}})))
        return st;
    if ((st = cline_run(mm, M, {{P, h[1], D1, -s * ab, 1.0, 0, -1}, {P, h[2], D1, -s * ab / asp, 1.0, 1, -1}})))
        return st;
    // (viii) u*, v*, w*: Dirichlet Helmholtz, alambda = 1/(b nu dt)
    {
        const double helco = 1.0 / (b * nu * *dt);
        const duct_devptr_t rhs[3] = {(duct_devptr_t)h[0], (duct_devptr_t)h[1], (duct_devptr_t)h[2]};
        const duct_devptr_t sol[3] = {*u, *v, *w};
        const double al[3] = {helco, helco, helco};
        if ((st = solve_fields(3, rhs, sol, "ddd", al, mm, M, 0, *alp, BND_DIR_ZERO))) return st;
    }
    // (ix) div3d(u*, v*, w*) / ((a+b) dt) -> h1
    const double sd = 1.0 / (ab * *dt);
    // (x + y) + z: the y and z products into ty, tz in one launch, then the x term adds them in
    // that order (the same rounding as accumulating them into h1 one after the other)
    if ((st = cline_run(mm, M, {{U[1], ty, D1, sd, 0.0, 0, -1}, {U[2], tz, D1, sd / asp, 0.0, 1, -1}}))) return st;
    {
        Timed t(PK_DERIV);
        HIPCHK(launch_xderiv(U[0], h[0], (long)mm * mm, M, 0, 1, *alp, sd, 0.0, ty, tz, L.stream));
    }
    // (x) phi: Neumann Poisson
    {
        const duct_devptr_t rhs[1] = {(duct_devptr_t)h[0]};
        const duct_devptr_t sol[1] = {(duct_devptr_t)phi};
        const double al[1] = {0.0};
        if ((st = solve_fields(1, rhs, sol, "n", al, mm, M, 0, *alp, BND_DIR_ZERO))) return st;
    }
    // (xi) update_uvwp: rhsphi = (a+b) dt nu lap(phi) -> h1; p += phi - rhsphi; u -= dt (a+b) grad phi.
    // The y and z parts of lap(phi) and of grad(phi) in one launch: all four read phi only, and the
    // x term and the p, u update that follow touch neither v nor w.
    const double sl = ab * *dt * nu, su = -*dt * ab;
    if ((st = cline_run(mm, M, {{phi, ty, D2, sl, 0.0, 0, 1}, {phi, tz, D2, sl / (asp * asp), 0.0, 1, 1},
                                {phi, U[1], D1, su, 1.0, 0, -1}, {phi, U[2], D1, su / asp, 1.0, 1, -1}})))
        return st;
    {
        Timed t(PK_DERIV);
        HIPCHK(launch_xderiv(phi, h[0], (long)mm * mm, M, 0, 2, *alp, sl, 0.0, ty, tz, L.stream));
    }
    {
        Timed t(PK_DERIV);
        HIPCHK(launch_rk3_update(P, U[0], phi, h[0], (long)mm * mm, M, 0, *alp, su, L.stream));
    }
    return OK;
}

// The GPU-upload tails of pseudomaty_temp / pseudomatz_temp (pseudomat_temp.f90:2-321): the
// caller's y basis (interior ny x ny, ny = my - 2) and z basis (nz x nz), column-major, their
// eigenvalues and the scaled derivative matrices (my x my, mz x mz); bounds = the four wall
// types y=-1, y=+1, z=-1, z=+1.  my and mz may differ (helm3d_temp.f90:35-36).
int duct_setup_temp_(const char* bounds, const int* my_, const int* mz_, const double* PY, const double* PYinv,
                     const double* ETY, const double* DY, const double* PZ, const double* PZinv, const double* ETZ,
                     const double* DZ, size_t bounds_len)
{
    if (!bounds || !my_ || !mz_ || !PY || !PYinv || !ETY || !DY || !PZ || !PZinv || !ETZ || !DZ)
        return fail(ERR_ARG, "null argument to duct_setup_temp_");
    if (bounds_len < 4) return fail(ERR_ARG, "bounds needs 4 characters (y=-1, y=+1, z=-1, z=+1)");
    int b[4];
    for (int w = 0; w < 4; ++w) {
        if (bounds[w] != 'd' && bounds[w] != 'n') return fail(ERR_ARG, "wall %d: bound '%c'", w + 1, bounds[w]);
        b[w] = bounds[w] == 'n';
    }
    const int my = *my_, mz = *mz_;
    for (int mm : {my, mz})
        if (mm < kMinM || mm > kMaxM) return fail(ERR_UNSUPPORTED, "m=%d outside [%d, %d]", mm, kMinM, kMaxM);
    const int ny = my - 2, nz = mz - 2, NP = padded_n(std::max(ny, nz));
    int st = ensure_init();
    if (st) return st;
    TempDev& T = L.temp;
    HIPCHK(hipStreamSynchronize(L.stream));
    free_temp();
    auto up = [&](double*& dst, const double* src, size_t count) -> int {
        HIPCHK(hipMalloc((void**)&dst, count * sizeof(double)));
        HIPCHK(hipMemcpy(dst, src, count * sizeof(double), hipMemcpyHostToDevice));
        return OK;
    };
    // both bases padded with zeros to the common NP: the passes of one direction then read zero
    // K columns beyond that direction's size, and the padded plane rows and columns stay zero
    std::vector<double> frag((size_t)NP * NP), lp(NP, 0.0), vec((size_t)4 * (my + mz), 0.0);
    build_frag(PYinv, ny, ny, NP, frag.data());
    if ((st = up(T.Ayi, frag.data(), frag.size()))) return st;
    build_frag(PZinv, nz, nz, NP, frag.data());
    if ((st = up(T.Azi, frag.data(), frag.size()))) return st;
    build_frag(PY, ny, ny, NP, frag.data());
    if ((st = up(T.Ay, frag.data(), frag.size()))) return st;
    build_frag(PZ, nz, nz, NP, frag.data());
    if ((st = up(T.Az, frag.data(), frag.size()))) return st;
    for (int i = 0; i < ny; ++i) lp[i] = ETY[i];
    if ((st = up(T.ly, lp.data(), lp.size()))) return st;
    std::fill(lp.begin(), lp.end(), 0.0);
    for (int i = 0; i < nz; ++i) lp[i] = ETZ[i];
    if ((st = up(T.lz, lp.data(), lp.size()))) return st;
    for (int d = 0; d < 2; ++d) {        // D2 = D D columns 1 and m (pseudomat_temp.f90:98-104), rows 1 and m of D
        const double* D = d ? DZ : DY;
        const int m = d ? mz : my;
        double* v = vec.data() + (d ? (size_t)4 * my : 0);
        for (int i = 0; i < m; ++i) {
            double s0 = 0.0, s1 = 0.0;
            for (int k = 0; k < m; ++k) {
                s0 += D[i + (size_t)k * m] * D[k];
                s1 += D[i + (size_t)k * m] * D[k + (size_t)(m - 1) * m];
            }
            v[i] = s0;
            v[(size_t)m + i] = s1;
            v[(size_t)2 * m + i] = D[(size_t)i * m];                   // D(1, i)
            v[(size_t)3 * m + i] = D[(m - 1) + (size_t)i * m];         // D(m, i)
        }
    }
    if ((st = up(T.vec, vec.data(), vec.size()))) return st;
    for (int w = 0; w < 4; ++w) T.b[w] = b[w];
    T.my = my;
    T.mz = mz;
    T.NP = NP;
    T.ready = true;
    return OK;
}

// helm3d_temp (helm3d_temp.f90:2-44, int2d3d_temp.f90, solvediag_temp.f90) on a device field:
// sol = the mixed-boundary solution of (D_yy + D_zz - xlambda) u = rhs, xlambda = alambda + (alp k)^2,
// fields (my, mz, nmodes), the wall data bcy [nmodes][2][mz] (over z) and bcz [nmodes][2][my]
// (over y); rhs is not modified.
int duct_helm3d_temp_dev_(const duct_devptr_t* rhs, const duct_devptr_t* sol, const duct_devptr_t* bcy,
                          const duct_devptr_t* bcz, const int* my, const int* mz, const int* nmodes, const int* ib,
                          const double* alp, const double* alambda)
{
    if (!rhs || !sol || !bcy || !bcz || !my || !mz || !nmodes || !ib || !alp || !alambda)
        return fail(ERR_ARG, "null argument to duct_helm3d_temp_dev_");
    if (!*rhs || !*sol || !*bcy || !*bcz) return fail(ERR_ARG, "null device pointer");
    if (*rhs == *sol) return fail(ERR_ARG, "rhs and sol must not alias (helm3d_temp keeps the rhs)");
    int st = ensure_init();
    if (st) return st;
    if ((st = check_common(*my, *nmodes)) || (st = check_common(*mz, *nmodes))) return st;
    if (*ib < 0) return fail(ERR_ARG, "ib=%d < 0", *ib);
    TempDev& T = L.temp;
    if (!T.ready || T.my != *my || T.mz != *mz)
        return fail(ERR_NOT_SETUP, "duct_setup_temp_ not called for my=%d, mz=%d", *my, *mz);
    const int NP = T.NP, ny = *my - 2, nz = *mz - 2, M = *nmodes;
    const size_t field = (size_t)M * *my * *mz;
    if ((st = ensure_buf(L.tfield, L.tfield_bytes, field * sizeof(d2v)))) return st;
    const long planes = (long)M * 2;
    if ((st = ensure_planar((size_t)planes * NP * NP * sizeof(double)))) return st;
    TempArgs ta{};
    ta.rhs = (const d2v*)*rhs;
    ta.out = L.tfield;
    ta.bcy = (const double*)*bcy;
    ta.bcz = (const double*)*bcz;
    ta.vec = T.vec;
    ta.my = *my; ta.mz = *mz; ta.nmodes = M; ta.mode0 = *ib;
    ta.by1 = T.b[0]; ta.by2 = T.b[1]; ta.bz1 = T.b[2]; ta.bz2 = T.b[3];
    {
        Timed t(PK_BCCORR);
        HIPCHK(launch_temp_rhs(ta, L.stream));
    }
    FieldGroup g{};
    g.nf = 1; g.nmodes = M; g.mode0 = *ib; g.m = *my; g.n = ny; g.alp = *alp;
    g.in[0] = L.tfield;
    g.out[0] = (d2v*)*sol;
    g.alambda[0] = *alambda;
    ProjArgs a{};
    a.g = g;
    a.lam = T.ly;
    a.lam2 = T.lz;
    a.gauge = -1;
    a.ldf = *my;
    a.fplane = (long)*my * *mz;
    {   // pass 1: (PTYM1 F)^T, K = y, columns z
        Timed t(PK_C2P);
        a.Afrag = T.Ayi; a.X = nullptr; a.Y = L.plan[0]; a.ncols = (long)M * NP;
        a.kn = ny; a.cn = nz;
        HIPCHK(launch_proj(NP, PROJ_C2P, a, L.stream));
    }
    {   // pass 2: PTYM1 F PTZM1^T ./ (ETY(i) + ETZ(j) - xlambda) (solvediag_temp.f90:84-96), K = z, columns y-eigen
        Timed t(PK_P2P_DIV);
        a.Afrag = T.Azi; a.X = L.plan[0]; a.Y = L.plan[1]; a.ncols = planes * NP;
        a.kn = nz; a.cn = ny;
        HIPCHK(launch_proj(NP, PROJ_P2P_DIV, a, L.stream));
    }
    {   // pass 3: (PTY G)^T, K = y-eigen, columns z-eigen
        Timed t(PK_P2P);
        a.Afrag = T.Ay; a.X = L.plan[1]; a.Y = L.plan[0]; a.ncols = planes * NP;
        a.kn = ny; a.cn = nz;
        HIPCHK(launch_proj(NP, PROJ_P2P, a, L.stream));
    }
    {   // pass 4: PTY G PTZ^T into the interior of sol, K = z-eigen, columns y
        Timed t(PK_P2C);
        a.Afrag = T.Az; a.X = L.plan[0]; a.Y = nullptr; a.ncols = (long)M * NP;
        a.kn = nz; a.cn = ny;
        HIPCHK(launch_proj(NP, PROJ_P2C, a, L.stream));
    }
    {   // edge_value_y / edge_value_z, corners from the rhs
        Timed t(PK_BOUNDARY);
        ta.out = (d2v*)*sol;
        HIPCHK(launch_temp_edges(ta, L.stream));
    }
    return OK;
}

// preshen (preshen.f90:2-330) for the device: the caller's (host-built) Galerkin matrices of one
// boundary type.  Ay (n x m) = P^T W, Az (n x m) = W, By (m x n) = R P, Bz (m x n) = R with
// n = m - 2, column-major; md, mu, sd the mass diagonal, mass (k, k+2) entries and stiffness
// diagonal; lam the y-eigenvalues of M x = lam S x (entry 0 unused for Neumann).
int duct_setup_shen_(const char* bound, const int* m_, const double* Ay, const double* Az, const double* By,
                     const double* Bz, const double* md, const double* mu, const double* sd, const double* lam,
                     size_t bound_len)
{
    (void)bound_len;
    if (!bound || !m_ || !Ay || !Az || !By || !Bz || !md || !mu || !sd || !lam)
        return fail(ERR_ARG, "null argument to duct_setup_shen_");
    const int bidx = bound_index(bound[0]);
    if (bidx < 0) return fail(ERR_ARG, "preshen: bound should be d or n, got '%c'", bound[0]);
    const int m = *m_, n = m - 2;
    if (m < kMinM || m > 16 * 3 * 6) return fail(ERR_UNSUPPORTED, "m=%d outside [%d, 288]", m, kMinM);
    int st = ensure_init();
    if (st) return st;
    ShenDev& S = L.shen[bidx];
    HIPCHK(hipStreamSynchronize(L.stream));
    free_shen(S);
    const int MP = (m + 15) / 16 * 16;
    std::vector<double> pad((size_t)m * m), frag((size_t)MP * MP);
    auto up = [&](double*& dst, const double* src, size_t count) -> int {
        HIPCHK(hipMalloc((void**)&dst, count * sizeof(double)));
        HIPCHK(hipMemcpy(dst, src, count * sizeof(double), hipMemcpyHostToDevice));
        return OK;
    };
    const double* src[4] = {Ay, Az, By, Bz};
    for (int i = 0; i < 4; ++i) {
        std::fill(pad.begin(), pad.end(), 0.0);
        const bool wide = i < 2;              // n x m (rows padded) or m x n (columns padded)
        for (int c = 0; c < (wide ? m : n); ++c)
            for (int r = 0; r < (wide ? n : m); ++r)
                pad[r + (size_t)c * m] = src[i][r + (size_t)c * (wide ? n : m)];
        build_frag(pad.data(), m, m, MP, frag.data());
        if ((st = up(S.frag[i], frag.data(), frag.size()))) return st;
    }
    std::vector<double> v((size_t)4 * n);
    for (int k = 0; k < n; ++k) {
        v[k] = md[k];
        v[n + k] = mu[k];
        v[2 * n + k] = sd[k];
        v[3 * n + k] = lam[k];
    }
    if ((st = up(S.vec, v.data(), v.size()))) return st;
    S.m = m;
    S.ready = true;
    return OK;
}

// helm3d_shen (helm_shen.f90:2-85, solveshen.f90:5-94) on a device field: sol = the Shen-Galerkin
// solution of -lap(u) + xlambda u = rhs, xlambda = alambda + (alp (ib+k))^2, nodal values at the
// CGL nodes in and out (the reference's transyz in / out included); rhs is kept.
int duct_helm3d_shen_dev_(const duct_devptr_t* rhs, const duct_devptr_t* sol, const int* m, const int* nmodes,
                          const int* ib, const double* alp, const double* alambda, const char* bound,
                          size_t bound_len)
{
    (void)bound_len;
    if (!rhs || !sol || !m || !nmodes || !ib || !alp || !alambda || !bound)
        return fail(ERR_ARG, "null argument to duct_helm3d_shen_dev_");
    if (!*rhs || !*sol) return fail(ERR_ARG, "null device pointer");
    if (*rhs == *sol) return fail(ERR_ARG, "rhs and sol must not alias");
    const int bidx = bound_index(bound[0]);
    if (bidx < 0) return fail(ERR_ARG, "bound '%c' is neither 'd' nor 'n'", bound[0]);
    int st = ensure_init();
    if (st) return st;
    if ((st = check_common(*m, *nmodes))) return st;
    ShenDev& S = L.shen[bidx];
    if (!S.ready || S.m != *m) return fail(ERR_NOT_SETUP, "duct_setup_shen_ not called for '%c', m=%d", bound[0], *m);
    const int mm = *m, M = *nmodes, n = mm - 2;
    const size_t field = (size_t)M * mm * mm;
    if ((st = ensure_buf(L.tfield, L.tfield_bytes, field * sizeof(d2v)))) return st;
    if ((st = ensure_buf(L.sgam, L.sgam_bytes, field * sizeof(double)))) return st;
    auto prod = [&](const d2v* in, d2v* out, int which, bool dirz) -> int {
        CollocArgs c{};
        c.g.nf = 1; c.g.nmodes = M; c.g.m = mm; c.g.n = n;
        c.g.in[0] = in;
        c.g.out[0] = out;
        c.Dfrag = S.frag[which];
        c.NP = (mm + 15) / 16 * 16;
        c.scal = 1.0;
        c.beta = 0.0;
        Timed t(PK_DERIV);
        HIPCHK(launch_cline(dirz, c, L.stream));
        return OK;
    };
    d2v* U = (d2v*)*sol;
    // F = A_y f A_z^T: Galerkin inner products with phi_p(y) phi_q(z), then inv(P) inv(S) along y
    if ((st = prod((const d2v*)*rhs, L.tfield, 0, false))) return st;
    if ((st = prod(L.tfield, U, 1, true))) return st;
    {   // solV: the tridiagonal systems along z (in place)
        ShenTriArgs ta{};
        ta.V = U;
        ta.gam = L.sgam;
        ta.md = S.vec;
        ta.mu = S.vec + n;
        ta.sd = S.vec + 2 * n;
        ta.lam = S.vec + 3 * n;
        ta.alambda = *alambda;
        ta.alp = *alp;
        ta.m = mm; ta.nmodes = M; ta.mode0 = *ib; ta.neumann = bidx == 1;
        Timed t(PK_TRIDIAG);
        HIPCHK(launch_shen_tri(ta, L.stream));
    }
    // u = B_y V B_z^T: back to Shen, Legendre, Chebyshev and the nodes
    if ((st = prod(U, L.tfield, 2, false))) return st;
    if ((st = prod(L.tfield, U, 3, true))) return st;
    return OK;
}

// ---- Fourier-space field files (HDF5) ------------------------------------------------------
// Host code only: these read files into host arrays and need no device.

int duct_h5_read_header_(const char* filename, int* ints, double* reals, size_t filename_len)
{
    if (!filename || !ints || !reals) return fail(ERR_ARG, "null argument to duct_h5_read_header_");
    try {
        duct_h5::File f(duct_h5::fortran_string(filename, filename_len));
        auto scalar = [&](const char* name, bool required, double& v) {
            if (!f.has(name)) {
                if (required) throw std::runtime_error(std::string("no header dataset '") + name + "'");
                v = std::nan("");
                return;
            }
            const duct_h5::Dataset d = f.dataset(name);
            if (d.count() < 1) throw std::runtime_error(std::string("empty header dataset '") + name + "'");
            f.read_doubles(d, 0, 1, &v);
        };
        double v[5];
        scalar("mgalx", true, v[0]);
        scalar("my", true, v[2]);
        scalar("mz", true, v[3]);
        scalar("nparams", false, v[4]);
        if (f.has("mx1")) scalar("mx1", true, v[1]);
        else v[1] = 2 * ((int)v[0] / 3) / 2 - 1;        // readheader_hdf5's fallback (:824-826)
        for (int i = 0; i < 5; ++i) ints[i] = std::isnan(v[i]) ? -1 : (int)v[i];
        const char* names[7] = {"alp", "aspect", "time", "fnu", "cfl", "dtfixed", "cmassflow"};
        for (int i = 0; i < 7; ++i) scalar(names[i], i == 0 || i == 2, reals[i]);
    } catch (const std::exception& e) {
        return fail(ERR_ARG, "%s", e.what());
    }
    return OK;
}

int duct_h5_read_field_(const char* filename, const char* name, double* out, const int* my, const int* mz,
                        const int* ib, const int* nmodes, size_t filename_len, size_t name_len)
{
    if (!filename || !name || !out || !my || !mz || !ib || !nmodes)
        return fail(ERR_ARG, "null argument to duct_h5_read_field_");
    if (*my < 2 || *mz < 2 || *ib < 0 || *nmodes < 1)
        return fail(ERR_ARG, "my=%d, mz=%d, ib=%d, nmodes=%d: need my, mz >= 2, ib >= 0, nmodes >= 1", *my, *mz,
                    *ib, *nmodes);
    const std::string fname = duct_h5::fortran_string(filename, filename_len);
    const std::string dname = duct_h5::fortran_string(name, name_len);
    try {
        duct_h5::File f(fname);
        const duct_h5::Dataset d = f.dataset(dname);
        if (d.dims.size() != 4 || d.dims[3] != 2 || d.kind != duct_h5::Kind::Float)
            return fail(ERR_ARG, "%s: '%s' is not a Fourier-space field (floating point, extents (2, my, mz, "
                        "modes) in Fortran order)", fname.c_str(), dname.c_str());
        const long myr = (long)d.dims[2], mzr = (long)d.dims[1], mfile = (long)d.dims[0];
        const bool same = myr == *my && mzr == *mz;
        if (!same && (myr == *my || mzr == *mz))
            return fail(ERR_UNSUPPORTED, "%s: the file's grid is %ld x %ld and the caller's %d x %d; only one of "
                        "my, mz differs, which the reference does not interpolate", fname.c_str(), myr, mzr,
                        *my, *mz);
        if (!same && (myr < 2 || mzr < 2))
            return fail(ERR_ARG, "%s: grid %ld x %ld too small to interpolate", fname.c_str(), myr, mzr);
        const long nread = std::max(0L, std::min<long>(*nmodes, mfile - *ib));   // modes past mx1 stay 0
        const size_t plane_out = (size_t)2 * *my * *mz, plane_in = (size_t)2 * myr * mzr;
        if (same) {
            if (nread) f.read_doubles(d, (uint64_t)*ib * plane_in, (uint64_t)nread * plane_in, out);
        } else {
            std::vector<double> tmp(plane_in);
            for (long k = 0; k < nread; ++k) {
                f.read_doubles(d, (uint64_t)(*ib + k) * plane_in, plane_in, tmp.data());
                duct_h5::cheb_interp((int)myr, (int)mzr, tmp.data(), *my, *mz, out + (size_t)k * plane_out);
            }
        }
        std::fill(out + (size_t)nread * plane_out, out + (size_t)*nmodes * plane_out, 0.0);
    } catch (const std::exception& e) {
        return fail(ERR_ARG, "%s", e.what());
    }
    return OK;
}

int duct_h5_read_doubles_(const char* filename, const char* name, double* out, const long long* count,
                          size_t filename_len, size_t name_len)
{
    if (!filename || !name || !out || !count) return fail(ERR_ARG, "null argument to duct_h5_read_doubles_");
    const std::string fname = duct_h5::fortran_string(filename, filename_len);
    const std::string dname = duct_h5::fortran_string(name, name_len);
    try {
        duct_h5::File f(fname);
        const duct_h5::Dataset d = f.dataset(dname);
        if ((long long)d.count() != *count)
            return fail(ERR_ARG, "%s: '%s' holds %llu elements, not %lld", fname.c_str(), dname.c_str(),
                        (unsigned long long)d.count(), *count);
        f.read_doubles(d, 0, d.count(), out);
    } catch (const std::exception& e) {
        return fail(ERR_ARG, "%s", e.what());
    }
    return OK;
}

int duct_h5_read_chars_(const char* filename, const char* name, char* out, const int* capacity, int* length,
                        size_t filename_len, size_t name_len)
{
    if (!filename || !name || !out || !capacity || !length)
        return fail(ERR_ARG, "null argument to duct_h5_read_chars_");
    const std::string fname = duct_h5::fortran_string(filename, filename_len);
    const std::string dname = duct_h5::fortran_string(name, name_len);
    try {
        duct_h5::File f(fname);
        const std::string s = f.read_string(f.dataset(dname));
        if ((long)s.size() > *capacity)
            return fail(ERR_ARG, "%s: '%s' holds %zu characters, more than %d", fname.c_str(), dname.c_str(),
                        s.size(), *capacity);
        std::memcpy(out, s.data(), s.size());
        *length = (int)s.size();
    } catch (const std::exception& e) {
        return fail(ERR_ARG, "%s", e.what());
    }
    return OK;
}

}  // extern "C"
