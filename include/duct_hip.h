/*
 * duct_hip.h - C-ABI of libduct_hip.so, the MI355X (gfx950) implementation of
 * duct's per-Fourier-mode cross-plane Helmholtz/Poisson solve stage.
 *
 * Conventions (the style of the reference's Fortran bindings, src/fortran.c
 * and src/mytranspose_cuda.cu of ando2800/duct):
 *   - lowercase names with a trailing underscore, every argument by pointer,
 *     so gfortran/flang can call them through implicit interfaces;
 *   - a Fortran CHARACTER argument adds a trailing hidden length (size_t, by
 *     value) at the end of the list;
 *   - device addresses travel as duct_devptr_t (an integer holding the
 *     address, like devptr_t in src/fortran.h:55);
 *   - every entry point returns an int status: 0 = success, otherwise one of
 *     DUCT_ERR_*; duct_last_error_ returns the message.  (The reference's GPU
 *     entry points always return 0, mytranspose_cuda.cu:149,178,195.)
 *   - complex fields are complex(8) (m, m, nmodes), y fastest, i.e. the
 *     reference's rhs(my,mz,ib:ie) (helm3d.f90:29).
 * One host thread drives the library; calls are ordered on the stream set by
 * duct_set_stream_ (default: the null stream).
 */
#ifndef DUCT_HIP_H
#define DUCT_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t duct_devptr_t;

#define DUCT_OK 0
#define DUCT_ERR_ARG 1
#define DUCT_ERR_NOT_SETUP 2
#define DUCT_ERR_HIP 3
#define DUCT_ERR_ALLOC 4
#define DUCT_ERR_UNSUPPORTED 5

/* ---- runtime ------------------------------------------------------------ */

/* Select the device and create the library state.
 * Replaces CUBLAS_INIT (src/fortran.c:71) as called at src/main1.f90:149-150. */
int duct_init_(const int *device);

/* Free every device buffer the library owns.  Replaces CUBLAS_SHUTDOWN. */
int duct_shutdown_(void);

/* Stream for all later launches (a hipStream_t held as an integer, 0 = null
 * stream).  Replaces cuda_stream_create_/cublas_set_stream_
 * (src/mytranspose_cuda.cu:198-207).  The library's scratch buffers are shared
 * by all calls, so a new stream is made to wait (hipStreamWaitEvent) for the
 * work queued on the previous one; when either stream is being captured into
 * a graph no wait is added and the caller orders the two streams. */
int duct_set_stream_(const duct_devptr_t *stream);

/* Wait for all work the library queued on its stream. */
int duct_synchronize_(void);

/* Copy the last error message (NUL-terminated, truncated to *len bytes). */
int duct_last_error_(char *buf, const int *len);

/* ---- one-time setup ----------------------------------------------------- */

/* Upload the eigenbasis of one boundary type after the host eigensolve.
 * Replaces the GPU-upload tails of pseudomat (src/pseudomat.f90:188-221 for
 * 'n', :265-298 for 'd'); the caller keeps its own LAPACK dgeev/dgetri.
 *   bound    'd' (Dirichlet) or 'n' (Neumann)
 *   m        CGL points per direction (my == mz), 4 <= m <= 514
 *   P        n x n right eigenvectors, n = m-2 (PD or PN)
 *   Pinv     n x n inverse (PDM1 or PNM1)
 *   lam      n eigenvalues (ED or EN)
 *   DC       m x m collocation derivative matrix (pseudomat.f90:90-107)
 *   AM1      4 x 4 inverse corner matrix (Neumann only, may be NULL for 'd')
 *   iofzero  1-based eigen-index of the Neumann null mode (ignored for 'd')
 *   wanted   gauge value at (iofzero+1, iofzero+1) (ignored for 'd')       */
int duct_setup_(const char *bound, const int *m, const double *P, const double *Pinv,
                const double *lam, const double *DC, const double *AM1, const int *iofzero,
                const double *wanted, size_t bound_len);

/* ---- the solve stage ---------------------------------------------------- */

/* helm3d on device-resident fields (src/helm3d.f90:2, ipro >= 1):
 * solve (D_y^2 + D_z^2 - xlambda) sol = rhs for modes ib .. ib+nmodes-1 with
 * xlambda = alambda + (alp*imode)^2 (helm3d.f90:67, fourier.F:201-203).
 * For bound 'd' the boundary of rhs is zeroed in place (helm3d.f90:48-60) and
 * sol's boundary is zero; for 'n' sol's boundary comes from the Neumann edge
 * and corner equations.  rhs and sol may be the same field.               */
int duct_helm3d_dev_(const duct_devptr_t *rhs, const duct_devptr_t *sol, const int *m,
                     const int *nmodes, const int *ib, const double *alp, const double *alambda,
                     const char *bound, size_t bound_len);

/* The same over nfields (<= 8) fields in one pass sequence: rhs[f], sol[f]
 * device fields, bounds[f] in {'d','n'} (a CHARACTER*(nfields)), alambda[f].
 * Fields sharing a boundary type share every launch.  This is the u, v, w
 * (Dirichlet) + p (Neumann) batch of src/main1.f90:617-643.                */
int duct_helm3d_batch_dev_(const int *nfields, const duct_devptr_t *rhs, const duct_devptr_t *sol,
                           const char *bounds, const double *alambda, const int *m,
                           const int *nmodes, const int *ib, const double *alp,
                           size_t bounds_len);

/* helm3d on host arrays (drop-in for the Fortran subroutine body): uploads
 * rhs, solves on the device, downloads sol and the boundary-zeroed rhs.   */
int duct_helm3d_(double *rhs, double *sol, const int *m, const int *nmodes, const int *ib,
                 const double *alp, const double *alambda, const char *bound, size_t bound_len);

/* int2d3d (src/int2d3d.f90:2) on host arrays: one complex plane, real part
 * solved always, imaginary part skipped when imode == 0 (sol imag = 0).    */
int duct_int2d3d_(double *sol, const double *rhs, const char *bound, const double *xlambda,
                  const int *imode, size_t bound_len);

/* solvediag (src/solvediag.f90:2) on one host real plane U(m,m), in place:
 * interior input = right-hand side, boundary input = Dirichlet data ('d',
 * with the boundary correction of :37-45) or ignored ('n'); output = the
 * solution including the Neumann boundary.  Also the entry for direct
 * callers such as precomp_mass_flow (src/operators.F:2276).               */
int duct_solvediag_(double *U, const char *bound, const double *xlambda, size_t bound_len);

/* duct_helm3d_batch_dev_ for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The scratch (two planar
 * buffers of 2 nfields nmodes NP^2 doubles for the larger boundary group, NP
 * the interior padded to 32/64/128/256/512, plus the Neumann edge sums) lives
 * in one static device array of 64 MiB per GPU, held until the process ends;
 * a batch that needs more is rejected with DUCT_ERR_ARG (m = 129: 0.5 MiB per
 * field-mode, about 120 field-modes; m = 257: 2 MiB, about 30).  Same results
 * as duct_helm3d_batch_dev_
 * with the unfolded FP64 passes (duct_set_fold_(0)); the two boundary groups
 * run one after the other on the library's stream.  Calls that overlap on one
 * GPU share the array and race.  Nothing is allocated: duct_init_ and
 * duct_setup_ must have run.                                               */
int duct_helm3d_batch_noalloc_dev_(const int *nfields, const duct_devptr_t *rhs,
                                   const duct_devptr_t *sol, const char *bounds,
                                   const double *alambda, const int *m, const int *nmodes,
                                   const int *ib, const double *alp, size_t bounds_len);

/* Mixed-precision refined variant of duct_helm3d_batch_dev_ (same fields,
 * boundary treatment and results to within the refinement tolerance).  The
 * reference solves in FP64 only (solvediag.f90:51-80, :140-219); here the four
 * eigenbasis products run on the fast matrix cores and FP64 is kept for the
 * residual r = f - (D2B_y u + u D2B_z^T - xlambda u), the eigenvalue divide,
 * the Neumann gauge, edges and corners:
 *   u = S[scheme0](f);  niter times: r = f - L u,  u += S[scheme1](r).
 * scheme: 1 = FP32 MFMA, 2 = bf16 split in 2 terms (3 products), 3 = bf16
 * split in 3 terms (6 products), 4 = plain bf16, 5 = the even/odd folded
 * solve: all four passes of a quarter fused in one workgroup, fp16 hi/lo
 * split (3 products), FP32 intermediates, the FP64 residual of the
 * centrosymmetric part of D2B in the folded basis; the leading Neumann modes
 * with xlambda < 64 (the singular k=0 plane among them) are solved by the
 * unfolded FP64 passes instead (scheme0 == scheme1 == 5, niter <= 2,
 * m with 32 < ceil((m-2)/2) <= 128), 6 = scheme 5 with the residual's two
 * products formed exactly from int8 digits on the int8 matrix cores (seven
 * base-256 digits of each row of the operator, three of the first solve's
 * result rounded to 2^-24 of its quarter's maximum) and summed to FP64
 * accuracy (scheme0 == scheme1 == 6, niter == 1).  niter in [0, 16].
 * stats (host, may be NULL): (niter+1) x nfields x 2 doubles = max|r| after
 * each step (the last one after the final correction) and max|f|, over the
 * refined modes (schemes 5 and 6: not the FP64-solved low Neumann modes).  */
int duct_helm3d_mixed_batch_dev_(const int *nfields, const duct_devptr_t *rhs,
                                 const duct_devptr_t *sol, const char *bounds,
                                 const double *alambda, const int *m, const int *nmodes,
                                 const int *ib, const double *alp, const int *scheme0,
                                 const int *scheme1, const int *niter, double *stats,
                                 size_t bounds_len);

/* Adjoint (transposed) solve of duct_helm3d_batch_dev_: rhsbar = S^T solbar for
 * every field, S the linear map rhs -> sol of the forward solve including its
 * boundary treatment (Dirichlet: rhs boundary ignored, sol boundary zero;
 * Neumann: sol boundary from the edge/corner equations, k=0 gauge).  The
 * reference has no adjoint; this is the backward of the solve for gradients
 * of flow functionals (solvediag.f90 with P -> P^-T and P^-1 -> P^T).
 * rhsbar's boundary is written as zero; solbar is not modified.            */
int duct_helm3d_adjoint_batch_dev_(const int *nfields, const duct_devptr_t *solbar,
                                   const duct_devptr_t *rhsbar, const char *bounds,
                                   const double *alambda, const int *m, const int *nmodes,
                                   const int *ib, const double *alp, size_t bounds_len);

/* Collocation derivative products on complex(8) fields (m, m, nmodes), full
 * planes: out = beta*out + scal * (D u) along y (cdir 'y') or (u D^T) along z
 * ('z'), D = DC (trans = 0) or DC^T (trans = 1), DC the CGL derivative matrix
 * uploaded by duct_setup_ (pseudomat.f90:90-107, cheb_collocation.F:19-75).
 * d/dy, d/dz of the nodal interpolant: the physical-space equivalent of
 * transyz + derivyz (chebyshev.f90:1-295, 458-540).  in and out must differ. */
int duct_colloc_deriv_dev_(const duct_devptr_t *in, const duct_devptr_t *out, const int *m,
                           const int *nmodes, const char *cdir, const int *trans,
                           const double *scal, const double *beta, size_t cdir_len);

/* fourx (src/fourier.F:1-115) on the reference's z-cut layouts:
 * fou(0:nmodes-1, my, nz) complex(8) and phys(mgalx+2, my, nz) real(8).
 * iopt >= 0: fou -> phys, modes above nmodes-1 zero (fill), FFTW c2r sum
 * (Im of mode 0 ignored), the two padding values written as 0; iopt < 0:
 * phys -> fou, the first nmodes modes of the r2c sum times 1/mgalx
 * (pack_fftw).  mgalx even, 2 nmodes <= mgalx.  Dense x-DFTs on the FP64
 * matrix cores for nmodes % 16 == 0, nmodes <= 64, mgalx % 16 == 0,
 * mgalx <= 192; batched rocFFT real transforms otherwise.                   */
int duct_fourx_dev_(const duct_devptr_t *fou, const duct_devptr_t *phys, const int *my, const int *nz,
                    const int *nmodes, const int *mgalx, const int *iopt);

/* Shen-Legendre spectral-Galerkin solve (-DHELM_SHEN: src/helm_shen.f90:2-85, solveshen.f90:5-513,
 * preshen.f90:2-535, legendre.f90), an independent cross-check of the collocation solve.
 * duct_setup_shen_ uploads one boundary type's Galerkin matrices, built by the caller (preshen):
 * with n = m-2, Ay (n x m) = P^T W and Az (n x m) = W map nodal values to the Galerkin inner
 * products (f, phi_p phi_q) (the y side through inv(P) inv(S) = P^T of M x = lam S x);
 * By (m x n) = R P and Bz (m x n) = R map coefficients back to nodal values; md, mu, sd the
 * mass diagonal, mass (k,k+2) entries and stiffness diagonal, lam the y-eigenvalues (n each).
 * duct_helm3d_shen_dev_ solves -lap(u) + xlambda u = rhs, xlambda = alambda + (alp (ib+k))^2,
 * nodal values in and out (rhs kept, unlike the reference, which leaves Chebyshev coefficients
 * in it): F = Ay f Az^T, the symmetric tridiagonal systems (lam_p (xlambda M + S) + M) v = h
 * along z per parity (solV), u = By V Bz^T.  Neumann at xlambda = 0: zero mean.  m <= 288.     */
int duct_setup_shen_(const char *bound, const int *m, const double *Ay, const double *Az, const double *By,
                     const double *Bz, const double *md, const double *mu, const double *sd, const double *lam,
                     size_t bound_len);
int duct_helm3d_shen_dev_(const duct_devptr_t *rhs, const duct_devptr_t *sol, const int *m, const int *nmodes,
                          const int *ib, const double *alp, const double *alambda, const char *bound,
                          size_t bound_len);

/* Mixed-boundary (temperature) solve, helm3d_temp (src/helm3d_temp.f90:2-44, int2d3d_temp.f90,
 * solvediag_temp.f90:2-352) on device fields.  duct_setup_temp_ is the GPU-upload tail of
 * pseudomaty_temp / pseudomatz_temp (src/pseudomat_temp.f90:2-321): bounds holds the four
 * wall types 'd'/'n' of y=-1, y=+1, z=-1, z=+1; PY, PYinv (PTY, PTYM1) are the interior
 * (my-2)^2 y eigenbasis, PZ, PZinv (PTZ, PTZM1) the (mz-2)^2 z eigenbasis, ETY / ETZ their
 * eigenvalues, DY (my x my) / DZ (mz x mz) the derivative matrices scaled by 1/sy, 1/sz (all
 * column-major); my and mz may differ (helm3d_temp.f90:35-36).  duct_helm3d_temp_dev_ solves
 * (D_yy + D_zz - xlambda) u = rhs, xlambda = alambda + (alp (ib+k))^2, for every mode k of the
 * complex(8) fields (my, mz, nmodes) with the wall data bcy (mz, 2, nmodes) over z and bcz
 * (my, 2, nmodes) over y (real(8); Dirichlet walls take the value, Neumann walls the y- or
 * z-derivative), used alike for Re and Im as int2d3d_temp does.  sol's interior and edges are
 * written, its corners are the rhs's, Im is zero at k=0; rhs is not modified (the reference
 * works on a copy).                                                                           */
int duct_setup_temp_(const char *bounds, const int *my, const int *mz, const double *PY, const double *PYinv,
                     const double *ETY, const double *DY, const double *PZ, const double *PZinv, const double *ETZ,
                     const double *DZ, size_t bounds_len);
int duct_helm3d_temp_dev_(const duct_devptr_t *rhs, const duct_devptr_t *sol, const duct_devptr_t *bcy,
                          const duct_devptr_t *bcz, const int *my, const int *mz, const int *nmodes, const int *ib,
                          const double *alp, const double *alambda);

/* One substep krk (1..3) of the low-storage RK3 for velocity and pressure, the
 * loop body of src/main1.f90:384-655 (items i-xi; icase = 1, no CMASSFLOW, forcing,
 * FILTER or temperature), on complex(8) device fields (m, m, nmodes), modes from 0:
 *   h    = -rkkc adv - gkkc advection3d(u) + fnu akkc lap(u) + u/dt
 *          - (akkc+bkkc) grad(p) - (akkc+bkkc) dpdx [x-momentum, mode 0]; h *= -1/(fnu bkkc)
 *   u,v,w <- helm3d(h, 'd', 1/(bkkc fnu dt));  phi = helm3d(div3d(u)/((akkc+bkkc) dt), 'n', 0)
 *   update_uvwp (operators.F:503-605): p += phi - (akkc+bkkc) dt fnu lap(phi),
 *                                       u -= dt (akkc+bkkc) grad(phi)
 * with the coefficients of mod.f90:51-58.  u, v, w, p are updated in place; adv1..3
 * hold advection3d of the previous substep's velocity on entry (read when krk > 1)
 * and of this substep's input velocity on exit.  dpdx is the caller's driving
 * gradient (get_dpdx, operators.F:2478-2522: -fnu dpval(aspect)); 0 for none.
 * The advection needs an even mgalx >= 3 nmodes (dense x-DFTs up to 64 modes and
 * 192 points, rocFFT past that).  Both bounds must be set up for m.               */
int duct_rk3_substep_dev_(const duct_devptr_t *u, const duct_devptr_t *v, const duct_devptr_t *w,
                          const duct_devptr_t *p, const duct_devptr_t *adv1, const duct_devptr_t *adv2,
                          const duct_devptr_t *adv3, const int *m, const int *nmodes, const int *mgalx,
                          const double *alp, const double *aspect, const double *fnu, const double *dt,
                          const int *krk, const double *dpdx);

/* Chebyshev transform of complex(8) fields (m, m, nmodes) along y or z
 * (transyz, src/chebyshev.f90:458-540, through tcheb_fftw, src/tchebw.f:62-150,
 * REDFT00 with the reference's end scalings and no dealiasing): iopt < 0
 * physical -> Chebyshev coefficients of cos(pi j/N), iopt >= 0 the inverse.
 * Applied as dense m x m products on the FP64 matrix cores; out-of-place
 * (in and out must differ; the reference transforms in place).           */
int duct_transyz_dev_(const duct_devptr_t *in, const duct_devptr_t *out, const int *m, const int *nmodes,
                      const char *cdir, const int *iopt, size_t cdir_len);

/* derivyz (src/chebyshev.f90:1-295) on Chebyshev coefficients: du = scal *
 * d^ider u / dy^ider (what 'n') or du += (any other what), z derivatives
 * divided by aspect^ider; ider 1 or 2; u and du must differ.  The reference's
 * backward coefficient recurrence, one line per thread.                     */
int duct_derivyz_dev_(const duct_devptr_t *u, const duct_devptr_t *du, const int *m, const int *nmodes,
                      const char *cdir, const int *ider, const double *scal, const char *what,
                      const double *aspect, size_t cdir_len, size_t what_len);

/* Fourier x-derivative: out = beta*out + scal * (i*alp*k) * in for modes
 * k = ib .. ib+nmodes-1 (derivx, src/fourier.F:117-188, ider = 1).          */
int duct_derivx_dev_(const duct_devptr_t *in, const duct_devptr_t *out, const int *m,
                     const int *nmodes, const int *ib, const double *alp, const double *scal,
                     const double *beta);

/* Advection term in divergence form (advection3d, src/operators.F:9-276) on
 * device fields u, v, w: complex(8) (m, m, nmodes), Fourier modes 0..nmodes-1
 * in x (the whole x spectrum, i.e. the reference's z-cut after chx2z), physical
 * (CGL) in y and z, y fastest.  Writes
 *   adv1 = d(uu)/dx + d(uv)/dy + d(uw)/dz,  adv2 = d(uv)/dx + d(vv)/dy + d(vw)/dz,
 *   adv3 = d(uw)/dx + d(vw)/dy + d(ww)/dz
 * with the products formed on mgalx physical x points by the 3/2 rule (fourx,
 * mult, pack_fftw: src/fourier.F:1-115, src/mult.F) and d/dx = i alp k,
 * d/dy = DC, d/dz = DC / aspect (the collocation form of transyz + derivyz,
 * src/chebyshev.f90).  Needs duct_setup_ for m and an even mgalx >= 3 nmodes
 * (dense x-DFTs up to 64 modes and 192 points, rocFFT past that).  The adv fields must
 * not alias u, v, w.  The reference's slip-velocity, CFL and statistics side
 * outputs of advection3d are not computed.                                  */
int duct_advection_dev_(const duct_devptr_t *u, const duct_devptr_t *v, const duct_devptr_t *w,
                        const duct_devptr_t *adv1, const duct_devptr_t *adv2, const duct_devptr_t *adv3,
                        const int *m, const int *nmodes, const int *mgalx, const double *alp,
                        const double *aspect);

/* ---- bandwidth-bound helpers (reference GPU glue) ----------------------- */

/* C = alpha * A^T for nbatch consecutive Nx x Ny FP64 matrices (column-major,
 * C is Ny x Nx).  Replaces transpose_gpu_ (src/mytranspose_cuda.cu:132-162,
 * kernel transposeCoalesced :69-88), which ignores alpha (callers pass 1.0)
 * and silently drops sizes that are not multiples of 32; any size works here.
 * A and C must not overlap.                                                 */
int duct_transpose_dev_(const int *Nx, const int *Ny, const int *nbatch, const double *alpha,
                        const duct_devptr_t *A, const duct_devptr_t *C);

/* Fourier <-> physical pencil transposes of complex(8) data:
 * phys(0:nx-1, my, nz) -> phys2(my, nz, 0:nx-1) and back.
 * Replace cxyz2yzx / cyzx2xyz (src/parallel.F:476-501, :504-529).           */
int duct_cxyz2yzx_dev_(const int *nx, const int *my, const int *nz, const duct_devptr_t *phys,
                       const duct_devptr_t *phys2);
int duct_cyzx2xyz_dev_(const int *nx, const int *my, const int *nz, const duct_devptr_t *phys2,
                       const duct_devptr_t *phys);

/* C = F ./ (ED2 - xlam) elementwise over Nx*Ny FP64 values.
 * Replaces mat_invdiag_gpu_ (src/mytranspose_cuda.cu:181-196, kernel :45-53). */
int duct_mat_invdiag_dev_(const int *Nx, const int *Ny, const double *xlam, const duct_devptr_t *F,
                          const duct_devptr_t *ED2, const duct_devptr_t *C);

/* ---- even/odd folded Dirichlet solve ------------------------------------ */

/* Even/odd folded eigenbasis for the FP64 solves (duct_helm3d_dev_, _batch_dev_,
 * duct_helm3d_, duct_int2d3d_, duct_solvediag_): half the FLOPs of the four
 * products.  D2B commutes with the reflection of the CGL nodes, so the
 * columns of PD / PN are even or odd; the folded solve uses their parity
 * projections.  mode 0 (default): the reference's unfolded products
 * (solvediag.f90:51-80, :140-219); 1: Dirichlet folded, agreeing with the
 * unfolded solve to ~1e-13 relative at m = 257; 2: Neumann folded as well,
 * except its leading modes with xlambda < 64 (the singular k=0 plane with its
 * gauge, :161-206, among them), which take the unfolded passes: on those
 * ill-conditioned Poisson planes the folded operator's solution differs from
 * D2B's by up to ~4e-11 of the field at m = 257, from xlambda = 64 up by
 * ~2e-13.  The reference has no such switch.                                */
int duct_set_fold_(const int *mode);

/* Concurrency of the boundary groups of a batched solve.  on != 0 (default):
 * when a call has both Dirichlet and Neumann fields, the Neumann group runs on
 * a second stream owned by the library (default priority), forked from and
 * joined back into the caller's stream with events, so its latency-bound
 * kernels overlap the Dirichlet passes; the results are the same to the bit.
 * Kernel timing (duct_prof_enable_) keeps the groups in order.  on = 0: both
 * groups on the caller's stream, one after the other.  The reference solves
 * one field per helm3d call and has no such switch.                         */
int duct_set_overlap_(const int *on);

/* Execution of the folded solve's middle passes (duct_set_fold_ 1 and 2).
 * on != 0 (default): where the folded size is 128 (m = 131 ... 259), the
 * undivided second product, the eigenvalue divide and the third product of
 * each 128 x 128 quarter run in one kernel, with the quarter held in LDS.
 * on = 0: two planar launches.  The results are the same to the bit.  The reference has no such switch
 * (its products are cuBLAS calls, solvediag.f90:51-80, :140-219).          */
int duct_set_fused_quarters_(const int *on);

/* Folded size NH (padded ceil(n/2)) of the basis set up for bound, 0 when it
 * does not fold (n < 65, or eigenvectors without a clean parity split).     */
int duct_fold_size_(const char *bound, int *NH, size_t bound_len);

/* ---- Fourier-space field files (HDF5) ---------------------------------- */
/* Host code: these read uvwp_NNNN.h5 / uvwpt_NNNN.h5 files into host arrays
 * and need no device (nor duct_init_).  The library parses the HDF5 format
 * itself (superblock versions 0-3; contiguous or compact datasets in the root
 * group, either byte order); chunked or filtered datasets and dense link
 * storage are DUCT_ERR_ARG with a message.  Character arguments are Fortran
 * strings: trailing blanks are ignored.                                      */

/* Header of a field file (readheader_hdf5, src/readwrite_hdf5.f90:798-850):
 * ints[5] = {mgalx, mx1, my, mz, nparams}; reals[7] = {alp, aspect, time,
 * fnu, cfl, dtfixed, cmassflow}.  A file without 'mx1' gives
 * 2*(mgalx/3)/2 - 1, as the reference does (:824-826); a missing optional
 * entry is NaN (reals) or -1 (nparams).                                      */
int duct_h5_read_header_(const char *filename, int *ints, double *reals, size_t filename_len);

/* One field ('u', 'v', 'w', 'p' or 't') into a host complex(8) array
 * out(my, mz, nmodes) holding the modes ib .. ib+nmodes-1 (replaces the body
 * of read_4fields_fou_hdf5 / read_5fields_fou_hdf5 and h5load_parallel,
 * src/readwrite_hdf5.f90:2-264).  Modes past the file's mx1 are zero.  A file
 * of another grid (my and mz both differ) is interpolated as the reference's
 * cheb_interp does (src/chebyshev.f90:725-765: the first 20 x 20 Chebyshev
 * coefficients evaluated on the new grid); a file where only one of my, mz
 * differs is DUCT_ERR_UNSUPPORTED (the reference copies a mis-shaped section
 * there, :134-140).                                                           */
int duct_h5_read_field_(const char *filename, const char *name, double *out, const int *my,
                        const int *mz, const int *ib, const int *nmodes, size_t filename_len,
                        size_t name_len);

/* Any numeric dataset of the root group (header scalars, the wall data bcy(mz,2,0:mx1)
 * and bcz(my,2,0:mx1) of a temperature file) as *count doubles in the file's
 * order; *count must equal the dataset's size (h5read_simple_double and
 * h5read_simple_int, src/readwrite_hdf5.f90:994-1038).                       */
int duct_h5_read_doubles_(const char *filename, const char *name, double *out,
                          const long long *count, size_t filename_len, size_t name_len);

/* A character dataset (the wall types boundy1 ... boundz2 of a temperature
 * file, written by h5write_simple_char, :970-992) into out[0 .. *length-1];
 * at most *capacity characters.                                              */
int duct_h5_read_chars_(const char *filename, const char *name, char *out, const int *capacity,
                        int *length, size_t filename_len, size_t name_len);

/* ---- kernel timing (hipEvents on the library stream) -------------------- */

/* Enable (1) / disable (0) per-launch event timing and clear the records. */
int duct_prof_enable_(const int *on);
/* Synchronise and return, for each kernel class k < *nclass, the summed
 * device time in ms and the launch count.  Classes: 0 proj_c2p,
 * 1 proj_p2p_div, 2 gauge, 3 proj_p2p, 4 proj_p2c, 5 boundary, 6 bc_correct,
 * 7 transpose, 8 invdiag, 9 lp_c2p, 10 lp_p2p_div, 11 lp_p2p, 12 lp_p2c,
 * 13 resid, 14 deriv, 15 adjoint (Neumann boundary adjoint), 16 fold_c2p,
 * 17 fold_p2p_div, 18 fold_p2p, 19 fold_p2c, 20 xprod (advection products), 21 tridiag (Shen),
 * 22 fold_mid (folded passes 2 and 3 on LDS-resident quarters).                                */
int duct_prof_read_(const int *nclass, double *ms, int *count);

#ifdef __cplusplus
}
#endif
#endif /* DUCT_HIP_H */
