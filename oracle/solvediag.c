/*
 * CPU oracle for the cross-plane Helmholtz/Poisson solve stage.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, statement by statement, the
 * reference's FP64 CPU path so that the HIP library can be checked against it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or as the timed CPU baseline; the product
 * library never links or calls it.
 *
 * Restated reference routines (files under src/ of ando2800/duct):
 *   oracle_solvediag  <- solvediag.f90:2-330   (ngpu=0 branch, both BCs)
 *   oracle_int2d3d    <- int2d3d.f90:2-53, com2re :55-80, re2com :82-109
 *   oracle_helm3d     <- helm3d.f90:36-73     (mode loop, xlambda, BC zeroing)
 * The basis arrays (PDT, PDM1T, ED2, ... as produced by pseudomat.f90) are
 * built by oracle/pseudomat.py and passed in through oracle_basis.
 *
 * All matrices are Fortran column-major.  The reference's padded scratch
 * F, FT, tmprt (m-1)x(m-1) (solvediag.f90:15-16, mod.f90:341) are kept,
 * including their zero row/column.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* BLAS dgemm with the Fortran calling convention (no hidden lengths needed). */
typedef void (*dgemm_fn)(const char *ta, const char *tb, const int *m, const int *n,
                         const int *k, const double *alpha, const double *a,
                         const int *lda, const double *b, const int *ldb,
                         const double *beta, double *c, const int *ldc);

typedef struct {
    int m;              /* CGL points per direction (my = mz)                  */
    /* Dirichlet (fulldiagdir, mod.f90:332-343) */
    const double *ED2;  /* (m-1)^2 padded, ED(i)+ED(j)            pseudomat.f90:247-251 */
    const double *PDT;  /* (m-1)^2 padded, transpose of PD         :258-263 */
    const double *PDM1T;/* (m-1)^2 padded, transpose of inv(PD)    :258-263 */
    /* Neumann (fulldiagneu, mod.f90:317-330) */
    const double *EN;   /* (m-2)                                   :131 */
    const double *PN;   /* (m-2)^2 right eigenvectors              :131 */
    const double *PNT;  /* (m-1)^2 padded                          :181-186 */
    const double *PNM1T;/* (m-1)^2 padded                          :181-186 */
    const double *AM1;  /* 4x4 inverse corner matrix               :158-178 */
    double wanted;      /* gauge value, 0                          :156 */
    int iofzero;        /* 1-based eigen-index of min |EN|         :146-152 */
    /* chebystuff (mod.f90:277-288) */
    const double *DC;   /* m x m collocation derivative            :90-107 */
    const double *D2bx; /* m x 2: D2(:,1), D2(:,m)                 :225-230 */
    const double *D2by; /* m x 2                                   :225-230 */
} oracle_basis;

static dgemm_fn g_dgemm = 0;

/* Select an external BLAS dgemm (NULL -> the built-in loop below). */
void oracle_set_dgemm(void *fn) { g_dgemm = (dgemm_fn)fn; }

/* C = A^T * B  (dgemm('t','n', M,N,K, 1, A,lda, B,ldb, 0, C,ldc)). */
static void gemm_tn(int M, int N, int K, const double *A, int lda, const double *B, int ldb,
                    double *C, int ldc)
{
    if (g_dgemm) {
        const double one = 1.0, zero = 0.0;
        g_dgemm("T", "N", &M, &N, &K, &one, A, &lda, B, &ldb, &zero, C, &ldc);
        return;
    }
    for (int j = 0; j < N; ++j) {
        const double *b = B + (size_t)j * ldb;
        for (int i = 0; i < M; ++i) {
            const double *a = A + (size_t)i * lda;
            double s = 0.0;
            for (int k = 0; k < K; ++k) s += a[k] * b[k];
            C[i + (size_t)j * ldc] = s;
        }
    }
}

#define IX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/*
 * solvediag.f90:2-330, ngpu=0 path.  U (m x m) in/out; work holds
 * F, FT, tmprt, each (m-1)^2.  Returns 0, or 1 on a bad bound character
 * (the reference prints and stops, :324-326).
 */
int oracle_solvediag(const oracle_basis *b, double *U, char bound, double xlambda, double *work)
{
    const int m = b->m, n1 = m - 1, mm2 = m - 2;
    double *F = work, *FT = work + (size_t)n1 * n1, *tmprt = work + 2 * (size_t)n1 * n1;
    memset(work, 0, 3 * sizeof(double) * (size_t)n1 * n1);           /* :24-26 */

    for (int j = 1; j < m - 1; ++j)                                   /* :28-32 */
        for (int i = 1; i < m - 1; ++i) F[IX(i - 1, j - 1, n1)] = U[IX(i, j, m)];

    if (bound == 'd') {
        for (int j = 0; j < mm2; ++j)                                  /* :37-43 */
            for (int i = 0; i < mm2; ++i) {
                double bcontv = b->D2bx[IX(i + 1, 0, m)] * U[IX(0, j + 1, m)] +
                                b->D2bx[IX(i + 1, 1, m)] * U[IX(m - 1, j + 1, m)];
                double bconth = b->D2by[IX(j + 1, 0, m)] * U[IX(i + 1, 0, m)] +
                                b->D2by[IX(j + 1, 1, m)] * U[IX(i + 1, m - 1, m)];
                F[IX(i, j, n1)] = F[IX(i, j, n1)] - bcontv - bconth;
            }
        gemm_tn(n1, n1, n1, b->PDM1T, n1, F, n1, FT, n1);             /* :51-52 */
        for (int i = 0; i < n1; ++i)                                   /* :54-58 */
            for (int j = 0; j < n1; ++j) tmprt[IX(i, j, n1)] = FT[IX(j, i, n1)];
        gemm_tn(n1, n1, n1, tmprt, n1, b->PDM1T, n1, F, n1);          /* :59-60 */
        for (int j = 0; j < mm2; ++j)                                  /* :61-66 */
            for (int i = 0; i < mm2; ++i)
                FT[IX(i, j, n1)] = F[IX(i, j, n1)] / (b->ED2[IX(i, j, n1)] - xlambda);
        gemm_tn(n1, n1, n1, b->PDT, n1, FT, n1, F, n1);               /* :71-72 */
        for (int i = 0; i < n1; ++i)                                   /* :74-78 */
            for (int j = 0; j < n1; ++j) tmprt[IX(i, j, n1)] = F[IX(j, i, n1)];
        gemm_tn(n1, n1, n1, tmprt, n1, b->PDT, n1, FT, n1);           /* :79-80 */
        for (int j = 0; j < mm2; ++j)                                  /* :131-135 */
            for (int i = 0; i < mm2; ++i) U[IX(i + 1, j + 1, m)] = FT[IX(i, j, n1)];
        return 0;
    }
    if (bound != 'n') return 1;

    /* Neumann, :137-323 */
    gemm_tn(n1, n1, n1, b->PNM1T, n1, F, n1, FT, n1);                 /* :140-141 */
    for (int i = 0; i < mm2; ++i)                                      /* :143-147 */
        for (int j = 0; j < mm2; ++j) tmprt[IX(i, j, n1)] = FT[IX(j, i, n1)];
    gemm_tn(n1, n1, n1, tmprt, n1, b->PNM1T, n1, F, n1);              /* :148-149 */

    if (fabs(xlambda) > 1.0e-10) {                                     /* :151-159 */
        for (int j = 0; j < mm2; ++j)
            for (int i = 0; i < mm2; ++i) {
                double den = (b->EN[i] + b->EN[j] - xlambda);
                FT[IX(i, j, n1)] = F[IX(i, j, n1)] / den;
            }
    } else {                                                           /* :161-206 */
        const int io = b->iofzero - 1; /* 0-based */
        for (int j = 0; j < io; ++j)
            for (int i = 0; i < mm2; ++i)
                FT[IX(i, j, n1)] = F[IX(i, j, n1)] / (b->EN[i] + b->EN[j] - xlambda);
        {
            const int j = io;
            for (int i = 0; i < io; ++i)
                FT[IX(i, j, n1)] = F[IX(i, j, n1)] / (b->EN[i] + b->EN[j] - xlambda);
            FT[IX(io, io, n1)] = 0.0;
            for (int i = io + 1; i < mm2; ++i)
                FT[IX(i, j, n1)] = F[IX(i, j, n1)] / (b->EN[i] + b->EN[j] - xlambda);
        }
        for (int j = io + 1; j < mm2; ++j)
            for (int i = 0; i < mm2; ++i)
                FT[IX(i, j, n1)] = F[IX(i, j, n1)] / (b->EN[i] + b->EN[j] - xlambda);
        double valore = 0.0;                                           /* :196-205 */
        for (int l = 0; l < mm2; ++l) {
            double val = 0.0;
            for (int k = 0; k < mm2; ++k) val = val + b->PN[IX(io, k, mm2)] * FT[IX(l, k, n1)];
            valore = valore + b->PN[IX(io, l, mm2)] * val;
        }
        const double p = b->PN[IX(io, io, mm2)];
        FT[IX(io, io, n1)] = (b->wanted - valore) / (p * p);
    }

    gemm_tn(n1, n1, n1, b->PNT, n1, FT, n1, F, n1);                   /* :210-211 */
    for (int i = 0; i < mm2; ++i)                                      /* :213-217 */
        for (int j = 0; j < mm2; ++j) tmprt[IX(i, j, n1)] = F[IX(j, i, n1)];
    gemm_tn(n1, n1, n1, tmprt, n1, b->PNT, n1, FT, n1);               /* :218-219 */

    for (int j = 0; j < mm2; ++j)                                      /* :244-249 */
        for (int i = 0; i < mm2; ++i) U[IX(i + 1, j + 1, m)] = FT[IX(i, j, n1)];

    const double *DC = b->DC;
    const double det = DC[IX(0, 0, m)] * DC[IX(m - 1, m - 1, m)] - DC[IX(0, m - 1, m)] * DC[IX(m - 1, 0, m)];
    for (int i = 1; i < m - 1; ++i) {                                  /* :254-265 */
        double sumb = 0.0, sumt = 0.0;
        for (int j = 0; j < mm2; ++j) {
            sumb = sumb + DC[IX(0, j + 1, m)] * FT[IX(i - 1, j, n1)];
            sumt = sumt + DC[IX(m - 1, j + 1, m)] * FT[IX(i - 1, j, n1)];
        }
        U[IX(i, 0, m)] = (-DC[IX(m - 1, m - 1, m)] * sumb + DC[IX(0, m - 1, m)] * sumt) / det;
        U[IX(i, m - 1, m)] = (DC[IX(m - 1, 0, m)] * sumb - DC[IX(0, 0, m)] * sumt) / det;
    }
    for (int j = 1; j < m - 1; ++j) {                                  /* :267-276 */
        double suml = 0.0, sumr = 0.0;
        for (int i = 0; i < mm2; ++i) {
            suml = suml + DC[IX(0, i + 1, m)] * FT[IX(i, j - 1, n1)];
            sumr = sumr + DC[IX(m - 1, i + 1, m)] * FT[IX(i, j - 1, n1)];
        }
        U[IX(0, j, m)] = (-DC[IX(m - 1, m - 1, m)] * suml + DC[IX(0, m - 1, m)] * sumr) / det;
        U[IX(m - 1, j, m)] = (DC[IX(m - 1, 0, m)] * suml - DC[IX(0, 0, m)] * sumr) / det;
    }
    double ff[4], s;                                                    /* :280-302 */
    s = 0.0;
    for (int i = 1; i < m - 1; ++i) s = s - DC[IX(0, i, m)] * (U[IX(i, 0, m)] + U[IX(0, i, m)]);
    ff[0] = s;
    s = 0.0;
    for (int i = 1; i < m - 1; ++i) s = s - DC[IX(0, i, m)] * U[IX(m - 1, i, m)] + DC[IX(m - 1, i, m)] * U[IX(i, 0, m)];
    ff[1] = s;
    s = 0.0;
    for (int i = 1; i < m - 1; ++i) s = s - DC[IX(m - 1, i, m)] * (U[IX(i, m - 1, m)] + U[IX(m - 1, i, m)]);
    ff[2] = s;
    s = 0.0;
    for (int i = 1; i < m - 1; ++i) s = s - DC[IX(0, i, m)] * U[IX(i, m - 1, m)] + DC[IX(m - 1, i, m)] * U[IX(0, i, m)];
    ff[3] = s;
    /* :304-322: rows 1..4 of AM1 give U(1,1), U(m,1), U(m,m), U(1,m) */
    const size_t corner[4] = {IX(0, 0, m), IX(m - 1, 0, m), IX(m - 1, m - 1, m), IX(0, m - 1, m)};
    for (int r = 0; r < 4; ++r) {
        double v = 0.0;
        for (int i = 0; i < 4; ++i) v = v + b->AM1[IX(r, i, 4)] * ff[i];
        U[corner[r]] = v;
    }
    return 0;
}

size_t oracle_work_doubles(int m) { return (size_t)m * m + 3 * (size_t)(m - 1) * (m - 1); }

/*
 * int2d3d.f90:36-47 (non-Shen branch).  rhs, sol: one complex(8) plane
 * (m x m, interleaved re/im).  The imaginary solve is skipped at imode=0 (:43).
 */
int oracle_int2d3d(const oracle_basis *b, double *sol, const double *rhs, char bound, double xlambda,
                   int imode, double *work)
{
    const int m = b->m;
    const size_t mm = (size_t)m * m;
    double *vr = work, *w2 = work + mm;
    for (size_t e = 0; e < mm; ++e) vr[e] = rhs[2 * e];             /* com2re 'r' */
    int st = oracle_solvediag(b, vr, bound, xlambda, w2);
    if (st) return st;
    for (size_t e = 0; e < mm; ++e) { sol[2 * e] = vr[e]; sol[2 * e + 1] = 0.0; } /* re2com 'r' */
    if (imode == 0) return 0;
    for (size_t e = 0; e < mm; ++e) vr[e] = rhs[2 * e + 1];         /* com2re 'i' */
    st = oracle_solvediag(b, vr, bound, xlambda, w2);
    if (st) return st;
    for (size_t e = 0; e < mm; ++e) sol[2 * e + 1] = sol[2 * e + 1] + vr[e]; /* re2com 'i' */
    return 0;
}

/*
 * helm3d.f90:48-73.  rhs, sol: complex(8) (m, m, nmodes), modes ib..ib+nmodes-1.
 * Dirichlet zeroes the boundary of rhs in place (:48-60).  xlambda =
 * alambda + (alp*imode)^2 with xalp(i)=alp*i (fourier.F:201-203).  Modes are
 * independent, so nthreads > 1 splits them over OpenMP threads, each with its
 * own work array (the reference's flat-MPI mode split, pointer.f90).
 */
int oracle_helm3d(const oracle_basis *b, double *rhs, double *sol, int nmodes, int ib, double alp,
                  double alambda, char bound, int nthreads)
{
    const int m = b->m;
    const size_t mm = (size_t)m * m;
    if (bound != 'd' && bound != 'n') return 1;
    if (bound == 'd') {
        for (int k = 0; k < nmodes; ++k) {
            double *p = rhs + 2 * mm * k;
            for (int j = 0; j < m; ++j) {
                p[2 * IX(j, 0, m)] = 0.0; p[2 * IX(j, 0, m) + 1] = 0.0;
                p[2 * IX(j, m - 1, m)] = 0.0; p[2 * IX(j, m - 1, m) + 1] = 0.0;
            }
            for (int kk = 0; kk < m; ++kk) {
                p[2 * IX(0, kk, m)] = 0.0; p[2 * IX(0, kk, m) + 1] = 0.0;
                p[2 * IX(m - 1, kk, m)] = 0.0; p[2 * IX(m - 1, kk, m) + 1] = 0.0;
            }
        }
    }
    int status = 0;
    if (nthreads < 1) nthreads = 1;
    const size_t wsz = mm + oracle_work_doubles(m);
#pragma omp parallel num_threads(nthreads) reduction(| : status)
    {
        double *work = (double *)malloc(sizeof(double) * wsz);
        if (!work) status |= 2;
        else {
#pragma omp for schedule(dynamic, 1)
            for (int k = 0; k < nmodes; ++k) {
                const int imode = ib + k;
                const double xa = alp * (double)imode;
                const double xlambda = alambda + xa * xa;
                status |= oracle_int2d3d(b, sol + 2 * mm * k, rhs + 2 * mm * k, bound, xlambda, imode, work);
            }
            free(work);
        }
    }
    return status;
}
