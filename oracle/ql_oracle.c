/*
 * ql_oracle.c -- CPU restatement of the reference QL time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path; see ql_oracle.h for
 * who may load it).  Each function below names the reference lines it
 * restates.  The arithmetic keeps the reference's expression order, with the
 * reference's "imaginary storage" of wavenumbers (kx, ky, kz held as i*k,
 * Auxiliary/Kdata.cpp:32-33,76-78) written as complex values with zero real
 * part, so every product is the one the reference's Eigen expressions form.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fcx-fortran-rules
 * -fopenmp); -ffp-contract=off keeps a*b+c unfused so that the remap
 * predicate (MHD_FullQlin.cpp:470-472) is evaluated exactly as written.
 */
#include "ql_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* General_Definitions.h:236 defines PI = atan(1)*4, which rounds to M_PI. */
static const double QL_PI = 3.14159265358979323846;

/* RK3 coefficients, Integrators/RK3CN.cpp:19-20 and RK3CN_Elin.cpp:19-20 */
static const double RK_A0 = 8.0 / 15.0, RK_A1 = 5.0 / 12.0, RK_A2 = 3.0 / 4.0;
static const double RK_B1 = -17.0 / 60.0, RK_B2 = -5.0 / 12.0;
static const double RK_P1 = 8.0 / 15.0, RK_P2 = 2.0 / 3.0;

static inline cplx mk(double re, double im) { return CMPLX(re, im); }

/* ------------------------------------------------------------------------ */
/* Radix-2 FFT (stands in for the FFTW plans of Auxiliary/fftwPlans.cpp:48-51)*/
/* ------------------------------------------------------------------------ */
typedef struct {
    int N;
    cplx *tw;  /* tw[k] = exp(-2 pi i k / N), k < N/2 */
    int *rev;  /* bit reversal permutation */
} fftplan;

static int ilog2(int n) { int l = 0; while ((1 << l) < n) ++l; return l; }

static void plan_init(fftplan *f, int N)
{
    int lg = ilog2(N);
    f->N = N;
    f->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(N / 2 > 0 ? N / 2 : 1));
    f->rev = (int *)malloc(sizeof(int) * (size_t)N);
    for (int k = 0; k < N / 2; ++k) {
        double a = 2.0 * QL_PI * (double)k / (double)N;
        f->tw[k] = mk(cos(a), -sin(a));
    }
    for (int i = 0; i < N; ++i) {
        int r = 0;
        for (int b = 0; b < lg; ++b)
            if (i & (1 << b)) r |= 1 << (lg - 1 - b);
        f->rev[i] = r;
    }
}

static void plan_free(fftplan *f) { free(f->tw); free(f->rev); }

/* in-place unnormalised transform: sign -1 -> exp(-i...), +1 -> exp(+i...) */
static void fft_exec(const fftplan *f, cplx *a, int sign)
{
    const int N = f->N;
    for (int i = 0; i < N; ++i) {
        int j = f->rev[i];
        if (i < j) { cplx t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int len = 2; len <= N; len <<= 1) {
        int hl = len >> 1, step = N / len;
        for (int i = 0; i < N; i += len) {
            for (int k = 0; k < hl; ++k) {
                cplx w = f->tw[k * step];
                if (sign > 0) w = conj(w);
                cplx u = a[i + k], v = a[i + k + hl] * w;
                a[i + k] = u + v;
                a[i + k + hl] = u - v;
            }
        }
    }
}

/* fftwPlans::inverse = pad_with_zeros (fftwPlans.h:66-82) + back_1D (:57-59) */
static void fft_inverse(const fftplan *f, int nz, const cplx *in, cplx *out)
{
    const int NZ = f->N;
    int deal_low = (nz - 1) / 2 + 1, deal_high = NZ - (nz - 1) / 2;
    int i = 0, j = deal_low;
    while (i < deal_low) { out[i] = in[i]; ++i; }
    while (j < deal_high) { out[j] = 0.0; ++j; }
    while (j < NZ) { out[j] = in[i]; ++i; ++j; }
    fft_exec(f, out, +1);
}

/* fftwPlans::forward = for_1D (fftwPlans.h:54-56, in place) + dealias (:84-97) */
static void fft_forward(const fftplan *f, int nz, cplx *inout, cplx *out)
{
    const int NZ = f->N;
    int deal_low = (nz - 1) / 2 + 1, deal_high = NZ - (nz - 1) / 2;
    fft_exec(f, inout, -1);
    int i = 0, j = deal_high;
    while (i < deal_low) { out[i] = inout[i]; ++i; }
    while (j < NZ) { out[i] = inout[j]; ++i; ++j; }
}

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
    cplx *u[9], *b[9];
    cplx *tmp1_z, *tmp2_z, *tmp3_z, *rey_z, *rey2_z;
    cplx *tmp1_k, *tmp2_k, *tmp3_k, *fy, *fz, *fy2, *fz2;
    double *lapF, *ilapF, *lap2t, *ilap2t;
    cplx *send;
    double diag[6];
} work_t;

struct orc {
    orc_params p;
    int nx, ny, nxy, nz, nzl, half, NZ, nMF, nLin, nthreads;
    double fac1D, rs_ft_fac, kxLfac, kyLfac, kzLfac, omega_eff;
    long totalN2;
    double *kx0, *ky;
    int *kx_index, *ky_index;
    cplx *kz;
    double *kz2;
    double *lap2, *ilap2; /* [ny][nz] */
    fftplan plan;
    cplx *lin, *mf;
    cplx *rhs1_lin, *rhs1_mf, *rhs2_lin, *rhs2_mf;
    double *linop, *linop_old, *linop_mf;
    cplx *drive;  /* [nMF][nz] */
    cplx *mfz;    /* [8][NZ] real-space mean-field profiles */
    work_t *work;
    /* DrivingNoise settings (MHD_FullQlin.cpp:18-22, 47-52, 65-66) */
    double f_noise, noise_range[2], mult_noise_fac;
    int drive_only_velocity, drive_only_magnetic, dont_drive_ky0;
    uint64_t noise_seed;
    int *ky0pkx, *ky0nkx; /* ky = 0 pairs, kx > 0 and kx < 0 (Kdata.cpp:122-146) */
};

static void *xcalloc(size_t n, size_t s) { void *p = calloc(n ? n : 1, s); return p; }

static int model_nMF(int model)
{
    switch (model) {
    case ORC_MODEL_MHD_BQLIN: return 2;
    case ORC_MODEL_HD_FULLU: return 2; /* Ux, Uy (HD_fullU.cpp:14) */
    case ORC_MODEL_CONSTANT_DAMPING: return 1;
    default: return 4;
    }
}

/* Kdata::Kdata (Auxiliary/Kdata.cpp:11-120) for a single rank */
static void build_tables(orc *o)
{
    const int nx = o->nx, ny = o->ny;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            int k = j + ny * i;
            o->kx0[k] = i * o->kxLfac;
            o->ky[k] = j * o->kyLfac;
            o->kx_index[k] = i;
            o->ky_index[k] = j;
        }
    for (int i = nx + 1; i < 2 * nx; ++i)
        for (int j = 0; j < ny; ++j) {
            int k = j + ny * (i - 1);
            o->kx0[k] = (-2 * nx + i) * o->kxLfac;
            o->ky[k] = j * o->kyLfac;
            o->kx_index[k] = -2 * nx + i;
            o->ky_index[k] = j;
        }
    const int nzl = o->nzl;
    for (int i = 0; i < nzl + 1; ++i) o->kz[i] = mk(0.0, i * o->kzLfac);
    for (int i = nzl + 1; i < 2 * nzl + 1; ++i) o->kz[i] = mk(0.0, (-2 * nzl + i - 1) * o->kzLfac);
    for (int m = 0; m < o->nz; ++m) o->kz2[m] = creal(o->kz[m] * o->kz[m]);
    for (int j = 0; j < ny; ++j) {
        double kyr = o->ky[j]; /* ky_tmp[i] for i < ny has index j */
        for (int m = 0; m < o->nz; ++m) {
            double l2 = -kyr * kyr + o->kz2[m];
            o->lap2[j * o->nz + m] = l2;
            o->ilap2[j * o->nz + m] = 1 / l2;
        }
        if (kyr == 0) {
            o->lap2[j * o->nz + 0] = 0;
            o->ilap2[j * o->nz + 0] = 1;
        }
    }
}

orc *orc_create(const orc_params *p)
{
    orc *o = (orc *)xcalloc(1, sizeof(orc));
    o->p = *p;
    o->NZ = p->NZ;
    o->nx = p->NX / 2;
    o->ny = p->NY / 2;
    o->nxy = (p->NX - 1) * (p->NY / 2);          /* Model.h:32-36 */
    o->nz = 1 + 2 * (int)(p->NZ / 3.0);          /* Model.h:34 */
    o->nzl = (o->nz - 1) / 2;
    o->half = (o->nz - 1) / 2 + 1;               /* MHD_FullQlin.cpp:86 */
    o->nMF = model_nMF(p->model);
    /* HD_fullU has 2 fluctuation fields (u, zeta; HD_fullU.cpp:14); they are
     * stored in the first 2 of 4 slots per mode, the other 2 stay zero */
    o->nLin = p->model == ORC_MODEL_HD_FULLU ? 2 : 4;
    o->fac1D = 1.0 / p->NZ;                      /* fftwPlans.cpp:30 */
    o->rs_ft_fac = 1.0 / (p->NY * p->NX) / (p->NY * p->NX); /* MHD_FullQlin.cpp:67 */
    o->totalN2 = (long)(p->NX * p->NY * p->NZ);
    o->totalN2 = o->totalN2 * o->totalN2;        /* MHD_FullQlin.cpp:63-64 */
    o->kxLfac = 2 * QL_PI / p->L[0];
    o->kyLfac = 2 * QL_PI / p->L[1];
    o->kzLfac = 2 * QL_PI / p->L[2];
    /* MHD_BQlin.cpp:19 hard-codes Omega = 2q/3 */
    o->omega_eff = (p->model == ORC_MODEL_MHD_BQLIN) ? 2.0 / 3.0 * p->q : p->omega;
#ifdef _OPENMP
    o->nthreads = p->nthreads > 0 ? p->nthreads : omp_get_max_threads();
#else
    o->nthreads = 1;
#endif
    if (o->nthreads > o->nxy) o->nthreads = o->nxy > 0 ? o->nxy : 1;

    const int nxy = o->nxy, nz = o->nz, NZ = o->NZ, nMF = o->nMF;
    o->kx0 = (double *)xcalloc(nxy, sizeof(double));
    o->ky = (double *)xcalloc(nxy, sizeof(double));
    o->kx_index = (int *)xcalloc(nxy, sizeof(int));
    o->ky_index = (int *)xcalloc(nxy, sizeof(int));
    o->kz = (cplx *)xcalloc(nz, sizeof(cplx));
    o->kz2 = (double *)xcalloc(nz, sizeof(double));
    o->lap2 = (double *)xcalloc((size_t)o->ny * nz, sizeof(double));
    o->ilap2 = (double *)xcalloc((size_t)o->ny * nz, sizeof(double));
    build_tables(o);
    plan_init(&o->plan, NZ);

    size_t nl = (size_t)nxy * 4 * nz;
    o->lin = (cplx *)xcalloc(nl, sizeof(cplx));
    o->rhs1_lin = (cplx *)xcalloc(nl, sizeof(cplx));
    o->rhs2_lin = (cplx *)xcalloc(nl, sizeof(cplx));
    o->mf = (cplx *)xcalloc((size_t)nMF * nz, sizeof(cplx));
    o->rhs1_mf = (cplx *)xcalloc((size_t)nMF * nz, sizeof(cplx));
    o->rhs2_mf = (cplx *)xcalloc((size_t)nMF * nz, sizeof(cplx));
    o->linop = (double *)xcalloc(nl, sizeof(double));
    o->linop_old = (double *)xcalloc(nl, sizeof(double));
    o->linop_mf = (double *)xcalloc((size_t)nMF * nz, sizeof(double));
    o->drive = (cplx *)xcalloc((size_t)nMF * nz, sizeof(cplx));
    o->mfz = (cplx *)xcalloc((size_t)8 * NZ, sizeof(cplx));

    o->work = (work_t *)xcalloc(o->nthreads, sizeof(work_t));
    for (int t = 0; t < o->nthreads; ++t) {
        work_t *w = &o->work[t];
        for (int k = 0; k < 9; ++k) {
            w->u[k] = (cplx *)xcalloc(NZ, sizeof(cplx));
            w->b[k] = (cplx *)xcalloc(NZ, sizeof(cplx));
        }
        w->tmp1_z = (cplx *)xcalloc(NZ, sizeof(cplx));
        w->tmp2_z = (cplx *)xcalloc(NZ, sizeof(cplx));
        w->tmp3_z = (cplx *)xcalloc(NZ, sizeof(cplx));
        w->rey_z = (cplx *)xcalloc(NZ, sizeof(cplx));
        w->rey2_z = (cplx *)xcalloc(NZ, sizeof(cplx));
        w->tmp1_k = (cplx *)xcalloc(nz, sizeof(cplx));
        w->tmp2_k = (cplx *)xcalloc(nz, sizeof(cplx));
        w->tmp3_k = (cplx *)xcalloc(nz, sizeof(cplx));
        w->fy = (cplx *)xcalloc(nz, sizeof(cplx));
        w->fz = (cplx *)xcalloc(nz, sizeof(cplx));
        w->fy2 = (cplx *)xcalloc(nz, sizeof(cplx));
        w->fz2 = (cplx *)xcalloc(nz, sizeof(cplx));
        w->lapF = (double *)xcalloc(nz, sizeof(double));
        w->ilapF = (double *)xcalloc(nz, sizeof(double));
        w->lap2t = (double *)xcalloc(nz, sizeof(double));
        w->ilap2t = (double *)xcalloc(nz, sizeof(double));
        w->send = (cplx *)xcalloc((size_t)nMF * o->half, sizeof(cplx));
    }
    /* noise: off by default (Input_parameters.cpp:107-111) */
    o->noise_range[0] = 0.0;
    o->noise_range[1] = 1e16 * 1e16;
    o->mult_noise_fac = 1.0 / (16 * 32 * 32);
    o->mult_noise_fac = o->mult_noise_fac * o->mult_noise_fac;
    o->dont_drive_ky0 = (p->model == ORC_MODEL_MHD_BQLIN); /* MHD_BQlin.cpp:18 */
    {
        /* Kdata.cpp:122-131 for a single rank */
        int nx = o->nx, ny = o->ny;
        o->ky0pkx = (int *)xcalloc(nx > 1 ? nx - 1 : 1, sizeof(int));
        o->ky0nkx = (int *)xcalloc(nx > 1 ? nx - 1 : 1, sizeof(int));
        for (int i = 0; i < 2 * nx - 2; ++i) {
            if (i < nx - 1) o->ky0pkx[i] = (i + 1) * ny;
            if (i >= nx - 1) o->ky0nkx[2 * (nx - 1) - i - 1] = (i + 1) * ny;
        }
    }
    orc_linear_ops_init(o, 0.0, NULL, NULL);
    return o;
}

void orc_destroy(orc *o)
{
    if (!o) return;
    for (int t = 0; t < o->nthreads; ++t) {
        work_t *w = &o->work[t];
        for (int k = 0; k < 9; ++k) { free(w->u[k]); free(w->b[k]); }
        free(w->tmp1_z); free(w->tmp2_z); free(w->tmp3_z); free(w->rey_z); free(w->rey2_z);
        free(w->tmp1_k); free(w->tmp2_k); free(w->tmp3_k);
        free(w->fy); free(w->fz); free(w->fy2); free(w->fz2);
        free(w->lapF); free(w->ilapF); free(w->lap2t); free(w->ilap2t); free(w->send);
    }
    free(o->work);
    free(o->kx0); free(o->ky); free(o->kx_index); free(o->ky_index);
    free(o->kz); free(o->kz2); free(o->lap2); free(o->ilap2);
    plan_free(&o->plan);
    free(o->lin); free(o->rhs1_lin); free(o->rhs2_lin);
    free(o->mf); free(o->rhs1_mf); free(o->rhs2_mf);
    free(o->linop); free(o->linop_old); free(o->linop_mf);
    free(o->drive); free(o->mfz);
    free(o->ky0pkx); free(o->ky0nkx);
    free(o);
}

void orc_dims(const orc *o, int *d)
{
    d[0] = o->nxy; d[1] = o->nz; d[2] = o->nMF; d[3] = o->nLin; d[4] = o->half; d[5] = o->NZ;
}

/* external [nxy][nLin][nz] <-> the 4-slot storage */
static void lin_in(const orc *o, cplx *dst, const cplx *src)
{
    const size_t nz = o->nz, nL = o->nLin;
    if (nL == 4) {
        memcpy(dst, src, sizeof(cplx) * o->nxy * 4 * nz);
        return;
    }
    memset(dst, 0, sizeof(cplx) * o->nxy * 4 * nz);
    for (int i = 0; i < o->nxy; ++i) memcpy(dst + (size_t)i * 4 * nz, src + (size_t)i * nL * nz, sizeof(cplx) * nL * nz);
}
static void lin_out(const orc *o, void *dst, const void *src, size_t esz)
{
    const size_t nz = o->nz, nL = o->nLin;
    for (int i = 0; i < o->nxy; ++i)
        memcpy((char *)dst + (size_t)i * nL * nz * esz, (const char *)src + (size_t)i * 4 * nz * esz, esz * nL * nz);
}

void orc_set_state(orc *o, const double *lin, const double *mf, const double *kx0)
{
    if (lin) lin_in(o, o->lin, (const cplx *)lin);
    if (mf) memcpy(o->mf, mf, sizeof(cplx) * (size_t)o->nMF * o->nz);
    if (kx0) memcpy(o->kx0, kx0, sizeof(double) * (size_t)o->nxy);
}

void orc_get_state(const orc *o, double *lin, double *mf, double *kx0)
{
    if (lin) lin_out(o, lin, o->lin, sizeof(cplx));
    if (mf) memcpy(mf, o->mf, sizeof(cplx) * (size_t)o->nMF * o->nz);
    if (kx0) memcpy(kx0, o->kx0, sizeof(double) * (size_t)o->nxy);
}

void orc_tables(const orc *o, double *kx0, double *ky, int *kxi, int *kyi, double *kz)
{
    if (kx0) memcpy(kx0, o->kx0, sizeof(double) * (size_t)o->nxy);
    if (ky) memcpy(ky, o->ky, sizeof(double) * (size_t)o->nxy);
    if (kxi) memcpy(kxi, o->kx_index, sizeof(int) * (size_t)o->nxy);
    if (kyi) memcpy(kyi, o->ky_index, sizeof(int) * (size_t)o->nxy);
    if (kz) for (int m = 0; m < o->nz; ++m) kz[m] = cimag(o->kz[m]);
}

/* ------------------------------------------------------------------------ */
/* assign_laplacians_ (MHD_FullQlin.cpp:722-738, MHD_BQlin.cpp:493-509)      */
/* ------------------------------------------------------------------------ */
typedef struct { cplx kxc, kyc; double kx, ky; int ind_ky; } kmode;

static void assign_laplacians(const orc *o, int i, double t, int need_inverse, work_t *w, kmode *k)
{
    const int nz = o->nz;
    k->ind_ky = o->ky_index[i];
    k->kyc = mk(0.0, o->ky[i]);
    k->ky = cimag(k->kyc);
    k->kxc = mk(0.0, o->kx0[i]) + o->p.q * t * k->kyc;
    k->kx = cimag(k->kxc);
    const double *l2 = o->lap2 + (size_t)k->ind_ky * nz;
    for (int m = 0; m < nz; ++m) {
        w->lap2t[m] = l2[m];
        w->lapF[m] = -k->kx * k->kx + w->lap2t[m];
    }
    if (need_inverse) {
        const double *il2 = o->ilap2 + (size_t)k->ind_ky * nz;
        for (int m = 0; m < nz; ++m) {
            w->ilap2t[m] = il2[m];
            w->ilapF[m] = 1 / w->lapF[m];
        }
        if (k->ind_ky == 0 && o->kx_index[i] == 0) w->ilapF[0] = 1.0;
    }
}

/* mode block of a thread, the contiguous split of MPIdata.cpp:30-33 */
static void block_of(int n, int nt, int tid, int *b0, int *b1)
{
    *b0 = (int)((long)n * tid / nt);
    *b1 = (int)((long)n * (tid + 1) / nt);
}

/* Allreduce (MHD_FullQlin.cpp:386) as an in-order sum, then the Hermitian
 * fill of :389-399 into o->drive */
static void reduce_and_fill(orc *o, cplx *recv)
{
    const int half = o->half, nMF = o->nMF, nz = o->nz;
    for (int k = 0; k < nMF * half; ++k) recv[k] = 0.0;
    for (int t = 0; t < o->nthreads; ++t)
        for (int k = 0; k < nMF * half; ++k) recv[k] += o->work[t].send[k];
    for (int f = 0; f < nMF; ++f) {
        cplx *d = o->drive + (size_t)f * nz;
        const cplx *r = recv + (size_t)f * half;
        for (int m = 0; m < half; ++m) d[m] = r[m];
        for (int m = 0; m < half - 1; ++m) d[half + m] = conj(r[half - 1 - m]);
    }
}

/* ------------------------------------------------------------------------ */
/* MHD_FullQlin::rhs (Models/MHD_FullQlin.cpp:171-427)                       */
/* ------------------------------------------------------------------------ */
static void rhs_fullqlin_mode(orc *o, work_t *w, int i, double t, double dt_lin,
                              const cplx *lin, cplx *out, double *linop)
{
    const int nz = o->nz, NZ = o->NZ, half = o->half;
    const double fac = o->fac1D, q = o->p.q, Om = o->omega_eff, B0z = o->p.B0z;
    const fftplan *pl = &o->plan;
    const cplx *kz = o->kz;
    const cplx *MBx = o->mfz + 0 * NZ, *dzMBx = o->mfz + 1 * NZ, *MBy = o->mfz + 2 * NZ,
               *dzMBy = o->mfz + 3 * NZ, *MUx = o->mfz + 4 * NZ, *dzMUx = o->mfz + 5 * NZ,
               *MUy = o->mfz + 6 * NZ, *dzMUy = o->mfz + 7 * NZ;
    const double Lm = 1.0; /* L_mult_ = !turn_off_fluct_Lorentz_force_ (:24,:111) */
    kmode k;
    assign_laplacians(o, i, t, 1, w, &k);
    const cplx kxc = k.kxc, kyc = k.kyc;
    const cplx *ux = lin + ((size_t)i * 4 + 0) * nz, *ze = lin + ((size_t)i * 4 + 1) * nz;
    const cplx *bx = lin + ((size_t)i * 4 + 2) * nz, *et = lin + ((size_t)i * 4 + 3) * nz;
    cplx *o0 = out + ((size_t)i * 4 + 0) * nz, *o1 = out + ((size_t)i * 4 + 1) * nz;
    cplx *o2 = out + ((size_t)i * 4 + 2) * nz, *o3 = out + ((size_t)i * 4 + 3) * nz;
    cplx *fy = w->fy, *fz = w->fz, *t1 = w->tmp1_k, *t2 = w->tmp2_k, *t3 = w->tmp3_k;
    cplx **u = w->u, **b = w->b;
    const double *ilap2 = w->ilap2t, *ilapF = w->ilapF;

    /* :220-221 */
    for (int m = 0; m < nz; ++m) {
        fy[m] = ((-kyc * kxc) * ux[m] + kz[m] * ze[m]) * ilap2[m];
        fz[m] = (-kxc * kz[m] * ux[m] - kyc * ze[m]) * ilap2[m];
    }
    /* :225-242 */
    for (int m = 0; m < nz; ++m) t1[m] = fac * ux[m];
    fft_inverse(pl, nz, t1, u[0]);
    for (int m = 0; m < nz; ++m) t1[m] = fac * fy[m];
    fft_inverse(pl, nz, t1, u[1]);
    for (int m = 0; m < nz; ++m) t1[m] = fac * fz[m];
    fft_inverse(pl, nz, t1, u[2]);
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kxc) * ux[m];
    fft_inverse(pl, nz, t1, u[3]);
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kxc) * fy[m];
    fft_inverse(pl, nz, t1, u[4]);
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kxc) * fz[m];
    fft_inverse(pl, nz, t1, u[5]);
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kyc) * ux[m];
    fft_inverse(pl, nz, t1, u[6]);
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kyc) * fy[m];
    fft_inverse(pl, nz, t1, u[7]);
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kyc) * fz[m];
    fft_inverse(pl, nz, t1, u[8]);
    /* :245-265 */
    for (int m = 0; m < nz; ++m) {
        fy[m] = ((-kyc * kxc) * bx[m] + kz[m] * et[m]) * ilap2[m];
        fz[m] = (-kxc * kz[m] * bx[m] - kyc * et[m]) * ilap2[m];
    }
    for (int m = 0; m < nz; ++m) t1[m] = fac * bx[m];
    fft_inverse(pl, nz, t1, b[0]);
    for (int m = 0; m < nz; ++m) t1[m] = fac * fy[m];
    fft_inverse(pl, nz, t1, b[1]);
    for (int m = 0; m < nz; ++m) t1[m] = fac * fz[m];
    fft_inverse(pl, nz, t1, b[2]);
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kxc) * bx[m];
    fft_inverse(pl, nz, t1, b[3]);
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kxc) * fy[m];
    fft_inverse(pl, nz, t1, b[4]);
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kxc) * fz[m];
    fft_inverse(pl, nz, t1, b[5]);
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kyc) * bx[m];
    fft_inverse(pl, nz, t1, b[6]);
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kyc) * fy[m];
    fft_inverse(pl, nz, t1, b[7]);
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kyc) * fz[m];
    fft_inverse(pl, nz, t1, b[8]);

    /* momentum, :271-285 */
    cplx *z1 = w->tmp1_z;
    for (int j = 0; j < NZ; ++j)
        z1[j] = MUx[j] * u[3][j] + MUy[j] * u[6][j] + u[2][j] * dzMUx[j] -
                Lm * (MBx[j] * b[3][j] + MBy[j] * b[6][j] + b[2][j] * dzMBx[j]);
    fft_forward(pl, nz, z1, t1);
    for (int j = 0; j < NZ; ++j)
        z1[j] = MUx[j] * u[4][j] + MUy[j] * u[7][j] + u[2][j] * dzMUy[j] -
                Lm * (MBx[j] * b[4][j] + MBy[j] * b[7][j] + b[2][j] * dzMBy[j]);
    fft_forward(pl, nz, z1, t2);
    for (int j = 0; j < NZ; ++j)
        z1[j] = MUx[j] * u[5][j] + MUy[j] * u[8][j] - Lm * (MBx[j] * b[5][j] + MBy[j] * b[8][j]);
    fft_forward(pl, nz, z1, t3);
    for (int m = 0; m < nz; ++m) {
        o0[m] = ((-2.0 * kxc * kyc * q) * ux[m] + 2 * Om * kz[m] * ze[m] + (kxc * kxc) * t1[m] +
                 (kxc * kyc) * t2[m] + kxc * kz[m] * t3[m]) * ilapF[m] +
                B0z * kz[m] * bx[m] - t1[m];
        o1[m] = (q - 2.0 * Om) * kz[m] * ux[m] + B0z * kz[m] * et[m] - (kz[m] * t2[m] - kyc * t3[m]);
    }
    /* induction, :291-305 */
    for (int j = 0; j < NZ; ++j)
        z1[j] = MUx[j] * b[3][j] + MUy[j] * b[6][j] + u[2][j] * dzMBx[j] -
                (MBx[j] * u[3][j] + MBy[j] * u[6][j] + b[2][j] * dzMUx[j]);
    fft_forward(pl, nz, z1, t1);
    for (int j = 0; j < NZ; ++j)
        z1[j] = MUx[j] * b[4][j] + MUy[j] * b[7][j] + u[2][j] * dzMBy[j] -
                (MBx[j] * u[4][j] + MBy[j] * u[7][j] + b[2][j] * dzMUy[j]);
    fft_forward(pl, nz, z1, t2);
    for (int j = 0; j < NZ; ++j)
        z1[j] = MUx[j] * b[5][j] + MUy[j] * b[8][j] - (MBx[j] * u[5][j] + MBy[j] * u[8][j]);
    fft_forward(pl, nz, z1, t3);
    for (int m = 0; m < nz; ++m) {
        o2[m] = B0z * kz[m] * ux[m] - t1[m];
        o3[m] = -q * kz[m] * bx[m] + B0z * kz[m] * ze[m] - (kz[m] * t2[m] - kyc * t3[m]);
    }

    /* EMF and stresses, :314-362 */
    double mult_fac = 2.0;
    if (k.ky == 0.0) mult_fac = 1.0;
    double ftfac = mult_fac * o->rs_ft_fac;
    cplx *rz = w->rey_z, *rz2 = w->rey2_z, *send = w->send;
    for (int j = 0; j < NZ; ++j)
        rz[j] = creal(b[2][j] * conj(u[0][j])) - creal(u[2][j] * conj(b[0][j]));
    fft_forward(pl, nz, rz, t1);
    for (int m = 0; m < half; ++m) send[m] += ftfac * (kz[m] * t1[m]);
    for (int j = 0; j < NZ; ++j)
        rz[j] = creal(b[2][j] * conj(u[1][j])) - creal(u[2][j] * conj(b[1][j]));
    fft_forward(pl, nz, rz, t1);
    for (int m = 0; m < half; ++m) send[half + m] += ftfac * (kz[m] * t1[m]);
    for (int m = 0; m < nz; ++m) {
        t1[m] = fac * kz[m] * ux[m];
        t2[m] = fac * kz[m] * bx[m];
    }
    fft_inverse(pl, nz, t1, z1);
    fft_inverse(pl, nz, t2, rz2);
    for (int j = 0; j < NZ; ++j)
        rz[j] = -(creal(u[0][j] * conj(u[3][j])) + creal(u[1][j] * conj(u[6][j])) + creal(u[2][j] * conj(z1[j]))) +
                (creal(b[0][j] * conj(b[3][j])) + creal(b[1][j] * conj(b[6][j])) + creal(b[2][j] * conj(rz2[j])));
    fft_forward(pl, nz, rz, t1);
    for (int m = 0; m < half; ++m) send[2 * half + m] += ftfac * t1[m];
    /* :352 recomputes uy into fluct_z_; fluct_y_ still holds by */
    for (int m = 0; m < nz; ++m) fz[m] = ((-kyc * kxc) * ux[m] + kz[m] * ze[m]) * ilap2[m];
    for (int m = 0; m < nz; ++m) {
        t1[m] = fac * kz[m] * fz[m];
        t2[m] = fac * kz[m] * fy[m];
    }
    fft_inverse(pl, nz, t1, z1);
    fft_inverse(pl, nz, t2, rz2);
    for (int j = 0; j < NZ; ++j)
        rz[j] = -(creal(u[0][j] * conj(u[4][j])) + creal(u[1][j] * conj(u[7][j])) + creal(u[2][j] * conj(z1[j]))) +
                (creal(b[0][j] * conj(b[4][j])) + creal(b[1][j] * conj(b[7][j])) + creal(b[2][j] * conj(rz2[j])));
    fft_forward(pl, nz, rz, t1);
    for (int m = 0; m < half; ++m) send[3 * half + m] += ftfac * t1[m];

    /* linear part, :371-376 */
    double kxt = k.kx + o->p.q * dt_lin * k.ky;
    double *L0 = linop + ((size_t)i * 4 + 0) * nz;
    for (int m = 0; m < nz; ++m) {
        double lapF = -pow(fabs((-kxt * kxt - k.ky * k.ky) + o->kz2[m]), (double)(o->p.viscosity_order / 2));
        L0[m] = o->p.nu * lapF;
        L0[nz + m] = L0[m];
        L0[2 * nz + m] = o->p.eta * lapF;
        L0[3 * nz + m] = L0[2 * nz + m];
    }
}

/* ------------------------------------------------------------------------ */
/* HD_fullU::rhs (Models/HD_fullU.cpp:145-309)                               */
/* ------------------------------------------------------------------------ */
static void rhs_hd_mode(orc *o, work_t *w, int i, double t, double dt_lin, const cplx *lin, cplx *out,
                        double *linop)
{
    const int nz = o->nz, NZ = o->NZ, half = o->half;
    const double fac = o->fac1D, q = o->p.q, Om = o->omega_eff;
    const fftplan *pl = &o->plan;
    const cplx *kz = o->kz;
    const cplx *MUx = o->mfz + 4 * NZ, *dzMUx = o->mfz + 5 * NZ, *MUy = o->mfz + 6 * NZ, *dzMUy = o->mfz + 7 * NZ;
    kmode k;
    assign_laplacians(o, i, t, 1, w, &k);
    const cplx kxc = k.kxc, kyc = k.kyc;
    const cplx *ux = lin + ((size_t)i * 4 + 0) * nz, *ze = lin + ((size_t)i * 4 + 1) * nz;
    cplx *o0 = out + ((size_t)i * 4 + 0) * nz, *o1 = out + ((size_t)i * 4 + 1) * nz;
    cplx *fy = w->fy, *fz = w->fz, *t1 = w->tmp1_k, *t2 = w->tmp2_k, *t3 = w->tmp3_k;
    cplx **u = w->u;
    const double *ilap2 = w->ilap2t, *ilapF = w->ilapF;

    /* :184-185 */
    for (int m = 0; m < nz; ++m) {
        fy[m] = ((-kyc * kxc) * ux[m] + kz[m] * ze[m]) * ilap2[m];
        fz[m] = (-kxc * kz[m] * ux[m] - kyc * ze[m]) * ilap2[m];
    }
    /* :189-206 */
    const cplx *src[3] = {ux, fy, fz};
    for (int d = 0; d < 3; ++d) {
        const cplx f = d == 0 ? fac : (d == 1 ? fac * kxc : fac * kyc);
        for (int c = 0; c < 3; ++c) {
            for (int m = 0; m < nz; ++m) t1[m] = f * src[c][m];
            fft_inverse(pl, nz, t1, u[3 * d + c]);
        }
    }
    /* :213-228 */
    cplx *z1 = w->tmp1_z;
    for (int j = 0; j < NZ; ++j) z1[j] = MUx[j] * u[3][j] + MUy[j] * u[6][j] + u[2][j] * dzMUx[j];
    fft_forward(pl, nz, z1, t1);
    for (int j = 0; j < NZ; ++j) z1[j] = MUx[j] * u[4][j] + MUy[j] * u[7][j] + u[2][j] * dzMUy[j];
    fft_forward(pl, nz, z1, t2);
    for (int j = 0; j < NZ; ++j) z1[j] = MUx[j] * u[5][j] + MUy[j] * u[8][j];
    fft_forward(pl, nz, z1, t3);
    for (int m = 0; m < nz; ++m) {
        o0[m] = ((-2.0 * kxc * kyc * q) * ux[m] + 2 * Om * kz[m] * ze[m] + (kxc * kxc) * t1[m] +
                 (kxc * kyc) * t2[m] + kxc * kz[m] * t3[m]) * ilapF[m] - t1[m];
        o1[m] = (q - 2.0 * Om) * kz[m] * ux[m] - (kz[m] * t2[m] - kyc * t3[m]);
    }
    /* Reynolds stresses, :235-264 */
    double mult_fac = 2.0;
    if (k.ky == 0.0) mult_fac = 1.0;
    double ftfac = mult_fac * o->rs_ft_fac;
    cplx *rz = w->rey_z, *send = w->send;
    for (int m = 0; m < nz; ++m) t1[m] = fac * kz[m] * ux[m];
    fft_inverse(pl, nz, t1, z1);
    for (int j = 0; j < NZ; ++j)
        rz[j] = -(creal(u[0][j] * conj(u[3][j])) + creal(u[1][j] * conj(u[6][j])) + creal(u[2][j] * conj(z1[j])));
    fft_forward(pl, nz, rz, t1);
    for (int m = 0; m < half; ++m) send[m] += ftfac * t1[m];
    for (int m = 0; m < nz; ++m) fz[m] = ((-kyc * kxc) * ux[m] + kz[m] * ze[m]) * ilap2[m];
    for (int m = 0; m < nz; ++m) t1[m] = fac * kz[m] * fz[m];
    fft_inverse(pl, nz, t1, z1);
    for (int j = 0; j < NZ; ++j)
        rz[j] = -(creal(u[0][j] * conj(u[4][j])) + creal(u[1][j] * conj(u[7][j])) + creal(u[2][j] * conj(z1[j])));
    fft_forward(pl, nz, rz, t1);
    for (int m = 0; m < half; ++m) send[half + m] += ftfac * t1[m];
    /* linear part, :270-273 */
    double kxt = k.kx + o->p.q * dt_lin * k.ky;
    double *L0 = linop + ((size_t)i * 4 + 0) * nz;
    for (int m = 0; m < nz; ++m) {
        L0[m] = o->p.nu * ((-kxt * kxt - k.ky * k.ky) + o->kz2[m]);
        L0[nz + m] = L0[m];
        L0[2 * nz + m] = 0.0;
        L0[3 * nz + m] = 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* MHD_BQlin::rhs (Models/MHD_BQlin.cpp:117-284)                             */
/* ------------------------------------------------------------------------ */
static void rhs_bqlin_mode(orc *o, work_t *w, int i, double t, double dt_lin,
                           const cplx *lin, cplx *out, double *linop)
{
    const int nz = o->nz, NZ = o->NZ, half = o->half;
    const double fac = o->fac1D, q = o->p.q, Om = o->omega_eff;
    const fftplan *pl = &o->plan;
    const cplx *kz = o->kz;
    const cplx *By = o->mfz + 0 * NZ, *dzBy = o->mfz + 1 * NZ, *dzdzBy = o->mfz + 2 * NZ;
    kmode k;
    assign_laplacians(o, i, t, 1, w, &k);
    const cplx kxc = k.kxc, kyc = k.kyc;
    const cplx *ux = lin + ((size_t)i * 4 + 0) * nz, *ze = lin + ((size_t)i * 4 + 1) * nz;
    const cplx *bx = lin + ((size_t)i * 4 + 2) * nz, *et = lin + ((size_t)i * 4 + 3) * nz;
    cplx *o0 = out + ((size_t)i * 4 + 0) * nz, *o1 = out + ((size_t)i * 4 + 1) * nz;
    cplx *o2 = out + ((size_t)i * 4 + 2) * nz, *o3 = out + ((size_t)i * 4 + 3) * nz;
    cplx *uy = w->fy, *uz = w->fz, *by = w->fy2, *bz = w->fz2;
    cplx *t1 = w->tmp1_k, *t2 = w->tmp2_k, *t3 = w->tmp3_k;
    cplx *z1 = w->tmp1_z, *z2 = w->tmp2_z, *z3 = w->tmp3_z;
    const double *ilap2 = w->ilap2t, *ilapF = w->ilapF;

    /* :143-146 */
    for (int m = 0; m < nz; ++m) {
        uy[m] = ((-kyc * kxc) * ux[m] + kz[m] * ze[m]) * ilap2[m];
        uz[m] = (-kxc * kz[m] * ux[m] - kyc * ze[m]) * ilap2[m];
        by[m] = ((-kyc * kxc) * bx[m] + kz[m] * et[m]) * ilap2[m];
        bz[m] = (-kxc * kz[m] * bx[m] - kyc * et[m]) * ilap2[m];
    }
    /* u, :153-162 */
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kyc) * bx[m];
    fft_inverse(pl, nz, t1, z1);
    for (int j = 0; j < NZ; ++j) z1[j] *= By[j];
    for (int m = 0; m < nz; ++m) t2[m] = (-2.0 * fac * kxc * kyc) * bz[m];
    fft_inverse(pl, nz, t2, z2);
    for (int j = 0; j < NZ; ++j) z2[j] *= dzBy[j];
    fft_forward(pl, nz, z1, o0);
    fft_forward(pl, nz, z2, t2);
    for (int m = 0; m < nz; ++m)
        o0[m] += ((-2.0 * kxc * kyc * q) * ux[m] + 2 * Om * kz[m] * ze[m] + t2[m]) * ilapF[m];
    /* zeta, :166-178 */
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kyc) * et[m];
    fft_inverse(pl, nz, t1, z1);
    for (int j = 0; j < NZ; ++j) z1[j] *= By[j];
    for (int m = 0; m < nz; ++m) t2[m] = (-fac * kxc) * bx[m];
    fft_inverse(pl, nz, t2, z2);
    for (int j = 0; j < NZ; ++j) z2[j] *= dzBy[j];
    for (int m = 0; m < nz; ++m) t3[m] = fac * bz[m];
    fft_inverse(pl, nz, t3, z3);
    for (int j = 0; j < NZ; ++j) z3[j] *= dzdzBy[j];
    for (int j = 0; j < NZ; ++j) z1[j] += z2[j] + z3[j];
    fft_forward(pl, nz, z1, o1);
    for (int m = 0; m < nz; ++m) o1[m] += (q - 2.0 * Om) * kz[m] * ux[m];
    /* b, :182-185 */
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kyc) * ux[m];
    fft_inverse(pl, nz, t1, z1);
    for (int j = 0; j < NZ; ++j) z1[j] *= By[j];
    fft_forward(pl, nz, z1, o2);
    /* eta, :189-201 */
    for (int m = 0; m < nz; ++m) t1[m] = (fac * kyc) * ze[m];
    fft_inverse(pl, nz, t1, z1);
    for (int j = 0; j < NZ; ++j) z1[j] *= By[j];
    for (int m = 0; m < nz; ++m) t2[m] = (-2 * fac * kz[m] * uz[m]) - (fac * kxc) * ux[m];
    fft_inverse(pl, nz, t2, z2);
    for (int j = 0; j < NZ; ++j) z2[j] *= dzBy[j];
    for (int m = 0; m < nz; ++m) t3[m] = -fac * uz[m];
    fft_inverse(pl, nz, t3, z3);
    for (int j = 0; j < NZ; ++j) z3[j] *= dzdzBy[j];
    for (int j = 0; j < NZ; ++j) z1[j] += z2[j] + z3[j];
    fft_forward(pl, nz, z1, o3);
    for (int m = 0; m < nz; ++m) o3[m] -= q * kz[m] * bx[m];

    /* EMF, :212-240 */
    double mult_fac = 2.0;
    if (k.ky == 0.0) mult_fac = 1.0;
    double ftfac = mult_fac * fac * fac * o->rs_ft_fac;
    cplx *rz = w->rey_z, *send = w->send;
    fft_inverse(pl, nz, uz, z1);
    fft_inverse(pl, nz, bz, z2);
    fft_inverse(pl, nz, ux, z3);
    for (int j = 0; j < NZ; ++j) rz[j] = creal(z2[j] * conj(z3[j]));
    fft_inverse(pl, nz, bx, z3);
    for (int j = 0; j < NZ; ++j) rz[j] -= creal(z1[j] * conj(z3[j]));
    fft_forward(pl, nz, rz, t1);
    for (int m = 0; m < half; ++m) send[m] += ftfac * (kz[m] * t1[m]);
    fft_inverse(pl, nz, uy, z3);
    for (int j = 0; j < NZ; ++j) rz[j] = creal(z2[j] * conj(z3[j]));
    fft_inverse(pl, nz, by, z3);
    for (int j = 0; j < NZ; ++j) rz[j] -= creal(z1[j] * conj(z3[j]));
    fft_forward(pl, nz, rz, t1);
    for (int m = 0; m < half; ++m) send[half + m] += ftfac * (kz[m] * t1[m]);

    /* linear part, :248-252 */
    double kxt = k.kx + o->p.q * dt_lin * k.ky;
    double *L0 = linop + ((size_t)i * 4 + 0) * nz;
    for (int m = 0; m < nz; ++m) {
        L0[m] = o->p.nu * ((-kxt * kxt - k.ky * k.ky) + o->kz2[m]);
        L0[nz + m] = L0[m];
        L0[2 * nz + m] = (o->p.eta / o->p.nu) * L0[m];
        L0[3 * nz + m] = L0[2 * nz + m];
    }
}

/* ConstantDamping::rhs (Models/ConstantDamping.cpp:73-111), dampFac = 0.5 */
static void rhs_damping_mode(orc *o, int i, const cplx *lin, cplx *out, double *linop)
{
    const int nz = o->nz;
    for (int v = 0; v < 4; ++v)
        for (int m = 0; m < nz; ++m) {
            out[((size_t)i * 4 + v) * nz + m] = -lin[((size_t)i * 4 + v) * nz + m] * 0.5;
            linop[((size_t)i * 4 + v) * nz + m] = 0.0;
        }
}

/* Oracle-only test model: zero nonlinear part, MHD_FullQlin's dissipation
 * operator (:371-376).  With no RHS the integrators reduce to their linear
 * factors, which is what it pins. */
static void rhs_linear_mode(orc *o, work_t *w, int i, double t, double dt_lin, cplx *out, double *linop)
{
    const int nz = o->nz;
    kmode k;
    assign_laplacians(o, i, t, 0, w, &k);
    double kxt = k.kx + o->p.q * dt_lin * k.ky;
    for (int m = 0; m < nz; ++m) {
        double lapF = -pow(fabs((-kxt * kxt - k.ky * k.ky) + o->kz2[m]), (double)(o->p.viscosity_order / 2));
        for (int v = 0; v < 4; ++v) out[((size_t)i * 4 + v) * nz + m] = 0.0;
        linop[((size_t)i * 4 + 0) * nz + m] = o->p.nu * lapF;
        linop[((size_t)i * 4 + 1) * nz + m] = o->p.nu * lapF;
        linop[((size_t)i * 4 + 2) * nz + m] = o->p.eta * lapF;
        linop[((size_t)i * 4 + 3) * nz + m] = o->p.eta * lapF;
    }
}

/* the whole Model::rhs for the configured model, on a given input state */
static void model_rhs(orc *o, double t, double dt_lin, const cplx *lin, const cplx *mf,
                      cplx *out_lin, cplx *out_mf, double *linop)
{
    const int nz = o->nz, NZ = o->NZ, half = o->half, nMF = o->nMF;
    const double fac = o->fac1D;
    const int model = o->p.model;
    cplx *tmp = (cplx *)xcalloc(nz, sizeof(cplx));
    if (model == ORC_MODEL_MHD_FULLQLIN) {
        /* MHD_FullQlin.cpp:180-198 */
        for (int f = 0; f < 4; ++f) {
            for (int m = 0; m < nz; ++m) tmp[m] = mf[(size_t)f * nz + m] * fac;
            fft_inverse(&o->plan, nz, tmp, o->mfz + (size_t)(2 * f) * NZ);
            for (int m = 0; m < nz; ++m) tmp[m] = fac * o->kz[m] * mf[(size_t)f * nz + m];
            fft_inverse(&o->plan, nz, tmp, o->mfz + (size_t)(2 * f + 1) * NZ);
        }
    } else if (model == ORC_MODEL_HD_FULLU) {
        /* HD_fullU.cpp:153-162: MUx, dzMUx, MUy, dzMUy (kept at profile lines 4-7) */
        for (int f = 0; f < 2; ++f) {
            for (int m = 0; m < nz; ++m) tmp[m] = mf[(size_t)f * nz + m] * fac;
            fft_inverse(&o->plan, nz, tmp, o->mfz + (size_t)(4 + 2 * f) * NZ);
            for (int m = 0; m < nz; ++m) tmp[m] = fac * o->kz[m] * mf[(size_t)f * nz + m];
            fft_inverse(&o->plan, nz, tmp, o->mfz + (size_t)(5 + 2 * f) * NZ);
        }
    } else if (model == ORC_MODEL_MHD_BQLIN) {
        /* MHD_BQlin.cpp:122-127 */
        const cplx *By = mf + nz;
        for (int m = 0; m < nz; ++m) tmp[m] = By[m] * fac;
        fft_inverse(&o->plan, nz, tmp, o->mfz + 0 * NZ);
        for (int m = 0; m < nz; ++m) tmp[m] = fac * o->kz[m] * By[m];
        fft_inverse(&o->plan, nz, tmp, o->mfz + 1 * NZ);
        for (int m = 0; m < nz; ++m) tmp[m] = fac * o->kz2[m] * By[m];
        fft_inverse(&o->plan, nz, tmp, o->mfz + 2 * NZ);
    }
    free(tmp);

#pragma omp parallel num_threads(o->nthreads)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        work_t *w = &o->work[tid];
        for (int k = 0; k < nMF * half; ++k) w->send[k] = 0.0;
        int b0, b1;
        block_of(o->nxy, o->nthreads, tid, &b0, &b1);
        for (int i = b0; i < b1; ++i) {
            switch (model) {
            case ORC_MODEL_MHD_FULLQLIN: rhs_fullqlin_mode(o, w, i, t, dt_lin, lin, out_lin, linop); break;
            case ORC_MODEL_MHD_BQLIN: rhs_bqlin_mode(o, w, i, t, dt_lin, lin, out_lin, linop); break;
            case ORC_MODEL_HD_FULLU: rhs_hd_mode(o, w, i, t, dt_lin, lin, out_lin, linop); break;
            case ORC_MODEL_CONSTANT_DAMPING: rhs_damping_mode(o, i, lin, out_lin, linop); break;
            default: rhs_linear_mode(o, w, i, t, dt_lin, out_lin, linop); break;
            }
        }
    }

    cplx *recv = (cplx *)xcalloc((size_t)nMF * half, sizeof(cplx));
    reduce_and_fill(o, recv);
    free(recv);

    const double q = o->p.q, Om = o->omega_eff, B0z = o->p.B0z;
    const cplx *kz = o->kz;
    if (model == ORC_MODEL_MHD_FULLQLIN && o->p.quasilinear) {
        /* MHD_FullQlin.cpp:411-417 */
        const cplx *M0 = mf, *M1 = mf + nz, *M2 = mf + 2 * nz, *M3 = mf + 3 * nz;
        const cplx *D0 = o->drive, *D1 = o->drive + nz, *D2 = o->drive + 2 * nz, *D3 = o->drive + 3 * nz;
        for (int m = 0; m < nz; ++m) {
            out_mf[nz + m] = -q * M0[m] + B0z * kz[m] * M3[m] + D1[m];
            out_mf[m] = B0z * kz[m] * M2[m] + D0[m];
            out_mf[2 * nz + m] = 2 * Om * M3[m] + B0z * kz[m] * M0[m] + D2[m];
            out_mf[3 * nz + m] = (q - 2.0 * Om) * M2[m] + B0z * kz[m] * M1[m] + D3[m];
        }
    } else if (model == ORC_MODEL_HD_FULLU && o->p.quasilinear) {
        /* HD_fullU.cpp:297-305 */
        for (int m = 0; m < nz; ++m) {
            out_mf[m] = 2 * Om * mf[nz + m] + o->drive[m];
            out_mf[nz + m] = (q - 2.0 * Om) * mf[m] + o->drive[nz + m];
        }
    } else if (model == ORC_MODEL_MHD_BQLIN && o->p.quasilinear) {
        /* MHD_BQlin.cpp:274-276 */
        for (int m = 0; m < nz; ++m) {
            out_mf[nz + m] = -q * mf[m] + o->drive[nz + m];
            out_mf[m] = o->drive[m];
        }
    } else {
        for (int k = 0; k < nMF * nz; ++k) out_mf[k] = 0.0;
    }
}

/* Model::linearOPs_Init; the integrator constructors call it with their
 * operator arrays (RK3CN.cpp:57 -> linearOp_fluct_old_, RK3CN_Elin.cpp:257). */
void orc_linear_ops_init(orc *o, double t0, double *linop_out, double *linop_mf_out)
{
    const int nz = o->nz, nxy = o->nxy, nMF = o->nMF;
    double *Lf = o->p.integrator == ORC_INT_RK3CN ? o->linop_old : o->linop;
    work_t *w = &o->work[0];
    for (int i = 0; i < nxy; ++i) {
        kmode k;
        double *L0 = Lf + (size_t)i * 4 * nz;
        if (o->p.model == ORC_MODEL_CONSTANT_DAMPING) {
            for (int m = 0; m < 4 * nz; ++m) L0[m] = 0.0;
            continue;
        }
        assign_laplacians(o, i, t0, 0, w, &k);
        for (int m = 0; m < nz; ++m) {
            if (o->p.model == ORC_MODEL_HD_FULLU) {
                /* HD_fullU.cpp:314-318 */
                L0[m] = o->p.nu * w->lapF[m];
                L0[nz + m] = L0[m];
                L0[2 * nz + m] = 0.0;
                L0[3 * nz + m] = 0.0;
            } else if (o->p.model == ORC_MODEL_MHD_BQLIN) {
                /* MHD_BQlin.cpp:289-295 */
                L0[m] = o->p.nu * w->lapF[m];
                L0[nz + m] = L0[m];
                L0[2 * nz + m] = (o->p.eta / o->p.nu) * L0[m];
                L0[3 * nz + m] = L0[2 * nz + m];
            } else {
                /* MHD_FullQlin.cpp:432-440 */
                double lapF = -pow(fabs(w->lapF[m]), (double)(o->p.viscosity_order / 2));
                L0[m] = o->p.nu * lapF;
                L0[nz + m] = L0[m];
                L0[2 * nz + m] = o->p.eta * lapF;
                L0[3 * nz + m] = L0[2 * nz + m];
            }
        }
    }
    for (int f = 0; f < nMF; ++f)
        for (int m = 0; m < nz; ++m) {
            double v = 0.0;
            if (o->p.quasilinear && (o->p.model == ORC_MODEL_MHD_FULLQLIN || o->p.model == ORC_MODEL_LINEAR_ONLY))
                v = (f < 2 ? o->p.eta : o->p.nu) * o->kz2[m]; /* MHD_FullQlin.cpp:444-447 */
            else if (o->p.quasilinear && o->p.model == ORC_MODEL_MHD_BQLIN)
                v = o->p.eta * o->kz2[m]; /* MHD_BQlin.cpp:298-299 */
            else if (o->p.quasilinear && o->p.model == ORC_MODEL_HD_FULLU)
                v = o->p.nu * o->kz2[m]; /* HD_fullU.cpp:321-323 */
            o->linop_mf[(size_t)f * nz + m] = v;
        }
    if (linop_out) lin_out(o, linop_out, Lf, sizeof(double));
    if (linop_mf_out) memcpy(linop_mf_out, o->linop_mf, sizeof(double) * (size_t)nMF * nz);
}

void orc_rhs(orc *o, double t, double dt_lin, double *rhs_lin, double *rhs_mf, double *linop)
{
    model_rhs(o, t, dt_lin, o->lin, o->mf, o->rhs1_lin, o->rhs1_mf, o->linop);
    if (rhs_lin) lin_out(o, rhs_lin, o->rhs1_lin, sizeof(cplx));
    if (rhs_mf) memcpy(rhs_mf, o->rhs1_mf, sizeof(cplx) * (size_t)o->nMF * o->nz);
    if (linop) lin_out(o, linop, o->linop, sizeof(double));
}

/* ------------------------------------------------------------------------ */
/* Integrators                                                               */
/* ------------------------------------------------------------------------ */

/* kmax of the CFL condition (MHD_FullQlin.cpp:92-97, MHD_BQlin.cpp:70-75) */
static double cfl_kmax(const orc *o)
{
    const double dealiasfac[3] = {2.0, 2.0, 3.0};
    const int N[3] = {o->p.NX, o->p.NY, o->p.NZ};
    double kmax = 0;
    for (int i = 0; i < 3; ++i)
        if (kmax < 2 * QL_PI / o->p.L[i] * (N[i] / dealiasfac[i])) kmax = 2 * QL_PI / o->p.L[i] * (N[i] / dealiasfac[i]);
    return kmax;
}

/* sqrt(max |z|^2) of a complex profile (Eigen's abs2().maxCoeff()) */
static double max_abs(const cplx *z, int n)
{
    double m = 0.0;
    for (int j = 0; j < n; ++j) {
        double a = creal(z[j]) * creal(z[j]) + cimag(z[j]) * cimag(z[j]);
        if (a > m) m = a;
    }
    return sqrt(m);
}

double orc_calculate_cfl(orc *o)
{
    const int NZ = o->NZ;
    const double kmax = cfl_kmax(o);
    if (o->p.model == ORC_MODEL_MHD_FULLQLIN) {
        /* MHD_FullQlin.cpp:773-778: profiles MBx_, MBy_, MUx_, MUy_ of the last rhs */
        double Bxmax = max_abs(o->mfz + 0 * NZ, NZ), Bymax = max_abs(o->mfz + 2 * NZ, NZ);
        double Uxmax = max_abs(o->mfz + 4 * NZ, NZ), Uymax = max_abs(o->mfz + 6 * NZ, NZ);
        return kmax * (Bymax + Bxmax + Uxmax + Uymax) + o->p.q;
    }
    if (o->p.model == ORC_MODEL_HD_FULLU) {
        /* HD_fullU.cpp:604-628: profiles and drives of the last rhs, Ux of the
         * current state */
        double Uxmax = max_abs(o->mfz + 4 * NZ, NZ), Uymax = max_abs(o->mfz + 6 * NZ, NZ);
        double MFcont = 0.;
        if (o->p.quasilinear) {
            double e = 0, e2 = 0;
            for (int j = 0; j < NZ; ++j) {
                cplx a = o->mfz[4 * NZ + j], b = o->mfz[6 * NZ + j];
                e += creal(a) * creal(a) + cimag(a) * cimag(a);
                e2 += creal(b) * creal(b) + cimag(b) * cimag(b);
            }
            if ((e + e2) * o->p.L[2] > 1e-6) {
                double dx = 0, dy = 0, ux = 0;
                for (int m = 0; m < o->nz; ++m) {
                    dx += cabs(o->drive[m] * o->kz2[m]);
                    dy += cabs(o->drive[o->nz + m] * o->kz2[m]);
                    ux += cabs(o->mf[m] * o->kz2[m]);
                }
                dx /= o->nz; dy /= o->nz; ux /= o->nz;
                MFcont = dx / ux + dy / ux;
            }
        }
        return kmax * (Uxmax + Uymax) + o->p.q + MFcont;
    }
    if (o->p.model == ORC_MODEL_MHD_BQLIN) {
        /* MHD_BQlin.cpp:544-550: By_ of the last rhs, Bx of the current state */
        double Bymax = max_abs(o->mfz, NZ);
        cplx *tmp = (cplx *)xcalloc(o->nz, sizeof(cplx)), *bx = (cplx *)xcalloc(NZ, sizeof(cplx));
        for (int m = 0; m < o->nz; ++m) tmp[m] = o->mf[m] * o->fac1D;
        fft_inverse(&o->plan, o->nz, tmp, bx);
        double Bxmax = max_abs(bx, NZ);
        free(tmp);
        free(bx);
        return kmax * Bymax + kmax * Bxmax + o->p.q;
    }
    return o->p.q;
}

/* the step's dt: fixed, or CFL / Calculate_CFL capped at -dt
 * (RK3CN.cpp:20-30, 79-83; RK3CN_Elin.cpp:76-80) */
static double step_dt(orc *o)
{
    if (o->p.dt > 0) return o->p.dt;
    double dt = o->p.CFL / orc_calculate_cfl(o);
    if (dt > -o->p.dt) dt = -o->p.dt;
    return dt;
}

/* the time step of the following steps: dt > 0 fixed, dt < 0 variable of at
 * most -dt (a step of a chosen size, e.g. to replay another run's steps) */
void orc_set_dt(orc *o, double dt) { o->p.dt = dt; }

void orc_check_sums(const orc *o, double *out)
{
    const int nz = o->nz, NZ = o->NZ;
    const cplx *Bx = o->mf, *By = o->mf + nz;
    cplx *r = (cplx *)xcalloc(NZ, sizeof(cplx));
    fft_inverse(&o->plan, nz, By, r);
    double sre = 0, sim = 0, mby = 0, mbx = 0, sby = 0;
    for (int j = 0; j < NZ; ++j) {
        sre += fabs(creal(r[j]));
        sim += fabs(cimag(r[j]));
    }
    for (int m = 0; m < nz; ++m) {
        if (cabs(By[m]) > mby) mby = cabs(By[m]);
        if (cabs(Bx[m]) > mbx) mbx = cabs(Bx[m]);
        sby += cabs(By[m]);
    }
    free(r);
    out[0] = sre; out[1] = sim; out[2] = mby; out[3] = mbx; out[4] = sby;
}

int orc_check_solution(const orc *o)
{
    double s[5];
    orc_check_sums(o, s);
    if (s[0] / s[1] < 1e10 && s[0] > 1e-2) return 1;  /* solution.cpp:174 */
    if (s[2] > 1e20 || s[3] > 1e20) return 2;          /* solution.cpp:184 */
    if (!isfinite(s[4])) return 3;                      /* solution.cpp:194 */
    return 0;
}

/* RK3CN_Elin::Step (Integrators/RK3CN_Elin.cpp:275-389) */
static double step_elin(orc *o, double t)
{
    const double dt = step_dt(o);
    const int nz = o->nz, nMF = o->nMF;
    const long nl = (long)o->nxy * 4 * nz;
    double linC, NLC1, NLC2, lin_dt;
    cplx *u = o->lin, *N1 = o->rhs1_lin, *N2 = o->rhs2_lin;
    double *L = o->linop;

    /* substep 1, :290-313 */
    linC = dt * (RK_A0); NLC1 = RK_A0 * dt; lin_dt = dt * RK_P1;
    model_rhs(o, t, lin_dt, o->lin, o->mf, N1, o->rhs1_mf, L);
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (long k = 0; k < nl; ++k) u[k] = exp(linC * L[k]) * (NLC1 * N1[k] + u[k]);
    for (int k = 0; k < nMF * nz; ++k)
        o->mf[k] = exp(linC * o->linop_mf[k]) * (NLC1 * o->rhs1_mf[k] + o->mf[k]);

    /* substep 2, :323-345 */
    linC = dt * (RK_A1 + RK_B1); NLC1 = RK_A1 * dt; NLC2 = RK_B1 * dt; lin_dt = dt * (RK_P2 - RK_P1);
    model_rhs(o, t + dt * RK_P1, lin_dt, o->lin, o->mf, N2, o->rhs2_mf, L);
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (long k = 0; k < nl; ++k) u[k] = exp(linC * L[k]) * (NLC1 * N2[k] + NLC2 * N1[k] + u[k]);
    for (int k = 0; k < nMF * nz; ++k)
        o->mf[k] = exp(linC * o->linop_mf[k]) * (NLC1 * o->rhs2_mf[k] + NLC2 * o->rhs1_mf[k] + o->mf[k]);

    /* substep 3, :356-378 */
    linC = dt * (RK_A2 + RK_B2); NLC1 = RK_A2 * dt; NLC2 = RK_B2 * dt; lin_dt = dt * (1.0 - RK_P2);
    model_rhs(o, t + dt * RK_P2, lin_dt, o->lin, o->mf, N1, o->rhs1_mf, L);
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (long k = 0; k < nl; ++k) u[k] = exp(linC * L[k]) * (NLC1 * N1[k] + NLC2 * N2[k] + u[k]);
    for (int k = 0; k < nMF * nz; ++k)
        o->mf[k] = exp(linC * o->linop_mf[k]) * (NLC1 * o->rhs1_mf[k] + NLC2 * o->rhs2_mf[k] + o->mf[k]);
    return dt;
}

/* RK3CN::Step (Integrators/RK3CN.cpp:77-201) */
static double step_cn(orc *o, double t)
{
    const double dt = step_dt(o);
    const int nz = o->nz, nMF = o->nMF;
    const long nl = (long)o->nxy * 4 * nz;
    double linC, NLC1, NLC2, lin_dt;
    cplx *u = o->lin, *N1 = o->rhs1_lin, *N2 = o->rhs2_lin;
    double *L = o->linop, *Lo = o->linop_old;

    /* substep 1, :92-122 */
    linC = dt / 2 * (RK_A0); NLC1 = RK_A0 * dt; lin_dt = dt * RK_P1;
    model_rhs(o, t, lin_dt, o->lin, o->mf, N1, o->rhs1_mf, L);
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (long k = 0; k < nl; ++k) {
        double denom = 1.0 / (1.0 - linC * L[k]);
        u[k] = NLC1 * denom * N1[k] + (1 + linC * Lo[k]) * denom * u[k];
    }
    for (int k = 0; k < nMF * nz; ++k) {
        double denom = 1.0 / (1.0 - linC * o->linop_mf[k]);
        o->mf[k] = NLC1 * denom * o->rhs1_mf[k] + (1 + linC * o->linop_mf[k]) * denom * o->mf[k];
    }
    memcpy(Lo, L, sizeof(double) * (size_t)nl);

    /* substep 2, :130-157 */
    linC = dt / 2 * (RK_A1 + RK_B1); NLC1 = RK_A1 * dt; NLC2 = RK_B1 * dt; lin_dt = dt * (RK_P2 - RK_P1);
    model_rhs(o, t + dt * RK_P1, lin_dt, o->lin, o->mf, N2, o->rhs2_mf, L);
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (long k = 0; k < nl; ++k) {
        double denom = 1.0 / (1.0 - linC * L[k]);
        u[k] = NLC1 * denom * N2[k] + NLC2 * denom * N1[k] + (1 + linC * Lo[k]) * denom * u[k];
    }
    for (int k = 0; k < nMF * nz; ++k) {
        double denom = 1.0 / (1.0 - linC * o->linop_mf[k]);
        o->mf[k] = NLC1 * denom * o->rhs2_mf[k] + NLC2 * denom * o->rhs1_mf[k] +
                   (1 + linC * o->linop_mf[k]) * denom * o->mf[k];
    }
    memcpy(Lo, L, sizeof(double) * (size_t)nl);

    /* substep 3, :166-193 */
    linC = dt / 2 * (RK_A2 + RK_B2); NLC1 = RK_A2 * dt; NLC2 = RK_B2 * dt; lin_dt = dt * (1.0 - RK_P2);
    model_rhs(o, t + dt * RK_P2, lin_dt, o->lin, o->mf, N1, o->rhs1_mf, L);
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (long k = 0; k < nl; ++k) {
        double denom = 1.0 / (1.0 - linC * L[k]);
        u[k] = NLC1 * denom * N1[k] + NLC2 * denom * N2[k] + (1 + linC * Lo[k]) * denom * u[k];
    }
    for (int k = 0; k < nMF * nz; ++k) {
        double denom = 1.0 / (1.0 - linC * o->linop_mf[k]);
        o->mf[k] = NLC1 * denom * o->rhs1_mf[k] + NLC2 * denom * o->rhs2_mf[k] +
                   (1 + linC * o->linop_mf[k]) * denom * o->mf[k];
    }
    memcpy(Lo, L, sizeof(double) * (size_t)nl);
    return dt;
}

double orc_step(orc *o, double t)
{
    return o->p.integrator == ORC_INT_RK3CN ? step_cn(o, t) : step_elin(o, t);
}

/* MHD_FullQlin::ShearingBox_Remap (Models/MHD_FullQlin.cpp:464-481); the
 * other models' copies are identical (MHD_BQlin.cpp:314-331). */
void orc_remap(orc *o, double qt)
{
    double kxLfac = 2 * QL_PI / o->p.L[0];
    int nx = o->p.NX - 1;
    double kxt;
    for (int i = 0; i < o->nxy; ++i) {
        kxt = o->kx0[i] + qt * o->ky[i];
        if (kxt > nx / 2.0 * kxLfac) {
            o->kx0[i] = o->kx0[i] - nx * kxLfac;
            memset(o->lin + (size_t)i * 4 * o->nz, 0, sizeof(cplx) * (size_t)4 * o->nz);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Driving noise                                                             */
/* ------------------------------------------------------------------------ */

/* Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC 2011): the counter-based generator that stands in for
 * the reference's clock-seeded mt19937 */
void orc_philox4x64(const uint64_t *ctr, const uint64_t *key, uint64_t *out)
{
    const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
    const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
    uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        unsigned __int128 p0 = (unsigned __int128)M0 * c0, p1 = (unsigned __int128)M1 * c2;
        uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0, hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Box-Muller: two standard normals from two 64-bit words, u1 in (0, 1) */
static void box_muller(uint64_t a, uint64_t b, double *z1, double *z2)
{
    double u1 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
    double u2 = (double)(b >> 11) * 0x1.0p-53;
    double r = sqrt(-2.0 * log(u1));
    double th = (2.0 * QL_PI) * u2;
    *z1 = r * cos(th);
    *z2 = r * sin(th);
}

/* the normals of entry (t, mode i, slot jj): z[2f], z[2f+1] for field f */
static void noise_normals(const orc *o, double t, int i, int jj, double *z)
{
    const uint64_t key[2] = {o->noise_seed, 0x514C4E4F49534500ull};
    uint64_t tb;
    memcpy(&tb, &t, sizeof tb);
    for (int k = 0; k < 2; ++k) {
        uint64_t c[4] = {tb, (uint64_t)i, (uint64_t)jj | ((uint64_t)k << 32), 0}, w[4];
        orc_philox4x64(c, key, w);
        box_muller(w[0], w[1], &z[4 * k + 0], &z[4 * k + 1]);
        box_muller(w[2], w[3], &z[4 * k + 2], &z[4 * k + 3]);
    }
}

void orc_noise_normals(const orc *o, double t, int i, int jj, double *z8) { noise_normals(o, t, i, jj, z8); }

void orc_set_noise(orc *o, double f_noise, double range_low, double range_high, int drive_only_velocity,
                   int drive_only_magnetic, uint64_t seed)
{
    /* ChangeNoiseRange (MHD_FullQlin.h:46-50) plus the input flags */
    o->f_noise = f_noise;
    o->noise_range[0] = range_low * range_low;
    o->noise_range[1] = range_high * range_high;
    /* only MHD_FullQlin reads the flags (MHD_FullQlin.cpp:21-22) */
    o->drive_only_velocity = o->p.model == ORC_MODEL_MHD_FULLQLIN ? drive_only_velocity : 0;
    o->drive_only_magnetic = o->p.model == ORC_MODEL_MHD_FULLQLIN ? drive_only_magnetic : 0;
    o->noise_seed = seed;
}

/* noise of one mode: slots jj0..nz-1; for a ky = 0 pair the conjugate goes to
 * slot nz - jj of mode ip (MHD_FullQlin.cpp:962-1046) */
static void noise_mode(orc *o, int i, int ip, int jj0, double t, double dt, work_t *w)
{
    const int nz = o->nz;
    kmode k;
    assign_laplacians(o, i, t, 0, w, &k);
    double noise_multfac = dt * o->totalN2 * o->mult_noise_fac;
    const double sigma = o->f_noise / sqrt(2.0); /* normal_distribution(0, f/sqrt 2), :47 */
    const int hd = o->p.model == ORC_MODEL_HD_FULLU; /* u and zeta only (HD_fullU.cpp:635-852) */
    const int on[4] = {!o->drive_only_magnetic, !o->drive_only_magnetic, !o->drive_only_velocity && !hd,
                       !o->drive_only_velocity && !hd};
    for (int jj = jj0; jj < nz; ++jj) {
        double lapF = w->lapF[jj], lap2 = w->lap2t[jj];
        if (!(lapF > -o->noise_range[1] && lapF < -o->noise_range[0])) continue; /* :806 */
        double multU = sqrt(noise_multfac * lap2 / lapF);                        /* :808-811 */
        double multZ = sqrt(-noise_multfac * lap2);
        const double mult[4] = {multU, multZ, multU, multZ};
        double z[8];
        noise_normals(o, t, i, jj, z);
        for (int f = 0; f < 4; ++f) {
            if (!on[f]) continue;
            cplx n = mult[f] * mk(sigma * z[2 * f], sigma * z[2 * f + 1]);
            o->lin[((size_t)i * 4 + f) * nz + jj] += n;
            if (ip >= 0) o->lin[((size_t)ip * 4 + f) * nz + (nz - jj)] += conj(n);
        }
    }
}

/* Model::DrivingNoise (MHD_FullQlin.cpp:786-1053; MHD_BQlin.cpp:558-740) */
void orc_driving_noise(orc *o, double t, double dt)
{
    if (o->f_noise == 0.0) return;
    work_t *w = &o->work[0];
    for (int i = 0; i < o->nxy; ++i)
        if (o->ky_index[i] != 0) noise_mode(o, i, -1, 0, t, dt, w); /* :796-845 */
    if (!o->dont_drive_ky0)
        for (int m = 0; m < o->nx - 1; ++m) /* the no-MPI branch, :962-1046 */
            noise_mode(o, o->ky0nkx[m], o->ky0pkx[m], 1, t, dt, w);
}

/* main.cpp:108-123 (no output): Step, DrivingNoise, t += dt, remap(q t) */
double orc_steps(orc *o, int nsteps, double t)
{
    for (int n = 0; n < nsteps; ++n) {
        double dt = orc_step(o, t);
        if (o->f_noise != 0.0) orc_driving_noise(o, t, dt);
        t += dt;
        if (o->p.remap) orc_remap(o, o->p.q * t);
    }
    return t;
}

/* Calc_Energy_AM_Diss (MHD_FullQlin.cpp:487-717, MHD_BQlin.cpp:337-488) */
void orc_diagnostics(orc *o, double t, double *out)
{
    const int nz = o->nz, nxy = o->nxy;
    const double nu = o->p.nu, eta = o->p.eta;
    for (int k = 0; k < 18; ++k) out[k] = 0.0;
    if (o->p.model == ORC_MODEL_CONSTANT_DAMPING || o->p.model == ORC_MODEL_LINEAR_ONLY) return;
#pragma omp parallel num_threads(o->nthreads)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        work_t *w = &o->work[tid];
        double eu = 0, eb = 0, amu = 0, amb = 0, du = 0, db = 0;
        int b0, b1;
        block_of(nxy, o->nthreads, tid, &b0, &b1);
        for (int i = b0; i < b1; ++i) {
            kmode k;
            assign_laplacians(o, i, t, 1, w, &k);
            int mult_fac = 2;
            if (k.ky == 0.0) mult_fac = 1;
            for (int m = 0; m < nz; ++m) w->lap2t[m] = w->lapF[m] * w->ilap2t[m];
            const cplx *ux = o->lin + ((size_t)i * 4 + 0) * nz, *ze = ux + nz, *bx = ux + 2 * nz, *et = ux + 3 * nz;
            double s_eu = 0, s_eb = 0, s_amu = 0, s_amb = 0, s_du = 0, s_db = 0;
            for (int m = 0; m < nz; ++m) {
                double a_u = w->lap2t[m] * (creal(ux[m]) * creal(ux[m]) + cimag(ux[m]) * cimag(ux[m])) -
                             w->ilap2t[m] * (creal(ze[m]) * creal(ze[m]) + cimag(ze[m]) * cimag(ze[m]));
                double a_b = w->lap2t[m] * (creal(bx[m]) * creal(bx[m]) + cimag(bx[m]) * cimag(bx[m])) -
                             w->ilap2t[m] * (creal(et[m]) * creal(et[m]) + cimag(et[m]) * cimag(et[m]));
                s_eu += a_u;
                s_eb += a_b;
                cplx fy = ((-k.kyc * k.kxc) * ux[m] + o->kz[m] * ze[m]) * w->ilap2t[m];
                cplx fz = ((-k.kyc * k.kxc) * bx[m] + o->kz[m] * et[m]) * w->ilap2t[m];
                s_amu += creal(ux[m] * conj(fy));
                s_amb += creal(bx[m] * conj(fz));
                s_du += -w->lapF[m] * a_u;
                s_db += -w->lapF[m] * a_b;
            }
            eu += mult_fac * s_eu;
            eb += mult_fac * s_eb;
            amu += mult_fac * s_amu;
            amb += mult_fac * s_amb;
            du += (mult_fac * nu) * s_du;
            db += (mult_fac * eta) * s_db;
        }
        w->diag[0] = eu; w->diag[1] = eb; w->diag[2] = amu; w->diag[3] = amb; w->diag[4] = du; w->diag[5] = db;
    }
    double rb[6] = {0, 0, 0, 0, 0, 0};
    for (int tt = 0; tt < o->nthreads; ++tt)
        for (int k = 0; k < 6; ++k) rb[k] += o->work[tt].diag[k];
    double divfac = o->p.L[2] / o->totalN2;
    double energy_u = rb[0] * divfac, energy_b = rb[1] * divfac, AM_u = rb[2] * divfac, AM_b = rb[3] * divfac;
    double diss_u = rb[4] * divfac, diss_b = rb[5] * divfac;
    double energy_MU = 0, energy_MB = 0, AM_MU = 0, AM_MB = 0, diss_MU = 0, diss_MB = 0;
    double divfavMF = o->p.L[2] / (o->NZ * o->NZ);
    const cplx *M = o->mf;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, am_b = 0, am_u = 0, d0 = 0, d1 = 0, d2 = 0, d3 = 0;
    for (int m = 0; m < nz; ++m) {
        double a0 = creal(M[m]) * creal(M[m]) + cimag(M[m]) * cimag(M[m]);
        double a1 = creal(M[nz + m]) * creal(M[nz + m]) + cimag(M[nz + m]) * cimag(M[nz + m]);
        s0 += a0; s1 += a1;
        am_b += creal(M[m] * conj(M[nz + m]));
        d0 += o->kz2[m] * a0; d1 += o->kz2[m] * a1;
        if (o->nMF == 4) {
            double a2 = creal(M[2 * nz + m]) * creal(M[2 * nz + m]) + cimag(M[2 * nz + m]) * cimag(M[2 * nz + m]);
            double a3 = creal(M[3 * nz + m]) * creal(M[3 * nz + m]) + cimag(M[3 * nz + m]) * cimag(M[3 * nz + m]);
            s2 += a2; s3 += a3;
            am_u += creal(M[2 * nz + m] * conj(M[3 * nz + m]));
            d2 += o->kz2[m] * a2; d3 += o->kz2[m] * a3;
        }
    }
    energy_MB = (s0 + s1) * divfavMF;
    AM_MB = divfavMF * am_b;
    diss_MB = (divfavMF * eta) * (-d0 - d1);
    if (o->p.model == ORC_MODEL_HD_FULLU) {
        /* HD_fullU.cpp:455-460: the two mean fields are Ux, Uy */
        energy_MU = energy_MB;
        AM_MU = AM_MB;
        diss_MU = (divfavMF * nu) * (-d0 - d1);
        energy_MB = AM_MB = diss_MB = 0;
    }
    if (o->nMF == 4) {
        energy_MU = (s2 + s3) * divfavMF;
        AM_MU = divfavMF * am_u;
        diss_MU = (divfavMF * nu) * (-d2 - d3);
    }
    out[0] = energy_MU / 2; out[1] = energy_MB / 2; out[2] = energy_u / 2; out[3] = energy_b / 2;
    out[4] = AM_MU; out[5] = AM_MB; out[6] = AM_u; out[7] = AM_b;
    out[8] = diss_MU; out[9] = diss_MB; out[10] = diss_u; out[11] = diss_b;
    double kzfac = 2 * QL_PI / o->p.L[2] * (o->NZ / 3);
    out[16] = pow(nu * nu * nu / (diss_MU + diss_u), 0.25) * kzfac;
    out[17] = o->p.model == ORC_MODEL_HD_FULLU ? 0.0 : pow(eta * eta * eta / (diss_MB + diss_b), 0.25) * kzfac;
    if (o->p.model == ORC_MODEL_MHD_FULLQLIN) {
        /* :669-676 */
        for (int f = 0; f < 4; ++f) {
            double num = 0, den = 0;
            for (int m = 0; m < nz; ++m) {
                cplx d = o->drive[(size_t)f * nz + m], v = M[(size_t)f * nz + m];
                num += creal(d * conj(v));
                den += creal(v) * creal(v) + cimag(v) * cimag(v);
            }
            out[12 + f] = num / sqrt(den);
        }
    }
}

void orc_get_drive(const orc *o, double *drive)
{
    memcpy(drive, o->drive, sizeof(cplx) * (size_t)o->nMF * o->nz);
}

/* TimeVariables::Save_Mean_Fields (Auxiliary/TimeVariables.cpp:202-228) */
void orc_mf_real(const orc *o, double *out)
{
    cplx *buf = (cplx *)xcalloc(o->NZ, sizeof(cplx));
    for (int f = 0; f < o->nMF; ++f) {
        fft_inverse(&o->plan, o->nz, o->mf + (size_t)f * o->nz, buf);
        for (int j = 0; j < o->NZ; ++j) out[(size_t)f * o->NZ + j] = o->fac1D * creal(buf[j]);
    }
    free(buf);
}

void orc_fft_inverse(int NZ, int nz, const double *in, double *out)
{
    fftplan f;
    plan_init(&f, NZ);
    fft_inverse(&f, nz, (const cplx *)in, (cplx *)out);
    plan_free(&f);
}

void orc_fft_forward(int NZ, int nz, const double *in, double *out)
{
    fftplan f;
    plan_init(&f, NZ);
    cplx *buf = (cplx *)xcalloc(NZ, sizeof(cplx));
    memcpy(buf, in, sizeof(cplx) * (size_t)NZ);
    fft_forward(&f, nz, buf, (cplx *)out);
    free(buf);
    plan_free(&f);
}

void orc_fft_plain(int N, int sign, const double *in, double *out)
{
    fftplan f;
    plan_init(&f, N);
    memcpy(out, in, sizeof(cplx) * (size_t)N);
    fft_exec(&f, (cplx *)out, sign);
    plan_free(&f);
}
