/*
 * ql_oracle.h -- CPU restatement of the reference quasilinear (QL) time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in ql_dns_amd/ links or calls it.
 *
 * It restates, in plain C99 double precision, the reference solver's
 *   Models/MHD_FullQlin.cpp   (rhs :171-427, linearOPs_Init :429-456,
 *                              ShearingBox_Remap :464-481,
 *                              Calc_Energy_AM_Diss :487-717,
 *                              assign_laplacians_ :722-738)
 *   Models/MHD_BQlin.cpp      (rhs :117-284, linearOPs_Init :286-306,
 *                              Calc_Energy_AM_Diss :337-488)
 *   Models/HD_fullU.cpp       (rhs :145-309, linearOPs_Init :311-327,
 *                              Calc_Energy_AM_Diss :360-550, CFL :604-628)
 *   Models/ConstantDamping.cpp (rhs :73-111)            -- integrator pin
 *   Integrators/RK3CN.cpp      (Step :77-201)
 *   Integrators/RK3CN_Elin.cpp (Step :74-188)
 *   Auxiliary/fftwPlans.h      (forward/inverse/pad/dealias :33-97)
 *   Auxiliary/Kdata.cpp        (mode order :28-46, kz :72-80, lap2 :100-120)
 *   Auxiliary/TimeVariables.cpp (Save_Mean_Fields :202-228)
 *   main.cpp                   (time loop :108-123)
 * with its own radix-2 FFT (FFTW is not available) and OpenMP threads over
 * contiguous mode blocks standing in for MPIdata::Split_NXY_Grid
 * (Auxiliary/MPIdata.cpp:17-41); per-thread partial sums are added in thread
 * order where the reference calls MPI Allreduce (MHD_FullQlin.cpp:386).
 *
 * Layouts (all little-endian doubles, complex = {re, im} pairs):
 *   lin   [nxy][nLin][nz] complex   -- solution::pLin(i, Vnum)
 *   mf    [nMF][nz] complex         -- solution::pMF(i)
 *   linop [nxy][nLin][nz] double    -- linOpFluct[i][Vnum]
 *   kx0   [nxy] double              -- imaginary part of Kdata::kx (remapped)
 */
#ifndef QL_ORACLE_H
#define QL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_MODEL_MHD_FULLQLIN = 0, ORC_MODEL_MHD_BQLIN = 1,
       ORC_MODEL_CONSTANT_DAMPING = 2, ORC_MODEL_LINEAR_ONLY = 3, ORC_MODEL_HD_FULLU = 4 };
enum { ORC_INT_RK3CN = 0, ORC_INT_RK3CN_ELIN = 1 };

typedef struct orc_params {
    int model;            /* ORC_MODEL_*                                   */
    int integrator;       /* ORC_INT_*                                     */
    int NX, NY, NZ;       /* Inputs::NXY[0], NXY[1], NZ                    */
    double L[3];          /* Lx, Ly, Lz                                    */
    double q, omega, nu, eta, B0z;
    int viscosity_order;
    int quasilinear;      /* QuasiLinear?                                  */
    int remap;            /* Remap?                                        */
    double dt;            /* fixed time step (> 0); < 0: variable steps of
                             at most -dt (RK3CN.cpp:20-30)                 */
    int nthreads;         /* OpenMP threads (<= 0: all)                    */
    double CFL;           /* Inputs::CFL (Input_parameters.cpp:77), used
                             when dt < 0                                   */
} orc_params;

typedef struct orc orc;

orc *orc_create(const orc_params *p);
void orc_destroy(orc *o);
/* nxy, nz, nMF, nLin, half, NZ */
void orc_dims(const orc *o, int *dims6);
void orc_set_state(orc *o, const double *lin, const double *mf, const double *kx0);
void orc_get_state(const orc *o, double *lin, double *mf, double *kx0);
/* Wavenumber tables: kx0[nxy], ky[nxy], kx_index[nxy], ky_index[nxy], kz[nz] */
void orc_tables(const orc *o, double *kx0, double *ky, int *kx_index, int *ky_index, double *kz);

/* Model::linearOPs_Init(t0, linOpFluct, linOpMF); also primes the integrator */
void orc_linear_ops_init(orc *o, double t0, double *linop, double *linop_mf);
/* Model::rhs(t, dt_lin, SolIn = current state, SolOut, linOpFluct) */
void orc_rhs(orc *o, double t, double dt_lin, double *rhs_lin, double *rhs_mf, double *linop);
/* Model::Calculate_CFL (MHD_FullQlin.cpp:769-780, MHD_BQlin.cpp:540-552):
 * CFL/dt from the profiles of the last rhs (and, for MHD_BQlin, Bx of the
 * current mean field) */
double orc_calculate_cfl(orc *o);
/* solution::Check_Solution (solution.cpp:169-203) reductions:
 * out[0] = sum |Re ifft(pad(By))|, out[1] = sum |Im ifft(pad(By))|,
 * out[2] = max |By|, out[3] = max |Bx| (spectral), out[4] = sum |By| */
void orc_check_sums(const orc *o, double *out5);
/* the verdict of Check_Solution on those sums: 0 passes, 1 imaginary mean
 * field, 2 very large, 3 NaN or Inf */
int orc_check_solution(const orc *o);
/* Model::ChangeNoiseRange with the noise inputs (Input_parameters.cpp:107-111)
 * and the generator key; f_noise = 0 (the default) turns the noise off */
void orc_set_noise(orc *o, double f_noise, double range_low, double range_high, int drive_only_velocity,
                   int drive_only_magnetic, uint64_t seed);
/* Model::DrivingNoise(t, dt, sol) (MHD_FullQlin.cpp:786-1053, MHD_BQlin.cpp:
 * 558-740) with normals from Philox4x64-10 counted by (t, mode, kz slot) in
 * place of the reference's clock-seeded mt19937 */
void orc_driving_noise(orc *o, double t, double dt);
/* the eight standard normals of entry (t, mode i, slot jj): re, im of fields
 * ux, zeta, bx, eta before the multipliers and f_noise / sqrt(2) */
void orc_noise_normals(const orc *o, double t, int i, int jj, double *z8);
/* the generator: out = Philox4x64-10(ctr[4], key[2]) */
void orc_philox4x64(const uint64_t *ctr, const uint64_t *key, uint64_t *out);
/* Integrator::Step(t, sol); returns dt */
double orc_step(orc *o, double t);
/* the time step of the following steps (orc_params.dt convention) */
void orc_set_dt(orc *o, double dt);
/* main.cpp:108-123 without output: Step, DrivingNoise (if f_noise != 0),
 * t += dt, remap(q t); returns t */
double orc_steps(orc *o, int nsteps, double t);
/* Model::ShearingBox_Remap(qt, sol) */
void orc_remap(orc *o, double qt);
/* Calc_Energy_AM_Diss with every save flag on:
 * out[0..3] energy (MU, MB, u, b), out[4..7] AM, out[8..11] dissipation,
 * out[12..15] Reynolds projections, out[16..17] etaK*kmax */
void orc_diagnostics(orc *o, double t, double *out18);
/* EMF / stress drives of the last rhs call: [nMF][nz] complex */
void orc_get_drive(const orc *o, double *drive);
/* TimeVariables::Save_Mean_Fields: fac1D * Re(ifft(pad(MF))) -> [nMF][NZ] */
void orc_mf_real(const orc *o, double *out);

/* Bare dealiased transforms of fftwPlans (unnormalised):
 * inverse: in[nz] complex -> out[NZ] complex (pad, e^{+i}),
 * forward: in[NZ] complex -> out[nz] complex (e^{-i}, truncate). */
void orc_fft_inverse(int NZ, int nz, const double *in, double *out);
void orc_fft_forward(int NZ, int nz, const double *in, double *out);
/* plain unnormalised length-N transform, sign -1 (forward) or +1 */
void orc_fft_plain(int N, int sign, const double *in, double *out);

#ifdef __cplusplus
}
#endif
#endif
