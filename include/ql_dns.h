/*
 * ql_dns.h -- C ABI of the MI355X (gfx950) quasilinear shearing-box time step.
 *
 * One context (ql_ctx) holds one run: its parameters, its device-resident
 * state and its HIP stream.  A run set (ql_batch) holds many runs of one model,
 * integrator, precision and dt, stepped together; a context is a set of one
 * (ql_batch_of).  Host buffers passed in are copied and never retained.
 * Every entry point returns QL_OK (0) or a negative QL_ERR_* code;
 * ql_last_error() gives the message of the last failure on the calling
 * thread.  Calls on one context are not thread-safe (as the reference's
 * objects are not); separate contexts are independent.
 *
 * Host layouts (little-endian doubles, complex = {re, im}):
 *   lin    [nxy][nLin][nz] complex  -- solution::pLin(i, Vnum), solution.h:31
 *   mf     [nMF][nz] complex        -- solution::pMF(i), solution.h:32
 *   linop  [nxy][nLin][nz] double   -- linOpFluct[i][Vnum], Model.h:46-50
 *   kx0    [nxy] double             -- Im(Kdata::kx[i]) after remaps, Kdata.h:212
 * with nxy = (NX-1)*NY/2, nz = 1 + 2*int(NZ/3) (Models/Model.h:32-36), mode
 * order j + ny*i' and kz order (0..nzl, -nzl..-1) (Auxiliary/Kdata.cpp:28-80),
 * fluctuations stored at NX*NY*NZ x and mean fields at NZ x the Fourier
 * coefficient, exactly as the reference stores them.
 */
#ifndef QL_DNS_H
#define QL_DNS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QL_OK 0
#define QL_ERR_ARG (-1)         /* invalid argument or unsupported size      */
#define QL_ERR_HIP (-2)         /* HIP runtime failure / no device           */
#define QL_ERR_UNSUPPORTED (-3) /* feature outside this build               */
#define QL_ERR_STATE (-4)       /* non-finite or runaway state detected      */

/* HD_fullU (Models/HD_fullU.cpp) has nLin = 2 fluctuation fields (ux, zeta)
 * and nMF = 2 mean fields (Ux, Uy) in every host layout below; it runs on the
 * MHD_FullQlin kernels with the magnetic fields held at zero. */
enum { QL_MODEL_MHD_FULLQLIN = 0, QL_MODEL_MHD_BQLIN = 1, QL_MODEL_HD_FULLU = 2 };
enum { QL_INT_RK3CN = 0, QL_INT_RK3CN_ELIN = 1 };
enum { QL_PREC_FP64 = 0, QL_PREC_MIXED = 1 };

/* Parameters of one run.  Field meanings follow the reference's Inputs
 * (Auxiliary/Input_parameters.cpp:68-138); model/integrator select what
 * main.cpp:58-74 and main.cpp:95 choose. */
typedef struct ql_params {
    int model;           /* QL_MODEL_*: equations_to_use_                    */
    int integrator;      /* QL_INT_*: RK3CN or RK3CN_Elin (main.cpp:95)       */
    int precision;       /* QL_PREC_*                                        */
    int NX, NY, NZ;      /* nx_, ny_, nz_ (NZ in {32, 64, 128, 256})          */
    double L[3];         /* lx_, ly_, lz_                                    */
    double q, omega;     /* q_, omega_ (MHD_BQlin uses 2q/3, MHD_BQlin.cpp:19)*/
    double nu, eta;      /* nu_, eta_                                        */
    double B0z;          /* B0z_                                             */
    int viscosity_order; /* viscosity_order_                                 */
    int quasilinear;     /* QuasiLinear?                                     */
    int remap;           /* Remap? (applied by ql_steps after each step)      */
    double dt;           /* dt_: > 0 fixed steps; < 0 variable steps of at most
                            -dt, CFL / Calculate_CFL (RK3CN.cpp:20-30, 79-83)  */
    int device;          /* HIP device ordinal                               */
    double CFL;          /* CFL number of variable steps (Input_parameters.cpp:77,
                            default 1.0); unused when dt > 0                  */
} ql_params;

typedef struct ql_ctx ql_ctx;

/* Model constructor + device allocation (replaces MHD_FullQlin::MHD_FullQlin,
 * Models/MHD_FullQlin.cpp:12-161, and MHD_BQlin::MHD_BQlin, MHD_BQlin.cpp:12-109;
 * the integrator's linearOPs_Init call is done with t0 = 0). */
int ql_create(const ql_params *p, ql_ctx **out);
int ql_destroy(ql_ctx *ctx);
/* dims[0..5] = nxy (Model::Dimxy, Model.h:56), nz (Model::NZ, :54),
 * nMF (num_MFs, :57), nLin (num_Lin, :58), half ((nz-1)/2+1), NZ (NZfull, :55) */
int ql_dims(const ql_ctx *ctx, int dims[6]);

/* solution upload/download (solution::pLin/pMF, solution.h:31-32).  Any
 * pointer may be NULL to leave that part unchanged / not read back. */
int ql_set_state(ql_ctx *ctx, const double *lin, const double *mf, const double *kx0);
int ql_get_state(ql_ctx *ctx, double *lin, double *mf, double *kx0);

/* Model::linearOPs_Init (Models/Model.h:50; MHD_FullQlin.cpp:429-456).
 * Also primes the integrator at t0 (RK3CN.cpp:57, RK3CN_Elin.cpp:257).
 * Outputs may be NULL. */
int ql_linear_ops_init(ql_ctx *ctx, double t0, double *linop, double *linop_mf);

/* Model::rhs(t, dt_lin, SolIn = current state, SolOut, linOpFluct)
 * (Models/Model.h:46-47; MHD_FullQlin.cpp:171-427; MHD_BQlin.cpp:117-284).
 * The state is not modified.  Outputs may be NULL. */
int ql_rhs(ql_ctx *ctx, double t, double dt_lin, double *rhs_lin, double *rhs_mf, double *linop);

/* Scratch of ql_rhs_noalloc, per device: 8 MiB. */
#define QL_NOALLOC_SCRATCH_BYTES ((size_t)8 << 20)

/* No counterpart in the reference: ql_rhs for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same computation as ql_rhs on the context's current
 * state, with the outputs in caller-owned DEVICE memory in the same layouts
 * (d_rhs_lin [nxy][4][nz] complex, d_rhs_mf [nMF][nz] complex, d_linop
 * [nxy][4][nz] double; any may be NULL) and the drive of ql_get_drive updated.
 * Its stress partial rows and column sums live in a fixed per-device scratch
 * area of QL_NOALLOC_SCRATCH_BYTES, so it calls no hipMalloc: launches with
 * (workgroups + 1) * nMF * NZ + 2 * nMF * nz doubles above that are refused
 * (QL_ERR_ARG).  The work is enqueued on the context's stream and the call
 * returns without waiting (ql_synchronize before reading the outputs).  Every
 * call on one device shares that scratch: calls that overlap in time on one
 * device (different streams or threads) race.  MHD_FullQlin and MHD_BQlin
 * contexts; HD_fullU returns QL_ERR_UNSUPPORTED. */
int ql_rhs_noalloc(ql_ctx *ctx, double t, double dt_lin, double *d_rhs_lin, double *d_rhs_mf, double *d_linop);

/* Integrator::Step(t, sol) (Integrators/Integrator.h:23; RK3CN.cpp:77-201,
 * RK3CN_Elin.cpp:275-389).  Writes the step size to *dt_out (may be NULL):
 * dt, or with dt < 0 the variable step CFL / Calculate_CFL capped at -dt
 * (RK3CN_Elin.cpp:76-80), computed on the device. */
int ql_step(ql_ctx *ctx, double t, double *dt_out);

/* nsteps of the reference main loop (main.cpp:108-123, noise off, no output):
 * Step, t += dt (accumulated in double, as main.cpp:116), then
 * ShearingBox_Remap(q t) if remap is on.  *t_inout is read and updated.
 * Variable steps (dt < 0) are chosen on the device step by step, so they
 * replay from the same captured graph; a run set with variable steps holds
 * one run. */
int ql_steps(ql_ctx *ctx, int nsteps, double *t_inout);

/* Model::ShearingBox_Remap(qt, sol) (MHD_FullQlin.cpp:464-481). */
int ql_remap(ql_ctx *ctx, double qt);

/* Model::Calc_Energy_AM_Diss with every save flag on (MHD_FullQlin.cpp:487-717):
 * out[0..3] energy (MU, MB, u, b), out[4..7] angular momentum, out[8..11]
 * dissipation, out[12..15] Reynolds projections (MHD_FullQlin only),
 * out[16..17] etaK*kmax.  Projections use the drive of the last rhs. */
int ql_diagnostics(ql_ctx *ctx, double t, double out[18]);

/* EMF / Reynolds-stress drives of the last rhs (Bx_drive_, By_drive_,
 * Ux_drive_, Uy_drive_ of MHD_FullQlin.cpp:389-399): [nMF][nz] complex. */
int ql_get_drive(ql_ctx *ctx, double *drive);

/* TimeVariables::Save_Mean_Fields values (TimeVariables.cpp:202-228):
 * fac1D * Re(ifft(pad(MF))) -> [nMF][NZ] doubles. */
int ql_mf_real(ql_ctx *ctx, double *out);

/* Full Reynolds-stress / EMF profiles of the last rhs, fac1D * Re(ifft(pad(
 * drive))) -> [nMF][NZ] doubles: HD_fullU's Reynolds saves
 * (HD_fullU.cpp:524-535, save_full_reynolds_). */
int ql_drive_real(ql_ctx *ctx, double *out);

/* Model::Calculate_CFL (MHD_FullQlin.cpp:769-780, MHD_BQlin.cpp:540-552,
 * HD_fullU.cpp:604-628 with its mean-field source term):
 * CFL/dt of the next variable step, kmax (max|By| + max|Bx| + max|Ux| +
 * max|Uy|) + q over the mean-field profiles the last rhs read (MHD_BQlin:
 * kmax max|By| + kmax max|Bx| + q, Bx of the current state).  With variable
 * steps the maxima are recorded on the device by every mean-field update; with
 * fixed steps (where the reference never calls it) nothing is recorded while
 * stepping and the value is that of the current state's profiles. */
int ql_calculate_cfl(ql_ctx *ctx, double *cfl_over_dt);

/* solution::Check_Solution (solution.cpp:169-203) as a device reduction of the
 * mean fields.  sums (may be NULL): [0] sum |Re ifft(pad(By))|, [1] sum
 * |Im ifft(pad(By))|, [2] max |By|, [3] max |Bx|, [4] sum |By|.  *verdict (may
 * be NULL): 0 passed, 1 the mean field has developed an imaginary part, 2 the
 * solution is very large (> 1e20), 3 it holds NaN or Inf.  A failed check
 * returns QL_ERR_STATE with the reference's message (where it aborts). */
int ql_check_solution(ql_ctx *ctx, double sums[5], int *verdict);

/* Driving noise of a run, as the Inputs set it and ChangeNoiseRange resets it
 * (Input_parameters.cpp:107-111; MHD_FullQlin.h:46-50).  Each driven entry
 * gets mult * (N(0, f_noise^2/2) + i N(0, f_noise^2/2)) with the reference's
 * multipliers, |k| range and ky = 0 Hermitian pairs (MHD_FullQlin.cpp:786-1053;
 * MHD_BQlin drives ky != 0 only and ignores the drive_only flags).  The
 * normals come from Philox4x64-10 keyed by seed and counted by (t, mode, kz
 * slot), so a replayed graph draws new noise at every t and a restart at the
 * same t draws the same noise; the reference's clock-seeded mt19937 stream is
 * matched in distribution, not in values. */
typedef struct ql_noise {
    double f_noise;                 /* f_noise_ (0: off, the default)          */
    double range_low, range_high;   /* noise_range_low_, noise_range_high_ in |k| */
    int drive_only_velocity;        /* drive_only_velocity?                    */
    int drive_only_magnetic;        /* drive_only_magnetic?                    */
    unsigned long long seed;        /* generator key                           */
} ql_noise;
int ql_set_noise(ql_ctx *ctx, const ql_noise *cfg);

/* Model::DrivingNoise(t, dt, sol): adds one step's noise at step-start time t.
 * ql_steps applies it inside every step when f_noise != 0, between Step and
 * the remap (main.cpp:111-123); ql_step does not (a caller mirroring main.cpp
 * calls this after it). */
int ql_driving_noise(ql_ctx *ctx, double t, double dt);

/* Waits for all work queued on the context's stream. */
int ql_synchronize(ql_ctx *ctx);

/* Runs nsteps like ql_steps with HIP events around every kernel launch and
 * returns, per kernel kind (0 fused per-mode kernel, 1 stress reduction +
 * mean-field kernel), the average duration in ms and the count.  For
 * measurement only: the events serialise the launches. */
int ql_profile_kernels(ql_ctx *ctx, int nsteps, double *t_inout, double avg_ms[2], int counts[2]);

/* Message of the last failed call on this thread ("" if none). */
const char *ql_last_error(void);

/* ------------------------------------------------------------------------
 * Run sets.  The reference runs one simulation per MPI job (main.cpp:20-160);
 * a parameter study is many such jobs.  A run set steps many runs of one
 * model / integrator / precision / dt (each with its own NX, NY, NZ, box, q,
 * omega, nu, eta, B0z, viscosity order, quasilinear and remap flags; runs of
 * one NZ share Lz) in the same kernel launches: runs are grouped by NZ, and a
 * whole step of every group is one replayed HIP graph.
 * ------------------------------------------------------------------------ */
typedef struct ql_batch ql_batch;

int ql_batch_create(const ql_params *runs, int nruns, ql_batch **out);
int ql_batch_destroy(ql_batch *b);
/* the run set of a context (owned by the context) */
ql_batch *ql_batch_of(ql_ctx *ctx);
int ql_batch_size(const ql_batch *b, int *nruns);
/* per-run counterparts of ql_dims / ql_set_state / ql_get_state /
 * ql_diagnostics / ql_get_drive, same layouts, run in [0, nruns) */
int ql_batch_dims(const ql_batch *b, int run, int dims[6]);
int ql_batch_set_state(ql_batch *b, int run, const double *lin, const double *mf, const double *kx0);
int ql_batch_get_state(ql_batch *b, int run, double *lin, double *mf, double *kx0);
int ql_batch_diagnostics(ql_batch *b, int run, double t, double out[18]);
int ql_batch_get_drive(ql_batch *b, int run, double *drive);
/* per-run ql_calculate_cfl / ql_check_solution */
int ql_batch_calculate_cfl(ql_batch *b, int run, double *cfl_over_dt);
int ql_batch_check_solution(ql_batch *b, int run, double sums[5], int *verdict);
/* per-run ql_set_noise; ql_driving_noise for every driven run of the set */
int ql_batch_set_noise(ql_batch *b, int run, const ql_noise *cfg);
int ql_batch_driving_noise(ql_batch *b, double t, double dt);
/* integrator priming at t0 of every run (RK3CN.cpp:57, RK3CN_Elin.cpp:257) */
int ql_batch_linear_ops_init(ql_batch *b, double t0);
/* nsteps of main.cpp:108-123 for every run (as ql_steps), one clock; a set
 * of several runs with variable steps (dt < 0) takes ql_batch_steps_runs */
int ql_batch_steps(ql_batch *b, int nsteps, double *t_inout);
/* nsteps for every run with each run's own time, t_runs[nruns] read and
 * updated: with variable steps each run chooses its steps from its own CFL
 * condition (RK3CN.cpp:79-83) on its own device clock; with fixed steps the
 * times must be equal (one shared clock) */
int ql_batch_steps_runs(ql_batch *b, int nsteps, double *t_runs);
/* Launch-shape tuner: for each NZ group, times every candidate shape (z
 * points per workgroup pass 128 or 256, or one wavefront per mode (64), and
 * a target workgroup count) over
 * trial_steps steps and keeps the fastest; the state is restored afterwards.
 * *step_ms (may be NULL) receives the sum over groups of the best step times.
 * The shape sets the summation order of the stress reduction: results of two
 * shapes agree to rounding, results of one shape are bitwise reproducible. */
int ql_batch_tune(ql_batch *b, int trial_steps, double *step_ms);
/* Transform of the workgroup mode kernel for the NZ group (fftwPlans.h:33-97):
 * the LDS radix passes (QL_TRANSFORM_FFT, the default) or dense complex
 * matrix products on the matrix cores (QL_TRANSFORM_DFT_MFMA: FP64, NZ 32 or
 * 64, line 128 or 256, else QL_ERR_ARG).  ql_batch_tune times both where both
 * exist.  A later ql_batch_set_config keeps it where the new shape has it. */
enum { QL_TRANSFORM_FFT = 0, QL_TRANSFORM_DFT_MFMA = 1 };
int ql_batch_set_transform(ql_batch *b, int NZ, int transform);
/* Substep tail of the NZ group: the column reduction and the mean-field
 * kernel as two launches (QL_TAIL_TWO_KERNELS, the default) or one launch with
 * a workgroup per mean field and run (QL_TAIL_FUSED: fixed steps, NZ <= 128,
 * at most 65536 / NZ partial rows per run, else QL_ERR_ARG).  Both sum the
 * stress / EMF partial rows in the same order: the results are bitwise equal.
 * No counterpart in the reference (its MPI Allreduce, MHD_FullQlin.cpp:386,
 * followed by the mean-field update, :386-423). */
enum { QL_TAIL_TWO_KERNELS = 0, QL_TAIL_FUSED = 1 };
int ql_batch_set_tail(ql_batch *b, int NZ, int tail);
/* Mixed precision (QL_PREC_MIXED): a mode whose largest |value| reaches tau
 * times the largest of its run runs in FP64 (a companion launch of the FP64
 * kernel before the FP32 one), the others in FP32; the choice is made again
 * every 16 steps of the clock and after each state upload.  Sets tau when
 * tau > 0 (default 0.03; tau <= 0 leaves it; tau > 1, which no mode reaches,
 * runs every mode in FP32 with neither the flags nor the FP64 launch) and
 * writes to *nfp64 (may be NULL) how many modes of the set reach it for the
 * current state. */
int ql_batch_mixed_split(ql_batch *b, double tau, int *nfp64);
/* current shape per group: cfg[7k..7k+6] = NZ, line, target, modes per
 * workgroup pass, workgroups, transform (QL_TRANSFORM_*), substep tail in use
 * (QL_TAIL_*); *ngroups receives the group count */
enum { QL_CONFIG_INTS = 7 };
int ql_batch_get_config(const ql_batch *b, int max_groups, int *ngroups, int *cfg);
/* sets the shape of the NZ group: line in {64, 128, 256}, target >= 1.
 * line 64 selects the kernel with one wavefront per mode (modes per pass 1;
 * every model; at NZ = 256 FP64 only except for HD_fullU, else QL_ERR_ARG) */
int ql_batch_set_config(ql_batch *b, int NZ, int line, int target);

/* ------------------------------------------------------------------------
 * Adjoint (reverse-mode gradient) of nsteps of ql_steps, for a single-run
 * MHD_FullQlin, MHD_BQlin or HD_fullU context in FP64, driven or not (either
 * integrator, remap on or off), with fixed or variable steps.  With variable
 * steps (dt < 0) the step sizes are those the forward chose and are held
 * fixed: the gradient is that of the forward with its step sequence frozen
 * (the CFL condition's own dependence on the state and the parameters is not
 * differentiated).  The reference has no adjoint;
 * these entry points differentiate the reference's step
 * (Models/MHD_FullQlin.cpp:171-427, Models/MHD_BQlin.cpp:117-306,
 * Integrators/RK3CN.cpp:77-201, Integrators/RK3CN_Elin.cpp:74-188,
 * main.cpp:108-123) as this library computes it.  Remap decisions are held
 * fixed (piecewise constant in q).  MHD_BQlin rotates at 2 q / 3
 * (MHD_BQlin.cpp:19), so its omega and B0z gradients are zero; HD_fullU has
 * no eta and no B0z (zero gradients) and its layouts are its own (nLin = 2,
 * nMF = 2, as ql_dims reports).
 *
 * ql_adjoint_forward: nsteps of ql_steps from the current state, storing a
 * checkpoint every `every` steps (device memory ~ (nsteps/every + 3 every + 8)
 * states).  ql_adjoint_backward: given dL/d(final state) (g_lin [nxy][nLin][nz]
 * complex, g_mf [nMF][nz] complex, either may be NULL for zero; complex
 * gradients as dL/dRe + i dL/dIm), returns dL/d(state before the forward) in
 * g_lin0 / g_mf0 (NULL to skip) and g_par[0..4] = dL/d(q, omega, nu, eta,
 * B0z).  Segments are recomputed from their checkpoints; the context is left
 * at the forward's final state. */
int ql_adjoint_forward(ql_ctx *ctx, int nsteps, int every, double *t_inout);
int ql_adjoint_backward(ql_ctx *ctx, const double *g_lin, const double *g_mf, double *g_lin0, double *g_mf0,
                        double g_par[5]);

#ifdef __cplusplus
}
#endif
#endif
