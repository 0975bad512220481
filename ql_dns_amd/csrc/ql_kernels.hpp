// ql_kernels.hpp -- shared structures and the non-modal kernels of the QL time
// step.
//
// Runs are processed in groups that share NZ (one kernel instantiation per NZ).
// A group's runs are stored one after the other in its arrays; every run has
// its own physical parameters (RunParams), mean fields and reductions.  One RK3
// substep of a group is three launches on one stream:
//   k_mode<..>    fused per-(kx,ky) RHS + substep update (ql_mode.hpp); each
//                 workgroup owns a contiguous range of one run's modes and
//                 writes one EMF / stress partial row for that run;
//   k_reduce_cols fixed-order, compensated sum of each run's partial rows;
//   k_meanfield   one workgroup per run: forward transform of the summed
//                 profiles, x i kz, Hermitian fill (the reference's MPI
//                 Allreduce + fill, MHD_FullQlin.cpp:386-399), mean-field RHS
//                 and update, and the real-space profiles the next substep reads.
// (A single launch in which the last reducing workgroup of a run goes on with
// the mean fields was measured slower: the device-scope fences and the atomic
// count cost more than the launch boundary of a replayed graph.)
// Step times live on the device (one Clock per group), advanced by the
// group's last mean-field kernel of each step, so a whole step can be captured
// once in a HIP graph and replayed.
#pragma once
#include "ql_fft.hpp"

namespace qlk {

enum { MODE_RHS = 0, MODE_ELIN = 1, MODE_CN = 2, MODE_PROFILES = 3 };

// Integrator coefficients (Integrators/RK3CN.cpp:19-20, RK3CN_Elin.cpp:19-20)
constexpr double A0 = 8.0 / 15.0, A1 = 5.0 / 12.0, A2 = 3.0 / 4.0;
constexpr double B1 = -17.0 / 60.0, B2 = -5.0 / 12.0;
constexpr double P1 = 8.0 / 15.0, P2 = 2.0 / 3.0;

// per-run record of the mean-field quantities that variable steps use:
// [0, 8) max |profile line| of the current state, [8, 16) those of the
// profiles the last right-hand side read (the reference's MBx_ ... arrays);
// for HD_fullU also the energy sum |MUx|^2 + |MUy|^2 of the current profiles
// and of the last rhs's, mean |kz^2 drive| of the Ux and Uy drives of the last
// rhs, and mean |kz^2 Ux| of the current state (HD_fullU.cpp:609-624)
constexpr int PM_CUR = 0, PM_RHS = 8, PM_E_CUR = 16, PM_E_RHS = 17, PM_DUX = 18, PM_DUY = 19, PM_UX = 20;
constexpr int PM_STRIDE = 24;

// physical parameters of one run
struct Phys {
    double q, omega, B0z, nu, eta;
    int vo2;  // viscosity_order / 2 (integer division, MHD_FullQlin.cpp:372)
    int quasilinear;
};

struct RunParams {
    Phys ph;
    double rsf;          // 1/(NX NY)^2 (MHD_FullQlin.cpp:67)
    double remap_thr;    // (NX-1)/2 * 2 pi/Lx (MHD_FullQlin.cpp:472)
    double remap_shift;  // (NX-1) * 2 pi/Lx (MHD_FullQlin.cpp:474)
    int remap;           // Remap?
    int nxy;             // modes of the run
    int mode_off;        // first mode of the run in the group arrays
    int wg_off, nwg;     // its mode-kernel workgroups = its partial rows
    int wg_off2, nwg2;   // its rows of the FP64 companion launch (mixed precision), after them
    double kmax;         // CFL wavenumber (MHD_FullQlin.cpp:92-97)
    double cfl, dtmax;   // CFL number and -dt of variable steps (RK3CN.cpp:20-30)
    double Lz;
    int hd;              // HD_fullU run (MHD_FullQlin kernels with b = B = 0)
    int pad;
};

// row r of a run's partial rows: its mode-kernel rows, then its companion rows
__device__ __forceinline__ int run_row(const RunParams &rp, int r)
{
    return r < rp.nwg ? rp.wg_off + r : rp.wg_off2 + (r - rp.nwg);
}

// one mode-kernel workgroup: its run, its mode-group range [g0, g1), and the
// run's first mode and mode count (one record, one scalar load)
struct WgInfo {
    int run, g0, g1, mode_off, nxy, pad0, pad1, pad2;
};

// time of a set of runs stepped together: t as main.cpp:116 accumulates it,
// t_prev the previous step start; old_init: the RK3CN "old" operator is the
// one of linearOPs_Init at t_init (RK3CN.cpp:57); dt the step about to be
// taken, dt_last the one taken last
struct Clock {
    double t, t_prev, t_init;
    int old_init, nstep;  // nstep: steps since the integrator was set up
    double dt, dt_last;
};

struct ModeArgs {
    double2 *u;             // state [modes][4][nz], updated in place (MODE_ELIN/CN)
    double2 *n_out;         // RHS out (nullable)
    const double2 *n_prev;  // previous substep RHS (nullable)
    double *lin_out;        // linear operator out (MODE_RHS, nullable)
    double *kx0;            // [modes] (remap updates it)
    const double *ky;       // [modes]
    const double *kzv;      // [nz] physical kz
    const double2 *tw;      // [NZ] exp(-2 pi i k/NZ)
    const float2 *tw32;     // the same, rounded to FP32
    const double *mfz;      // real-space mean-field profiles [runs][8][NZ]
    double *partials;       // [workgroups][nacc][NZ]
    const RunParams *runs;  // [runs]
    const WgInfo *wg;       // per workgroup
    const Clock *clock;
    int s;                  // substep 1..3, or 0 for the explicit times below
    int mode, do_remap;
    int vdt;                // variable steps: dt and its coefficients from the clock
    double ts_x, dtl_x;     // explicit rhs time and dt_lin (s == 0)
    double dt, linC, nlc1, nlc2;
    double dtP1, dtP2, dtl1, dtl2, dtl3;  // dt*8/15, dt*2/3, and the three dt_lin
    double fac;             // 1/NZ
    // mixed-precision split (split == SPLIT_OFF: not used): a mode flagged in
    // hi (its amplitude reached a fraction of its run's largest when the flags
    // were last set, k_amp_flags) runs in FP64, the others in FP32: the FP32
    // launch (SPLIT_LO) skips the flagged modes and an FP64 companion launch
    // (SPLIT_HI) takes only them
    const unsigned char *hi;  // [modes]
    int split;
};

enum { SPLIT_OFF = 0, SPLIT_LO = 1, SPLIT_HI = 2 };
constexpr int COMP_LIST = 512;  // candidate modes per companion workgroup, at most

// max(|re|, |im|) of two complex values, as a float
__device__ __forceinline__ float amp2(double2 x, double2 y)
{
    return (float)fmax(fmax(fabs(x.x), fabs(x.y)), fmax(fabs(y.x), fabs(y.y)));
}

// substep coefficients of RK3CN_Elin::Step (RK3CN_Elin.cpp:88-90, :118-121,
// :151-154) and RK3CN::Step (linC with dt/2, RK3CN.cpp:92-95, :130-133,
// :166-169), the host's products rounded one by one
struct StepCoefs {
    double linC, nlc1, nlc2;
};
__device__ __forceinline__ StepCoefs step_coefs(double dt, int s, bool cn)
{
    const double h = cn ? dt * 0.5 : dt;
    const double a = s == 1 ? A0 : (s == 2 ? A1 : A2);
    const double ab = s == 1 ? A0 : (s == 2 ? A1 + B1 : A2 + B2);
    const double b = s == 1 ? 0.0 : (s == 2 ? B1 : B2);
    return StepCoefs{rn_mul(h, ab), rn_mul(a, dt), rn_mul(b, dt)};
}

// substep times from the clock, with the host's roundings (unfused): ts,
// dt_lin, where the RK3CN "old" operator was evaluated, and t + dt; with
// variable steps also the substep's coefficients (fixed steps take them from
// the arguments)
struct SubTimes {
    double ts, dtl, tso, dtlo, tnew;
    double linC, nlc1, nlc2;
    bool oinit;
};
template <bool CN>
__device__ __forceinline__ SubTimes sub_times(const ModeArgs &a, int run)
{
    // scalars first, the struct once at the end (partial stores of a local
    // struct in separate branches can leave it in scratch memory)
    if (a.s == 0) return SubTimes{a.ts_x, a.dtl_x, 0.0, 0.0, 0.0, a.linC, a.nlc1, a.nlc2, false};
    const Clock *ck = a.clock + run;
    const double t = ck->t;
    double dt = a.dt, dtP1 = a.dtP1, dtP2 = a.dtP2, dtl2 = a.dtl2, dtl3 = a.dtl3;
    double dtP2o = a.dtP2, dtl3o = a.dtl3;  // the previous step's (old operator of substep 1)
    StepCoefs c{a.linC, a.nlc1, a.nlc2};
    if (a.vdt) {
        dt = ck->dt;
        dtP1 = rn_mul(dt, P1);
        dtP2 = rn_mul(dt, P2);
        dtl2 = rn_mul(dt, P2 - P1);
        dtl3 = rn_mul(dt, 1.0 - P2);
        const double dl = ck->dt_last;
        dtP2o = rn_mul(dl, P2);
        dtl3o = rn_mul(dl, 1.0 - P2);
        c = step_coefs(dt, a.s, CN);
    }
    const double dtl1 = dtP1;  // dt * p1 (RK3CN_Elin.cpp:90)
    const double tnew = rn_add(t, dt);
    const bool s1 = a.s == 1, s2 = a.s == 2;
    const bool oinit = s1 && ck->old_init != 0;
    const double ts = s1 ? t : rn_add(t, s2 ? dtP1 : dtP2);
    const double dtl = s1 ? dtl1 : (s2 ? dtl2 : dtl3);
    const double tso = s1 ? (oinit ? ck->t_init : rn_add(ck->t_prev, dtP2o))
                          : (s2 ? t : rn_add(t, dtP1));
    const double dtlo = s1 ? (oinit ? 0.0 : dtl3o) : (s2 ? dtl1 : dtl2);
    return SubTimes{ts, dtl, tso, dtlo, tnew, c.linC, c.nlc1, c.nlc2, oinit};
}

// dt of the next variable step (RK3CN_Elin.cpp:76-80) from the recorded
// profile maxima: CFL / Calculate_CFL, at most dtmax.  Calculate_CFL is
// kmax (max|By| + max|Bx| + max|Ux| + max|Uy|) + q over the profiles of the
// last rhs (MHD_FullQlin.cpp:773-778), or kmax max|By| + kmax max|Bx| + q with
// By of the last rhs and Bx of the current state (MHD_BQlin.cpp:544-550).
// Profile lines: Bx 0, By 2, Ux 4, Uy 6 (MHD_FullQlin); By 0, Bx 3 (MHD_BQlin).
__host__ __device__ inline double cfl_value(const double *pm, const RunParams &rp, int model)
{
#pragma clang fp contract(off)
    const double *cur = pm + PM_CUR, *rhs = pm + PM_RHS;
    if (model == 0) {
        // with B = 0 (HD_fullU) the sum is Ux + Uy exactly (HD_fullU.cpp:628)
        const double s = ((rhs[2] + rhs[0]) + rhs[4]) + rhs[6];
        const double k = rp.kmax * s;
        double v = k + rp.ph.q;
        if (rp.hd) {
            // HD_fullU.cpp:614-624: mean-field source term above an energy cutoff
            double mf = 0.0;
            if (rp.ph.quasilinear && pm[PM_E_RHS] * rp.Lz > 1e-6) mf = pm[PM_DUX] / pm[PM_UX] + pm[PM_DUY] / pm[PM_UX];
            v = v + mf;
        }
        return v;
    }
    const double ky = rp.kmax * rhs[0], kx = rp.kmax * cur[3];
    const double s = ky + kx;
    return s + rp.ph.q;
}
__host__ __device__ inline double cfl_dt(const double *pm, const RunParams &rp, int model)
{
#pragma clang fp contract(off)
    double dt = rp.cfl / cfl_value(pm, rp, model);
    if (dt > rp.dtmax) dt = rp.dtmax;
    return dt;
}

// -|x|^(order/2): the dissipation exponent uses integer division
// (MHD_FullQlin.cpp:372); order 2 and 4 avoid pow.
__device__ __forceinline__ double neg_abs_pow(double x, int vo2)
{
    const double a = fabs(x);
    if (vo2 == 1) return -a;
    if (vo2 == 2) return -(a * a);
    if (vo2 == 0) return -1.0;
    return -pow(a, (double)vo2);
}

// remap predicate of ShearingBox_Remap (MHD_FullQlin.cpp:470-472), evaluated
// with unfused roundings so it agrees bit for bit with the host expression
__device__ __forceinline__ bool remap_hit(double qt, double thr, double kx0, double ky)
{
    return rn_add(kx0, rn_mul(qt, ky)) > thr;
}

// every run's clock from the host's view v (passed by value, no host buffer
// involved); with variable steps dt is each run's next step, from its
// profile maxima
[[maybe_unused]] static __global__ void k_clock_fill(Clock *c, int nruns, Clock v, const double *pm,
                                                     const RunParams *rp, int model, int vdt)
{
    for (int r = threadIdx.x; r < nruns; r += blockDim.x) {
        Clock w = v;
        if (vdt) w.dt = cfl_dt(pm + (size_t)r * PM_STRIDE, rp[r], model);
        c[r] = w;
    }
}

// each run's next variable step dt after its clock was uploaded
[[maybe_unused]] static __global__ void k_clock_dt(Clock *c, int nruns, const double *pm, const RunParams *rp,
                                                   int model)
{
    for (int r = threadIdx.x; r < nruns; r += blockDim.x) c[r].dt = cfl_dt(pm + (size_t)r * PM_STRIDE, rp[r], model);
}

// t_prev <- t, t <- t + dt (main.cpp:116), once per step
__device__ __forceinline__ void clock_advance(Clock *c)
{
    const double t = c->t, dt = c->dt;
    c->t_prev = t;
    c->t = rn_add(t, dt);
    c->dt_last = dt;
    c->old_init = 0;
    c->nstep += 1;
}

// ---------------------------------------------------------------------------
// stress / EMF reduction: column c of run r's partial rows summed with
// Neumaier compensation.  A block (blockIdx.y = run) owns RED_CW columns; its
// NT / RED_CW row chunks are summed in row order by separate threads and then
// added in chunk order, so the result does not depend on scheduling.  No
// cross-block communication (no fences, no atomics).
// ---------------------------------------------------------------------------
constexpr int RED_CW = 16;  // columns per block (4, 8 and 32 measured slower)

[[maybe_unused]] static __global__ __launch_bounds__(NT) void k_reduce_cols(const double *__restrict__ part, const RunParams *__restrict__ runs,
                                                    int ncols, double *__restrict__ sums)
{
    constexpr int CW = RED_CW, CH = NT / CW;
    __shared__ double red[CH][CW + 1], red2[CH][CW + 1];
    const RunParams &rp = runs[blockIdx.y];
    const int cl = threadIdx.x % CW, chunk = threadIdx.x / CW;
    const int c = blockIdx.x * CW + cl;
    const int nrows = rp.nwg + rp.nwg2;
    const int per = (nrows + CH - 1) / CH;
    const int r0 = chunk * per, r1 = min(nrows, chunk * per + per);
    constexpr int B = 256 / CH;  // rows in flight per thread (one batch up to 256 rows)
    double s = 0.0, e = 0.0;
    if (c < ncols) {
        for (int r = r0; r < r1; r += B) {
            double v[B];
#pragma unroll
            for (int k = 0; k < B; ++k) v[k] = (r + k < r1) ? part[(size_t)run_row(rp, r + k) * ncols + c] : 0.0;
#pragma unroll
            for (int k = 0; k < B; ++k) {
                const double t = s + v[k];
                e += (fabs(s) >= fabs(v[k])) ? ((s - t) + v[k]) : ((v[k] - t) + s);
                s = t;
            }
        }
    }
    red[chunk][cl] = s;
    red2[chunk][cl] = e;
    __syncthreads();
    if (chunk == 0 && c < ncols) {
        double t = 0.0, te = 0.0;
        for (int k = 0; k < CH; ++k) {
            const double v = red[k][cl];
            const double u = t + v;
            te += (fabs(t) >= fabs(v)) ? ((t - u) + v) : ((v - u) + t);
            te += red2[k][cl];
            t = u;
        }
        sums[(size_t)blockIdx.y * ncols + c] = t + te;
    }
}

// mixed precision: every `every` steps of the clock (always when clock is
// null), max |value| of each mode's state, one wavefront per mode
[[maybe_unused]] static __global__ __launch_bounds__(64) void k_amp_modes(const double2 *__restrict__ u, int nz,
                                                                          float *__restrict__ amp,
                                                                          const Clock *__restrict__ clock, int every)
{
    if (clock && clock->nstep % every != 0) return;
    const size_t i = blockIdx.x;
    float m = 0.0f;
    for (int k = threadIdx.x; k < 2 * nz; k += 64) m = fmaxf(m, amp2(u[i * 4 * nz + k], u[i * 4 * nz + 2 * nz + k]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (threadIdx.x == 0) amp[i] = m;
}

// then, per run (one workgroup each), its largest amplitude and the flag of
// each mode that reaches tau times it (the modes the FP64 companion takes)
constexpr int AMP_NT = 1024;
[[maybe_unused]] static __global__ __launch_bounds__(AMP_NT) void k_amp_flags(const float *__restrict__ amp,
                                                                             const RunParams *__restrict__ runs,
                                                                             double tau, unsigned char *__restrict__ hi,
                                                                             const Clock *__restrict__ clock, int every)
{
    if (clock && clock->nstep % every != 0) return;
    __shared__ float red[AMP_NT / 64];
    const RunParams &rp = runs[blockIdx.x];
    const int tid = threadIdx.x;
    float m = 0.0f;
    for (int l = tid; l < rp.nxy; l += AMP_NT) m = fmaxf(m, amp[rp.mode_off + l]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (tid % 64 == 0) red[tid / 64] = m;
    __syncthreads();
    float top = 0.0f;
#pragma unroll
    for (int w = 0; w < AMP_NT / 64; ++w) top = fmaxf(top, red[w]);
    const double thr = top > 0.0f ? tau * (double)top : __builtin_inf();
    for (int l = tid; l < rp.nxy; l += AMP_NT) hi[rp.mode_off + l] = (double)amp[rp.mode_off + l] >= thr ? 1 : 0;
}

// sum over the workgroup's threads in a fixed order (every thread gets it)
__device__ __forceinline__ double block_sum(double v, double *red)
{
    const int tid = threadIdx.x;
    __syncthreads();
    red[tid] = v;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    return red[0];
}

// ---------------------------------------------------------------------------
// mean fields of one run, by one workgroup
// ---------------------------------------------------------------------------
struct MFArgs {
    const double *sums;       // reduced stress / EMF z-profiles [runs][nMF][NZ]
    double2 *mf;              // [runs][nMF][nz]: the updated mean fields
    const double2 *mf_in;     // the mean fields at the substep start (null: mf, updated in place)
    double2 *mfn_out;         // mean-field RHS out
    const double2 *mfn_prev;  // previous substep mean-field RHS (nullable)
    double2 *drive;           // [runs][nMF][nz]
    double *mfz;              // profiles out [runs][8][NZ]
    const RunParams *runs;
    const double *kzv;
    const double2 *tw;
    Clock *clock;  // advanced by dt when advance != 0 (last substep of a step)
    double *pm;    // profile maxima [runs][PM_STRIDE]
    int model, nMF, mode, advance;
    int s, vdt;    // substep; variable steps (coefficients from the clock)
    int rec;       // record the CFL quantities in pm (variable steps, or on request)
    double fac, linC, nlc1, nlc2;
};

// mean fields of a model: MHD_FullQlin (and HD_fullU on its layout) 4, MHD_BQlin 2
constexpr int NMF_OF(int model) { return model == 0 ? 4 : 2; }

template <int NZ>
struct MFShared {
    static constexpr int NZD = Dealias<NZ>::nz;
    double2 pad[8 * PadIO<NZ, double2>::LP];  // between the two passes of a transform
    double2 line[8 * NZ];                     // stress spectra, then the profile lines
    double2 tw[NZ];
    double skz[NZD];
    double ssum[4 * NZ];      // summed stress / EMF z-profiles
    double2 smf[4 * NZD];     // mean fields at the start of the substep
    double2 sprev[4 * NZD];   // previous substep's mean-field RHS
    double2 sdrive[4 * NZD];  // drives (HD_fullU records)
    double2 smfn[4 * NZD];    // updated mean fields
    double red[NT];           // block sums
};

// Inputs of the mean-field kernel (twiddles, wavenumbers, mean fields,
// previous RHS, summed profiles) into LDS, every load issued before the first
// store so that they are in flight together; meanfield_body starts with the
// barrier that makes them visible.
template <int NZ, int MODE, int MODEL>
__device__ __forceinline__ void meanfield_load(const MFArgs &a, int run, MFShared<NZ> &sh)
{
    constexpr int NZD = Dealias<NZ>::nz;
    constexpr int TW = (NZ + NT - 1) / NT, MW = (4 * NZD + NT - 1) / NT, SW = (4 * NZ + NT - 1) / NT;
    static_assert(NZD <= NT, "one wavenumber per thread");
    const int tid = threadIdx.x;
    const int nslot = NMF_OF(MODEL) * NZD, nsum = NMF_OF(MODEL) * NZ;
    const size_t mo = (size_t)run * nslot;
    const bool upd = MODE != MODE_PROFILES;
    const double2 *mf_in = a.mf_in ? a.mf_in : a.mf;
    double2 t[TW], m[MW], p[MW];
    double su[SW];
#pragma unroll
    for (int w = 0; w < TW; ++w) t[w] = (tid + w * NT < NZ) ? a.tw[tid + w * NT] : make_double2(0.0, 0.0);
    const double kz = tid < NZD ? a.kzv[tid] : 0.0;
#pragma unroll
    for (int w = 0; w < MW; ++w) {
        const int k = tid + w * NT;
        m[w] = (k < nslot) ? mf_in[mo + k] : make_double2(0.0, 0.0);
        p[w] = (upd && k < nslot && a.mfn_prev) ? a.mfn_prev[mo + k] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int w = 0; w < SW; ++w) {
        const int k = tid + w * NT;
        su[w] = (upd && k < nsum) ? a.sums[(size_t)run * nsum + k] : 0.0;
    }
#pragma unroll
    for (int w = 0; w < TW; ++w)
        if (tid + w * NT < NZ) sh.tw[tid + w * NT] = t[w];
    if (tid < NZD) sh.skz[tid] = kz;
#pragma unroll
    for (int w = 0; w < MW; ++w) {
        const int k = tid + w * NT;
        if (k < nslot) {
            sh.smf[k] = m[w];
            sh.sprev[k] = p[w];
        }
    }
#pragma unroll
    for (int w = 0; w < SW; ++w)
        if (tid + w * NT < nsum) sh.ssum[tid + w * NT] = su[w];
}

// Slot m of mean field f: its drive, d/dz = i kz of the field's transformed
// stress / EMF line S (the EMF, MHD_FullQlin.cpp:326-336, and the stresses,
// accumulated before their z-derivative by the mode kernel; the non-negative
// half is kept and the negative half filled by conjugate symmetry, :386-399),
// the mean-field RHS (:411-423, MHD_BQlin.cpp:274-280) from the substep's mean
// fields M [field][NZD], and for MODE_ELIN / MODE_CN the update
// (RK3CN_Elin.cpp:109-112 / RK3CN.cpp:112-115) with P, the slot's previous
// substep RHS (used when has_prev)
template <int NZD, int HALF, int MODE, int MODEL>
__device__ __forceinline__ void mf_slot(int f, int m, const double2 *S, const double *skz, const double2 *M,
                                        const Phys &ph, bool has_prev, double2 P, double linC, double nlc1,
                                        double nlc2, double2 &Dv, double2 &N, double2 &Mn)
{
    const bool cj = m >= HALF;
    const int src = cj ? NZD - m : m;
    const double2 Dv0 = imul(skz[src], S[src]);
    Dv = cj ? make_double2(Dv0.x, -Dv0.y) : Dv0;
    const double kz = skz[m], kz2 = -(kz * kz);
    N = make_double2(0.0, 0.0);
    double Lmf = 0.0;
    if (ph.quasilinear) {
        if (MODEL == 0) {
            const double2 M0 = M[m], M1 = M[NZD + m], M2 = M[2 * NZD + m], M3 = M[3 * NZD + m];
            if (f == 0) N = cadd(imul(ph.B0z * kz, M2), Dv);
            else if (f == 1) N = cadd(cadd(cscale(-ph.q, M0), imul(ph.B0z * kz, M3)), Dv);
            else if (f == 2) N = cadd(cadd(cscale(2 * ph.omega, M3), imul(ph.B0z * kz, M0)), Dv);
            else N = cadd(cadd(cscale(ph.q - 2.0 * ph.omega, M2), imul(ph.B0z * kz, M1)), Dv);
            Lmf = (f < 2 ? ph.eta : ph.nu) * kz2;
        } else {
            if (f == 0) N = Dv;
            else N = cadd(cscale(-ph.q, M[m]), Dv);
            Lmf = ph.eta * kz2;
        }
    }
    Mn = make_double2(0.0, 0.0);
    if (MODE == MODE_RHS) return;
    const double2 Mo = M[f * NZD + m];
    if (MODE == MODE_ELIN) {
        double2 sacc = cscale(nlc1, N);
        if (has_prev) sacc = cadd(sacc, cscale(nlc2, P));
        Mn = cscale(exp(linC * Lmf), cadd(sacc, Mo));
    } else {
        const double denom = 1.0 / (1.0 - linC * Lmf);
        double2 sacc = cscale(nlc1 * denom, N);
        if (has_prev) sacc = cadd(sacc, cscale(nlc2 * denom, P));
        Mn = cadd(sacc, cscale((1 + linC * Lmf) * denom, Mo));
    }
}

// Both transforms are two passes (radix 16 x 8 at NZ = 128) between separate
// LDS regions; the inverse one builds its padded input lines from the updated
// spectra in its first pass and stores the real profiles in its second, so a
// substep takes six barriers.  meanfield_load must have run.
template <int NZ, int MODE, int MODEL, bool REC>
__device__ __forceinline__ void meanfield_body(const MFArgs &a, int run, MFShared<NZ> &sh)
{
    using D = Dealias<NZ>;
    using PL = Plan2<NZ>;
    constexpr int NZD = D::nz, HALF = D::half;
    constexpr int MW = (4 * NZD + NT - 1) / NT;  // mean-field slots per thread
    const int tid = threadIdx.x;
    double *pm = a.pm + (size_t)run * PM_STRIDE;
    StepCoefs cf{a.linC, a.nlc1, a.nlc2};
    if (REC && a.vdt && MODE != MODE_PROFILES && a.s > 0) cf = step_coefs(a.clock[run].dt, a.s, MODE == MODE_CN);
    const double linC = cf.linC, nlc1 = cf.nlc1, nlc2 = cf.nlc2;
    // the right-hand side of this substep read the profiles of the current
    // state (Calculate_CFL reads those, MHD_FullQlin.cpp:772)
    const bool rec = REC && a.rec != 0;
    if (rec && MODE != MODE_PROFILES && tid < 8) pm[PM_RHS + tid] = pm[PM_CUR + tid];
    if (rec && MODE != MODE_PROFILES && tid == 8) pm[PM_E_RHS] = pm[PM_E_CUR];
    const bool hd = rec && a.runs[run].hd != 0;
    // mean |kz^2 Ux_drive|, mean |kz^2 Uy_drive| of a drive (HD_fullU.cpp:619)
    auto drive_means = [&](const double2 *d) {
        double sx = 0.0, sy = 0.0;
        for (int m = tid; m < NZD; m += NT) {
            const double k2 = -(sh.skz[m] * sh.skz[m]);
            const double2 dx = d[2 * NZD + m], dy = d[3 * NZD + m];
            sx += hypot(dx.x * k2, dx.y * k2);
            sy += hypot(dy.x * k2, dy.y * k2);
        }
        sx = block_sum(sx, sh.red);
        sy = block_sum(sy, sh.red);
        if (tid == 0) {
            pm[PM_DUX] = sx / NZD;
            pm[PM_DUY] = sy / NZD;
        }
    };
    // the substep's mode kernel, the only reader of the clock, has finished;
    // dt is read above, before this block can change it (the next step's dt is
    // written at the end)
    if (a.advance && tid == 0) clock_advance(a.clock + run);
    const Phys ph = a.runs[run].ph;
    const int nMF = NMF_OF(MODEL), nslot = nMF * NZD;
    const size_t mo = (size_t)run * nslot;
    const PadIO<NZ, double2> pad{sh.pad};
    const BufIO<NZ, double2> lines{sh.line};
    __syncthreads();  // the inputs meanfield_load stored

    if (MODE != MODE_PROFILES) {
        // forward transform of the summed z-profiles (real lines)
        const double *S = sh.ssum;
        const auto ld = [S](int l, int j) { return make_double2(S[l * NZ + j], 0.0); };
        if (nMF == 4) {
            fft_pass_io<NZ, 4, false, PL::R1, 1, false>(sh.tw, ld, pad);
            fft_pass_io<NZ, 4, false, PL::R2, PL::R1, false>(sh.tw, pad, lines);
        } else {
            fft_pass_io<NZ, 2, false, PL::R1, 1, false>(sh.tw, ld, pad);
            fft_pass_io<NZ, 2, false, PL::R2, PL::R1, false>(sh.tw, pad, lines);
        }
#pragma unroll
        for (int w = 0; w < MW; ++w) {
            const int k = tid + w * NT;
            if (k >= nslot) continue;
            const int f = k / NZD, m = k % NZD;
            double2 Dv, N, Mn;
            mf_slot<NZD, HALF, MODE, MODEL>(f, m, sh.line + f * NZ, sh.skz, sh.smf, ph, a.mfn_prev != nullptr,
                                            sh.sprev[k], linC, nlc1, nlc2, Dv, N, Mn);
            a.drive[mo + k] = Dv;
            if (hd) sh.sdrive[k] = Dv;
            a.mfn_out[mo + k] = N;
            if (MODE == MODE_RHS) continue;
            a.mf[mo + k] = Mn;
            sh.smfn[k] = Mn;
        }
        __syncthreads();
        if (hd) drive_means(sh.sdrive);
        if (MODE == MODE_RHS) return;
    } else {
        for (int k = tid; k < nslot; k += NT) sh.smfn[k] = sh.smf[k];
        __syncthreads();
        if (hd) drive_means(a.drive + mo);  // the last rhs's drive
    }
    // real-space profiles for the next rhs (MHD_FullQlin.cpp:180-198,
    // MHD_BQlin.cpp:122-127): fac1D * ifft(pad(...)), real part; for
    // MHD_BQlin line 3 is Bx, which only the CFL condition reads
    // (MHD_BQlin.cpp:546-548)
    const int nlines = (MODEL == 0) ? 8 : 3;
    const double fac = a.fac;
    const int model = MODEL;
    const auto padded = [&](int l, int j) {
        double2 v = make_double2(0.0, 0.0);
        if (D::kept(j)) {
            const int m = D::slot(j);
            const double kz = sh.skz[m];
            if (model == 0) {
                const double2 M = sh.smfn[(l >> 1) * NZD + m];
                v = (l & 1) ? imul(fac * kz, M) : cscale(fac, M);
            } else if (l < 4) {
                const double2 M = sh.smfn[(l == 3 ? 0 : NZD) + m];  // By, or Bx for line 3
                if (l == 0 || l == 3) v = cscale(fac, M);
                else if (l == 1) v = imul(fac * kz, M);
                else v = cscale(fac * (-(kz * kz)), M);
            }
        }
        return v;
    };
    double *Z = a.mfz + (size_t)run * 8 * NZ;
    const auto out = [&](int l, int i, double2 v) {
        sh.line[l * NZ + i] = v;
        if (l < nlines) Z[l * NZ + i] = v.x;
    };
    fft_pass_io<NZ, 8, true, PL::R1, 1, false>(sh.tw, padded, pad);
    fft_pass_io<NZ, 8, true, PL::R2, PL::R1, false>(sh.tw, pad, out);
    // max |profile| per line, sqrt(max |z|^2) as Eigen's abs2().maxCoeff()
    // (MHD_FullQlin.cpp:773-776): 32 threads per line, fixed order
    if (rec) {
        static_assert(NT == 8 * 32, "one 32-lane group per profile line");
        const int l = tid / 32, j0 = tid % 32;
        double m2 = 0.0;
        for (int j = j0; j < NZ; j += 32) {
            const double2 v = sh.line[l * NZ + j];
            m2 = fmax(m2, rn_add(rn_mul(v.x, v.x), rn_mul(v.y, v.y)));
        }
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) m2 = fmax(m2, __shfl_xor(m2, o, 32));
        if (j0 == 0) pm[PM_CUR + l] = sqrt(m2);
    }
    if (hd) {
        // |MUx|^2 + |MUy|^2 summed over z (HD_fullU.cpp:618) and mean |kz^2 Ux|
        // of the current state (:619)
        double ex = 0.0, ey = 0.0, su = 0.0;
        for (int j = tid; j < NZ; j += NT) {
            const double2 x = sh.line[4 * NZ + j], y = sh.line[6 * NZ + j];
            ex += x.x * x.x + x.y * x.y;
            ey += y.x * y.x + y.y * y.y;
        }
        for (int m = tid; m < NZD; m += NT) {
            const double k2 = -(sh.skz[m] * sh.skz[m]);
            const double2 u = sh.smfn[2 * NZD + m];
            su += hypot(u.x * k2, u.y * k2);
        }
        ex = block_sum(ex, sh.red);
        ey = block_sum(ey, sh.red);
        su = block_sum(su, sh.red);
        if (tid == 0) {
            pm[PM_E_CUR] = ex + ey;
            pm[PM_UX] = su / NZD;
        }
    }
    if (REC && a.vdt && a.advance) {
        __syncthreads();
        // the run's next step dt (RK3CN_Elin.cpp:76-80)
        if (tid == 0) a.clock[run].dt = cfl_dt(pm, a.runs[run], MODEL);
    }
}

// one workgroup per run (blockIdx.x); the substep mode, the model and
// whether CFL quantities are recorded or variable steps taken (REC) are
// template parameters, so each instantiation carries only its own code
// (which the one workgroup fetches once: at 64x128x128 this took the column
// reduction and mean fields from 13.4 to 12.0 us per substep)
template <int NZ, int MODE, int MODEL, bool REC>
__global__ __launch_bounds__(NT) void k_meanfield(MFArgs a)
{
    __shared__ MFShared<NZ> sh;
    meanfield_load<NZ, MODE, MODEL>(a, blockIdx.x, sh);
    meanfield_body<NZ, MODE, MODEL, REC>(a, blockIdx.x, sh);
}

// ---------------------------------------------------------------------------
// The substep tail in one launch: one workgroup per (mean field f, run), grid
// (nMF, runs).  It sums its field's NZ columns of the run's partial rows in
// the order of k_reduce_cols (RED_CH row chunks per column, each summed with
// Neumaier compensation, the chunks then combined in order: the same bits),
// transforms the line, forms the field's drive, RHS and update (mf_slot, as
// k_meanfield) and the field's real-space profile lines.  The couplings read
// the other fields at the substep start, which other workgroups of the launch
// update, so the update is written to another buffer than the one read
// (a.mf_in -> a.mf).  Substep updates with fixed steps and no CFL records only
// (the profile maxima and HD_fullU's records combine the fields).
// ---------------------------------------------------------------------------
constexpr int MFF_T = 1024;          // threads per workgroup
constexpr int RED_CH = NT / RED_CW;  // row chunks per column, as k_reduce_cols

template <int NZ>
struct MFFShared {
    static constexpr int NZD = Dealias<NZ>::nz;
    double red[RED_CH][NZ], red2[RED_CH][NZ];  // chunk sums and their compensations
    double ssum[NZ];                           // the field's summed z-profile
    double2 pad[3 * PadIO<NZ, double2>::LP];   // between the two passes of a transform
    double2 line[NZ];                          // the field's stress / EMF spectrum
    double2 tw[NZ];
    double skz[NZD];
    double2 smf[4 * NZD];  // every mean field at the substep start
    double2 sprev[NZD];    // the field's previous substep RHS
    double2 smfn[NZD];     // the field's updated mean field
};

template <int NZ, int MODE, int MODEL>
__global__ __launch_bounds__(MFF_T) void k_mf_fused(MFArgs a, const double *__restrict__ part)
{
    static_assert(MODE == MODE_ELIN || MODE == MODE_CN, "substep updates");
    __shared__ MFFShared<NZ> sh;
    using D = Dealias<NZ>;
    using PL = Plan2<NZ>;
    constexpr int NZD = D::nz, HALF = D::half;
    constexpr int NPAIR = NZ * RED_CH;                    // (column, chunk) pairs of the field
    constexpr int NP = (NPAIR + MFF_T - 1) / MFF_T;       // pairs per thread
    constexpr int B = 16;                                 // rows of a pair in flight
    static_assert(4 * NZD <= MFF_T && NZ <= MFF_T, "one table entry per thread");
    const int f = blockIdx.x, run = blockIdx.y, tid = threadIdx.x;
    const RunParams &rp = a.runs[run];
    const Phys ph = rp.ph;
    const int nMF = NMF_OF(MODEL), ncols = nMF * NZ, nslot = nMF * NZD;
    const size_t mo = (size_t)run * nslot;
    const double2 *mf_in = a.mf_in ? a.mf_in : a.mf;
    // tables first, so that their loads are in flight with the rows'
    const double2 twv = tid < NZ ? a.tw[tid] : make_double2(0.0, 0.0);
    const double kzv = tid < NZD ? a.kzv[tid] : 0.0;
    const double2 mv = tid < nslot ? mf_in[mo + tid] : make_double2(0.0, 0.0);
    const double2 pv = (tid < NZD && a.mfn_prev) ? a.mfn_prev[mo + (size_t)f * NZD + tid] : make_double2(0.0, 0.0);
    // the field's columns, every pair of the thread advancing together
    const int nrows = rp.nwg + rp.nwg2, per = (nrows + RED_CH - 1) / RED_CH;
    double s[NP], e[NP];
    int r0[NP], r1[NP], col[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const int pq = tid + q * MFF_T;
        const bool act = NPAIR % MFF_T == 0 || pq < NPAIR;
        const int z = pq % NZ, chunk = pq / NZ;
        col[q] = f * NZ + z;
        r0[q] = chunk * per;
        r1[q] = act ? min(nrows, r0[q] + per) : r0[q];
        s[q] = e[q] = 0.0;
    }
    for (int j = 0; j < per; j += B) {
        double v[NP][B];
#pragma unroll
        for (int q = 0; q < NP; ++q)
#pragma unroll
            for (int k = 0; k < B; ++k) {
                const int r = r0[q] + j + k;
                v[q][k] = r < r1[q] ? part[(size_t)run_row(rp, r) * ncols + col[q]] : 0.0;
            }
#pragma unroll
        for (int q = 0; q < NP; ++q)
#pragma unroll
            for (int k = 0; k < B; ++k) {
                const double t = s[q] + v[q][k];
                e[q] += (fabs(s[q]) >= fabs(v[q][k])) ? ((s[q] - t) + v[q][k]) : ((v[q][k] - t) + s[q]);
                s[q] = t;
            }
    }
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const int pq = tid + q * MFF_T;
        if (NPAIR % MFF_T == 0 || pq < NPAIR) {
            sh.red[pq / NZ][pq % NZ] = s[q];
            sh.red2[pq / NZ][pq % NZ] = e[q];
        }
    }
    if (tid < NZ) sh.tw[tid] = twv;
    if (tid < NZD) {
        sh.skz[tid] = kzv;
        sh.sprev[tid] = pv;
    }
    if (tid < nslot) sh.smf[tid] = mv;
    if (a.advance && f == 0 && tid == 0) clock_advance(a.clock + run);
    __syncthreads();
    if (tid < NZ) {
        double t = 0.0, te = 0.0;
        for (int k = 0; k < RED_CH; ++k) {
            const double v = sh.red[k][tid];
            const double u = t + v;
            te += (fabs(t) >= fabs(v)) ? ((t - u) + v) : ((v - u) + t);
            te += sh.red2[k][tid];
            t = u;
        }
        sh.ssum[tid] = t + te;
    }
    __syncthreads();
    const PadIO<NZ, double2> pad{sh.pad};
    {
        const double *S = sh.ssum;
        const auto ld = [S](int, int j) { return make_double2(S[j], 0.0); };
        fft_pass_io<NZ, 1, false, PL::R1, 1, false, MFF_T>(sh.tw, ld, pad);
        fft_pass_io<NZ, 1, false, PL::R2, PL::R1, false, MFF_T>(sh.tw, pad, BufIO<NZ, double2>{sh.line});
    }
    if (tid < NZD) {
        double2 Dv, N, Mn;
        mf_slot<NZD, HALF, MODE, MODEL>(f, tid, sh.line, sh.skz, sh.smf, ph, a.mfn_prev != nullptr, sh.sprev[tid],
                                        a.linC, a.nlc1, a.nlc2, Dv, N, Mn);
        const size_t k = mo + (size_t)f * NZD + tid;
        a.drive[k] = Dv;
        a.mfn_out[k] = N;
        a.mf[k] = Mn;
        sh.smfn[tid] = Mn;
    }
    __syncthreads();
    // the field's real-space profile lines (k_meanfield's lines 2f, 2f + 1;
    // MHD_BQlin: By, dz By, dzz By from field 1, and field 0's Bx line, which
    // only the CFL records read, is not formed)
    if (MODEL == 0 || f == 1) {
        constexpr int NL = MODEL == 0 ? 2 : 3;
        const double fac = a.fac;
        const auto padded = [&](int l, int j) {
            double2 v = make_double2(0.0, 0.0);
            if (D::kept(j)) {
                const int m = D::slot(j);
                const double kz = sh.skz[m];
                const double2 M = sh.smfn[m];
                if (MODEL == 0) v = (l & 1) ? imul(fac * kz, M) : cscale(fac, M);
                else if (l == 0) v = cscale(fac, M);
                else if (l == 1) v = imul(fac * kz, M);
                else v = cscale(fac * (-(kz * kz)), M);
            }
            return v;
        };
        double *Z = a.mfz + (size_t)run * 8 * NZ + (size_t)(MODEL == 0 ? 2 * f : 0) * NZ;
        const auto out = [Z](int l, int i, double2 v) { Z[l * NZ + i] = v.x; };
        fft_pass_io<NZ, NL, true, PL::R1, 1, false, MFF_T>(sh.tw, padded, pad);
        fft_pass_io<NZ, NL, true, PL::R2, PL::R1, false, MFF_T>(sh.tw, pad, out);
    }
}

// ---------------------------------------------------------------------------
// ShearingBox_Remap (MHD_FullQlin.cpp:464-481) over modes [m0, m0 + nxy)
// ---------------------------------------------------------------------------
[[maybe_unused]] static __global__ __launch_bounds__(NT) void k_remap(double2 *u, double *kx0, const double *ky, int m0, int nxy, int nz,
                                              double qt, double thr, double shift)
{
    for (int i = m0 + blockIdx.x; i < m0 + nxy; i += gridDim.x) {
        const double k0 = kx0[i], kyv = ky[i];
        const bool hit = remap_hit(qt, thr, k0, kyv);
        __syncthreads();
        if (hit) {
            for (int k = threadIdx.x; k < 4 * nz; k += NT) u[(size_t)i * 4 * nz + k] = make_double2(0.0, 0.0);
            if (threadIdx.x == 0) kx0[i] = rn_sub(k0, shift);
        }
    }
}

// ---------------------------------------------------------------------------
// Calc_Energy_AM_Diss (MHD_FullQlin.cpp:515-559): per-workgroup partial sums
// of energy, angular momentum and dissipation over a run's modes, fixed order
// ---------------------------------------------------------------------------
struct DiagArgs {
    const double2 *u;
    const double *kx0, *ky, *kzv;
    int m0, nxy, nz;
    double qt, nu, eta;
    double *part;  // [gridDim.x][6]
};

[[maybe_unused]] static __global__ __launch_bounds__(NT) void k_diag_modes(DiagArgs a)
{
    __shared__ double sred[6][NT];
    double s[6] = {0, 0, 0, 0, 0, 0};
    const int nz = a.nz;
    for (int i = a.m0 + blockIdx.x; i < a.m0 + a.nxy; i += gridDim.x) {
        const double ky = a.ky[i];
        const double kx = rn_add(a.kx0[i], rn_mul(a.qt, ky));
        const double mult = (ky == 0.0) ? 1.0 : 2.0;
        for (int m = threadIdx.x; m < nz; m += NT) {
            const double kz = a.kzv[m], kz2 = -(kz * kz);
            const double lap2 = (ky == 0.0 && m == 0) ? 0.0 : (-ky * ky + kz2);
            const double ilap2 = (ky == 0.0 && m == 0) ? 1.0 : 1.0 / lap2;
            const double lapF = -kx * kx + lap2;
            const double l2t = lapF * ilap2;
            const double2 *U = a.u + (size_t)i * 4 * nz;
            const double2 ux = U[m], ze = U[nz + m], bx = U[2 * nz + m], et = U[3 * nz + m];
            const double eu = l2t * redot(ux, ux) - ilap2 * redot(ze, ze);
            const double eb = l2t * redot(bx, bx) - ilap2 * redot(et, et);
            const double sy = ky * kx;
            const double2 fy = make_double2((sy * ux.x - kz * ze.y) * ilap2, (sy * ux.y + kz * ze.x) * ilap2);
            const double2 by = make_double2((sy * bx.x - kz * et.y) * ilap2, (sy * bx.y + kz * et.x) * ilap2);
            s[0] += mult * eu;
            s[1] += mult * eb;
            s[2] += mult * redot(ux, fy);
            s[3] += mult * redot(bx, by);
            s[4] += mult * a.nu * (-lapF * eu);
            s[5] += mult * a.eta * (-lapF * eb);
        }
    }
    for (int c = 0; c < 6; ++c) sred[c][threadIdx.x] = s[c];
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 6; ++c) sred[c][threadIdx.x] += sred[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) a.part[blockIdx.x * 6 + threadIdx.x] = sred[threadIdx.x][0];
}

struct DiagFinalArgs {
    const double *part;
    int nparts;
    const double2 *mf, *drive;
    const double *kzv;
    int nz, nMF, model, NZ;
    double Lz, totalN2, nu, eta, kzfac;
    double *out;  // [18]
};

// MHD_FullQlin.cpp:560-676 (single thread: a few hundred terms)
[[maybe_unused]] static __global__ void k_diag_final(DiagFinalArgs a)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double rb[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < a.nparts; ++b)
        for (int c = 0; c < 6; ++c) rb[c] += a.part[b * 6 + c];
    const double divfac = a.Lz / a.totalN2;
    const double eu = rb[0] * divfac, eb = rb[1] * divfac, amu = rb[2] * divfac, amb = rb[3] * divfac;
    const double du = rb[4] * divfac, db = rb[5] * divfac;
    const double divMF = a.Lz / ((double)a.NZ * a.NZ);
    const int nz = a.nz;
    double s[4] = {0, 0, 0, 0}, d[4] = {0, 0, 0, 0}, amB = 0, amU = 0;
    for (int f = 0; f < a.nMF; ++f)
        for (int m = 0; m < nz; ++m) {
            const double2 v = a.mf[f * nz + m];
            const double kz = a.kzv[m];
            s[f] += redot(v, v);
            d[f] += -(kz * kz) * redot(v, v);
        }
    for (int m = 0; m < nz; ++m) {
        amB += redot(a.mf[m], a.mf[nz + m]);
        if (a.nMF == 4) amU += redot(a.mf[2 * nz + m], a.mf[3 * nz + m]);
    }
    const double eMB = (s[0] + s[1]) * divMF, eMU = (s[2] + s[3]) * divMF;
    const double dMB = (divMF * a.eta) * (-d[0] - d[1]), dMU = (divMF * a.nu) * (-d[2] - d[3]);
    double *o = a.out;
    o[0] = eMU / 2; o[1] = eMB / 2; o[2] = eu / 2; o[3] = eb / 2;
    o[4] = divMF * amU; o[5] = divMF * amB; o[6] = amu; o[7] = amb;
    o[8] = dMU; o[9] = dMB; o[10] = du; o[11] = db;
    for (int f = 0; f < 4; ++f) o[12 + f] = 0.0;
    if (a.model == 0) {
        for (int f = 0; f < 4; ++f) {
            double num = 0, den = 0;
            for (int m = 0; m < nz; ++m) {
                num += redot(a.drive[f * nz + m], a.mf[f * nz + m]);
                den += redot(a.mf[f * nz + m], a.mf[f * nz + m]);
            }
            o[12 + f] = num / sqrt(den);
        }
    }
    o[16] = pow(a.nu * a.nu * a.nu / (dMU + du), 0.25) * a.kzfac;
    o[17] = pow(a.eta * a.eta * a.eta / (dMB + db), 0.25) * a.kzfac;
}

// TimeVariables::Save_Mean_Fields (TimeVariables.cpp:202-228)
template <int NZ>
__global__ __launch_bounds__(NT) void k_mf_real(const double2 *mf, int nMF, const double2 *twg, double fac,
                                                double *out)
{
    using D = Dealias<NZ>;
    __shared__ double2 buf[4 * NZ];
    __shared__ double2 tw[NZ];
    for (int k = threadIdx.x; k < NZ; k += NT) tw[k] = twg[k];
    for (int idx = threadIdx.x; idx < 4 * NZ; idx += NT) {
        const int l = idx / NZ, j = idx % NZ;
        buf[idx] = (l < nMF && D::kept(j)) ? mf[l * D::nz + D::slot(j)] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lines<NZ, 4, true>(buf, tw);
    for (int idx = threadIdx.x; idx < nMF * NZ; idx += NT) out[idx] = fac * buf[idx].x;
}

// solution::Check_Solution (solution.cpp:169-203) reductions over one run's
// mean fields: out[0] = sum |Re ifft(pad(By))|, out[1] = sum |Im ...| (the
// transform unnormalised, as fftwPlans::inverse), out[2] = max |By|,
// out[3] = max |Bx|, out[4] = sum |By| over the dealiased coefficients.
// One workgroup; sums in a fixed order.
template <int NZ>
__global__ __launch_bounds__(NT) void k_check_sums(const double2 *mf, const double2 *twg, double *out)
{
    using D = Dealias<NZ>;
    __shared__ double2 buf[NZ];
    __shared__ double2 tw[NZ];
    __shared__ double red[5][NT];
    const int tid = threadIdx.x;
    for (int k = tid; k < NZ; k += NT) tw[k] = twg[k];
    for (int j = tid; j < NZ; j += NT) buf[j] = D::kept(j) ? mf[D::nz + D::slot(j)] : make_double2(0.0, 0.0);
    __syncthreads();
    fft_lines<NZ, 1, true>(buf, tw);
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int j = tid; j < NZ; j += NT) {
        s[0] += fabs(buf[j].x);
        s[1] += fabs(buf[j].y);
    }
    for (int m = tid; m < D::nz; m += NT) {
        const double2 by = mf[D::nz + m], bx = mf[m];
        const double aby = hypot(by.x, by.y);
        s[2] = fmax(s[2], aby);
        s[3] = fmax(s[3], hypot(bx.x, bx.y));
        s[4] += aby;
    }
    for (int c = 0; c < 5; ++c) red[c][tid] = s[c];
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (tid < w) {
            red[0][tid] += red[0][tid + w];
            red[1][tid] += red[1][tid + w];
            red[2][tid] = fmax(red[2][tid], red[2][tid + w]);
            red[3][tid] = fmax(red[3][tid], red[3][tid + w]);
            red[4][tid] += red[4][tid + w];
        }
        __syncthreads();
    }
    if (tid < 5) out[tid] = red[tid][0];
}

}  // namespace qlk
