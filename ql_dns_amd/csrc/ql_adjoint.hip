// ql_adjoint.hip -- adjoint (reverse-mode gradient) of n steps of the QL main
// loop (main.cpp:108-123: Step, t += dt, ShearingBox_Remap) for one
// MHD_FullQlin or MHD_BQlin run in FP64, with checkpointed recomputation.
//
// ql_adjoint_forward runs the steps exactly as ql_steps does (the same kernels
// and launch shape, so the same bits) and stores the state every `every`
// steps.  ql_adjoint_backward walks the segments backwards: it restores a
// checkpoint, recomputes the segment's substeps eagerly while keeping every
// substep state (3 every + 1 states), then runs the reverse substeps
// (ql_adjoint.hpp) from the last to the first.  It returns dL/d(state at the
// start of the forward) and dL/d(q, omega, nu, eta, B0z) for the caller's
// dL/d(final state); the context is left at the final state of the forward.
#include <algorithm>
#include <cstdio>
#include <cstring>

#include "ql_adjoint.hpp"
#include "ql_runtime.hpp"

namespace qlh {

struct AdjState {
    int nsteps = 0, every = 0, nseg = 0;
    bool noise = false;  // the forward was driven
    double t_end = 0.0;
    // checkpoints at segment starts (index nseg: the final state); with
    // variable steps also the step taken last and the profile maxima the next
    // step's dt comes from
    struct Ck {
        double2 *u = nullptr, *mf = nullptr;
        double *kx0 = nullptr, *pm = nullptr;
        double t = 0.0, t_prev = 0.0, dt_last = 0.0;
        int old_init = 0;
    };
    std::vector<Ck> ck;
    double t_init = 0.0;
    // substep states of one segment; with driving noise also each step's
    // state after substep 3 before the noise and the remap
    double2 *st_u = nullptr, *st_mf = nullptr, *st_pre = nullptr;
    double *st_kx0 = nullptr;
    // adjoint buffers
    double2 *gu = nullptr, *gnp = nullptr, *gM = nullptr, *gnmp = nullptr;
    double *gsums = nullptr, *gpart = nullptr, *grows = nullptr, *gpar_mf = nullptr, *gpar = nullptr;
    double *mfz = nullptr;
    qlk::RunParams *rp1 = nullptr;  // one run, rows [0, W) (column reduction)
    int W = 0;
};

void adj_release(Batch &b)
{
    AdjState *A = b.adj;
    if (!A) return;
    for (auto &c : A->ck) dfree(c.u), dfree(c.mf), dfree(c.kx0), dfree(c.pm);
    dfree(A->st_u), dfree(A->st_mf), dfree(A->st_kx0), dfree(A->st_pre);
    dfree(A->gu), dfree(A->gnp), dfree(A->gM), dfree(A->gnmp);
    dfree(A->gsums), dfree(A->gpart), dfree(A->grows), dfree(A->gpar_mf), dfree(A->gpar), dfree(A->mfz);
    dfree(A->rp1);
    delete A;
    b.adj = nullptr;
}

namespace {

constexpr int ADJ_GMAX = 1024;

int adj_line_mb(int NZ) { return NZ >= 128 ? 1 : 128 / NZ; }

int check_adjoint(const Batch &b, const char *who)
{
    if (b.runs.size() != 1) return fail(QL_ERR_UNSUPPORTED, "%s: the adjoint is for single-run contexts", who);
    if (b.model != QL_MODEL_MHD_FULLQLIN && b.model != QL_MODEL_MHD_BQLIN && !b.hd)
        return fail(QL_ERR_UNSUPPORTED, "%s: the adjoint is implemented for MHD_FullQlin, MHD_BQlin and HD_fullU", who);
    if (b.precision != QL_PREC_FP64) return fail(QL_ERR_UNSUPPORTED, "%s: the adjoint runs in FP64 only", who);
    return QL_OK;
}

int save_ck(Batch &b, AdjState::Ck &c, double t)
{
    Group &g = b.groups[0];
    const size_t nl = (size_t)g.modes * 4 * g.nz, nm = (size_t)b.nMF * g.nz;
    if (!c.u) {
        QL_RC(dalloc(&c.u, nl));
        QL_RC(dalloc(&c.mf, nm));
        QL_RC(dalloc(&c.kx0, g.modes));
        QL_RC(dalloc(&c.pm, (size_t)qlk::PM_STRIDE));
    }
    QL_HIP(hipMemcpyAsync(c.u, g.u, sizeof(double2) * nl, hipMemcpyDeviceToDevice, b.s));
    QL_HIP(hipMemcpyAsync(c.mf, g.mf, sizeof(double2) * nm, hipMemcpyDeviceToDevice, b.s));
    QL_HIP(hipMemcpyAsync(c.kx0, g.kx0, sizeof(double) * g.modes, hipMemcpyDeviceToDevice, b.s));
    QL_HIP(hipMemcpyAsync(c.pm, g.pm, sizeof(double) * qlk::PM_STRIDE, hipMemcpyDeviceToDevice, b.s));
    c.t = t;
    c.t_prev = b.t_prev;
    c.dt_last = b.dt_last;
    c.old_init = b.old_init;
    return QL_OK;
}

int load_ck(Batch &b, const AdjState::Ck &c)
{
    Group &g = b.groups[0];
    const size_t nl = (size_t)g.modes * 4 * g.nz, nm = (size_t)b.nMF * g.nz;
    QL_HIP(hipMemcpyAsync(g.u, c.u, sizeof(double2) * nl, hipMemcpyDeviceToDevice, b.s));
    QL_HIP(hipMemcpyAsync(g.mf, c.mf, sizeof(double2) * nm, hipMemcpyDeviceToDevice, b.s));
    QL_HIP(hipMemcpyAsync(g.kx0, c.kx0, sizeof(double) * g.modes, hipMemcpyDeviceToDevice, b.s));
    QL_HIP(hipMemcpyAsync(g.pm, c.pm, sizeof(double) * qlk::PM_STRIDE, hipMemcpyDeviceToDevice, b.s));
    b.t_prev = c.t_prev;
    b.dt_last = c.dt_last;
    b.old_init = c.old_init;
    return refresh_profiles(b, g, b.s);
}

// substep times of the forward (sub_times in ql_kernels.hpp, same roundings);
// dt_last, the step before (equal to dt with fixed steps), places the RK3CN
// "old" operator of substep 1
struct HostTimes {
    double ts, dtl, tso, dtlo, tnew;
    int oinit;
};
HostTimes host_times(double t, double t_prev, int old_init, double t_init, double dt, double dt_last, int s)
{
    HostTimes r{};
    const double dtP1 = dt * P1, dtP2 = dt * P2;
    const double dtl1 = dt * P1, dtl2 = dt * (P2 - P1), dtl3 = dt * (1.0 - P2);
    const double dtP2o = dt_last * P2, dtl3o = dt_last * (1.0 - P2);
    r.tnew = t + dt;
    if (s == 1) {
        r.ts = t;
        r.dtl = dtl1;
        r.oinit = old_init;
        r.tso = old_init ? t_init : t_prev + dtP2o;
        r.dtlo = old_init ? 0.0 : dtl3o;
    } else if (s == 2) {
        r.ts = t + dtP1;
        r.dtl = dtl2;
        r.tso = t;
        r.dtlo = dtl1;
    } else {
        r.ts = t + dtP2;
        r.dtl = dtl3;
        r.tso = t + dtP1;
        r.dtlo = dtl2;
    }
    return r;
}

template <int NZ, int MODEL>
void launch_adj_nz(const qlk::AdjModeArgs &am, const qlk::AdjMFArgs &aa, const qlk::AdjMF2Args &ab, int W,
                   const double *gpart, const qlk::RunParams *rp1, double *grows, hipStream_t st, int stage)
{
    constexpr int MB = NZ >= 128 ? 1 : 128 / NZ;
    using P = typename std::conditional<MODEL == 0, qlk::ModelFull, qlk::ModelBQ>::type;
    if (stage == 0) {
        qlk::k_adj_mf_a<NZ, MODEL><<<1, qlk::NT, 0, st>>>(aa);
    } else if (stage == 1) {
        qlk::k_adj_mode<NZ, MB, P><<<W, qlk::NT, 0, st>>>(am);
        const int ncols = 8 * NZ + qlk::NGP;
        qlk::k_reduce_cols<<<dim3((ncols + qlk::RED_CW - 1) / qlk::RED_CW, 1), qlk::NT, 0, st>>>(gpart, rp1, ncols, grows);
    } else {
        qlk::k_adj_mf_b<NZ, MODEL><<<1, qlk::NT, 0, st>>>(ab);
    }
}

template <int MODEL>
void launch_adj_m(int NZ, const qlk::AdjModeArgs &am, const qlk::AdjMFArgs &aa, const qlk::AdjMF2Args &ab, int W,
                  const double *gpart, const qlk::RunParams *rp1, double *grows, hipStream_t st, int stage)
{
    switch (NZ) {
    case 32: launch_adj_nz<32, MODEL>(am, aa, ab, W, gpart, rp1, grows, st, stage); break;
    case 64: launch_adj_nz<64, MODEL>(am, aa, ab, W, gpart, rp1, grows, st, stage); break;
    case 128: launch_adj_nz<128, MODEL>(am, aa, ab, W, gpart, rp1, grows, st, stage); break;
    default: launch_adj_nz<256, MODEL>(am, aa, ab, W, gpart, rp1, grows, st, stage); break;
    }
}

void launch_adj(int NZ, int model, const qlk::AdjModeArgs &am, const qlk::AdjMFArgs &aa, const qlk::AdjMF2Args &ab,
                int W, const double *gpart, const qlk::RunParams *rp1, double *grows, hipStream_t st, int stage)
{
    if (model == QL_MODEL_MHD_BQLIN) launch_adj_m<1>(NZ, am, aa, ab, W, gpart, rp1, grows, st, stage);
    else launch_adj_m<0>(NZ, am, aa, ab, W, gpart, rp1, grows, st, stage);
}

int adjoint_forward(Batch &b, int nsteps, int every, double *t_inout)
{
    QL_RC(check_adjoint(b, "ql_adjoint_forward"));
    if (nsteps < 1 || every < 1) return fail(QL_ERR_ARG, "ql_adjoint_forward: nsteps=%d, every=%d", nsteps, every);
    adj_release(b);
    AdjState *A = new AdjState();
    b.adj = A;
    A->nsteps = nsteps;
    A->every = every;
    A->nseg = (nsteps + every - 1) / every;
    A->noise = b.groups[0].noise;
    A->t_init = b.t_init;
    A->ck.resize(A->nseg + 1);
    double t = *t_inout;
    for (int sg = 0; sg < A->nseg; ++sg) {
        QL_RC(save_ck(b, A->ck[sg], t));
        const int n = std::min(every, nsteps - sg * every);
        QL_RC(run_steps(b, n, &t));
    }
    QL_RC(save_ck(b, A->ck[A->nseg], t));
    QL_HIP(hipStreamSynchronize(b.s));
    A->t_end = t;
    *t_inout = t;
    return QL_OK;
}

int adjoint_backward(Batch &b, const double *g_lin, const double *g_mf, double *g_lin0, double *g_mf0,
                     double *g_par)
{
    QL_RC(check_adjoint(b, "ql_adjoint_backward"));
    AdjState *A = b.adj;
    if (!A) return fail(QL_ERR_STATE, "ql_adjoint_backward: no ql_adjoint_forward on this context");
    if (A->noise != b.groups[0].noise)
        return fail(QL_ERR_STATE, "ql_adjoint_backward: the driving noise changed since ql_adjoint_forward");
    Group &g = b.groups[0];
    const RunInfo &r = b.runs[0];
    const int NZ = g.NZ, nz = g.nz, nxy = r.nxy, K = A->every;
    const size_t nl = (size_t)nxy * 4 * nz, nm = (size_t)b.nMF * nz;
    const int MB = adj_line_mb(NZ);
    const int W = std::min((nxy + MB - 1) / MB, ADJ_GMAX);
    const int ncols = 8 * NZ + qlk::NGP;
    if (!A->gu) {
        QL_RC(dalloc(&A->st_u, nl * (3 * K + 1)));
        QL_RC(dalloc(&A->st_mf, nm * (3 * K + 1)));
        QL_RC(dalloc(&A->st_kx0, (size_t)nxy * K));
        if (g.noise) QL_RC(dalloc(&A->st_pre, nl * K));
        QL_RC(dalloc(&A->gu, nl));
        QL_RC(dalloc(&A->gnp, nl));
        QL_RC(dalloc(&A->gM, nm));
        QL_RC(dalloc(&A->gnmp, nm));
        QL_RC(dalloc(&A->gsums, (size_t)4 * NZ));
        QL_HIP(hipMemset(A->gsums, 0, sizeof(double) * 4 * NZ));  // MHD_BQlin fills 2 of the 4 lines
        QL_RC(dalloc(&A->gpart, (size_t)W * ncols));
        QL_RC(dalloc(&A->grows, (size_t)ncols));
        QL_RC(dalloc(&A->gpar_mf, (size_t)qlk::NGP));
        QL_RC(dalloc(&A->gpar, (size_t)qlk::NGP));
        QL_RC(dalloc(&A->mfz, (size_t)8 * NZ));
        // MHD_BQlin writes 3 of the 8 profile lines; the others stay zero
        QL_HIP(hipMemset(A->mfz, 0, sizeof(double) * 8 * NZ));
        QL_RC(dalloc(&A->rp1, 1));
        qlk::RunParams rp = g.h_rp[0];
        rp.wg_off = 0;
        rp.nwg = W;
        QL_HIP(hipMemcpy(A->rp1, &rp, sizeof(rp), hipMemcpyHostToDevice));
        A->W = W;
    }
    // HD_fullU runs the MHD_FullQlin kernels with b = B = 0 on the internal
    // four-field layout: its caller's (ux, zeta) and (Ux, Uy) gradients sit at
    // fields 0, 1 and mean-field slots 2, 3, the others are zero
    std::vector<double2> hl, hm;
    if (b.hd) {
        hl.assign(nl, make_double2(0.0, 0.0));
        hm.assign(nm, make_double2(0.0, 0.0));
        if (g_lin)
            for (int i = 0; i < nxy; ++i)
                std::memcpy(&hl[(size_t)i * 4 * nz], g_lin + (size_t)i * 2 * nz * 2, sizeof(double2) * 2 * nz);
        if (g_mf) std::memcpy(&hm[(size_t)2 * nz], g_mf, sizeof(double2) * 2 * nz);
        QL_HIP(hipMemcpy(A->gu, hl.data(), sizeof(double2) * nl, hipMemcpyHostToDevice));
        QL_HIP(hipMemcpy(A->gM, hm.data(), sizeof(double2) * nm, hipMemcpyHostToDevice));
    } else {
        if (g_lin) QL_HIP(hipMemcpyAsync(A->gu, g_lin, sizeof(double2) * nl, hipMemcpyHostToDevice, b.s));
        else QL_HIP(hipMemsetAsync(A->gu, 0, sizeof(double2) * nl, b.s));
        if (g_mf) QL_HIP(hipMemcpyAsync(A->gM, g_mf, sizeof(double2) * nm, hipMemcpyHostToDevice, b.s));
        else QL_HIP(hipMemsetAsync(A->gM, 0, sizeof(double2) * nm, b.s));
    }
    QL_HIP(hipMemsetAsync(A->gnp, 0, sizeof(double2) * nl, b.s));
    QL_HIP(hipMemsetAsync(A->gnmp, 0, sizeof(double2) * nm, b.s));
    QL_HIP(hipMemsetAsync(A->gpar, 0, sizeof(double) * qlk::NGP, b.s));

    const bool cn = (b.integrator == QL_INT_RK3CN);
    for (int sg = A->nseg - 1; sg >= 0; --sg) {
        const int n = std::min(K, A->nsteps - sg * K);
        // recompute the segment, keeping every substep state
        QL_RC(load_ck(b, A->ck[sg]));
        double t = A->ck[sg].t;
        std::vector<double> st_t(n), st_tp(n), st_dt(n), st_dtl(n);
        std::vector<int> st_oi(n);
        QL_HIP(hipMemcpyAsync(A->st_u, g.u, sizeof(double2) * nl, hipMemcpyDeviceToDevice, b.s));
        QL_HIP(hipMemcpyAsync(A->st_mf, g.mf, sizeof(double2) * nm, hipMemcpyDeviceToDevice, b.s));
        for (int k = 0; k < n; ++k) {
            st_t[k] = t;
            st_tp[k] = b.t_prev;
            st_oi[k] = b.old_init;
            QL_HIP(hipMemcpyAsync(A->st_kx0 + (size_t)k * nxy, g.kx0, sizeof(double) * nxy, hipMemcpyDeviceToDevice,
                                  b.s));
            QL_RC(set_clock(b, t));
            // with variable steps the device chose the step from the profile
            // maxima, as in the forward (ql_step); the reverse takes it as given
            double dt = b.dt;
            if (b.vdt) {
                qlk::Clock ck{};
                QL_RC(read_clock(b, &ck));
                dt = ck.dt;
            }
            st_dt[k] = dt;
            st_dtl[k] = b.vdt ? b.dt_last : dt;
            for (int s = 1; s <= 3; ++s) {
                // with noise the step's remap follows the noise (main.cpp:111-123)
                QL_RC(enqueue_substep(b, g, s, s == 3 && !g.noise, false, b.s));
                if (s == 3 && g.noise) {
                    QL_HIP(hipMemcpyAsync(A->st_pre + (size_t)k * nl, g.u, sizeof(double2) * nl,
                                          hipMemcpyDeviceToDevice, b.s));
                    QL_RC(launch_noise(g, false, t, dt, true, r.p.remap != 0, b.s));
                }
                const size_t j = (size_t)3 * k + s;
                QL_HIP(hipMemcpyAsync(A->st_u + j * nl, g.u, sizeof(double2) * nl, hipMemcpyDeviceToDevice, b.s));
                QL_HIP(hipMemcpyAsync(A->st_mf + j * nm, mf_after_substep(b, g, s), sizeof(double2) * nm,
                                      hipMemcpyDeviceToDevice, b.s));
            }
            b.t_prev = t;
            b.dt_last = dt;
            b.old_init = 0;
            t = t + dt;
        }
        // reverse substeps
        for (int k = n - 1; k >= 0; --k) {
            for (int s = 3; s >= 1; --s) {
                const size_t jin = (size_t)3 * k + s - 1, jout = jin + 1;
                const double dt = st_dt[k];
                const HostTimes ht = host_times(st_t[k], st_tp[k], st_oi[k], A->t_init, dt, st_dtl[k], s);
                double lc, nlc1, nlc2 = 0.0;
                if (s == 1) {
                    lc = cn ? dt / 2 * (A0) : dt * (A0);
                    nlc1 = A0 * dt;
                } else if (s == 2) {
                    lc = cn ? dt / 2 * (A1 + B1) : dt * (A1 + B1);
                    nlc1 = A1 * dt;
                    nlc2 = B1 * dt;
                } else {
                    lc = cn ? dt / 2 * (A2 + B2) : dt * (A2 + B2);
                    nlc1 = A2 * dt;
                    nlc2 = B2 * dt;
                }
                // z-profiles of the substep's mean fields
                qlk::MFArgs pm = mf_args(b, g);
                pm.mode = qlk::MODE_PROFILES;
                pm.mf = A->st_mf + jin * nm;
                pm.mfz = A->mfz;
                pm.mfn_out = g.mfN1;
                pm.advance = 0;
                pm.vdt = pm.rec = 0;  // profiles only: no CFL records, no clock
                launch_mf(g, pm, b.s);
                QL_HIP(hipGetLastError());

                qlk::AdjMFArgs aa{};
                aa.gM = A->gM;
                aa.gnmp = A->gnmp;
                aa.M_in = A->st_mf + jin * nm;
                aa.M_out = A->st_mf + jout * nm;
                aa.kzv = g.kzv;
                aa.tw = g.tw;
                aa.gsums = A->gsums;
                aa.gpar = A->gpar_mf;
                aa.ph = g.h_rp[0].ph;
                aa.mode = cn ? qlk::MODE_CN : qlk::MODE_ELIN;
                aa.s = s;
                aa.linC = lc;
                aa.nlc1 = nlc1;
                aa.nlc2 = nlc2;

                qlk::AdjModeArgs am{};
                am.u_in = A->st_u + jin * nl;
                // the substep's own output: before the step's noise at s = 3
                am.u_out = (s == 3 && g.noise) ? A->st_pre + (size_t)k * nl : A->st_u + jout * nl;
                am.gu = A->gu;
                am.gnp = A->gnp;
                am.kx0 = A->st_kx0 + (size_t)k * nxy;
                am.ky = g.ky;
                am.kzv = g.kzv;
                am.tw = g.tw;
                am.mfz = A->mfz;
                am.gsums = A->gsums;
                am.gpart = A->gpart;
                am.ph = g.h_rp[0].ph;
                am.nxy = nxy;
                am.s = s;
                am.mode = aa.mode;
                am.do_remap = r.p.remap;
                am.oinit = ht.oinit;
                am.rsf = r.rsf;
                am.remap_thr = r.remap_thr;
                am.fac = 1.0 / NZ;
                am.linC = lc;
                am.nlc1 = nlc1;
                am.nlc2 = nlc2;
                am.ts = ht.ts;
                am.dtl = ht.dtl;
                am.tso = ht.tso;
                am.dtlo = ht.dtlo;
                am.tnew = ht.tnew;
                am.noise = g.noise ? 1 : 0;
                if (g.noise) am.np = g.h_np[0];
                am.t_step = st_t[k];
                am.dt_step = dt;

                qlk::AdjMF2Args ab{};
                ab.gM = A->gM;
                ab.rows = A->grows;
                ab.gpar_mf = A->gpar_mf;
                ab.gpar_total = A->gpar;
                ab.kzv = g.kzv;
                ab.tw = g.tw;
                ab.fac = 1.0 / NZ;
                for (int stage = 0; stage < 3; ++stage) {
                    launch_adj(NZ, b.model, am, aa, ab, A->W, A->gpart, A->rp1, A->grows, b.s, stage);
                    QL_HIP(hipGetLastError());
                }
            }
        }
    }
    // leave the context at the forward's final state
    QL_RC(load_ck(b, A->ck[A->nseg]));
    if (b.hd) {
        QL_HIP(hipStreamSynchronize(b.s));
        QL_HIP(hipMemcpy(hl.data(), A->gu, sizeof(double2) * nl, hipMemcpyDeviceToHost));
        QL_HIP(hipMemcpy(hm.data(), A->gM, sizeof(double2) * nm, hipMemcpyDeviceToHost));
        if (g_lin0)
            for (int i = 0; i < nxy; ++i)
                std::memcpy(g_lin0 + (size_t)i * 2 * nz * 2, &hl[(size_t)i * 4 * nz], sizeof(double2) * 2 * nz);
        if (g_mf0) std::memcpy(g_mf0, &hm[(size_t)2 * nz], sizeof(double2) * 2 * nz);
    } else {
        if (g_lin0) QL_HIP(hipMemcpyAsync(g_lin0, A->gu, sizeof(double2) * nl, hipMemcpyDeviceToHost, b.s));
        if (g_mf0) QL_HIP(hipMemcpyAsync(g_mf0, A->gM, sizeof(double2) * nm, hipMemcpyDeviceToHost, b.s));
    }
    double gp[qlk::NGP];
    QL_HIP(hipMemcpyAsync(gp, A->gpar, sizeof(double) * qlk::NGP, hipMemcpyDeviceToHost, b.s));
    QL_HIP(hipStreamSynchronize(b.s));
    if (b.hd) gp[3] = gp[4] = 0.0;  // HD_fullU has no eta and no B0z
    if (b.model == QL_MODEL_MHD_BQLIN) {
        // MHD_BQlin rotates at 2 q / 3 whatever omega is (MHD_BQlin.cpp:19): the
        // kernels' rotation gradient belongs to q
        gp[0] += 2.0 / 3.0 * gp[1];
        gp[1] = 0.0;
    }
    if (g_par)
        for (int c = 0; c < 5; ++c) g_par[c] = gp[c];
    return QL_OK;
}

}  // namespace
}  // namespace qlh

using namespace qlh;

extern "C" {

int ql_adjoint_forward(ql_ctx *c, int nsteps, int every, double *t_inout)
{
    if (!c || !t_inout) return fail(QL_ERR_ARG, "ql_adjoint_forward: null argument");
    return adjoint_forward(c->bt.b, nsteps, every, t_inout);
}

int ql_adjoint_backward(ql_ctx *c, const double *g_lin, const double *g_mf, double *g_lin0, double *g_mf0,
                        double g_par[5])
{
    if (!c) return fail(QL_ERR_ARG, "ql_adjoint_backward: null context");
    return adjoint_backward(c->bt.b, g_lin, g_mf, g_lin0, g_mf0, g_par);
}

}  // extern "C"
