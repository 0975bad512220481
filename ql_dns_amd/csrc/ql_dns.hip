// ql_dns.hip -- host side of the MI355X quasilinear time step: run sets,
// the per-substep launch sequence, HIP-graph replay of whole steps, the
// launch-shape tuner and the C ABI declared in include/ql_dns.h.
//
// Reference counterparts: the Model objects (Models/MHD_FullQlin.cpp,
// Models/MHD_BQlin.cpp) become kernels plus device-resident tables; the
// Integrator objects (Integrators/RK3CN.cpp, Integrators/RK3CN_Elin.cpp) become
// the substep sequence in enqueue_substep(); the MPI Allreduce of the stresses
// (MHD_FullQlin.cpp:386) becomes the on-device fixed-order reduction; the
// driver loop (main.cpp:108-123) becomes a replayed HIP graph.
//
// Every handle holds a set of runs (a context is a set of one).  Runs are
// grouped by NZ; a group's runs share one kernel instantiation and are laid
// out one after another in the group's arrays, each with its own parameters,
// mean fields and reductions.  Groups of different NZ run on their own streams,
// forked from and joined back into the handle's stream at every substep.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include <hip/hip_ext.h>

#include "ql_mode.hpp"
#include "ql_mode_wave.hpp"
#include "ql_runtime.hpp"

namespace qlh {

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

void drop_graphs(Batch &b)
{
    for (hipGraphExec_t &g : b.graphs) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

void release(Batch &b)
{
    adj_release(b);
    drop_graphs(b);
    for (Group &g : b.groups) {
        dfree(g.u), dfree(g.N1), dfree(g.N2), dfree(g.mf), dfree(g.mfN1), dfree(g.mfN2), dfree(g.drive);
        dfree(g.mfc1), dfree(g.mfc2);
        dfree(g.tw), dfree(g.tw32), dfree(g.kx0), dfree(g.ky), dfree(g.kzv), dfree(g.mfz), dfree(g.part);
        dfree(g.sums), dfree(g.lin_tmp), dfree(g.rp), dfree(g.wg);
        dfree(g.clock), dfree(g.pm), dfree(g.mode_run), dfree(g.np);
        dfree(g.wg2), dfree(g.amp), dfree(g.hi);
        if (g.join) (void)hipEventDestroy(g.join);
        if (g.s) (void)hipStreamDestroy(g.s);
        g.join = nullptr;
        g.s = nullptr;
    }
    dfree(b.dpart), dfree(b.dout), dfree(b.mfr_tmp);
    for (auto &e : b.ev)
        if (e) (void)hipEventDestroy(e), e = nullptr;
    for (hipEvent_t e : b.tpool) (void)hipEventDestroy(e);
    b.tpool.clear();
    b.tused = 0;
    if (b.fork) (void)hipEventDestroy(b.fork);
    if (b.s) (void)hipStreamDestroy(b.s);
    b.fork = nullptr;
    b.s = nullptr;
}

// waits for everything queued on the handle's streams
int sync_all(Batch &b)
{
    if (b.s) QL_HIP(hipStreamSynchronize(b.s));
    for (Group &g : b.groups)
        if (g.s) QL_HIP(hipStreamSynchronize(g.s));
    return QL_OK;
}

// ---------------------------------------------------------------------------
// validation and tables
// ---------------------------------------------------------------------------
int check_params(const ql_params *p, const char *who)
{
    if (p->model != QL_MODEL_MHD_FULLQLIN && p->model != QL_MODEL_MHD_BQLIN && p->model != QL_MODEL_HD_FULLU)
        return fail(QL_ERR_ARG, "%s: unknown model %d", who, p->model);
    if (p->integrator != QL_INT_RK3CN && p->integrator != QL_INT_RK3CN_ELIN)
        return fail(QL_ERR_ARG, "%s: unknown integrator %d", who, p->integrator);
    if (p->precision != QL_PREC_FP64 && p->precision != QL_PREC_MIXED)
        return fail(QL_ERR_UNSUPPORTED, "%s: precision %d is not available in this build", who, p->precision);
    if (!valid_NZ(p->NZ)) return fail(QL_ERR_ARG, "%s: NZ=%d not in {32,64,128,256}", who, p->NZ);
    if (p->NX < 2 || p->NY < 2 || (p->NX & 1) || (p->NY & 1))
        return fail(QL_ERR_ARG, "%s: NX=%d, NY=%d must be even and >= 2", who, p->NX, p->NY);
    if (p->NX > 4096 || p->NY > 4096) return fail(QL_ERR_ARG, "%s: NX=%d, NY=%d above 4096", who, p->NX, p->NY);
    if (!(p->L[0] > 0) || !(p->L[1] > 0) || !(p->L[2] > 0)) return fail(QL_ERR_ARG, "%s: box sizes must be > 0", who);
    if (!(p->dt != 0) || !std::isfinite(p->dt))
        return fail(QL_ERR_ARG, "%s: dt=%g (> 0 fixed steps, < 0 variable steps of at most -dt)", who, p->dt);
    if (p->dt < 0 && !(p->CFL > 0 && std::isfinite(p->CFL)))
        return fail(QL_ERR_ARG, "%s: variable steps need CFL > 0 (CFL=%g)", who, p->CFL);
    if (p->viscosity_order < 0) return fail(QL_ERR_ARG, "%s: viscosity_order < 0", who);
    return QL_OK;
}

// HD_fullU (HD_fullU.cpp:12-21, 145-327) is MHD_FullQlin with bx = eta = 0,
// Bx = By = 0, B0z = 0 and viscosity order 2 (its operator nu (-kx^2 - ky^2 +
// kz2), :272): every magnetic term is then a product with an exact zero and
// the magnetic fields stay zero
ql_params kernel_params(const ql_params &p)
{
    ql_params k = p;
    if (p.model == QL_MODEL_HD_FULLU) {
        k.model = QL_MODEL_MHD_FULLQLIN;
        k.B0z = 0.0;
        k.viscosity_order = 2;
        k.eta = p.nu;
    }
    return k;
}

// sizes and wavenumber tables of one run (Model.h:32-36, Kdata.cpp:28-80)
void fill_run(RunInfo &r, const ql_params &pu)
{
    const ql_params p = kernel_params(pu);
    r.hd = pu.model == QL_MODEL_HD_FULLU;
    r.p = p;
    r.nxy = (p.NX - 1) * (p.NY / 2);
    r.nz = 1 + 2 * (int)(p.NZ / 3.0);
    r.nzl = (r.nz - 1) / 2;
    r.half = r.nzl + 1;
    r.rsf = 1.0 / (p.NY * p.NX) / (p.NY * p.NX);  // MHD_FullQlin.cpp:67
    r.omega_eff = (p.model == QL_MODEL_MHD_BQLIN) ? 2.0 / 3.0 * p.q : p.omega;  // MHD_BQlin.cpp:19
    const int nxr = p.NX - 1;                                                      // MHD_FullQlin.cpp:466
    r.remap_thr = nxr / 2.0 * (2 * QL_PI / p.L[0]);
    r.remap_shift = nxr * (2 * QL_PI / p.L[0]);
    long tn = (long)p.NX * p.NY * p.NZ;
    r.totalN2 = (double)(tn * tn);
    r.kzfac = 2 * QL_PI / p.L[2] * (p.NZ / 3);
    // MHD_FullQlin.cpp:92-97 (MHD_BQlin.cpp:70-75)
    const double dealiasfac[3] = {2.0, 2.0, 3.0};
    const int N[3] = {p.NX, p.NY, p.NZ};
    r.kmax = 0;
    for (int i = 0; i < 3; ++i)
        if (r.kmax < 2 * QL_PI / p.L[i] * (N[i] / dealiasfac[i])) r.kmax = 2 * QL_PI / p.L[i] * (N[i] / dealiasfac[i]);
    const double kxL = 2 * QL_PI / p.L[0], kyL = 2 * QL_PI / p.L[1], kzL = 2 * QL_PI / p.L[2];
    const int nx = p.NX / 2, ny = p.NY / 2;
    r.h_kx0.assign(r.nxy, 0.0);
    r.h_ky.assign(r.nxy, 0.0);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            r.h_kx0[j + ny * i] = i * kxL;
            r.h_ky[j + ny * i] = j * kyL;
        }
    for (int i = nx + 1; i < 2 * nx; ++i)
        for (int j = 0; j < ny; ++j) {
            r.h_kx0[j + ny * (i - 1)] = (-2 * nx + i) * kxL;
            r.h_ky[j + ny * (i - 1)] = j * kyL;
        }
    r.h_kz.assign(r.nz, 0.0);
    for (int i = 0; i < r.nzl + 1; ++i) r.h_kz[i] = i * kzL;
    for (int i = r.nzl + 1; i < 2 * r.nzl + 1; ++i) r.h_kz[i] = (-2 * r.nzl + i - 1) * kzL;
}

qlk::RunParams run_params(const RunInfo &r)
{
    qlk::RunParams rp{};
    rp.ph.q = r.p.q;
    rp.ph.omega = r.omega_eff;
    rp.ph.B0z = r.p.B0z;
    rp.ph.nu = r.p.nu;
    rp.ph.eta = r.p.eta;
    rp.ph.vo2 = r.p.viscosity_order / 2;
    rp.ph.quasilinear = r.p.quasilinear;
    rp.rsf = r.rsf;
    rp.remap_thr = r.remap_thr;
    rp.remap_shift = r.remap_shift;
    rp.remap = r.p.remap;
    rp.nxy = r.nxy;
    rp.mode_off = r.mode_off;
    rp.kmax = r.kmax;
    rp.cfl = r.p.CFL;
    rp.dtmax = r.p.dt < 0 ? -r.p.dt : r.p.dt;
    rp.Lz = r.p.L[2];
    rp.hd = r.hd ? 1 : 0;
    return rp;
}

// workgroups of a run's FP64 companion (mixed precision): a prime near
// COMP_TARGET_WG that does not divide the mode count, so that the stride
// dealing does not line up with the kx rows (the large modes sit at the start
// of the rows), and at least enough that no workgroup has more than
// qlk::COMP_LIST candidates
int comp_nwg(int nxy)
{
    auto prime = [](int n) {
        if (n < 2) return false;
        for (int d = 2; d * d <= n; ++d)
            if (n % d == 0) return false;
        return true;
    };
    int n = std::max((nxy + qlk::COMP_LIST - 1) / qlk::COMP_LIST, std::min(COMP_TARGET_WG, nxy));
    while (!(prime(n) && nxy % n != 0) && n < nxy) ++n;
    return std::min(n, nxy);
}

// workgroup -> (run, mode-group range) tables for a launch shape: modes per
// workgroup pass MB = line/NZ; consecutive passes per workgroup chosen so that
// the group launches about `target` workgroups; each run's passes split into
// near-equal contiguous ranges
int configure(Batch &b, Group &g, int line, int target)
{
    if (g.dft && !dft_ok(g.NZ, b.precision, line)) g.dft = 0;  // the shape has no matrix-core variant
    // the tables are rewritten below: nothing queued may still read them
    QL_RC(sync_all(b));
    const int MB = mb_for(g.NZ, line);
    long total = 0;
    for (int id : g.run_ids) total += (b.runs[id].nxy + MB - 1) / MB;
    const long gpw = std::max<long>(1, (total + target - 1) / target);
    std::vector<qlk::WgInfo> wi;
    for (int k = 0; k < g.nruns; ++k) {
        qlk::RunParams &rp = g.h_rp[k];
        const int ng = (rp.nxy + MB - 1) / MB;
        const int nw = (int)((ng + gpw - 1) / gpw);
        rp.wg_off = (int)wi.size();
        rp.nwg = nw;
        for (int j = 0; j < nw; ++j) {
            qlk::WgInfo w{};
            w.run = k;
            w.g0 = (int)((long)j * ng / nw);
            w.g1 = (int)((long)(j + 1) * ng / nw);
            w.mode_off = rp.mode_off;
            w.nxy = rp.nxy;
            wi.push_back(w);
        }
    }
    const int W = (int)wi.size();
    // mixed precision: the FP64 companion (one-wavefront-per-mode kernel);
    // workgroup j of a run's nw2 has g0 = j, g1 = nw2 and strides over the
    // run's modes
    // (the wavefront kernel takes its flagged modes in FP64 itself,
    // k_wave_mix, and has no companion)
    std::vector<qlk::WgInfo> wi2;
    for (qlk::RunParams &rp : g.h_rp) rp.wg_off2 = rp.nwg2 = 0;
    if (g.split && line != LINE_WAVE) {
        for (int k = 0; k < g.nruns; ++k) {
            qlk::RunParams &rp = g.h_rp[k];
            const int nw2 = comp_nwg(rp.nxy);
            rp.wg_off2 = W + (int)wi2.size();
            rp.nwg2 = nw2;
            for (int j = 0; j < nw2; ++j) {
                qlk::WgInfo w{};
                w.run = k;
                w.g0 = j;
                w.g1 = nw2;
                w.mode_off = rp.mode_off;
                w.nxy = rp.nxy;
                wi2.push_back(w);
            }
        }
    }
    const int W2 = (int)wi2.size();
    const int ncols = (b.nMF) * g.NZ;
    if (W + W2 > g.wg_cap) {
        dfree(g.wg), dfree(g.wg2), dfree(g.part);
        QL_RC(dalloc(&g.wg, W + W2));
        QL_RC(dalloc(&g.wg2, W + W2));
        QL_RC(dalloc(&g.part, (size_t)(W + W2) * ncols));
        g.wg_cap = W + W2;
    }
    QL_HIP(hipMemcpy(g.wg, wi.data(), sizeof(qlk::WgInfo) * W, hipMemcpyHostToDevice));
    if (W2) QL_HIP(hipMemcpy(g.wg2, wi2.data(), sizeof(qlk::WgInfo) * W2, hipMemcpyHostToDevice));
    g.W2 = W2;
    QL_HIP(hipMemcpy(g.rp, g.h_rp.data(), sizeof(qlk::RunParams) * g.nruns, hipMemcpyHostToDevice));
    g.line = line;
    g.target = target;
    g.MB = MB;
    g.W = W;
    drop_graphs(b);
    return QL_OK;
}

// DrivingNoise settings of one run (MHD_FullQlin.cpp:20-22, 47-52, 65-66;
// MHD_BQlin.cpp:18, 54-55)
void noise_params(qlk::NoiseParams &n, const RunInfo &r, const ql_noise &c)
{
    n = qlk::NoiseParams{};
    n.sigma = c.f_noise / std::sqrt(2.0);
    n.lo2 = c.range_low * c.range_low;
    n.hi2 = c.range_high * c.range_high;
    // totalN2 * (1/(16 32 32))^2; the factor is 2^-28, so the product with dt
    // rounds as dt * totalN2 * mult_noise_fac does
    const double mnf = 1.0 / (16 * 32 * 32);
    n.nfac = r.totalN2 * (mnf * mnf);
    n.seed = c.seed;
    // the flags are MHD_FullQlin's; HD_fullU drives u and zeta, ky = 0 included
    // (HD_fullU.cpp:635-852); MHD_BQlin everything but ky = 0
    const bool full = r.p.model == QL_MODEL_MHD_FULLQLIN && !r.hd;
    n.drive_u = full ? !c.drive_only_magnetic : 1;
    n.drive_b = r.hd ? 0 : (full ? !c.drive_only_velocity : 1);
    n.drive_ky0 = r.p.model == QL_MODEL_MHD_FULLQLIN ? 1 : 0;
    n.nyh = r.p.NY / 2;
    n.nxh = r.p.NX / 2;
}

int create_batch(Batch &b, const ql_params *ps, int nruns)
{
    const ql_params p0 = kernel_params(ps[0]);
    b.hd = ps[0].model == QL_MODEL_HD_FULLU;
    b.model = p0.model;
    b.integrator = p0.integrator;
    b.precision = p0.precision;
    b.device = p0.device;
    b.vdt = p0.dt < 0 ? 1 : 0;
    b.dt = b.vdt ? -p0.dt : p0.dt;
    b.dt_last = b.dt;
    b.rdt_last.assign(nruns, b.dt);
    b.rt_prev.assign(nruns, 0.0);
    b.nMF = (p0.model == QL_MODEL_MHD_BQLIN) ? 2 : 4;
    b.runs.resize(nruns);
    std::vector<int> nzs;
    for (int r = 0; r < nruns; ++r) {
        fill_run(b.runs[r], ps[r]);
        if (std::find(nzs.begin(), nzs.end(), ps[r].NZ) == nzs.end()) nzs.push_back(ps[r].NZ);
    }
    std::sort(nzs.begin(), nzs.end());
    b.groups.resize(nzs.size());
    for (size_t gi = 0; gi < nzs.size(); ++gi) {
        Group &g = b.groups[gi];
        g.NZ = nzs[gi];
        for (int r = 0; r < nruns; ++r)
            if (ps[r].NZ == g.NZ) {
                RunInfo &ri = b.runs[r];
                ri.group = (int)gi;
                ri.gi = g.nruns++;
                ri.mode_off = g.modes;
                g.modes += ri.nxy;
                g.nz = ri.nz;
                g.run_ids.push_back(r);
                g.h_rp.push_back(run_params(ri));
            }
    }

    QL_HIP(hipStreamCreateWithFlags(&b.s, hipStreamNonBlocking));
    QL_HIP(hipEventCreateWithFlags(&b.fork, hipEventDisableTiming));
    for (auto &e : b.ev) QL_HIP(hipEventCreate(&e));
    QL_RC(dalloc(&b.dpart, (size_t)256 * 6));
    QL_RC(dalloc(&b.dout, (size_t)18));
    QL_RC(dalloc(&b.mfr_tmp, (size_t)4 * 256));
    for (Group &g : b.groups) {
        if (b.groups.size() > 1) {
            QL_HIP(hipStreamCreateWithFlags(&g.s, hipStreamNonBlocking));
            QL_HIP(hipEventCreateWithFlags(&g.join, hipEventDisableTiming));
        }
        const size_t nl = (size_t)g.modes * 4 * g.nz, nm = (size_t)g.nruns * b.nMF * g.nz;
        QL_RC(dalloc(&g.u, nl));
        QL_RC(dalloc(&g.N1, nl));
        QL_RC(dalloc(&g.N2, nl));
        QL_RC(dalloc(&g.mf, nm));
        QL_RC(dalloc(&g.mfc1, nm));
        QL_RC(dalloc(&g.mfc2, nm));
        QL_RC(dalloc(&g.mfN1, nm));
        QL_RC(dalloc(&g.mfN2, nm));
        QL_RC(dalloc(&g.drive, nm));
        QL_RC(dalloc(&g.tw, g.NZ));
        QL_RC(dalloc(&g.tw32, g.NZ));
        QL_RC(dalloc(&g.kx0, g.modes));
        QL_RC(dalloc(&g.ky, g.modes));
        QL_RC(dalloc(&g.kzv, g.nz));
        QL_RC(dalloc(&g.mfz, (size_t)g.nruns * 8 * g.NZ));
        QL_RC(dalloc(&g.sums, (size_t)g.nruns * b.nMF * g.NZ));
        QL_RC(dalloc(&g.rp, g.nruns));
        QL_RC(dalloc(&g.pm, (size_t)g.nruns * qlk::PM_STRIDE));
        QL_RC(dalloc(&g.mode_run, g.modes));
        QL_RC(dalloc(&g.np, g.nruns));
        QL_RC(dalloc(&g.clock, g.nruns));
        g.split = b.precision == QL_PREC_MIXED;
        if (g.split) {
            QL_RC(dalloc(&g.amp, g.modes));
            QL_RC(dalloc(&g.hi, g.modes));
        }
        std::vector<double> kx0(g.modes), ky(g.modes);
        std::vector<int> mrun(g.modes);
        g.h_np.assign(g.nruns, qlk::NoiseParams{});
        for (int id : g.run_ids) {
            const RunInfo &ri = b.runs[id];
            std::copy(ri.h_kx0.begin(), ri.h_kx0.end(), kx0.begin() + ri.mode_off);
            std::copy(ri.h_ky.begin(), ri.h_ky.end(), ky.begin() + ri.mode_off);
            std::fill(mrun.begin() + ri.mode_off, mrun.begin() + ri.mode_off + ri.nxy, ri.gi);
            noise_params(g.h_np[ri.gi], ri, ql_noise{0.0, 0.0, 1e16, 0, 0, 0ull});
        }
        QL_HIP(hipMemcpy(g.mode_run, mrun.data(), sizeof(int) * g.modes, hipMemcpyHostToDevice));
        QL_HIP(hipMemcpy(g.np, g.h_np.data(), sizeof(qlk::NoiseParams) * g.nruns, hipMemcpyHostToDevice));
        std::vector<double2> htw(g.NZ);
        std::vector<float2> htw32(g.NZ);
        for (int k = 0; k < g.NZ; ++k) {
            const double ang = 2.0 * QL_PI * (double)k / (double)g.NZ;
            htw[k] = make_double2(cos(ang), -sin(ang));
            htw32[k] = make_float2((float)htw[k].x, (float)htw[k].y);
        }
        const RunInfo &r0 = b.runs[g.run_ids[0]];
        QL_HIP(hipMemcpy(g.kx0, kx0.data(), sizeof(double) * g.modes, hipMemcpyHostToDevice));
        QL_HIP(hipMemcpy(g.ky, ky.data(), sizeof(double) * g.modes, hipMemcpyHostToDevice));
        // kz depends on Lz: runs of a group share it (checked by the caller)
        QL_HIP(hipMemcpy(g.kzv, r0.h_kz.data(), sizeof(double) * g.nz, hipMemcpyHostToDevice));
        QL_HIP(hipMemcpy(g.tw, htw.data(), sizeof(double2) * g.NZ, hipMemcpyHostToDevice));
        QL_HIP(hipMemcpy(g.tw32, htw32.data(), sizeof(float2) * g.NZ, hipMemcpyHostToDevice));
        const int line = default_line(g.NZ, b.precision, b.hd);
        QL_RC(configure(b, g, line, default_target(line)));
    }
    return QL_OK;
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------
// A mode-kernel launch.  In the per-kernel timing pass the launch carries its
// start and / or stop event itself (hipExtLaunchKernel), so they hold the
// dispatch's own begin and end times rather than those of separate event
// packets around it; a null event is not set (the companion launch of the
// mixed path sets the start, the FP32 launch after it the stop).
struct LaunchTiming {
    hipEvent_t start = nullptr, stop = nullptr;
};
// set by enqueue_sequence for the launches of one timed sequence
thread_local LaunchTiming tm_cur;

template <class K>
void launch_k(K kernel, int grid, int block, hipStream_t st, const LaunchTiming &tm, const qlk::ModeArgs &a)
{
    if (tm.start || tm.stop)
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, tm.start, tm.stop, 0, a);
    else
        kernel<<<grid, block, 0, st>>>(a);
}

template <int NZ, int MB, int MODE>
void launch_mode_m(const Batch &b, const Group &g, const qlk::ModeArgs &a, hipStream_t st)
{
    const LaunchTiming tm = tm_cur;
    const bool full = (b.model == QL_MODEL_MHD_FULLQLIN);
    constexpr int T = qlk::mode_threads(NZ, MB);
    // HD_fullU: the velocity-only policy (its magnetic fields are zero)
    if (b.precision == QL_PREC_MIXED) {
        if (b.hd) launch_k(qlk::k_mode32<NZ, MB, qlk::ModelHD, MODE>, g.W, T, st, tm, a);
        else if (full) launch_k(qlk::k_mode32<NZ, MB, qlk::ModelFull, MODE>, g.W, T, st, tm, a);
        else launch_k(qlk::k_mode32<NZ, MB, qlk::ModelBQ, MODE>, g.W, T, st, tm, a);
    } else if (g.dft) {
        if constexpr (NZ <= 64) {  // dft_ok
            if (b.hd) launch_k(qlk::k_mode64d<NZ, MB, qlk::ModelHD, MODE>, g.W, T, st, tm, a);
            else if (full) launch_k(qlk::k_mode64d<NZ, MB, qlk::ModelFull, MODE>, g.W, T, st, tm, a);
            else launch_k(qlk::k_mode64d<NZ, MB, qlk::ModelBQ, MODE>, g.W, T, st, tm, a);
        }
    } else {
        if (b.hd) launch_k(qlk::k_mode64<NZ, MB, qlk::ModelHD, MODE>, g.W, T, st, tm, a);
        else if (full) launch_k(qlk::k_mode64<NZ, MB, qlk::ModelFull, MODE>, g.W, T, st, tm, a);
        else launch_k(qlk::k_mode64<NZ, MB, qlk::ModelBQ, MODE>, g.W, T, st, tm, a);
    }
}

template <int NZ, int MODE>
void launch_wave_m(const Batch &b, const Group &g, const qlk::ModeArgs &a, hipStream_t st)
{
    const LaunchTiming tm = tm_cur;
    constexpr int TF = qlk::mode_waves<qlk::ModelFull>(NZ) * qlk::WAVE;
    constexpr int TH = qlk::mode_waves<qlk::ModelHD>(NZ) * qlk::WAVE;
    constexpr int TB = qlk::mode_waves<qlk::ModelBQ>(NZ) * qlk::WAVE;
    const bool bq = b.model == QL_MODEL_MHD_BQLIN;
    // HD_fullU: the velocity-only policy (its magnetic fields are zero)
    if (b.precision == QL_PREC_MIXED && a.split == qlk::SPLIT_LO) {
        // flagged modes in FP64, the others in FP32, in one launch
        if (b.hd) launch_k(qlk::k_wave_mix<NZ, qlk::ModelHD, MODE>, g.W, TH, st, tm, a);
        else if (bq) launch_k(qlk::k_wave_mix<NZ, qlk::ModelBQ, MODE>, g.W, TB, st, tm, a);
        else launch_k(qlk::k_wave_mix<NZ, qlk::ModelFull, MODE>, g.W, TF, st, tm, a);
    } else if (b.precision == QL_PREC_MIXED) {
        if (b.hd) launch_k(qlk::k_wave32<NZ, qlk::ModelHD, MODE>, g.W, TH, st, tm, a);
        else if (bq) launch_k(qlk::k_wave32<NZ, qlk::ModelBQ, MODE>, g.W, TB, st, tm, a);
        else launch_k(qlk::k_wave32<NZ, qlk::ModelFull, MODE>, g.W, TF, st, tm, a);
    } else {
        if (b.hd) launch_k(qlk::k_wave64<NZ, qlk::ModelHD, MODE>, g.W, TH, st, tm, a);
        else if (bq) launch_k(qlk::k_wave64<NZ, qlk::ModelBQ, MODE>, g.W, TB, st, tm, a);
        else launch_k(qlk::k_wave64<NZ, qlk::ModelFull, MODE>, g.W, TF, st, tm, a);
    }
}

template <int NZ>
void launch_wave_t(const Batch &b, const Group &g, const qlk::ModeArgs &a, hipStream_t st)
{
    if (a.mode == qlk::MODE_ELIN) launch_wave_m<NZ, qlk::MODE_ELIN>(b, g, a, st);
    else if (a.mode == qlk::MODE_CN) launch_wave_m<NZ, qlk::MODE_CN>(b, g, a, st);
    else launch_wave_m<NZ, qlk::MODE_RHS>(b, g, a, st);
}

// the mixed-precision FP64 companion: the FP64 workgroup kernel over the
// group's companion table, at the smallest pass of each NZ
template <int NZ, int MODE>
void launch_comp_m(const Batch &b, const Group &g, const qlk::ModeArgs &a, hipStream_t st)
{
    const LaunchTiming tm = tm_cur;
    constexpr int MB = NZ == 32 ? 4 : (NZ == 64 ? 2 : 1);
    constexpr int T = qlk::mode_threads(NZ, MB);
    if (b.hd) launch_k(qlk::k_mode64s<NZ, MB, qlk::ModelHD, MODE>, g.W2, T, st, tm, a);
    else if (b.model == QL_MODEL_MHD_BQLIN) launch_k(qlk::k_mode64s<NZ, MB, qlk::ModelBQ, MODE>, g.W2, T, st, tm, a);
    else launch_k(qlk::k_mode64s<NZ, MB, qlk::ModelFull, MODE>, g.W2, T, st, tm, a);
}

template <int NZ>
void launch_comp_t(const Batch &b, const Group &g, const qlk::ModeArgs &a, hipStream_t st)
{
    if (a.mode == qlk::MODE_ELIN) launch_comp_m<NZ, qlk::MODE_ELIN>(b, g, a, st);
    else if (a.mode == qlk::MODE_CN) launch_comp_m<NZ, qlk::MODE_CN>(b, g, a, st);
    else launch_comp_m<NZ, qlk::MODE_RHS>(b, g, a, st);
}

void launch_comp(const Batch &b, const Group &g, const qlk::ModeArgs &a, hipStream_t st)
{
    switch (g.NZ) {
    case 32: launch_comp_t<32>(b, g, a, st); break;
    case 64: launch_comp_t<64>(b, g, a, st); break;
    case 128: launch_comp_t<128>(b, g, a, st); break;
    default: launch_comp_t<256>(b, g, a, st); break;
    }
}

template <int NZ, int MB>
void launch_mode_t(const Batch &b, const Group &g, const qlk::ModeArgs &a, hipStream_t st)
{
    if (a.mode == qlk::MODE_ELIN) launch_mode_m<NZ, MB, qlk::MODE_ELIN>(b, g, a, st);
    else if (a.mode == qlk::MODE_CN) launch_mode_m<NZ, MB, qlk::MODE_CN>(b, g, a, st);
    else launch_mode_m<NZ, MB, qlk::MODE_RHS>(b, g, a, st);
}

void launch_mode(const Batch &b, const Group &g, const qlk::ModeArgs &a, hipStream_t st)
{
    if (g.line == LINE_WAVE) {
        switch (g.NZ) {
        case 32: launch_wave_t<32>(b, g, a, st); break;
        case 64: launch_wave_t<64>(b, g, a, st); break;
        case 128: launch_wave_t<128>(b, g, a, st); break;
        default: launch_wave_t<256>(b, g, a, st); break;
        }
        return;
    }
    switch (g.NZ) {
    case 32: g.MB == 4 ? launch_mode_t<32, 4>(b, g, a, st) : launch_mode_t<32, 8>(b, g, a, st); break;
    case 64: g.MB == 2 ? launch_mode_t<64, 2>(b, g, a, st) : launch_mode_t<64, 4>(b, g, a, st); break;
    case 128: g.MB == 1 ? launch_mode_t<128, 1>(b, g, a, st) : launch_mode_t<128, 2>(b, g, a, st); break;
    default: launch_mode_t<256, 1>(b, g, a, st); break;
    }
}

template <int NZ, int MODE>
void launch_mf_m(const Group &g, const qlk::MFArgs &m, hipStream_t st)
{
    const bool rec = m.rec || m.vdt;
    if (m.model == QL_MODEL_MHD_BQLIN) {
        if (rec) qlk::k_meanfield<NZ, MODE, 1, true><<<g.nruns, qlk::NT, 0, st>>>(m);
        else qlk::k_meanfield<NZ, MODE, 1, false><<<g.nruns, qlk::NT, 0, st>>>(m);
    } else {
        if (rec) qlk::k_meanfield<NZ, MODE, 0, true><<<g.nruns, qlk::NT, 0, st>>>(m);
        else qlk::k_meanfield<NZ, MODE, 0, false><<<g.nruns, qlk::NT, 0, st>>>(m);
    }
}

template <int NZ>
void launch_mf_t(const Group &g, const qlk::MFArgs &m, hipStream_t st)
{
    switch (m.mode) {
    case qlk::MODE_RHS: launch_mf_m<NZ, qlk::MODE_RHS>(g, m, st); break;
    case qlk::MODE_ELIN: launch_mf_m<NZ, qlk::MODE_ELIN>(g, m, st); break;
    case qlk::MODE_CN: launch_mf_m<NZ, qlk::MODE_CN>(g, m, st); break;
    default: launch_mf_m<NZ, qlk::MODE_PROFILES>(g, m, st); break;
    }
}

void launch_mf(const Group &g, const qlk::MFArgs &m, hipStream_t st)
{
    switch (g.NZ) {
    case 32: launch_mf_t<32>(g, m, st); break;
    case 64: launch_mf_t<64>(g, m, st); break;
    case 128: launch_mf_t<128>(g, m, st); break;
    default: launch_mf_t<256>(g, m, st); break;
    }
}

qlk::ModeArgs mode_args(const Batch &b, const Group &g)
{
    qlk::ModeArgs a{};
    a.u = g.u;
    a.kx0 = g.kx0;
    a.ky = g.ky;
    a.kzv = g.kzv;
    a.tw = g.tw;
    a.tw32 = g.tw32;
    a.mfz = g.mfz;
    a.partials = g.part;
    a.runs = g.rp;
    a.wg = g.wg;
    a.clock = g.clock;
    a.vdt = b.vdt;
    const double dt = b.dt;
    a.dt = dt;
    a.dtP1 = dt * P1;
    a.dtP2 = dt * P2;
    a.dtl1 = dt * P1;
    a.dtl2 = dt * (P2 - P1);
    a.dtl3 = dt * (1.0 - P2);
    a.fac = 1.0 / g.NZ;  // fftwPlans.cpp:30
    return a;
}

// mixed-precision split of a mode launch: the FP32 launch skips the flagged
// modes, which the FP64 companion takes
// mixed precision with tau > 1: no mode can reach it, so every mode runs in
// FP32 and neither the flags nor the companion are launched
bool split_active(const Batch &b, const Group &g) { return g.split && b.amp_tau <= 1.0; }

void set_split(const Batch &b, const Group &g, qlk::ModeArgs &a)
{
    if (!split_active(b, g)) return;
    a.split = qlk::SPLIT_LO;
    a.hi = g.hi;
}

// the precision flags from the mode amplitudes of the current state: always
// (clocked = false: after the state changed outside the steps), or inside a
// step when the group's clock is at a multiple of SPLIT_EVERY steps
void launch_flags(const Batch &b, const Group &g, bool clocked, hipStream_t st)
{
    const qlk::Clock *ck = clocked ? g.clock : nullptr;
    qlk::k_amp_modes<<<g.modes, 64, 0, st>>>(g.u, g.nz, g.amp, ck, SPLIT_EVERY);
    qlk::k_amp_flags<<<g.nruns, qlk::AMP_NT, 0, st>>>(g.amp, g.rp, b.amp_tau, g.hi, ck, SPLIT_EVERY);
}

int prime_split(Batch &b, Group &g, hipStream_t st)
{
    if (!split_active(b, g) || !g.amp_stale) return QL_OK;
    launch_flags(b, g, false, st);
    QL_HIP(hipGetLastError());
    g.amp_stale = false;
    return QL_OK;
}

int prime_split_all(Batch &b)
{
    for (Group &g : b.groups) QL_RC(prime_split(b, g, b.s));
    return QL_OK;
}

qlk::MFArgs mf_args(const Batch &b, const Group &g)
{
    qlk::MFArgs m{};
    m.sums = g.sums;
    m.mf = g.mf;
    m.drive = g.drive;
    m.mfz = g.mfz;
    m.runs = g.rp;
    m.kzv = g.kzv;
    m.tw = g.tw;
    m.clock = g.clock;
    m.pm = g.pm;
    m.vdt = b.vdt;
    m.rec = b.vdt;
    m.model = b.model;
    m.nMF = b.nMF;
    m.fac = 1.0 / g.NZ;
    return m;
}

int launch_noise(Group &g, bool use_clock, double t, double dt, bool noise, bool remap, hipStream_t st)
{
    qlk::NoiseArgs a{};
    a.u = g.u;
    a.kx0 = g.kx0;
    a.ky = g.ky;
    a.kzv = g.kzv;
    a.runs = g.rp;
    a.np = g.np;
    a.mode_run = g.mode_run;
    a.clock = g.clock;
    a.nz = g.nz;
    a.use_clock = use_clock ? 1 : 0;
    a.noise = noise ? 1 : 0;
    a.do_remap = remap ? 1 : 0;
    a.t_x = t;
    a.dt_x = dt;
    const int threads = (g.nz + 63) / 64 * 64;  // one thread per kz slot (nz <= 171)
    qlk::k_noise_remap<<<g.modes, threads, 0, st>>>(a);
    QL_HIP(hipGetLastError());
    return QL_OK;
}

// The substep tail as one launch (k_mf_fused) applies to fixed steps (no CFL
// records), NZ <= 128 and runs of at most TAIL_MAX_ROWS_NZ / NZ partial rows
// (each of its workgroups reads one field's columns of every row)
constexpr long TAIL_MAX_ROWS_NZ = 65536;
bool fused_tail_fits(const Batch &b, const Group &g)
{
    if (b.vdt || g.NZ > 128) return false;
    for (const qlk::RunParams &rp : g.h_rp)
        if ((long)(rp.nwg + rp.nwg2) * g.NZ > TAIL_MAX_ROWS_NZ) return false;
    return true;
}
bool fused_tail(const Batch &b, const Group &g) { return g.tail && fused_tail_fits(b, g); }

// the buffer holding the group's mean fields after substep s of a step
const double2 *mf_after_substep(const Batch &b, const Group &g, int s)
{
    if (!fused_tail(b, g) || s == 3) return g.mf;
    return s == 1 ? g.mfc1 : g.mfc2;
}

template <int NZ>
void launch_fused_tail_t(const Group &g, const qlk::MFArgs &m, const double *part, hipStream_t st)
{
    const dim3 grid(m.nMF, g.nruns);
    if (m.model == QL_MODEL_MHD_BQLIN) {
        if (m.mode == qlk::MODE_CN) qlk::k_mf_fused<NZ, qlk::MODE_CN, 1><<<grid, qlk::MFF_T, 0, st>>>(m, part);
        else qlk::k_mf_fused<NZ, qlk::MODE_ELIN, 1><<<grid, qlk::MFF_T, 0, st>>>(m, part);
    } else {
        if (m.mode == qlk::MODE_CN) qlk::k_mf_fused<NZ, qlk::MODE_CN, 0><<<grid, qlk::MFF_T, 0, st>>>(m, part);
        else qlk::k_mf_fused<NZ, qlk::MODE_ELIN, 0><<<grid, qlk::MFF_T, 0, st>>>(m, part);
    }
}

void launch_fused_tail(const Group &g, const qlk::MFArgs &m, const double *part, hipStream_t st)
{
    switch (g.NZ) {
    case 32: launch_fused_tail_t<32>(g, m, part, st); break;
    case 64: launch_fused_tail_t<64>(g, m, part, st); break;
    default: launch_fused_tail_t<128>(g, m, part, st); break;
    }
}

// mode kernel -> column reduction of its partials -> mean-field kernel; the
// partial rows and column sums are a.partials and m.sums (the group's own
// buffers, or ql_rhs_noalloc's scratch)
int enqueue_sequence(Batch &b, Group &g, const qlk::ModeArgs &a, const qlk::MFArgs &m, hipStream_t st)
{
    // timing pass: e[0] = the first mode launch's start, e[1] = the last
    // one's end (dispatch times), e[2] = after the mean-field kernel; read by
    // ql_profile_kernels when the pass has finished
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    if (b.timing) {
        while (b.tpool.size() < b.tused + 3) {
            hipEvent_t ev = nullptr;
            QL_HIP(hipEventCreate(&ev));
            b.tpool.push_back(ev);
        }
        for (int k = 0; k < 3; ++k) e[k] = b.tpool[b.tused + k];
        b.tused += 3;
    }
    const int ncols = b.nMF * g.NZ;
    if (a.split == qlk::SPLIT_LO && g.line != LINE_WAVE) {
        // the FP64 companion first, then the FP32 launch (the FP32 launch
        // fills every CU in one round of workgroups, so running them side by
        // side would only delay the workgroups the companion displaces)
        qlk::ModeArgs a2 = a;
        a2.split = qlk::SPLIT_HI;
        a2.wg = g.wg2;
        a2.partials = a.partials + (size_t)g.W * ncols;
        tm_cur = LaunchTiming{e[0], nullptr};
        launch_comp(b, g, a2, st);
        QL_HIP(hipGetLastError());
        tm_cur = LaunchTiming{nullptr, e[1]};
        launch_mode(b, g, a, st);
        QL_HIP(hipGetLastError());
    } else {
        tm_cur = LaunchTiming{e[0], e[1]};
        launch_mode(b, g, a, st);
        QL_HIP(hipGetLastError());
    }
    tm_cur = LaunchTiming{};
    if (fused_tail(b, g) && (m.mode == qlk::MODE_ELIN || m.mode == qlk::MODE_CN)) {
        launch_fused_tail(g, m, a.partials, st);
        QL_HIP(hipGetLastError());
    } else {
        qlk::k_reduce_cols<<<dim3((ncols + qlk::RED_CW - 1) / qlk::RED_CW, g.nruns), qlk::NT, 0, st>>>(
            a.partials, g.rp, ncols, const_cast<double *>(m.sums));
        QL_HIP(hipGetLastError());
        launch_mf(g, m, st);
        QL_HIP(hipGetLastError());
    }
    if (b.timing) QL_HIP(hipEventRecord(e[2], st));
    return QL_OK;
}

int refresh_profiles(Batch &b, Group &g, hipStream_t st, bool rec)
{
    qlk::MFArgs m = mf_args(b, g);
    m.mode = qlk::MODE_PROFILES;
    if (rec) m.rec = 1;
    m.mfn_out = g.mfN1;
    launch_mf(g, m, st);
    QL_HIP(hipGetLastError());
    return QL_OK;
}

// One substep s (1..3) of RK3CN_Elin::Step (RK3CN_Elin.cpp:290-378) or
// RK3CN::Step (RK3CN.cpp:92-193) for one group; times come from the group's
// clock, which the last substep advances when `advance` is set.
int enqueue_substep(Batch &b, Group &g, int s, bool remap_after, bool advance, hipStream_t st)
{
    const bool cn = (b.integrator == QL_INT_RK3CN);
    const double dt = b.dt;
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
    // with driving noise the step ends with the noise kernel, which remaps
    // after it (main.cpp:112-123)
    const bool noisy = g.noise && advance && s == 3;
    qlk::ModeArgs a = mode_args(b, g);
    a.s = s;
    a.mode = cn ? qlk::MODE_CN : qlk::MODE_ELIN;
    set_split(b, g, a);
    a.linC = lc;
    a.nlc1 = nlc1;
    a.nlc2 = nlc2;
    a.n_prev = (s == 1) ? nullptr : (s == 2 ? g.N1 : g.N2);
    a.n_out = (s == 1) ? g.N1 : (s == 2 ? g.N2 : nullptr);
    a.do_remap = (remap_after && !noisy) ? 1 : 0;

    qlk::MFArgs m = mf_args(b, g);
    m.mode = a.mode;
    m.linC = lc;
    m.nlc1 = nlc1;
    m.nlc2 = nlc2;
    m.mfn_prev = (s == 1) ? nullptr : (s == 2 ? g.mfN1 : g.mfN2);
    m.mfn_out = (s == 2) ? g.mfN2 : g.mfN1;
    m.advance = (advance && s == 3) ? 1 : 0;
    m.s = s;
    if (fused_tail(b, g)) {
        m.mf_in = mf_after_substep(b, g, s == 1 ? 3 : s - 1);
        m.mf = const_cast<double2 *>(mf_after_substep(b, g, s));
    }
    QL_RC(enqueue_sequence(b, g, a, m, st));
    if (noisy) QL_RC(launch_noise(g, true, 0.0, 0.0, true, remap_after, st));
    return QL_OK;
}

// substep s of every group: groups fork from the handle's stream and join it
int enqueue_substep_all(Batch &b, int s, bool remap_after, bool advance)
{
    if (b.groups.size() == 1) return enqueue_substep(b, b.groups[0], s, remap_after, advance, b.s);
    QL_HIP(hipEventRecord(b.fork, b.s));
    for (Group &g : b.groups) {
        QL_HIP(hipStreamWaitEvent(g.s, b.fork, 0));
        QL_RC(enqueue_substep(b, g, s, remap_after, advance, g.s));
        QL_HIP(hipEventRecord(g.join, g.s));
    }
    for (Group &g : b.groups) QL_HIP(hipStreamWaitEvent(b.s, g.join, 0));
    return QL_OK;
}

// mixed precision: the clocked flag refresh at the start of a step
int enqueue_flags(Batch &b)
{
    for (Group &g : b.groups)
        if (split_active(b, g)) launch_flags(b, g, true, b.s);
    QL_HIP(hipGetLastError());
    return QL_OK;
}

// one step of main.cpp:108-123 (noise off): Step, t += dt, remap
int enqueue_step(Batch &b, bool remap)
{
    QL_RC(enqueue_flags(b));
    for (int s = 1; s <= 3; ++s) QL_RC(enqueue_substep_all(b, s, remap && s == 3, true));
    return QL_OK;
}

// every run at time t, the step history shared (fixed steps, or one run)
int set_clock(Batch &b, double t)
{
    qlk::Clock c{};
    c.t = t;
    c.t_prev = b.t_prev;
    c.t_init = b.t_init;
    c.old_init = b.old_init;
    c.nstep = b.nstep;
    c.dt = b.dt;
    c.dt_last = b.dt_last;
    for (Group &g : b.groups)
        qlk::k_clock_fill<<<1, 256, 0, b.s>>>(g.clock, g.nruns, c, g.pm, g.rp, b.model, b.vdt);
    QL_HIP(hipGetLastError());
    return QL_OK;
}

// each run at its own time with its own step history (variable steps in a set
// of several runs); the device then sets each run's next dt
int set_clock_runs(Batch &b, const double *t_runs)
{
    for (Group &g : b.groups) {
        std::vector<qlk::Clock> &h = g.h_clock;
        h.resize(g.nruns);
        for (int k = 0; k < g.nruns; ++k) {
            const int r = g.run_ids[k];
            qlk::Clock &c = h[k];
            c = qlk::Clock{};
            c.t = t_runs[r];
            c.t_prev = b.rt_prev[r];
            c.t_init = b.t_init;
            c.old_init = b.old_init;
            c.nstep = b.nstep;
            c.dt = b.dt;
            c.dt_last = b.rdt_last[r];
        }
        QL_HIP(hipMemcpyAsync(g.clock, h.data(), sizeof(qlk::Clock) * g.nruns, hipMemcpyHostToDevice, b.s));
        qlk::k_clock_dt<<<1, 256, 0, b.s>>>(g.clock, g.nruns, g.pm, g.rp, b.model);
        QL_HIP(hipGetLastError());
    }
    // the copies read g.h_clock: let them finish before it can change
    QL_HIP(hipStreamSynchronize(b.s));
    return QL_OK;
}

// the device clock of every run (run order)
int read_clocks(Batch &b, std::vector<qlk::Clock> &out)
{
    out.resize(b.runs.size());
    for (Group &g : b.groups) {
        g.h_clock.resize(g.nruns);
        QL_HIP(hipMemcpyAsync(g.h_clock.data(), g.clock, sizeof(qlk::Clock) * g.nruns, hipMemcpyDeviceToHost, b.s));
    }
    QL_HIP(hipStreamSynchronize(b.s));
    for (Group &g : b.groups)
        for (int k = 0; k < g.nruns; ++k) out[g.run_ids[k]] = g.h_clock[k];
    return QL_OK;
}

// the host's view of run 0's clock after variable steps
int read_clock(Batch &b, qlk::Clock *c)
{
    std::vector<qlk::Clock> all;
    QL_RC(read_clocks(b, all));
    *c = all[0];
    return QL_OK;
}

int capture(Batch &b, int nsteps, hipGraphExec_t *out)
{
    hipGraph_t graph = nullptr;
    QL_HIP(hipStreamBeginCapture(b.s, hipStreamCaptureModeThreadLocal));
    int rc = QL_OK;
    for (int n = 0; n < nsteps && rc == QL_OK; ++n) rc = enqueue_step(b, true);
    hipError_t e = hipStreamEndCapture(b.s, &graph);
    if (rc != QL_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(QL_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(QL_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    return QL_OK;
}

// the replay graphs of every level for the current shapes
int ensure_graphs(Batch &b)
{
    for (int k = 0; k < 5; ++k)
        if (!b.graphs[k]) QL_RC(capture(b, GRAPH_LEVELS[k], &b.graphs[k]));
    return QL_OK;
}

// nsteps of the driver loop with remap as each run asks; graph replay unless
// per-kernel timing is on.  The replay graphs (16, 8, 4, 2 and 1 steps) are
// all captured by the first call that steps, whatever its step count, so that
// later calls (a timed loop after a warm-up) only launch.
int run_steps(Batch &b, int nsteps, double *t_runs)
{
    const bool per_run = b.vdt && b.runs.size() > 1;
    double t = t_runs[0];
    if (per_run) QL_RC(set_clock_runs(b, t_runs));
    else QL_RC(set_clock(b, t));
    QL_RC(prime_split_all(b));
    if (b.timing) {
        for (int n = 0; n < nsteps; ++n) QL_RC(enqueue_step(b, true));
    } else if (nsteps > 0) {
        QL_RC(ensure_graphs(b));
        int left = nsteps;
        for (int k = 0; k < 5; ++k)
            for (; left >= GRAPH_LEVELS[k]; left -= GRAPH_LEVELS[k]) QL_HIP(hipGraphLaunch(b.graphs[k], b.s));
    }
    QL_HIP(hipStreamSynchronize(b.s));
    if (b.vdt) {
        if (nsteps > 0) {
            std::vector<qlk::Clock> cs;
            QL_RC(read_clocks(b, cs));
            for (size_t r = 0; r < cs.size(); ++r) {
                t_runs[r] = cs[r].t;
                b.rt_prev[r] = cs[r].t_prev;
                b.rdt_last[r] = cs[r].dt_last;
            }
            b.t_prev = cs[0].t_prev;
            b.dt_last = cs[0].dt_last;
        }
    } else {
        // the same accumulation as the device clock (main.cpp:116)
        for (int n = 0; n < nsteps; ++n) {
            b.t_prev = t;
            t = t + b.dt;
        }
        for (size_t r = 0; r < b.runs.size(); ++r) t_runs[r] = t;
    }
    if (nsteps > 0) b.old_init = 0;
    b.nstep += nsteps;
    return QL_OK;
}

// the caller's layout: fluctuation fields per mode, mean fields, and the
// first internal mean-field slot the caller's mean field 0 is
int ext_nlin(const Batch &b) { return b.hd ? 2 : 4; }
int ext_nmf(const Batch &b) { return b.hd ? 2 : b.nMF; }
int ext_mf0(const Batch &b) { return b.hd ? 2 : 0; }

// [modes][4][nz] <-> [modes][nL][nz] (the internal slots past nL are zero)
template <class T>
void lin_to_int(const T *src, T *dst, int modes, int nL, int nz)
{
    for (int i = 0; i < modes; ++i) {
        std::copy(src + (size_t)i * nL * nz, src + (size_t)(i + 1) * nL * nz, dst + (size_t)i * 4 * nz);
        std::fill(dst + (size_t)i * 4 * nz + (size_t)nL * nz, dst + (size_t)(i + 1) * 4 * nz, T{});
    }
}
template <class T>
void lin_to_ext(const T *src, T *dst, int modes, int nL, int nz)
{
    for (int i = 0; i < modes; ++i)
        std::copy(src + (size_t)i * 4 * nz, src + (size_t)i * 4 * nz + (size_t)nL * nz, dst + (size_t)i * nL * nz);
}

int check_run(const Batch &b, int run, const char *who)
{
    if (run < 0 || run >= (int)b.runs.size()) return fail(QL_ERR_ARG, "%s: run %d of %d", who, run, (int)b.runs.size());
    return QL_OK;
}

// MHD_FullQlin and HD_fullU form their stresses as z-derivatives, which holds
// for a divergence-free state (see ModelFull in ql_mode.hpp).  The one slot
// the (ux, zeta) representation leaves unconstrained, kz = 0 of a ky = 0 mode,
// is then ux = bx = 0 (u parallel to k there is a compressive component).  A
// state uploaded with more than rounding there is refused: the reference keeps
// such a component frozen and would give different stresses.
int check_solenoidal(const Batch &b, const RunInfo &r, const double2 *ul)
{
    if (b.model == QL_MODEL_MHD_BQLIN) return QL_OK;  // exact for any state
    const int ny = r.p.NY / 2, nz = r.nz;
    for (int pr = 0; pr < 2; ++pr) {
        double big = 0.0;
        for (int i = 0; i < r.nxy; ++i)
            for (int f = 2 * pr; f < 2 * pr + 2; ++f)
                for (int m = 0; m < nz; ++m) {
                    const double2 v = ul[((size_t)i * 4 + f) * nz + m];
                    big = std::max(big, std::max(std::fabs(v.x), std::fabs(v.y)));
                }
        for (int i = 0; i < r.nxy; i += ny) {
            const double2 v = ul[((size_t)i * 4 + 2 * pr) * nz];
            if (std::max(std::fabs(v.x), std::fabs(v.y)) > 1e-12 * big)
                return fail(QL_ERR_ARG,
                            "set_state: %s of mode %d (ky = 0) is %g + %gi at kz = 0, the compressive slot a "
                            "divergence-free state holds at zero (largest field value %g); zero it before upload",
                            pr == 0 ? "ux" : "bx", i, v.x, v.y, big);
        }
    }
    return QL_OK;
}

int set_state(Batch &b, int run, const double *lin, const double *mf, const double *kx0)
{
    const RunInfo &r = b.runs[run];
    Group &g = b.groups[r.group];
    const size_t mo = (size_t)r.gi * b.nMF * g.nz;
    std::vector<double2> tl, tm;  // internal layout of an HD_fullU state
    const double2 *ul = reinterpret_cast<const double2 *>(lin), *um = reinterpret_cast<const double2 *>(mf);
    if (b.hd && ul) {
        tl.resize((size_t)r.nxy * 4 * g.nz);
        lin_to_int(ul, tl.data(), r.nxy, 2, g.nz);
        ul = tl.data();
    }
    if (ul) QL_RC(check_solenoidal(b, r, ul));
    if (b.hd && um) {
        tm.assign((size_t)b.nMF * g.nz, make_double2(0.0, 0.0));
        std::copy(um, um + 2 * g.nz, tm.begin() + 2 * g.nz);
        um = tm.data();
    }
    if (ul) {
        QL_HIP(hipMemcpyAsync(g.u + (size_t)r.mode_off * 4 * g.nz, ul, sizeof(double2) * (size_t)r.nxy * 4 * g.nz,
                              hipMemcpyHostToDevice, b.s));
        g.amp_stale = true;
    }
    if (um) QL_HIP(hipMemcpyAsync(g.mf + mo, um, sizeof(double2) * b.nMF * g.nz, hipMemcpyHostToDevice, b.s));
    if (kx0)
        QL_HIP(hipMemcpyAsync(g.kx0 + r.mode_off, kx0, sizeof(double) * r.nxy, hipMemcpyHostToDevice, b.s));
    QL_RC(refresh_profiles(b, g, b.s));
    QL_HIP(hipStreamSynchronize(b.s));
    return QL_OK;
}

int get_state(Batch &b, int run, double *lin, double *mf, double *kx0)
{
    const RunInfo &r = b.runs[run];
    Group &g = b.groups[r.group];
    const size_t mo = (size_t)r.gi * b.nMF * g.nz;
    std::vector<double2> tl(b.hd && lin ? (size_t)r.nxy * 4 * g.nz : 0);
    double2 *dl = b.hd ? tl.data() : reinterpret_cast<double2 *>(lin);
    if (lin)
        QL_HIP(hipMemcpyAsync(dl, g.u + (size_t)r.mode_off * 4 * g.nz, sizeof(double2) * (size_t)r.nxy * 4 * g.nz,
                              hipMemcpyDeviceToHost, b.s));
    if (mf)
        QL_HIP(hipMemcpyAsync(mf, g.mf + mo + (size_t)ext_mf0(b) * g.nz, sizeof(double2) * ext_nmf(b) * g.nz,
                              hipMemcpyDeviceToHost, b.s));
    if (kx0)
        QL_HIP(hipMemcpyAsync(kx0, g.kx0 + r.mode_off, sizeof(double) * r.nxy, hipMemcpyDeviceToHost, b.s));
    QL_HIP(hipStreamSynchronize(b.s));
    if (b.hd && lin) lin_to_ext(tl.data(), reinterpret_cast<double2 *>(lin), r.nxy, 2, g.nz);
    return QL_OK;
}

int diagnostics(Batch &b, int run, double t, double out[18])
{
    const RunInfo &r = b.runs[run];
    Group &g = b.groups[r.group];
    qlk::DiagArgs d{};
    d.u = g.u;
    d.kx0 = g.kx0;
    d.ky = g.ky;
    d.kzv = g.kzv;
    d.m0 = r.mode_off;
    d.nxy = r.nxy;
    d.nz = g.nz;
    d.qt = r.p.q * t;
    d.nu = r.p.nu;
    d.eta = r.p.eta;
    d.part = b.dpart;
    const int grid = r.nxy < 256 ? r.nxy : 256;
    qlk::k_diag_modes<<<grid, qlk::NT, 0, b.s>>>(d);
    QL_HIP(hipGetLastError());
    const size_t mo = (size_t)r.gi * b.nMF * g.nz;
    qlk::DiagFinalArgs f{};
    f.part = b.dpart;
    f.nparts = grid;
    f.mf = g.mf + mo;
    f.drive = g.drive + mo;
    f.kzv = g.kzv;
    f.nz = g.nz;
    f.nMF = b.nMF;
    f.model = b.model;
    f.NZ = g.NZ;
    f.Lz = r.p.L[2];
    f.totalN2 = r.totalN2;
    f.nu = r.p.nu;
    f.eta = r.p.eta;
    f.kzfac = r.kzfac;
    f.out = b.dout;
    qlk::k_diag_final<<<1, 64, 0, b.s>>>(f);
    QL_HIP(hipGetLastError());
    QL_HIP(hipMemcpyAsync(out, b.dout, sizeof(double) * 18, hipMemcpyDeviceToHost, b.s));
    QL_HIP(hipStreamSynchronize(b.s));
    if (r.hd) {
        // HD_fullU.cpp:463-488: no magnetic quantities, Reynolds stresses are
        // saved as full profiles (ql_drive_real) instead of projections
        for (int k = 12; k < 16; ++k) out[k] = 0.0;
        out[17] = 0.0;
    }
    return QL_OK;
}

void linear_ops_host(const Batch &b, const RunInfo &r, const std::vector<double> &k0, double qt,
                     double *linop, double *linop_mf)
{
    const int nz = r.nz;
    const ql_params &p = r.p;
    if (linop) {
        // MHD_FullQlin.cpp:432-440 / MHD_BQlin.cpp:289-295 on the current kx0
        for (int i = 0; i < r.nxy; ++i) {
            const double ky = r.h_ky[i];
            const double kx = k0[i] + qt * ky;
            for (int m = 0; m < nz; ++m) {
                const double kz2 = -(r.h_kz[m] * r.h_kz[m]);
                const double lap2 = (ky == 0.0 && m == 0) ? 0.0 : (-ky * ky + kz2);
                const double lapF = -kx * kx + lap2;
                double L0, L2;
                if (p.model == QL_MODEL_MHD_BQLIN) {
                    L0 = p.nu * lapF;
                    L2 = (p.eta / p.nu) * L0;
                } else {
                    const double lp = -pow(fabs(lapF), (double)(p.viscosity_order / 2));
                    L0 = p.nu * lp;
                    L2 = p.eta * lp;
                }
                double *o = linop + (size_t)i * 4 * nz;
                o[m] = L0;
                o[nz + m] = L0;
                o[2 * nz + m] = L2;
                o[3 * nz + m] = L2;
            }
        }
    }
    if (linop_mf) {
        for (int f = 0; f < b.nMF; ++f)
            for (int m = 0; m < nz; ++m) {
                const double kz2 = -(r.h_kz[m] * r.h_kz[m]);
                double v = 0.0;
                if (p.quasilinear) v = (p.model == QL_MODEL_MHD_BQLIN || f < 2 ? p.eta : p.nu) * kz2;
                linop_mf[(size_t)f * nz + m] = v;
            }
    }
}

// launch-shape tuner: per group, time each candidate shape over trial_steps
// steps of that group alone and keep the fastest; the state is restored
// afterwards.  Shapes change the summation order of the stress reduction, so
// results agree across shapes to rounding, and bitwise for one shape.
// TUNE_WARM_MS of device time on the current shapes, on a state the tuner
// restores afterwards
int tune_warm(Batch &b)
{
    double warm_ms = 0.0;
    for (int batch = 0; batch < 256 && warm_ms < TUNE_WARM_MS; ++batch) {
        QL_HIP(hipEventRecord(b.ev[0], b.s));
        for (int n = 0; n < 8; ++n)
            for (Group &g : b.groups)
                for (int s = 1; s <= 3; ++s) QL_RC(enqueue_substep(b, g, s, false, false, b.s));
        float ms = 0.f;
        QL_HIP(hipEventRecord(b.ev[1], b.s));
        QL_HIP(hipEventSynchronize(b.ev[1]));
        QL_HIP(hipEventElapsedTime(&ms, b.ev[0], b.ev[1]));
        warm_ms += ms;
    }
    return QL_OK;
}

int tune(Batch &b, int trial_steps, double *best_ms_out)
{
    if (trial_steps < 1) return fail(QL_ERR_ARG, "tune: trial_steps < 1");
    QL_HIP(hipStreamSynchronize(b.s));
    // snapshot of what the trial steps overwrite
    struct Snap {
        double2 *u = nullptr, *mf = nullptr, *drive = nullptr;
        double *kx0 = nullptr, *pm = nullptr;
    };
    std::vector<Snap> snaps(b.groups.size());
    int rc = QL_OK;
    for (size_t gi = 0; gi < b.groups.size() && rc == QL_OK; ++gi) {
        Group &g = b.groups[gi];
        Snap &sn = snaps[gi];
        const size_t nl = (size_t)g.modes * 4 * g.nz, nm = (size_t)g.nruns * b.nMF * g.nz;
        const size_t npm = (size_t)g.nruns * qlk::PM_STRIDE;
        if ((rc = dalloc(&sn.u, nl)) || (rc = dalloc(&sn.mf, nm)) || (rc = dalloc(&sn.drive, nm)) ||
            (rc = dalloc(&sn.kx0, g.modes)) || (rc = dalloc(&sn.pm, npm)))
            break;
        if (hipMemcpyAsync(sn.pm, g.pm, sizeof(double) * npm, hipMemcpyDeviceToDevice, b.s) != hipSuccess ||
            hipMemcpyAsync(sn.u, g.u, sizeof(double2) * nl, hipMemcpyDeviceToDevice, b.s) != hipSuccess ||
            hipMemcpyAsync(sn.mf, g.mf, sizeof(double2) * nm, hipMemcpyDeviceToDevice, b.s) != hipSuccess ||
            hipMemcpyAsync(sn.drive, g.drive, sizeof(double2) * nm, hipMemcpyDeviceToDevice, b.s) != hipSuccess ||
            hipMemcpyAsync(sn.kx0, g.kx0, sizeof(double) * g.modes, hipMemcpyDeviceToDevice, b.s) != hipSuccess)
            rc = fail(QL_ERR_HIP, "tune: state snapshot failed");
    }
    double total_best = 0.0;
    const bool timing = b.timing;
    b.timing = false;
    if (rc == QL_OK) rc = set_clock(b, 0.0);
    // The GPU raises its clocks over the first ~20 ms of sustained work after
    // an idle spell of a few ms (measured at 64x128x128: 120 us per step
    // falling to 114 us over about 100 steps, from any state), so candidates
    // timed from a cold start would be compared at different clocks: the
    // current shapes run first
    if (rc == QL_OK) rc = tune_warm(b);
    for (size_t gi = 0; gi < b.groups.size() && rc == QL_OK; ++gi) {
        Group &g = b.groups[gi];
        double best = 1e300;
        int best_line = g.line, best_target = g.target, best_dft = g.dft, best_tail = g.tail;
        std::vector<std::pair<int, int>> seen;  // (kernel and MB, W) already timed
        for (int cand = 0; cand < 2 * (int)(sizeof(LINES) / sizeof(LINES[0])); ++cand) {
            const int line = LINES[cand / 2], dft = cand % 2;
            if (dft && !dft_ok(g.NZ, b.precision, line)) continue;
            for (int target : TARGETS) {
                g.dft = dft;
                if ((rc = configure(b, g, line, target))) break;
                const int kind = line == LINE_WAVE ? -1 : g.MB + 100 * dft;
                bool dup = false;
                for (auto &sw : seen) dup = dup || (sw.first == kind && sw.second == g.W);
                if (dup) continue;
                seen.emplace_back(kind, g.W);
                // both substep tails where the one-launch tail applies
                for (int tail = 0; tail <= 1 && rc == QL_OK; ++tail) {
                    if (tail && !fused_tail_fits(b, g)) continue;
                    g.tail = tail;
                    for (int s = 1; s <= 3 && rc == QL_OK; ++s) rc = enqueue_substep(b, g, s, false, false, b.s);
                    if (rc) break;
                    if (hipEventRecord(b.ev[0], b.s) != hipSuccess) {
                        rc = fail(QL_ERR_HIP, "tune: event record failed");
                        break;
                    }
                    for (int n = 0; n < trial_steps && rc == QL_OK; ++n)
                        for (int s = 1; s <= 3 && rc == QL_OK; ++s) rc = enqueue_substep(b, g, s, false, false, b.s);
                    if (rc) break;
                    float ms = 0.f;
                    if (hipEventRecord(b.ev[1], b.s) != hipSuccess || hipEventSynchronize(b.ev[1]) != hipSuccess ||
                        hipEventElapsedTime(&ms, b.ev[0], b.ev[1]) != hipSuccess) {
                        rc = fail(QL_ERR_HIP, "tune: event timing failed");
                        break;
                    }
                    const double per = ms / trial_steps;
                    if (per < best) {
                        best = per;
                        best_line = line;
                        best_target = target;
                        best_dft = dft;
                        best_tail = tail;
                    }
                }
                if (rc) break;
            }
            if (rc) break;
        }
        g.dft = best_dft;
        g.tail = best_tail;
        if (rc == QL_OK) rc = configure(b, g, best_line, best_target);
        total_best += best;
    }
    // the replay graphs of the chosen shapes, captured here, so that the
    // first stepping call after the tuner only launches
    if (rc == QL_OK) rc = ensure_graphs(b);
    b.timing = timing;
    // restore
    for (size_t gi = 0; gi < b.groups.size(); ++gi) {
        Group &g = b.groups[gi];
        Snap &sn = snaps[gi];
        const size_t nl = (size_t)g.modes * 4 * g.nz, nm = (size_t)g.nruns * b.nMF * g.nz;
        if (sn.u && sn.mf && sn.drive && sn.kx0) {
            if (hipMemcpyAsync(g.u, sn.u, sizeof(double2) * nl, hipMemcpyDeviceToDevice, b.s) != hipSuccess ||
                hipMemcpyAsync(g.mf, sn.mf, sizeof(double2) * nm, hipMemcpyDeviceToDevice, b.s) != hipSuccess ||
                hipMemcpyAsync(g.drive, sn.drive, sizeof(double2) * nm, hipMemcpyDeviceToDevice, b.s) != hipSuccess ||
                hipMemcpyAsync(g.kx0, sn.kx0, sizeof(double) * g.modes, hipMemcpyDeviceToDevice, b.s) != hipSuccess)
                if (rc == QL_OK) rc = fail(QL_ERR_HIP, "tune: state restore failed");
            g.amp_stale = true;
            if (rc == QL_OK) rc = refresh_profiles(b, g, b.s);
            // after the refresh, which rewrites the current-state maxima
            if (sn.pm && hipMemcpyAsync(g.pm, sn.pm, sizeof(double) * g.nruns * qlk::PM_STRIDE,
                                        hipMemcpyDeviceToDevice, b.s) != hipSuccess)
                if (rc == QL_OK) rc = fail(QL_ERR_HIP, "tune: state restore failed");
        }
    }
    if (hipStreamSynchronize(b.s) != hipSuccess && rc == QL_OK) rc = fail(QL_ERR_HIP, "tune: synchronize failed");
    for (Snap &sn : snaps) dfree(sn.u), dfree(sn.mf), dfree(sn.drive), dfree(sn.kx0), dfree(sn.pm);
    if (rc == QL_OK && best_ms_out) *best_ms_out = total_best;
    return rc;
}

// Scratch of ql_rhs_noalloc (the one entry point that allocates nothing): the
// mode kernel's stress partial rows, their column sums and a mean-field RHS
// slot.  Each device holds its own copy from module load until the process
// ends; every ql_rhs_noalloc call on a device uses it.
constexpr size_t NOALLOC_DOUBLES = QL_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// the device address of `scratch` on the current device (once per device):
// the array's name is a host-side stand-in that must never reach a kernel
int resolve_scratch(int dev)
{
    if (dev < 0 || dev >= MAX_DEVICES) return fail(QL_ERR_ARG, "device %d beyond %d", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev]) return QL_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(QL_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch[dev] = static_cast<double *>(ptr);
    return QL_OK;
}

int init_batch(Batch &b, const ql_params *ps, int nruns, const char *who)
{
    if (nruns < 1) return fail(QL_ERR_ARG, "%s: nruns=%d", who, nruns);
    for (int r = 0; r < nruns; ++r) {
        QL_RC(check_params(&ps[r], who));
        const ql_params &a = ps[0], &c = ps[r];
        if (c.model != a.model || c.integrator != a.integrator || c.precision != a.precision || c.device != a.device)
            return fail(QL_ERR_ARG, "%s: run %d: runs of a set share model, integrator, precision and device", who, r);
        if (c.dt != a.dt) return fail(QL_ERR_ARG, "%s: run %d: runs of a set share dt", who, r);
        for (int q = 0; q < r; ++q)
            if (ps[q].NZ == c.NZ && ps[q].L[2] != c.L[2])
                return fail(QL_ERR_ARG, "%s: runs %d and %d: runs of one NZ share Lz", who, q, r);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(QL_ERR_HIP, "%s: no HIP device", who);
    if (ps[0].device < 0 || ps[0].device >= ndev) return fail(QL_ERR_ARG, "%s: device %d of %d", who, ps[0].device, ndev);
    QL_HIP(hipSetDevice(ps[0].device));
    QL_RC(resolve_scratch(ps[0].device));
    int rc = create_batch(b, ps, nruns);
    if (rc) release(b);
    return rc;
}

// fac1D Re(ifft(pad(v))) of the caller's mean-field slots of v -> [nMF][NZ]
// (TimeVariables.cpp:202-228, HD_fullU.cpp:524-535)
int profiles_real(Batch &b, const double2 *v, const Group &g, double *out)
{
    const int n = ext_nmf(b);
    const double fac = 1.0 / g.NZ;
    switch (g.NZ) {
    case 32: qlk::k_mf_real<32><<<1, qlk::NT, 0, b.s>>>(v, n, g.tw, fac, b.mfr_tmp); break;
    case 64: qlk::k_mf_real<64><<<1, qlk::NT, 0, b.s>>>(v, n, g.tw, fac, b.mfr_tmp); break;
    case 128: qlk::k_mf_real<128><<<1, qlk::NT, 0, b.s>>>(v, n, g.tw, fac, b.mfr_tmp); break;
    default: qlk::k_mf_real<256><<<1, qlk::NT, 0, b.s>>>(v, n, g.tw, fac, b.mfr_tmp); break;
    }
    QL_HIP(hipGetLastError());
    QL_HIP(hipMemcpyAsync(out, b.mfr_tmp, sizeof(double) * n * g.NZ, hipMemcpyDeviceToHost, b.s));
    QL_HIP(hipStreamSynchronize(b.s));
    return QL_OK;
}

}  // namespace qlh

using namespace qlh;

extern "C" {

const char *ql_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// single run
// ---------------------------------------------------------------------------
int ql_create(const ql_params *p, ql_ctx **out)
{
    if (!p || !out) return fail(QL_ERR_ARG, "ql_create: null argument");
    *out = nullptr;
    ql_ctx *c = new ql_ctx();
    const int rc = init_batch(c->bt.b, p, 1, "ql_create");
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return QL_OK;
}

int ql_destroy(ql_ctx *c)
{
    if (!c) return QL_OK;
    (void)sync_all(c->bt.b);
    release(c->bt.b);
    delete c;
    return QL_OK;
}

ql_batch *ql_batch_of(ql_ctx *c) { return c ? &c->bt : nullptr; }

int ql_dims(const ql_ctx *c, int d[6])
{
    if (!c || !d) return fail(QL_ERR_ARG, "ql_dims: null argument");
    return ql_batch_dims(&c->bt, 0, d);
}

int ql_set_state(ql_ctx *c, const double *lin, const double *mf, const double *kx0)
{
    if (!c) return fail(QL_ERR_ARG, "ql_set_state: null context");
    return set_state(c->bt.b, 0, lin, mf, kx0);
}

int ql_get_state(ql_ctx *c, double *lin, double *mf, double *kx0)
{
    if (!c) return fail(QL_ERR_ARG, "ql_get_state: null context");
    return get_state(c->bt.b, 0, lin, mf, kx0);
}

int ql_linear_ops_init(ql_ctx *c, double t0, double *linop, double *linop_mf)
{
    if (!c) return fail(QL_ERR_ARG, "ql_linear_ops_init: null context");
    Batch &b = c->bt.b;
    QL_RC(ql_batch_linear_ops_init(&c->bt, t0));
    if (linop || linop_mf) {
        const RunInfo &r = b.runs[0];
        std::vector<double> k0(r.nxy);
        QL_RC(get_state(b, 0, nullptr, nullptr, k0.data()));
        std::vector<double> L(linop ? (size_t)r.nxy * 4 * r.nz : 0), Lm(linop_mf ? (size_t)b.nMF * r.nz : 0);
        linear_ops_host(b, r, k0, r.p.q * t0, linop ? L.data() : nullptr, linop_mf ? Lm.data() : nullptr);
        if (linop) lin_to_ext(L.data(), linop, r.nxy, ext_nlin(b), r.nz);
        if (linop_mf)
            std::copy(Lm.begin() + (size_t)ext_mf0(b) * r.nz, Lm.begin() + (size_t)(ext_mf0(b) + ext_nmf(b)) * r.nz,
                      linop_mf);
    }
    return QL_OK;
}

int ql_rhs(ql_ctx *c, double t, double dt_lin, double *rhs_lin, double *rhs_mf, double *linop)
{
    if (!c) return fail(QL_ERR_ARG, "ql_rhs: null context");
    Batch &b = c->bt.b;
    Group &g = b.groups[0];
    const size_t nl = (size_t)g.modes * 4 * g.nz;
    if (!g.lin_tmp) QL_RC(dalloc(&g.lin_tmp, nl));
    QL_RC(refresh_profiles(b, g, b.s));
    QL_RC(prime_split(b, g, b.s));
    qlk::ModeArgs a = mode_args(b, g);
    a.s = 0;
    a.mode = qlk::MODE_RHS;
    set_split(b, g, a);
    a.ts_x = t;
    a.dtl_x = dt_lin;
    a.n_out = g.N1;
    a.lin_out = g.lin_tmp;
    qlk::MFArgs m = mf_args(b, g);
    m.mode = qlk::MODE_RHS;
    m.mfn_out = g.mfN1;
    QL_RC(enqueue_sequence(b, g, a, m, b.s));
    const int nL = ext_nlin(b);
    std::vector<double2> tl(rhs_lin && b.hd ? nl : 0);
    std::vector<double> tL(linop && b.hd ? nl : 0);
    double2 *dl = b.hd ? tl.data() : reinterpret_cast<double2 *>(rhs_lin);
    double *dL = b.hd ? tL.data() : linop;
    if (rhs_lin) QL_HIP(hipMemcpyAsync(dl, g.N1, sizeof(double2) * nl, hipMemcpyDeviceToHost, b.s));
    if (rhs_mf)
        QL_HIP(hipMemcpyAsync(rhs_mf, g.mfN1 + (size_t)ext_mf0(b) * g.nz, sizeof(double2) * ext_nmf(b) * g.nz,
                              hipMemcpyDeviceToHost, b.s));
    if (linop) QL_HIP(hipMemcpyAsync(dL, g.lin_tmp, sizeof(double) * nl, hipMemcpyDeviceToHost, b.s));
    QL_HIP(hipStreamSynchronize(b.s));
    if (b.hd && rhs_lin) lin_to_ext(tl.data(), reinterpret_cast<double2 *>(rhs_lin), g.modes, nL, g.nz);
    if (b.hd && linop) lin_to_ext(tL.data(), linop, g.modes, nL, g.nz);
    return QL_OK;
}

int ql_rhs_noalloc(ql_ctx *c, double t, double dt_lin, double *d_rhs_lin, double *d_rhs_mf, double *d_linop)
{
    if (!c) return fail(QL_ERR_ARG, "ql_rhs_noalloc: null context");
    Batch &b = c->bt.b;
    if (b.hd) return fail(QL_ERR_UNSUPPORTED, "ql_rhs_noalloc: HD_fullU (its host layout differs from the device's)");
    Group &g = b.groups[0];
    double *const sp = (b.device >= 0 && b.device < MAX_DEVICES) ? g_scratch[b.device] : nullptr;
    if (!sp) return fail(QL_ERR_HIP, "ql_rhs_noalloc: scratch of device %d not resolved", b.device);
    const size_t ncols = (size_t)b.nMF * g.NZ;
    const size_t need = (size_t)(g.W + g.W2) * ncols + ncols + (size_t)2 * b.nMF * g.nz;
    if (need > NOALLOC_DOUBLES)
        return fail(QL_ERR_ARG,
                    "ql_rhs_noalloc: %d workgroups x %zu stress columns need %zu bytes of scratch, above the %zu "
                    "available (use a launch shape with fewer workgroups, ql_batch_set_config)",
                    g.W, ncols, need * sizeof(double), (size_t)QL_NOALLOC_SCRATCH_BYTES);
    QL_HIP(hipSetDevice(b.device));
    // every partial row and column sum is written in full before it is read,
    // so the contents left by an earlier call need no clearing
    double *part = sp, *sums = sp + (size_t)(g.W + g.W2) * ncols;
    double2 *mf_slot = reinterpret_cast<double2 *>(sums + ncols);
    QL_RC(refresh_profiles(b, g, b.s));
    QL_RC(prime_split(b, g, b.s));
    qlk::ModeArgs a = mode_args(b, g);
    a.s = 0;
    a.mode = qlk::MODE_RHS;
    set_split(b, g, a);
    a.ts_x = t;
    a.dtl_x = dt_lin;
    a.n_out = reinterpret_cast<double2 *>(d_rhs_lin);
    a.lin_out = d_linop;
    a.partials = part;
    qlk::MFArgs m = mf_args(b, g);
    m.mode = qlk::MODE_RHS;
    m.sums = sums;
    m.mfn_out = d_rhs_mf ? reinterpret_cast<double2 *>(d_rhs_mf) : mf_slot;
    return enqueue_sequence(b, g, a, m, b.s);
}

int ql_step(ql_ctx *c, double t, double *dt_out)
{
    if (!c) return fail(QL_ERR_ARG, "ql_step: null context");
    Batch &b = c->bt.b;
    QL_RC(set_clock(b, t));
    QL_RC(prime_split_all(b));
    QL_RC(enqueue_flags(b));
    for (int s = 1; s <= 3; ++s) QL_RC(enqueue_substep_all(b, s, false, false));
    b.t_prev = t;
    b.old_init = 0;
    b.nstep += 1;
    QL_HIP(hipStreamSynchronize(b.s));
    double dt = b.dt;
    if (b.vdt) {
        qlk::Clock ck{};
        QL_RC(read_clock(b, &ck));
        dt = ck.dt;  // the step's (set_clock computed it; no advance here)
    }
    b.dt_last = dt;
    if (dt_out) *dt_out = dt;
    return QL_OK;
}

int ql_steps(ql_ctx *c, int nsteps, double *t_inout)
{
    if (!c || !t_inout) return fail(QL_ERR_ARG, "ql_steps: null argument");
    return ql_batch_steps(&c->bt, nsteps, t_inout);
}

int ql_remap(ql_ctx *c, double qt)
{
    if (!c) return fail(QL_ERR_ARG, "ql_remap: null context");
    Batch &b = c->bt.b;
    const RunInfo &r = b.runs[0];
    Group &g = b.groups[r.group];
    const int grid = r.nxy < 1024 ? r.nxy : 1024;
    qlk::k_remap<<<grid, qlk::NT, 0, b.s>>>(g.u, g.kx0, g.ky, r.mode_off, r.nxy, g.nz, qt, r.remap_thr, r.remap_shift);
    QL_HIP(hipGetLastError());
    QL_HIP(hipStreamSynchronize(b.s));
    return QL_OK;
}

int ql_diagnostics(ql_ctx *c, double t, double out[18])
{
    if (!c || !out) return fail(QL_ERR_ARG, "ql_diagnostics: null argument");
    return diagnostics(c->bt.b, 0, t, out);
}

int ql_get_drive(ql_ctx *c, double *drive)
{
    if (!c || !drive) return fail(QL_ERR_ARG, "ql_get_drive: null argument");
    return ql_batch_get_drive(&c->bt, 0, drive);
}

int ql_mf_real(ql_ctx *c, double *out)
{
    if (!c || !out) return fail(QL_ERR_ARG, "ql_mf_real: null argument");
    Batch &b = c->bt.b;
    const RunInfo &r = b.runs[0];
    Group &g = b.groups[r.group];
    return profiles_real(b, g.mf + (size_t)(r.gi * b.nMF + ext_mf0(b)) * g.nz, g, out);
}

int ql_drive_real(ql_ctx *c, double *out)
{
    if (!c || !out) return fail(QL_ERR_ARG, "ql_drive_real: null argument");
    Batch &b = c->bt.b;
    const RunInfo &r = b.runs[0];
    Group &g = b.groups[r.group];
    return profiles_real(b, g.drive + (size_t)(r.gi * b.nMF + ext_mf0(b)) * g.nz, g, out);
}

int ql_calculate_cfl(ql_ctx *c, double *cfl_over_dt)
{
    if (!c || !cfl_over_dt) return fail(QL_ERR_ARG, "ql_calculate_cfl: null argument");
    return ql_batch_calculate_cfl(&c->bt, 0, cfl_over_dt);
}

int ql_check_solution(ql_ctx *c, double sums[5], int *verdict)
{
    if (!c) return fail(QL_ERR_ARG, "ql_check_solution: null context");
    return ql_batch_check_solution(&c->bt, 0, sums, verdict);
}

int ql_set_noise(ql_ctx *c, const ql_noise *cfg)
{
    if (!c) return fail(QL_ERR_ARG, "ql_set_noise: null context");
    return ql_batch_set_noise(&c->bt, 0, cfg);
}

int ql_driving_noise(ql_ctx *c, double t, double dt)
{
    if (!c) return fail(QL_ERR_ARG, "ql_driving_noise: null context");
    return ql_batch_driving_noise(&c->bt, t, dt);
}

int ql_synchronize(ql_ctx *c)
{
    if (!c) return fail(QL_ERR_ARG, "ql_synchronize: null context");
    QL_HIP(hipStreamSynchronize(c->bt.b.s));
    return QL_OK;
}

int ql_profile_kernels(ql_ctx *c, int nsteps, double *t_inout, double avg_ms[2], int counts[2])
{
    if (!c || !t_inout || !avg_ms || !counts) return fail(QL_ERR_ARG, "ql_profile_kernels: null argument");
    if (nsteps < 0) return fail(QL_ERR_ARG, "ql_profile_kernels: nsteps < 0");
    Batch &b = c->bt.b;
    for (int k = 0; k < 2; ++k) {
        b.t_ms[k] = 0.0;
        b.t_cnt[k] = 0;
    }
    b.timing = true;
    b.tused = 0;
    int rc = run_steps(b, nsteps, t_inout);
    b.timing = false;
    if (rc) return rc;
    QL_RC(sync_all(b));
    for (size_t k = 0; k + 3 <= b.tused; k += 3) {
        float ms0 = 0.f, ms1 = 0.f;
        QL_HIP(hipEventElapsedTime(&ms0, b.tpool[k], b.tpool[k + 1]));
        QL_HIP(hipEventElapsedTime(&ms1, b.tpool[k + 1], b.tpool[k + 2]));
        b.t_ms[0] += ms0;
        b.t_ms[1] += ms1;
        b.t_cnt[0] += 1;
        b.t_cnt[1] += 1;
    }
    b.tused = 0;
    for (int k = 0; k < 2; ++k) {
        counts[k] = b.t_cnt[k];
        avg_ms[k] = b.t_cnt[k] ? b.t_ms[k] / b.t_cnt[k] : 0.0;
    }
    return QL_OK;
}

// ---------------------------------------------------------------------------
// sets of runs
// ---------------------------------------------------------------------------
int ql_batch_create(const ql_params *runs, int nruns, ql_batch **out)
{
    if (!runs || !out) return fail(QL_ERR_ARG, "ql_batch_create: null argument");
    *out = nullptr;
    ql_batch *bt = new ql_batch();
    const int rc = init_batch(bt->b, runs, nruns, "ql_batch_create");
    if (rc) {
        delete bt;
        return rc;
    }
    *out = bt;
    return QL_OK;
}

int ql_batch_destroy(ql_batch *bt)
{
    if (!bt) return QL_OK;
    (void)sync_all(bt->b);
    release(bt->b);
    delete bt;
    return QL_OK;
}

int ql_batch_size(const ql_batch *bt, int *nruns)
{
    if (!bt || !nruns) return fail(QL_ERR_ARG, "ql_batch_size: null argument");
    *nruns = (int)bt->b.runs.size();
    return QL_OK;
}

int ql_batch_dims(const ql_batch *bt, int run, int d[6])
{
    if (!bt || !d) return fail(QL_ERR_ARG, "ql_batch_dims: null argument");
    QL_RC(check_run(bt->b, run, "ql_batch_dims"));
    const RunInfo &r = bt->b.runs[run];
    d[0] = r.nxy;
    d[1] = r.nz;
    d[2] = ext_nmf(bt->b);
    d[3] = ext_nlin(bt->b);
    d[4] = r.half;
    d[5] = r.p.NZ;
    return QL_OK;
}

int ql_batch_set_state(ql_batch *bt, int run, const double *lin, const double *mf, const double *kx0)
{
    if (!bt) return fail(QL_ERR_ARG, "ql_batch_set_state: null handle");
    QL_RC(check_run(bt->b, run, "ql_batch_set_state"));
    return set_state(bt->b, run, lin, mf, kx0);
}

int ql_batch_get_state(ql_batch *bt, int run, double *lin, double *mf, double *kx0)
{
    if (!bt) return fail(QL_ERR_ARG, "ql_batch_get_state: null handle");
    QL_RC(check_run(bt->b, run, "ql_batch_get_state"));
    return get_state(bt->b, run, lin, mf, kx0);
}

int ql_batch_linear_ops_init(ql_batch *bt, double t0)
{
    if (!bt) return fail(QL_ERR_ARG, "ql_batch_linear_ops_init: null handle");
    bt->b.old_init = 1;
    bt->b.nstep = 0;
    bt->b.t_init = t0;
    bt->b.t_prev = t0;
    bt->b.rt_prev.assign(bt->b.runs.size(), t0);
    bt->b.rdt_last.assign(bt->b.runs.size(), bt->b.dt_last);
    return QL_OK;
}

int ql_batch_steps(ql_batch *bt, int nsteps, double *t_inout)
{
    if (!bt || !t_inout) return fail(QL_ERR_ARG, "ql_batch_steps: null argument");
    if (nsteps < 0) return fail(QL_ERR_ARG, "ql_batch_steps: nsteps < 0");
    Batch &b = bt->b;
    if (b.vdt && b.runs.size() > 1)
        return fail(QL_ERR_ARG, "ql_batch_steps: the runs of a variable-step set keep their own times "
                                "(ql_batch_steps_runs)");
    std::vector<double> t(b.runs.size(), *t_inout);
    QL_RC(run_steps(b, nsteps, t.data()));
    *t_inout = t[0];
    return QL_OK;
}

int ql_batch_steps_runs(ql_batch *bt, int nsteps, double *t_runs)
{
    if (!bt || !t_runs) return fail(QL_ERR_ARG, "ql_batch_steps_runs: null argument");
    if (nsteps < 0) return fail(QL_ERR_ARG, "ql_batch_steps_runs: nsteps < 0");
    Batch &b = bt->b;
    if (!b.vdt)
        for (size_t r = 1; r < b.runs.size(); ++r)
            if (t_runs[r] != t_runs[0])
                return fail(QL_ERR_ARG, "ql_batch_steps_runs: the runs of a fixed-step set share one time");
    return run_steps(b, nsteps, t_runs);
}

int ql_batch_diagnostics(ql_batch *bt, int run, double t, double out[18])
{
    if (!bt || !out) return fail(QL_ERR_ARG, "ql_batch_diagnostics: null argument");
    QL_RC(check_run(bt->b, run, "ql_batch_diagnostics"));
    return diagnostics(bt->b, run, t, out);
}

int ql_batch_get_drive(ql_batch *bt, int run, double *drive)
{
    if (!bt || !drive) return fail(QL_ERR_ARG, "ql_batch_get_drive: null argument");
    QL_RC(check_run(bt->b, run, "ql_batch_get_drive"));
    Batch &b = bt->b;
    const RunInfo &r = b.runs[run];
    Group &g = b.groups[r.group];
    QL_HIP(hipMemcpyAsync(drive, g.drive + (size_t)(r.gi * b.nMF + ext_mf0(b)) * g.nz,
                          sizeof(double2) * ext_nmf(b) * g.nz, hipMemcpyDeviceToHost, b.s));
    QL_HIP(hipStreamSynchronize(b.s));
    return QL_OK;
}

int ql_batch_calculate_cfl(ql_batch *bt, int run, double *cfl_over_dt)
{
    if (!bt || !cfl_over_dt) return fail(QL_ERR_ARG, "ql_batch_calculate_cfl: null argument");
    QL_RC(check_run(bt->b, run, "ql_batch_calculate_cfl"));
    Batch &b = bt->b;
    const RunInfo &r = b.runs[run];
    Group &g = b.groups[r.group];
    // with fixed steps nothing is recorded while stepping (the reference calls
    // Calculate_CFL for variable steps only): evaluate the current state's
    // profiles and the last rhs's drives now
    if (!b.vdt) QL_RC(refresh_profiles(b, g, b.s, true));
    double pm[qlk::PM_STRIDE];
    QL_HIP(hipMemcpyAsync(pm, g.pm + (size_t)r.gi * qlk::PM_STRIDE, sizeof(pm), hipMemcpyDeviceToHost, b.s));
    QL_HIP(hipStreamSynchronize(b.s));
    if (!b.vdt) {
        std::copy(pm + qlk::PM_CUR, pm + qlk::PM_CUR + 8, pm + qlk::PM_RHS);
        pm[qlk::PM_E_RHS] = pm[qlk::PM_E_CUR];
    }
    *cfl_over_dt = qlk::cfl_value(pm, g.h_rp[r.gi], b.model);
    return QL_OK;
}

int ql_batch_check_solution(ql_batch *bt, int run, double sums[5], int *verdict)
{
    if (!bt) return fail(QL_ERR_ARG, "ql_batch_check_solution: null handle");
    QL_RC(check_run(bt->b, run, "ql_batch_check_solution"));
    Batch &b = bt->b;
    const RunInfo &r = b.runs[run];
    Group &g = b.groups[r.group];
    // solution.cpp:170: the caller's mean fields 1 (checked) and 0
    const double2 *mf = g.mf + (size_t)(r.gi * b.nMF + ext_mf0(b)) * g.nz;
    switch (g.NZ) {
    case 32: qlk::k_check_sums<32><<<1, qlk::NT, 0, b.s>>>(mf, g.tw, b.dout); break;
    case 64: qlk::k_check_sums<64><<<1, qlk::NT, 0, b.s>>>(mf, g.tw, b.dout); break;
    case 128: qlk::k_check_sums<128><<<1, qlk::NT, 0, b.s>>>(mf, g.tw, b.dout); break;
    default: qlk::k_check_sums<256><<<1, qlk::NT, 0, b.s>>>(mf, g.tw, b.dout); break;
    }
    QL_HIP(hipGetLastError());
    double sm[5];
    QL_HIP(hipMemcpyAsync(sm, b.dout, sizeof(sm), hipMemcpyDeviceToHost, b.s));
    QL_HIP(hipStreamSynchronize(b.s));
    if (sums) std::memcpy(sums, sm, sizeof(sm));
    // solution.cpp:174, :184, :194, in that order
    int v = 0;
    if (sm[0] / sm[1] < 1e10 && sm[0] > 1e-2) v = 1;
    else if (sm[2] > 1e20 || sm[3] > 1e20) v = 2;
    else if (!std::isfinite(sm[4])) v = 3;
    if (verdict) *verdict = v;
    if (v == 1) return fail(QL_ERR_STATE, "run %d: ERROR: Mean field has developed an imaginary part!!", run);
    if (v == 2) return fail(QL_ERR_STATE, "run %d: ERROR: Solution is very large, probably unstable!!", run);
    if (v == 3) return fail(QL_ERR_STATE, "run %d: ERROR: Solution contains NaN or Inf!!", run);
    return QL_OK;
}

int ql_batch_set_noise(ql_batch *bt, int run, const ql_noise *cfg)
{
    if (!bt || !cfg) return fail(QL_ERR_ARG, "ql_batch_set_noise: null argument");
    QL_RC(check_run(bt->b, run, "ql_batch_set_noise"));
    if (!(cfg->f_noise >= 0) || !std::isfinite(cfg->f_noise))
        return fail(QL_ERR_ARG, "ql_batch_set_noise: f_noise=%g must be finite and >= 0", cfg->f_noise);
    if (std::isnan(cfg->range_low) || std::isnan(cfg->range_high))
        return fail(QL_ERR_ARG, "ql_batch_set_noise: NaN noise range");
    Batch &b = bt->b;
    const RunInfo &r = b.runs[run];
    Group &g = b.groups[r.group];
    QL_RC(sync_all(b));
    noise_params(g.h_np[r.gi], r, *cfg);
    g.noise = false;
    for (const qlk::NoiseParams &n : g.h_np) g.noise = g.noise || n.sigma != 0.0;
    QL_HIP(hipMemcpy(g.np, g.h_np.data(), sizeof(qlk::NoiseParams) * g.nruns, hipMemcpyHostToDevice));
    drop_graphs(b);  // the noise kernel joins or leaves the captured step
    return QL_OK;
}

int ql_batch_driving_noise(ql_batch *bt, double t, double dt)
{
    if (!bt) return fail(QL_ERR_ARG, "ql_batch_driving_noise: null handle");
    Batch &b = bt->b;
    for (Group &g : b.groups)
        if (g.noise) QL_RC(launch_noise(g, false, t, dt, true, false, b.s));
    QL_HIP(hipStreamSynchronize(b.s));
    return QL_OK;
}

int ql_batch_set_transform(ql_batch *bt, int NZ, int transform)
{
    if (!bt) return fail(QL_ERR_ARG, "ql_batch_set_transform: null handle");
    Batch &b = bt->b;
    for (Group &g : b.groups) {
        if (g.NZ != NZ) continue;
        if (transform != QL_TRANSFORM_FFT && transform != QL_TRANSFORM_DFT_MFMA)
            return fail(QL_ERR_ARG, "ql_batch_set_transform: unknown transform %d", transform);
        if (transform == QL_TRANSFORM_DFT_MFMA && !dft_ok(g.NZ, b.precision, g.line))
            return fail(QL_ERR_ARG,
                        "ql_batch_set_transform: the matrix-core transform needs FP64, NZ 32 or 64 and line 128 "
                        "or 256 (NZ %d, line %d)", g.NZ, g.line);
        g.dft = transform;
        return configure(b, g, g.line, g.target);
    }
    return fail(QL_ERR_ARG, "ql_batch_set_transform: no group with NZ = %d", NZ);
}

int ql_batch_set_tail(ql_batch *bt, int NZ, int tail)
{
    if (!bt) return fail(QL_ERR_ARG, "ql_batch_set_tail: null handle");
    Batch &b = bt->b;
    for (Group &g : b.groups) {
        if (g.NZ != NZ) continue;
        if (tail != QL_TAIL_TWO_KERNELS && tail != QL_TAIL_FUSED)
            return fail(QL_ERR_ARG, "ql_batch_set_tail: unknown tail %d", tail);
        if (tail == QL_TAIL_FUSED && !fused_tail_fits(b, g))
            return fail(QL_ERR_ARG,
                        "ql_batch_set_tail: the one-launch tail needs fixed steps, NZ <= 128 and at most %ld "
                        "partial rows per run (NZ %d)", TAIL_MAX_ROWS_NZ / g.NZ, g.NZ);
        g.tail = tail;
        return configure(b, g, g.line, g.target);
    }
    return fail(QL_ERR_ARG, "ql_batch_set_tail: no group with NZ = %d", NZ);
}

int ql_batch_mixed_split(ql_batch *bt, double tau, int *nfp64)
{
    if (!bt) return fail(QL_ERR_ARG, "ql_batch_mixed_split: null handle");
    Batch &b = bt->b;
    if (b.precision != QL_PREC_MIXED) return fail(QL_ERR_ARG, "ql_batch_mixed_split: not a mixed-precision set");
    if (!(tau == tau) || std::isinf(tau)) return fail(QL_ERR_ARG, "ql_batch_mixed_split: tau %g not finite", tau);
    if (tau > 0.0) {
        b.amp_tau = tau;
        drop_graphs(b);  // the fraction is a launch argument of the captured steps
        for (Group &g : b.groups) g.amp_stale = true;
    }
    if (nfp64) {
        QL_RC(prime_split_all(b));
        int n = 0;
        for (Group &g : b.groups) {
            if (!split_active(b, g)) continue;  // tau > 1: every mode in FP32
            std::vector<unsigned char> hi(g.modes);
            QL_HIP(hipMemcpyAsync(hi.data(), g.hi, g.modes, hipMemcpyDeviceToHost, b.s));
            QL_HIP(hipStreamSynchronize(b.s));
            for (int i = 0; i < g.modes; ++i) n += hi[i] ? 1 : 0;
        }
        *nfp64 = n;
    }
    return QL_OK;
}

int ql_batch_tune(ql_batch *bt, int trial_steps, double *step_ms)
{
    if (!bt) return fail(QL_ERR_ARG, "ql_batch_tune: null handle");
    return tune(bt->b, trial_steps, step_ms);
}

int ql_batch_get_config(const ql_batch *bt, int max_groups, int *ngroups, int *cfg)
{
    if (!bt || !ngroups) return fail(QL_ERR_ARG, "ql_batch_get_config: null argument");
    const Batch &b = bt->b;
    *ngroups = (int)b.groups.size();
    if (cfg)
        for (int k = 0; k < max_groups && k < (int)b.groups.size(); ++k) {
            const Group &g = b.groups[k];
            int *c = cfg + (size_t)QL_CONFIG_INTS * k;
            c[0] = g.NZ;
            c[1] = g.line;
            c[2] = g.target;
            c[3] = g.MB;
            c[4] = g.W;
            c[5] = g.dft;
            c[6] = fused_tail(b, g) ? QL_TAIL_FUSED : QL_TAIL_TWO_KERNELS;
        }
    return QL_OK;
}

int ql_batch_set_config(ql_batch *bt, int NZ, int line, int target)
{
    if (!bt) return fail(QL_ERR_ARG, "ql_batch_set_config: null handle");
    if (line != LINE_WAVE && line != 128 && line != 256)
        return fail(QL_ERR_ARG, "ql_batch_set_config: line=%d not in {64, 128, 256}", line);
    if (target < 1) return fail(QL_ERR_ARG, "ql_batch_set_config: target=%d < 1", target);
    Batch &b = bt->b;
    for (Group &g : b.groups)
        if (g.NZ == NZ) return configure(b, g, line, target);
    return fail(QL_ERR_ARG, "ql_batch_set_config: no runs with NZ=%d", NZ);
}

}  // extern "C"
