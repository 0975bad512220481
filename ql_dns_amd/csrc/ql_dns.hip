// ql_dns.hip -- host side of the MI355X quasilinear time step: the run context,
// the per-substep launch sequence and the C ABI declared in include/ql_dns.h.
//
// Reference counterparts: the Model objects (Models/MHD_FullQlin.cpp,
// Models/MHD_BQlin.cpp) become kernels plus device-resident tables; the
// Integrator objects (Integrators/RK3CN.cpp, Integrators/RK3CN_Elin.cpp) become
// the substep sequence in run_substep(); the MPI Allreduce of the stresses
// (MHD_FullQlin.cpp:386) becomes the on-device fixed-order reduction.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ql_dns.h"
#include "ql_kernels.hpp"
#include "ql_mode.hpp"

namespace {

thread_local std::string g_err;

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

#define QL_HIP(call)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return fail(QL_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// Integrator coefficients (Integrators/RK3CN.cpp:19-20, RK3CN_Elin.cpp:19-20)
constexpr double A0 = 8.0 / 15.0, A1 = 5.0 / 12.0, A2 = 3.0 / 4.0;
constexpr double B1 = -17.0 / 60.0, B2 = -5.0 / 12.0;
constexpr double P1 = 8.0 / 15.0, P2 = 2.0 / 3.0;
// General_Definitions.h:236: PI = atan(1)*4, which rounds to this double
constexpr double QL_PI = 3.14159265358979323846;

}  // namespace

struct ql_ctx {
    ql_params p;
    int nxy, nz, nzl, half, NZ, nMF, MB, G, vo2;
    double fac, rsf, kxLfac, kyLfac, kzLfac, omega_eff;
    double remap_thr, remap_shift, totalN2, kzfac;
    hipStream_t s = nullptr;
    // device state and tables
    double2 *u = nullptr, *N1 = nullptr, *N2 = nullptr;
    double2 *mf = nullptr, *mfN1 = nullptr, *mfN2 = nullptr, *drive = nullptr, *tw = nullptr;
    float2 *tw32 = nullptr;
    double *kx0 = nullptr, *ky = nullptr, *kzv = nullptr, *mfz = nullptr;
    double *part = nullptr, *part2 = nullptr, *lin_tmp = nullptr, *dpart = nullptr, *dout = nullptr;
    double *mfr_tmp = nullptr;
    // integrator bookkeeping: where the RK3CN "old" operator was evaluated
    int old_init = 1;
    double qt_old = 0.0, qdtl_old = 0.0;
    // host copies of the tables
    std::vector<double> h_kx0, h_ky, h_kz;
    // optional per-kernel timing
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double t_ms[2] = {0, 0};
    int t_cnt[2] = {0, 0};
};

namespace {

// modes per workgroup of the per-mode kernels: MB * NZ = QL_MB_LINE points
#ifndef QL_MB_LINE
#define QL_MB_LINE 256
#endif
// upper bound on the per-mode kernel grid
#ifndef QL_GMAX
#define QL_GMAX 512
#endif

bool valid_NZ(int NZ) { return NZ == 32 || NZ == 64 || NZ == 128 || NZ == 256; }

constexpr int mb_of(int NZ) { return NZ >= QL_MB_LINE ? 1 : QL_MB_LINE / NZ; }
int mb_for(int NZ) { return mb_of(NZ); }

int begin_timing(ql_ctx *c)
{
    if (c->timing) QL_HIP(hipEventRecord(c->ev[0], c->s));
    return QL_OK;
}

int end_timing(ql_ctx *c, int kind)
{
    if (!c->timing) return QL_OK;
    QL_HIP(hipEventRecord(c->ev[1], c->s));
    QL_HIP(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    QL_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->t_ms[kind] += ms;
    c->t_cnt[kind] += 1;
    return QL_OK;
}

template <int NZ>
void launch_mode(ql_ctx *c, const qlk::ModeArgs &a)
{
    constexpr int MB = mb_of(NZ);
    const bool full = (c->p.model == QL_MODEL_MHD_FULLQLIN);
    if (c->p.precision == QL_PREC_MIXED) {
        if (full) qlk::k_mode32<NZ, MB, qlk::ModelFull><<<c->G, qlk::NT, 0, c->s>>>(a);
        else qlk::k_mode32<NZ, MB, qlk::ModelBQ><<<c->G, qlk::NT, 0, c->s>>>(a);
    } else {
        if (full) qlk::k_mode64<NZ, MB, qlk::ModelFull><<<c->G, qlk::NT, 0, c->s>>>(a);
        else qlk::k_mode64<NZ, MB, qlk::ModelBQ><<<c->G, qlk::NT, 0, c->s>>>(a);
    }
}

template <int NZ>
void launch_mf(ql_ctx *c, const qlk::MFArgs &a)
{
    qlk::k_meanfield<NZ><<<1, qlk::NT, 0, c->s>>>(a);
}

template <int NZ>
void launch_mfreal(ql_ctx *c, double *out)
{
    qlk::k_mf_real<NZ><<<1, qlk::NT, 0, c->s>>>(c->mf, c->nMF, c->tw, c->fac, out);
}

#define QL_DISPATCH_NZ(NZv, FN, ...)               \
    switch (NZv) {                                 \
    case 32: FN<32>(__VA_ARGS__); break;           \
    case 64: FN<64>(__VA_ARGS__); break;           \
    case 128: FN<128>(__VA_ARGS__); break;         \
    default: FN<256>(__VA_ARGS__); break;          \
    }

qlk::ModeArgs mode_args(ql_ctx *c)
{
    qlk::ModeArgs a{};
    a.u = c->u;
    a.kx0 = c->kx0;
    a.ky = c->ky;
    a.kzv = c->kzv;
    a.tw = c->tw;
    a.tw32 = c->tw32;
    a.mfz = c->mfz;
    a.partials = c->part;
    a.nxy = c->nxy;
    a.vo2 = c->vo2;
    a.q = c->p.q;
    a.omega = c->omega_eff;
    a.B0z = c->p.B0z;
    a.nu = c->p.nu;
    a.eta = c->p.eta;
    a.fac = c->fac;
    a.rsf = c->rsf;
    a.remap_thr = c->remap_thr;
    a.remap_shift = c->remap_shift;
    return a;
}

qlk::MFArgs mf_args(ql_ctx *c)
{
    qlk::MFArgs m{};
    m.sums = c->part2;
    m.mf = c->mf;
    m.drive = c->drive;
    m.mfz = c->mfz;
    m.kzv = c->kzv;
    m.tw = c->tw;
    m.model = c->p.model;
    m.nMF = c->nMF;
    m.quasilinear = c->p.quasilinear;
    m.q = c->p.q;
    m.omega = c->omega_eff;
    m.B0z = c->p.B0z;
    m.nu = c->p.nu;
    m.eta = c->p.eta;
    m.fac = c->fac;
    return m;
}

// mode kernel -> column reduction of its partials -> mean-field kernel
int run_sequence(ql_ctx *c, const qlk::ModeArgs &a, const qlk::MFArgs &m)
{
    int rc;
    if ((rc = begin_timing(c))) return rc;
    QL_DISPATCH_NZ(c->NZ, launch_mode, c, a);
    QL_HIP(hipGetLastError());
    if ((rc = end_timing(c, 0))) return rc;
    if ((rc = begin_timing(c))) return rc;
    const int ncols = c->nMF * c->NZ;
    qlk::k_reduce_cols<<<(ncols + 15) / 16, qlk::NT, 0, c->s>>>(c->part, c->G, ncols, c->part2);
    QL_HIP(hipGetLastError());
    QL_DISPATCH_NZ(c->NZ, launch_mf, c, m);
    QL_HIP(hipGetLastError());
    return end_timing(c, 1);
}

int refresh_profiles(ql_ctx *c)
{
    qlk::MFArgs m = mf_args(c);
    m.mode = qlk::MODE_PROFILES;
    m.mfn_out = c->mfN1;
    QL_DISPATCH_NZ(c->NZ, launch_mf, c, m);
    QL_HIP(hipGetLastError());
    return QL_OK;
}

// One substep s (1..3) of RK3CN_Elin::Step (RK3CN_Elin.cpp:290-378) or
// RK3CN::Step (RK3CN.cpp:92-193) at step start t with step dt.
int run_substep(ql_ctx *c, int s, double t, double dt, bool remap_after, double t_new)
{
    const bool cn = (c->p.integrator == QL_INT_RK3CN);
    double lc, nlc1, nlc2 = 0.0, ts, dtl;
    if (s == 1) {
        lc = cn ? dt / 2 * (A0) : dt * (A0);
        nlc1 = A0 * dt;
        ts = t;
        dtl = dt * P1;
    } else if (s == 2) {
        lc = cn ? dt / 2 * (A1 + B1) : dt * (A1 + B1);
        nlc1 = A1 * dt;
        nlc2 = B1 * dt;
        ts = t + dt * P1;
        dtl = dt * (P2 - P1);
    } else {
        lc = cn ? dt / 2 * (A2 + B2) : dt * (A2 + B2);
        nlc1 = A2 * dt;
        nlc2 = B2 * dt;
        ts = t + dt * P2;
        dtl = dt * (1.0 - P2);
    }
    qlk::ModeArgs a = mode_args(c);
    a.mode = cn ? qlk::MODE_CN : qlk::MODE_ELIN;
    a.qt = c->p.q * ts;
    a.qdtl = c->p.q * dtl;
    a.linC = lc;
    a.nlc1 = nlc1;
    a.nlc2 = nlc2;
    a.n_prev = (s == 1) ? nullptr : (s == 2 ? c->N1 : c->N2);
    a.n_out = (s == 1) ? c->N1 : (s == 2 ? c->N2 : nullptr);
    a.old_init = c->old_init;
    a.qt_old = c->qt_old;
    a.qdtl_old = c->qdtl_old;
    a.do_remap = remap_after ? 1 : 0;
    a.remap_qt = c->p.q * t_new;

    qlk::MFArgs m = mf_args(c);
    m.mode = a.mode;
    m.linC = lc;
    m.nlc1 = nlc1;
    m.nlc2 = nlc2;
    m.mfn_prev = (s == 1) ? nullptr : (s == 2 ? c->mfN1 : c->mfN2);
    m.mfn_out = (s == 2) ? c->mfN2 : c->mfN1;
    int rc = run_sequence(c, a, m);
    if (rc) return rc;
    c->old_init = 0;
    c->qt_old = a.qt;
    c->qdtl_old = a.qdtl;
    return QL_OK;
}

int do_step(ql_ctx *c, double t, bool remap, double t_new)
{
    int rc;
    for (int s = 1; s <= 3; ++s)
        if ((rc = run_substep(c, s, t, c->p.dt, remap && s == 3, t_new))) return rc;
    return QL_OK;
}

template <typename T>
int dalloc(T **p, size_t n)
{
    QL_HIP(hipMalloc((void **)p, sizeof(T) * (n ? n : 1)));
    QL_HIP(hipMemset(*p, 0, sizeof(T) * (n ? n : 1)));
    return QL_OK;
}

void release(ql_ctx *c)
{
    void *bufs[] = {c->u, c->N1, c->N2, c->mf, c->mfN1, c->mfN2, c->drive, c->tw, c->tw32, c->kx0, c->ky,
                    c->kzv, c->mfz, c->part, c->part2, c->lin_tmp, c->dpart, c->dout, c->mfr_tmp};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->s) (void)hipStreamDestroy(c->s);
}

}  // namespace

extern "C" {

const char *ql_last_error(void) { return g_err.c_str(); }

int ql_create(const ql_params *p, ql_ctx **out)
{
    if (!p || !out) return fail(QL_ERR_ARG, "ql_create: null argument");
    *out = nullptr;
    if (p->model != QL_MODEL_MHD_FULLQLIN && p->model != QL_MODEL_MHD_BQLIN)
        return fail(QL_ERR_ARG, "ql_create: unknown model %d", p->model);
    if (p->integrator != QL_INT_RK3CN && p->integrator != QL_INT_RK3CN_ELIN)
        return fail(QL_ERR_ARG, "ql_create: unknown integrator %d", p->integrator);
    if (p->precision != QL_PREC_FP64 && p->precision != QL_PREC_MIXED)
        return fail(QL_ERR_UNSUPPORTED, "ql_create: precision %d is not available in this build", p->precision);
    if (!valid_NZ(p->NZ)) return fail(QL_ERR_ARG, "ql_create: NZ=%d not in {32,64,128,256}", p->NZ);
    if (p->NX < 2 || p->NY < 2 || (p->NX & 1) || (p->NY & 1))
        return fail(QL_ERR_ARG, "ql_create: NX=%d, NY=%d must be even and >= 2", p->NX, p->NY);
    if (!(p->L[0] > 0) || !(p->L[1] > 0) || !(p->L[2] > 0)) return fail(QL_ERR_ARG, "ql_create: box sizes must be > 0");
    if (!(p->dt > 0))
        return fail(QL_ERR_UNSUPPORTED, "ql_create: dt=%g; variable (CFL) time steps are not supported", p->dt);
    if (p->viscosity_order < 0) return fail(QL_ERR_ARG, "ql_create: viscosity_order < 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(QL_ERR_HIP, "ql_create: no HIP device");
    if (p->device < 0 || p->device >= ndev) return fail(QL_ERR_ARG, "ql_create: device %d of %d", p->device, ndev);
    QL_HIP(hipSetDevice(p->device));

    ql_ctx *c = new ql_ctx();
    c->p = *p;
    c->NZ = p->NZ;
    c->nxy = (p->NX - 1) * (p->NY / 2);         // Model.h:32-36
    c->nz = 1 + 2 * (int)(p->NZ / 3.0);         // Model.h:34
    c->nzl = (c->nz - 1) / 2;
    c->half = c->nzl + 1;
    c->nMF = (p->model == QL_MODEL_MHD_BQLIN) ? 2 : 4;
    c->vo2 = p->viscosity_order / 2;
    c->fac = 1.0 / p->NZ;                        // fftwPlans.cpp:30
    c->rsf = 1.0 / (p->NY * p->NX) / (p->NY * p->NX);  // MHD_FullQlin.cpp:67
    c->kxLfac = 2 * QL_PI / p->L[0];
    c->kyLfac = 2 * QL_PI / p->L[1];
    c->kzLfac = 2 * QL_PI / p->L[2];
    c->omega_eff = (p->model == QL_MODEL_MHD_BQLIN) ? 2.0 / 3.0 * p->q : p->omega;  // MHD_BQlin.cpp:19
    {
        const int nxr = p->NX - 1;                                  // MHD_FullQlin.cpp:466
        c->remap_thr = nxr / 2.0 * (2 * QL_PI / p->L[0]);
        c->remap_shift = nxr * (2 * QL_PI / p->L[0]);
    }
    {
        long tn = (long)(p->NX * p->NY * p->NZ);
        tn = tn * tn;
        c->totalN2 = (double)tn;
    }
    c->kzfac = 2 * QL_PI / p->L[2] * (p->NZ / 3);
    c->MB = mb_for(p->NZ);
    const int ngroups = (c->nxy + c->MB - 1) / c->MB;
    c->G = ngroups < QL_GMAX ? ngroups : QL_GMAX;

    // wavenumber tables, Kdata.cpp:28-80
    const int nx = p->NX / 2, ny = p->NY / 2;
    c->h_kx0.assign(c->nxy, 0.0);
    c->h_ky.assign(c->nxy, 0.0);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            c->h_kx0[j + ny * i] = i * c->kxLfac;
            c->h_ky[j + ny * i] = j * c->kyLfac;
        }
    for (int i = nx + 1; i < 2 * nx; ++i)
        for (int j = 0; j < ny; ++j) {
            c->h_kx0[j + ny * (i - 1)] = (-2 * nx + i) * c->kxLfac;
            c->h_ky[j + ny * (i - 1)] = j * c->kyLfac;
        }
    c->h_kz.assign(c->nz, 0.0);
    for (int i = 0; i < c->nzl + 1; ++i) c->h_kz[i] = i * c->kzLfac;
    for (int i = c->nzl + 1; i < 2 * c->nzl + 1; ++i) c->h_kz[i] = (-2 * c->nzl + i - 1) * c->kzLfac;
    std::vector<double2> htw(c->NZ);
    std::vector<float2> htw32(c->NZ);
    for (int k = 0; k < c->NZ; ++k) {
        const double ang = 2.0 * QL_PI * (double)k / (double)c->NZ;
        htw[k] = make_double2(cos(ang), -sin(ang));
        htw32[k] = make_float2((float)htw[k].x, (float)htw[k].y);
    }

    int rc = QL_OK;
    const size_t nl = (size_t)c->nxy * 4 * c->nz;
    if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) rc = fail(QL_ERR_HIP, "stream create failed");
    if (!rc) rc = dalloc(&c->u, nl);
    if (!rc) rc = dalloc(&c->N1, nl);
    if (!rc) rc = dalloc(&c->N2, nl);
    if (!rc) rc = dalloc(&c->lin_tmp, nl);
    if (!rc) rc = dalloc(&c->mf, (size_t)c->nMF * c->nz);
    if (!rc) rc = dalloc(&c->mfN1, (size_t)c->nMF * c->nz);
    if (!rc) rc = dalloc(&c->mfN2, (size_t)c->nMF * c->nz);
    if (!rc) rc = dalloc(&c->drive, (size_t)c->nMF * c->nz);
    if (!rc) rc = dalloc(&c->tw, (size_t)c->NZ);
    if (!rc) rc = dalloc(&c->tw32, (size_t)c->NZ);
    if (!rc) rc = dalloc(&c->kx0, (size_t)c->nxy);
    if (!rc) rc = dalloc(&c->ky, (size_t)c->nxy);
    if (!rc) rc = dalloc(&c->kzv, (size_t)c->nz);
    if (!rc) rc = dalloc(&c->mfz, (size_t)8 * c->NZ);
    if (!rc) rc = dalloc(&c->part, (size_t)c->G * c->nMF * c->NZ);
    if (!rc) rc = dalloc(&c->part2, (size_t)c->nMF * c->NZ);
    if (!rc) rc = dalloc(&c->dpart, (size_t)256 * 6);
    if (!rc) rc = dalloc(&c->dout, (size_t)18);
    if (!rc) rc = dalloc(&c->mfr_tmp, (size_t)4 * c->NZ);
    if (!rc && hipMemcpy(c->kx0, c->h_kx0.data(), sizeof(double) * c->nxy, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(QL_ERR_HIP, "table upload failed");
    if (!rc && hipMemcpy(c->ky, c->h_ky.data(), sizeof(double) * c->nxy, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(QL_ERR_HIP, "table upload failed");
    if (!rc && hipMemcpy(c->kzv, c->h_kz.data(), sizeof(double) * c->nz, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(QL_ERR_HIP, "table upload failed");
    if (!rc && hipMemcpy(c->tw, htw.data(), sizeof(double2) * c->NZ, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(QL_ERR_HIP, "table upload failed");
    if (!rc && hipMemcpy(c->tw32, htw32.data(), sizeof(float2) * c->NZ, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(QL_ERR_HIP, "table upload failed");
    if (!rc && (hipEventCreate(&c->ev[0]) != hipSuccess || hipEventCreate(&c->ev[1]) != hipSuccess))
        rc = fail(QL_ERR_HIP, "event create failed");
    if (rc) {
        release(c);
        delete c;
        return rc;
    }
    *out = c;
    return QL_OK;
}

int ql_destroy(ql_ctx *c)
{
    if (!c) return QL_OK;
    if (c->s) (void)hipStreamSynchronize(c->s);
    release(c);
    delete c;
    return QL_OK;
}

int ql_dims(const ql_ctx *c, int d[6])
{
    if (!c || !d) return fail(QL_ERR_ARG, "ql_dims: null argument");
    d[0] = c->nxy;
    d[1] = c->nz;
    d[2] = c->nMF;
    d[3] = 4;
    d[4] = c->half;
    d[5] = c->NZ;
    return QL_OK;
}

int ql_set_state(ql_ctx *c, const double *lin, const double *mf, const double *kx0)
{
    if (!c) return fail(QL_ERR_ARG, "ql_set_state: null context");
    if (lin)
        QL_HIP(hipMemcpyAsync(c->u, lin, sizeof(double2) * (size_t)c->nxy * 4 * c->nz, hipMemcpyHostToDevice, c->s));
    if (mf) QL_HIP(hipMemcpyAsync(c->mf, mf, sizeof(double2) * (size_t)c->nMF * c->nz, hipMemcpyHostToDevice, c->s));
    if (kx0) QL_HIP(hipMemcpyAsync(c->kx0, kx0, sizeof(double) * (size_t)c->nxy, hipMemcpyHostToDevice, c->s));
    int rc = refresh_profiles(c);
    if (rc) return rc;
    QL_HIP(hipStreamSynchronize(c->s));
    return QL_OK;
}

int ql_get_state(ql_ctx *c, double *lin, double *mf, double *kx0)
{
    if (!c) return fail(QL_ERR_ARG, "ql_get_state: null context");
    if (lin)
        QL_HIP(hipMemcpyAsync(lin, c->u, sizeof(double2) * (size_t)c->nxy * 4 * c->nz, hipMemcpyDeviceToHost, c->s));
    if (mf) QL_HIP(hipMemcpyAsync(mf, c->mf, sizeof(double2) * (size_t)c->nMF * c->nz, hipMemcpyDeviceToHost, c->s));
    if (kx0) QL_HIP(hipMemcpyAsync(kx0, c->kx0, sizeof(double) * (size_t)c->nxy, hipMemcpyDeviceToHost, c->s));
    QL_HIP(hipStreamSynchronize(c->s));
    return QL_OK;
}

int ql_linear_ops_init(ql_ctx *c, double t0, double *linop, double *linop_mf)
{
    if (!c) return fail(QL_ERR_ARG, "ql_linear_ops_init: null context");
    c->old_init = 1;
    c->qt_old = c->p.q * t0;
    c->qdtl_old = 0.0;
    const int nz = c->nz;
    if (linop) {
        // MHD_FullQlin.cpp:432-440 / MHD_BQlin.cpp:289-295 on the current kx0
        std::vector<double> k0(c->nxy);
        QL_HIP(hipMemcpyAsync(k0.data(), c->kx0, sizeof(double) * c->nxy, hipMemcpyDeviceToHost, c->s));
        QL_HIP(hipStreamSynchronize(c->s));
        for (int i = 0; i < c->nxy; ++i) {
            const double ky = c->h_ky[i];
            const double kx = k0[i] + c->qt_old * ky;
            for (int m = 0; m < nz; ++m) {
                const double kz2 = -(c->h_kz[m] * c->h_kz[m]);
                const double lap2 = (ky == 0.0 && m == 0) ? 0.0 : (-ky * ky + kz2);
                const double lapF = -kx * kx + lap2;
                double L0, L2;
                if (c->p.model == QL_MODEL_MHD_BQLIN) {
                    L0 = c->p.nu * lapF;
                    L2 = (c->p.eta / c->p.nu) * L0;
                } else {
                    const double lp = -pow(fabs(lapF), (double)c->vo2);
                    L0 = c->p.nu * lp;
                    L2 = c->p.eta * lp;
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
        for (int f = 0; f < c->nMF; ++f)
            for (int m = 0; m < nz; ++m) {
                const double kz2 = -(c->h_kz[m] * c->h_kz[m]);
                double v = 0.0;
                if (c->p.quasilinear) v = (c->p.model == QL_MODEL_MHD_BQLIN || f < 2 ? c->p.eta : c->p.nu) * kz2;
                linop_mf[(size_t)f * nz + m] = v;
            }
    }
    return QL_OK;
}

int ql_rhs(ql_ctx *c, double t, double dt_lin, double *rhs_lin, double *rhs_mf, double *linop)
{
    if (!c) return fail(QL_ERR_ARG, "ql_rhs: null context");
    int rc = refresh_profiles(c);
    if (rc) return rc;
    qlk::ModeArgs a = mode_args(c);
    a.mode = qlk::MODE_RHS;
    a.qt = c->p.q * t;
    a.qdtl = c->p.q * dt_lin;
    a.n_out = c->N1;
    a.lin_out = c->lin_tmp;
    qlk::MFArgs m = mf_args(c);
    m.mode = qlk::MODE_RHS;
    m.mfn_out = c->mfN1;
    if ((rc = run_sequence(c, a, m))) return rc;
    const size_t nl = (size_t)c->nxy * 4 * c->nz;
    if (rhs_lin) QL_HIP(hipMemcpyAsync(rhs_lin, c->N1, sizeof(double2) * nl, hipMemcpyDeviceToHost, c->s));
    if (rhs_mf) QL_HIP(hipMemcpyAsync(rhs_mf, c->mfN1, sizeof(double2) * c->nMF * c->nz, hipMemcpyDeviceToHost, c->s));
    if (linop) QL_HIP(hipMemcpyAsync(linop, c->lin_tmp, sizeof(double) * nl, hipMemcpyDeviceToHost, c->s));
    QL_HIP(hipStreamSynchronize(c->s));
    return QL_OK;
}

int ql_step(ql_ctx *c, double t, double *dt_out)
{
    if (!c) return fail(QL_ERR_ARG, "ql_step: null context");
    int rc = do_step(c, t, false, t);
    if (rc) return rc;
    QL_HIP(hipStreamSynchronize(c->s));
    if (dt_out) *dt_out = c->p.dt;
    return QL_OK;
}

int ql_steps(ql_ctx *c, int nsteps, double *t_inout)
{
    if (!c || !t_inout) return fail(QL_ERR_ARG, "ql_steps: null argument");
    if (nsteps < 0) return fail(QL_ERR_ARG, "ql_steps: nsteps < 0");
    double t = *t_inout;
    for (int n = 0; n < nsteps; ++n) {
        const double dt = c->p.dt;
        const double t_new = t + dt;  // main.cpp:116
        int rc = do_step(c, t, c->p.remap != 0, t_new);
        if (rc) return rc;
        t = t_new;
    }
    QL_HIP(hipStreamSynchronize(c->s));
    *t_inout = t;
    return QL_OK;
}

int ql_remap(ql_ctx *c, double qt)
{
    if (!c) return fail(QL_ERR_ARG, "ql_remap: null context");
    const int grid = c->nxy < 1024 ? c->nxy : 1024;
    qlk::k_remap<<<grid, qlk::NT, 0, c->s>>>(c->u, c->kx0, c->ky, c->nxy, c->nz, qt, c->remap_thr, c->remap_shift);
    QL_HIP(hipGetLastError());
    QL_HIP(hipStreamSynchronize(c->s));
    return QL_OK;
}

int ql_diagnostics(ql_ctx *c, double t, double out[18])
{
    if (!c || !out) return fail(QL_ERR_ARG, "ql_diagnostics: null argument");
    qlk::DiagArgs d{};
    d.u = c->u;
    d.kx0 = c->kx0;
    d.ky = c->ky;
    d.kzv = c->kzv;
    d.nxy = c->nxy;
    d.nz = c->nz;
    d.qt = c->p.q * t;
    d.nu = c->p.nu;
    d.eta = c->p.eta;
    d.part = c->dpart;
    const int grid = c->nxy < 256 ? c->nxy : 256;
    qlk::k_diag_modes<<<grid, qlk::NT, 0, c->s>>>(d);
    QL_HIP(hipGetLastError());
    qlk::DiagFinalArgs f{};
    f.part = c->dpart;
    f.nparts = grid;
    f.mf = c->mf;
    f.drive = c->drive;
    f.kzv = c->kzv;
    f.nz = c->nz;
    f.nMF = c->nMF;
    f.model = c->p.model;
    f.NZ = c->NZ;
    f.Lz = c->p.L[2];
    f.totalN2 = c->totalN2;
    f.nu = c->p.nu;
    f.eta = c->p.eta;
    f.kzfac = c->kzfac;
    f.out = c->dout;
    qlk::k_diag_final<<<1, 64, 0, c->s>>>(f);
    QL_HIP(hipGetLastError());
    QL_HIP(hipMemcpyAsync(out, c->dout, sizeof(double) * 18, hipMemcpyDeviceToHost, c->s));
    QL_HIP(hipStreamSynchronize(c->s));
    return QL_OK;
}

int ql_get_drive(ql_ctx *c, double *drive)
{
    if (!c || !drive) return fail(QL_ERR_ARG, "ql_get_drive: null argument");
    QL_HIP(hipMemcpyAsync(drive, c->drive, sizeof(double2) * c->nMF * c->nz, hipMemcpyDeviceToHost, c->s));
    QL_HIP(hipStreamSynchronize(c->s));
    return QL_OK;
}

int ql_mf_real(ql_ctx *c, double *out)
{
    if (!c || !out) return fail(QL_ERR_ARG, "ql_mf_real: null argument");
    QL_DISPATCH_NZ(c->NZ, launch_mfreal, c, c->mfr_tmp);
    QL_HIP(hipGetLastError());
    QL_HIP(hipMemcpyAsync(out, c->mfr_tmp, sizeof(double) * c->nMF * c->NZ, hipMemcpyDeviceToHost, c->s));
    QL_HIP(hipStreamSynchronize(c->s));
    return QL_OK;
}

int ql_synchronize(ql_ctx *c)
{
    if (!c) return fail(QL_ERR_ARG, "ql_synchronize: null context");
    QL_HIP(hipStreamSynchronize(c->s));
    return QL_OK;
}

int ql_profile_kernels(ql_ctx *c, int nsteps, double *t_inout, double avg_ms[2], int counts[2])
{
    if (!c || !t_inout || !avg_ms || !counts) return fail(QL_ERR_ARG, "ql_profile_kernels: null argument");
    for (int k = 0; k < 2; ++k) {
        c->t_ms[k] = 0.0;
        c->t_cnt[k] = 0;
    }
    c->timing = true;
    int rc = ql_steps(c, nsteps, t_inout);
    c->timing = false;
    if (rc) return rc;
    for (int k = 0; k < 2; ++k) {
        counts[k] = c->t_cnt[k];
        avg_ms[k] = c->t_cnt[k] ? c->t_ms[k] / c->t_cnt[k] : 0.0;
    }
    return QL_OK;
}

}  // extern "C"
