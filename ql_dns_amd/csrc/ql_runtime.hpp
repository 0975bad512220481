// ql_runtime.hpp -- host-side structures of the QL library shared by its
// translation units (ql_dns.hip: run sets and the forward step; ql_adjoint.hip:
// the adjoint).  Not part of the public interface (include/ql_dns.h).
//
// Every handle holds a set of runs (a context is a set of one).  Runs are
// grouped by NZ; a group's runs share one kernel instantiation and are laid
// out one after another in the group's arrays, each with its own parameters,
// mean fields and reductions.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ql_dns.h"
#include "ql_noise.hpp"

namespace qlh {

// message of the last failure on this thread (ql_last_error)
inline thread_local std::string g_err;
int fail(int code, const char *fmt, ...);

#define QL_HIP(call)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) return qlh::fail(QL_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define QL_RC(call)                   \
    do {                              \
        int rc_ = (call);             \
        if (rc_ != QL_OK) return rc_; \
    } while (0)

// integrator coefficients, shared with the kernels
using qlk::A0, qlk::A1, qlk::A2, qlk::B1, qlk::B2, qlk::P1, qlk::P2;
// General_Definitions.h:236: PI = atan(1)*4, which rounds to this double
constexpr double QL_PI = 3.14159265358979323846;

// steps captured in the long replay graph
constexpr int GRAPH_STEPS = 16;
// step counts of the replay graphs: any step count is at most 4 launches past
// its multiples of GRAPH_STEPS
constexpr int GRAPH_LEVELS[5] = {GRAPH_STEPS, 8, 4, 2, 1};
// launch-shape candidates of the tuner: z points per workgroup pass
// (modes per workgroup MB = line / NZ) and the target workgroup count; line
// LINE_WAVE selects the kernel with one wavefront per mode (MB = 1,
// ql_mode_wave.hpp)
constexpr int LINE_WAVE = 64;
constexpr int LINES[] = {LINE_WAVE, 128, 256};
constexpr int TARGETS[] = {256, 512, 1024, 2048, 4096, 1 << 30};
constexpr int DEFAULT_LINE = 256, DEFAULT_TARGET = 512;

inline bool valid_NZ(int NZ) { return NZ == 32 || NZ == 64 || NZ == 128 || NZ == 256; }
inline int mb_for(int NZ, int line) { return (line == LINE_WAVE || NZ >= line) ? 1 : line / NZ; }
// shape before tuning: the wavefront kernel, which measured fastest at NZ =
// 128 and 256 for every model and precision (in mixed precision at NZ = 256 it
// runs the flagged modes in FP64 itself, which beats the workgroup kernel plus
// its companion: 2672 against 2625 steps/s at 128x128x256), with about one
// workgroup per CU
inline int default_line(int NZ, int, bool) { return NZ >= 128 ? LINE_WAVE : DEFAULT_LINE; }
inline int default_target(int line) { return line == LINE_WAVE ? 256 : DEFAULT_TARGET; }
// the matrix-core transform (QL_TRANSFORM_DFT_MFMA): FP64 workgroup kernel, NZ 32 and 64
inline bool dft_ok(int NZ, int precision, int line)
{
    return precision == QL_PREC_FP64 && (NZ == 32 || NZ == 64) && line != LINE_WAVE;
}

// one run: parameters, sizes and its place in its group
struct RunInfo {
    ql_params p{};  // as the kernels see it (HD_fullU: MHD_FullQlin with b = B = 0)
    bool hd = false;
    int group = 0, gi = 0, mode_off = 0;
    int nxy = 0, nz = 0, nzl = 0, half = 0;
    double omega_eff = 0, rsf = 0, remap_thr = 0, remap_shift = 0, totalN2 = 0, kzfac = 0, kmax = 0;
    std::vector<double> h_kx0, h_ky, h_kz;
};

// runs of one NZ
struct Group {
    int NZ = 0, nz = 0, nruns = 0, modes = 0;
    std::vector<int> run_ids;
    int line = DEFAULT_LINE, target = DEFAULT_TARGET;
    int dft = 0;  // 1: the workgroup kernel's transforms as matrix-core products (ql_dft.hpp)
    int MB = 1, W = 0, part_cap = 0, wg_cap = 0;
    hipStream_t s = nullptr;  // own stream (sets with several groups)
    hipEvent_t join = nullptr;
    double2 *u = nullptr, *N1 = nullptr, *N2 = nullptr;
    double2 *mf = nullptr, *mfN1 = nullptr, *mfN2 = nullptr, *drive = nullptr, *tw = nullptr;
    // the substep tail as one launch (k_mf_fused) where it applies; it writes
    // the mean fields of substeps 1 and 2 to mfc1, mfc2 and those of substep 3
    // to mf, so mf holds them at the end of every step
    int tail = 0;
    double2 *mfc1 = nullptr, *mfc2 = nullptr;
    float2 *tw32 = nullptr;
    double *kx0 = nullptr, *ky = nullptr, *kzv = nullptr, *mfz = nullptr, *part = nullptr, *sums = nullptr;
    double *lin_tmp = nullptr;
    double *pm = nullptr;  // profile maxima of the CFL condition [runs][PM_STRIDE]
    int *mode_run = nullptr;        // [modes] run (within the group) of each mode
    qlk::NoiseParams *np = nullptr;  // [runs] driving noise
    std::vector<qlk::NoiseParams> h_np;
    bool noise = false;              // some run of the group is driven
    qlk::RunParams *rp = nullptr;
    qlk::WgInfo *wg = nullptr;
    qlk::Clock *clock = nullptr;  // [runs] each run's step times
    std::vector<qlk::Clock> h_clock;  // host staging of per-run clocks
    std::vector<qlk::RunParams> h_rp;
    // mixed precision: the FP64 companion launch of the modes that carry the
    // norm (ModeArgs::split)
    bool split = false;
    int W2 = 0;                      // companion workgroups (rows after the W)
    qlk::WgInfo *wg2 = nullptr;      // companion table
    float *amp = nullptr;            // [modes] mode amplitudes when the flags were last set
    unsigned char *hi = nullptr;     // [modes] the modes the companion takes
    bool amp_stale = true;           // a state was uploaded (or the fraction changed) since the flags were set
};

// mixed precision: a mode runs in FP64 when its amplitude reaches this
// fraction of its run's largest (measured at 128x128x256 over 2000 steps: 29
// of 8128 modes in FP64 and the mean fields, drives and energies within 3e-9
// of the FP64 run, against 1.7e-6 with every mode in FP32); the flags are
// set again every SPLIT_EVERY steps of the clock (a mode's amplitude changes
// by a few per cent over that many steps) and after a state upload.  Remaps
// and driving noise leave them: a zeroed mode gives zeros in either
// precision, and the noise is small against the modes that set the flags
constexpr double MIXED_AMP_TAU = 0.03;
constexpr int SPLIT_EVERY = 16;
// companion launch: about this many workgroups per run
constexpr int COMP_TARGET_WG = 128;
// device time the tuner runs the current shapes before timing candidates
constexpr double TUNE_WARM_MS = 25.0;

struct AdjState;  // adjoint checkpoints and buffers (ql_adjoint.hip)

struct Batch {
    int model = 0, integrator = 0, precision = 0, device = 0, nMF = 4;
    double amp_tau = MIXED_AMP_TAU;  // mixed precision: FP64 from this fraction of a run's largest mode
    int nstep = 0;                   // steps since the integrator was set up (the device clock's count)
    // HD_fullU runs on the MHD_FullQlin kernels with the magnetic fields zero;
    // at the boundary they have 2 fluctuation fields (u, zeta) and 2 mean
    // fields (Ux, Uy, internal slots 2 and 3)
    bool hd = false;
    double dt = 0;   // the fixed step, or the largest variable one
    int vdt = 0;     // variable steps (RK3CN.cpp:20-30): dt and coefficients on the device
    double dt_last = 0;  // the step taken last
    std::vector<RunInfo> runs;
    std::vector<Group> groups;
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr;
    // host view of the clock: start of the previous step, linearOPs_Init time
    double t_prev = 0.0, t_init = 0.0;
    // per run, for sets of several runs with variable steps (each its own clock)
    std::vector<double> rt_prev, rdt_last;
    int old_init = 1;
    // diagnostics scratch
    double *dpart = nullptr, *dout = nullptr, *mfr_tmp = nullptr;
    // replay graphs of GRAPH_LEVELS[k] steps (all captured by the first stepping call)
    hipGraphExec_t graphs[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // per-kernel timing (ql_profile_kernels)
    bool timing = false;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    double t_ms[2] = {0, 0};
    int t_cnt[2] = {0, 0};
    // timing pass: three events per substep sequence (mode start, mode end,
    // after the mean fields), read once the pass has finished, so that the
    // pass runs without host round trips between launches
    std::vector<hipEvent_t> tpool;
    size_t tused = 0;
    AdjState *adj = nullptr;
};

template <typename T>
int dalloc(T **p, size_t n)
{
    QL_HIP(hipMalloc((void **)p, sizeof(T) * (n ? n : 1)));
    QL_HIP(hipMemset(*p, 0, sizeof(T) * (n ? n : 1)));
    return QL_OK;
}

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// shared by the translation units (defined in ql_dns.hip)
int sync_all(Batch &b);
int set_clock(Batch &b, double t);
int set_clock_runs(Batch &b, const double *t_runs);
int read_clock(Batch &b, qlk::Clock *c);
int enqueue_substep(Batch &b, Group &g, int s, bool remap_after, bool advance, hipStream_t st);
bool fused_tail(const Batch &b, const Group &g);
const double2 *mf_after_substep(const Batch &b, const Group &g, int s);
void launch_mf(const Group &g, const qlk::MFArgs &m, hipStream_t st);
int launch_noise(Group &g, bool use_clock, double t, double dt, bool noise, bool remap, hipStream_t st);
qlk::MFArgs mf_args(const Batch &b, const Group &g);
int refresh_profiles(Batch &b, Group &g, hipStream_t st, bool rec = false);
int run_steps(Batch &b, int nsteps, double *t_inout);
// defined in ql_adjoint.hip
void adj_release(Batch &b);

}  // namespace qlh

struct ql_batch {
    qlh::Batch b;
};
struct ql_ctx {
    ql_batch bt;
};
