// ql_noise.hpp -- Model::DrivingNoise on the device (MHD_FullQlin.cpp:786-1053,
// MHD_BQlin.cpp:558-740), fused with the ShearingBox_Remap that follows it in
// the driver loop (main.cpp:111-123: Step, DrivingNoise(t, dt), t += dt,
// remap(q t)).
//
// Each fluctuation entry (mode i, field f, kz slot j) inside the driven |k|
// range gets mult * (sigma z1 + i sigma z2), sigma = f_noise / sqrt(2), with
// mult = sqrt(nf lap2 / lapF) for ux and bx and sqrt(-nf lap2) for zeta and eta,
// nf = dt (NX NY NZ)^2 / (16 32 32)^2, lapF and lap2 at the step start time
// (MHD_FullQlin.cpp:800-845).  ky = 0 modes are driven in Hermitian pairs: the
// kx < 0 mode draws the noise for slots j >= 1 and its kx > 0 partner gets the
// conjugate at slot nz - j (MHD_FullQlin.cpp:962-1046); MHD_BQlin does not
// drive ky = 0 (MHD_BQlin.cpp:18).
//
// The reference draws from a clock-seeded mt19937 stream whose order is
// sequential over modes.  Here z1, z2 come from Philox4x64-10 (Salmon et al.,
// SC 2011) keyed by the run's seed and counted by (t, mode, slot), so every
// entry is independent of launch order, a replayed graph draws fresh noise at
// each new t, and a restart at the same t redraws the same noise; pairs of
// uniforms become normals by the Box-Muller transform.
#pragma once
#include "ql_kernels.hpp"

namespace qlk {

struct NoiseParams {
    double sigma;       // f_noise / sqrt(2) (MHD_FullQlin.cpp:47)
    double lo2, hi2;    // noise_range_low^2, noise_range_high^2 (:51-52)
    double nfac;        // (NX NY NZ)^2 * (1/(16 32 32))^2, times dt per call (:65-66, :804)
    unsigned long long seed;
    int drive_u, drive_b;  // !drive_only_magnetic, !drive_only_velocity (:21-22)
    int drive_ky0;         // !dont_drive_ky0_modes_ (:20; MHD_BQlin.cpp:18)
    int nyh;               // NY/2: ky values per kx (Kdata.cpp:29-46)
    int nxh;               // NX/2
    int pad;
};

// Philox4x64 with 10 rounds (multipliers and Weyl increments of the
// published generator)
struct U64x4 {
    unsigned long long v[4];
};
__device__ __forceinline__ U64x4 philox4x64_10(U64x4 c, unsigned long long k0, unsigned long long k1)
{
    constexpr unsigned long long M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
    constexpr unsigned long long W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const unsigned long long hi0 = __umul64hi(M0, c.v[0]), lo0 = M0 * c.v[0];
        const unsigned long long hi1 = __umul64hi(M1, c.v[2]), lo1 = M1 * c.v[2];
        c = U64x4{{hi1 ^ c.v[1] ^ k0, lo1, hi0 ^ c.v[3] ^ k1, lo0}};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// stream constant of the second key word ("QLNOISE\0")
constexpr unsigned long long NOISE_KEY1 = 0x514C4E4F49534500ull;

// two standard normals from two 64-bit words (Box-Muller; u1 in (0, 1))
__device__ __forceinline__ double2 box_muller(unsigned long long a, unsigned long long b)
{
    const double u1 = ((double)(a >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = (double)(b >> 11) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1));
    const double th = rn_mul(2.0 * 3.14159265358979323846, u2);
    return make_double2(rn_mul(r, cos(th)), rn_mul(r, sin(th)));
}

// the eight normals (4 fields x re, im) of entry (t, mode, slot)
__device__ __forceinline__ void noise_normals(unsigned long long seed, double t, int mode, int slot, double2 z[4])
{
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const U64x4 c{{(unsigned long long)__double_as_longlong(t), (unsigned long long)mode,
                       (unsigned long long)slot | ((unsigned long long)k << 32), 0ull}};
        const U64x4 w = philox4x64_10(c, seed, NOISE_KEY1);
        z[2 * k] = box_muller(w.v[0], w.v[1]);
        z[2 * k + 1] = box_muller(w.v[2], w.v[3]);
    }
}

struct NoiseArgs {
    double2 *u;             // state [modes][4][nz]
    double *kx0;            // [modes] (the remap updates it)
    const double *ky;       // [modes]
    const double *kzv;      // [nz]
    const RunParams *runs;
    const NoiseParams *np;  // [runs]
    const int *mode_run;    // [modes] run of each group mode
    const Clock *clock;     // [runs] step times after the step (t_prev = t, dt_last = dt)
    int nz;
    int use_clock;          // times from the clock, else t_x, dt_x
    int noise;              // apply the noise (else remap only)
    int do_remap;           // remap at q (t + dt) afterwards
    double t_x, dt_x;
};

// one workgroup per mode of the group, one thread per kz slot
[[maybe_unused]] static __global__ __launch_bounds__(NT) void k_noise_remap(NoiseArgs a)
{
    const int m = blockIdx.x, j = threadIdx.x;
    const int run = a.mode_run[m];
    const RunParams &rp = a.runs[run];
    const NoiseParams &np = a.np[run];
    const int nz = a.nz;
    const double t = a.use_clock ? a.clock[run].t_prev : a.t_x;
    const double dt = a.use_clock ? a.clock[run].dt_last : a.dt_x;
    const double kx0 = a.kx0[m], ky = a.ky[m];
    const int il = m - rp.mode_off;  // the mode's index in its run
    double2 *U = a.u + (size_t)m * 4 * nz;
    if (a.noise && np.sigma != 0.0 && j < nz) {
        const bool ky0 = (ky == 0.0);
        const int ix = il / np.nyh;      // kx index position (Kdata.cpp:29-46)
        const bool gen = !ky0 || (np.drive_ky0 && ix >= np.nxh && j >= 1);
        if (gen) {
            // assign_laplacians_(i, t, 0) (MHD_FullQlin.cpp:722-738)
            const double kz = a.kzv[j];
            const double kx = rn_add(kx0, rn_mul(rn_mul(rp.ph.q, t), ky));
            const double lap2 = rn_add(rn_mul(-ky, ky), -(kz * kz));
            const double lapF = rn_add(rn_mul(-kx, kx), lap2);
            if (lapF > -np.hi2 && lapF < -np.lo2) {
                const double nf = rn_mul(dt, np.nfac);
                const double mu = sqrt(rn_mul(nf, lap2) / lapF);
                const double mz = sqrt(rn_mul(-nf, lap2));
                double2 z[4];
                noise_normals(np.seed, t, il, j, z);
                const double mult[4] = {mu, mz, mu, mz};
                const bool on[4] = {np.drive_u != 0, np.drive_u != 0, np.drive_b != 0, np.drive_b != 0};
                // the ky = 0, kx > 0 partner of this kx < 0 mode (Kdata.cpp:122-146)
                double2 *P = ky0 ? a.u + (size_t)(m + np.nyh * (2 * np.nxh - 1) - 2 * il) * 4 * nz : nullptr;
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    if (!on[f]) continue;
                    const double2 n = make_double2(rn_mul(mult[f], rn_mul(np.sigma, z[f].x)),
                                                   rn_mul(mult[f], rn_mul(np.sigma, z[f].y)));
                    double2 &e = U[f * nz + j];
                    e = make_double2(rn_add(e.x, n.x), rn_add(e.y, n.y));
                    if (ky0) {
                        double2 &ep = P[f * nz + (nz - j)];
                        ep = make_double2(rn_add(ep.x, n.x), rn_add(ep.y, -n.y));
                    }
                }
            }
        }
    }
    if (a.do_remap && rp.remap) {
        // ShearingBox_Remap(q (t + dt)) (MHD_FullQlin.cpp:464-481); never a
        // ky = 0 mode, so no partner write above touches a remapped mode
        const double tn = a.use_clock ? a.clock[run].t : rn_add(t, dt);
        const bool hit = remap_hit(rn_mul(rp.ph.q, tn), rp.remap_thr, kx0, ky);
        __syncthreads();  // every thread has read kx0
        if (hit) {
            for (int k = j; k < 4 * nz; k += blockDim.x) U[k] = make_double2(0.0, 0.0);
            if (j == 0) a.kx0[m] = rn_sub(kx0, rp.remap_shift);
        }
    }
}

}  // namespace qlk
