// ql_kernels.hpp -- the per-substep kernels of the QL time step (FP64).
//
// One RK3 substep is two launches on one stream:
//   k_mode_*      fused per-(kx,ky) RHS + substep update.  A workgroup takes MB
//                 modes at a time: builds the padded spectral lines of u, b and
//                 their z-derivatives, inverse-transforms them in LDS, forms the
//                 mean-field x fluctuation products on the z grid, forward-
//                 transforms them, assembles the RHS, applies the exp / CN
//                 linear factor computed on the fly from kx(t), and writes the
//                 new state (and the RHS when a later substep needs it).  The
//                 EMF / Reynolds-stress z-profiles of its modes are summed in
//                 registers in a fixed order and written as one partial per WG.
//   k_meanfield   fixed-order reduction of the partials (row blocks, then the
//                 last block to finish adds the block sums), forward transform, x i kz,
//                 Hermitian fill (the reference's MPI Allreduce + fill,
//                 MHD_FullQlin.cpp:386-399), mean-field RHS and update, and the
//                 real-space mean-field profiles the next substep reads.
#pragma once
#include "ql_fft.hpp"

namespace qlk {

enum { MODE_RHS = 0, MODE_ELIN = 1, MODE_CN = 2, MODE_PROFILES = 3 };

// Occupancy request for the per-mode kernels (waves per SIMD); bounds their
// VGPR budget.  0 leaves the choice to the compiler.
#ifndef QL_MODE_WAVES
#define QL_MODE_WAVES 0
#endif
#if QL_MODE_WAVES > 0
#define QL_MODE_ATTR __attribute__((amdgpu_waves_per_eu(QL_MODE_WAVES)))
#else
#define QL_MODE_ATTR
#endif

struct ModeArgs {
    double2 *u;             // state [nxy][4][nz], read and (MODE_ELIN/CN) updated in place
    double2 *n_out;         // RHS out (nullable)
    const double2 *n_prev;  // previous substep RHS (nullable)
    double *lin_out;        // linear operator out (MODE_RHS, nullable)
    double *kx0;            // [nxy] (remap updates it)
    const double *ky;       // [nxy]
    const double *kzv;      // [nz] physical kz
    const double2 *tw;      // [NZ] exp(-2 pi i k/NZ)
    const double *mfz;      // real-space mean-field profiles [.][NZ]
    double *partials;       // [gridDim.x][nMF][NZ]
    int nxy, vo2, mode, old_init, do_remap;
    double qt, qdtl;        // q * t_s and q * dt_lin
    double q, omega, B0z, nu, eta, fac, rsf;
    double linC, nlc1, nlc2;
    double qt_old, qdtl_old;  // where the CN "old" operator was evaluated
    double remap_qt, remap_thr, remap_shift;
};

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
__device__ __forceinline__ bool remap_hit(const ModeArgs &a, double kx0, double ky)
{
    return __dadd_rn(kx0, __dmul_rn(a.remap_qt, ky)) > a.remap_thr;
}

// Per-mode inverse Laplacians of assign_laplacians_ (MHD_FullQlin.cpp:722-738,
// Kdata.cpp:100-120): ilap2 = 1/(-ky^2 - kz^2) with the (ky, kz) = (0, 0)
// entry set to 1, ilapF = 1/(-kx(t)^2 + lap2) with the (0, 0, 0) entry set to 1.
template <int NZD, int MB>
__device__ __forceinline__ void mode_tables(const ModeArgs &a, const double *skx, const double *sky,
                                            const int *sidx, double *sil2, double *silF)
{
    for (int k = threadIdx.x; k < MB * NZD; k += NT) {
        const int mb = k / NZD, m = k % NZD;
        const double kx = skx[mb], ky = sky[mb], kz = a.kzv[m];
        const bool zero = (ky == 0.0 && m == 0);
        const double lap2 = zero ? 0.0 : (-ky * ky + -(kz * kz));
        sil2[k] = zero ? 1.0 : 1.0 / lap2;
        silF[k] = (sidx[mb] == 0 && m == 0) ? 1.0 : 1.0 / (-kx * kx + lap2);
    }
}

// Substep update of one (mode, kz) slot for a pair of fields sharing one
// dissipation coefficient: RK3CN_Elin (RK3CN_Elin.cpp:303-305 and :335-337,
// :368-370) or RK3CN (RK3CN.cpp:105-107, :141-143, :177-179).  Ln is the
// operator at the end of the substep; Lo is only used by RK3CN.
__device__ __forceinline__ void update_pair(const ModeArgs &a, size_t off0, size_t off1, double2 u0, double2 u1,
                                            double2 N0, double2 N1, double Ln, double Lo, bool zero_out)
{
    double2 un0, un1;
    if (a.mode == MODE_ELIN) {
        const double e = exp(a.linC * Ln);
        double2 s0 = cscale(a.nlc1, N0), s1 = cscale(a.nlc1, N1);
        if (a.n_prev) {
            s0 = cadd(s0, cscale(a.nlc2, a.n_prev[off0]));
            s1 = cadd(s1, cscale(a.nlc2, a.n_prev[off1]));
        }
        un0 = cscale(e, cadd(s0, u0));
        un1 = cscale(e, cadd(s1, u1));
    } else {
        const double denom = 1.0 / (1.0 - a.linC * Ln);
        const double c1 = a.nlc1 * denom, co = (1 + a.linC * Lo) * denom;
        double2 s0 = cscale(c1, N0), s1 = cscale(c1, N1);
        if (a.n_prev) {
            const double c2 = a.nlc2 * denom;
            s0 = cadd(s0, cscale(c2, a.n_prev[off0]));
            s1 = cadd(s1, cscale(c2, a.n_prev[off1]));
        }
        un0 = cadd(s0, cscale(co, u0));
        un1 = cadd(s1, cscale(co, u1));
    }
    if (zero_out) un0 = un1 = make_double2(0.0, 0.0);
    a.u[off0] = un0;
    a.u[off1] = un1;
    if (a.n_out) {
        a.n_out[off0] = N0;
        a.n_out[off1] = N1;
    }
}

// ---------------------------------------------------------------------------
// MHD_FullQlin (Models/MHD_FullQlin.cpp:171-427)
// ---------------------------------------------------------------------------
template <int NZ, int MB>
__global__ __launch_bounds__(NT) QL_MODE_ATTR void k_mode_full(ModeArgs a)
{
    using D = Dealias<NZ>;
    constexpr int NZD = D::nz, NZL = D::nzl;
    constexpr int NL = 10 * MB;
    constexpr int PI_ = MB * 2 * NZ;
    constexpr int PW = (PI_ + NT - 1) / NT;
    __shared__ double2 buf[NL * NZ];
    __shared__ double2 uin[MB * 4 * NZD];
    __shared__ double2 tw[NZ];
    __shared__ double skx0[MB], sky[MB], skx[MB];
    __shared__ int sidx[MB];
    __shared__ double sil2[MB * NZD], silF[MB * NZD];
    const int tid = threadIdx.x;

    for (int k = tid; k < NZ; k += NT) tw[k] = a.tw[k];

    // each thread owns fixed (mb, part, z) pointwise items for the whole kernel
    double mfr[PW][8];
    double acc[PW][2];
#pragma unroll
    for (int p = 0; p < PW; ++p) {
        const int item = tid + p * NT;
        const int z = item % NZ;
        acc[p][0] = acc[p][1] = 0.0;
#pragma unroll
        for (int c = 0; c < 8; ++c) mfr[p][c] = (item < PI_) ? a.mfz[c * NZ + z] : 0.0;
    }

    const int ngroups = (a.nxy + MB - 1) / MB;
    for (int g = blockIdx.x; g < ngroups; g += gridDim.x) {
        if (tid < MB) {
            const int i = g * MB + tid;
            sidx[tid] = i;
            if (i < a.nxy) {
                const double kx0 = a.kx0[i], ky = a.ky[i];
                skx0[tid] = kx0;
                sky[tid] = ky;
                skx[tid] = __dadd_rn(kx0, __dmul_rn(a.qt, ky));  // kx(t), MHD_FullQlin.cpp:727
            } else {
                skx0[tid] = sky[tid] = skx[tid] = 0.0;
            }
        }
        for (int k = tid; k < MB * 4 * NZD; k += NT) {
            const int mb = k / (4 * NZD);
            const int i = g * MB + mb;
            uin[k] = (i < a.nxy) ? a.u[(size_t)i * 4 * NZD + (k - mb * 4 * NZD)] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        mode_tables<NZD, MB>(a, skx, sky, sidx, sil2, silF);
        __syncthreads();

        // padded spectral lines, fac1D folded in (MHD_FullQlin.cpp:220-265, :338-356):
        // 0 ux 1 uy 2 uz 3 bx 4 by 5 bz 6 dz ux 7 dz bx 8 dz uy 9 dz by
        for (int idx = tid; idx < NL * NZ; idx += NT) {
            const int line = idx / NZ, j = idx % NZ;
            const int mb = line / 10, l = line % 10;
            double2 val = make_double2(0.0, 0.0);
            if (D::kept(j)) {
                const int m = D::slot(j);
                const double2 *U = uin + mb * 4 * NZD;
                const double kx = skx[mb], ky = sky[mb], kz = a.kzv[m];
                const double ilap2 = sil2[mb * NZD + m];
                const bool isb = (l == 3 || l == 4 || l == 5 || l == 7 || l == 9);
                const double2 f1 = isb ? U[2 * NZD + m] : U[m];        // ux / bx
                const double2 f2 = isb ? U[3 * NZD + m] : U[NZD + m];  // zeta / eta
                // y and z components: ((ky kx) f1 + i kz f2) ilap2, ((kx kz) f1 - i ky f2) ilap2
                const double sy = ky * kx, sz = kx * kz;
                double2 fy = make_double2((sy * f1.x - kz * f2.y) * ilap2, (sy * f1.y + kz * f2.x) * ilap2);
                double2 fzv = make_double2((sz * f1.x + ky * f2.y) * ilap2, (sz * f1.y - ky * f2.x) * ilap2);
                const double fac = a.fac;
                switch (l) {
                case 0: case 3: val = cscale(fac, f1); break;
                case 1: case 4: val = cscale(fac, fy); break;
                case 2: case 5: val = cscale(fac, fzv); break;
                case 6: case 7: val = imul(fac * kz, f1); break;
                default: val = imul(fac * kz, fy); break;  // 8, 9
                }
            }
            buf[idx] = val;
        }
        __syncthreads();
        fft_lines<NZ, NL, true>(buf, tw);

        // products on the z grid (MHD_FullQlin.cpp:271-300) and EMF / stresses
        // (:322-358); x and y derivatives are i kx, i ky times the z profile.
        double2 prod[PW][3];
#pragma unroll
        for (int p = 0; p < PW; ++p) {
            const int item = tid + p * NT;
            if (PI_ % NT != 0 && item >= PI_) continue;
            const int mb = item / (2 * NZ), rem = item % (2 * NZ);
            const int part = rem / NZ, z = rem % NZ;
            const double2 *X = buf + mb * 10 * NZ + z;
            const double2 u0 = X[0], u1 = X[NZ], u2 = X[2 * NZ];
            const double2 b0 = X[3 * NZ], b1 = X[4 * NZ], b2 = X[5 * NZ];
            const double MBx = mfr[p][0], dzMBx = mfr[p][1], MBy = mfr[p][2], dzMBy = mfr[p][3];
            const double MUx = mfr[p][4], dzMUx = mfr[p][5], MUy = mfr[p][6], dzMUy = mfr[p][7];
            const double kx = skx[mb], ky = sky[mb];
            const double aU = kx * MUx + ky * MUy, aB = kx * MBx + ky * MBy;
            const bool valid = sidx[mb] < a.nxy;
            const double ftfac = (ky == 0.0 ? 1.0 : 2.0) * a.rsf;
            if (part == 0) {
                // u.grad terms: N = U.grad(u) + uz dzU - (B.grad(b) + bz dzB)
                prod[p][0] = csub(cadd(imul(aU, u0), cscale(dzMUx, u2)), cadd(imul(aB, b0), cscale(dzMBx, b2)));
                prod[p][1] = csub(cadd(imul(aU, u1), cscale(dzMUy, u2)), cadd(imul(aB, b1), cscale(dzMBy, b2)));
                prod[p][2] = csub(imul(aU, u2), imul(aB, b2));
                if (valid) {
                    acc[p][0] += ftfac * (redot(b2, u0) - redot(u2, b0));
                    acc[p][1] += ftfac * (redot(b2, u1) - redot(u2, b1));
                }
            } else {
                // induction: N = U.grad(b) + uz dzB - (B.grad(u) + bz dzU)
                prod[p][0] = csub(cadd(imul(aU, b0), cscale(dzMBx, u2)), cadd(imul(aB, u0), cscale(dzMUx, b2)));
                prod[p][1] = csub(cadd(imul(aU, b1), cscale(dzMBy, u2)), cadd(imul(aB, u1), cscale(dzMUy, b2)));
                prod[p][2] = csub(imul(aU, b2), imul(aB, u2));
                if (valid) {
                    const double2 dzux = X[6 * NZ], dzbx = X[7 * NZ], dzuy = X[8 * NZ], dzby = X[9 * NZ];
                    const double2 u3 = imul(kx, u0), u6 = imul(ky, u0), b3 = imul(kx, b0), b6 = imul(ky, b0);
                    const double2 u4 = imul(kx, u1), u7 = imul(ky, u1), b4 = imul(kx, b1), b7 = imul(ky, b1);
                    const double rx = -(redot(u0, u3) + redot(u1, u6) + redot(u2, dzux)) +
                                      (redot(b0, b3) + redot(b1, b6) + redot(b2, dzbx));
                    const double ry = -(redot(u0, u4) + redot(u1, u7) + redot(u2, dzuy)) +
                                      (redot(b0, b4) + redot(b1, b7) + redot(b2, dzby));
                    acc[p][0] += ftfac * rx;
                    acc[p][1] += ftfac * ry;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PW; ++p) {
            const int item = tid + p * NT;
            if (PI_ % NT != 0 && item >= PI_) continue;
            const int mb = item / (2 * NZ), rem = item % (2 * NZ);
            const int part = rem / NZ, z = rem % NZ;
            double2 *Y = buf + (mb * 6 + part * 3) * NZ + z;
            Y[0] = prod[p][0];
            Y[NZ] = prod[p][1];
            Y[2 * NZ] = prod[p][2];
        }
        __syncthreads();
        fft_lines<NZ, 6 * MB, false>(buf, tw);

        // RHS assembly (MHD_FullQlin.cpp:283-305), linear operator (:371-376)
        // and the substep update, per (mode, kz) slot and field pair
        for (int k = tid; k < MB * 2 * NZD; k += NT) {
            const int mb = k / (2 * NZD), r2 = k % (2 * NZD);
            const int pr = r2 / NZD, m = r2 % NZD;
            const int i = sidx[mb];
            if (i >= a.nxy) continue;
            const double2 *T = buf + mb * 6 * NZ + D::pad(m);
            const double2 *U = uin + mb * 4 * NZD;
            const double kx = skx[mb], ky = sky[mb], kz = a.kzv[m];
            const double kz2 = -(kz * kz);
            const double q = a.q, Om = a.omega, B0z = a.B0z;
            const double2 ux = U[m], ze = U[NZD + m], bx = U[2 * NZD + m], et = U[3 * NZD + m];
            double2 N0, N1;
            if (pr == 0) {
                const double ilapF = silF[mb * NZD + m];
                const double2 T1 = T[0], T2 = T[NZ], T3 = T[2 * NZ];
                double2 s = cscale(2.0 * kx * ky * q, ux);
                s = cadd(s, imul(2.0 * Om * kz, ze));
                s = csub(s, cscale(kx * kx, T1));
                s = csub(s, cscale(kx * ky, T2));
                s = csub(s, cscale(kx * kz, T3));
                N0 = csub(cadd(cscale(ilapF, s), imul(B0z * kz, bx)), T1);
                N1 = cadd(imul((q - 2.0 * Om) * kz, ux), imul(B0z * kz, et));
                N1 = csub(N1, csub(imul(kz, T2), imul(ky, T3)));
            } else {
                const double2 T2 = T[4 * NZ], T3 = T[5 * NZ];
                N0 = csub(imul(B0z * kz, ux), T[3 * NZ]);
                N1 = cadd(imul(-q * kz, bx), imul(B0z * kz, ze));
                N1 = csub(N1, csub(imul(kz, T2), imul(ky, T3)));
            }
            const double coef = (pr == 0) ? a.nu : a.eta;
            const double kxn = kx + a.qdtl * ky;
            const double Ln = coef * neg_abs_pow((-kxn * kxn - ky * ky) + kz2, a.vo2);
            const size_t off0 = (size_t)i * 4 * NZD + (2 * pr) * NZD + m, off1 = off0 + NZD;
            if (a.mode == MODE_RHS) {
                if (a.n_out) {
                    a.n_out[off0] = N0;
                    a.n_out[off1] = N1;
                }
                if (a.lin_out) a.lin_out[off0] = a.lin_out[off1] = Ln;
                continue;
            }
            double Lo = 0.0;
            if (a.mode == MODE_CN) {
                if (a.old_init) {
                    const double kxo = __dadd_rn(skx0[mb], __dmul_rn(a.qt_old, ky));
                    const double lap2 = (ky == 0.0 && m == 0) ? 0.0 : (-ky * ky + kz2);
                    Lo = coef * neg_abs_pow(-kxo * kxo + lap2, a.vo2);
                } else {
                    const double kxo = __dadd_rn(skx0[mb], __dmul_rn(a.qt_old, ky)) + a.qdtl_old * ky;
                    Lo = coef * neg_abs_pow((-kxo * kxo - ky * ky) + kz2, a.vo2);
                }
            }
            const bool zero_out = a.do_remap && remap_hit(a, skx0[mb], ky);
            update_pair(a, off0, off1, U[2 * pr * NZD + m], U[(2 * pr + 1) * NZD + m], N0, N1, Ln, Lo, zero_out);
        }
        if (a.do_remap && tid < MB) {
            const int i = sidx[tid];
            if (i < a.nxy && remap_hit(a, skx0[tid], sky[tid])) a.kx0[i] = __dsub_rn(skx0[tid], a.remap_shift);
        }
        __syncthreads();
    }

    // fixed-order reduction of the per-thread EMF/stress sums over the MB mode
    // slots, one partial [4][NZ] per workgroup
    double *red = reinterpret_cast<double *>(buf);
#pragma unroll
    for (int p = 0; p < PW; ++p) {
        const int item = tid + p * NT;
        if (PI_ % NT != 0 && item >= PI_) continue;
        const int mb = item / (2 * NZ), rem = item % (2 * NZ);
        const int part = rem / NZ, z = rem % NZ;
        red[(mb * 4 + part * 2 + 0) * NZ + z] = acc[p][0];
        red[(mb * 4 + part * 2 + 1) * NZ + z] = acc[p][1];
    }
    __syncthreads();
    for (int k = tid; k < 4 * NZ; k += NT) {
        double s = 0.0;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) s += red[mb * 4 * NZ + k];
        a.partials[(size_t)blockIdx.x * 4 * NZ + k] = s;
    }
}

// ---------------------------------------------------------------------------
// MHD_BQlin (Models/MHD_BQlin.cpp:117-284): fluctuations see By(z) only,
// Omega = 2q/3, plain Laplacian dissipation.
// ---------------------------------------------------------------------------
template <int NZ, int MB>
__global__ __launch_bounds__(NT) QL_MODE_ATTR void k_mode_bq(ModeArgs a)
{
    using D = Dealias<NZ>;
    constexpr int NZD = D::nz;
    constexpr int NL = 9 * MB;
    constexpr int PI_ = MB * 2 * NZ;
    constexpr int PW = (PI_ + NT - 1) / NT;
    __shared__ double2 buf[NL * NZ];
    __shared__ double2 uin[MB * 4 * NZD];
    __shared__ double2 tw[NZ];
    __shared__ double skx0[MB], sky[MB], skx[MB];
    __shared__ int sidx[MB];
    __shared__ double sil2[MB * NZD], silF[MB * NZD];
    const int tid = threadIdx.x;
    for (int k = tid; k < NZ; k += NT) tw[k] = a.tw[k];

    double mfr[PW][3];
    double acc[PW];
#pragma unroll
    for (int p = 0; p < PW; ++p) {
        const int item = tid + p * NT;
        const int z = item % NZ;
        acc[p] = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) mfr[p][c] = (item < PI_) ? a.mfz[c * NZ + z] : 0.0;
    }

    const int ngroups = (a.nxy + MB - 1) / MB;
    for (int g = blockIdx.x; g < ngroups; g += gridDim.x) {
        if (tid < MB) {
            const int i = g * MB + tid;
            sidx[tid] = i;
            if (i < a.nxy) {
                const double kx0 = a.kx0[i], ky = a.ky[i];
                skx0[tid] = kx0;
                sky[tid] = ky;
                skx[tid] = __dadd_rn(kx0, __dmul_rn(a.qt, ky));
            } else {
                skx0[tid] = sky[tid] = skx[tid] = 0.0;
            }
        }
        for (int k = tid; k < MB * 4 * NZD; k += NT) {
            const int mb = k / (4 * NZD);
            const int i = g * MB + mb;
            uin[k] = (i < a.nxy) ? a.u[(size_t)i * 4 * NZD + (k - mb * 4 * NZD)] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        mode_tables<NZD, MB>(a, skx, sky, sidx, sil2, silF);
        __syncthreads();
        // lines (fac1D folded in): 0 ux 1 zeta 2 bx 3 eta 4 uy 5 uz 6 by 7 bz 8 dz uz
        for (int idx = tid; idx < NL * NZ; idx += NT) {
            const int line = idx / NZ, j = idx % NZ;
            const int mb = line / 9, l = line % 9;
            double2 val = make_double2(0.0, 0.0);
            if (D::kept(j)) {
                const int m = D::slot(j);
                const double2 *U = uin + mb * 4 * NZD;
                const double kx = skx[mb], ky = sky[mb], kz = a.kzv[m];
                const double ilap2 = sil2[mb * NZD + m];
                const double fac = a.fac;
                if (l < 4) {
                    val = cscale(fac, U[l * NZD + m]);
                } else {
                    const bool isb = (l == 6 || l == 7);
                    const double2 f1 = isb ? U[2 * NZD + m] : U[m];
                    const double2 f2 = isb ? U[3 * NZD + m] : U[NZD + m];
                    const double sy = ky * kx, sz = kx * kz;
                    const double2 fy = make_double2((sy * f1.x - kz * f2.y) * ilap2, (sy * f1.y + kz * f2.x) * ilap2);
                    const double2 fzv = make_double2((sz * f1.x + ky * f2.y) * ilap2, (sz * f1.y - ky * f2.x) * ilap2);
                    if (l == 4 || l == 6) val = cscale(fac, fy);
                    else if (l == 5 || l == 7) val = cscale(fac, fzv);
                    else val = imul(fac * kz, fzv);  // dz uz
                }
            }
            buf[idx] = val;
        }
        __syncthreads();
        fft_lines<NZ, NL, true>(buf, tw);

        // products (MHD_BQlin.cpp:153-201) and EMF (:219-240); 5 output lines
        double2 prod[PW][3];
#pragma unroll
        for (int p = 0; p < PW; ++p) {
            const int item = tid + p * NT;
            if (PI_ % NT != 0 && item >= PI_) continue;
            const int mb = item / (2 * NZ), rem = item % (2 * NZ);
            const int part = rem / NZ, z = rem % NZ;
            const double2 *X = buf + mb * 9 * NZ + z;
            const double2 ux = X[0], ze = X[NZ], bx = X[2 * NZ], et = X[3 * NZ];
            const double2 uy = X[4 * NZ], uz = X[5 * NZ], by = X[6 * NZ], bz = X[7 * NZ], dzuz = X[8 * NZ];
            const double By = mfr[p][0], dzBy = mfr[p][1], dzdzBy = mfr[p][2];
            const double kx = skx[mb], ky = sky[mb];
            const bool valid = sidx[mb] < a.nxy;
            const double ftfac = (ky == 0.0 ? 1.0 : 2.0) * a.rsf;
            if (part == 0) {
                prod[p][0] = cscale(By, imul(ky, bx));                 // u, first part
                prod[p][1] = cscale(dzBy, cscale(2.0 * kx * ky, bz));  // u, part under ilapF
                prod[p][2] = cscale(By, imul(ky, ux));                 // b
                if (valid) acc[p] += ftfac * (redot(bz, ux) - redot(uz, bx));
            } else {
                // zeta: By (i ky eta) + dzBy (-i kx bx) + dzdzBy bz
                prod[p][0] = cadd(cadd(cscale(By, imul(ky, et)), cscale(dzBy, imul(-kx, bx))), cscale(dzdzBy, bz));
                // eta: By (i ky zeta) + dzBy (-2 dz uz - i kx ux) + dzdzBy (-uz)
                const double2 t2 = csub(cscale(-2.0, dzuz), imul(kx, ux));
                prod[p][1] = cadd(cadd(cscale(By, imul(ky, ze)), cscale(dzBy, t2)), cscale(-dzdzBy, uz));
                prod[p][2] = make_double2(0.0, 0.0);
                if (valid) acc[p] += ftfac * (redot(bz, uy) - redot(uz, by));
            }
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PW; ++p) {
            const int item = tid + p * NT;
            if (PI_ % NT != 0 && item >= PI_) continue;
            const int mb = item / (2 * NZ), rem = item % (2 * NZ);
            const int part = rem / NZ, z = rem % NZ;
            double2 *Y = buf + mb * 6 * NZ + z;
            if (part == 0) {
                Y[0] = prod[p][0];
                Y[NZ] = prod[p][1];
                Y[2 * NZ] = prod[p][2];
            } else {
                Y[3 * NZ] = prod[p][0];
                Y[4 * NZ] = prod[p][1];
            }
        }
        __syncthreads();
        fft_lines<NZ, 6 * MB, false>(buf, tw);

        // RHS assembly (MHD_BQlin.cpp:162-201), linear operator (:248-252) and
        // the substep update, per (mode, kz) slot and field pair
        for (int k = tid; k < MB * 2 * NZD; k += NT) {
            const int mb = k / (2 * NZD), r2 = k % (2 * NZD);
            const int pr = r2 / NZD, m = r2 % NZD;
            const int i = sidx[mb];
            if (i >= a.nxy) continue;
            const double2 *T = buf + mb * 6 * NZ + D::pad(m);
            const double2 *U = uin + mb * 4 * NZD;
            const double kx = skx[mb], ky = sky[mb], kz = a.kzv[m];
            const double kz2 = -(kz * kz);
            const double q = a.q, Om = a.omega;
            double2 N0, N1;
            if (pr == 0) {
                const double ilapF = silF[mb * NZD + m];
                const double2 s = cadd(cscale(2.0 * kx * ky * q, U[m]), imul(2.0 * Om * kz, U[NZD + m]));
                N0 = cadd(T[0], cscale(ilapF, cadd(s, T[NZ])));
                N1 = cadd(T[3 * NZ], imul((q - 2.0 * Om) * kz, U[m]));
            } else {
                N0 = T[2 * NZ];
                N1 = csub(T[4 * NZ], imul(q * kz, U[2 * NZD + m]));
            }
            const double kxn = kx + a.qdtl * ky;
            const double L0 = a.nu * ((-kxn * kxn - ky * ky) + kz2);
            const double Ln = (pr == 0) ? L0 : (a.eta / a.nu) * L0;
            const size_t off0 = (size_t)i * 4 * NZD + (2 * pr) * NZD + m, off1 = off0 + NZD;
            if (a.mode == MODE_RHS) {
                if (a.n_out) {
                    a.n_out[off0] = N0;
                    a.n_out[off1] = N1;
                }
                if (a.lin_out) a.lin_out[off0] = a.lin_out[off1] = Ln;
                continue;
            }
            double Lo = 0.0;
            if (a.mode == MODE_CN) {
                double L0o;
                if (a.old_init) {
                    const double kxo = __dadd_rn(skx0[mb], __dmul_rn(a.qt_old, ky));
                    const double lap2 = (ky == 0.0 && m == 0) ? 0.0 : (-ky * ky + kz2);
                    L0o = a.nu * (-kxo * kxo + lap2);
                } else {
                    const double kxo = __dadd_rn(skx0[mb], __dmul_rn(a.qt_old, ky)) + a.qdtl_old * ky;
                    L0o = a.nu * ((-kxo * kxo - ky * ky) + kz2);
                }
                Lo = (pr == 0) ? L0o : (a.eta / a.nu) * L0o;
            }
            const bool zero_out = a.do_remap && remap_hit(a, skx0[mb], ky);
            update_pair(a, off0, off1, U[2 * pr * NZD + m], U[(2 * pr + 1) * NZD + m], N0, N1, Ln, Lo, zero_out);
        }
        if (a.do_remap && tid < MB) {
            const int i = sidx[tid];
            if (i < a.nxy && remap_hit(a, skx0[tid], sky[tid])) a.kx0[i] = __dsub_rn(skx0[tid], a.remap_shift);
        }
        __syncthreads();
    }
    double *red = reinterpret_cast<double *>(buf);
#pragma unroll
    for (int p = 0; p < PW; ++p) {
        const int item = tid + p * NT;
        if (PI_ % NT != 0 && item >= PI_) continue;
        const int mb = item / (2 * NZ), rem = item % (2 * NZ);
        red[mb * 2 * NZ + rem] = acc[p];
    }
    __syncthreads();
    for (int k = tid; k < 2 * NZ; k += NT) {
        double s = 0.0;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) s += red[mb * 2 * NZ + k];
        a.partials[(size_t)blockIdx.x * 2 * NZ + k] = s;
    }
}

// ---------------------------------------------------------------------------
// mean fields: fixed-order reduction of the mode-kernel partials, drive, RHS,
// update, next profiles.  Launched with gridDim.x blocks: each adds a block of
// partial rows; the block that finishes last (atomic ticket, no waiting) adds
// the block sums in index order and does the rest.  The ticket counter is
// reset by that last block, so the kernel needs no memset between launches.
// ---------------------------------------------------------------------------
struct MFArgs {
    const double *part;   // mode-kernel partials [nrows][nMF][NZ]
    int nrows, rows_per;
    double *part2;        // block sums [gridDim.x][nMF][NZ]
    unsigned *counter;    // ticket, 0 between launches
    double2 *mf;              // [nMF][nz]
    double2 *mfn_out;         // mean-field RHS out [nMF][nz]
    const double2 *mfn_prev;  // previous substep mean-field RHS (nullable)
    double2 *drive;           // [nMF][nz]
    double *mfz;              // profiles out
    const double *kzv;
    const double2 *tw;
    int model, nMF, quasilinear, mode;
    double q, omega, B0z, nu, eta, fac, linC, nlc1, nlc2;
};

// sum of rows [r0, r1) of column c, in row order, loads issued 32 at a time
__device__ __forceinline__ double sum_rows(const double *p, int r0, int r1, int ncols, int c)
{
    constexpr int B = 32;
    double s = 0.0;
    for (int r = r0; r < r1; r += B) {
        double v[B];
#pragma unroll
        for (int k = 0; k < B; ++k) v[k] = (r + k < r1) ? p[(size_t)(r + k) * ncols + c] : 0.0;
#pragma unroll
        for (int k = 0; k < B; ++k) s += v[k];
    }
    return s;
}

template <int NZ>
__global__ __launch_bounds__(NT) void k_meanfield(MFArgs a)
{
    using D = Dealias<NZ>;
    constexpr int NZD = D::nz, HALF = D::half;
    __shared__ double2 buf[8 * NZ];
    __shared__ double2 tw[NZ];
    __shared__ double2 smf[4 * NZD];   // mean fields at the start of the substep
    __shared__ double2 smfn[4 * NZD];  // updated mean fields
    const int tid = threadIdx.x;
    const int nMF = a.nMF;
    for (int k = tid; k < NZ; k += NT) tw[k] = a.tw[k];

    if (a.mode != MODE_PROFILES) {
        __shared__ int amlast;
        const int ncols = nMF * NZ;
        const int r0 = blockIdx.x * a.rows_per;
        const int r1 = min(a.nrows, r0 + a.rows_per);
        for (int c = tid; c < ncols; c += NT) a.part2[(size_t)blockIdx.x * ncols + c] = sum_rows(a.part, r0, r1, ncols, c);
        __threadfence();
        __syncthreads();
        if (tid == 0) amlast = (atomicAdd(a.counter, 1u) == gridDim.x - 1);
        __syncthreads();
        if (!amlast) return;
        __threadfence();
        // fixed-order sum of the block sums -> real z-profiles
        for (int c = tid; c < ncols; c += NT) buf[c] = make_double2(sum_rows(a.part2, 0, gridDim.x, ncols, c), 0.0);
        if (tid == 0) *a.counter = 0u;
        __syncthreads();
        if (nMF == 4) fft_lines<NZ, 4, false>(buf, tw);
        else fft_lines<NZ, 2, false>(buf, tw);
        // EMF lines get d/dz = i kz; keep the non-negative half and fill the
        // negative half by conjugate symmetry (MHD_FullQlin.cpp:326-399)
        for (int k = tid; k < nMF * NZD; k += NT) {
            const int f = k / NZD, m = k % NZD;
            const bool emf = (f < 2);
            int src = m;
            bool cj = false;
            if (m >= HALF) {
                src = HALF - 1 - (m - HALF);
                cj = true;
            }
            double2 v = buf[f * NZ + src];
            if (emf) v = imul(a.kzv[src], v);
            if (cj) v.y = -v.y;
            a.drive[k] = v;
        }
        for (int k = tid; k < nMF * NZD; k += NT) smf[k] = a.mf[k];
        __syncthreads();
        // mean-field RHS (MHD_FullQlin.cpp:411-423, MHD_BQlin.cpp:274-280) and
        // update (RK3CN_Elin.cpp:109-112 / RK3CN.cpp:112-115)
        for (int k = tid; k < nMF * NZD; k += NT) {
            const int f = k / NZD, m = k % NZD;
            const double kz = a.kzv[m], kz2 = -(kz * kz);
            double2 N = make_double2(0.0, 0.0);
            double Lmf = 0.0;
            if (a.quasilinear) {
                const double2 Dv = a.drive[k];
                if (a.model == 0) {
                    const double2 M0 = smf[m], M1 = smf[NZD + m], M2 = smf[2 * NZD + m], M3 = smf[3 * NZD + m];
                    if (f == 0) N = cadd(imul(a.B0z * kz, M2), Dv);
                    else if (f == 1) N = cadd(cadd(cscale(-a.q, M0), imul(a.B0z * kz, M3)), Dv);
                    else if (f == 2) N = cadd(cadd(cscale(2 * a.omega, M3), imul(a.B0z * kz, M0)), Dv);
                    else N = cadd(cadd(cscale(a.q - 2.0 * a.omega, M2), imul(a.B0z * kz, M1)), Dv);
                    Lmf = (f < 2 ? a.eta : a.nu) * kz2;
                } else {
                    if (f == 0) N = Dv;
                    else N = cadd(cscale(-a.q, smf[m]), Dv);
                    Lmf = a.eta * kz2;
                }
            }
            a.mfn_out[k] = N;
            if (a.mode == MODE_RHS) continue;
            const double2 M = smf[k];
            double2 Mn;
            if (a.mode == MODE_ELIN) {
                double2 s = cscale(a.nlc1, N);
                if (a.mfn_prev) s = cadd(s, cscale(a.nlc2, a.mfn_prev[k]));
                Mn = cscale(exp(a.linC * Lmf), cadd(s, M));
            } else {
                const double denom = 1.0 / (1.0 - a.linC * Lmf);
                double2 s = cscale(a.nlc1 * denom, N);
                if (a.mfn_prev) s = cadd(s, cscale(a.nlc2 * denom, a.mfn_prev[k]));
                Mn = cadd(s, cscale((1 + a.linC * Lmf) * denom, M));
            }
            a.mf[k] = Mn;
            smfn[k] = Mn;
        }
        __syncthreads();
        if (a.mode == MODE_RHS) return;
    } else {
        for (int k = tid; k < nMF * NZD; k += NT) smfn[k] = a.mf[k];
        __syncthreads();
    }
    // real-space profiles for the next rhs (MHD_FullQlin.cpp:180-198,
    // MHD_BQlin.cpp:122-127): fac1D * ifft(pad(...)), real part
    const int nlines = (a.model == 0) ? 8 : 3;
    for (int idx = tid; idx < 8 * NZ; idx += NT) {
        const int l = idx / NZ, j = idx % NZ;
        double2 v = make_double2(0.0, 0.0);
        if (l < nlines && D::kept(j)) {
            const int m = D::slot(j);
            const double kz = a.kzv[m];
            if (a.model == 0) {
                const double2 M = smfn[(l >> 1) * NZD + m];
                v = (l & 1) ? imul(a.fac * kz, M) : cscale(a.fac, M);
            } else {
                const double2 M = smfn[NZD + m];  // By
                if (l == 0) v = cscale(a.fac, M);
                else if (l == 1) v = imul(a.fac * kz, M);
                else v = cscale(a.fac * (-(kz * kz)), M);
            }
        }
        buf[idx] = v;
    }
    __syncthreads();
    fft_lines<NZ, 8, true>(buf, tw);
    for (int idx = tid; idx < nlines * NZ; idx += NT) a.mfz[idx] = buf[idx].x;
}

// ---------------------------------------------------------------------------
// ShearingBox_Remap (MHD_FullQlin.cpp:464-481) as a standalone pass
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_remap(double2 *u, double *kx0, const double *ky, int nxy, int nz,
                                              double qt, double thr, double shift)
{
    for (int i = blockIdx.x; i < nxy; i += gridDim.x) {
        const double k0 = kx0[i], kyv = ky[i];
        const bool hit = __dadd_rn(k0, __dmul_rn(qt, kyv)) > thr;
        __syncthreads();
        if (hit) {
            for (int k = threadIdx.x; k < 4 * nz; k += NT) u[(size_t)i * 4 * nz + k] = make_double2(0.0, 0.0);
            if (threadIdx.x == 0) kx0[i] = __dsub_rn(k0, shift);
        }
    }
}

// ---------------------------------------------------------------------------
// Calc_Energy_AM_Diss (MHD_FullQlin.cpp:515-559): per-workgroup partial sums
// of energy, angular momentum and dissipation over modes, in a fixed order
// ---------------------------------------------------------------------------
struct DiagArgs {
    const double2 *u;
    const double *kx0, *ky, *kzv;
    int nxy, nz;
    double qt, nu, eta;
    double *part;  // [gridDim.x][6]
};

__global__ __launch_bounds__(NT) void k_diag_modes(DiagArgs a)
{
    __shared__ double sred[6][NT];
    double s[6] = {0, 0, 0, 0, 0, 0};
    const int nz = a.nz;
    for (int i = blockIdx.x; i < a.nxy; i += gridDim.x) {
        const double ky = a.ky[i];
        const double kx = __dadd_rn(a.kx0[i], __dmul_rn(a.qt, ky));
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
__global__ void k_diag_final(DiagFinalArgs a)
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

}  // namespace qlk
