// ql_adjoint.hpp -- reverse-mode (adjoint) kernels of one RK3 substep of the
// MHD_FullQlin and MHD_BQlin quasilinear steps in FP64
// (Models/MHD_FullQlin.cpp:171-427, Models/MHD_BQlin.cpp:117-306, with the
// substep updates of Integrators/RK3CN_Elin.cpp:290-378 and
// Integrators/RK3CN.cpp:92-193).
//
// Gradients of a real loss L with respect to a complex value z are carried as
// G = dL/dRe z + i dL/dIm z, so dL = Re(conj(G) dz).  For z' = a z the
// reverse rule is G_z += conj(a) G_z'; for r = Re(a conj b) it is
// G_a += dL/dr b and G_b += dL/dr a.
//
// The transposes of the dealiased transforms are the other transform: the
// adjoint of "zero-pad nz -> NZ, unnormalised exp(+i) transform" is the
// unnormalised exp(-i) transform followed by truncation NZ -> nz, and vice
// versa, so the reverse pass reuses the forward transform code unchanged.
//
// One reverse substep (host sequence, ql_adjoint.hip):
//   k_adj_mf_a    one workgroup: mean-field update and mean-field RHS in
//                 reverse, drive (EMF / stress) reverse down to dL/d(stress
//                 z-profiles) by an inverse transform;
//   k_adj_mode    per mode: recompute the forward lines and products, then
//                 reverse the update, the RHS assembly, the forward transforms,
//                 the mean-field x fluctuation products and the EMF / stress
//                 products, the inverse transforms and the base fields; per
//                 workgroup partial rows of dL/d(mean-field z-profiles) and of
//                 the parameter gradients;
//   k_reduce_cols column sum of those rows;
//   k_adj_mf_b    one workgroup: z-profile adjoint into dL/dM, parameter
//                 gradient accumulation.
// The remap decision (MHD_FullQlin.cpp:470-472) is piecewise constant in q and
// enters as a fixed mask.
#pragma once
#include <type_traits>

#include "ql_mode.hpp"
#include "ql_noise.hpp"

namespace qlk {

// parameter-gradient slots (columns after the 8 z-profiles of a partial row)
enum { GP_Q = 0, GP_OMEGA = 1, GP_NU = 2, GP_ETA = 3, GP_B0Z = 4, NGP = 8 };

// d/dx of -|x|^p (neg_abs_pow)
__device__ __forceinline__ double neg_abs_pow_d(double x, int p)
{
    if (p == 0) return 0.0;
    const double a = fabs(x), sg = x < 0 ? -1.0 : (x > 0 ? 1.0 : 0.0);
    if (p == 1) return -sg;
    if (p == 2) return -2.0 * x;
    return -(double)p * pow(a, (double)(p - 1)) * sg;
}

struct AdjModeArgs {
    const double2 *u_in;   // state at the substep start [nxy][4][nz]
    const double2 *u_out;  // state after the substep (after the remap at s = 3)
    double2 *gu;           // in: dL/du_out; out: dL/du_in
    double2 *gnp;          // in (s < 3): dL/dN_s via substep s+1; out (s > 1): dL/dN_{s-1}
    const double *kx0, *ky, *kzv;
    const double2 *tw;
    const double *mfz;     // z-profiles of the substep's mean fields [8][NZ]
    const double *gsums;   // dL/d(EMF / stress z-profiles) [4][NZ]
    double *gpart;         // [gridDim.x][8 NZ + NGP]
    Phys ph;
    int nxy, s, mode, do_remap, oinit;
    double rsf, remap_thr, fac;
    double linC, nlc1, nlc2;
    double ts, dtl, tso, dtlo, tnew;  // substep times (sub_times of the forward)
    // driving noise added after substep 3 (before the remap): its multipliers
    // depend on q through kx(t) at the step start t_step
    NoiseParams np;
    int noise;
    double t_step, dt_step;
};

// first-pass source of the reverse transforms: compact kz-slot lines
template <int NZ>
struct CompactSrc {
    const double2 *v;  // [line][nz]
    __device__ double2 operator()(int line, int g) const
    {
        using D = Dealias<NZ>;
        return D::kept(g) ? v[line * D::nz + D::slot(g)] : make_double2(0.0, 0.0);
    }
};

// P = ModelFull (MHD_FullQlin) or ModelBQ (MHD_BQlin.cpp:117-306: 8 base
// lines, 5 product lines, By profiles only, no B0z terms, and the dz uz
// constant of the ky = 0 modes, whose gradient is summed over z per mode)
template <int NZ, int MB, class P>
__global__ __launch_bounds__(NT) void k_adj_mode(AdjModeArgs a)
{
    constexpr bool BQ = std::is_same<P, ModelBQ>::value;
    using D = Dealias<NZ>;
    constexpr int NZD = D::nz;
    constexpr int NLX = P::NLINE * MB;  // 6 (8) lines per mode
    constexpr int NLT = P::NOUT * MB;   // 6 (5) lines per mode
    constexpr int HI = MB * 2 * NZ;     // product items (mode, half, z)
    constexpr int HW = (HI + NT - 1) / NT;
    constexpr int NBS = P::NB > P::NOUT ? P::NB : P::NOUT;
    constexpr int GL = BQ ? 4 : 3;      // lines whose gradient one half-item forms
    __shared__ double2 bufX[NLX * NZ];  // X lines, then dL/dX lines
    __shared__ double2 bufT[NLT * NZ];  // P -> T, then dL/dP
    __shared__ double2 uin[MB * 4 * NZD];
    __shared__ double2 sgu[MB * 4 * NZD];      // dL/du_out rows, then dL/du_in (direct terms)
    __shared__ double2 sbase[MB * NBS * NZD];  // bases, then dL/dT (compact)
    __shared__ double2 sdz[MB], sgdz[MB];      // ModelBQ: fac ux at kz = 0 of a ky = 0 mode, and its gradient
    __shared__ double2 gdzz[BQ ? MB * NZ : 1]; // ModelBQ: per-z terms of that gradient
    __shared__ double2 tw[NZ];
    __shared__ double skz[NZD];
    __shared__ double silF[MB * NZD];
    __shared__ double skx0[MB], sky[MB], skx[MB];
    __shared__ int sidx[MB];
    __shared__ double smfz[NZ * 8];  // [z][profile]
    __shared__ double sgs[4 * NZ];   // dL/d(stress sums) [slot][z]
    const int tid = threadIdx.x;
    const Phys ph = a.ph;
    const double q = ph.q, Om = ph.omega, B0z = ph.B0z;
    // d L / d coefficient and d L / d x of a dissipation operator coef * op(x),
    // x = -kx^2 - ky^2 + kz2: -|x|^p for MHD_FullQlin, x itself for MHD_BQlin
    const auto opv = [&](double x) { return BQ ? x : neg_abs_pow(x, ph.vo2); };
    const auto opd = [&](double x) { return BQ ? 1.0 : neg_abs_pow_d(x, ph.vo2); };
    const double qt = to_sgpr(rn_mul(q, a.ts)), qdtl = to_sgpr(rn_mul(q, a.dtl));
    const double qt_old = to_sgpr(rn_mul(q, a.tso)), qdtl_old = to_sgpr(rn_mul(q, a.dtlo));
    const double remap_qt = to_sgpr(rn_mul(q, a.tnew));
    const bool remap_now = a.s == 3 && a.do_remap;
    for (int k = tid; k < NZ; k += NT) tw[k] = a.tw[k];
    for (int k = tid; k < NZD; k += NT) skz[k] = a.kzv[k];
    for (int k = tid; k < NZ * 8; k += NT) smfz[(k % NZ) * 8 + k / NZ] = a.mfz[k];
    for (int k = tid; k < 4 * NZ; k += NT) sgs[k] = a.gsums[k];

    // per-thread gradient accumulators: parameters and z-profile rows
    double gq = 0, gom = 0, gnu = 0, geta = 0, gb0 = 0;
    double accm[HW][8];
#pragma unroll
    for (int p = 0; p < HW; ++p)
#pragma unroll
        for (int c = 0; c < 8; ++c) accm[p][c] = 0.0;

    const int ngroups = (a.nxy + MB - 1) / MB;
    for (int g = blockIdx.x; g < ngroups; g += gridDim.x) {
        // A: rows of this group
        if (tid < MB) {
            const int i = g * MB + tid;
            sidx[tid] = i < a.nxy ? i : -1;
            if (i < a.nxy) {
                const double kx0 = a.kx0[i], ky = a.ky[i];
                skx0[tid] = kx0;
                sky[tid] = ky;
                skx[tid] = rn_add(kx0, rn_mul(qt, ky));
                sdz[tid] = (BQ && ky == 0.0) ? cscale(a.fac, a.u_in[(size_t)i * 4 * NZD]) : make_double2(0.0, 0.0);
            } else {
                skx0[tid] = sky[tid] = skx[tid] = 0.0;
                sdz[tid] = make_double2(0.0, 0.0);
            }
        }
        for (int k = tid; k < MB * 4 * NZD; k += NT) {
            const int mb = k / (4 * NZD), i = g * MB + mb;
            const bool ok = i < a.nxy;
            const size_t off = (size_t)i * 4 * NZD + (k - mb * 4 * NZD);
            uin[k] = ok ? a.u_in[off] : make_double2(0.0, 0.0);
            sgu[k] = ok ? a.gu[off] : make_double2(0.0, 0.0);
        }
        __syncthreads();

        // B, C: bases and inverse transforms (forward recomputation)
        for (int k = tid; k < MB * NZD; k += NT) {
            const int mb = k / NZD, m = k % NZD;
            const double kx = skx[mb], ky = sky[mb], kz = skz[m];
            const bool zero = (ky == 0.0 && m == 0);
            const double lap2 = zero ? 0.0 : (-ky * ky + -(kz * kz));
            const double ilap2 = zero ? 1.0 : 1.0 / lap2;
            silF[k] = (sidx[mb] == 0 && m == 0) ? 1.0 : 1.0 / (-kx * kx + lap2);
            const double2 *U = uin + mb * 4 * NZD;
            P::bases(U[m], U[NZD + m], U[2 * NZD + m], U[3 * NZD + m], kx, ky, kz, ilap2, a.fac,
                     sbase + mb * P::NB * NZD + m, NZD);
        }
        __syncthreads();
        fft_lines_from<NZ, NLX, true>(bufX, tw, LineSrc<NZ, P, double2>{sbase, skz});

        // D, E: products and forward transforms (forward recomputation)
        for (int item = tid; item < HI; item += NT) {
            const int mb = item / (2 * NZ), rem = item % (2 * NZ);
            const int part = rem / NZ, z = rem % NZ;
            double2 prod[3];
            double racc[2];
            P::products(part, bufX + mb * P::NLINE * NZ + z, NZ, smfz + z * 8, skx[mb], sky[mb],
                        imul(skx[mb], sdz[mb]), prod, racc);
#pragma unroll
            for (int c = 0; c < 3; ++c)
                if (c < P::nprod(part)) bufT[(mb * P::NOUT + P::out_line(part, c)) * NZ + z] = prod[c];
        }
        __syncthreads();
        fft_lines<NZ, NLT, false>(bufT, tw);

        // F: substep update and RHS assembly in reverse, per (mode, kz slot)
        for (int k = tid; k < MB * NZD; k += NT) {
            const int mb = k / NZD, m = k % NZD;
            const int i = sidx[mb];
            double2 *GT = sbase + mb * P::NOUT * NZD + m;  // dL/dT, compact [line][nz]
            if (i < 0) {
                for (int o = 0; o < P::NOUT; ++o) GT[o * NZD] = make_double2(0.0, 0.0);
                continue;
            }
            const double kx = skx[mb], ky = sky[mb], kz = skz[m];
            const double kz2 = -(kz * kz);
            double2 *Gu = sgu + mb * 4 * NZD + m;
            const double2 *U = uin + mb * 4 * NZD + m;
            const double2 ux = U[0], ze = U[NZD], bx = U[2 * NZD], et = U[3 * NZD];
            double2 gout[4], GN[4], gdir[4];
            const bool hit = remap_now && remap_hit(remap_qt, a.remap_thr, skx0[mb], ky);
            for (int f = 0; f < 4; ++f) gout[f] = hit ? make_double2(0.0, 0.0) : Gu[f * NZD];
            // the step's driving noise (ql_noise.hpp) is state independent; its
            // ux and bx multipliers sqrt(nf lap2 / lapF) depend on q through
            // lapF = -kx(t)^2 + lap2, d/dq = mult kx t ky / lapF (ky = 0 modes:
            // none); the band mask is held fixed like the remap decision
            if (a.noise && a.s == 3 && !hit && a.np.sigma != 0.0 && ky != 0.0) {
                const NoiseParams &np = a.np;
                const double tn = a.t_step;
                const double kxn = rn_add(skx0[mb], rn_mul(rn_mul(q, tn), ky));
                const double lap2n = rn_add(rn_mul(-ky, ky), kz2);
                const double lapF = rn_add(rn_mul(-kxn, kxn), lap2n);
                if (lapF > -np.hi2 && lapF < -np.lo2) {
                    const double nf = rn_mul(a.dt_step, np.nfac);
                    const double mu = sqrt(rn_mul(nf, lap2n) / lapF);
                    double2 z[4];
                    noise_normals(np.seed, tn, i, m, z);
                    double gz = 0.0;
                    if (np.drive_u) gz += redot(z[0], gout[0]);
                    if (np.drive_b) gz += redot(z[2], gout[2]);
                    gq += mu * kxn * (tn * ky) / lapF * np.sigma * gz;
                }
            }
            const size_t off = (size_t)i * 4 * NZD + m;
            for (int pr = 0; pr < 2; ++pr) {
                const double coef = pr == 0 ? ph.nu : ph.eta;
                const double kxn = kx + qdtl * ky;
                const double xn = (-kxn * kxn - ky * ky) + kz2;
                const double Ln = P::lin(pr, kxn, ky, kz2, ph);
                double gLn = 0.0, gLo = 0.0, xo = 0.0, dxo_dq = 0.0, kxo_e = 0.0;
                for (int h = 0; h < 2; ++h) {
                    const int f = 2 * pr + h;
                    const double2 go = gout[f];
                    const double2 gin = (a.s < 3) ? a.gnp[off + f * NZD] : make_double2(0.0, 0.0);
                    const double2 uo = a.u_out[off + f * NZD];
                    if (a.mode == MODE_ELIN) {
                        const double e = exp(a.linC * Ln);
                        const double2 mu = cscale(e, go);
                        GN[f] = cadd(cscale(a.nlc1, mu), gin);
                        if (a.s > 1) a.gnp[off + f * NZD] = cscale(a.nlc2, mu);
                        gdir[f] = mu;
                        gLn += a.linC * redot(uo, go);
                    } else {
                        double Lo;
                        const double kxo = rn_add(skx0[mb], rn_mul(qt_old, ky));
                        if (a.oinit) {
                            const double lap2 = (ky == 0.0 && m == 0) ? 0.0 : (-ky * ky + kz2);
                            xo = -kxo * kxo + lap2;
                            kxo_e = kxo;
                            dxo_dq = a.tso * ky;
                            Lo = P::lin_init(pr, kxo, lap2, ph);
                        } else {
                            kxo_e = kxo + qdtl_old * ky;
                            xo = (-kxo_e * kxo_e - ky * ky) + kz2;
                            dxo_dq = (a.tso + a.dtlo) * ky;
                            Lo = P::lin(pr, kxo_e, ky, kz2, ph);
                        }
                        const double d = 1.0 / (1.0 - a.linC * Ln);
                        GN[f] = cadd(cscale(a.nlc1 * d, go), gin);
                        if (a.s > 1) a.gnp[off + f * NZD] = cscale(a.nlc2 * d, go);
                        gdir[f] = cscale((1 + a.linC * Lo) * d, go);
                        gLn += a.linC * d * redot(uo, go);
                        gLo += a.linC * d * redot(U[f * NZD], go);
                    }
                }
                // L = coef * op(x): coefficient and kx(t) = kx0 + q t ky dependence
                const double gc = gLn * opv(xn) + (a.mode == MODE_CN ? gLo * opv(xo) : 0.0);
                if (pr == 0) gnu += gc;
                else geta += gc;
                gq += gLn * coef * opd(xn) * (-2.0 * kxn) * ((a.ts + a.dtl) * ky);
                if (a.mode == MODE_CN) gq += gLo * coef * opd(xo) * (-2.0 * kxo_e) * dxo_dq;
            }
            if constexpr (!BQ) {
                // RHS assembly in reverse (MHD_FullQlin.cpp:283-285, :303-305)
                const double2 *Tm = bufT + mb * P::NOUT * NZ + D::pad(m);
                const double2 T1 = Tm[0], T2 = Tm[NZ], T3 = Tm[2 * NZ];
                const double ilapF = silF[k];
                double gkx = 0.0;
                // N_ux = ilapF s + i B0z kz bx - T1,  s = 2 kx ky q ux + i 2 Om kz ze - kx^2 T1 - kx ky T2 - kx kz T3
                double2 s = cscale(2.0 * kx * ky * q, ux);
                s = cadd(s, imul(2.0 * Om * kz, ze));
                s = csub(s, cscale(kx * kx, T1));
                s = csub(s, cscale(kx * ky, T2));
                s = csub(s, cscale(kx * kz, T3));
                const double2 Gs = cscale(ilapF, GN[0]);
                double2 GT0 = csub(cscale(-kx * kx, Gs), GN[0]);
                double2 GT1 = cscale(-kx * ky, Gs);
                double2 GT2 = cscale(-kx * kz, Gs);
                gdir[0] = cadd(gdir[0], cscale(2.0 * kx * ky * q, Gs));
                gdir[1] = cadd(gdir[1], imul(-2.0 * Om * kz, Gs));
                gdir[2] = cadd(gdir[2], imul(-B0z * kz, GN[0]));
                {
                    double2 dsk = cscale(2.0 * ky * q, ux);
                    dsk = csub(dsk, cscale(2.0 * kx, T1));
                    dsk = csub(dsk, cscale(ky, T2));
                    dsk = csub(dsk, cscale(kz, T3));
                    gkx += redot(dsk, Gs);
                }
                if (!(i == 0 && m == 0)) gkx += redot(s, GN[0]) * 2.0 * kx * ilapF * ilapF;
                gq += redot(cscale(2.0 * kx * ky, ux), Gs);
                gom += redot(imul(2.0 * kz, ze), Gs);
                gb0 += redot(imul(kz, bx), GN[0]);
                // N_ze = i (q - 2 Om) kz ux + i B0z kz et - i kz T2 + i ky T3
                gdir[0] = cadd(gdir[0], imul(-(q - 2.0 * Om) * kz, GN[1]));
                gdir[3] = cadd(gdir[3], imul(-B0z * kz, GN[1]));
                GT1 = cadd(GT1, imul(kz, GN[1]));
                GT2 = cadd(GT2, imul(-ky, GN[1]));
                gq += redot(imul(kz, ux), GN[1]);
                gom += redot(imul(-2.0 * kz, ux), GN[1]);
                gb0 += redot(imul(kz, et), GN[1]);
                // N_bx = i B0z kz ux - T4
                gdir[0] = cadd(gdir[0], imul(-B0z * kz, GN[2]));
                const double2 GT3 = cscale(-1.0, GN[2]);
                gb0 += redot(imul(kz, ux), GN[2]);
                // N_et = -i q kz bx + i B0z kz ze - i kz T5 + i ky T6
                gdir[2] = cadd(gdir[2], imul(q * kz, GN[3]));
                gdir[1] = cadd(gdir[1], imul(-B0z * kz, GN[3]));
                const double2 GT4 = imul(kz, GN[3]);
                const double2 GT5 = imul(-ky, GN[3]);
                gq += redot(imul(-kz, bx), GN[3]);
                gb0 += redot(imul(kz, ze), GN[3]);
                gq += gkx * (a.ts * ky);
                GT[0] = GT0;
                GT[NZD] = GT1;
                GT[2 * NZD] = GT2;
                GT[3 * NZD] = GT3;
                GT[4 * NZD] = GT4;
                GT[5 * NZD] = GT5;
            } else {
                // MHD_BQlin.cpp:160-162, :177-178, :185, :200-201 in reverse:
                // N_ux = T1 + ilapF (s + T2), s = 2 kx ky q ux + i 2 Om kz ze;
                // N_ze = T4 + i (q - 2 Om) kz ux; N_bx = T3; N_et = T5 - i q kz bx
                const double2 *Tm = bufT + mb * P::NOUT * NZ + D::pad(m);
                const double2 T2 = Tm[NZ];
                const double ilapF = silF[k];
                double gkx = 0.0;
                const double2 s = cadd(cscale(2.0 * kx * ky * q, ux), imul(2.0 * Om * kz, ze));
                const double2 Gs = cscale(ilapF, GN[0]);
                gdir[0] = cadd(gdir[0], cscale(2.0 * kx * ky * q, Gs));
                gdir[1] = cadd(gdir[1], imul(-2.0 * Om * kz, Gs));
                gkx += redot(cscale(2.0 * ky * q, ux), Gs);
                if (!(i == 0 && m == 0)) gkx += redot(cadd(s, T2), GN[0]) * 2.0 * kx * ilapF * ilapF;
                gq += redot(cscale(2.0 * kx * ky, ux), Gs);
                gom += redot(imul(2.0 * kz, ze), Gs);
                gdir[0] = cadd(gdir[0], imul(-(q - 2.0 * Om) * kz, GN[1]));
                gq += redot(imul(kz, ux), GN[1]);
                gom += redot(imul(-2.0 * kz, ux), GN[1]);
                gdir[2] = cadd(gdir[2], imul(q * kz, GN[3]));
                gq += redot(imul(-kz, bx), GN[3]);
                gq += gkx * (a.ts * ky);
                GT[0] = GN[0];
                GT[NZD] = Gs;
                GT[2 * NZD] = GN[2];
                GT[3 * NZD] = GN[1];
                GT[4 * NZD] = GN[3];
            }
            for (int f = 0; f < 4; ++f) Gu[f * NZD] = gdir[f];
        }
        __syncthreads();

        // G: adjoint of the forward transforms = padded inverse transforms
        fft_lines_from<NZ, NLT, true>(bufT, tw, CompactSrc<NZ>{sbase});

        // H: products and EMF / stress terms in reverse, per (mode, half, z):
        // half 0 owns the u lines (ux uy uz), half 1 the b lines (bx by bz)
        double2 gX[HW][GL];
#pragma unroll
        for (int p = 0; p < HW; ++p) {
            const int item = tid + p * NT;
            if (HI % NT != 0 && item >= HI) continue;
            const int mb = item / (2 * NZ), rem = item % (2 * NZ);
            const int h = rem / NZ, z = rem % NZ;
#pragma unroll
            for (int c = 0; c < GL; ++c) gX[p][c] = make_double2(0.0, 0.0);
            if (BQ && h == 0) gdzz[mb * NZ + z] = make_double2(0.0, 0.0);
            if (sidx[mb] < 0) continue;
            if constexpr (BQ) {
                // MHD_BQlin products (MHD_BQlin.cpp:153-201, :219-240) in reverse;
                // X lines ux ze bx et uy uz by bz, P lines: By i ky bx,
                // dzBy 2 kx ky bz, By i ky ux, (zeta), (eta)
                const double2 *X = bufX + mb * P::NLINE * NZ + z;
                const double2 ux = X[0], ze = X[NZ], bx = X[2 * NZ], et = X[3 * NZ];
                const double2 uy = X[4 * NZ], uz = X[5 * NZ], by = X[6 * NZ], bz = X[7 * NZ];
                const double2 *GPp = bufT + mb * P::NOUT * NZ + z;
                const double2 G0 = GPp[0], G1 = GPp[NZ], G2 = GPp[2 * NZ], G3 = GPp[3 * NZ], G4 = GPp[4 * NZ];
                const double *mf = smfz + z * 8;
                const double By = mf[0], dzBy = mf[1], dzdzBy = mf[2];
                const double kx = skx[mb], ky = sky[mb];
                const double ftfac = (ky == 0.0 ? 1.0 : 2.0) * a.rsf;
                const double g0 = ftfac * sgs[z], g1 = ftfac * sgs[NZ + z];  // EMF x, EMF y
                double gkx = 0.0;
                if (h == 0) {
                    // ux, ze, uy, uz; dz uz = -i kx ux - i ky uy + dzc enters the
                    // eta product as dzBy (-2 dz uz - i kx ux)
                    const double2 dzuz = cadd(cadd(imul(-kx, ux), imul(-ky, uy)), imul(kx, sdz[mb]));
                    const double2 t2 = csub(cscale(-2.0, dzuz), imul(kx, ux));
                    gX[p][0] = cadd(cadd(imul(-ky * By, G2), imul(-kx * dzBy, G4)), cscale(g0, bz));
                    gX[p][1] = imul(-ky * By, G4);
                    gX[p][2] = cadd(imul(-2.0 * ky * dzBy, G4), cscale(g1, bz));
                    gX[p][3] = csub(cscale(-dzdzBy, G4), cadd(cscale(g0, bx), cscale(g1, by)));
                    gdzz[mb * NZ + z] = cscale(-2.0 * dzBy, G4);
                    accm[p][0] += redot(imul(ky, ux), G2) + redot(imul(ky, ze), G4);
                    accm[p][1] += redot(t2, G4);
                    accm[p][2] -= redot(uz, G4);
                    gkx += redot(imul(dzBy, ux), G4);
                } else {
                    // bx, et, by, bz
                    gX[p][0] = csub(cadd(imul(-ky * By, G0), imul(kx * dzBy, G3)), cscale(g0, uz));
                    gX[p][1] = imul(-ky * By, G3);
                    gX[p][2] = cscale(-g1, uz);
                    gX[p][3] = cadd(cadd(cscale(2.0 * kx * ky * dzBy, G1), cscale(dzdzBy, G3)),
                                    cadd(cscale(g0, ux), cscale(g1, uy)));
                    accm[p][0] += redot(imul(ky, bx), G0) + redot(imul(ky, et), G3);
                    accm[p][1] += redot(cscale(2.0 * kx * ky, bz), G1) + redot(imul(-kx, bx), G3);
                    accm[p][2] += redot(bz, G3);
                    gkx += redot(cscale(2.0 * ky * dzBy, bz), G1) + redot(imul(-dzBy, bx), G3);
                }
                gq += gkx * (a.ts * ky);
                continue;
            }
            const double2 *X = bufX + mb * P::NLINE * NZ + z;
            const double2 u0 = X[0], u1 = X[NZ], u2 = X[2 * NZ], b0 = X[3 * NZ], b1 = X[4 * NZ], b2 = X[5 * NZ];
            const double2 *GPp = bufT + mb * P::NOUT * NZ + z;
            const double2 G0 = GPp[0], G1 = GPp[NZ], G2 = GPp[2 * NZ], G3 = GPp[3 * NZ], G4 = GPp[4 * NZ],
                          G5 = GPp[5 * NZ];
            const double *mf = smfz + z * 8;
            const double MBx = mf[0], dzMBx = mf[1], MBy = mf[2], dzMBy = mf[3];
            const double MUx = mf[4], dzMUx = mf[5], MUy = mf[6], dzMUy = mf[7];
            const double kx = skx[mb], ky = sky[mb];
            const double aU = kx * MUx + ky * MUy, aB = kx * MBx + ky * MBy;
            const double ftfac = (ky == 0.0 ? 1.0 : 2.0) * a.rsf;
            const double g0 = ftfac * sgs[z], g1 = ftfac * sgs[NZ + z], g2 = ftfac * sgs[2 * NZ + z],
                         g3 = ftfac * sgs[3 * NZ + z];
            double gkx = 0.0;
            if (h == 0) {
                // u0 (0), u1 (1), u2 (2)
                double2 Gu0 = cadd(imul(-aU, G0), imul(aB, G3));
                double2 Gu1 = cadd(imul(-aU, G1), imul(aB, G4));
                double2 Gu2 = cadd(cadd(cscale(dzMUx, G0), cscale(dzMUy, G1)), imul(-aU, G2));
                Gu2 = cadd(Gu2, cadd(cadd(cscale(dzMBx, G3), cscale(dzMBy, G4)), imul(aB, G5)));
                // EMF x, y (slots 0, 1) and stresses (slots 2, 3:
                // -Re(uz conj(ux)), -Re(uz conj(uy)) before their z-derivative)
                Gu0 = cadd(Gu0, cscale(g0, b2));
                Gu2 = csub(Gu2, cadd(cscale(g0, b0), cscale(g1, b1)));
                Gu1 = cadd(Gu1, cscale(g1, b2));
                Gu0 = csub(Gu0, cscale(g2, u2));
                Gu1 = csub(Gu1, cscale(g3, u2));
                Gu2 = csub(Gu2, cadd(cscale(g2, u0), cscale(g3, u1)));
                gX[p][0] = Gu0;
                gX[p][1] = Gu1;
                gX[p][2] = Gu2;
                // mean-field profile and kx gradients from the first three products
                const double gaU = redot(imul(1.0, u0), G0) + redot(imul(1.0, u1), G1) + redot(imul(1.0, u2), G2);
                const double gaB = redot(imul(-1.0, b0), G0) + redot(imul(-1.0, b1), G1) + redot(imul(-1.0, b2), G2);
                accm[p][5] += redot(u2, G0);   // dz MUx
                accm[p][1] -= redot(b2, G0);   // dz MBx
                accm[p][7] += redot(u2, G1);   // dz MUy
                accm[p][3] -= redot(b2, G1);   // dz MBy
                accm[p][4] += kx * gaU;
                accm[p][6] += ky * gaU;
                accm[p][0] += kx * gaB;
                accm[p][2] += ky * gaB;
                gkx += MUx * gaU + MBx * gaB;
            } else {
                // b0 (0), b1 (1), b2 (2)
                double2 Gb0 = cadd(imul(aB, G0), imul(-aU, G3));
                double2 Gb1 = cadd(imul(aB, G1), imul(-aU, G4));
                double2 Gb2 = cadd(cadd(cscale(-dzMBx, G0), cscale(-dzMBy, G1)), imul(aB, G2));
                Gb2 = cadd(Gb2, cadd(cadd(cscale(-dzMUx, G3), cscale(-dzMUy, G4)), imul(-aU, G5)));
                Gb2 = cadd(Gb2, cadd(cscale(g0, u0), cscale(g1, u1)));
                Gb0 = csub(Gb0, cscale(g0, u2));
                Gb1 = csub(Gb1, cscale(g1, u2));
                // stresses: +Re(bz conj(bx)), +Re(bz conj(by))
                Gb0 = cadd(Gb0, cscale(g2, b2));
                Gb1 = cadd(Gb1, cscale(g3, b2));
                Gb2 = cadd(Gb2, cadd(cscale(g2, b0), cscale(g3, b1)));
                gX[p][0] = Gb0;
                gX[p][1] = Gb1;
                gX[p][2] = Gb2;
                const double gaU = redot(imul(1.0, b0), G3) + redot(imul(1.0, b1), G4) + redot(imul(1.0, b2), G5);
                const double gaB = redot(imul(-1.0, u0), G3) + redot(imul(-1.0, u1), G4) + redot(imul(-1.0, u2), G5);
                accm[p][1] += redot(u2, G3);   // dz MBx
                accm[p][5] -= redot(b2, G3);   // dz MUx
                accm[p][3] += redot(u2, G4);   // dz MBy
                accm[p][7] -= redot(b2, G4);   // dz MUy
                accm[p][4] += kx * gaU;
                accm[p][6] += ky * gaU;
                accm[p][0] += kx * gaB;
                accm[p][2] += ky * gaB;
                gkx += MUx * gaU + MBx * gaB;
            }
            gq += gkx * (a.ts * ky);
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < HW; ++p) {
            const int item = tid + p * NT;
            if (HI % NT != 0 && item >= HI) continue;
            const int mb = item / (2 * NZ), rem = item % (2 * NZ);
            const int h = rem / NZ, z = rem % NZ;
            double2 *Y = bufX + mb * P::NLINE * NZ + z;
            if constexpr (BQ) {
                // lines: ux ze bx et uy uz by bz; half 0 the u side, half 1 the b side
                const int l0 = h ? 2 : 0;
                Y[l0 * NZ] = gX[p][0];
                Y[(l0 + 1) * NZ] = gX[p][1];
                Y[(l0 + 4) * NZ] = gX[p][2];
                Y[(l0 + 5) * NZ] = gX[p][3];
            } else {
                // lines: u0 u1 u2 b0 b1 b2
                const int l0 = h ? 3 : 0;
                Y[l0 * NZ] = gX[p][0];
                Y[(l0 + 1) * NZ] = gX[p][1];
                Y[(l0 + 2) * NZ] = gX[p][2];
            }
        }
        // ModelBQ: the dz uz constant's gradient of each mode, summed over z
        // in order
        if (BQ && tid < MB) {
            double2 sg = make_double2(0.0, 0.0);
            for (int z = 0; z < NZ; ++z) sg = cadd(sg, gdzz[tid * NZ + z]);
            sgdz[tid] = sg;
        }
        __syncthreads();

        // I: adjoint of the padded inverse transforms = truncated forward transforms
        fft_lines<NZ, NLX, false>(bufX, tw);

        // J: base fields in reverse; total dL/du_in
        for (int k = tid; k < MB * NZD; k += NT) {
            const int mb = k / NZD, m = k % NZD;
            const int i = sidx[mb];
            if (i < 0) continue;
            const double kx = skx[mb], ky = sky[mb], kz = skz[m];
            const bool zero = (ky == 0.0 && m == 0);
            const double ilap2 = zero ? 1.0 : 1.0 / (-ky * ky + -(kz * kz));
            if constexpr (!BQ) {
                const double2 *GL = bufX + mb * P::NLINE * NZ + D::pad(m);
                const double2 GB0 = GL[0], GB1 = GL[NZ], GB2 = GL[2 * NZ];
                const double2 GB3 = GL[3 * NZ], GB4 = GL[4 * NZ], GB5 = GL[5 * NZ];
                const double2 *U = uin + mb * 4 * NZD + m;
                const double2 ux = U[0], bx = U[2 * NZD];
                const double f = a.fac;
                double2 *Gu = sgu + mb * 4 * NZD + m;
                // uy = (ky kx ux + i kz ze) ilap2, uz = (kx kz ux - i ky ze) ilap2 (and b likewise)
                const double2 Gux = cscale(f, cadd(GB0, cadd(cscale(ky * kx * ilap2, GB1), cscale(kx * kz * ilap2, GB2))));
                const double2 Gze = cscale(f, cadd(imul(-kz * ilap2, GB1), imul(ky * ilap2, GB2)));
                const double2 Gbx = cscale(f, cadd(GB3, cadd(cscale(ky * kx * ilap2, GB4), cscale(kx * kz * ilap2, GB5))));
                const double2 Get = cscale(f, cadd(imul(-kz * ilap2, GB4), imul(ky * ilap2, GB5)));
                const double gkx = f * (redot(cscale(ky * ilap2, ux), GB1) + redot(cscale(kz * ilap2, ux), GB2) +
                                        redot(cscale(ky * ilap2, bx), GB4) + redot(cscale(kz * ilap2, bx), GB5));
                gq += gkx * (a.ts * ky);
                const size_t off = (size_t)i * 4 * NZD + m;
                a.gu[off] = cadd(Gu[0], Gux);
                a.gu[off + NZD] = cadd(Gu[NZD], Gze);
                a.gu[off + 2 * NZD] = cadd(Gu[2 * NZD], Gbx);
                a.gu[off + 3 * NZD] = cadd(Gu[3 * NZD], Get);
            } else {
                // bases ux ze bx et uy uz by bz: uy = (ky kx ux + i kz ze) ilap2,
                // uz = (kx kz ux - i ky ze) ilap2 (and b likewise), all times fac
                const double2 *GL = bufX + mb * P::NLINE * NZ + D::pad(m);
                const double2 GB0 = GL[0], GB1 = GL[NZ], GB2 = GL[2 * NZ], GB3 = GL[3 * NZ];
                const double2 GB4 = GL[4 * NZ], GB5 = GL[5 * NZ], GB6 = GL[6 * NZ], GB7 = GL[7 * NZ];
                const double2 *U = uin + mb * 4 * NZD + m;
                const double2 ux = U[0], bx = U[2 * NZD];
                const double f = a.fac;
                double2 *Gu = sgu + mb * 4 * NZD + m;
                double2 Gux = cscale(f, cadd(GB0, cadd(cscale(ky * kx * ilap2, GB4), cscale(kx * kz * ilap2, GB5))));
                const double2 Gze = cscale(f, cadd(GB1, cadd(imul(-kz * ilap2, GB4), imul(ky * ilap2, GB5))));
                const double2 Gbx = cscale(f, cadd(GB2, cadd(cscale(ky * kx * ilap2, GB6), cscale(kx * kz * ilap2, GB7))));
                const double2 Get = cscale(f, cadd(GB3, cadd(imul(-kz * ilap2, GB6), imul(ky * ilap2, GB7))));
                double gkx = f * (redot(cscale(ky * ilap2, ux), GB4) + redot(cscale(kz * ilap2, ux), GB5) +
                                  redot(cscale(ky * ilap2, bx), GB6) + redot(cscale(kz * ilap2, bx), GB7));
                if (zero) {
                    // dzc = i kx (fac ux at kz = 0) of a ky = 0 mode
                    const double2 Gd = sgdz[mb];
                    Gux = cadd(Gux, imul(-kx * f, Gd));
                    gkx += redot(imul(1.0, sdz[mb]), Gd);
                }
                gq += gkx * (a.ts * ky);
                const size_t off = (size_t)i * 4 * NZD + m;
                a.gu[off] = cadd(Gu[0], Gux);
                a.gu[off + NZD] = cadd(Gu[NZD], Gze);
                a.gu[off + 2 * NZD] = cadd(Gu[2 * NZD], Gbx);
                a.gu[off + 3 * NZD] = cadd(Gu[3 * NZD], Get);
            }
        }
        __syncthreads();
    }

    // partial row of this workgroup: z-profile gradients summed over mode slots
    // and halves in a fixed order, then the parameter gradients
    // (four profile lines at a time: 2 MB x 4 x NZ doubles fit the X lines)
    double *red = reinterpret_cast<double *>(bufX);
    static_assert(2 * MB * 4 * NZ * sizeof(double) <= sizeof(bufX), "reduction fits the X lines");
    double *row = a.gpart + (size_t)blockIdx.x * (8 * NZ + NGP);
#pragma unroll
    for (int c0 = 0; c0 < 8; c0 += 4) {
#pragma unroll
        for (int p = 0; p < HW; ++p) {
            const int item = tid + p * NT;
            if (HI % NT != 0 && item >= HI) continue;
            const int mbh = item / NZ, z = item % NZ;  // (mode slot, half) pair, z
#pragma unroll
            for (int c = 0; c < 4; ++c) red[(mbh * 4 + c) * NZ + z] = accm[p][c0 + c];
        }
        __syncthreads();
        for (int k = tid; k < 4 * NZ; k += NT) {
            double s = 0.0;
            for (int mbh = 0; mbh < 2 * MB; ++mbh) s += red[mbh * 4 * NZ + k];
            row[c0 * NZ + k] = s;
        }
        __syncthreads();
    }
    red[0 * NT + tid] = gq;
    red[1 * NT + tid] = gom;
    red[2 * NT + tid] = gnu;
    red[3 * NT + tid] = geta;
    red[4 * NT + tid] = gb0;
    __syncthreads();
    if (tid < NGP) {
        double s = 0.0;
        if (tid < 5)
            for (int k = 0; k < NT; ++k) s += red[tid * NT + k];
        row[8 * NZ + tid] = s;
    }
}

// ---------------------------------------------------------------------------
// mean fields in reverse, part A (one workgroup)
// ---------------------------------------------------------------------------
struct AdjMFArgs {
    double2 *gM;           // in: dL/dM_out; out: dL/dM_in without the z-profile path
    double2 *gnmp;         // in (s < 3): dL/dNM_s; out (s > 1): dL/dNM_{s-1}
    const double2 *M_in, *M_out;
    const double *kzv;
    const double2 *tw;
    double *gsums;         // out: dL/d(EMF / stress z-profiles) [4][NZ]
    double *gpar;          // out: parameter-gradient partials [NGP]
    Phys ph;
    int mode, s;
    double linC, nlc1, nlc2;
};

// MODEL 0: MHD_FullQlin (Bx, By, Ux, Uy); 1: MHD_BQlin (Bx, By: N_Bx = EMF x,
// N_By = -q Bx + EMF y, both with eta kz^2, MHD_BQlin.cpp:274-280)
template <int NZ, int MODEL>
__global__ __launch_bounds__(NT) void k_adj_mf_a(AdjMFArgs a)
{
    using D = Dealias<NZ>;
    constexpr int NZD = D::nz, HALF = D::half;
    constexpr int NM = MODEL == 0 ? 4 : 2;
    __shared__ double2 buf[4 * NZ];
    __shared__ double2 tw[NZ];
    __shared__ double skz[NZD];
    __shared__ double2 sGN[4 * NZD], sdir[4 * NZD];
    __shared__ double red[5][NT];
    const int tid = threadIdx.x;
    const Phys ph = a.ph;
    for (int k = tid; k < NZ; k += NT) tw[k] = a.tw[k];
    for (int k = tid; k < NZD; k += NT) skz[k] = a.kzv[k];
    __syncthreads();
    double gq = 0, gom = 0, gnu = 0, geta = 0, gb0 = 0;
    // update in reverse (RK3CN_Elin.cpp:109-112 / RK3CN.cpp:112-115)
    for (int k = tid; k < NM * NZD; k += NT) {
        const int f = k / NZD, m = k % NZD;
        const double kz2 = -(skz[m] * skz[m]);
        const double coef = f < 2 ? ph.eta : ph.nu;
        const double L = ph.quasilinear ? coef * kz2 : 0.0;
        const double2 g = a.gM[k];
        const double2 gin = (a.s < 3) ? a.gnmp[k] : make_double2(0.0, 0.0);
        double2 GN, dir;
        double gL;
        if (a.mode == MODE_ELIN) {
            const double2 mu = cscale(exp(a.linC * L), g);
            GN = cadd(cscale(a.nlc1, mu), gin);
            if (a.s > 1) a.gnmp[k] = cscale(a.nlc2, mu);
            dir = mu;
            gL = a.linC * redot(a.M_out[k], g);
        } else {
            const double d = 1.0 / (1.0 - a.linC * L);
            GN = cadd(cscale(a.nlc1 * d, g), gin);
            if (a.s > 1) a.gnmp[k] = cscale(a.nlc2 * d, g);
            dir = cscale((1 + a.linC * L) * d, g);
            gL = a.linC * d * (redot(a.M_in[k], g) + redot(a.M_out[k], g));
        }
        if (ph.quasilinear) {
            if (f < 2) geta += gL * kz2;
            else gnu += gL * kz2;
        }
        sGN[k] = ph.quasilinear ? GN : make_double2(0.0, 0.0);
        sdir[k] = dir;
    }
    __syncthreads();
    // mean-field RHS in reverse (MHD_FullQlin.cpp:411-423)
    for (int k = tid; k < NM * NZD; k += NT) {
        const int f = k / NZD, m = k % NZD;
        const double kz = skz[m];
        if constexpr (MODEL == 1) {
            double2 G = sdir[k];
            if (f == 0) {
                G = cadd(G, cscale(-ph.q, sGN[NZD + m]));
                gq += redot(cscale(-1.0, a.M_in[m]), sGN[NZD + m]);
            }
            a.gM[k] = G;
            continue;
        }
        const double2 N0 = sGN[m], N1 = sGN[NZD + m], N2 = sGN[2 * NZD + m], N3 = sGN[3 * NZD + m];
        double2 G = sdir[k];
        if (f == 0) G = cadd(G, cadd(cscale(-ph.q, N1), imul(-ph.B0z * kz, N2)));
        else if (f == 1) G = cadd(G, imul(-ph.B0z * kz, N3));
        else if (f == 2) G = cadd(G, cadd(imul(-ph.B0z * kz, N0), cscale(ph.q - 2.0 * ph.omega, N3)));
        else G = cadd(G, cadd(imul(-ph.B0z * kz, N1), cscale(2.0 * ph.omega, N2)));
        a.gM[k] = G;
        if (f == 0) {
            const double2 M0 = a.M_in[m], M1 = a.M_in[NZD + m], M2 = a.M_in[2 * NZD + m], M3 = a.M_in[3 * NZD + m];
            gq += redot(cscale(-1.0, M0), N1) + redot(M2, N3);
            gom += redot(cscale(2.0, M3), N2) + redot(cscale(-2.0, M2), N3);
            gb0 += redot(imul(kz, M2), N0) + redot(imul(kz, M3), N1) + redot(imul(kz, M0), N2) +
                   redot(imul(kz, M1), N3);
        }
    }
    // drive in reverse: Hermitian fill and x i kz of every line (the EMF,
    // MHD_FullQlin.cpp:326-336, and the stresses before their z-derivative),
    // then the forward transform's adjoint (an inverse transform) of the
    // real stress profiles
    for (int k = tid; k < NM * NZ; k += NT) {
        const int f = k / NZ, n = k % NZ;
        double2 v = make_double2(0.0, 0.0);
        if (n < HALF) {
            const double c = skz[n];
            const double2 d0 = sGN[f * NZD + n];
            v = imul(-c, d0);
            const int mm = 2 * HALF - 1 - n;  // slot filled with conj of this one
            if (n >= 1 && mm < NZD) {
                double2 d1 = sGN[f * NZD + mm];
                d1.y = -d1.y;
                v = cadd(v, imul(-c, d1));
            }
        }
        buf[k] = v;
    }
    __syncthreads();
    fft_lines<NZ, NM, true>(buf, tw);
    for (int k = tid; k < NM * NZ; k += NT) a.gsums[k] = buf[k].x;
    red[0][tid] = gq;
    red[1][tid] = gom;
    red[2][tid] = gnu;
    red[3][tid] = geta;
    red[4][tid] = gb0;
    __syncthreads();
    if (tid < NGP) {
        double s = 0.0;
        if (tid < 5)
            for (int k = 0; k < NT; ++k) s += red[tid][k];
        a.gpar[tid] = s;
    }
}

// ---------------------------------------------------------------------------
// mean fields in reverse, part B (one workgroup): the z-profiles
// mfz_l = Re(inverse(pad(c_l M))) with c = fac (l even) or i fac kz (l odd)
// (MHD_FullQlin.cpp:180-198), and the parameter-gradient sum of the substep
// ---------------------------------------------------------------------------
struct AdjMF2Args {
    double2 *gM;
    const double *rows;     // reduced partial rows [8 NZ + NGP]
    const double *gpar_mf;  // [NGP] from k_adj_mf_a
    double *gpar_total;     // [NGP], accumulated
    const double *kzv;
    const double2 *tw;
    double fac;
};

// MHD_BQlin: profiles By, dz By, dz^2 By = Re(inverse(pad(c_l By))) with
// c = fac, i fac kz, -fac kz^2 (MHD_BQlin.cpp:122-127); Bx has none
template <int NZ, int MODEL>
__global__ __launch_bounds__(NT) void k_adj_mf_b(AdjMF2Args a)
{
    using D = Dealias<NZ>;
    constexpr int NZD = D::nz;
    __shared__ double2 buf[8 * NZ];
    __shared__ double2 tw[NZ];
    const int tid = threadIdx.x;
    for (int k = tid; k < NZ; k += NT) tw[k] = a.tw[k];
    for (int k = tid; k < 8 * NZ; k += NT) buf[k] = make_double2(a.rows[k], 0.0);
    __syncthreads();
    fft_lines<NZ, 8, false>(buf, tw);
    if constexpr (MODEL == 1) {
        for (int m = tid; m < NZD; m += NT) {
            const double kz = a.kzv[m];
            const int j = D::pad(m);
            double2 G = cadd(cscale(a.fac, buf[j]), imul(-a.fac * kz, buf[NZ + j]));
            G = cadd(G, cscale(-a.fac * kz * kz, buf[2 * NZ + j]));
            a.gM[NZD + m] = cadd(a.gM[NZD + m], G);
        }
    } else {
        for (int k = tid; k < 4 * NZD; k += NT) {
            const int f = k / NZD, m = k % NZD;
            const double kz = a.kzv[m];
            const int j = D::pad(m);
            const double2 G = cadd(cscale(a.fac, buf[(2 * f) * NZ + j]), imul(-a.fac * kz, buf[(2 * f + 1) * NZ + j]));
            a.gM[k] = cadd(a.gM[k], G);
        }
    }
    if (tid < NGP) a.gpar_total[tid] += a.gpar_mf[tid] + a.rows[8 * NZ + tid];
}

}  // namespace qlk
