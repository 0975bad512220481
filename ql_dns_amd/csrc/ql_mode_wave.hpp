// ql_mode_wave.hpp -- the per-(kx,ky) substep kernel of MHD_FullQlin, MHD_BQlin
// and HD_fullU (through its velocity-only policy, ModelHD) with one wavefront
// per mode (Models/MHD_FullQlin.cpp:171-427, Models/MHD_BQlin.cpp:117-284,
// Models/HD_fullU.cpp:145-309 and the
// substep updates of Integrators/RK3CN_Elin.cpp:290-378,
// Integrators/RK3CN.cpp:92-193).
//
// Same stages as mode_body (ql_mode.hpp) -- bases, inverse transforms,
// products and EMF / stress sums, forward transforms, RHS and update -- but a
// wavefront takes a whole mode in a private LDS region, so the stages are
// ordered by wavefront-local fences instead of workgroup barriers and the
// wavefronts of a workgroup run independently; the other wavefronts of a SIMD
// hide a wavefront's memory and LDS latency.  The workgroup (NW wavefronts)
// shares the twiddle, wavenumber and profile tables and reduces its
// wavefronts' EMF / stress sums into one partial row at the end.  Every
// transform pass works in place in the wavefront's buffer: a pass loads all
// its inputs before its first store, and a fence separates the two.
#pragma once
#include "ql_mode.hpp"

namespace qlk {

// Orders the LDS accesses of one wavefront: its LDS instructions execute in
// issue order, so waiting for the outstanding ones and keeping the compiler
// from moving memory accesses across is enough for its lanes to see each
// other's stores.  Global memory counters are not waited for.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt unchanged
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int WAVE = 64;

// fft_pass_io for the 64 lanes of one wavefront
template <int NZ, int NL, bool INV, int R, int NS, class C, class Ld, class St>
__device__ __forceinline__ void wave_pass(int lane, const C *tw, const Ld &ld, const St &st)
{
    constexpr int M = NZ / R;
    constexpr int TOT = NL * M;
    constexpr int IT = (TOT + WAVE - 1) / WAVE;
    C v[IT][R];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int item = lane + it * WAVE;
        if (TOT % WAVE == 0 || item < TOT) {
            const int line = item / M, j = item % M;
#pragma unroll
            for (int r = 0; r < R; ++r) v[it][r] = ld(line, j + r * M);
        }
    }
    wave_sync();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        if (it > 0) __builtin_amdgcn_sched_barrier(0);  // one butterfly at a time (registers)
        const int item = lane + it * WAVE;
        if (TOT % WAVE == 0 || item < TOT) {
            const int line = item / M, j = item % M;
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    C w = tw[k * r * (NZ / (NS * R))];
                    if (INV) w.y = -w.y;
                    v[it][r] = cmul(v[it][r], w);
                }
            }
            if constexpr (R == 16) {
                dft16<INV>(v[it]);
            } else if constexpr (R == 8) {
                dft8<INV>(v[it]);
            } else if constexpr (R == 4) {
                dft4<INV>(v[it][0], v[it][1], v[it][2], v[it][3]);
            } else {
                const C a = v[it][0], b = v[it][1];
                v[it][0] = cadd(a, b);
                v[it][1] = csub(a, b);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) st(line, (j - k) * R + k + r * NS, v[it][r]);
        }
    }
    wave_sync();
}

// LDS layout of one wavefront's buffer (one mode): base table [base][NZDP],
// z lines [line][LS], transform intermediate [line][LP], transformed products
// [line][nz], all from slot 0 (each stage's inputs are dead once it stores);
// at the end the wavefront's EMF / stress sums [NACC][NZ] (FP64)
template <int NZ, class P, class C>
struct WaveLayout {
    static constexpr int NZD = Dealias<NZ>::nz;
    static constexpr int NZDP = bank_stride(NZD);
    static constexpr int LS = bank_stride(NZ);
    static constexpr int LP = PadIO<NZ, C>::LP;
    static constexpr int cmax(int x, int y) { return x > y ? x : y; }
    static constexpr int RED = (P::NACC * NZ * (int)sizeof(double) + (int)sizeof(C) - 1) / (int)sizeof(C);
    static constexpr int BUF = cmax(cmax(cmax(P::NB * NZDP, P::NLINE * cmax(LS, LP)), P::NOUT * NZD), RED);
};

template <int NZ, class P, class C>
struct WaveLineSrc {
    const C *sbase;  // [base][NZDP]
    const RealOf<C> *skz;
    __device__ C operator()(int line, int g) const
    {
        using D = Dealias<NZ>;
        if (!D::kept(g)) return Cx<C>::make(0, 0);
        const int m = D::slot(g);
        const C v = sbase[P::base_of(line) * WaveLayout<NZ, P, C>::NZDP + m];
        return P::dz_of(line) ? imul(skz[m], v) : v;
    }
};

template <int NZ, class C>
struct WaveCompactDst {
    C *ct;  // [line][nz]
    __device__ void operator()(int line, int j, C v) const
    {
        using D = Dealias<NZ>;
        if (D::kept(j)) ct[line * D::nz + D::slot(j)] = v;
    }
};

// MIXW (C = float2): mixed precision in one launch, each mode in FP64 when
// its flag is set and in FP32 otherwise (no companion launch)
template <int NZ, class P, class C, int MODE, int NW, bool SPL, bool MIXW = false>
__device__ __forceinline__ void mode_wave_body(const ModeArgs &a)
{
    typedef RealOf<C> R;
    using D = Dealias<NZ>;
    using L = WaveLayout<NZ, P, C>;
    using PL = Plan2<NZ>;
    constexpr int NZD = D::nz;
    constexpr int KI = (NZD + WAVE - 1) / WAVE;  // kz slots per lane
    constexpr int ZI = (NZ + WAVE - 1) / WAVE;   // z points per lane
    static_assert(KI <= 4 && ZI <= 4, "slots per lane");
    __shared__ C tw[NZ];
    __shared__ R skzr[NZD];
    __shared__ double skz[NZD];  // FP64 wavenumbers of B and F
    __shared__ R smfz[NZ * P::NMFZ];
    // MIXW: the buffer holds an FP64 mode, and FP64 copies of the tables
    constexpr int BUFC = MIXW ? ((int)(WaveLayout<NZ, P, double2>::BUF * sizeof(double2)) + (int)sizeof(C) - 1) / (int)sizeof(C)
                              : L::BUF;
    __shared__ C wbuf[NW][BUFC > L::BUF ? BUFC : L::BUF];
    __shared__ double2 tw64[MIXW ? NZ : 1];
    __shared__ double smfz64[MIXW ? NZ * P::NMFZ : 1];
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid / WAVE), lane = tid % WAVE;
    C *buf = wbuf[wave];

    const WgInfo wgi = a.wg[blockIdx.x];
    const int run = wgi.run, g0 = wgi.g0, g1 = wgi.g1, mode_off = wgi.mode_off;
    const RunParams rp = a.runs[run];
    const Phys ph = rp.ph;
    const SubTimes st = sub_times<MODE == MODE_CN>(a, run);
    const double linC = to_sgpr(st.linC), nlc1 = to_sgpr(st.nlc1), nlc2 = to_sgpr(st.nlc2);
    const double qt = to_sgpr(rn_mul(ph.q, st.ts)), qdtl = to_sgpr(rn_mul(ph.q, st.dtl));
    const double qt_old = to_sgpr(rn_mul(ph.q, st.tso)), qdtl_old = to_sgpr(rn_mul(ph.q, st.dtlo));
    const double remap_qt = to_sgpr(rn_mul(ph.q, st.tnew));
    const bool do_remap = a.do_remap && rp.remap;

    // shared tables, every load before the first LDS store
    {
        const C *twg = reinterpret_cast<const C *>(sizeof(C) == 8 ? a.tw32 : (const void *)a.tw);
        const double *mfz = a.mfz + (size_t)run * 8 * NZ;
        constexpr int NTH = NW * WAVE;
        constexpr int TWN = (NZ + NTH - 1) / NTH, MFN = (NZ * P::NMFZ + NTH - 1) / NTH;
        static_assert(NZD <= NTH, "one wavenumber per thread");
        C twv[TWN];
        double mfv[MFN];
#pragma unroll
        for (int w = 0; w < TWN; ++w) twv[w] = (tid + w * NTH < NZ) ? twg[tid + w * NTH] : Cx<C>::make(0, 0);
        const double kzt = tid < NZD ? a.kzv[tid] : 0.0;
#pragma unroll
        for (int w = 0; w < MFN; ++w) mfv[w] = (tid + w * NTH < NZ * P::NMFZ) ? mfz[tid + w * NTH] : 0.0;
#pragma unroll
        for (int w = 0; w < TWN; ++w)
            if (tid + w * NTH < NZ) tw[tid + w * NTH] = twv[w];
        if (tid < NZD) {
            skzr[tid] = (R)kzt;
            skz[tid] = kzt;
        }
#pragma unroll
        for (int w = 0; w < MFN; ++w)
            if (tid + w * NTH < NZ * P::NMFZ) smfz[tid + w * NTH] = (R)mfv[w];
        if constexpr (MIXW) {
#pragma unroll
            for (int w = 0; w < TWN; ++w)
                if (tid + w * NTH < NZ) tw64[tid + w * NTH] = a.tw[tid + w * NTH];
#pragma unroll
            for (int w = 0; w < MFN; ++w)
                if (tid + w * NTH < NZ * P::NMFZ) smfz64[tid + w * NTH] = mfv[w];
        }
    }
    double acc[ZI][P::NACC];
#pragma unroll
    for (int k = 0; k < ZI; ++k)
#pragma unroll
        for (int c = 0; c < P::NACC; ++c) acc[k][c] = 0.0;
    __syncthreads();

    // the workgroup's modes (run-local l in [g0, g1), one per pass), dealt
    // round robin to its wavefronts; with the mixed-precision split
    // (SPLIT_LO) the FP64 companion takes the modes whose amplitude reaches
    // the threshold
    unsigned long long skip = 0ull;  // SPL: which of the chunk's 64 modes the companion takes
    int cbase = g0 + wave;
    for (int l = g0 + wave; l < g1; l += NW) {
        const int i = mode_off + l;
        if constexpr (SPL && !MIXW) {
            if (a.split == SPLIT_LO) {
                // the flags of this wavefront's next 64 modes, one per lane in
                // one round trip, then read from a ballot
                if ((l - g0 - wave) % (64 * NW) == 0) {
                    const int lq = l + lane * NW;
                    skip = __ballot(lq < g1 && a.hi[mode_off + lq]);
                    cbase = l;
                }
                if ((skip >> ((l - cbase) / NW)) & 1ull) continue;
            }
        }
        // one mode in precision CC (the FP64 or the FP32 transforms), its
        // line buffer and tables
        auto pass = [&](auto ctag, auto *bufc, const auto *twc, const auto *skzrc, const auto *smfzc) {
            using CC = decltype(ctag);
            typedef RealOf<CC> RR;
            using LL = WaveLayout<NZ, P, CC>;
            const double kx0 = to_sgpr(a.kx0[i]), ky = to_sgpr(a.ky[i]);
            const double kx = to_sgpr(rn_add(kx0, rn_mul(qt, ky)));  // kx(t), MHD_FullQlin.cpp:727
            const size_t row = (size_t)i * 4 * NZD;

            // B: 1/lap2, 1/lapF and the base spectral fields (MHD_FullQlin.cpp:220-221,
            // 722-738), fac1D folded in
    #pragma unroll
            for (int k = 0; k < KI; ++k) {
                if (k > 0) __builtin_amdgcn_sched_barrier(0);  // one slot at a time (registers)
                const int m = lane + k * WAVE;
                if (KI * WAVE == NZD || m < NZD) {
                    const double2 u0 = a.u[row + m], u1 = a.u[row + NZD + m];
                    double2 u2 = make_double2(0.0, 0.0), u3 = u2;
                    if constexpr (P::NPAIR == 2) {
                        u2 = a.u[row + 2 * NZD + m];
                        u3 = a.u[row + 3 * NZD + m];
                    }
                    const double kzk = skz[m], kz2 = -(kzk * kzk);
                    const bool zero = (ky == 0.0 && m == 0);
                    const double lap2 = zero ? 0.0 : (-ky * ky + kz2);
                    const double ilap2 = zero ? 1.0 : 1.0 / lap2;
                    P::bases(u0, u1, u2, u3, kx, ky, kzk, ilap2, a.fac, bufc + m, LL::NZDP);
                }
            }
            wave_sync();

            // CC: inverse transforms of the padded lines
            {
                const PadIO<NZ, CC> pad{bufc};
                wave_pass<NZ, P::NLINE, true, PL::R1, 1>(lane, twc, WaveLineSrc<NZ, P, CC>{bufc, skzrc}, pad);
                wave_pass<NZ, P::NLINE, true, PL::R2, PL::R1>(lane, twc, pad, BufIO<LL::LS, CC>{bufc});
            }

            // D: products on the z grid, in place, and the EMF / stress sums
            {
                const double ftfac = (ky == 0.0 ? 1.0 : 2.0) * rp.rsf;  // MHD_FullQlin.cpp:314-318
                // ModelBQ's dz uz constant of a ky = 0 mode: i kx fac1D ux at kz = 0
                const CC dzc = from_d<CC>(ky == 0.0 ? imul(kx, cscale(a.fac, a.u[row])) : make_double2(0.0, 0.0));
    #pragma unroll
                for (int k = 0; k < ZI; ++k) {
                    if (k > 0) __builtin_amdgcn_sched_barrier(0);  // one z point at a time (registers)
                    const int zz = lane + k * WAVE;
                    if (ZI * WAVE == NZ || zz < NZ) {
                        CC *X = bufc + zz;
                        RR mfl[P::NMFZ];
    #pragma unroll
                        for (int c = 0; c < P::NMFZ; ++c) mfl[c] = smfzc[c * NZ + zz];
                        CC prod[2][3];
                        double racc[2][2];
                        P::products(0, X, LL::LS, mfl, (RR)kx, (RR)ky, dzc, prod[0], racc[0]);
                        P::products(1, X, LL::LS, mfl, (RR)kx, (RR)ky, dzc, prod[1], racc[1]);
    #pragma unroll
                        for (int part = 0; part < 2; ++part) {
                            acc[k][P::acc_slot(part, 0)] += ftfac * racc[part][0];
                            if (P::NACC == 4) acc[k][P::acc_slot(part, 1)] += ftfac * racc[part][1];
                        }
    #pragma unroll
                        for (int part = 0; part < 2; ++part)
    #pragma unroll
                            for (int c = 0; c < 3; ++c)
                                if (c < P::nprod(part)) X[P::out_line(part, c) * LL::LS] = prod[part][c];
                    }
                }
            }
            wave_sync();

            // E: forward transforms of the product lines, dealiased band kept
            {
                const PadIO<NZ, CC> pad{bufc};
                wave_pass<NZ, P::NOUT, false, PL::R1, 1>(lane, twc, BufIO<LL::LS, CC>{bufc}, pad);
                wave_pass<NZ, P::NOUT, false, PL::R2, PL::R1>(lane, twc, pad, WaveCompactDst<NZ, CC>{bufc});
            }

            // F: RHS, dissipation operator at the substep end, update
            bool hit = false;
            if (do_remap) hit = remap_hit(remap_qt, rp.remap_thr, kx0, ky);
    #pragma unroll
            for (int k = 0; k < KI; ++k) {
                // one slot at a time: the scheduler would otherwise hoist the
                // loads of every slot of the lane and run out of registers
                if (k > 0) __builtin_amdgcn_sched_barrier(0);
                const int m = lane + k * WAVE;
                if (!(KI * WAVE == NZD || m < NZD)) continue;
                const size_t off = row + m;
                const double kzk = skz[m], kz2 = -(kzk * kzk);
                // 1/lapF as in B (recomputed rather than held through the transforms)
                const double lap2 = (ky == 0.0 && m == 0) ? 0.0 : (-ky * ky + kz2);
                const double ilapF = (l == 0 && m == 0) ? 1.0 : 1.0 / (-kx * kx + lap2);
                const double2 U0 = a.u[off], U1 = a.u[off + NZD];
                double2 U2 = make_double2(0.0, 0.0), U3 = U2;
                if constexpr (P::NPAIR == 2) {
                    U2 = a.u[off + 2 * NZD];
                    U3 = a.u[off + 3 * NZD];
                }
                double2 np[4];
    #pragma unroll
                for (int f = 0; f < 2 * P::NPAIR; ++f) np[f] = a.n_prev ? a.n_prev[off + f * NZD] : make_double2(0.0, 0.0);
                const CC *T = bufc + m;
    #pragma unroll
                for (int pr = 0; pr < P::NPAIR; ++pr) {
                    double2 N0, N1;
                    P::rhs_pair(pr, T, NZD, U0, U1, U2, U3, kx, ky, kzk, ilapF, ph, N0, N1);
                    const double Ln = P::lin(pr, kx + qdtl * ky, ky, kz2, ph);
                    const size_t off0 = off + (2 * pr) * NZD, off1 = off0 + NZD;
                    if constexpr (MODE == MODE_RHS) {
                        if (a.n_out) {
                            a.n_out[off0] = N0;
                            a.n_out[off1] = N1;
                        }
                        if (a.lin_out) a.lin_out[off0] = a.lin_out[off1] = Ln;
                        continue;
                    }
                    const double2 u0 = pr ? U2 : U0, u1 = pr ? U3 : U1;
                    const double2 p0 = np[2 * pr], p1 = np[2 * pr + 1];
                    double2 un0, un1;
                    if constexpr (MODE == MODE_ELIN) {
                        // RK3CN_Elin.cpp:303-305, :335-337, :368-370
                        const double e = exp(linC * Ln);
                        double2 s0 = cscale(nlc1, N0), s1 = cscale(nlc1, N1);
                        if (a.n_prev) {
                            s0 = cadd(s0, cscale(nlc2, p0));
                            s1 = cadd(s1, cscale(nlc2, p1));
                        }
                        un0 = cscale(e, cadd(s0, u0));
                        un1 = cscale(e, cadd(s1, u1));
                    } else {
                        // RK3CN.cpp:105-107, :141-143, :177-179; the "old" operator is
                        // the one of the previous substep (or linearOPs_Init at t0)
                        double Lo;
                        const double kxo = rn_add(kx0, rn_mul(qt_old, ky));
                        if (st.oinit) {
                            Lo = P::lin_init(pr, kxo, lap2, ph);
                        } else {
                            Lo = P::lin(pr, kxo + qdtl_old * ky, ky, kz2, ph);
                        }
                        const double denom = 1.0 / (1.0 - linC * Ln);
                        const double c1 = nlc1 * denom, co = (1 + linC * Lo) * denom;
                        double2 s0 = cscale(c1, N0), s1 = cscale(c1, N1);
                        if (a.n_prev) {
                            const double c2 = nlc2 * denom;
                            s0 = cadd(s0, cscale(c2, p0));
                            s1 = cadd(s1, cscale(c2, p1));
                        }
                        un0 = cadd(s0, cscale(co, u0));
                        un1 = cadd(s1, cscale(co, u1));
                    }
                    if (hit) un0 = un1 = make_double2(0.0, 0.0);
                    a.u[off0] = un0;
                    a.u[off1] = un1;
                    if (a.n_out) {
                        a.n_out[off0] = N0;
                        a.n_out[off1] = N1;
                    }
                }
            }
            if (hit && lane == 0) a.kx0[i] = rn_sub(kx0, rp.remap_shift);
        };
        if constexpr (MIXW) {
            // mixed precision per mode: the flagged modes in FP64 in this
            // same launch, the others in FP32, both in the wavefront's buffer
            if (a.split == SPLIT_LO && a.hi[i])
                pass(double2{}, reinterpret_cast<double2 *>(buf), tw64, skz, smfz64);
            else
                pass(C{}, buf, tw, skzr, smfz);
        } else {
            pass(C{}, buf, tw, skzr, smfz);
        }
        // the next mode's B overwrites the transformed products F read
        wave_sync();
    }

    // fixed-order reduction of the wavefronts' sums: one partial [NACC][NZ]
    // per workgroup
    __syncthreads();
    double *red = reinterpret_cast<double *>(&wbuf[0][0]);
    static_assert(NW * P::NACC * NZ * sizeof(double) <= sizeof(wbuf), "reduction fits the buffers");
#pragma unroll
    for (int k = 0; k < ZI; ++k) {
        const int zz = lane + k * WAVE;
        if (ZI * WAVE == NZ || zz < NZ)
#pragma unroll
            for (int c = 0; c < P::NACC; ++c) red[(wave * P::NACC + c) * NZ + zz] = acc[k][c];
    }
    __syncthreads();
    for (int k = tid; k < P::NACC * NZ; k += NW * WAVE) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) s += red[w * P::NACC * NZ + k];
        a.partials[(size_t)blockIdx.x * P::NACC * NZ + k] = s;
    }
}

// MHD_FullQlin (ModelFull), MHD_BQlin (ModelBQ) and HD_fullU (ModelHD); workgroups of
// mode_waves(NZ) wavefronts: 8 at NZ <= 128 (one workgroup per CU at NZ = 128
// in FP64: 116 KB of LDS; measured at 64x128x128 in an earlier build: 8
// wavefronts and 252 workgroups 31.1 us, 4 and 504 31.9 us, 2 and 1008 31.6 us and twice the
// partial rows for the column reduction), 4 at NZ = 256 (26 KB per wavefront
// in FP64, one wavefront per SIMD with the whole register file)
// (the velocity-only policy, with 3 lines per transform, keeps 8 at NZ = 256)
template <class P>
constexpr int mode_waves(int NZ) { return NZ >= 256 && P::NPAIR == 2 ? 4 : 8; }
template <class P>
constexpr int wave_occ(int NZ) { return NZ >= 256 && P::NPAIR == 2 ? 1 : 2; }
template <int NZ, class P, int MODE>
__global__ __launch_bounds__(mode_waves<P>(NZ) * WAVE) __attribute__((amdgpu_waves_per_eu(wave_occ<P>(NZ))))
void k_wave64(ModeArgs a)
{
    mode_wave_body<NZ, P, double2, MODE, mode_waves<P>(NZ), false>(a);
}
template <int NZ, class P, int MODE>
__global__ __launch_bounds__(mode_waves<P>(NZ) * WAVE) __attribute__((amdgpu_waves_per_eu(wave_occ<P>(NZ))))
void k_wave32(ModeArgs a)
{
    mode_wave_body<NZ, P, float2, MODE, mode_waves<P>(NZ), true>(a);
}
template <int NZ, class P, int MODE>
__global__ __launch_bounds__(mode_waves<P>(NZ) * WAVE) __attribute__((amdgpu_waves_per_eu(wave_occ<P>(NZ))))
void k_wave_mix(ModeArgs a)
{
    mode_wave_body<NZ, P, float2, MODE, mode_waves<P>(NZ), true, true>(a);
}

}  // namespace qlk
