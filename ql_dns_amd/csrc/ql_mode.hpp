// ql_mode.hpp -- the fused per-(kx,ky) kernel of one RK3 substep (FP64).
//
// One kernel skeleton, two model policies:
//   ModelFull  MHD_FullQlin (Models/MHD_FullQlin.cpp:171-427): fluctuations
//              coupled to Bx, By, Ux, Uy(z); EMF and both stresses reduced.
//   ModelBQ    MHD_BQlin (Models/MHD_BQlin.cpp:117-284): fluctuations see
//              By(z), dBy/dz, d2By/dz2 only; the two EMF components reduced.
//
// A 256-thread workgroup takes a contiguous range of one run's groups of MB
// modes.  Per group:
//   A  the group's state rows -> LDS (prefetched into registers during the
//      previous group), the next group's rows -> registers;
//   B  per (mode, kz slot): 1/lap2, 1/lapF and the base spectral fields
//      (fac1D folded in; y, z components rebuilt from (ux, zeta), (bx, eta));
//   C  inverse transforms of the padded lines, the first pass reading the
//      base table directly (lines = bases or i kz * bases);
//   D  mean-field x fluctuation products on the z grid, x and y derivatives
//      being i kx and i ky times the z profile; EMF / stress z-profiles summed
//      in registers over the workgroup's modes;
//   E  forward transforms of the product lines;
//   F  RHS assembly, dissipation operator at the substep end, substep update
//      per (mode, kz slot, field pair), remap predicate on the new state.
#pragma once
#include "ql_kernels.hpp"

// register prefetch of the next group's state rows (1 = on)
#ifndef QL_PREFETCH_ROWS
#define QL_PREFETCH_ROWS 1
#endif
// occupancy requests (waves per SIMD, which bound the VGPR budget) for
// 256-point passes and for 128-point passes (half the modes per pass)
#ifndef QL_WAVES64
#define QL_WAVES64 2
#endif
#ifndef QL_WAVES32
#define QL_WAVES32 2
#endif
#ifndef QL_WAVES64_HALF
#define QL_WAVES64_HALF 2
#endif
#ifndef QL_WAVES32_HALF
#define QL_WAVES32_HALF 2
#endif

namespace qlk {

// A wave-uniform double moved to scalar registers (both halves read from the
// first active lane).  FP64 arithmetic runs on the vector ALU only, so a
// uniform value computed in the kernel would otherwise hold a VGPR pair for
// the kernel's lifetime.
__device__ __forceinline__ double to_sgpr(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Re(a conj b) formed in FP64 (exact for FP32 inputs)
template <class C>
__device__ __forceinline__ double redot64(C a, C b)
{
    return (double)a.x * (double)b.x + (double)a.y * (double)b.y;
}

// y and z components of a divergence-free field from (fx, zeta-like):
// ((ky kx) f + i kz z) ilap2 and ((kx kz) f - i ky z) ilap2
// (MHD_FullQlin.cpp:220-221 with kx, ky, kz stored as i*k)
__device__ __forceinline__ void yz_from(double2 f, double2 zt, double kx, double ky, double kz, double ilap2,
                                        double2 &fy, double2 &fz)
{
    const double sy = ky * kx, sz = kx * kz;
    fy = make_double2((sy * f.x - kz * zt.y) * ilap2, (sy * f.y + kz * zt.x) * ilap2);
    fz = make_double2((sz * f.x + ky * zt.y) * ilap2, (sz * f.y - ky * zt.x) * ilap2);
}

struct ModelFull {
    static constexpr int NB = 6;     // bases: ux uy uz bx by bz
    static constexpr int NLINE = 10; // + dz ux, dz bx, dz uy, dz by
    static constexpr int NOUT = 6;   // forward lines (3 momentum, 3 induction)
    static constexpr int NMFZ = 8;   // MBx dzMBx MBy dzMBy MUx dzMUx MUy dzMUy
    static constexpr int NACC = 4;   // EMF x, EMF y, stress x, stress y
    __device__ static int base_of(int l) { return l < 6 ? l : (((l & 1) ? 3 : 0) + (l >= 8 ? 1 : 0)); }
    __device__ static bool dz_of(int l) { return l >= 6; }
    template <class C>
    __device__ static void bases(double2 ux, double2 ze, double2 bx, double2 et, double kx, double ky, double kz,
                                 double ilap2, double fac, C *out, int stride)
    {
        double2 uy, uz, by, bz;
        yz_from(ux, ze, kx, ky, kz, ilap2, uy, uz);
        yz_from(bx, et, kx, ky, kz, ilap2, by, bz);
        out[0] = from_d<C>(cscale(fac, ux));
        out[stride] = from_d<C>(cscale(fac, uy));
        out[2 * stride] = from_d<C>(cscale(fac, uz));
        out[3 * stride] = from_d<C>(cscale(fac, bx));
        out[4 * stride] = from_d<C>(cscale(fac, by));
        out[5 * stride] = from_d<C>(cscale(fac, bz));
    }
    __device__ static int nprod(int) { return 3; }
    __device__ static int out_line(int part, int c) { return part * 3 + c; }
    __device__ static int acc_slot(int part, int c) { return part * 2 + c; }
    // MHD_FullQlin.cpp:271-300 (products) and :322-358 (EMF, stresses)
    template <class C>
    __device__ static void products(int part, const C *X, int ld, const RealOf<C> *mf, RealOf<C> kx, RealOf<C> ky,
                                    C *prod, double *racc)
    {
        typedef RealOf<C> R;
        const C u0 = X[0], u1 = X[ld], u2 = X[2 * ld];
        const C b0 = X[3 * ld], b1 = X[4 * ld], b2 = X[5 * ld];
        const R MBx = mf[0], dzMBx = mf[1], MBy = mf[2], dzMBy = mf[3];
        const R MUx = mf[4], dzMUx = mf[5], MUy = mf[6], dzMUy = mf[7];
        const R aU = kx * MUx + ky * MUy, aB = kx * MBx + ky * MBy;
        if (part == 0) {
            // U.grad(u) + uz dzU - (B.grad(b) + bz dzB)
            prod[0] = csub(cadd(imul(aU, u0), cscale(dzMUx, u2)), cadd(imul(aB, b0), cscale(dzMBx, b2)));
            prod[1] = csub(cadd(imul(aU, u1), cscale(dzMUy, u2)), cadd(imul(aB, b1), cscale(dzMBy, b2)));
            prod[2] = csub(imul(aU, u2), imul(aB, b2));
            racc[0] = redot64(b2, u0) - redot64(u2, b0);
            racc[1] = redot64(b2, u1) - redot64(u2, b1);
        } else {
            // U.grad(b) + uz dzB - (B.grad(u) + bz dzU)
            prod[0] = csub(cadd(imul(aU, b0), cscale(dzMBx, u2)), cadd(imul(aB, u0), cscale(dzMUx, b2)));
            prod[1] = csub(cadd(imul(aU, b1), cscale(dzMBy, u2)), cadd(imul(aB, u1), cscale(dzMUy, b2)));
            prod[2] = csub(imul(aU, b2), imul(aB, u2));
            const C dzux = X[6 * ld], dzbx = X[7 * ld], dzuy = X[8 * ld], dzby = X[9 * ld];
            // x and y derivatives i kx f, i ky f formed in FP64
            const double kxd = kx, kyd = ky;
            const double2 U0 = to_d(u0), U1 = to_d(u1), B0 = to_d(b0), B1 = to_d(b1);
            const double2 u3 = imul(kxd, U0), u6 = imul(kyd, U0), b3 = imul(kxd, B0), b6 = imul(kyd, B0);
            const double2 u4 = imul(kxd, U1), u7 = imul(kyd, U1), b4 = imul(kxd, B1), b7 = imul(kyd, B1);
            const double2 U2 = to_d(u2), B2 = to_d(b2);
            racc[0] = -(redot(U0, u3) + redot(U1, u6) + redot64(u2, dzux)) + (redot(B0, b3) + redot(B1, b6) + redot64(b2, dzbx));
            racc[1] = -(redot(U0, u4) + redot(U1, u7) + redot64(u2, dzuy)) + (redot(B0, b4) + redot(B1, b7) + redot64(b2, dzby));
            (void)U2;
            (void)B2;
        }
    }
    // MHD_FullQlin.cpp:283-285 (pair 0: ux, zeta) and :303-305 (pair 1: bx, eta)
    template <class C>
    __device__ static void rhs_pair(int pr, const C *T, int ld, double2 ux, double2 ze, double2 bx, double2 et,
                                    double kx, double ky, double kz, double ilapF, const Phys &a, double2 &N0,
                                    double2 &N1)
    {
        const double q = a.q, Om = a.omega, B0z = a.B0z;
        if (pr == 0) {
            const double2 T1 = to_d(T[0]), T2 = to_d(T[ld]), T3 = to_d(T[2 * ld]);
            double2 s = cscale(2.0 * kx * ky * q, ux);
            s = cadd(s, imul(2.0 * Om * kz, ze));
            s = csub(s, cscale(kx * kx, T1));
            s = csub(s, cscale(kx * ky, T2));
            s = csub(s, cscale(kx * kz, T3));
            N0 = csub(cadd(cscale(ilapF, s), imul(B0z * kz, bx)), T1);
            N1 = cadd(imul((q - 2.0 * Om) * kz, ux), imul(B0z * kz, et));
            N1 = csub(N1, csub(imul(kz, T2), imul(ky, T3)));
        } else {
            const double2 T2 = to_d(T[4 * ld]), T3 = to_d(T[5 * ld]);
            N0 = csub(imul(B0z * kz, ux), to_d(T[3 * ld]));
            N1 = cadd(imul(-q * kz, bx), imul(B0z * kz, ze));
            N1 = csub(N1, csub(imul(kz, T2), imul(ky, T3)));
        }
    }
    // -coef |k^2|^(order/2) at the substep end (:371-376) and at t0 (:432-440)
    __device__ static double lin(int pr, double kxn, double ky, double kz2, const Phys &a)
    {
        return (pr == 0 ? a.nu : a.eta) * neg_abs_pow((-kxn * kxn - ky * ky) + kz2, a.vo2);
    }
    __device__ static double lin_init(int pr, double kxo, double lap2, const Phys &a)
    {
        return (pr == 0 ? a.nu : a.eta) * neg_abs_pow(-kxo * kxo + lap2, a.vo2);
    }
};

struct ModelBQ {
    static constexpr int NB = 8;    // bases: ux zeta bx eta uy uz by bz
    static constexpr int NLINE = 9; // + dz uz
    static constexpr int NOUT = 5;  // u (two parts), b, zeta, eta
    static constexpr int NMFZ = 3;  // By dzBy dzdzBy
    static constexpr int NACC = 2;  // EMF x, EMF y
    __device__ static int base_of(int l) { return l < 8 ? l : 5; }
    __device__ static bool dz_of(int l) { return l == 8; }
    template <class C>
    __device__ static void bases(double2 ux, double2 ze, double2 bx, double2 et, double kx, double ky, double kz,
                                 double ilap2, double fac, C *out, int stride)
    {
        double2 uy, uz, by, bz;
        yz_from(ux, ze, kx, ky, kz, ilap2, uy, uz);  // MHD_BQlin.cpp:143-146
        yz_from(bx, et, kx, ky, kz, ilap2, by, bz);
        out[0] = from_d<C>(cscale(fac, ux));
        out[stride] = from_d<C>(cscale(fac, ze));
        out[2 * stride] = from_d<C>(cscale(fac, bx));
        out[3 * stride] = from_d<C>(cscale(fac, et));
        out[4 * stride] = from_d<C>(cscale(fac, uy));
        out[5 * stride] = from_d<C>(cscale(fac, uz));
        out[6 * stride] = from_d<C>(cscale(fac, by));
        out[7 * stride] = from_d<C>(cscale(fac, bz));
    }
    __device__ static int nprod(int part) { return part == 0 ? 3 : 2; }
    __device__ static int out_line(int part, int c) { return part * 3 + c; }
    __device__ static int acc_slot(int part, int) { return part; }
    // MHD_BQlin.cpp:153-201 (products) and :219-240 (EMF)
    template <class C>
    __device__ static void products(int part, const C *X, int ld, const RealOf<C> *mf, RealOf<C> kx, RealOf<C> ky,
                                    C *prod, double *racc)
    {
        typedef RealOf<C> R;
        const C ux = X[0], ze = X[ld], bx = X[2 * ld], et = X[3 * ld];
        const C uy = X[4 * ld], uz = X[5 * ld], by = X[6 * ld], bz = X[7 * ld], dzuz = X[8 * ld];
        const R By = mf[0], dzBy = mf[1], dzdzBy = mf[2];
        if (part == 0) {
            prod[0] = cscale(By, imul(ky, bx));                 // u, first part
            prod[1] = cscale(dzBy, cscale((R)2 * kx * ky, bz));  // u, part under 1/lapF
            prod[2] = cscale(By, imul(ky, ux));                 // b
            racc[0] = redot64(bz, ux) - redot64(uz, bx);
        } else {
            // zeta: By (i ky eta) + dzBy (-i kx bx) + dzdzBy bz
            prod[0] = cadd(cadd(cscale(By, imul(ky, et)), cscale(dzBy, imul(-kx, bx))), cscale(dzdzBy, bz));
            // eta: By (i ky zeta) + dzBy (-2 dz uz - i kx ux) + dzdzBy (-uz)
            const C t2 = csub(cscale((R)-2, dzuz), imul(kx, ux));
            prod[1] = cadd(cadd(cscale(By, imul(ky, ze)), cscale(dzBy, t2)), cscale(-dzdzBy, uz));
            prod[2] = Cx<C>::make(0, 0);
            racc[0] = redot64(bz, uy) - redot64(uz, by);
        }
        racc[1] = 0;
    }
    // MHD_BQlin.cpp:160-162, :177-178 (pair 0), :185, :200-201 (pair 1)
    template <class C>
    __device__ static void rhs_pair(int pr, const C *T, int ld, double2 ux, double2 ze, double2 bx, double2,
                                    double kx, double ky, double kz, double ilapF, const Phys &a, double2 &N0,
                                    double2 &N1)
    {
        const double q = a.q, Om = a.omega;
        if (pr == 0) {
            const double2 s = cadd(cscale(2.0 * kx * ky * q, ux), imul(2.0 * Om * kz, ze));
            N0 = cadd(to_d(T[0]), cscale(ilapF, cadd(s, to_d(T[ld]))));
            N1 = cadd(to_d(T[3 * ld]), imul((q - 2.0 * Om) * kz, ux));
        } else {
            N0 = to_d(T[2 * ld]);
            N1 = csub(to_d(T[4 * ld]), imul(q * kz, bx));
        }
    }
    // nu (-k^2) for u, zeta; (eta/nu) times that for b, eta (:248-252, :289-295)
    __device__ static double lin(int pr, double kxn, double ky, double kz2, const Phys &a)
    {
        const double L0 = a.nu * ((-kxn * kxn - ky * ky) + kz2);
        return pr == 0 ? L0 : (a.eta / a.nu) * L0;
    }
    __device__ static double lin_init(int pr, double kxo, double lap2, const Phys &a)
    {
        const double L0 = a.nu * (-kxo * kxo + lap2);
        return pr == 0 ? L0 : (a.eta / a.nu) * L0;
    }
};

// inputs of the first inverse-transform pass: padded line `line` at grid
// index g, from the base table (zero in the dealiased band)
template <int NZ, class P, class C>
struct LineSrc {
    const C *sbase;
    const RealOf<C> *skz;
    __device__ C operator()(int line, int g) const
    {
        using D = Dealias<NZ>;
        if (!D::kept(g)) return Cx<C>::make(0, 0);
        const int m = D::slot(g);
        const int mb = line / P::NLINE, l = line - mb * P::NLINE;
        const C v = sbase[(mb * P::NB + P::base_of(l)) * D::nz + m];
        return P::dz_of(l) ? imul(skz[m], v) : v;
    }
};

template <int NZ, int MB, class P, class C>
__device__ __forceinline__ void mode_body(const ModeArgs &a)
{
    typedef RealOf<C> R;
    using D = Dealias<NZ>;
    constexpr int NZD = D::nz;
    constexpr int NL = P::NLINE * MB;
    constexpr int PI_ = MB * 2 * NZ;  // pointwise items: (mode slot, part, z)
    constexpr int PW = (PI_ + NT - 1) / NT;
    constexpr int NU = MB * 4 * NZD;  // state values of a group
    constexpr int UW = (NU + NT - 1) / NT;
    constexpr int AI = MB * 2 * NZD;  // update items: (mode slot, field pair, kz slot)
    constexpr int AW = (AI + NT - 1) / NT;
    __shared__ C buf[NL * NZ];
    __shared__ double2 uin[NU];
    __shared__ C sbase[MB * P::NB * NZD];
    __shared__ C tw[NZ];
    __shared__ double skz[NZD];
    __shared__ R skzr[NZD];
    __shared__ double silF[MB * NZD];
    __shared__ double skx0[MB], sky[MB], skx[MB];
    __shared__ int sidx[MB];   // mode index in the group arrays (-1: empty slot)
    __shared__ int sloc[MB];   // mode index within its run
    // mean-field profiles as [z][profile]: in LDS up to NZ = 128, in
    // registers (the thread's fixed z points) at NZ = 256 where LDS is short
    constexpr bool MF_LDS = (NZ <= 128) || sizeof(C) == 8;
    __shared__ R smfz[MF_LDS ? NZ * P::NMFZ : 1];
    R mfr[MF_LDS ? 1 : PW][P::NMFZ];
    const int tid = threadIdx.x;

    // this workgroup's run, its mode groups [g0, g1), the substep times
    const int run = a.wg_run[blockIdx.x];
    const RunParams rp = a.runs[run];
    const Phys ph = rp.ph;
    const int g0 = a.wg_g0[blockIdx.x], g1 = a.wg_g1[blockIdx.x];
    // q t products: wave-uniform, kept in scalar registers
    const SubTimes st = sub_times(a);
    const double qt = to_sgpr(__dmul_rn(ph.q, st.ts)), qdtl = to_sgpr(__dmul_rn(ph.q, st.dtl));
    const double qt_old = to_sgpr(__dmul_rn(ph.q, st.tso)), qdtl_old = to_sgpr(__dmul_rn(ph.q, st.dtlo));
    const double remap_qt = to_sgpr(__dmul_rn(ph.q, st.tnew));
    const bool do_remap = a.do_remap && rp.remap;
    const double *mfz = a.mfz + (size_t)run * 8 * NZ;

    const C *twg = reinterpret_cast<const C *>(sizeof(C) == 8 ? a.tw32 : (const void *)a.tw);
    for (int k = tid; k < NZ; k += NT) tw[k] = twg[k];
    for (int k = tid; k < NZD; k += NT) {
        skz[k] = a.kzv[k];
        skzr[k] = (R)a.kzv[k];
    }
    if constexpr (MF_LDS) {
        for (int k = tid; k < NZ * P::NMFZ; k += NT) smfz[(k % NZ) * P::NMFZ + k / NZ] = (R)mfz[k];
    } else {
#pragma unroll
        for (int p = 0; p < PW; ++p) {
            const int item = tid + p * NT;
#pragma unroll
            for (int c = 0; c < P::NMFZ; ++c) mfr[p][c] = (item < PI_) ? (R)mfz[c * NZ + item % NZ] : (R)0;
        }
    }

    double acc[PW][2];
#pragma unroll
    for (int p = 0; p < PW; ++p) acc[p][0] = acc[p][1] = 0.0;

    // run-local mode index of slot mb of group g, or -1
    auto local_mode = [&](int g, int mb) { const int l = g * MB + mb; return l < rp.nxy ? l : -1; };
    double2 upre[UW];
    auto fetch_rows = [&](int g) {
#pragma unroll
        for (int w = 0; w < UW; ++w) {
            const int k = tid + w * NT;
            const int mb = k / (4 * NZD);
            const int l = (k < NU && g < g1) ? local_mode(g, mb) : -1;
            upre[w] = (l >= 0) ? a.u[(size_t)(rp.mode_off + l) * 4 * NZD + (k - mb * 4 * NZD)] : make_double2(0.0, 0.0);
        }
    };
    fetch_rows(g0);

    for (int g = g0; g < g1; ++g) {
        // A: this group's rows to LDS, the next group's rows to registers
        if (tid < MB) {
            const int l = local_mode(g, tid);
            sloc[tid] = l;
            sidx[tid] = l >= 0 ? rp.mode_off + l : -1;
            if (l >= 0) {
                const double kx0 = a.kx0[rp.mode_off + l], ky = a.ky[rp.mode_off + l];
                skx0[tid] = kx0;
                sky[tid] = ky;
                skx[tid] = __dadd_rn(kx0, __dmul_rn(qt, ky));  // kx(t), MHD_FullQlin.cpp:727
            } else {
                skx0[tid] = sky[tid] = skx[tid] = 0.0;
            }
        }
        if (QL_PREFETCH_ROWS || g == g0) {
#pragma unroll
            for (int w = 0; w < UW; ++w) {
                const int k = tid + w * NT;
                if (k < NU) uin[k] = upre[w];
            }
        } else {
            for (int k = tid; k < NU; k += NT) {
                const int mb = k / (4 * NZD);
                const int l = local_mode(g, mb);
                uin[k] = (l >= 0) ? a.u[(size_t)(rp.mode_off + l) * 4 * NZD + (k - mb * 4 * NZD)] : make_double2(0.0, 0.0);
            }
        }
        if (QL_PREFETCH_ROWS) fetch_rows(g + 1);
        __syncthreads();

        // B: per-slot inverse Laplacians (assign_laplacians_, MHD_FullQlin.cpp:722-738,
        // Kdata.cpp:100-120) and the base spectral fields
        for (int k = tid; k < MB * NZD; k += NT) {
            const int mb = k / NZD, m = k % NZD;
            const double kx = skx[mb], ky = sky[mb], kz = skz[m];
            const bool zero = (ky == 0.0 && m == 0);
            const double lap2 = zero ? 0.0 : (-ky * ky + -(kz * kz));
            const double ilap2 = zero ? 1.0 : 1.0 / lap2;
            silF[k] = (sloc[mb] == 0 && m == 0) ? 1.0 : 1.0 / (-kx * kx + lap2);
            const double2 *U = uin + mb * 4 * NZD;
            P::bases(U[m], U[NZD + m], U[2 * NZD + m], U[3 * NZD + m], kx, ky, kz, ilap2, a.fac,
                     sbase + mb * P::NB * NZD + m, NZD);
        }
        __syncthreads();

        // C: inverse transforms
        fft_lines_from<NZ, NL, true>(buf, tw, LineSrc<NZ, P, C>{sbase, skzr});

        // D: products on the z grid and EMF / stress sums
        C prod[PW][3];
#pragma unroll
        for (int p = 0; p < PW; ++p) {
            const int item = tid + p * NT;
            if (PI_ % NT != 0 && item >= PI_) continue;
            const int mb = item / (2 * NZ), rem = item % (2 * NZ);
            const int part = rem / NZ, z = rem % NZ;
            const double ky = sky[mb];
            double racc[2];
            const R *mfp;
            if constexpr (MF_LDS) mfp = smfz + z * P::NMFZ;
            else mfp = mfr[p];
            P::products(part, buf + mb * P::NLINE * NZ + z, NZ, mfp, (R)skx[mb], (R)ky, prod[p], racc);
            if (sidx[mb] >= 0) {
                const double ftfac = (ky == 0.0 ? 1.0 : 2.0) * rp.rsf;  // MHD_FullQlin.cpp:314-318
                acc[p][0] += ftfac * racc[0];
                acc[p][1] += ftfac * racc[1];
            }
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PW; ++p) {
            const int item = tid + p * NT;
            if (PI_ % NT != 0 && item >= PI_) continue;
            const int mb = item / (2 * NZ), rem = item % (2 * NZ);
            const int part = rem / NZ, z = rem % NZ;
            C *Y = buf + mb * P::NOUT * NZ + z;
#pragma unroll
            for (int c = 0; c < 3; ++c)
                if (c < P::nprod(part)) Y[P::out_line(part, c) * NZ] = prod[p][c];
        }
        __syncthreads();

        // E: forward transforms
        fft_lines<NZ, P::NOUT * MB, false>(buf, tw);

        // F: RHS, dissipation operator, update
#pragma unroll
        for (int w = 0; w < AW; ++w) {
            const int k = tid + w * NT;
            if (AI % NT != 0 && k >= AI) continue;
            const int mb = k / (2 * NZD), r2 = k % (2 * NZD);
            const int pr = r2 / NZD, m = r2 % NZD;
            const int i = sidx[mb];
            if (i < 0) continue;
            const double2 *U = uin + mb * 4 * NZD;
            const double kx = skx[mb], ky = sky[mb], kz = skz[m];
            const double kz2 = -(kz * kz);
            double2 N0, N1;
            P::rhs_pair(pr, buf + mb * P::NOUT * NZ + D::pad(m), NZ, U[m], U[NZD + m], U[2 * NZD + m],
                        U[3 * NZD + m], kx, ky, kz, silF[mb * NZD + m], ph, N0, N1);
            const double Ln = P::lin(pr, kx + qdtl * ky, ky, kz2, ph);
            const size_t off0 = (size_t)i * 4 * NZD + (2 * pr) * NZD + m, off1 = off0 + NZD;
            if (a.mode == MODE_RHS) {
                if (a.n_out) {
                    a.n_out[off0] = N0;
                    a.n_out[off1] = N1;
                }
                if (a.lin_out) a.lin_out[off0] = a.lin_out[off1] = Ln;
                continue;
            }
            const double2 u0 = U[2 * pr * NZD + m], u1 = U[(2 * pr + 1) * NZD + m];
            double2 un0, un1;
            if (a.mode == MODE_ELIN) {
                // RK3CN_Elin.cpp:303-305, :335-337, :368-370
                const double e = exp(a.linC * Ln);
                double2 s0 = cscale(a.nlc1, N0), s1 = cscale(a.nlc1, N1);
                if (a.n_prev) {
                    s0 = cadd(s0, cscale(a.nlc2, a.n_prev[off0]));
                    s1 = cadd(s1, cscale(a.nlc2, a.n_prev[off1]));
                }
                un0 = cscale(e, cadd(s0, u0));
                un1 = cscale(e, cadd(s1, u1));
            } else {
                // RK3CN.cpp:105-107, :141-143, :177-179; the "old" operator is
                // the one of the previous substep (or linearOPs_Init at t0)
                double Lo;
                const double kxo = __dadd_rn(skx0[mb], __dmul_rn(qt_old, ky));
                if (st.oinit) {
                    const double lap2 = (ky == 0.0 && m == 0) ? 0.0 : (-ky * ky + kz2);
                    Lo = P::lin_init(pr, kxo, lap2, ph);
                } else {
                    Lo = P::lin(pr, kxo + qdtl_old * ky, ky, kz2, ph);
                }
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
            if (do_remap && remap_hit(remap_qt, rp.remap_thr, skx0[mb], ky)) un0 = un1 = make_double2(0.0, 0.0);
            a.u[off0] = un0;
            a.u[off1] = un1;
            if (a.n_out) {
                a.n_out[off0] = N0;
                a.n_out[off1] = N1;
            }
        }
        if (do_remap && tid < MB) {
            const int i = sidx[tid];
            if (i >= 0 && remap_hit(remap_qt, rp.remap_thr, skx0[tid], sky[tid]))
                a.kx0[i] = __dsub_rn(skx0[tid], rp.remap_shift);
        }
        __syncthreads();
    }

    // fixed-order reduction of the per-thread sums over the MB mode slots:
    // one partial [NACC][NZ] per workgroup
    double *red = reinterpret_cast<double *>(buf);
#pragma unroll
    for (int p = 0; p < PW; ++p) {
        const int item = tid + p * NT;
        if (PI_ % NT != 0 && item >= PI_) continue;
        const int mb = item / (2 * NZ), rem = item % (2 * NZ);
        const int part = rem / NZ, z = rem % NZ;
        red[(mb * P::NACC + P::acc_slot(part, 0)) * NZ + z] = acc[p][0];
        if (P::NACC == 4) red[(mb * P::NACC + P::acc_slot(part, 1)) * NZ + z] = acc[p][1];
    }
    __syncthreads();
    for (int k = tid; k < P::NACC * NZ; k += NT) {
        double s = 0.0;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) s += red[mb * P::NACC * NZ + k];
        a.partials[(size_t)blockIdx.x * P::NACC * NZ + k] = s;
    }
}

// FP64 path and mixed path (FP32 transforms); each with its own occupancy
// request, by pass size
constexpr bool half_pass(int NZ, int MB) { return NZ * MB < 256; }
template <int NZ, int MB, class P>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(half_pass(NZ, MB) ? QL_WAVES64_HALF : QL_WAVES64)))
void k_mode64(ModeArgs a)
{
    mode_body<NZ, MB, P, double2>(a);
}
template <int NZ, int MB, class P>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(half_pass(NZ, MB) ? QL_WAVES32_HALF : QL_WAVES32)))
void k_mode32(ModeArgs a)
{
    mode_body<NZ, MB, P, float2>(a);
}

}  // namespace qlk
