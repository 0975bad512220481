// ql_mode.hpp -- the fused per-(kx,ky) kernel of one RK3 substep (FP64).
//
// One kernel skeleton, three model policies:
//   ModelFull  MHD_FullQlin (Models/MHD_FullQlin.cpp:171-427): fluctuations
//              coupled to Bx, By, Ux, Uy(z); EMF and both stresses reduced.
//   ModelHD    HD_fullU (Models/HD_fullU.cpp:145-309): ModelFull with the
//              magnetic fields at zero; only (ux, zeta) read and written.
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
//      base table directly (lines = bases, for MHD_BQlin also i kz uz);
//   D  mean-field x fluctuation products on the z grid, x and y derivatives
//      being i kx and i ky times the z profile; EMF / stress z-profiles summed
//      in registers over the workgroup's modes;
//   E  forward transforms of the product lines;
//   F  RHS assembly, dissipation operator at the substep end, substep update
//      per (mode, kz slot, field pair), remap predicate on the new state.
#pragma once
#include "ql_dft.hpp"
#include "ql_kernels.hpp"

// occupancy requests (waves per SIMD, which bound the VGPR budget) for
// 256-point passes and for 128-point passes (half the modes per pass)
#ifndef QL_WAVES64
#define QL_WAVES64 2
#endif
// FP64 at NZ = 256: fetch the state at the start of each group (1) or one
// group ahead (0)
#ifndef QL_NZ256_FETCH_AT_USE
#define QL_NZ256_FETCH_AT_USE 1
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

// The stress terms of MHD_FullQlin (MHD_FullQlin.cpp:339-362) are
// -Re(u . grad conj(ux)) + Re(b . grad conj(bx)) and likewise for y, with
// x and y derivatives i kx, i ky.  The reconstructed fields are divergence
// free slot by slot (i kx ux + i ky uy + i kz uz = 0 wherever lap2 != 0, by
// MHD_FullQlin.cpp:220-221), so Re(u . grad conj(ux)) = d/dz Re(uz conj(ux)):
// the stresses are z-derivatives, accumulated here as
// Re(bz conj(bx)) - Re(uz conj(ux)) and differentiated by the mean-field
// kernel (x i kz) exactly as the EMF is, which takes the four dz lines out of
// the inverse transforms.  The one slot the reconstruction leaves alone,
// kz = 0 of a ky = 0 mode, holds ux = bx = 0 in a divergence-free state and
// keeps it (its RHS is ilapF (-kx^2 T1) - T1 = 0).
struct ModelFull {
    static constexpr int NPAIR = 2;  // field pairs updated: (ux, zeta), (bx, eta)
    static constexpr int NB = 6;     // bases: ux uy uz bx by bz
    static constexpr int NLINE = 6;  // inverse-transformed lines: the bases
    static constexpr int NOUT = 6;   // forward lines (3 momentum, 3 induction)
    static constexpr int NMFZ = 8;   // MBx dzMBx MBy dzMBy MUx dzMUx MUy dzMUy
    static constexpr int NACC = 4;   // EMF x, EMF y, stress x, stress y (as above)
    __device__ static int base_of(int l) { return l; }
    __device__ static bool dz_of(int) { return false; }
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
    // MHD_FullQlin.cpp:271-300 (products) and :322-362 (EMF, stresses)
    template <class C>
    __device__ static void products(int part, const C *X, int ld, const RealOf<C> *mf, RealOf<C> kx, RealOf<C> ky,
                                    C dzc, C *prod, double *racc)
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
            // stresses before the z-derivative (see above)
            racc[0] = redot64(b2, b0) - redot64(u2, u0);
            racc[1] = redot64(b2, b1) - redot64(u2, u1);
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

// HD_fullU (Models/HD_fullU.cpp:145-309) on the MHD_FullQlin state layout,
// whose magnetic fields it holds at zero (b = eta = 0, B0z = 0): the velocity
// bases and the momentum products only, each exactly as ModelFull computes
// them with b = 0 (the terms dropped are products with zero), and the
// Reynolds stresses -Re(uz conj(ux)), -Re(uz conj(uy)) before their
// z-derivative in the stress slots 2, 3 (the EMF slots stay zero).  Only the
// (ux, zeta) pair is read, updated and written.  Both mode kernels use it.
struct ModelHD {
    static constexpr int NPAIR = 1;
    static constexpr int NB = 3;     // bases: ux uy uz
    static constexpr int NLINE = 3;
    static constexpr int NOUT = 3;   // momentum lines
    static constexpr int NMFZ = 8;   // the MHD_FullQlin profile table (Ux, Uy at 4-7)
    static constexpr int NACC = 4;
    __device__ static int base_of(int l) { return l; }
    __device__ static bool dz_of(int) { return false; }
    template <class C>
    __device__ static void bases(double2 ux, double2 ze, double2, double2, double kx, double ky, double kz,
                                 double ilap2, double fac, C *out, int stride)
    {
        double2 uy, uz;
        yz_from(ux, ze, kx, ky, kz, ilap2, uy, uz);
        out[0] = from_d<C>(cscale(fac, ux));
        out[stride] = from_d<C>(cscale(fac, uy));
        out[2 * stride] = from_d<C>(cscale(fac, uz));
    }
    __device__ static int nprod(int part) { return part == 0 ? 3 : 0; }
    __device__ static int out_line(int, int c) { return c; }
    __device__ static int acc_slot(int, int c) { return 2 + c; }
    template <class C>
    __device__ static void products(int part, const C *X, int ld, const RealOf<C> *mf, RealOf<C> kx, RealOf<C> ky,
                                    C dzc, C *prod, double *racc)
    {
        typedef RealOf<C> R;
        racc[0] = racc[1] = 0.0;
        if (part != 0) return;
        const C u0 = X[0], u1 = X[ld], u2 = X[2 * ld];
        const R MUx = mf[4], dzMUx = mf[5], MUy = mf[6], dzMUy = mf[7];
        const R aU = kx * MUx + ky * MUy;
        // U.grad(u) + uz dzU (MHD_FullQlin.cpp:271-280 with B = b = 0)
        prod[0] = cadd(imul(aU, u0), cscale(dzMUx, u2));
        prod[1] = cadd(imul(aU, u1), cscale(dzMUy, u2));
        prod[2] = imul(aU, u2);
        racc[0] = -redot64(u2, u0);
        racc[1] = -redot64(u2, u1);
    }
    template <class C>
    __device__ static void rhs_pair(int, const C *T, int ld, double2 ux, double2 ze, double2, double2, double kx,
                                    double ky, double kz, double ilapF, const Phys &a, double2 &N0, double2 &N1)
    {
        // HD_fullU.cpp:283-286 (ModelFull's pair 0 with b = 0)
        const double q = a.q, Om = a.omega;
        const double2 T1 = to_d(T[0]), T2 = to_d(T[ld]), T3 = to_d(T[2 * ld]);
        double2 s = cscale(2.0 * kx * ky * q, ux);
        s = cadd(s, imul(2.0 * Om * kz, ze));
        s = csub(s, cscale(kx * kx, T1));
        s = csub(s, cscale(kx * ky, T2));
        s = csub(s, cscale(kx * kz, T3));
        N0 = csub(cscale(ilapF, s), T1);
        N1 = imul((q - 2.0 * Om) * kz, ux);
        N1 = csub(N1, csub(imul(kz, T2), imul(ky, T3)));
    }
    __device__ static double lin(int pr, double kxn, double ky, double kz2, const Phys &a)
    {
        return ModelFull::lin(pr, kxn, ky, kz2, a);
    }
    __device__ static double lin_init(int pr, double kxo, double lap2, const Phys &a)
    {
        return ModelFull::lin_init(pr, kxo, lap2, a);
    }
};

// The reference transforms dz uz as a line of its own (MHD_BQlin.cpp:195);
// the reconstruction is divergence free slot by slot (as for ModelFull), so
// on the z grid dz uz = -i (kx ux + ky uy) from the ux and uy lines, except
// for the kz = 0 slot of a ky = 0 mode, where uz = 0 and ux is free: there
// -i kx ux is a z-constant that the reference's line does not hold.  The
// kernels pass it back as dzc = i kx (fac1D ux at kz = 0) for ky = 0 modes
// (zero otherwise), so the result is the reference's for any state.
struct ModelBQ {
    static constexpr int NPAIR = 2;
    static constexpr int NB = 8;    // bases: ux zeta bx eta uy uz by bz
    static constexpr int NLINE = 8; // the bases
    static constexpr int NOUT = 5;  // u (two parts), b, zeta, eta
    static constexpr int NMFZ = 3;  // By dzBy dzdzBy
    static constexpr int NACC = 2;  // EMF x, EMF y
    __device__ static int base_of(int l) { return l; }
    __device__ static bool dz_of(int) { return false; }
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
                                    C dzc, C *prod, double *racc)
    {
        typedef RealOf<C> R;
        const C ux = X[0], ze = X[ld], bx = X[2 * ld], et = X[3 * ld];
        const C uy = X[4 * ld], uz = X[5 * ld], by = X[6 * ld], bz = X[7 * ld];
        const C dzuz = cadd(cadd(imul(-kx, ux), imul(-ky, uy)), dzc);
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

// LDS layout of a mode group: line l of mode slot mb at line l*MB + mb, so the
// product lines (l < NOUT) are the first NOUT*MB lines; the transformed
// products are stored compactly ([line][kz slot]) behind the slots their
// transform uses, in the space of the remaining input lines.
template <int NZ, int MB, class P>
struct GroupLayout {
    static constexpr int NZD = Dealias<NZ>::nz;
    static constexpr int NZDP = bank_stride(NZD);  // base-table line stride
    static constexpr int LS = bank_stride(NZ);     // z-grid line stride
    static constexpr int NL = P::NLINE * MB;
    static constexpr int NLO = P::NOUT * MB;
    static constexpr int CT_OFF = Plan2Size<NZ, NLO, LS>::SLOTS;
    static constexpr int CT = NLO * NZD;
    static constexpr int A = Plan2Size<NZ, NL, LS>::SLOTS, B = CT_OFF + CT;
    static constexpr int BUF = A > B ? A : B;
};

// first-pass source of the inverse transforms in the interleaved layout
template <int NZ, int MB, class P, class C>
struct LineSrcI {
    const C *sbase;  // [(base * MB + mb)][NZDP]
    const RealOf<C> *skz;
    __device__ C operator()(int line, int g) const
    {
        using D = Dealias<NZ>;
        if (!D::kept(g)) return Cx<C>::make(0, 0);
        const int m = D::slot(g);
        const int l = line / MB, mb = line - l * MB;
        const C v = sbase[(P::base_of(l) * MB + mb) * GroupLayout<NZ, MB, P>::NZDP + m];
        return P::dz_of(l) ? imul(skz[m], v) : v;
    }
};

// slot-m input of line `line` (interleaved layout) for the matrix-core
// inverse transform (ql_dft.hpp)
template <int NZ, int MB, class P>
struct SlotSrc {
    const double2 *sbase;  // [(base * MB + mb)][NZDP]
    const double *skz;
    __device__ double2 operator()(int line, int m) const
    {
        const int l = line / MB, mb = line - l * MB;
        const double2 v = sbase[(P::base_of(l) * MB + mb) * GroupLayout<NZ, MB, P>::NZDP + m];
        return P::dz_of(l) ? imul(skz[m], v) : v;
    }
};

// band sink of the matrix-core forward transform: [line][nz], one part
struct SlotDst {
    double2 *ct;
    int nz;
    __device__ void operator()(int line, int m, int part, double v) const
    {
        reinterpret_cast<double *>(&ct[line * nz + m])[part] = v;
    }
};

// last-pass sink of the forward transforms: keep the dealiased band, compact
template <int NZ, class C>
struct CompactDst {
    C *ct;  // [line][nz]
    __device__ void operator()(int line, int j, C v) const
    {
        using D = Dealias<NZ>;
        if (D::kept(j)) ct[line * D::nz + D::slot(j)] = v;
    }
};

// threads of a mode-kernel workgroup: one per z point of a pass
constexpr int mode_threads(int NZ, int MB) { return NZ * MB < NT ? NZ * MB : NT; }

template <int NZ, int MB, class P, class C, int MODE, bool SPL, bool DFT = false>
__device__ __forceinline__ void mode_body(const ModeArgs &a)
{
    typedef RealOf<C> R;
    using D = Dealias<NZ>;
    using L = GroupLayout<NZ, MB, P>;
    constexpr int NZD = D::nz;
    constexpr int NIT = MB * NZD;  // (mode slot, kz slot) items: bases, RHS, update
    constexpr int NZI = MB * NZ;   // (mode slot, z) items: products
    constexpr int TPB = mode_threads(NZ, MB);
    static_assert(NIT <= TPB && NZI <= TPB, "one item per thread");
    __shared__ C buf[L::BUF];
    __shared__ C sbase[MB * P::NB * L::NZDP];
    __shared__ C tw[NZ];
    __shared__ R skzr[NZD];
    __shared__ double skx[MB], sky[MB];
    __shared__ int sok[MB];
    __shared__ double2 sdz[MB];  // fac1D ux at kz = 0 of a ky = 0 mode (ModelBQ's dz uz)
    // mean-field z-profiles [profile][z]: in registers of the product threads
    // for the FP32 path and at NZ = 256 (where LDS is short), in LDS otherwise
    // (FP64 registers are the scarcer resource)
    constexpr bool MF_REG = sizeof(C) == 8 || NZ >= 256;
    __shared__ R smfz[MF_REG ? 1 : NZ * P::NMFZ];
    const int tid = threadIdx.x;

    // this workgroup's run, its mode groups [g0, g1), the substep times
    const WgInfo wgi = a.wg[blockIdx.x];
    const int run = wgi.run, g0 = wgi.g0, g1 = wgi.g1, mode_off = wgi.mode_off, nxy = wgi.nxy;
    const RunParams rp = a.runs[run];
    const Phys ph = rp.ph;
    // q t products: wave-uniform, kept in scalar registers
    const SubTimes st = sub_times<MODE == MODE_CN>(a, run);
    const double linC = to_sgpr(st.linC), nlc1 = to_sgpr(st.nlc1), nlc2 = to_sgpr(st.nlc2);
    const double qt = to_sgpr(rn_mul(ph.q, st.ts)), qdtl = to_sgpr(rn_mul(ph.q, st.dtl));
    const double qt_old = to_sgpr(rn_mul(ph.q, st.tso)), qdtl_old = to_sgpr(rn_mul(ph.q, st.dtlo));
    const double remap_qt = to_sgpr(rn_mul(ph.q, st.tnew));
    const bool do_remap = a.do_remap && rp.remap;

    // the thread's items: (imb, im) for the bases and the update, (zmb, zz)
    // for the products; both fixed for the whole kernel
    const bool has_it = tid < NIT, has_z = tid < NZI;
    const int imb = tid / NZD, im = tid - imb * NZD;
    const int zmb = tid / NZ, zz = tid - zmb * NZ;

    // tables: every load is issued before the first LDS store, so that they
    // are in flight together
    const C *twg = reinterpret_cast<const C *>(sizeof(C) == 8 ? a.tw32 : (const void *)a.tw);
    const double *mfz = a.mfz + (size_t)run * 8 * NZ;
    constexpr int TWN = (NZ + TPB - 1) / TPB, MFN = MF_REG ? 1 : (NZ * P::NMFZ + TPB - 1) / TPB;
    static_assert(NZD <= TPB, "one wavenumber per thread");
    C twv[TWN];
    double mfv[MFN];
#pragma unroll
    for (int w = 0; w < TWN; ++w) twv[w] = (tid + w * TPB < NZ) ? twg[tid + w * TPB] : Cx<C>::make(0, 0);
    const double kzt = tid < NZD ? a.kzv[tid] : 0.0;
    const double kz = has_it ? a.kzv[im] : 0.0, kz2 = -(kz * kz);
    R mfr[MF_REG ? P::NMFZ : 1];
    if constexpr (MF_REG) {
#pragma unroll
        for (int c = 0; c < P::NMFZ; ++c) mfr[c] = has_z ? (R)mfz[c * NZ + zz] : (R)0;
    } else {
#pragma unroll
        for (int w = 0; w < MFN; ++w) mfv[w] = (tid + w * TPB < NZ * P::NMFZ) ? mfz[tid + w * TPB] : 0.0;
    }

    // the FP64 companion of a mixed-precision launch (SPLIT_HI): workgroup g0
    // of the g1 its run has takes those of the run's modes l = g0 + k g1 whose
    // amplitude reaches the threshold, listed in mode order (the host sizes
    // g1 so that no workgroup has more than COMP_LIST candidates)
    const bool comp = SPL && a.split == SPLIT_HI;
    __shared__ int slist[SPL ? COMP_LIST : 1];
    __shared__ int scnt[SPL ? (TPB + 63) / 64 : 1];
    int ncomp = 0;
    if constexpr (SPL) {
        if (comp) {
            constexpr int NWV = (TPB + 63) / 64;
            const int wv = tid / 64, ln = tid % 64;
            int base = 0;
            for (int c0 = 0; g0 + c0 * g1 < nxy; c0 += TPB) {
                const int l = g0 + (c0 + tid) * g1;
                const bool hi = l < nxy && a.hi[mode_off + l];
                const unsigned long long bal = __ballot(hi);
                if (ln == 0) scnt[wv] = __popcll(bal);
                __syncthreads();
                int off = base, tot = 0;
#pragma unroll
                for (int w = 0; w < NWV; ++w) {
                    off += w < wv ? scnt[w] : 0;
                    tot += scnt[w];
                }
                if (hi) slist[off + __popcll(bal & ((1ull << ln) - 1ull))] = l;
                base += tot;
                __syncthreads();
            }
            ncomp = base;
        }
    }
    const int gbeg = comp ? 0 : g0, gend = comp ? (ncomp + MB - 1) / MB : g1;
    // run-local mode index of this thread's item in group g, or -1
    auto local_mode = [&](int g) {
        if (comp) {
            const int k = g * MB + imb;
            return (has_it && g < gend && k < ncomp) ? slist[k] : -1;
        }
        const int l = g * MB + imb;
        return (has_it && g < g1 && l < nxy) ? l : -1;
    };
    // the item's state row and wavenumbers, fetched one group ahead (except
    // for FP64 at NZ = 256, where the registers are needed by the transforms)
    constexpr bool AHEAD = !(sizeof(C) == 16 && NZ >= 256 && QL_NZ256_FETCH_AT_USE);
    double2 upre[4];
    double kpre[2];
    // SPL: the mode's precision flag, fetched with its state and tested where
    // the state is used (a test here would wait for the whole prefetch)
    unsigned char hpre = 0;
    auto fetch = [&](int g) {
        const int l = local_mode(g);
        const int i = mode_off + (l < 0 ? 0 : l);
        const size_t off = (size_t)i * 4 * NZD + im;
#pragma unroll
        for (int f = 0; f < 4; ++f)
            upre[f] = (l >= 0 && f < 2 * P::NPAIR) ? a.u[off + f * NZD] : make_double2(0.0, 0.0);
        kpre[0] = (l >= 0) ? a.kx0[i] : 0.0;
        kpre[1] = (l >= 0) ? a.ky[i] : 0.0;
        if constexpr (SPL) hpre = (l >= 0 && a.split == SPLIT_LO) ? a.hi[i] : 0;
    };
    if (AHEAD) fetch(gbeg);
#pragma unroll
    for (int w = 0; w < TWN; ++w)
        if (tid + w * TPB < NZ) tw[tid + w * TPB] = twv[w];
    if (tid < NZD) skzr[tid] = (R)kzt;
    if constexpr (!MF_REG) {
#pragma unroll
        for (int w = 0; w < MFN; ++w)
            if (tid + w * TPB < NZ * P::NMFZ) smfz[tid + w * TPB] = (R)mfv[w];
    }
    double acc[P::NACC];
#pragma unroll
    for (int c = 0; c < P::NACC; ++c) acc[c] = 0.0;

    for (int g = gbeg; g < gend; ++g) {
        // B: state of the item (prefetched), next group's state, previous RHS;
        // kx(t), 1/lap2, 1/lapF (assign_laplacians_, MHD_FullQlin.cpp:722-738,
        // Kdata.cpp:100-120) and the base spectral fields
        int l = local_mode(g);
        if (!AHEAD) fetch(g);
        const double2 B0 = upre[0], B1 = upre[1], B2 = upre[2], B3 = upre[3];
        const double kx0 = kpre[0], ky = kpre[1];
        const unsigned char hcur = hpre;
        if (AHEAD) fetch(g + 1);
        if (SPL && hcur) l = -1;  // the FP64 companion takes this mode
        const size_t off = (size_t)(mode_off + (l < 0 ? 0 : l)) * 4 * NZD + im;
        double kx = 0.0, ilapF = 1.0;
        if (has_it) {
            if (l >= 0) kx = rn_add(kx0, rn_mul(qt, ky));  // kx(t), MHD_FullQlin.cpp:727
            const bool zero = (ky == 0.0 && im == 0);
            const double lap2 = zero ? 0.0 : (-ky * ky + kz2);
            const double ilap2 = zero ? 1.0 : 1.0 / lap2;
            if (l >= 0) ilapF = (l == 0 && im == 0) ? 1.0 : 1.0 / (-kx * kx + lap2);
            P::bases(B0, B1, B2, B3, kx, ky, kz, ilap2, a.fac, sbase + imb * L::NZDP + im, MB * L::NZDP);
            if (im == 0) {
                skx[imb] = kx;
                sky[imb] = ky;
                sok[imb] = l >= 0;
                sdz[imb] = ky == 0.0 ? cscale(a.fac, B0) : make_double2(0.0, 0.0);
            }
        }
        __syncthreads();

        // C: inverse transforms of the padded lines
        if constexpr (DFT) dft_inv_mfma<NZ, L::NL, L::LS, TPB>(buf, tw, SlotSrc<NZ, MB, P>{sbase, skzr});
        else fft2_inv_from<NZ, L::NL, L::LS, TPB>(buf, tw, LineSrcI<NZ, MB, P, C>{sbase, skzr});

        // D: products on the z grid, in place, and the EMF / stress sums
        if (has_z) {
            C *X = buf + zmb * L::LS + zz;  // line l at X[l * MB * LS]
            const R kxr = (R)skx[zmb], kyr = (R)sky[zmb];
            R mfl[P::NMFZ];
            if constexpr (!MF_REG) {
#pragma unroll
                for (int c = 0; c < P::NMFZ; ++c) mfl[c] = smfz[c * NZ + zz];
            }
            const R *mfp = MF_REG ? mfr : mfl;
            C prod[2][3];
            double racc[2][2];
            const C dzc = from_d<C>(imul(skx[zmb], sdz[zmb]));
            P::products(0, X, MB * L::LS, mfp, kxr, kyr, dzc, prod[0], racc[0]);
            P::products(1, X, MB * L::LS, mfp, kxr, kyr, dzc, prod[1], racc[1]);
            if (sok[zmb]) {
                const double ftfac = (sky[zmb] == 0.0 ? 1.0 : 2.0) * rp.rsf;  // MHD_FullQlin.cpp:314-318
#pragma unroll
                for (int part = 0; part < 2; ++part) {
                    acc[P::acc_slot(part, 0)] += ftfac * racc[part][0];
                    if (P::NACC == 4) acc[P::acc_slot(part, 1)] += ftfac * racc[part][1];
                }
            }
#pragma unroll
            for (int part = 0; part < 2; ++part)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (c < P::nprod(part)) X[P::out_line(part, c) * MB * L::LS] = prod[part][c];
        }
        __syncthreads();
        // previous substep's RHS of the item, needed in F
        double2 np[4];
#pragma unroll
        for (int f = 0; f < 4; ++f)
            np[f] = (l >= 0 && a.n_prev && f < 2 * P::NPAIR) ? a.n_prev[off + f * NZD] : make_double2(0.0, 0.0);

        // E: forward transforms of the product lines, dealiased band kept
        C *ct = buf + L::CT_OFF;  // [line][nz]
        if constexpr (DFT) dft_fwd_mfma<NZ, L::NLO, L::LS, TPB>(buf, tw, SlotDst{ct, NZD});
        else fft2_fwd_to<NZ, L::NLO, L::LS, TPB>(buf, tw, CompactDst<NZ, C>{ct});

        // F: RHS, dissipation operator at the substep end, update
        if (l >= 0) {
            const C *T = ct + imb * NZD + im;
            // without prefetch (FP64, NZ = 256) the item's state is read again
            // here instead of being kept in registers through both transforms,
            // which removes that kernel's spills
            double2 U0 = B0, U1 = B1, U2 = B2, U3 = B3;
            if constexpr (!AHEAD) {
                U0 = a.u[off];
                U1 = a.u[off + NZD];
                if constexpr (P::NPAIR == 2) {
                    U2 = a.u[off + 2 * NZD];
                    U3 = a.u[off + 3 * NZD];
                }
            }
            bool hit = false;
            if (do_remap) hit = remap_hit(remap_qt, rp.remap_thr, kx0, ky);
#pragma unroll
            for (int pr = 0; pr < P::NPAIR; ++pr) {
                double2 N0, N1;
                P::rhs_pair(pr, T, MB * NZD, U0, U1, U2, U3, kx, ky, kz, ilapF, ph, N0, N1);
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
                        const double lap2 = (ky == 0.0 && im == 0) ? 0.0 : (-ky * ky + kz2);
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
            if (hit && im == 0) a.kx0[mode_off + l] = rn_sub(kx0, rp.remap_shift);
        }
        // no barrier here: the next group's B touches only sbase and the mode
        // tables, which nothing in F reads; its C starts after B's barrier
    }

    // fixed-order reduction of the per-thread sums over the MB mode slots:
    // one partial [NACC][NZ] per workgroup
    __syncthreads();
    double *red = reinterpret_cast<double *>(buf);
    if (has_z)
#pragma unroll
        for (int c = 0; c < P::NACC; ++c) red[(zmb * P::NACC + c) * NZ + zz] = acc[c];
    __syncthreads();
    for (int k = tid; k < P::NACC * NZ; k += TPB) {
        double s = 0.0;
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) s += red[mb * P::NACC * NZ + k];
        a.partials[(size_t)blockIdx.x * P::NACC * NZ + k] = s;
    }
}

// FP64 path and mixed path (FP32 transforms); each with its own occupancy
// request, by pass size; the substep mode (MODE_RHS, MODE_ELIN, MODE_CN) is a
// template parameter, so each instantiation carries only its own update.  A
// workgroup has one thread per z point of a pass (MB NZ, at most NT).
constexpr bool half_pass(int NZ, int MB) { return NZ * MB < 256; }
template <int NZ, int MB, class P, int MODE>
__global__ __launch_bounds__(mode_threads(NZ, MB)) __attribute__((amdgpu_waves_per_eu(half_pass(NZ, MB) ? QL_WAVES64_HALF : QL_WAVES64)))
void k_mode64(ModeArgs a)
{
    mode_body<NZ, MB, P, double2, MODE, false>(a);
}
// FP64 with the dealiased transforms as matrix-core products (ql_dft.hpp;
// NZ = 32 and 64)
template <int NZ, int MB, class P, int MODE>
__global__ __launch_bounds__(mode_threads(NZ, MB)) __attribute__((amdgpu_waves_per_eu(half_pass(NZ, MB) ? QL_WAVES64_HALF : QL_WAVES64)))
void k_mode64d(ModeArgs a)
{
    mode_body<NZ, MB, P, double2, MODE, false, true>(a);
}
// the FP64 companion of a mixed-precision launch (ModeArgs::split)
template <int NZ, int MB, class P, int MODE>
__global__ __launch_bounds__(mode_threads(NZ, MB)) __attribute__((amdgpu_waves_per_eu(half_pass(NZ, MB) ? QL_WAVES64_HALF : QL_WAVES64)))
void k_mode64s(ModeArgs a)
{
    mode_body<NZ, MB, P, double2, MODE, true>(a);
}
template <int NZ, int MB, class P, int MODE>
__global__ __launch_bounds__(mode_threads(NZ, MB)) __attribute__((amdgpu_waves_per_eu(half_pass(NZ, MB) ? QL_WAVES32_HALF : QL_WAVES32)))
void k_mode32(ModeArgs a)
{
    mode_body<NZ, MB, P, float2, MODE, true>(a);
}

}  // namespace qlk
