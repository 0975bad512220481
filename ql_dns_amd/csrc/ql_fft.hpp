// ql_fft.hpp -- batched dealiased z-transforms in LDS for gfx950.
//
// Replaces the reference's per-line FFTW calls (Auxiliary/fftwPlans.h:33-59:
// `inverse` = zero-pad nz -> NZ then unnormalised exp(+i) transform, `forward`
// = unnormalised exp(-i) transform then truncate NZ -> nz).  Here a workgroup
// transforms NL lines of length NZ at once, in place in LDS, with a Stockham
// autosort schedule of radix-8 passes (then one radix-4 or radix-2 pass for
// the remaining factor).  Each pass reads every element of every line into registers, waits at
// a barrier, and writes the butterflies back, so a single LDS buffer suffices.
// Twiddles come from a table tw[k] = exp(-2 pi i k / NZ) staged in LDS.  The
// complex type C is double2 (FP64 path) or float2 (FP32 transforms of the
// mixed-precision path).
#pragma once
#include <hip/hip_runtime.h>

namespace qlk {

constexpr int NT = 256;  // threads per workgroup of every kernel here

// IEEE double operations the compiler may not contract into an FMA with a
// neighbouring multiply or add.  HIP contracts across statements by default
// and its __dadd_rn / __dmul_rn are plain operators, so values that must round
// exactly as the reference's host expressions do (times, wavenumbers, remap
// thresholds, step coefficients) go through these.
__device__ __forceinline__ double rn_add(double a, double b)
{
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double rn_sub(double a, double b)
{
#pragma clang fp contract(off)
    return a - b;
}
__device__ __forceinline__ double rn_mul(double a, double b)
{
#pragma clang fp contract(off)
    return a * b;
}

// complex arithmetic on double2 (FP64 path) and float2 (FP32 transforms)
template <class C>
struct Cx;
template <>
struct Cx<double2> {
    typedef double R;
    __device__ static __forceinline__ double2 make(double x, double y) { return make_double2(x, y); }
};
template <>
struct Cx<float2> {
    typedef float R;
    __device__ static __forceinline__ float2 make(float x, float y) { return make_float2(x, y); }
};
template <class C>
using RealOf = typename Cx<C>::R;

template <class C>
__device__ __forceinline__ C cadd(C a, C b) { return Cx<C>::make(a.x + b.x, a.y + b.y); }
template <class C>
__device__ __forceinline__ C csub(C a, C b) { return Cx<C>::make(a.x - b.x, a.y - b.y); }
template <class C>
__device__ __forceinline__ C cmul(C a, C b) { return Cx<C>::make(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <class C>
__device__ __forceinline__ C cscale(RealOf<C> s, C a) { return Cx<C>::make(s * a.x, s * a.y); }
// (i k) * a
template <class C>
__device__ __forceinline__ C imul(RealOf<C> k, C a) { return Cx<C>::make(-k * a.y, k * a.x); }
// Re(a * conj(b))
template <class C>
__device__ __forceinline__ RealOf<C> redot(C a, C b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ double2 to_d(float2 a) { return make_double2(a.x, a.y); }
__device__ __forceinline__ double2 to_d(double2 a) { return a; }
template <class C>
__device__ __forceinline__ C from_d(double2 a) { return Cx<C>::make((RealOf<C>)a.x, (RealOf<C>)a.y); }

template <bool INV, class C>
__device__ __forceinline__ void dft4(C &v0, C &v1, C &v2, C &v3)
{
    const C a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), a3 = csub(v1, v3);
    // forward: -i * a3, inverse: +i * a3
    const C ja3 = INV ? Cx<C>::make(-a3.y, a3.x) : Cx<C>::make(a3.y, -a3.x);
    v0 = cadd(a0, a2);
    v2 = csub(a0, a2);
    v1 = cadd(a1, ja3);
    v3 = csub(a1, ja3);
}

// 8-point DFT as two 4-point DFTs (radix-2 split): even outputs from
// v[k] + v[k+4], odd outputs from (v[k] - v[k+4]) * w8^k.
template <bool INV, class C>
__device__ __forceinline__ void dft8(C *v)
{
    typedef RealOf<C> R;
    const R H = (R)0.70710678118654752440;  // 1/sqrt(2)
    C a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        a[k] = cadd(v[k], v[k + 4]);
        b[k] = csub(v[k], v[k + 4]);
    }
    // w8^1 = (1 -/+ i)/sqrt2, w8^2 = -/+ i, w8^3 = (-1 -/+ i)/sqrt2
    if (INV) {
        b[1] = Cx<C>::make(H * (b[1].x - b[1].y), H * (b[1].x + b[1].y));
        b[2] = Cx<C>::make(-b[2].y, b[2].x);
        b[3] = Cx<C>::make(-H * (b[3].x + b[3].y), H * (b[3].x - b[3].y));
    } else {
        b[1] = Cx<C>::make(H * (b[1].x + b[1].y), H * (b[1].y - b[1].x));
        b[2] = Cx<C>::make(b[2].y, -b[2].x);
        b[3] = Cx<C>::make(H * (b[3].y - b[3].x), -H * (b[3].x + b[3].y));
    }
    dft4<INV>(a[0], a[1], a[2], a[3]);
    dft4<INV>(b[0], b[1], b[2], b[3]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[2 * k] = a[k];
        v[2 * k + 1] = b[k];
    }
}

// 16-point DFT as two 8-point DFTs (radix-2 split): even outputs from
// v[k] + v[k+8], odd outputs from (v[k] - v[k+8]) * w16^k.
template <bool INV, class C>
__device__ __forceinline__ void dft16(C *v)
{
    typedef RealOf<C> R;
    const R c1 = (R)0.92387953251128675613;  // cos(pi/8)
    const R s1 = (R)0.38268343236508977173;  // sin(pi/8)
    const R H = (R)0.70710678118654752440;   // cos(pi/4)
    const R sg = INV ? (R)1 : (R)-1;         // sign of Im w16
    C a[8], b[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        a[k] = cadd(v[k], v[k + 8]);
        b[k] = csub(v[k], v[k + 8]);
    }
    b[1] = cmul(b[1], Cx<C>::make(c1, sg * s1));
    b[2] = cmul(b[2], Cx<C>::make(H, sg * H));
    b[3] = cmul(b[3], Cx<C>::make(s1, sg * c1));
    b[4] = Cx<C>::make(-sg * b[4].y, sg * b[4].x);
    b[5] = cmul(b[5], Cx<C>::make(-s1, sg * c1));
    b[6] = cmul(b[6], Cx<C>::make(-H, sg * H));
    b[7] = cmul(b[7], Cx<C>::make(-c1, sg * s1));
    dft8<INV>(a);
    dft8<INV>(b);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        v[2 * k] = a[k];
        v[2 * k + 1] = b[k];
    }
}

// One Stockham pass of radix R at stride NS over NL lines held in buf.
template <int NZ, int NL, bool INV, int R, int NS, class C>
__device__ __forceinline__ void fft_pass(C *buf, const C *tw)
{
    constexpr int M = NZ / R;
    constexpr int TOT = NL * M;
    constexpr int IT = (TOT + NT - 1) / NT;
    C v[IT][R];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int item = threadIdx.x + it * NT;
        if (TOT % NT == 0 || item < TOT) {
            const int line = item / M, j = item % M;
            const C *src = buf + line * NZ + j;
#pragma unroll
            for (int r = 0; r < R; ++r) v[it][r] = src[r * M];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int item = threadIdx.x + it * NT;
        if (TOT % NT == 0 || item < TOT) {
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
            C *dst = buf + line * NZ + (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[r * NS] = v[it][r];
        }
    }
    __syncthreads();
}

// Full length-NZ transform of NL lines (INV: exp(+i), else exp(-i)),
// unnormalised.  The caller must have synchronised after filling buf.
template <int NZ, int NL, bool INV, int NS = 1, class C>
__device__ __forceinline__ void fft_lines(C *buf, const C *tw)
{
    if constexpr (NS < NZ) {
        constexpr int R = (NZ / NS >= 8) ? 8 : ((NZ / NS >= 4) ? 4 : 2);
        fft_pass<NZ, NL, INV, R, NS>(buf, tw);
        fft_lines<NZ, NL, INV, NS * R>(buf, tw);
    }
}

// As fft_lines, but the first pass takes its inputs from src(line, grid index)
// instead of buf (so the input lines never have to be materialised in LDS).
// The caller must have synchronised before buf may be overwritten.
template <int NZ, int NL, bool INV, class C, class Src>
__device__ __forceinline__ void fft_lines_from(C *buf, const C *tw, const Src &src)
{
    constexpr int R = (NZ >= 8) ? 8 : ((NZ >= 4) ? 4 : 2);
    constexpr int M = NZ / R;
    constexpr int TOT = NL * M;
    constexpr int IT = (TOT + NT - 1) / NT;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int item = threadIdx.x + it * NT;
        if (TOT % NT == 0 || item < TOT) {
            const int line = item / M, j = item % M;
            C v[R];
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = src(line, j + r * M);
            if constexpr (R == 8) {
                dft8<INV>(v);
            } else if constexpr (R == 4) {
                dft4<INV>(v[0], v[1], v[2], v[3]);
            } else {
                const C x = v[0], y = v[1];
                v[0] = cadd(x, y);
                v[1] = csub(x, y);
            }
            C *dst = buf + line * NZ + j * R;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[r] = v[r];
        }
    }
    __syncthreads();
    fft_lines<NZ, NL, INV, R>(buf, tw);
}

// ---------------------------------------------------------------------------
// Two-pass transforms (radix 8x4 at NZ = 32, 8x8 at 64, 16x8 at 128, 16x16 at
// 256).  A pass reads its inputs through ld(line, index) and writes through
// st(line, index, value); SYNC_MID separates the reads from the writes with a
// barrier when both address the same buffer.
// ---------------------------------------------------------------------------
template <int NZ>
struct Plan2 {
    static constexpr int R1 = NZ >= 128 ? 16 : 8;
    static constexpr int R2 = NZ / R1;
};

template <int NZ, int NL, bool INV, int R, int NS, bool SYNC_MID, int TPB = NT, class C, class Ld, class St>
__device__ __forceinline__ void fft_pass_io(const C *tw, const Ld &ld, const St &st)
{
    constexpr int M = NZ / R;
    constexpr int TOT = NL * M;
    constexpr int IT = (TOT + TPB - 1) / TPB;
    C v[IT][R];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int item = threadIdx.x + it * TPB;
        if (TOT % TPB == 0 || item < TOT) {
            const int line = item / M, j = item % M;
#pragma unroll
            for (int r = 0; r < R; ++r) v[it][r] = ld(line, j + r * M);
        }
    }
    if constexpr (SYNC_MID) __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int item = threadIdx.x + it * TPB;
        if (TOT % TPB == 0 || item < TOT) {
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
    __syncthreads();
}

// lines of LS slots (LS >= NZ)
template <int LS, class C>
struct BufIO {
    C *buf;
    __device__ C operator()(int line, int i) const { return buf[line * LS + i]; }
    __device__ void operator()(int line, int i, C v) const { buf[line * LS + i] = v; }
};

// The intermediate between the two passes.  The first pass writes index
// j*R1 + s from lane j, a stride of R1 complex values across lanes, which
// would put every lane of a group on the same LDS banks; one padding slot per
// R1 values (index i -> i + i/R1, lines of LP slots) spreads them over all
// banks, and the second pass still reads consecutive slots across lanes.
template <int NZ, class C>
struct PadIO {
    static constexpr int R1 = Plan2<NZ>::R1;
    static constexpr int LP = NZ + NZ / R1;
    C *buf;
    __device__ C operator()(int line, int i) const { return buf[line * LP + i + i / R1]; }
    __device__ void operator()(int line, int i, C v) const { buf[line * LP + i + i / R1] = v; }
};

// Line strides that keep 8-lane groups on different LDS banks: consecutive
// lines start 8 slots apart modulo 16 (half the banks apart for 16-byte
// values, a quarter for 8-byte values).
constexpr int bank_stride(int n) { return n + ((8 - n % 16) + 16) % 16; }

// slots of buf used by a two-pass transform of NL lines of stride LS
template <int NZ, int NL, int LS>
struct Plan2Size {
    static constexpr int LP = NZ + NZ / Plan2<NZ>::R1;
    static constexpr int SLOTS = NL * (LP > LS ? LP : LS);
};

// Inverse (exp(+i)) transform of NL lines whose first pass reads src(line, g),
// result in buf (lines of LS slots), by a workgroup of TPB threads.  The
// caller must have synchronised before buf may be overwritten and before src
// may be read.  Three barriers.
template <int NZ, int NL, int LS, int TPB = NT, class C, class Src>
__device__ __forceinline__ void fft2_inv_from(C *buf, const C *tw, const Src &src)
{
    using PL = Plan2<NZ>;
    const BufIO<LS, C> io{buf};
    const PadIO<NZ, C> pad{buf};
    fft_pass_io<NZ, NL, true, PL::R1, 1, false, TPB>(tw, src, pad);
    fft_pass_io<NZ, NL, true, PL::R2, PL::R1, true, TPB>(tw, pad, io);
}

// Forward (exp(-i)) transform of NL lines held in buf (lines of LS slots); the
// last pass writes through dst(line, index, value), which must not alias the
// first Plan2Size<NZ, NL, LS>::SLOTS slots of buf.  The caller must have
// synchronised after filling buf.  Three barriers.
template <int NZ, int NL, int LS, int TPB = NT, class C, class Dst>
__device__ __forceinline__ void fft2_fwd_to(C *buf, const C *tw, const Dst &dst)
{
    using PL = Plan2<NZ>;
    const BufIO<LS, C> io{buf};
    const PadIO<NZ, C> pad{buf};
    fft_pass_io<NZ, NL, false, PL::R1, 1, true, TPB>(tw, io, pad);
    fft_pass_io<NZ, NL, false, PL::R2, PL::R1, false, TPB>(tw, pad, dst);
}

// dealiased sizes fixed by NZ (Models/Model.h:34: nz = 1 + 2*int(NZ/3))
template <int NZ>
struct Dealias {
    static constexpr int nz = 1 + 2 * (NZ / 3);
    static constexpr int nzl = (nz - 1) / 2;
    static constexpr int half = nzl + 1;
    // padded index of retained slot m, kz order (0..nzl, -nzl..-1)
    __device__ static __forceinline__ int pad(int m) { return m <= nzl ? m : m + (NZ - nz); }
    __device__ static __forceinline__ bool kept(int j) { return j <= nzl || j >= NZ - nzl; }
    __device__ static __forceinline__ int slot(int j) { return j <= nzl ? j : j - (NZ - nz); }
};

}  // namespace qlk
