// ql_dft.hpp -- the dealiased z transforms of the workgroup mode kernel as
// dense matrix products on the matrix cores (v_mfma_f64_16x16x4f64), a
// launch-shape candidate beside the LDS radix transforms of ql_fft.hpp for
// small NZ, where the transform really is a small dense contraction.
//
// Inverse (fftwPlans.h:33-82: zero-pad nz -> NZ, unnormalised exp(+i)):
//     Z[z][line] = sum_m exp(+2 pi i k_m z / NZ) X[m][line]
// Forward (fftwPlans.h:84-97: exp(-i), keep the dealiased band):
//     T[m][line] = sum_z exp(-2 pi i k_m z / NZ) Y[z][line]
// with k_m the kz index of slot m (0..nzl, -nzl..-1).  Each complex product is
// two real ones over an interleaved real K index kk = 2 m + c (c = 0 real,
// c = 1 imaginary part of the input): the real part of the output takes
// (Re W, -Im W) and the imaginary part (Im W, Re W).  Output tiles of 16 x 16
// (16 z points or band slots by 16 lines, real or imaginary part) are dealt
// to the workgroup's wavefronts; each accumulates over K in steps of 4 with
// the twiddles read from the LDS table tw[j] = exp(-2 pi i j / NZ).
//
// Lane layout of v_mfma_f64_16x16x4f64 (checked on gfx950): lane l holds
// A[l % 16][l / 16], B[l / 16][l % 16] and D[l / 16 + 4 v][l % 16], v = 0..3.
#pragma once
#include "ql_fft.hpp"

namespace qlk {

typedef double dft_d4 __attribute__((ext_vector_type(4)));

// Inverse transform of NL lines, src(line, m) the slot-m input of a line,
// into buf[line * LS + z].  The caller has synchronised after filling the
// inputs; ends with a barrier.
template <int NZ, int NL, int LS, int TPB, class Src>
__device__ __forceinline__ void dft_inv_mfma(double2 *buf, const double2 *tw, const Src &src)
{
    using D = Dealias<NZ>;
    constexpr int NZD = D::nz, KS = (2 * NZD + 3) / 4, MT = NZ / 16, NTL = (NL + 15) / 16;
    constexpr int TILES = 2 * MT * NTL, NWV = TPB / 64;
    static_assert(NZ % 16 == 0 && TPB % 64 == 0, "whole tiles and wavefronts");
    const int wv = threadIdx.x / 64, ln = threadIdx.x % 64, r = ln % 16, q = ln / 16;
    for (int t = wv; t < TILES; t += NWV) {
        const int part = t % 2, mt = (t / 2) % MT, nt = t / (2 * MT);
        const int z = mt * 16 + r, line = nt * 16 + r;
        dft_d4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int ks = 0; ks < KS; ++ks) {
            const int kk = ks * 4 + q, m = kk >> 1, c = kk & 1;
            double a = 0.0, bv = 0.0;
            if (m < NZD) {
                const int km = m <= D::nzl ? m : m - NZD;
                const double2 w = tw[(km * z) & (NZ - 1)];  // exp(-i ...); the inverse takes its conjugate
                a = part == 0 ? (c == 0 ? w.x : w.y) : (c == 0 ? -w.y : w.x);
                if (line < NL) {
                    const double2 x = src(line, m);
                    bv = c == 0 ? x.x : x.y;
                }
            }
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc, 0, 0, 0);
        }
        if (line < NL)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int zz = mt * 16 + q + 4 * v, ll = nt * 16 + r;
                reinterpret_cast<double *>(&buf[ll * LS + zz])[part] = acc[v];
            }
    }
    __syncthreads();
}

// Forward transform of NL lines held in buf[line * LS + z], the dealiased
// band written through dst(line, m, part, value) (part 0 real, 1 imaginary),
// which must not alias the lines.  The caller has synchronised after filling
// buf; ends with a barrier.
template <int NZ, int NL, int LS, int TPB, class Dst>
__device__ __forceinline__ void dft_fwd_mfma(const double2 *buf, const double2 *tw, const Dst &dst)
{
    using D = Dealias<NZ>;
    constexpr int NZD = D::nz, KS = (2 * NZ) / 4, MT = (NZD + 15) / 16, NTL = (NL + 15) / 16;
    constexpr int TILES = 2 * MT * NTL, NWV = TPB / 64;
    static_assert(TPB % 64 == 0, "whole wavefronts");
    const int wv = threadIdx.x / 64, ln = threadIdx.x % 64, r = ln % 16, q = ln / 16;
    for (int t = wv; t < TILES; t += NWV) {
        const int part = t % 2, mt = (t / 2) % MT, nt = t / (2 * MT);
        const int m = mt * 16 + r, line = nt * 16 + r;
        const int km = m <= D::nzl ? m : m - NZD;
        dft_d4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int ks = 0; ks < KS; ++ks) {
            const int kk = ks * 4 + q, z = kk >> 1, c = kk & 1;
            double a = 0.0, bv = 0.0;
            if (m < NZD) {
                const double2 w = tw[(km * z) & (NZ - 1)];
                a = part == 0 ? (c == 0 ? w.x : -w.y) : (c == 0 ? w.y : w.x);
            }
            if (line < NL) {
                const double2 y = buf[line * LS + z];
                bv = c == 0 ? y.x : y.y;
            }
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc, 0, 0, 0);
        }
        if (line < NL)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int mm = mt * 16 + q + 4 * v;
                if (mm < NZD) dst(line, mm, part, acc[v]);
            }
    }
    __syncthreads();
}

}  // namespace qlk
