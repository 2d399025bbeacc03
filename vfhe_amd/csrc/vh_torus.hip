// vh_torus.hip -- CGGI functional bootstrapping over the 32-bit discretised torus
// (Z / 2^32) with the external product computed through an FP64 negacyclic FFT.
//
// No counterpart in the reference: its CGGI16 (modules/fhe/python/src/vfhe/fhe/
// cggi16.py:46-140) runs over RNS primes and the NTT.  This keeps that bootstrap's
// structure -- rotation amounts rounded as cggi16.py:54-86 does, the test vector
// rotated by -rot_b, one CMUX per key index with a non-zero amount (cggi16.py:76-90),
// extraction of coefficient 0 (cggi16.py:118-140) -- on the torus, with the external
// product and signed gadget decomposition of TFHE (Chillotti, Gama, Georgieva,
// Izabachene, "TFHE: fast fully homomorphic encryption over the torus",
// J. Cryptology 33 (2020): Definition 3.6 / Lemma 3.7 (decomposition), Theorem 3.13
// (external product), Algorithm 3 (blind rotation)).
//
// Negacyclic FFT.  A real polynomial p of N = 1024 coefficients is evaluated at the
// N/2 roots x of X^N = -1 with x^(N/2) = i: x = zeta * w^m, zeta = e^(i pi / N),
// w = e^(-2 pi i / M), M = N/2.  Since x^(N/2) = i,
//     p(zeta w^m) = sum_{j < M} (p_j + i p_{j+M}) zeta^j w^(jm),
// an M-point DFT of the folded and twisted sequence z_j = (p_j + i p_{j+M}) zeta^j.
// Real polynomials take conjugate values at the other N/2 roots, so products of
// these M values determine the negacyclic product; its inverse gives
// (r_j + i r_{j+M}) zeta^j, untwisted and unfolded.
//
// One wave computes one M = 512-point transform, 8 points per lane, as three
// radix-8 passes (input index j = 64 j2 + 8 j1 + j0, output k = k2 + 8 k1 + 64 k0):
//   layout A  lane = j mod 64, slot = j2       forward input, inverse output
//   layout B  lane = 8 k2 + j0, slot = j1      between the passes
//   layout C  lane = 8 k2 + k1, slot = k0      forward output (digit-reversed order),
//                                              the key's storage order, inverse input
// pass 1: DFT-8 over j2, twiddle w^((8 j1 + j0) k2); pass 2: DFT-8 over j1, twiddle
// w^(8 j0 k1); pass 3: DFT-8 over j0.  The two exchanges go through a private LDS
// slice of the wave with row pitches chosen so that every access spreads evenly over
// the banks (pitch 72 for the A<->B exchange, 65 for B<->C).
// The twist zeta^j, j = 64 j2 + r, splits as zeta^r zeta^(64 j2): the slot factor
// c_j2 = zeta^(64 j2) multiplies the inputs (uniform constants), and the lane factor
// zeta^r commutes with the DFT-8 over j2 and joins pass 1's twiddle, which becomes
// zeta^r w^(r k2) for every k2 (k2 = 0 included); the inverse mirrors both.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vh_kernels.h"

namespace {

constexpr int VT_N = 1024, VT_M = 512;
constexpr int VT_P1 = 72, VT_P2 = 65;          // exchange row pitches (doubles)
constexpr int VT_XS = 8 * VT_P1;               // doubles per exchange array (>= 8 * VT_P2)

// S = -1: X_k = sum_j x_j e^(-2 pi i jk / 8); S = +1: the same with +i (no scaling).
template <int S>
__device__ __forceinline__ void vt_dft4(double &r0, double &i0, double &r1, double &i1, double &r2, double &i2,
                                        double &r3, double &i3)
{
    const double c0r = r0 + r2, c0i = i0 + i2, d0r = r0 - r2, d0i = i0 - i2;
    const double c1r = r1 + r3, c1i = i1 + i3, er = r1 - r3, ei = i1 - i3;
    const double d1r = S > 0 ? -ei : ei, d1i = S > 0 ? er : -er;   // (r1 - r3) * (S i)
    r0 = c0r + c1r;
    i0 = c0i + c1i;
    r2 = c0r - c1r;
    i2 = c0i - c1i;
    r1 = d0r + d1r;
    i1 = d0i + d1i;
    r3 = d0r - d1r;
    i3 = d0i - d1i;
}

template <int S>
__device__ __forceinline__ void vt_dft8(double *re, double *im)
{
    const double h = 0.70710678118654752440;
    double ar[4], ai[4], br[4], bi[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        ar[j] = re[j] + re[j + 4];
        ai[j] = im[j] + im[j + 4];
    }
    const double d0r = re[0] - re[4], d0i = im[0] - im[4];
    const double d1r = re[1] - re[5], d1i = im[1] - im[5];
    const double d2r = re[2] - re[6], d2i = im[2] - im[6];
    const double d3r = re[3] - re[7], d3i = im[3] - im[7];
    br[0] = d0r;
    bi[0] = d0i;
    br[1] = (S > 0 ? d1r - d1i : d1r + d1i) * h;          // * (1 + S i) / sqrt 2
    bi[1] = (S > 0 ? d1i + d1r : d1i - d1r) * h;
    br[2] = S > 0 ? -d2i : d2i;                           // * (S i)
    bi[2] = S > 0 ? d2r : -d2r;
    br[3] = (S > 0 ? -d3r - d3i : d3i - d3r) * h;         // * (-1 + S i) / sqrt 2
    bi[3] = (S > 0 ? d3r - d3i : -d3r - d3i) * h;
    vt_dft4<S>(ar[0], ai[0], ar[1], ai[1], ar[2], ai[2], ar[3], ai[3]);
    vt_dft4<S>(br[0], bi[0], br[1], bi[1], br[2], bi[2], br[3], bi[3]);
#pragma unroll
    for (int m = 0; m < 4; m++) {
        re[2 * m] = ar[m];
        im[2 * m] = ai[m];
        re[2 * m + 1] = br[m];
        im[2 * m + 1] = bi[m];
    }
}

__device__ __forceinline__ void vt_cmul(double &r, double &i, double wr, double wi)
{
    const double t = __fma_rn(r, wr, -(i * wi));
    i = __fma_rn(r, wi, i * wr);
    r = t;
}

// Twiddle tables in LDS, filled once per workgroup: t1[k 64 + r] = zeta^r w^(r k)
// (pass 1 with the lane's twist factor, lane r, k = 0..7), t2[(s-1) 8 + a] = w^(8 a s)
// (pass 2, a = lane mod 8, s = 1..7).  Per lane the reads are consecutive 16-byte words
// (t1) or broadcasts (t2).
constexpr int VT_T1 = 8 * 64, VT_T2 = 7 * 8;

__device__ __forceinline__ void vt_fill_tw(double2 *t1, double2 *t2, const double2 *consts, int tid, int nthreads)
{
    for (int e = tid; e < VT_T1 + VT_T2; e += nthreads) {
        if (e < VT_T1) {
            t1[e] = consts[e];
        } else {
            const int s = (e - VT_T1) / 8 + 1, a = (e - VT_T1) % 8;
            t2[e - VT_T1] = consts[512 + a * 8 + s];
        }
    }
}

// Forward transform: layout A in, layout C out.
// Pass 1's twiddles zeta^r w^(r k) are formed by powers of the lane's w^r from zeta^r
// (z = {zeta^r, w^r}): seven products instead of eight LDS reads of 16 bytes.
__device__ __forceinline__ void vt_fwd(double *re, double *im, double *xr, double *xi, const double2 (&z)[2],
                                       const double2 *t2, int lane)
{
    const int hi = lane >> 3, lo = lane & 7;
    vt_dft8<-1>(re, im);
    {
        double cr = z[0].x, ci = z[0].y;
#pragma unroll
        for (int s = 0; s < 8; s++) {
            vt_cmul(re[s], im[s], cr, ci);
            if (s < 7) vt_cmul(cr, ci, z[1].x, z[1].y);
        }
    }
#pragma unroll
    for (int s = 0; s < 8; s++) {
        xr[VT_P1 * s + lane] = re[s];
        xi[VT_P1 * s + lane] = im[s];
    }
    vh_wave_sync();
#pragma unroll
    for (int s = 0; s < 8; s++) {
        re[s] = xr[VT_P1 * hi + 8 * s + lo];
        im[s] = xi[VT_P1 * hi + 8 * s + lo];
    }
    vh_wave_sync();
    vt_dft8<-1>(re, im);
#pragma unroll
    for (int s = 1; s < 8; s++) {
        const double2 t = t2[(s - 1) * 8 + lo];
        vt_cmul(re[s], im[s], t.x, t.y);
    }
#pragma unroll
    for (int s = 0; s < 8; s++) {
        xr[VT_P2 * lo + 8 * hi + s] = re[s];
        xi[VT_P2 * lo + 8 * hi + s] = im[s];
    }
    vh_wave_sync();
#pragma unroll
    for (int s = 0; s < 8; s++) {
        re[s] = xr[VT_P2 * s + lane];
        im[s] = xi[VT_P2 * s + lane];
    }
    vh_wave_sync();
    vt_dft8<-1>(re, im);
}

// Inverse transform (unscaled): layout C in, layout A out.
__device__ __forceinline__ void vt_inv(double *re, double *im, double *xr, double *xi, const double2 (&z)[2],
                                       const double2 *t2, int lane)
{
    const int hi = lane >> 3, lo = lane & 7;
    vt_dft8<1>(re, im);
#pragma unroll
    for (int s = 1; s < 8; s++) {
        const double2 t = t2[(s - 1) * 8 + lo];
        vt_cmul(re[s], im[s], t.x, -t.y);
    }
#pragma unroll
    for (int s = 0; s < 8; s++) {
        xr[VT_P2 * s + lane] = re[s];
        xi[VT_P2 * s + lane] = im[s];
    }
    vh_wave_sync();
#pragma unroll
    for (int s = 0; s < 8; s++) {
        re[s] = xr[VT_P2 * lo + 8 * hi + s];
        im[s] = xi[VT_P2 * lo + 8 * hi + s];
    }
    vh_wave_sync();
    vt_dft8<1>(re, im);
#pragma unroll
    for (int s = 0; s < 8; s++) {
        xr[VT_P1 * hi + 8 * s + lo] = re[s];
        xi[VT_P1 * hi + 8 * s + lo] = im[s];
    }
    vh_wave_sync();
#pragma unroll
    for (int s = 0; s < 8; s++) {
        re[s] = xr[VT_P1 * s + lane];
        im[s] = xi[VT_P1 * s + lane];
    }
    vh_wave_sync();
    {
        double cr = z[0].x, ci = z[0].y;
#pragma unroll
        for (int s = 0; s < 8; s++) {
            vt_cmul(re[s], im[s], cr, -ci);
            if (s < 7) vt_cmul(cr, ci, z[1].x, z[1].y);
        }
    }
    vt_dft8<1>(re, im);
}

// {zeta^r, w^r} of lane r from the pass-1 table (w^r = t1[64 + r] conj(t1[r]), |zeta^r| = 1).
__device__ __forceinline__ void vt_lane_roots(double2 (&z)[2], const double2 *t1, int lane)
{
    const double2 a = t1[lane], b = t1[64 + lane];
    z[0] = a;
    double r = b.x, i = b.y;
    vt_cmul(r, i, a.x, -a.y);
    z[1] = make_double2(r, i);
}

// The slot factors c_s = zeta^(64 s) = e^(i pi s / 16), correctly rounded (uniform
// constants, no loads): applied to the forward input, conjugated to the inverse output.
__device__ constexpr double VT_CS[8][2] = {
    {1.0, 0.0},
    {0x1.f6297cff75cb0p-1, 0x1.8f8b83c69a60bp-3},
    {0x1.d906bcf328d46p-1, 0x1.87de2a6aea963p-2},
    {0x1.a9b66290ea1a3p-1, 0x1.1c73b39ae68c8p-1},
    {0x1.6a09e667f3bcdp-1, 0x1.6a09e667f3bcdp-1},
    {0x1.1c73b39ae68c8p-1, 0x1.a9b66290ea1a3p-1},
    {0x1.87de2a6aea963p-2, 0x1.d906bcf328d46p-1},
    {0x1.8f8b83c69a60bp-3, 0x1.f6297cff75cb0p-1},
};

__device__ __forceinline__ void vt_slot_twist(double *re, double *im, bool conj)
{
#pragma unroll
    for (int s = 1; s < 8; s++) vt_cmul(re[s], im[s], VT_CS[s][0], conj ? -VT_CS[s][1] : VT_CS[s][1]);
}

// Layout A input of a polynomial given by its coefficients as signed 32-bit values:
// lane r, slot s holds (p_j + i p_(j+M)) c_s, j = 64 s + r (the lane's twist factor
// comes with pass 1).
__device__ __forceinline__ void vt_fold(double *re, double *im, const uint32_t *p, int lane)
{
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int j = 64 * s + lane;
        re[s] = (double)(int32_t)p[j];
        im[s] = (double)(int32_t)p[j + VT_M];
    }
    vt_slot_twist(re, im, false);
}

// x mod 2^32 of an integer-valued double with |x| < 2^51 (1.5 * 2^52 puts the
// rounded integer's low bits at the bottom of the mantissa).
__device__ __forceinline__ uint32_t vt_round_u32(double x)
{
    return (uint32_t)__double_as_longlong(x + 6755399441055744.0);
}

// round(x * 2N / 2^32) with ties to even, as a double quotient rounds in Python
// (cggi16.py:62,77; exact here since 2N / 2^32 is a power of two), mod 2N.
__device__ __forceinline__ uint32_t vt_mod_switch(uint64_t x, int shift, uint32_t two_n)
{
    const uint64_t q = x >> shift, rem = x & ((1ull << shift) - 1), half = 1ull << (shift - 1);
    const uint64_t r = q + ((rem > half || (rem == half && (q & 1))) ? 1 : 0);
    return (uint32_t)(r & (two_n - 1));
}

// coefficient j of X^a * p (a in [0, 2N)), mod 2^32
__device__ __forceinline__ uint32_t vt_rot(const uint32_t *p, int j, uint32_t a)
{
    const uint32_t t = ((uint32_t)j - a) & (2 * VT_N - 1);
    const uint32_t v = p[t & (VT_N - 1)], neg = 0u - (t >> 10);   // t >> 10: 1 when t >= N
    return (v ^ neg) - neg;
}

}  // namespace

// ---------------------------------------------------------------------------
// Key transform: one wave per polynomial (signed 32-bit coefficients) to layout C,
// stored [poly][slot][lane] (double2) and scaled by 1/M (exact: a power of two), so
// the blind rotation's inverse transforms need no scaling.  natural != 0: unscaled,
// in natural order [poly][k] (the transform check).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_torus_fft(const uint32_t *__restrict__ in, double2 *__restrict__ out,
                                                   const double2 *__restrict__ consts, uint64_t count, int natural)
{
    __shared__ double lds[4][2][VT_XS];
    __shared__ double2 t1[VT_T1], t2[VT_T2];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    vt_fill_tw(t1, t2, consts, threadIdx.x, 256);
    __syncthreads();
    const uint64_t poly = (uint64_t)blockIdx.x * 4 + w;
    if (poly >= count) return;   // whole waves leave; the exchanges are wave-private
    double2 z[2];
    vt_lane_roots(z, t1, lane);
    double re[8], im[8];
    vt_fold(re, im, in + poly * VT_N, lane);
    vt_fwd(re, im, lds[w][0], lds[w][1], z, t2, lane);
    double2 *o = out + poly * VT_M;
    const double scale = natural ? 1.0 : 1.0 / VT_M;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int idx = natural ? (lane >> 3) + 8 * (lane & 7) + 64 * s : s * 64 + lane;
        o[idx] = make_double2(re[s] * scale, im[s] * scale);
    }
}

hipError_t vh_launch_torus_fft(hipStream_t st, const uint32_t *in, double2 *out, const double2 *consts, uint64_t count,
                               int natural)
{
    if (!count) return hipSuccess;
    VH_KL(k_torus_fft, dim3((unsigned)((count + 3) / 4)), dim3(256), 0, st, in, out, consts, count, natural);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Blind rotation + extraction: one workgroup of two waves per LWE sample; wave c owns
// component c of the accumulator (0: the mask a, 1: the body b), kept in LDS as
// 32-bit words.  Per key index i with a non-zero amount:
//   diff = X^a_i * acc_c - acc_c  (registers, 16 coefficients per lane)
//   for each level l: digit polynomial -> forward transform -> accumulate its products
//     with key rows (c * ell + l, output component 0 and 1) into two partial sums;
//   the waves swap the partial of the other's component through LDS (two barriers),
//   so wave c holds the sum over all 2 * ell rows for its component in a fixed order
//   (wave 0's rows, then wave 1's); inverse transform, untwist (the key rows carry
//   the 1/M of the inverse), round, and add into acc_c.
// The key: n x (2 ell rows) x (2 components) x [slot][lane] double2 (layout C).
// ---------------------------------------------------------------------------
template <int ELL>   // levels (1..4)
__global__ void __launch_bounds__(128, 2) k_torus_blind_rotate(vt_br_args A)
{
    __shared__ uint32_t acc[2][VT_N];
    __shared__ double lds[2][2][VT_XS];
    __shared__ double2 t1[VT_T1], t2[VT_T2];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const uint64_t k = blockIdx.x;
    const uint32_t *in = A.lwe + k * (uint64_t)(A.n + 1);
    const uint32_t two_n = 2 * VT_N;
    const int shift = 32 - 11;   // log2(2^32 / 2N)
    double *xr = lds[w][0], *xi = lds[w][1];

    vt_fill_tw(t1, t2, A.consts, threadIdx.x, 128);

    // acc = (0, X^-rot_b * tv): rot_b from b + floor(2^32 / (4 torus_base)) (cggi16.py:60-71)
    const uint32_t rot_b = vt_mod_switch((uint64_t)in[A.n] + A.offset, shift, two_n);
    const uint32_t rot_tv = (two_n - rot_b) & (two_n - 1);
    uint32_t *mine = acc[w];
    for (int j = lane; j < VT_N; j += 64) mine[j] = w == 0 ? 0u : vt_rot(A.tv, j, rot_tv);
    __syncthreads();
    double2 z[2];
    vt_lane_roots(z, t1, lane);

    // Digits: with y = x + dec_off, field_l = (y >> sh_l) mod 2^B and d_l = field_l - 2^(B-1),
    // which is the B-bit sign extension of field_l ^ 2^(B-1): one xor per coefficient, then
    // one signed bit-field extract per level.
    const uint32_t B = A.base_log;
    uint32_t xmask = 0;
#pragma unroll
    for (int l = 1; l <= ELL; l++) xmask |= (1u << (B - 1)) << (32 - B * l);
    const int rows = 2 * ELL;
    for (uint32_t i = 0; i < A.n; i++) {
        const uint32_t ai = vt_mod_switch(in[i], shift, two_n);
        if (ai == 0) continue;
        uint32_t dlo[8], dhi[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int j = 64 * s + lane;
            dlo[s] = (vt_rot(mine, j, ai) - mine[j] + A.dec_off) ^ xmask;
            dhi[s] = (vt_rot(mine, j + VT_M, ai) - mine[j + VT_M] + A.dec_off) ^ xmask;
        }
        // own: this wave's output component (w), other: the other wave's (1 - w)
        double por[8], poi[8], pxr[8], pxi[8];
#pragma unroll
        for (int s = 0; s < 8; s++) por[s] = poi[s] = pxr[s] = pxi[s] = 0.0;
#pragma unroll
        for (int l = 0; l < ELL; l++) {
            // the level's key rows are loaded first, their latency covered by the transform
            const double2 *K = A.bsk + (((uint64_t)i * rows + (uint64_t)w * ELL + l) * 2) * VT_M + lane;
            const double2 *Ko = K + (w ? VT_M : 0), *Kx = K + (w ? 0 : VT_M);
            double2 kor[8], kxr[8];
#pragma unroll
            for (int s = 0; s < 8; s++) {
                kor[s] = Ko[s * 64];
                kxr[s] = Kx[s * 64];
            }
            const uint32_t sh = 32 - B * (l + 1);
            double re[8], im[8];
#pragma unroll
            for (int s = 0; s < 8; s++) {
                re[s] = (double)(int)__builtin_amdgcn_sbfe((int)dlo[s], sh, B);
                im[s] = (double)(int)__builtin_amdgcn_sbfe((int)dhi[s], sh, B);
            }
            vt_slot_twist(re, im, false);
            vt_fwd(re, im, xr, xi, z, t2, lane);
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const double2 ko = kor[s], kx = kxr[s];
                por[s] = __fma_rn(re[s], ko.x, por[s]);
                por[s] = __fma_rn(-im[s], ko.y, por[s]);
                poi[s] = __fma_rn(re[s], ko.y, poi[s]);
                poi[s] = __fma_rn(im[s], ko.x, poi[s]);
                pxr[s] = __fma_rn(re[s], kx.x, pxr[s]);
                pxr[s] = __fma_rn(-im[s], kx.y, pxr[s]);
                pxi[s] = __fma_rn(re[s], kx.y, pxi[s]);
                pxi[s] = __fma_rn(im[s], kx.x, pxi[s]);
            }
        }
        // swap partials; each output is wave 0's partial plus wave 1's (addition commutes)
#pragma unroll
        for (int s = 0; s < 8; s++) {
            xr[s * 64 + lane] = pxr[s];
            xi[s * 64 + lane] = pxi[s];
        }
        __syncthreads();
        const double *yr = lds[1 - w][0], *yi = lds[1 - w][1];
        double re[8], im[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            re[s] = por[s] + yr[s * 64 + lane];
            im[s] = poi[s] + yi[s * 64 + lane];
        }
        __syncthreads();
        vt_inv(re, im, xr, xi, z, t2, lane);
        vt_slot_twist(re, im, true);
#pragma unroll
        for (int s = 0; s < 8; s++) {   // the key carries the 1/M of the inverse
            const int j = 64 * s + lane;
            mine[j] += vt_round_u32(re[s]);
            mine[j + VT_M] += vt_round_u32(im[s]);
        }
        vh_wave_sync();
    }

    if (A.acc_out) {
        uint32_t *o = A.acc_out + k * 2 * VT_N + (uint64_t)w * VT_N;
        for (int j = lane; j < VT_N; j += 64) o[j] = mine[j];
    }
    if (A.out) {
        // coefficient 0 of the phase b - a s: a'_0 = a_0, a'_j = -a_(N-j); b' = b_0
        uint32_t *o = A.out + k * (VT_N + 1);
        if (w == 0)
            for (int j = lane; j < VT_N; j += 64) o[j] = j == 0 ? mine[0] : 0u - mine[VT_N - j];
        else if (lane == 0)
            o[VT_N] = mine[0];
    }
}

hipError_t vh_launch_torus_blind_rotate(hipStream_t st, const vt_br_args &A, uint64_t count)
{
    if (!count) return hipSuccess;
    switch (A.ell) {
    case 1: VH_KL(k_torus_blind_rotate<1>, dim3((unsigned)count), dim3(128), 0, st, A); break;
    case 2: VH_KL(k_torus_blind_rotate<2>, dim3((unsigned)count), dim3(128), 0, st, A); break;
    case 3: VH_KL(k_torus_blind_rotate<3>, dim3((unsigned)count), dim3(128), 0, st, A); break;
    case 4: VH_KL(k_torus_blind_rotate<4>, dim3((unsigned)count), dim3(128), 0, st, A); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
