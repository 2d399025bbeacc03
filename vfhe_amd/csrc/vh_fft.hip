// Complex FP64 FFT of the CKKS encoder (modules/arith/c/src/complex.c) on gfx950.
//
// A row is 2n doubles, real parts then imaginary parts (the reference's
// layout).  CT_NR (natural in, bit-reversed out) and GS_RN (bit-reversed in,
// natural out) use the reference's twiddle table ws[m + i] for the i-th group
// of the stage with m groups (complex.c:389-453).  Every butterfly computes the
// reference's expression in the reference's order, so results are bit-identical
// to the reference engine whose complex product they follow:
//   FMA = 1: the avx512ifma engine's product (complex.c:52-59):
//            im = fma(a.re, w.im, w.re * a.im), re = fma(a.re, w.re, -(a.im * w.im));
//   FMA = 0: the portable engine's product (complex.c:355-361), two roundings each.
// The file is compiled with -ffp-contract=off, so no other product is fused.
// Stages are independent of each other's schedule, so passes may group them
// freely: n <= 2048 runs in one LDS pass, larger rows in a column pass and a row
// pass.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vh_kernels.h"

template <int FMA>
__device__ __forceinline__ double2 vh_cmul(double2 a, double2 w)
{
    double2 c;
    if (FMA) {
        c.y = __fma_rn(a.x, w.y, w.x * a.y);
        c.x = __fma_rn(a.x, w.x, -(a.y * w.y));
    } else {
        c.y = a.x * w.y + w.x * a.y;
        c.x = a.x * w.x - a.y * w.y;
    }
    return c;
}

// One pass of R stages over G sub-transforms of S = 2^R points per block.
// DIR 0: CT_NR stages s0 .. s0+R-1; DIR 1: GS_RN stages s0+R-1 .. s0.
template <int DIR, int FMA>
__global__ void __launch_bounds__(256)
k_fft_pass(double *const *__restrict__ rows, const double2 *__restrict__ ws, int logn, int s0, int R, int col, int G)
{
    extern __shared__ double2 vh_fsm[];
    double *re = rows[blockIdx.y];
    const size_t n = (size_t)1 << logn;
    double *im = re + n;
    const int S = 1 << R, halfS = S >> 1;
    const size_t nsub = (size_t)1 << (logn - R);
    const size_t sub0 = (size_t)blockIdx.x * G;

    for (int e = threadIdx.x; e < G * S; e += blockDim.x) {
        int g, k;
        if (col) { g = e % G; k = e / G; } else { g = e / S; k = e % S; }
        const size_t j = col ? (sub0 + g) + (size_t)k * nsub : (sub0 + g) * S + k;
        vh_fsm[g * S + k] = make_double2(re[j], im[j]);
    }
    __syncthreads();
    const int nbf = G * halfS;
    for (int c = 0; c < R; c++) {
        const int st = DIR == 0 ? c : (R - 1 - c);
        const int s = s0 + st;
        const int half = S >> (st + 1);
        for (int b = threadIdx.x; b < nbf; b += blockDim.x) {
            const int g = b / halfS, bi = b % halfS;
            const int k = ((bi & ~(half - 1)) << 1) | (bi & (half - 1));
            const int e0 = g * S + k, e1 = e0 + half;
            const size_t j = col ? (sub0 + g) + (size_t)k * nsub : (sub0 + g) * S + k;
            const double2 w = ws[((size_t)1 << s) + (j >> (logn - s))];
            const double2 u = vh_fsm[e0], v = vh_fsm[e1];
            if (DIR == 0) {
                const double2 V = vh_cmul<FMA>(v, w);
                vh_fsm[e1] = make_double2(u.x - V.x, u.y - V.y);
                vh_fsm[e0] = make_double2(u.x + V.x, u.y + V.y);
            } else {
                const double2 V = make_double2(u.x - v.x, u.y - v.y);
                vh_fsm[e0] = make_double2(u.x + v.x, u.y + v.y);
                vh_fsm[e1] = vh_cmul<FMA>(V, w);
            }
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < G * S; e += blockDim.x) {
        int g, k;
        if (col) { g = e % G; k = e / G; } else { g = e / S; k = e % S; }
        const size_t j = col ? (sub0 + g) + (size_t)k * nsub : (sub0 + g) * S + k;
        const double2 v = vh_fsm[g * S + k];
        re[j] = v.x;
        im[j] = v.y;
    }
}

// complex.c:13-37: v'[brv_prec(i)] = v[i] on the real and the imaginary halves,
// with the reference's 20-bit reversal (x reversed over 20 bits, >> (20 - prec)).
__device__ __forceinline__ uint32_t vh_brv20(uint32_t x, uint32_t prec)
{
    return __builtin_bitreverse32(x << 12) >> (20 - prec);
}

__global__ void __launch_bounds__(256)
k_fft_bitrev(double *const *__restrict__ rows, double *const *__restrict__ tmp, uint64_t n, uint32_t prec)
{
    const double *src = rows[blockIdx.y];
    double *dst = tmp[blockIdx.y];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = vh_brv20((uint32_t)i, prec);
        dst[r] = src[i];
        dst[r + n] = src[i + n];
    }
}

__global__ void __launch_bounds__(256)
k_fft_scale(double *const *__restrict__ rows, double scale, uint64_t count2n)
{
    double *v = rows[blockIdx.y];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count2n; i += (uint64_t)gridDim.x * blockDim.x)
        v[i] *= scale;
}

// complex.c:39-50: (int64) llrint(v) of the first n doubles (round half to even).
__global__ void __launch_bounds__(256)
k_fft_round(double *const *__restrict__ rows, uint64_t *const *__restrict__ outs, uint64_t n)
{
    const double *v = rows[blockIdx.y];
    uint64_t *o = outs[blockIdx.y];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        o[i] = (uint64_t)__double2ll_rn(v[i]);
}

static inline unsigned vh_fblocks(uint64_t n)
{
    const uint64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 256 ? 256 : b));
}

hipError_t vh_launch_fft(hipStream_t st, double *const *rows, unsigned nrows, const double2 *ws, int logn, int dir,
                         int fma)
{
    if (logn < 1 || logn > 20) return hipErrorInvalidValue;
    auto pass = [&](int s0, int R, int col, int G) {
        const int S = 1 << R;
        const unsigned nsub = 1u << (logn - R);
        dim3 grid(nsub / G, nrows);
        int thr = G * S / 2;
        thr = thr < 64 ? 64 : (thr > 256 ? 256 : thr);
        const size_t lds = (size_t)G * S * sizeof(double2);
        if (dir == 0) {
            if (fma) VH_KL((k_fft_pass<0, 1>), grid, dim3(thr), lds, st, rows, ws, logn, s0, R, col, G);
            else VH_KL((k_fft_pass<0, 0>), grid, dim3(thr), lds, st, rows, ws, logn, s0, R, col, G);
        } else {
            if (fma) VH_KL((k_fft_pass<1, 1>), grid, dim3(thr), lds, st, rows, ws, logn, s0, R, col, G);
            else VH_KL((k_fft_pass<1, 0>), grid, dim3(thr), lds, st, rows, ws, logn, s0, R, col, G);
        }
    };
    if (logn <= 11) {
        pass(0, logn, 0, 1);
    } else {
        // column pass over the first R1 stages (strided sub-transforms), row pass over the rest
        const int R2 = 11, R1 = logn - R2;
        const int Gc = (1 << (logn - R1)) < (2048 >> R1) ? (1 << (logn - R1)) : (2048 >> R1);
        if (dir == 0) {
            pass(0, R1, 1, Gc);
            pass(R1, R2, 0, 1);
        } else {
            pass(R1, R2, 0, 1);
            pass(0, R1, 1, Gc);
        }
    }
    return hipGetLastError();
}

hipError_t vh_launch_fft_bitrev(hipStream_t st, double *const *rows, double *const *tmp, unsigned nrows, uint64_t n,
                                uint32_t prec)
{
    VH_KL(k_fft_bitrev, dim3(vh_fblocks(n), nrows), dim3(256), 0, st, rows, tmp, n, prec);
    return hipGetLastError();
}

hipError_t vh_launch_fft_scale(hipStream_t st, double *const *rows, unsigned nrows, double scale, uint64_t count)
{
    VH_KL(k_fft_scale, dim3(vh_fblocks(count), nrows), dim3(256), 0, st, rows, scale, count);
    return hipGetLastError();
}

hipError_t vh_launch_fft_round(hipStream_t st, double *const *rows, uint64_t *const *outs, unsigned nrows, uint64_t n)
{
    VH_KL(k_fft_round, dim3(vh_fblocks(n), nrows), dim3(256), 0, st, rows, outs, n);
    return hipGetLastError();
}
