// vh_device.h -- device-side modular arithmetic and the structs shared by the
// host runtime (vh_engine.cpp) and the kernels (vh_kernels.hip) of the hip engine.
//
// Two exact arithmetics are used, both producing the canonical residue in
// [0, q) that the reference produces (modules/arith/c/src/mod_portable.c:7-161):
//
//  * FP64 exact-integer arithmetic for q < 2^50 (every prime Ring.gen_prime makes
//    at 50 bits, the benchmarked rings').  Residues are held as integer-valued doubles.  A product
//    b*w (|b| < 2^52, |w| <= q/2) is split exactly into hi + lo with one FMA
//    (TwoProduct); the quotient estimate qt = round(b * (w/q)) comes from one
//    FMA against the 1.5*2^52 rounding constant, so |b*w/q - qt| <= 3/4 and
//    r = (hi - qt*q) + lo is an exact integer with |r| <= 3q/4.  On gfx950 the
//    FP64 FMA issues at full VALU rate, so a butterfly costs ~8 full-rate
//    instructions versus ~25 for a 64-bit Shoup butterfly built from 32-bit
//    integer multiplies.
//  * 64-bit integer Shoup / Barrett arithmetic for any q < 2^62 (the wider
//    primes the reference's tests sweep, modules/arith/c/test/test_mod.c:190-205).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// An exchange through LDS among the lanes of one wave: a wave's LDS instructions
// execute in order, so only the compiler has to be kept from moving loads and stores
// across the exchange, which the fences do.
__device__ __forceinline__ void vh_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-limb constants, device resident (one per prime of a pool).
struct vh_limb {
    uint64_t q;
    uint64_t mu64;        // floor(2^64 / q): reduction of any uint64
    uint64_t bar_mu;      // floor(2^(2s) / q), s = bit length of q: Barrett for q^2-sized products
    uint32_t bar_s;       // s
    uint32_t fp64;        // 1 when q < 2^50 (FP64 exact path for the NTT)
    double qd, qinv;      // q and RN(1/q)
    uint64_t ninv, ninv_sh;   // n^-1 mod q and floor(ninv * 2^64 / q)
    double ninv_d, ninv_q;    // n^-1 as a centred residue (|.| <= q/2) and RN(ninv_d / q)
    const double2 *tw_fwd;    // FP64 path: {w, RN(w/q)} with w = psi^brv(i) centred, reference ws order
    const double2 *tw_inv;    // FP64 path: {w, RN(w/q)} with w = psi^-brv(i) centred
    const ulonglong2 *twi_fwd;  // integer path: {w, floor(w 2^64 / q)}
    const ulonglong2 *twi_inv;
    // FP64 path, n = 2^13 .. 2^16, s0 = log2(n) - 8: row factors {b, RN(b/q)} (b centred) with
    //   tw[2^s + (sub << s') + local] == tw[2^s' + local] * b[sub][s']  (mod q),  s = s0 + s',
    // for sub < 2^s0, s' < 8, local < 2^s'; [sub][s'] at index 8 sub + s'.  Null otherwise.
    const double2 *rowf_fwd;
    const double2 *rowf_inv;
};

// One unit of work: a row of n coefficients of one limb.
struct vh_item {
    const uint64_t *src;
    uint64_t *dst;
    uint32_t limb;
    uint32_t aux;         // op-specific (e.g. a second limb index)
    const uint64_t *src2; // op-specific second operand
    uint64_t scalar;      // op-specific scalar (already reduced mod q where relevant)
    uint64_t scalar_sh;   // Shoup companion of `scalar` when used as a multiplier
};

// Global address space: pointers read from work lists are otherwise generic
// (flat) pointers, whose loads wait on both the vector and the LDS counters.
#define VH_GLOBAL __attribute__((address_space(1)))
template <typename T> __device__ __forceinline__ const VH_GLOBAL T *vh_global(const T *p)
{
    return (const VH_GLOBAL T *)p;
}
template <typename T> __device__ __forceinline__ VH_GLOBAL T *vh_global(T *p) { return (VH_GLOBAL T *)p; }

#define VH_MAGIC 6755399441055744.0   /* 1.5 * 2^52 */
#define VH_TWO52 4503599627370496.0   /* 2^52 */

// ---------------------------------------------------------------- FP64 path

// x < 2^52 -> exact double (exponent splice: 2^52 + x, minus 2^52).
__device__ __forceinline__ double vh_u2d(uint64_t x)
{
    return __longlong_as_double((long long)(x | 0x4330000000000000ULL)) - VH_TWO52;
}

// r integer-valued double in [0, 2^52) -> uint64.
__device__ __forceinline__ uint64_t vh_d2u(double r)
{
    return (uint64_t)__double_as_longlong(r + VH_TWO52) & 0x000FFFFFFFFFFFFFULL;
}

// b * w mod q up to a multiple of q: |b| < 2^52, |w| <= q/2, wq = RN(w/q).
// Returns an exact integer r == b*w (mod q) with |r| <= 3q/4.
__device__ __forceinline__ double vh_mulmod_tw(double b, double w, double wq, double q)
{
    double hi = b * w;
    double lo = __fma_rn(b, w, -hi);
    double qt = __fma_rn(b, wq, VH_MAGIC) - VH_MAGIC;
    double r = __fma_rn(-qt, q, hi);
    return r + lo;
}

// Centre: |x| < 2^53 -> r == x (mod q), |r| <= q/2 + 1.
__device__ __forceinline__ double vh_center(double x, double q, double qinv)
{
    double qt = __fma_rn(x, qinv, VH_MAGIC) - VH_MAGIC;
    return __fma_rn(-qt, q, x);
}

// Canonical: |x| < 2^53 -> x mod q in [0, q).
__device__ __forceinline__ double vh_canon(double x, double q, double qinv)
{
    double r = vh_center(x, q, qinv);
    return r < 0.0 ? r + q : r;
}

// ------------------------------------------------------------- integer path

__device__ __forceinline__ uint64_t vh_mulhi(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// a * w mod q with Shoup's precomputed wsh = floor(w 2^64 / q); any a < 2^64, w < q.
__device__ __forceinline__ uint64_t vh_shoup(uint64_t a, uint64_t w, uint64_t wsh, uint64_t q)
{
    uint64_t r = a * w - vh_mulhi(a, wsh) * q;
    return r >= q ? r - q : r;
}

// x mod q for any uint64 x (mu64 = floor(2^64/q); the estimate is short by at most one q).
__device__ __forceinline__ uint64_t vh_reduce64(uint64_t x, uint64_t q, uint64_t mu64)
{
    uint64_t r = x - vh_mulhi(x, mu64) * q;
    return r >= q ? r - q : r;
}

// (hi:lo) mod q for hi:lo < q^2 (Barrett with mu = floor(2^(2s)/q), s = bit
// length of q, 2 <= s <= 63, 3q < 2^64).  The quotient estimate is short by at
// most 2, so two conditional subtractions finish the reduction.
__device__ __forceinline__ uint64_t vh_barrett128(uint64_t hi, uint64_t lo, uint64_t q, uint64_t mu, uint32_t s)
{
    const uint64_t t = (lo >> (s - 1)) | (hi << (65 - s));          // x >> (s - 1) < 2^(s+1)
    const uint64_t ph = vh_mulhi(t, mu), pl = t * mu;
    const uint64_t qh = (s >= 63) ? ph : ((ph << (63 - s)) | (pl >> (s + 1)));  // (t mu) >> (s + 1)
    uint64_t r = lo - qh * q;
    if (r >= q) r -= q;
    if (r >= q) r -= q;
    return r;
}

__device__ __forceinline__ uint64_t vh_mulmod(uint64_t a, uint64_t b, const vh_limb &L)
{
    return vh_barrett128(vh_mulhi(a, b), a * b, L.q, L.bar_mu, L.bar_s);
}

__device__ __forceinline__ uint64_t vh_addmod(uint64_t a, uint64_t b, uint64_t q)
{
    uint64_t s = a + b;
    return s >= q ? s - q : s;
}

__device__ __forceinline__ uint64_t vh_submod(uint64_t a, uint64_t b, uint64_t q)
{
    return a >= b ? a - b : a + q - b;
}
