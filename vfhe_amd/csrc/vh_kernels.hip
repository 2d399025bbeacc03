// vh_kernels.hip -- HIP kernels of the hip engine (gfx950).
//
// Every kernel works on a list of vh_item rows so that one launch covers any
// number of (polynomial, limb) rows: the reference's per-limb loops
// (modules/arith/c/src/rns_polynomial.c:582-614 and friends) become the y
// dimension of a single grid.
#include <climits>
#include <type_traits>
#include "vh_device.h"
#include "vh_kernels.h"

int g_vh_suppress_launch = 0;

// ---------------------------------------------------------------------------
// NTT: one pass of R consecutive stages of the merged negacyclic transform.
//
// Forward = Cooley-Tukey, natural order in, bit-reversed order out; inverse =
// Gentleman-Sande with n^-1 applied at the end (modules/arith/c/src/
// ntt_portable.c:76-135).  Global stage s (0 = largest stride) pairs element j
// with j + n/2^(s+1) and uses twiddle index 2^s + (j >> (logn - s)) of the
// reference's ws table (ntt_portable.c:18-62): tw[idx] = psi^(+-brv(idx)).
//
// A pass owns G sub-transforms of S = 2^R points.  In column mode sub-transform
// c holds elements c + (n/S)*k (the first stages of a two-pass transform);
// otherwise it holds the contiguous block c*S .. c*S + S - 1.  Data sit in LDS
// between stages.  FP64 path: values are exact integers; CT outputs are
// re-centred every 4 stages and GS sums every 2, which keeps every operand of
// vh_mulmod_tw below 2^52 (q < 2^50).  Between passes the data are stored as
// centred doubles (bit patterns) in the destination row.
// ---------------------------------------------------------------------------
template <int DIR, bool FP>
__global__ void __launch_bounds__(256)
k_ntt_pass(const vh_item *__restrict__ items, const vh_limb *__restrict__ limbs,
           int logn, int s0, int R, int col, int G, int first, int last)
{
    extern __shared__ double vh_smem[];
    using T = typename std::conditional<FP, double, uint64_t>::type;
    T *sm = reinterpret_cast<T *>(vh_smem);

    const vh_item it = items[blockIdx.y];
    const vh_limb L = limbs[it.limb];
    const int S = 1 << R, halfS = S >> 1;
    const size_t nsub = (size_t)1 << (logn - R);
    const size_t sub0 = (size_t)blockIdx.x * G;
    const uint64_t *in = first ? it.src : it.dst;
    uint64_t *out = it.dst;
    const double qd = L.qd, qinv = L.qinv;

    for (int e = threadIdx.x; e < G * S; e += blockDim.x) {
        int g, k;
        if (col) { g = e % G; k = e / G; } else { g = e / S; k = e % S; }
        size_t j = col ? (sub0 + g) + (size_t)k * nsub : (sub0 + g) * S + k;
        uint64_t raw = in[j];
        T v;
        if constexpr (FP) v = first ? vh_u2d(raw) : __longlong_as_double((long long)raw);
        else v = raw;
        sm[g * S + k] = v;
    }
    __syncthreads();

    const int nbf = G * halfS;
    for (int c = 0; c < R; c++) {
        const int st = (DIR == 0) ? c : (R - 1 - c);
        const int s = s0 + st;
        const int half = S >> (st + 1);
        for (int b = threadIdx.x; b < nbf; b += blockDim.x) {
            const int g = b / halfS, bi = b % halfS;
            const int k = ((bi & ~(half - 1)) << 1) | (bi & (half - 1));
            const int e0 = g * S + k, e1 = e0 + half;
            const size_t j = col ? (sub0 + g) + (size_t)k * nsub : (sub0 + g) * S + k;
            const size_t widx = ((size_t)1 << s) + (j >> (logn - s));
            T u = sm[e0], v = sm[e1];
            if constexpr (FP) {
                if (DIR == 0) {
                    const double2 tw = L.tw_fwd[widx];
                    double t = vh_mulmod_tw(v, tw.x, tw.y, qd);
                    double a = u + t, d = u - t;
                    if ((c & 3) == 3) { a = vh_center(a, qd, qinv); d = vh_center(d, qd, qinv); }
                    sm[e0] = a; sm[e1] = d;
                } else {
                    const double2 tw = L.tw_inv[widx];
                    double a = u + v;
                    double d = vh_mulmod_tw(u - v, tw.x, tw.y, qd);
                    if (c & 1) a = vh_center(a, qd, qinv);
                    sm[e0] = a; sm[e1] = d;
                }
            } else {
                const uint64_t q = L.q;
                if (DIR == 0) {
                    const ulonglong2 tw = L.twi_fwd[widx];
                    uint64_t t = vh_shoup(v, tw.x, tw.y, q);
                    sm[e0] = vh_addmod(u, t, q);
                    sm[e1] = vh_submod(u, t, q);
                } else {
                    const ulonglong2 tw = L.twi_inv[widx];
                    sm[e0] = vh_addmod(u, v, q);
                    sm[e1] = vh_shoup(vh_submod(u, v, q), tw.x, tw.y, q);
                }
            }
        }
        __syncthreads();
    }

    for (int e = threadIdx.x; e < G * S; e += blockDim.x) {
        int g, k;
        if (col) { g = e % G; k = e / G; } else { g = e / S; k = e % S; }
        size_t j = col ? (sub0 + g) + (size_t)k * nsub : (sub0 + g) * S + k;
        T v = sm[g * S + k];
        uint64_t raw;
        if constexpr (FP) {
            if (last) {
                if (DIR == 1) v = vh_mulmod_tw(v, L.ninv_d, L.ninv_q, qd);
                raw = vh_d2u(vh_canon(v, qd, qinv));
            } else {
                raw = (uint64_t)__double_as_longlong(vh_center(v, qd, qinv));
            }
        } else {
            if (last && DIR == 1) v = vh_shoup(v, L.ninv, L.ninv_sh, L.q);
            raw = v;
        }
        out[j] = raw;
    }
}

// ---------------------------------------------------------------------------
// Register radix-16 NTT pass (FP64 path), the hot kernel for N = 2^13 .. 2^16.
//
// One pass runs R (4 <= R <= 8) consecutive stages of S = 2^R-point
// sub-transforms; TPS = S/16 threads own one sub-transform, 16 values each, and
// a workgroup of 256 threads owns G = 4096/S sub-transforms.  The first four
// stages (halves S/2 .. S/16) run in registers on the stride-TPS element set
// k = t + TPS*m; one LDS exchange regroups the values into contiguous sets
// k = 16t + m for the remaining R-4 stages (halves <= 8).  The inverse runs the
// same schedule backwards.  Global stage s = s0 + s' uses the reference's ws
// index 2^s + toff*2^s' + (k >> (R - s')), toff = 0 for column passes and the
// row index for row passes (see k_ntt_pass for the derivation of the index).
//
// Column passes: lanes run over sub-transforms (adjacent columns), so every
// global access is a contiguous run of G columns.  Row passes: lanes run over
// the 16 threads of a row, loads in the stride ownership are contiguous; the
// contiguous-ownership side goes through an LDS transpose.
// ---------------------------------------------------------------------------
// LDS doubles one task needs: column passes index [k][g] (lanes run over g), row
// passes [g][k] with one pad slot per 16 values; both keep the exchange at the
// 2-way minimum for 64-bit lanes.
// Row passes keep each row's 16 lanes inside one wave and each row's LDS slice
// private to it, so their exchanges only need the wave's own LDS operations to
// stay in program order (vh_wave_sync).

// Streaming accesses of the passes: every row is read once and written once per
// pass, so loads and stores are non-temporal (measured: ~5% on a round trip
// over 64 x 24 rows of 2^16 against ordinary accesses).
#ifndef VH_NT_STORE
#define VH_NT_STORE 1
#endif
#ifndef VH_NT_LOAD
#define VH_NT_LOAD 1
#endif
__device__ __forceinline__ void vh_st(VH_GLOBAL uint64_t *p, uint64_t v)
{
    if (VH_NT_STORE) __builtin_nontemporal_store(v, p);
    else *p = v;
}
__device__ __forceinline__ uint64_t vh_ld(const VH_GLOBAL uint64_t *p)
{
    if (VH_NT_LOAD) return __builtin_nontemporal_load(p);
    return *p;
}

template <int R, bool COL> constexpr int vh_ntt16_lds()
{
    return COL ? (1 << R) * (256 / ((1 << R) / 16) + 1) : (256 / ((1 << R) / 16)) * ((1 << R) + (1 << R) / 16);
}

// One task: the G sub-transforms starting at sub-transform blk*G of one row item.
// COHERENT: read the input with L1-bypassing loads (sc0: a relaxed workgroup-scope
// atomic load on gfx950), served by the L2: it was written by other workgroups of
// the same launch on the same XCD, and this CU's L1 may hold stale lines of it.
// SHARED: the input row is read by many items (the ModUp lift: one digit row
// transformed modulo every extended prime), so it is loaded through the caches.
// NTS: rows loaded and results stored non-temporally (streamed past the caches);
// false loads and stores them as ordinary L2-allocating accesses, so the rows stay
// in the L2 for a reader on the same XCD (a non-temporal load leaves its line at
// streaming priority, and an ordinary store into that line does not raise it).
template <int DIR, int R, bool COL, bool FIRST, bool LAST, bool COHERENT, bool SHARED = false, bool NTS = true>
__device__ __forceinline__ void ntt16_task(const vh_item &it, const vh_limb &L, int logn, int s0, size_t blk,
                                           double *__restrict__ lds)
{
    constexpr int S = 1 << R, TPS = S / 16, G = 256 / TPS;
    const int tid = threadIdx.x;
    const int g = COL ? (tid % G) : (tid / TPS);
    const int t = COL ? (tid / G) : (tid % TPS);
    const double q = L.qd, qinv = L.qinv;
    const size_t nsub = (size_t)1 << (logn - R);
    const size_t sub = blk * G + g;
    const VH_GLOBAL uint64_t *in = vh_global(FIRST ? it.src : it.dst);
    VH_GLOBAL uint64_t *out = vh_global(it.dst);
    const VH_GLOBAL double *tw = vh_global(reinterpret_cast<const double *>(DIR == 0 ? L.tw_fwd : L.tw_inv));
    const size_t toff = COL ? 0 : sub;

    auto gidx = [&](int k) -> size_t { return COL ? sub + (size_t)k * nsub : sub * S + (size_t)k; };
    auto lidx = [&](int k) -> int { return COL ? k * (G + 1) + g : g * (S + S / 16) + k + (k >> 4); };
    auto load = [&](size_t j) -> double {
        const uint64_t raw = COHERENT ? __scoped_atomic_load_n(in + j, __ATOMIC_RELAXED, __MEMORY_SCOPE_WRKGRP)
                                      : ((SHARED || !NTS) ? in[j] : vh_ld(in + j));
        return FIRST ? vh_u2d(raw) : __longlong_as_double((long long)raw);
    };
    auto st = [&](VH_GLOBAL uint64_t *p, uint64_t x) {
        if (NTS) vh_st(p, x);
        else *p = x;
    };
    auto finish = [&](double x) -> uint64_t {
        return LAST ? vh_d2u(vh_canon(x, q, qinv)) : (uint64_t)__double_as_longlong(vh_center(x, q, qinv));
    };
    auto twid = [&](int sp, size_t local) -> double2 {
        const size_t i = ((size_t)1 << (s0 + sp)) + (toff << sp) + local;
        return make_double2(tw[2 * i], tw[2 * i + 1]);
    };

    double v[16];
    if (DIR == 0) {
        // ---- stride ownership, stages s' = 0..3 ----
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = load(gidx(t + TPS * m));
#pragma unroll
        for (int sp = 0; sp < 4; sp++) {
            const int hm = 8 >> sp;
#pragma unroll
            for (int m = 0; m < 16; m++) {
                if (m & hm) continue;
                const double2 w = twid(sp, (size_t)(m >> (4 - sp)));
                const double x = vh_mulmod_tw(v[m + hm], w.x, w.y, q);
                const double u = v[m];
                v[m] = u + x;
                v[m + hm] = u - x;
            }
        }
        if (R > 4) {
#pragma unroll
            for (int m = 0; m < 16; m++) lds[lidx(t + TPS * m)] = vh_center(v[m], q, qinv);
            if (COL) __syncthreads(); else vh_wave_sync();
#pragma unroll
            for (int m = 0; m < 16; m++) v[m] = lds[lidx(16 * t + m)];
            // ---- contiguous ownership, stages s' = 4..R-1 ----
#pragma unroll
            for (int sp = 4; sp < R; sp++) {
                const int hm = 1 << (R - sp - 1);
#pragma unroll
                for (int m = 0; m < 16; m++) {
                    if (m & hm) continue;
                    const double2 w = twid(sp, (size_t)((16 * t + m) >> (R - sp)));
                    const double x = vh_mulmod_tw(v[m + hm], w.x, w.y, q);
                    const double u = v[m];
                    v[m] = u + x;
                    v[m + hm] = u - x;
                }
            }
        }
        if (R == 4) {
#pragma unroll
            for (int m = 0; m < 16; m++) st(out + gidx(t + TPS * m), finish(v[m]));
        } else if (COL) {
#pragma unroll
            for (int m = 0; m < 16; m++) st(out + gidx(16 * t + m), finish(v[m]));
        } else {
            // row pass: back to the stride ownership through LDS, so each store
            // instruction writes whole 128-B runs
            vh_wave_sync();
#pragma unroll
            for (int m = 0; m < 16; m++) lds[lidx(16 * t + m)] = __longlong_as_double((long long)finish(v[m]));
            vh_wave_sync();
#pragma unroll
            for (int m = 0; m < 16; m++)
                st(out + gidx(t + TPS * m), (uint64_t)__double_as_longlong(lds[lidx(t + TPS * m)]));
        }
    } else {
        // ---- load into contiguous ownership ----
        if (R == 4) {
#pragma unroll
            for (int m = 0; m < 16; m++) v[m] = load(gidx(t + TPS * m));
        } else if (COL) {
#pragma unroll
            for (int m = 0; m < 16; m++) v[m] = load(gidx(16 * t + m));
        } else {
            // row pass: whole 128-B runs per load instruction, regrouped through LDS
#pragma unroll
            for (int m = 0; m < 16; m++) lds[lidx(t + TPS * m)] = load(gidx(t + TPS * m));
            vh_wave_sync();
#pragma unroll
            for (int m = 0; m < 16; m++) v[m] = lds[lidx(16 * t + m)];
        }
        int c = 0;  // stages executed in this pass: GS sums are re-centred after every second one
        if (R > 4) {
#pragma unroll
            for (int sp = R - 1; sp >= 4; sp--) {
                const int hm = 1 << (R - sp - 1);
#pragma unroll
                for (int m = 0; m < 16; m++) {
                    if (m & hm) continue;
                    const double2 w = twid(sp, (size_t)((16 * t + m) >> (R - sp)));
                    const double u = v[m], x = v[m + hm];
                    double a = u + x;
                    if (c & 1) a = vh_center(a, q, qinv);
                    v[m] = a;
                    v[m + hm] = vh_mulmod_tw(u - x, w.x, w.y, q);
                }
                c++;
            }
            if (!COL) vh_wave_sync();  // LDS is reused: every read of the load transpose is done
#pragma unroll
            for (int m = 0; m < 16; m++) lds[lidx(16 * t + m)] = v[m];
            if (COL) __syncthreads(); else vh_wave_sync();
#pragma unroll
            for (int m = 0; m < 16; m++) v[m] = lds[lidx(t + TPS * m)];
        }
#pragma unroll
        for (int sp = 3; sp >= 0; sp--) {
            const int hm = 8 >> sp;
#pragma unroll
            for (int m = 0; m < 16; m++) {
                if (m & hm) continue;
                const double2 w = twid(sp, (size_t)(m >> (4 - sp)));
                const double u = v[m], x = v[m + hm];
                double a = u + x;
                if (c & 1) a = vh_center(a, q, qinv);
                v[m] = a;
                v[m + hm] = vh_mulmod_tw(u - x, w.x, w.y, q);
            }
            c++;
        }
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const double x = LAST ? vh_mulmod_tw(v[m], L.ninv_d, L.ninv_q, q) : v[m];
            st(out + gidx(t + TPS * m), finish(x));
        }
    }
}

template <int DIR, int R, bool COL, bool FIRST, bool LAST, bool SHARED>
__global__ void __launch_bounds__(256)
k_ntt16(const vh_item *__restrict__ items, const vh_limb *__restrict__ limbs, int logn, int s0)
{
    __shared__ double lds[vh_ntt16_lds<R, COL>()];
    const vh_item &it = items[blockIdx.y];
    ntt16_task<DIR, R, COL, FIRST, LAST, false, SHARED>(it, limbs[it.limb], logn, s0, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// Pipelined pass (FP64 path, N = 2^13 .. 2^16): the same stages, ownership and
// arithmetic as ntt16_task, but a workgroup walks a contiguous range of the
// launch's tiles (tile u = G sub-transforms of row item u / tiles_per_item),
// issuing the loads of its next tile before transforming the current one, and
// takes no twiddle from global memory inside the transform, so those loads stay
// in flight through it.  With a grid of one workgroup per tile the range is one
// tile: the workgroups in flight then stream neighbouring tiles, which measures
// faster than persistent ranges for the row passes and the 512-thread column
// passes (the host picks the grid).
// NT = 512 (column passes) doubles the columns of a tile, so each row segment a
// load instruction touches is 256 bytes instead of 128.  Twiddles come from the table head tw[0..255] staged in
// LDS (column passes: s0 = 0 and sub-independent indices) and, in row passes,
// from the product of a head entry and one of the row's 8 factors
// (vh_limb::rowf_*): tw[2^s + (sub << s') + local] = tw[2^s' + local] * b[sub][s'].
// A product is centred and paired with RN(w * RN(1/q)), whose relative error
// (<= 2^-52) moves the quotient estimate by at most |b| / 2^53 for a multiplied
// operand b.  Every multiplied operand stays below 4q <= 2^52 (forward sums grow
// by less than q per stage from q/2; inverse sums are re-centred every other
// stage, as in the column passes), so each quotient estimate is below 2^51 and
// each product within q of zero.  The 15 twiddles of a row's stages s' = 0..3
// are the same for its 16 threads: thread e < 15 forms twiddle e once per tile
// into the row's padding slots of the exchange buffer, which no exchange writes.
// ---------------------------------------------------------------------------
template <int DIR, int R, bool COL, bool FIRST, bool LAST, bool SHARED = false, int NT = 256>
__global__ void __launch_bounds__(NT, 1024 / NT)
k_ntt16p(const vh_item *__restrict__ items, const vh_limb *__restrict__ limbs, int logn, int s0,
         unsigned tiles_per_item, unsigned ntiles)
{
    // NT threads: 256, or 512 for a column pass over twice as many columns per tile
    constexpr int S = 1 << R, TPS = S / 16, G = NT / TPS;
    static_assert(COL || (R == 8 && NT == 256), "row passes: 256-point rows (15 padding slots per row)");
    static_assert(NT == 256 || NT == 512, "256 or 512 threads");
    __shared__ double lds[COL ? S * (G + 1) : G * (S + S / 16)];
    // Table head tw[0..255]: column passes use (w, w/q) pairs; row passes multiply the
    // entry by a row factor and use w alone, which keeps their LDS at 40 KiB (four
    // workgroups per CU).
    __shared__ __align__(16) double headw[COL ? 512 : 256];
    const int tid = threadIdx.x;
    const int g = COL ? (tid % G) : (tid / TPS);
    const int t = COL ? (tid / G) : (tid % TPS);
    const size_t nsub = (size_t)1 << (logn - R);
    // Element k of sub-transform sub lives at k * nsub + sub (column) or sub * S + k (row).  An
    // access is split into a uniform row pointer (k's multiple of the ownership stride) and a
    // 32-bit per-lane byte offset, so the 16 accesses of a thread share one offset register.
    const unsigned kstep = COL ? (unsigned)nsub : 1u;
    auto at = [&](auto *base, size_t sub, int kl, int ku) {
        const uint32_t off = (uint32_t)((COL ? sub + (size_t)kl * nsub : sub * S + (size_t)kl) * 8);
        using P = decltype(base);
        return (P)((typename std::conditional<std::is_const<typename std::remove_pointer<P>::type>::value,
                                              const VH_GLOBAL char, VH_GLOBAL char>::type *)(base + (size_t)ku * kstep) + off);
    };
    auto lidx = [&](int k) -> int { return COL ? k * (G + 1) + g : g * (S + S / 16) + k + (k >> 4); };
    // raw[m] holds element 16t + m (inverse column pass: contiguous set) or t + TPS m (stride set)
    constexpr bool CLOAD = DIR == 1 && COL;

    uint64_t raw[16];
    // Row passes: the 8 factors of each of the tile's 16 rows; thread 16 r + k (k < 8) loads
    // factor k of row r, which lies in the wave that consumes it.
    __shared__ double2 rowb[2][G][8];
    double2 fnext = make_double2(0.0, 0.0);
    auto fetch = [&](unsigned uu) {
        const vh_item &iu = items[uu / tiles_per_item];
        const size_t su = (size_t)(uu % tiles_per_item) * G + g;
        if (!COL && t < 8) {   // issued first: landed whenever the data has
            const vh_limb &Lu = limbs[iu.limb];
            const VH_GLOBAL double *rf = vh_global(reinterpret_cast<const double *>(DIR == 0 ? Lu.rowf_fwd : Lu.rowf_inv));
            fnext = make_double2(rf[16 * su + 2 * t], rf[16 * su + 2 * t + 1]);
        }
        const VH_GLOBAL uint64_t *in = vh_global(FIRST ? iu.src : iu.dst);
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const VH_GLOBAL uint64_t *pm = CLOAD ? at(in, su, 16 * t, m) : at(in, su, t, TPS * m);
            raw[m] = SHARED ? *pm : vh_ld(pm);   // a shared source row stays in the caches
        }
    };
    // A workgroup walks a contiguous range of tiles (items are sorted limb-major, so the
    // limb, and with it the table head in LDS, changes rarely).  Nothing in the loop waits
    // on a global load issued after the prefetch except at a change of limb.
    const unsigned per = (ntiles + gridDim.x - 1) / gridDim.x;
    const unsigned u0 = blockIdx.x * per, u1 = u0 + per < ntiles ? u0 + per : ntiles;
    unsigned cur_limb = ~0u;
    if (u0 < u1) fetch(u0);
    for (unsigned u = u0; u < u1; u++) {
        const vh_item &it = items[u / tiles_per_item];
        const vh_limb &L = limbs[it.limb];
        const double q = L.qd, qinv = L.qinv;
        const size_t sub = (size_t)(u % tiles_per_item) * G + g;
        VH_GLOBAL uint64_t *out = vh_global(it.dst);
        const int buf = (u - u0) & 1;

        double v[16];
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = FIRST ? vh_u2d(raw[m]) : __longlong_as_double((long long)raw[m]);
        if (!COL && t < 8) rowb[buf][g][t] = fnext;
        __syncthreads();   // the previous tile's LDS reads are done; rowb[buf] is visible
        if (it.limb != cur_limb) {   // uniform: reload the table head
            cur_limb = it.limb;
            const VH_GLOBAL double *tw = vh_global(reinterpret_cast<const double *>(DIR == 0 ? L.tw_fwd : L.tw_inv));
            if (tid < 256) {
                if (COL) reinterpret_cast<double2 *>(headw)[tid] = make_double2(tw[2 * tid], tw[2 * tid + 1]);
                else headw[tid] = tw[2 * tid];
            }
            __syncthreads();
        }
        __builtin_amdgcn_sched_barrier(0);
        if (u + 1 < u1) fetch(u + 1);
        __builtin_amdgcn_sched_barrier(0);
        // padding slot of row-twiddle e (e < 15): 17 e + 16 (lidx skips 16, 33, ..., 271)
        auto rslot = [&](int e) -> int { return g * (S + S / 16) + 17 * e + 16; };
        if (!COL) {
            if (t < 15) {
                const int sp = 31 - __builtin_clz(t + 1), local = t + 1 - (1 << sp);
                const double2 f = rowb[buf][g][sp];
                lds[rslot(t)] = vh_center(vh_mulmod_tw(headw[(1 << sp) + local], f.x, f.y, q), q, qinv);
            }
            vh_wave_sync();   // a row's 16 threads are in one wave
        }

        auto twid = [&](int sp, int local) -> double2 {
            if (COL) return reinterpret_cast<const double2 *>(headw)[(1 << sp) + local];
            if (sp < 4) {
                const double w = lds[rslot((1 << sp) - 1 + local)];
                return make_double2(w, w * qinv);
            }
            // the factor is read at its point of use: held in registers across the tile,
            // the 8 factors would take 32 VGPRs
            const double2 f = rowb[buf][g][sp];
            const double w = vh_center(vh_mulmod_tw(headw[(1 << sp) + local], f.x, f.y, q), q, qinv);
            return make_double2(w, w * qinv);
        };
        auto finish = [&](double x) -> uint64_t {
            return LAST ? vh_d2u(vh_canon(x, q, qinv)) : (uint64_t)__double_as_longlong(vh_center(x, q, qinv));
        };
        if (DIR == 0) {
#pragma unroll
            for (int sp = 0; sp < 4; sp++) {
                const int hm = 8 >> sp;
#pragma unroll
                for (int m = 0; m < 16; m++) {
                    if (m & hm) continue;
                    const double2 w = twid(sp, m >> (4 - sp));
                    const double x = vh_mulmod_tw(v[m + hm], w.x, w.y, q);
                    const double u0 = v[m];
                    v[m] = u0 + x;
                    v[m + hm] = u0 - x;
                }
            }
#pragma unroll
            for (int m = 0; m < 16; m++) lds[lidx(t + TPS * m)] = vh_center(v[m], q, qinv);
            if (COL) __syncthreads(); else vh_wave_sync();
#pragma unroll
            for (int m = 0; m < 16; m++) v[m] = lds[lidx(16 * t + m)];
#pragma unroll
            for (int sp = 4; sp < R; sp++) {
                const int hm = 1 << (R - sp - 1);
#pragma unroll
                for (int m = 0; m < 16; m++) {
                    if (m & hm) continue;
                    const double2 w = twid(sp, (16 * t + m) >> (R - sp));
                    const double x = vh_mulmod_tw(v[m + hm], w.x, w.y, q);
                    const double u0 = v[m];
                    v[m] = u0 + x;
                    v[m + hm] = u0 - x;
                }
            }
            if (COL) {
#pragma unroll
                for (int m = 0; m < 16; m++) vh_st(at(out, sub, 16 * t, m), finish(v[m]));
            } else {
                vh_wave_sync();
#pragma unroll
                for (int m = 0; m < 16; m++) lds[lidx(16 * t + m)] = __longlong_as_double((long long)finish(v[m]));
                vh_wave_sync();
#pragma unroll
                for (int m = 0; m < 16; m++)
                    vh_st(at(out, sub, t, TPS * m), (uint64_t)__double_as_longlong(lds[lidx(t + TPS * m)]));
            }
        } else {
            if (!COL) {   // row pass: regroup the stride loads into contiguous ownership
#pragma unroll
                for (int m = 0; m < 16; m++) lds[lidx(t + TPS * m)] = v[m];
                vh_wave_sync();
#pragma unroll
                for (int m = 0; m < 16; m++) v[m] = lds[lidx(16 * t + m)];
            }
            int c = 0;
#pragma unroll
            for (int sp = R - 1; sp >= 4; sp--) {
                const int hm = 1 << (R - sp - 1);
#pragma unroll
                for (int m = 0; m < 16; m++) {
                    if (m & hm) continue;
                    const double2 w = twid(sp, (16 * t + m) >> (R - sp));
                    const double u0 = v[m], x = v[m + hm];
                    double a = u0 + x;
                    if (c & 1) a = vh_center(a, q, qinv);
                    v[m] = a;
                    v[m + hm] = vh_mulmod_tw(u0 - x, w.x, w.y, q);
                }
                c++;
            }
            if (!COL) vh_wave_sync();
#pragma unroll
            for (int m = 0; m < 16; m++) lds[lidx(16 * t + m)] = v[m];
            if (COL) __syncthreads(); else vh_wave_sync();
#pragma unroll
            for (int m = 0; m < 16; m++) v[m] = lds[lidx(t + TPS * m)];
#pragma unroll
            for (int sp = 3; sp >= 0; sp--) {
                const int hm = 8 >> sp;
#pragma unroll
                for (int m = 0; m < 16; m++) {
                    if (m & hm) continue;
                    const double2 w = twid(sp, m >> (4 - sp));
                    const double u0 = v[m], x = v[m + hm];
                    double a = u0 + x;
                    if (c & 1) a = vh_center(a, q, qinv);
                    v[m] = a;
                    v[m + hm] = vh_mulmod_tw(u0 - x, w.x, w.y, q);
                }
                c++;
            }
#pragma unroll
            for (int m = 0; m < 16; m++) {
                const double x = LAST ? vh_mulmod_tw(v[m], L.ninv_d, L.ninv_q, q) : v[m];
                vh_st(at(out, sub, t, TPS * m), finish(x));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Key-switch inner product fused into the forward row pass (FP64 path,
// N = 2^13..2^16, at most 2 output components).  The column pass has left the
// lifted digit rows, transformed over their first logn-8 stages, in jb.modup;
// this kernel finishes the last 8 stages of every digit for one 4096-element
// block and accumulates   out_c = - sum_d NTT(digit_d) * key[d][c]   (the
// mul_subto of gadget.c:22-37) in registers, so the transformed digits never
// return to memory.  Grid: x = ciphertext, y = 4096-element block, z = extended
// limb; a key block is (y, z).  Consecutive workgroups are dealt round-robin
// over the 8 XCDs, so with `xcd` set the workgroup index is remapped to put all
// ciphertexts of one key block on one XCD, one after another: each XCD's L2 then
// fetches a key block once instead of every XCD fetching it.
// ---------------------------------------------------------------------------
template <int C>
__global__ void __launch_bounds__(256)
k_ntt16_row_mac(const vh_mac_job *__restrict__ jobs, const vh_limb *__restrict__ limbs,
                const uint64_t *const *__restrict__ kptr, int logn, int xcd)
{
    constexpr int R = 8, S = 256, TPS = 16, G = 16;
    __shared__ double lds[vh_ntt16_lds<R, false>()];
    unsigned bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (xcd) {   // bijective when gridDim.y * gridDim.z is a multiple of 8 (checked at launch)
        const unsigned lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        const unsigned slot = lin >> 3, kb = (slot / gridDim.x) * 8 + (lin & 7);
        bx = slot % gridDim.x;
        by = kb % gridDim.y;
        bz = kb / gridDim.y;
    }
    const vh_mac_job &jb = jobs[bx];
    const int ti = (int)bz, D = jb.digits;
    if (ti >= jb.next) return;   // jobs of a batch may have fewer extended limbs (whole block)
    const int limb = jb.limb[ti];
    const vh_limb &L = limbs[limb];
    const double q = L.qd, qinv = L.qinv;
    const size_t n = (size_t)1 << logn;
    const int s0 = logn - R;
    const int tid = threadIdx.x, g = tid / TPS, t = tid % TPS;
    const size_t sub = (size_t)by * G + g;
    const VH_GLOBAL double *tw = vh_global(reinterpret_cast<const double *>(L.tw_fwd));
    auto lidx = [&](int k) -> int { return g * (S + S / 16) + k + (k >> 4); };
    // Twiddles do not depend on the digit: they are staged in LDS once, so the digit
    // loop issues no twiddle load behind the next digit's prefetch.  rowt[g]: the 15
    // stride-stage twiddles of row g; head: tw[0..255]; rowb[g]: row g's factors of
    // stages 4..7 (contiguous-stage twiddle = head entry * factor, as in k_ntt16p).
    __shared__ double2 head[256], rowt[G][15], rowb[G][4];
    {
        const int tid0 = threadIdx.x;
        head[tid0] = make_double2(tw[2 * tid0], tw[2 * tid0 + 1]);
        if (tid0 < G * 15) {
            const int r = tid0 / 15, e = tid0 % 15;           // e = 2^sp - 1 + local
            const int sp = 31 - __builtin_clz(e + 1), local = e + 1 - (1 << sp);
            const size_t i = ((size_t)1 << (s0 + sp)) + (((size_t)by * G + r) << sp) + local;
            rowt[r][e] = make_double2(tw[2 * i], tw[2 * i + 1]);
        }
        if (tid0 < G * 4) {
            const int r = tid0 / 4, k = tid0 % 4;
            const VH_GLOBAL double *rf = vh_global(reinterpret_cast<const double *>(L.rowf_fwd));
            const size_t i = ((size_t)by * G + r) * 8 + 4 + k;
            rowb[r][k] = make_double2(rf[2 * i], rf[2 * i + 1]);
        }
        __syncthreads();
    }
    auto twid = [&](int sp, size_t local) -> double2 {
        if (sp < 4) return rowt[g][(1 << sp) - 1 + (int)local];
        const double2 h = head[(1 << sp) + (int)local], b = rowb[g][sp - 4];
        const double w = vh_center(vh_mulmod_tw(h.x, b.x, b.y, q), q, qinv);
        return make_double2(w, w * qinv);
    };
    auto digit_row = [&](int d) -> const VH_GLOBAL uint64_t * {
        return vh_global(jb.modup + ((size_t)d * jb.next + ti) * n + sub * S);
    };
    // After the last stage a thread owns the contiguous elements x0 .. x0+15 of its
    // row: key rows are read and outputs written as 16 consecutive words per lane.
    const size_t x0 = sub * S + 16 * t;

    double acc[C][16];
#pragma unroll
    for (int c = 0; c < C; c++)
#pragma unroll
        for (int m = 0; m < 16; m++) acc[c][m] = 0.0;

    // One output component: the next digit is prefetched and the key words load with it,
    // so both land during the butterflies.  Two components: the 64 accumulator registers
    // leave room for neither (the prefetching form needs 238 VGPRs, two waves per SIMD);
    // the digit and then each component's key words load when needed, at 162 VGPRs and
    // three waves per SIMD, whose extra wave hides those loads (3% faster, measured).
    constexpr bool PREF = C == 1, KEARLY = C == 1;
    uint64_t raw[16];
    if (D > 0 && PREF) {
        const VH_GLOBAL uint64_t *in = digit_row(0);
#pragma unroll
        for (int m = 0; m < 16; m++) raw[m] = vh_ld(in + t + TPS * m);
    }
    for (int d = 0; d < D; d++) {
        double v[16];
        if (!PREF) {
            const VH_GLOBAL uint64_t *in = digit_row(d);
#pragma unroll
            for (int m = 0; m < 16; m++) raw[m] = vh_ld(in + t + TPS * m);
        }
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = __longlong_as_double((long long)raw[m]);
        if (PREF && d + 1 < D) {   // next digit's loads overlap this digit's butterflies
            const VH_GLOBAL uint64_t *in = digit_row(d + 1);
#pragma unroll
            for (int m = 0; m < 16; m++) raw[m] = vh_ld(in + t + TPS * m);
        }
        const uint64_t *const *kp = kptr + jb.kptr_off + ((size_t)ti * D + d) * C;
        uint64_t kw[C][16];
        if (KEARLY) {
#pragma unroll
            for (int c = 0; c < C; c++) {
                const VH_GLOBAL uint64_t *kr = vh_global(kp[c]) + x0;
#pragma unroll
                for (int m = 0; m < 16; m++) kw[c][m] = kr[m];
            }
        }
        // stages s' = 0..3, stride ownership
#pragma unroll
        for (int sp = 0; sp < 4; sp++) {
            const int hm = 8 >> sp;
#pragma unroll
            for (int m = 0; m < 16; m++) {
                if (m & hm) continue;
                const double2 w = twid(sp, (size_t)(m >> (4 - sp)));
                const double x = vh_mulmod_tw(v[m + hm], w.x, w.y, q);
                const double u = v[m];
                v[m] = u + x;
                v[m + hm] = u - x;
            }
        }
        vh_wave_sync();  // every read of the previous digit's exchange is done
#pragma unroll
        for (int m = 0; m < 16; m++) lds[lidx(t + TPS * m)] = vh_center(v[m], q, qinv);
        vh_wave_sync();
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = lds[lidx(16 * t + m)];
        // stages s' = 4..7, contiguous ownership
#pragma unroll
        for (int sp = 4; sp < R; sp++) {
            const int hm = 1 << (R - sp - 1);
#pragma unroll
            for (int m = 0; m < 16; m++) {
                if (m & hm) continue;
                const double2 w = twid(sp, (size_t)((16 * t + m) >> (R - sp)));
                const double x = vh_mulmod_tw(v[m + hm], w.x, w.y, q);
                const double u = v[m];
                v[m] = u + x;
                v[m + hm] = u - x;
            }
        }
#pragma unroll
        for (int c = 0; c < C; c++) {
            if (!KEARLY) {
                const VH_GLOBAL uint64_t *kr = vh_global(kp[c]) + x0;
#pragma unroll
                for (int m = 0; m < 16; m++) kw[c][m] = kr[m];
            }
#pragma unroll
            for (int h = 0; h < 8; h++) {
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const int m = 2 * h + e;
                    const double a = vh_center(v[m], q, qinv);          // |a| <= q/2 + 1
                    const double k = vh_u2d(kw[c][m]);
                    const double hi = a * k;
                    const double lo = __fma_rn(a, k, -hi);
                    const double qt = __fma_rn(a, k * qinv, VH_MAGIC) - VH_MAGIC;
                    const double r = __fma_rn(-qt, q, hi) + lo;   // |r| <= q + 2^47
                    // centred every 4th digit: |acc| <= q/2 + 4 (q + 2^47) < 2^53 stays exact
                    acc[c][m] = (d & 3) == 3 ? vh_center(acc[c][m] + r, q, qinv) : acc[c][m] + r;
                }
            }
        }
    }
    const bool fold = (jb.pass_mask >> limb) & 1;
#pragma unroll
    for (int c = 0; c < C; c++) {
        VH_GLOBAL uint64_t *out = vh_global(jb.out[c] + (size_t)limb * n + x0);
        if (fold && jb.pass[c]) {
            // |acc| <= q/2 + 4 (q + 2^47) and |P d| <= 3q/4: the sum stays below 2^53.  A lane
            // reads 16 consecutive words (ordinary loads, merged into wide ones, like the key words).
            const VH_GLOBAL uint64_t *ps = vh_global(jb.pass[c] + (size_t)limb * n + x0);
            const double2 pm = jb.pmod[limb];
            uint64_t pw[16];
#pragma unroll
            for (int m = 0; m < 16; m++) pw[m] = ps[m];
#pragma unroll
            for (int m = 0; m < 16; m++) {
                const double y = (jb.add ? acc[c][m] : -acc[c][m]) + vh_mulmod_tw(vh_u2d(pw[m]), pm.x, pm.y, q);
                vh_st(out + m, vh_d2u(vh_canon(y, q, qinv)));
            }
        } else {
#pragma unroll
            for (int m = 0; m < 16; m++) vh_st(out + m, vh_d2u(vh_canon(jb.add ? acc[c][m] : -acc[c][m], q, qinv)));
        }
    }
}

hipError_t vh_launch_ntt16_row_mac(hipStream_t st, const vh_mac_job *jobs, unsigned njobs, unsigned next,
                                   const vh_limb *limbs, const uint64_t *const *kptr, int logn, int comps)
{
    if (logn < 13 || logn > 16 || comps < 1 || comps > 2) return hipErrorInvalidValue;
    const dim3 grid(njobs, 1u << (logn - 12), next);
    const int xcd = (grid.y * grid.z) % 8 == 0;
    if (comps == 1) VH_KL(k_ntt16_row_mac<1>, grid, dim3(256), 0, st, jobs, limbs, kptr, logn, xcd);
    else VH_KL(k_ntt16_row_mac<2>, grid, dim3(256), 0, st, jobs, limbs, kptr, logn, xcd);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// z = INTT(NTT(x) * NTT(y)) middle step (FP64 path, N = 2^13..2^16): after the
// forward column passes of x (in it.src) and y (in it.src2), one kernel runs the
// forward row stages of both, the pointwise product and the inverse row stages,
// and writes the input of the inverse column pass to it.dst.  The transformed
// operands never return to memory.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_ntt16_row_mul_inv(const vh_item *__restrict__ items, const vh_limb *__restrict__ limbs, int logn)
{
    constexpr int R = 8, S = 256, TPS = 16, G = 16;
    __shared__ double lds[vh_ntt16_lds<R, false>()];
    const vh_item &it = items[blockIdx.y];
    const vh_limb &L = limbs[it.limb];
    const double q = L.qd, qinv = L.qinv;
    const int s0 = logn - R;
    const int tid = threadIdx.x, g = tid / TPS, t = tid % TPS;
    const size_t sub = (size_t)blockIdx.x * G + g;
    const VH_GLOBAL double *twf = vh_global(reinterpret_cast<const double *>(L.tw_fwd));
    const VH_GLOBAL double *twi = vh_global(reinterpret_cast<const double *>(L.tw_inv));
    auto lidx = [&](int k) -> int { return g * (S + S / 16) + k + (k >> 4); };
    auto tw = [&](const VH_GLOBAL double *tab, int sp, size_t local) -> double2 {
        const size_t i = ((size_t)1 << (s0 + sp)) + (sub << sp) + local;
        return make_double2(tab[2 * i], tab[2 * i + 1]);
    };
    const VH_GLOBAL uint64_t *xa = vh_global(it.src) + sub * S;
    const VH_GLOBAL uint64_t *xb = vh_global(it.src2) + sub * S;
    VH_GLOBAL uint64_t *out = vh_global(it.dst) + sub * S;

    double a[16], b[16];
#pragma unroll
    for (int m = 0; m < 16; m++) {
        a[m] = __longlong_as_double((long long)vh_ld(xa + t + TPS * m));
        b[m] = __longlong_as_double((long long)vh_ld(xb + t + TPS * m));
    }
    // forward stages s' = 0..3 (stride ownership) on both operands
#pragma unroll
    for (int sp = 0; sp < 4; sp++) {
        const int hm = 8 >> sp;
#pragma unroll
        for (int m = 0; m < 16; m++) {
            if (m & hm) continue;
            const double2 w = tw(twf, sp, (size_t)(m >> (4 - sp)));
            const double xa_ = vh_mulmod_tw(a[m + hm], w.x, w.y, q), xb_ = vh_mulmod_tw(b[m + hm], w.x, w.y, q);
            const double ua = a[m], ub = b[m];
            a[m] = ua + xa_; a[m + hm] = ua - xa_;
            b[m] = ub + xb_; b[m + hm] = ub - xb_;
        }
    }
    // regroup both into contiguous ownership through LDS
#pragma unroll
    for (int m = 0; m < 16; m++) lds[lidx(t + TPS * m)] = vh_center(a[m], q, qinv);
    vh_wave_sync();
#pragma unroll
    for (int m = 0; m < 16; m++) a[m] = lds[lidx(16 * t + m)];
    vh_wave_sync();
#pragma unroll
    for (int m = 0; m < 16; m++) lds[lidx(t + TPS * m)] = vh_center(b[m], q, qinv);
    vh_wave_sync();
#pragma unroll
    for (int m = 0; m < 16; m++) b[m] = lds[lidx(16 * t + m)];
    // forward stages s' = 4..7
#pragma unroll
    for (int sp = 4; sp < R; sp++) {
        const int hm = 1 << (R - sp - 1);
#pragma unroll
        for (int m = 0; m < 16; m++) {
            if (m & hm) continue;
            const double2 w = tw(twf, sp, (size_t)((16 * t + m) >> (R - sp)));
            const double xa_ = vh_mulmod_tw(a[m + hm], w.x, w.y, q), xb_ = vh_mulmod_tw(b[m + hm], w.x, w.y, q);
            const double ua = a[m], ub = b[m];
            a[m] = ua + xa_; a[m + hm] = ua - xa_;
            b[m] = ub + xb_; b[m + hm] = ub - xb_;
        }
    }
    // pointwise product of centred operands (|.| <= q/2 + 1): exact product,
    // quotient within 1, |r| <= q + 2^47, then centred
#pragma unroll
    for (int m = 0; m < 16; m++) {
        const double u = vh_center(a[m], q, qinv), v = vh_center(b[m], q, qinv);
        const double hi = u * v;
        const double lo = __fma_rn(u, v, -hi);
        const double qt = __fma_rn(u, v * qinv, VH_MAGIC) - VH_MAGIC;
        a[m] = vh_center(__fma_rn(-qt, q, hi) + lo, q, qinv);
    }
    // inverse stages s' = 7..4 (contiguous ownership), GS sums re-centred every second stage
    int c = 0;
#pragma unroll
    for (int sp = R - 1; sp >= 4; sp--) {
        const int hm = 1 << (R - sp - 1);
#pragma unroll
        for (int m = 0; m < 16; m++) {
            if (m & hm) continue;
            const double2 w = tw(twi, sp, (size_t)((16 * t + m) >> (R - sp)));
            const double u = a[m], x = a[m + hm];
            double s_ = u + x;
            if (c & 1) s_ = vh_center(s_, q, qinv);
            a[m] = s_;
            a[m + hm] = vh_mulmod_tw(u - x, w.x, w.y, q);
        }
        c++;
    }
    vh_wave_sync();   // every read of b's regrouping is done
#pragma unroll
    for (int m = 0; m < 16; m++) lds[lidx(16 * t + m)] = a[m];
    vh_wave_sync();
#pragma unroll
    for (int m = 0; m < 16; m++) a[m] = lds[lidx(t + TPS * m)];
    // inverse stages s' = 3..0 (stride ownership)
#pragma unroll
    for (int sp = 3; sp >= 0; sp--) {
        const int hm = 8 >> sp;
#pragma unroll
        for (int m = 0; m < 16; m++) {
            if (m & hm) continue;
            const double2 w = tw(twi, sp, (size_t)(m >> (4 - sp)));
            const double u = a[m], x = a[m + hm];
            double s_ = u + x;
            if (c & 1) s_ = vh_center(s_, q, qinv);
            a[m] = s_;
            a[m + hm] = vh_mulmod_tw(u - x, w.x, w.y, q);
        }
        c++;
    }
#pragma unroll
    for (int m = 0; m < 16; m++) vh_st(out + t + TPS * m, (uint64_t)__double_as_longlong(vh_center(a[m], q, qinv)));
}

// ---------------------------------------------------------------------------
// Pipelined form of k_ntt16_row_mul_inv (FP64 path, N = 2^13..2^16): persistent
// workgroups walk contiguous ranges of 4096-element tiles (16 rows of x and y),
// loading the current tile's y and then the next tile's x before the transform
// (y lands while x's first stages run).  Twiddles are formed as in k_ntt16p's row passes: head entry (LDS) times
// the row's factor, the 15 stage-0..3 twiddles of a row once per tile into the
// padding slots of the exchange buffer (forward ones first, then the inverse
// ones after the forward stages have used them).  Every multiplied operand stays
// below 4q (inverse sums re-centred every other stage), as in k_ntt16p.  234 VGPRs,
// two workgroups per CU: 5% faster than the same kernel without the prefetch at
// three (which spills), and than k_ntt16_row_mul_inv.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256, 2)
k_ntt16p_row_mul(const vh_item *__restrict__ items, const vh_limb *__restrict__ limbs, int logn,
                 unsigned tiles_per_item, unsigned ntiles)
{
    constexpr int R = 8, S = 256, TPS = 16, G = 16;
    __shared__ double lds[vh_ntt16_lds<R, false>()];
    __shared__ double headf[256], headi[256];
    __shared__ double2 rowbf[2][G][8], rowbi[2][G][8];
    const int tid = threadIdx.x, g = tid / TPS, t = tid % TPS;
    auto lidx = [&](int k) -> int { return g * (S + S / 16) + k + (k >> 4); };
    // padding slot of row-twiddle e (e < 15): 17 e + 16 (lidx skips 16, 33, ..., 271)
    auto rslot = [&](int e) -> int { return g * (S + S / 16) + 17 * e + 16; };

    // x of the next tile is prefetched; y of the current tile is loaded at its start, ahead
    // of that prefetch, and lands while x is transformed.
    uint64_t rx[16];
    double2 fnf = make_double2(0.0, 0.0), fni = make_double2(0.0, 0.0);
    auto fetch = [&](unsigned uu) {
        const vh_item &iu = items[uu / tiles_per_item];
        const size_t su = (size_t)(uu % tiles_per_item) * G + g;
        if (t < 8) {   // issued first: landed whenever the data has
            const vh_limb &Lu = limbs[iu.limb];
            const VH_GLOBAL double *rf = vh_global(reinterpret_cast<const double *>(Lu.rowf_fwd));
            const VH_GLOBAL double *ri = vh_global(reinterpret_cast<const double *>(Lu.rowf_inv));
            fnf = make_double2(rf[16 * su + 2 * t], rf[16 * su + 2 * t + 1]);
            fni = make_double2(ri[16 * su + 2 * t], ri[16 * su + 2 * t + 1]);
        }
        const VH_GLOBAL uint64_t *xa = vh_global(iu.src) + su * S;
#pragma unroll
        for (int m = 0; m < 16; m++) rx[m] = vh_ld(xa + t + TPS * m);
    };
    const unsigned per = (ntiles + gridDim.x - 1) / gridDim.x;
    const unsigned u0 = blockIdx.x * per, u1 = u0 + per < ntiles ? u0 + per : ntiles;
    unsigned cur_limb = ~0u;
    if (u0 < u1) fetch(u0);
    for (unsigned u = u0; u < u1; u++) {
        const vh_item &it = items[u / tiles_per_item];
        const vh_limb &L = limbs[it.limb];
        const double q = L.qd, qinv = L.qinv;
        const size_t sub = (size_t)(u % tiles_per_item) * G + g;
        VH_GLOBAL uint64_t *out = vh_global(it.dst) + sub * S;
        const int buf = (u - u0) & 1;

        double a[16];
#pragma unroll
        for (int m = 0; m < 16; m++) a[m] = __longlong_as_double((long long)rx[m]);
        if (t < 8) {
            rowbf[buf][g][t] = fnf;
            rowbi[buf][g][t] = fni;
        }
        __syncthreads();   // the previous tile's LDS reads are done; rowb*[buf] are visible
        if (it.limb != cur_limb) {   // uniform: reload the table heads
            cur_limb = it.limb;
            headf[tid] = vh_global(reinterpret_cast<const double *>(L.tw_fwd))[2 * tid];
            headi[tid] = vh_global(reinterpret_cast<const double *>(L.tw_inv))[2 * tid];
            __syncthreads();
        }
        __builtin_amdgcn_sched_barrier(0);
        uint64_t ry[16];
        {
            const VH_GLOBAL uint64_t *xb = vh_global(it.src2) + sub * S;
#pragma unroll
            for (int m = 0; m < 16; m++) ry[m] = vh_ld(xb + t + TPS * m);
        }
        if (u + 1 < u1) fetch(u + 1);
        __builtin_amdgcn_sched_barrier(0);

        // INV selects the inverse table; the LDS arrays are named directly (a pointer
        // parameter would be a generic address, read with flat loads)
        auto slot_twiddles = [&](auto inv) {
            constexpr bool INV = decltype(inv)::value;
            if (t < 15) {
                const int sp = 31 - __builtin_clz(t + 1), local = t + 1 - (1 << sp);
                const double2 f = INV ? rowbi[buf][g][sp] : rowbf[buf][g][sp];
                const double h = INV ? headi[(1 << sp) + local] : headf[(1 << sp) + local];
                lds[rslot(t)] = vh_center(vh_mulmod_tw(h, f.x, f.y, q), q, qinv);
            }
        };
        auto twid = [&](auto inv, int sp, int local) -> double2 {
            constexpr bool INV = decltype(inv)::value;
            if (sp < 4) {
                const double w = lds[rslot((1 << sp) - 1 + local)];
                return make_double2(w, w * qinv);
            }
            const double2 f = INV ? rowbi[buf][g][sp] : rowbf[buf][g][sp];
            const double h = INV ? headi[(1 << sp) + local] : headf[(1 << sp) + local];
            const double w = vh_center(vh_mulmod_tw(h, f.x, f.y, q), q, qinv);
            return make_double2(w, w * qinv);
        };
        const std::false_type FWD{};
        const std::true_type INVT{};
        // forward row stages: 0..3 in stride ownership on each operand (x first, while y
        // lands), both regrouped through the data slots of LDS, then 4..7 on both at once
        // (each twiddle formed once and used for both)
        auto stages03 = [&](double (&v)[16]) {
#pragma unroll
            for (int sp = 0; sp < 4; sp++) {
                const int hm = 8 >> sp;
#pragma unroll
                for (int m = 0; m < 16; m++) {
                    if (m & hm) continue;
                    const double2 w = twid(FWD, sp, m >> (4 - sp));
                    const double x = vh_mulmod_tw(v[m + hm], w.x, w.y, q);
                    const double uu = v[m];
                    v[m] = uu + x;
                    v[m + hm] = uu - x;
                }
            }
        };
        auto regroup = [&](double (&v)[16]) {
            vh_wave_sync();   // the previous reads of the data slots are done
#pragma unroll
            for (int m = 0; m < 16; m++) lds[lidx(t + TPS * m)] = vh_center(v[m], q, qinv);
            vh_wave_sync();
#pragma unroll
            for (int m = 0; m < 16; m++) v[m] = lds[lidx(16 * t + m)];
        };
        slot_twiddles(FWD);
        vh_wave_sync();   // a row's 16 threads are in one wave
        stages03(a);
        // compiler barrier: the twiddles are re-read for y rather than held in registers
        asm volatile("" ::: "memory");
        double b[16];
#pragma unroll
        for (int m = 0; m < 16; m++) b[m] = __longlong_as_double((long long)ry[m]);
        stages03(b);
        regroup(a);
        regroup(b);
#pragma unroll
        for (int sp = 4; sp < R; sp++) {
            const int hm = 1 << (R - sp - 1);
#pragma unroll
            for (int m = 0; m < 16; m++) {
                if (m & hm) continue;
                const double2 w = twid(FWD, sp, (16 * t + m) >> (R - sp));
                const double xa_ = vh_mulmod_tw(a[m + hm], w.x, w.y, q), xb_ = vh_mulmod_tw(b[m + hm], w.x, w.y, q);
                const double ua = a[m], ub = b[m];
                a[m] = ua + xa_; a[m + hm] = ua - xa_;
                b[m] = ub + xb_; b[m + hm] = ub - xb_;
            }
        }
        // pointwise product of centred operands (|.| <= q/2 + 1), centred
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const double x = vh_center(a[m], q, qinv), y = vh_center(b[m], q, qinv);
            const double hi = x * y;
            const double lo = __fma_rn(x, y, -hi);
            const double qt = __fma_rn(x, y * qinv, VH_MAGIC) - VH_MAGIC;
            a[m] = vh_center(__fma_rn(-qt, q, hi) + lo, q, qinv);
        }
        // inverse stages s' = 7..4 (contiguous ownership), sums re-centred every other stage
        int c = 0;
#pragma unroll
        for (int sp = R - 1; sp >= 4; sp--) {
            const int hm = 1 << (R - sp - 1);
#pragma unroll
            for (int m = 0; m < 16; m++) {
                if (m & hm) continue;
                const double2 w = twid(INVT, sp, (16 * t + m) >> (R - sp));
                const double uu = a[m], x = a[m + hm];
                double s_ = uu + x;
                if (c & 1) s_ = vh_center(s_, q, qinv);
                a[m] = s_;
                a[m + hm] = vh_mulmod_tw(uu - x, w.x, w.y, q);
            }
            c++;
        }
        vh_wave_sync();   // every read of the forward twiddle slots and of b's regrouping is done
        slot_twiddles(INVT);   // the inverse stage-0..3 twiddles replace the forward ones
#pragma unroll
        for (int m = 0; m < 16; m++) lds[lidx(16 * t + m)] = a[m];
        vh_wave_sync();
#pragma unroll
        for (int m = 0; m < 16; m++) a[m] = lds[lidx(t + TPS * m)];
        // inverse stages s' = 3..0 (stride ownership)
#pragma unroll
        for (int sp = 3; sp >= 0; sp--) {
            const int hm = 8 >> sp;
#pragma unroll
            for (int m = 0; m < 16; m++) {
                if (m & hm) continue;
                const double2 w = twid(INVT, sp, m >> (4 - sp));
                const double uu = a[m], x = a[m + hm];
                double s_ = uu + x;
                if (c & 1) s_ = vh_center(s_, q, qinv);
                a[m] = s_;
                a[m + hm] = vh_mulmod_tw(uu - x, w.x, w.y, q);
            }
            c++;
        }
#pragma unroll
        for (int m = 0; m < 16; m++) vh_st(out + t + TPS * m, (uint64_t)__double_as_longlong(vh_center(a[m], q, qinv)));
    }
}

hipError_t vh_launch_ntt16p_row_mul(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                                    int logn, unsigned cus, int one_tile)
{
    if (logn < 13 || logn > 16 || cus == 0) return hipErrorInvalidValue;
    static int occ = -1;
    if (occ < 0 && hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_ntt16p_row_mul, 256, 0) != hipSuccess) occ = -1;
    if (occ <= 0) return hipErrorInvalidValue;
    const unsigned tpi = 1u << (logn - 12), ntiles = tpi * nitems, nb = one_tile ? ntiles : cus * (unsigned)occ;
    VH_KL(k_ntt16p_row_mul, dim3(nb < ntiles ? nb : ntiles), dim3(256), 0, st, items, limbs, logn, tpi,
                       ntiles);
    return hipGetLastError();
}

hipError_t vh_launch_ntt16_row_mul_inv(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                                       int logn)
{
    if (logn < 13 || logn > 16) return hipErrorInvalidValue;
    VH_KL(k_ntt16_row_mul_inv, dim3(1u << (logn - 12), nitems), dim3(256), 0, st, items, limbs, logn);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Persistent two-pass NTT with per-XCD work queues (FP64 path, N = 2^13..2^16).
//
// Item i belongs to XCD i mod 8; each XCD drains its own ticket counter.  Slot j
// of an XCD's queue holds the T first-pass tasks of its j-th item followed by the
// T second-pass tasks of its (j - lag)-th item (T = N/4096 tasks per pass).  A
// second-pass task waits until the first-pass tasks of its item have signalled;
// those hold smaller tickets of the same queue, were taken by running workgroups
// of the same XCD and never wait themselves, so every wait ends.  The
// intermediate therefore moves between the two passes through the XCD's own L2:
// stores complete to L2 before the signal (s_waitcnt), and the second pass reads
// with device-coherent loads that miss the L1.  A bounded spin sets an error flag
// instead of waiting forever.
//
// ctl layout (uint32): [0..7] tickets per XCD, [16 + i] first-pass tasks done for
// item i, [16 + nitems + i] second-pass tasks done; a failed wait ORs 1 into *err
// and k_ntt_xcd_verify ORs 2 for any item not fully transformed (both are checked
// by the host at its next synchronisation).
// ---------------------------------------------------------------------------
template <int DIR, int R1, int OCC, bool L2X>
__global__ void __launch_bounds__(256, OCC)
k_ntt_xcd(const vh_item *__restrict__ items, unsigned nitems, const vh_limb *__restrict__ limbs, int logn,
          unsigned *__restrict__ ctl, unsigned *__restrict__ err, unsigned lag, unsigned spin_limit)
{
    constexpr int LDS_N = vh_ntt16_lds<R1, true>() > vh_ntt16_lds<8, false>() ? vh_ntt16_lds<R1, true>()
                                                                             : vh_ntt16_lds<8, false>();
    __shared__ double lds[LDS_N];
    __shared__ unsigned s_ticket, s_ok;
    const unsigned xcd = __builtin_amdgcn_s_getreg((3 << 11) | 20) & 7;  // HW_REG_XCC_ID
    const unsigned mine = nitems > xcd ? (nitems - xcd + 7) / 8 : 0;
    const unsigned T = 1u << (logn - 12);
    const unsigned total = (mine + lag) * 2 * T;
    unsigned *done = ctl + 16;
    for (;;) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(ctl + xcd, 1u);
        __syncthreads();
        const unsigned tk = s_ticket;
        __syncthreads();
        if (tk >= total) break;
        const unsigned j = tk / (2 * T), r = tk % (2 * T);
        if (r < T) {
            if (j >= mine) continue;
            const unsigned i = xcd + 8 * j;
            const vh_item &it = items[i];
            if (DIR == 0) ntt16_task<0, R1, true, true, false, false, false, !L2X>(it, limbs[it.limb], logn, 0, r, lds);
            else ntt16_task<1, 8, false, true, false, false, false, !L2X>(it, limbs[it.limb], logn, R1, r, lds);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_fetch_add(done + i, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (j < lag || j - lag >= mine) continue;
            const unsigned i = xcd + 8 * (j - lag);
            if (threadIdx.x == 0) {
                unsigned spins = 0;
                while (__hip_atomic_load(done + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < T &&
                       ++spins < spin_limit)
                    __builtin_amdgcn_s_sleep(2);
                s_ok = spins < spin_limit;
                if (!s_ok) __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
            if (!s_ok) break;
            const vh_item &it = items[i];
            if (DIR == 0) ntt16_task<0, 8, false, false, true, true>(it, limbs[it.limb], logn, R1, r - T, lds);
            else ntt16_task<1, R1, true, false, true, true>(it, limbs[it.limb], logn, 0, r - T, lds);
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_fetch_add(done + nitems + i, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------------------
// Elementwise ops (modules/arith/c/src/mod_portable.c:55-161 semantics, exact
// for every uint64 input, canonical output).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_eltwise(const vh_item *__restrict__ items, const vh_limb *__restrict__ limbs, int op, uint64_t n)
{
    const vh_item it = items[blockIdx.y];
    const vh_limb L = limbs[it.limb];
    const uint64_t q = L.q, mu = L.mu64;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t r;
        switch (op) {
        case VH_OP_MUL:
            r = vh_mulmod(vh_reduce64(it.src[i], q, mu), vh_reduce64(it.src2[i], q, mu), L);
            break;
        case VH_OP_MUL_ADDTO:
            r = vh_addmod(vh_reduce64(it.dst[i], q, mu),
                          vh_mulmod(vh_reduce64(it.src[i], q, mu), vh_reduce64(it.src2[i], q, mu), L), q);
            break;
        case VH_OP_MUL_SUBTO:
            r = vh_submod(vh_reduce64(it.dst[i], q, mu),
                          vh_mulmod(vh_reduce64(it.src[i], q, mu), vh_reduce64(it.src2[i], q, mu), L), q);
            break;
        case VH_OP_SCALE:
            r = vh_shoup(it.src[i], it.scalar, it.scalar_sh, q);
            break;
        case VH_OP_FMA:
            r = vh_addmod(vh_reduce64(it.dst[i], q, mu), vh_shoup(it.src[i], it.scalar, it.scalar_sh, q), q);
            break;
        case VH_OP_ADD:
            r = vh_reduce64(it.src[i] + it.src2[i], q, mu);
            break;
        case VH_OP_SUB:
            r = vh_reduce64(it.src[i] + q - it.src2[i], q, mu);
            break;
        case VH_OP_NEGATE: {
            uint64_t a = vh_reduce64(it.src[i], q, mu);
            r = a ? q - a : 0;
            break;
        }
        case VH_OP_REDUCE:
            r = vh_reduce64(it.src[i], q, mu);
            break;
        case VH_OP_REDUCE_SIGNED: {
            int64_t v = (int64_t)it.src[i];
            uint64_t a = vh_reduce64(v < 0 ? (uint64_t)0 - (uint64_t)v : (uint64_t)v, q, mu);
            r = (v < 0 && a) ? q - a : a;
            break;
        }
        case VH_OP_ADD_SCALAR:
            r = vh_reduce64(it.src[i] + it.scalar, q, mu);
            break;
        case VH_OP_SUB_SCALAR:
            r = vh_reduce64(it.src[i] + q - it.scalar, q, mu);
            break;
        case VH_OP_COPY:
            r = it.src[i];
            break;
        case VH_OP_ZERO:
            r = 0;
            break;
        case VH_OP_REDUCE_MP:  // (src2[i] : src[i]) = hi : lo; scalar = 2^64 mod q
            r = vh_addmod(vh_shoup(vh_reduce64(it.src2[i], q, mu), it.scalar, it.scalar_sh, q),
                          vh_reduce64(it.src[i], q, mu), q);
            break;
        case VH_OP_ADD_AT0:    // add scalar to coefficient 0 only, copy the rest
            r = (i == 0) ? vh_addmod(vh_reduce64(it.src[i], q, mu), it.scalar, q) : it.src[i];
            break;
        default:
            r = 0;
        }
        it.dst[i] = r;
    }
}

// ---------------------------------------------------------------------------
// Split-degree product with the X^split wrap (rns_polynomial.c:311-435 for
// split_degree > 1): out block d = sum_{j+k = d (mod S)} in1_j * in2_k * (w if
// j + k >= S).  mode 0: =, 1: +=, 2: -=.  One thread per (block position x,
// output block d); items: src = in1 row, src2 = in2 row, dst = out row, aux = S,
// scalar = device address of the w row.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_split_mul(const vh_item *__restrict__ items, const vh_limb *__restrict__ limbs, int mode, uint64_t n)
{
    const vh_item it = items[blockIdx.y];
    const vh_limb L = limbs[it.limb];
    const uint64_t q = L.q, mu = L.mu64;
    const int S = it.aux;
    const uint64_t *w = reinterpret_cast<const uint64_t *>(it.scalar);
    const uint64_t total = n * S;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t d = e / n, x = e % n;
        uint64_t acc = 0;
        for (int j = 0; j < S; j++) {
            const int k = (int)((d + S - j) % S);
            uint64_t p = vh_mulmod(vh_reduce64(it.src[j * n + x], q, mu), vh_reduce64(it.src2[k * n + x], q, mu), L);
            if (j + k >= S) p = vh_mulmod(p, w[x], L);
            acc = vh_addmod(acc, p, q);
        }
        uint64_t o = it.dst[d * n + x];
        it.dst[d * n + x] = mode == 0 ? acc : (mode == 1 ? vh_addmod(vh_reduce64(o, q, mu), acc, q)
                                                         : vh_submod(vh_reduce64(o, q, mu), acc, q));
    }
}

// ---------------------------------------------------------------------------
// Sequential rounded / floor division by K primes (rns_polynomial.c:905-976),
// all K passes fused: one thread per (coefficient x, output limb).  For pass k
// (divide prime p_k, ascending limb order) the reference computes
// t_k = (x_{p_k} + h_k) mod p_k and, for every still-active limb i,
// x_i <- (x_i + h_k mod q_i - t_k mod q_i) * p_k^-1 mod q_i; later divide
// primes are updated like any other limb before their own pass.
// consts[k * nl + limb] = {h_k mod q_limb, p_k^-1 mod q_limb, Shoup(p_k^-1)}.
// ---------------------------------------------------------------------------
struct vh_divc { uint64_t h, inv, inv_sh; };

// One thread per coefficient x of one polynomial: the K-prime chain is computed
// once, then every output limb row is updated (coalesced row accesses).
template <int KMAX>
__global__ void __launch_bounds__(256)
k_division(const vh_div_job *__restrict__ jobs, const vh_limb *__restrict__ limbs,
           const vh_divc *__restrict__ consts, uint64_t n)
{
    const vh_div_job &jb = jobs[blockIdx.y];
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int K = jb.K, nl = jb.nl;
    uint64_t t[KMAX];
    uint64_t v[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; k++)
        if (k < K) v[k] = jb.base[(uint64_t)jb.div_limb[k] * jb.stride + x];
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        if (k >= K) continue;   // (continue, not break: the loop unrolls fully)
        const int lk = jb.div_limb[k];
        const uint64_t qk = limbs[lk].q, muk = limbs[lk].mu64;
        t[k] = vh_reduce64(vh_reduce64(v[k], qk, muk) + consts[k * nl + lk].h, qk, muk);
#pragma unroll
        for (int k2 = k + 1; k2 < KMAX; k2++) {
            if (k2 >= K) continue;
            const int l2 = jb.div_limb[k2];
            const uint64_t q2 = limbs[l2].q, mu2 = limbs[l2].mu64;
            const vh_divc c2 = consts[k * nl + l2];
            uint64_t a = vh_reduce64(vh_reduce64(v[k2], q2, mu2) + c2.h, q2, mu2);
            a = vh_submod(a, vh_reduce64(t[k], q2, mu2), q2);
            v[k2] = vh_shoup(a, c2.inv, c2.inv_sh, q2);
        }
    }
    for (int o = 0; o < jb.nout; o++) {
        const int li = jb.out_limb[o];
        const uint64_t qi = limbs[li].q, mui = limbs[li].mu64;
        uint64_t *row = jb.base + (uint64_t)li * jb.stride;
        uint64_t a = vh_reduce64(row[x], qi, mui);
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (k >= K) continue;   // (continue, not break: the loop unrolls fully)
            const vh_divc ck = consts[k * nl + li];
            a = vh_addmod(a, ck.h, qi);
            a = vh_submod(a, vh_reduce64(t[k], qi, mui), qi);
            a = vh_shoup(a, ck.inv, ck.inv_sh, qi);
        }
        row[x] = a;
    }
}

// ModDown, pass-through add and CKKS rescale of one key-switch output component
// in one pass (the sequence mlwe.c:583-597 then rns_polynomial.c:948-973 for the
// top prime): per coefficient, the K special limbs' chain is formed as in
// k_division; limb `resc` is divided, gets the pass-through value added and
// becomes the one-prime chain of the rescale; every other output limb is
// divided, gets its pass-through value, and is divided by the rescale prime.
// Every step is the sequential kernels' arithmetic in their order, so results
// are identical.  The divided rows (specials and `resc`) are not written; the
// caller clears them, as the sequential divisions do.
template <int KMAX>
__global__ void __launch_bounds__(256)
k_moddown_rescale(const vh_mdr_job *__restrict__ jobs, const vh_limb *__restrict__ limbs,
                  const vh_divc *__restrict__ cdown, const vh_divc *__restrict__ cresc, uint64_t n)
{
    const vh_mdr_job &jb = jobs[blockIdx.y];
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int K = jb.K, nl = jb.nl;
    uint64_t t[KMAX], v[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; k++)
        if (k < K) v[k] = jb.base[(uint64_t)jb.div_limb[k] * jb.stride + x];
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        if (k >= K) continue;   // (continue, not break: the loop unrolls fully)
        const int lk = jb.div_limb[k];
        const uint64_t qk = limbs[lk].q, muk = limbs[lk].mu64;
        t[k] = vh_reduce64(vh_reduce64(v[k], qk, muk) + cdown[k * nl + lk].h, qk, muk);
#pragma unroll
        for (int k2 = k + 1; k2 < KMAX; k2++) {
            if (k2 >= K) continue;
            const int l2 = jb.div_limb[k2];
            const uint64_t q2 = limbs[l2].q, mu2 = limbs[l2].mu64;
            const vh_divc c2 = cdown[k * nl + l2];
            uint64_t a = vh_reduce64(vh_reduce64(v[k2], q2, mu2) + c2.h, q2, mu2);
            a = vh_submod(a, vh_reduce64(t[k], q2, mu2), q2);
            v[k2] = vh_shoup(a, c2.inv, c2.inv_sh, q2);
        }
    }
    auto down_add = [&](int li) -> uint64_t {
        const uint64_t qi = limbs[li].q, mui = limbs[li].mu64;
        uint64_t a = vh_reduce64(jb.base[(uint64_t)li * jb.stride + x], qi, mui);
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (k >= K) continue;   // (continue, not break: the loop unrolls fully)
            const vh_divc ck = cdown[k * nl + li];
            a = vh_addmod(a, ck.h, qi);
            a = vh_submod(a, vh_reduce64(t[k], qi, mui), qi);
            a = vh_shoup(a, ck.inv, ck.inv_sh, qi);
        }
        if (jb.add) a = vh_reduce64(a + jb.add[(uint64_t)li * jb.stride + x], qi, mui);
        return a;
    };
    const int lr = jb.resc;
    const uint64_t ar = down_add(lr);
    const uint64_t qr = limbs[lr].q, mur = limbs[lr].mu64;
    const uint64_t tr = vh_reduce64(vh_reduce64(ar, qr, mur) + cresc[lr].h, qr, mur);
    for (int o = 0; o < jb.nout; o++) {
        const int li = jb.out_limb[o];
        const uint64_t qi = limbs[li].q, mui = limbs[li].mu64;
        uint64_t a = down_add(li);
        const vh_divc c = cresc[li];
        a = vh_addmod(a, c.h, qi);
        a = vh_submod(a, vh_reduce64(tr, qi, mui), qi);
        jb.base[(uint64_t)li * jb.stride + x] = vh_shoup(a, c.inv, c.inv_sh, qi);
    }
}

// The same ModDown + pass-through + rescale on the FP64 path (every limb below 2^50,
// at most 4 special primes).  The K sequential round divisions of the reference
// (rns_polynomial.c:939-976, ascending special primes, then the rescale prime) are
// affine in the row values and in the carried t_k = (special row k after the
// divisions before it + floor(p_k/2)) mod p_k, so with the products of inverses
// folded into per-limb constants (built on the host, exact):
//   t_k   = canon(v_k * G_k + c_k + sum_{j<k} t_j * g_kj)                 mod p_k
//   t_r   = canon(x_r * X_r + add_r + C_r + sum_k t_k * T_rk)             mod q_r
//   out_i = canon(x_i * X_i + add_i * A_i + C_i + t_r * R_i + sum_k t_k * T_ik)  mod q_i
// Each product is vh_mulmod_tw against a centred constant ({w, RN(w/q)}): |term| <= 3q/4,
// so with K <= 4 every sum stays below 6q < 2^53 and the result is the reference's,
// bit for bit.  Constant rows (stride VH_MDR_W double2): rows 0..3 the special primes
// [G, c, g_k0..g_k3], then row 4 + limb for every limb of the pool [X, A, C, R, T_0..T_3]
// (C in .x only).
__global__ void __launch_bounds__(256)
k_moddown_rescale_fp(const vh_mdr_job *__restrict__ jobs, const vh_limb *__restrict__ limbs,
                     const double2 *__restrict__ cs, uint64_t n)
{
    constexpr int KMAX = 4, W = VH_MDR_W;
    const vh_mdr_job &jb = jobs[blockIdx.y];
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int K = jb.K;
    const VH_GLOBAL uint64_t *base = vh_global(jb.base);
    const VH_GLOBAL uint64_t *add = jb.add ? vh_global(jb.add) : nullptr;
    auto row = [&](const VH_GLOBAL uint64_t *b, int limb) -> double { return vh_u2d(b[(uint64_t)limb * jb.stride + x]); };
    double t[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        if (k >= K) continue;   // (continue, not break: the loop unrolls fully)
        const int lk = jb.div_limb[k];
        const vh_limb &L = limbs[lk];
        const double2 *c = cs + k * W;
        double s = c[1].x + vh_mulmod_tw(row(base, lk), c[0].x, c[0].y, L.qd);
#pragma unroll
        for (int j = 0; j < KMAX; j++)
            if (j < k) s += vh_mulmod_tw(t[j], c[2 + j].x, c[2 + j].y, L.qd);
        t[k] = vh_canon(s, L.qd, L.qinv);
    }
    const int lr = jb.resc;
    double tr;
    {
        const vh_limb &L = limbs[lr];
        const double2 *c = cs + (KMAX + lr) * W;
        double s = c[2].x + vh_mulmod_tw(row(base, lr), c[0].x, c[0].y, L.qd);
        if (add) s += row(add, lr);
#pragma unroll
        for (int k = 0; k < KMAX; k++)
            if (k < K) s += vh_mulmod_tw(t[k], c[4 + k].x, c[4 + k].y, L.qd);
        tr = vh_canon(s, L.qd, L.qinv);
    }
    VH_GLOBAL uint64_t *outb = vh_global(jb.base);
    for (int o = 0; o < jb.nout; o++) {
        const int li = jb.out_limb[o];
        const vh_limb &L = limbs[li];
        const double2 *c = cs + (KMAX + li) * W;
        double s = c[2].x + vh_mulmod_tw(row(base, li), c[0].x, c[0].y, L.qd) + vh_mulmod_tw(tr, c[3].x, c[3].y, L.qd);
        if (add) s += vh_mulmod_tw(row(add, li), c[1].x, c[1].y, L.qd);
#pragma unroll
        for (int k = 0; k < KMAX; k++)
            if (k < K) s += vh_mulmod_tw(t[k], c[4 + k].x, c[4 + k].y, L.qd);
        outb[(uint64_t)li * jb.stride + x] = vh_d2u(vh_canon(s, L.qd, L.qinv));
    }
}

hipError_t vh_launch_moddown_rescale_fp(hipStream_t st, const vh_mdr_job *jobs, unsigned njobs, const vh_limb *limbs,
                                        const void *cs, uint64_t n)
{
    const dim3 grid((unsigned)((n + 255) / 256), njobs);
    VH_KL(k_moddown_rescale_fp, grid, dim3(256), 0, st, jobs, limbs, (const double2 *)cs, n);
    return hipGetLastError();
}

hipError_t vh_launch_moddown_rescale(hipStream_t st, const vh_mdr_job *jobs, unsigned njobs, const vh_limb *limbs,
                                     const void *cdown, const void *cresc, uint64_t n, int kmax)
{
    const dim3 grid((unsigned)((n + 255) / 256), njobs);
    if (kmax <= 4)
        VH_KL(k_moddown_rescale<4>, grid, dim3(256), 0, st, jobs, limbs, (const vh_divc *)cdown,
                           (const vh_divc *)cresc, n);
    else
        VH_KL(k_moddown_rescale<VH_MAX_DIV>, grid, dim3(256), 0, st, jobs, limbs, (const vh_divc *)cdown,
                           (const vh_divc *)cresc, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fast base extension (rns_polynomial.c:738-822): for output limb p,
// out_p = sum_j [x_j * Dhat_j]_{q_j} * D_mod_p[p][j] mod p.  One thread per
// (x, output limb).  consts: Dhat (w entries, {val, shoup}) then D_mod_p (v*w).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_base_conversion(const vh_bconv_job *__restrict__ jobs, const vh_limb *__restrict__ limbs,
                  const ulonglong2 *__restrict__ consts, uint64_t n)
{
    // by reference: the job's limb lists are indexed with loop variables, which a private
    // copy would place in scratch memory
    const vh_bconv_job &jb = jobs[blockIdx.y];
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int t = blockIdx.z;
    const int P = jb.P[t];
    const vh_limb &Lp = limbs[P];
    uint64_t acc = 0;
    for (int j = 0; j < jb.w; j++) {
        const vh_limb &Lj = limbs[jb.D[j]];
        const ulonglong2 dh = consts[j];
        uint64_t y = vh_shoup(jb.in[(uint64_t)jb.D[j] * jb.stride + x], dh.x, dh.y, Lj.q);
        const ulonglong2 dm = consts[jb.w + t * jb.w + j];
        acc = vh_addmod(acc, vh_shoup(vh_reduce64(y, Lp.q, Lp.mu64), dm.x, dm.y, Lp.q), Lp.q);
    }
    jb.out[(uint64_t)P * jb.stride + x] = acc;
}

// ---------------------------------------------------------------------------
// Automorphism X -> X^gen (rns_polynomial.c:996-1035) in the split block
// layout: the coefficient at block i, position i2 is X^(i + i2*S).  One thread
// per source coefficient; items: src, dst, aux = gen, scalar = S.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_permute(const vh_item *__restrict__ items, const vh_limb *__restrict__ limbs, uint64_t N)
{
    const vh_item it = items[blockIdx.y];
    const vh_limb L = limbs[it.limb];
    const uint64_t S = it.scalar, n = N / S, gen = it.aux;
    int slog = 0;
    while ((1ULL << slog) < S) slog++;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = e / n, i2 = e % n;
        const uint64_t idx = (i + (i2 << slog)) * gen;
        const uint64_t dst = (idx & (S - 1)) * n + ((idx & (N - 1)) >> slog);
        int64_t v = (int64_t)it.src[e];
        if (idx & N) v = (int64_t)((uint64_t)0 - (uint64_t)v);
        uint64_t a = vh_reduce64(v < 0 ? (uint64_t)0 - (uint64_t)v : (uint64_t)v, L.q, L.mu64);
        it.dst[dst] = (v < 0 && a) ? L.q - a : a;
    }
}

// ---------------------------------------------------------------------------
// Key-switch accumulation (gadget.c:22-37 + mlwe.c:272-279 for every digit):
// out_c[t][x] = -sum_d ksk_d.c[t][x] * modup_d[t][x] mod q_t, all digits d of
// every non-NULL key at once.  One thread per (x, extended limb t); it reads the
// lifted-and-transformed digit once and every component's key row.
// ---------------------------------------------------------------------------
template <bool FP>
__global__ void __launch_bounds__(256)
k_keyswitch_mac(const vh_mac_job *__restrict__ jobs, const vh_limb *__restrict__ limbs,
                const uint64_t *const *__restrict__ kptr, uint64_t n)
{
    // blockIdx.x runs over the batch, so consecutive workgroups read the same key rows
    // (same x range, same limb): the evaluation key is fetched once per batch.
    const vh_mac_job &jb = jobs[blockIdx.x];
    const uint64_t x = (uint64_t)blockIdx.y * blockDim.x + threadIdx.x;
    const int ti = blockIdx.z;
    if (x >= n || ti >= jb.next) return;
    const int limb = jb.limb[ti];
    const vh_limb &L = limbs[limb];
    const int C = jb.comps, D = jb.digits;
    if constexpr (FP) {
        // Canonical operands below 2^50: exact product (hi + lo), quotient
        // rint(m * (k/q)) within 1 of m*k/q, so r = m*k - qt*q is exact with
        // |r| <= q + 2^47; the running sum is re-centred after every term.
        const double q = L.qd, qinv = L.qinv;
        double acc[VH_MAX_COMPS];
#pragma unroll
        for (int c = 0; c < VH_MAX_COMPS; c++) acc[c] = 0.0;
        for (int d = 0; d < D; d++) {
            const double m = vh_u2d(jb.modup[((uint64_t)d * jb.next + ti) * n + x]);
            const uint64_t *const *kp = kptr + jb.kptr_off + ((uint64_t)ti * D + d) * C;
#pragma unroll
            for (int c = 0; c < VH_MAX_COMPS; c++) {
                if (c >= C) break;
                const double k = vh_u2d(kp[c][x]);
                const double hi = m * k;
                const double lo = __fma_rn(m, k, -hi);
                const double qt = __fma_rn(m, k * qinv, VH_MAGIC) - VH_MAGIC;
                const double r = __fma_rn(-qt, q, hi) + lo;
                acc[c] = vh_center(acc[c] + r, q, qinv);
            }
        }
#pragma unroll
        for (int c = 0; c < VH_MAX_COMPS; c++) {
            if (c >= C) break;
            jb.out[c][(uint64_t)limb * n + x] = vh_d2u(vh_canon(jb.add ? acc[c] : -acc[c], q, qinv));
        }
    } else {
        const uint64_t q = L.q, mu = L.mu64;
        uint64_t acc[VH_MAX_COMPS];
#pragma unroll
        for (int c = 0; c < VH_MAX_COMPS; c++) acc[c] = 0;
        for (int d = 0; d < D; d++) {
            const uint64_t m = jb.modup[((uint64_t)d * jb.next + ti) * n + x];
            const uint64_t *const *kp = kptr + jb.kptr_off + ((uint64_t)ti * D + d) * C;
#pragma unroll
            for (int c = 0; c < VH_MAX_COMPS; c++)
                if (c < C) acc[c] = vh_addmod(acc[c], vh_mulmod(vh_reduce64(kp[c][x], q, mu), m, L), q);
        }
#pragma unroll
        for (int c = 0; c < VH_MAX_COMPS; c++) {
            if (c >= C) break;
            const uint64_t a = acc[c];
            jb.out[c][(uint64_t)limb * n + x] = jb.add ? a : (a ? q - a : 0);
        }
    }
}

// ---------------------------------------------------------------------------
// Lift with split placement for int_array_to_RNS (rns_polynomial.c:204-223):
// dst[(j mod S) * n + j / S] = reduce_signed(src[j]).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_scatter_split(const vh_item *__restrict__ items, const vh_limb *__restrict__ limbs, uint64_t N, int signed_in)
{
    const vh_item it = items[blockIdx.y];
    const vh_limb L = limbs[it.limb];
    const uint64_t S = it.scalar, n = N / S;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t r;
        if (signed_in) {
            int64_t v = (int64_t)it.src[j];
            uint64_t a = vh_reduce64(v < 0 ? (uint64_t)0 - (uint64_t)v : (uint64_t)v, L.q, L.mu64);
            r = (v < 0 && a) ? L.q - a : a;
        } else {
            r = it.src[j];
        }
        it.dst[(j & (S - 1)) * n + j / S] = r;
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
static inline unsigned vh_blocks(uint64_t n, unsigned per = 256, unsigned cap = 1024)
{
    uint64_t b = (n + per - 1) / per;
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

// ---------------------------------------------------------------------------
// Whole transforms of N = 2^9 .. 2^12 points on the FP64 path, 16 values per thread in
// registers: T = N/16 threads per transform, 4096/N transforms per 256-thread workgroup.
// The reference's butterfly network (merged CT forward, natural in, bit-reversed out;
// GS inverse with n^-1, ntt.c / ntt50.c) in three register phases:
//   A  stages 0..3         stride ownership   j = t + T m            (uniform twiddles)
//   B  stages 4..7         middle ownership   j = 16U g + u + U m    (t = u + U g, U = T/16)
//   C  stages 8..logN-1    contiguous         j = 16 t + m
// with an LDS exchange between phases (one pad slot per 16 values: the contiguous reads
// stride 17 words).  A stage-s butterfly on (j, j + N/2^(s+1)) takes the twiddle of index
// 2^s + (j >> (logN - s)), so every product is the reference's and the canonical outputs
// are identical.  Rows are loaded and stored in the stride ownership (whole 128-byte runs
// per instruction); the forward pass regroups its output through LDS for that.
// Bounds as in the two-pass kernels: at most four stages between centrings, |x| < 2^52.
// ---------------------------------------------------------------------------
// Register cores of the transform for T = N/16 threads (thread index t < T); lds holds
// N + N/16 doubles private to the transform.  WAVE: the transform's threads are one wave
// (or part of one), so the exchanges need only the wave's own ordering.
template <int LOGN> struct vh_reg_ntt {
    static constexpr int N = 1 << LOGN, T = N / 16, U = T / 16;
    static __device__ __forceinline__ int li(int j) { return j + (j >> 4); }
    static __device__ __forceinline__ int jA(int t, int m) { return t + T * m; }
    static __device__ __forceinline__ int jB(int t, int m) { return 16 * U * (t / U) + t % U + U * m; }
    static __device__ __forceinline__ int jC(int t, int m) { return 16 * t + m; }
    template <bool WAVE> static __device__ __forceinline__ void sync()
    {
        if (WAVE) vh_wave_sync();
        else __syncthreads();
    }
    template <bool WAVE, typename FW, typename FR>
    static __device__ __forceinline__ void exchange(double (&v)[16], double *lds, FW jw, FR jr, bool centre, double q,
                                                    double qinv)
    {
#pragma unroll
        for (int m = 0; m < 16; m++) lds[li(jw(m))] = centre ? vh_center(v[m], q, qinv) : v[m];
        sync<WAVE>();
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = lds[li(jr(m))];
        sync<WAVE>();
    }
    // forward: v in the stride ownership (any |v| < 2^51) -> contiguous ownership, |v| < 2^52
    template <bool WAVE>
    static __device__ __forceinline__ void fwd(double (&v)[16], double *lds, int t, const VH_GLOBAL double *tw,
                                               double q, double qinv)
    {
        const int g = t / U;
        auto ct = [&](int s, int hm, auto blk) {
#pragma unroll
            for (int m = 0; m < 16; m++) {
                if (m & hm) continue;
                const int i = (1 << s) + blk(m);
                const double x = vh_mulmod_tw(v[m + hm], tw[2 * i], tw[2 * i + 1], q);
                const double a = v[m];
                v[m] = a + x;
                v[m + hm] = a - x;
            }
        };
#pragma unroll
        for (int s = 0; s < 4; s++) ct(s, 8 >> s, [&](int m) { return m >> (4 - s); });
        exchange<WAVE>(v, lds, [&](int m) { return jA(t, m); }, [&](int m) { return jB(t, m); }, true, q, qinv);
#pragma unroll
        for (int s = 4; s < 8; s++) ct(s, 8 >> (s - 4), [&](int m) { return (g << (s - 4)) + (m >> (8 - s)); });
        exchange<WAVE>(v, lds, [&](int m) { return jB(t, m); }, [&](int m) { return jC(t, m); }, true, q, qinv);
#pragma unroll
        for (int s = 8; s < LOGN; s++) ct(s, 1 << (LOGN - s - 1), [&](int m) { return (16 * t + m) >> (LOGN - s); });
    }
    // inverse (before the n^-1 factor): v in the contiguous ownership (|v| <= q) -> stride
    // ownership, |v| < 2^52; GS sums re-centred after every second stage
    template <bool WAVE>
    static __device__ __forceinline__ void inv(double (&v)[16], double *lds, int t, const VH_GLOBAL double *tw,
                                               double q, double qinv)
    {
        const int g = t / U;
        int c = 0;
        auto gs = [&](int s, int hm, auto blk) {
#pragma unroll
            for (int m = 0; m < 16; m++) {
                if (m & hm) continue;
                const int i = (1 << s) + blk(m);
                const double a = v[m], x = v[m + hm];
                double sum = a + x;
                if (c & 1) sum = vh_center(sum, q, qinv);
                v[m] = sum;
                v[m + hm] = vh_mulmod_tw(a - x, tw[2 * i], tw[2 * i + 1], q);
            }
            c++;
        };
#pragma unroll
        for (int s = LOGN - 1; s >= 8; s--) gs(s, 1 << (LOGN - s - 1), [&](int m) { return (16 * t + m) >> (LOGN - s); });
        exchange<WAVE>(v, lds, [&](int m) { return jC(t, m); }, [&](int m) { return jB(t, m); }, false, q, qinv);
#pragma unroll
        for (int s = 7; s >= 4; s--) gs(s, 8 >> (s - 4), [&](int m) { return (g << (s - 4)) + (m >> (8 - s)); });
        exchange<WAVE>(v, lds, [&](int m) { return jB(t, m); }, [&](int m) { return jA(t, m); }, false, q, qinv);
#pragma unroll
        for (int s = 3; s >= 0; s--) gs(s, 8 >> s, [&](int m) { return m >> (4 - s); });
    }
};

template <int DIR, int LOGN>
__global__ void __launch_bounds__(256)
k_ntt_small(const vh_item *__restrict__ items, unsigned nitems, const vh_limb *__restrict__ limbs)
{
    using R = vh_reg_ntt<LOGN>;
    constexpr int N = R::N, T = R::T, PER = 256 / T, PAD = N + N / 16;
    constexpr bool WAVE = T <= 64;
    static_assert(LOGN >= 9 && LOGN <= 12, "k_ntt_small: N = 2^9 .. 2^12");
    __shared__ double lds_all[PER * PAD];
    const int tid = threadIdx.x, slot = tid / T, t = tid % T;
    const unsigned idx = blockIdx.x * PER + slot;
    const bool live = idx < nitems;
    // a slot past the end transforms the last item again (every thread reaches the
    // barriers) and stores nothing
    const vh_item &it = items[live ? idx : nitems - 1];
    const vh_limb &L = limbs[it.limb];
    const double q = L.qd, qinv = L.qinv;
    const VH_GLOBAL double *tw = vh_global(reinterpret_cast<const double *>(DIR == 0 ? L.tw_fwd : L.tw_inv));
    double *lds = lds_all + slot * PAD;
    const VH_GLOBAL uint64_t *in = vh_global(it.src);
    VH_GLOBAL uint64_t *out = vh_global(it.dst);
    double v[16];
#pragma unroll
    for (int m = 0; m < 16; m++) v[m] = vh_u2d(vh_ld(in + R::jA(t, m)));
    if (DIR == 0) {
        R::template fwd<WAVE>(v, lds, t, tw, q, qinv);
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = __longlong_as_double((long long)vh_d2u(vh_canon(v[m], q, qinv)));
        R::template exchange<WAVE>(v, lds, [&](int m) { return R::jC(t, m); }, [&](int m) { return R::jA(t, m); },
                                   false, q, qinv);
        if (live) {
#pragma unroll
            for (int m = 0; m < 16; m++) vh_st(out + R::jA(t, m), (uint64_t)__double_as_longlong(v[m]));
        }
    } else {
        R::template exchange<WAVE>(v, lds, [&](int m) { return R::jA(t, m); }, [&](int m) { return R::jC(t, m); },
                                   false, q, qinv);
        R::template inv<WAVE>(v, lds, t, tw, q, qinv);
        if (live) {
#pragma unroll
            for (int m = 0; m < 16; m++)
                vh_st(out + R::jA(t, m), vh_d2u(vh_canon(vh_mulmod_tw(v[m], L.ninv_d, L.ninv_q, q), q, qinv)));
        }
    }
}

// ---------------------------------------------------------------------------
// CGGI16 blind rotation of a batch, every key index in one launch (cggi16.py:70-90 with
// mgsw.c:51-71 per index).  N = 1024, FP64 limbs, one special prime.  Workgroup k holds
// accumulator k's level rows in LDS (coefficient form) for the whole loop; wave w works
// modulo extended limb E[w].  Per key index i with a = amounts[k n + i] != 0:
//   digit d = (component c, level limb J[jj]):  (X^a - 1) acc_c  at limb J[jj], canonical,
//     read from LDS straight into the forward transform's registers (the reference's
//     mlwe_RNSc_mul_by_xai, subtraction and gadget lift; the lift is the identity on
//     values below q_J[jj] < 2^50, which the FP64 transform accepts for any E[w]);
//   out_c' += NTT_E[w](digit) * bk[i][d] component c' (FP64 products, centred every 4th
//     digit), the 16 key words per lane contiguous;
//   then per c': inverse transform, the special wave publishes t = (x + floor(p/2)) mod p
//     and every level wave adds (x + floor(p/2) - t) p^-1 mod q to its accumulator row
//     (round division by p, then the CMUX's  acc + ExtProd(X^a acc - acc)).
// A zero exponent leaves the accumulator as it is (the reference skips that index).
// Every wave takes part in every barrier; the exponent is uniform over the workgroup.
// ---------------------------------------------------------------------------
template <int C, int NJ>
__global__ void __launch_bounds__(256, 2)
k_blind_rotate(const vh_br_params P, const vh_limb *__restrict__ limbs, uint64_t *const *__restrict__ accp,
               const uint64_t *const *__restrict__ bkp, const uint64_t *__restrict__ amounts,
               const uint64_t *const *__restrict__ tvp, const uint64_t *__restrict__ rot)
{
    using R = vh_reg_ntt<10>;
    constexpr int N = 1024, PAD = N + N / 16, nJ = NJ, next = NJ + 1, nd = C * NJ;
    // C NJ rows of the accumulator, one exchange buffer per wave, the special limb's t:
    // 65.5 KiB for C = 2, NJ = 2 (two workgroups per CU)
    __shared__ double dyn[C * NJ * N + next * PAD + N];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned k = blockIdx.x;
    double *accd = dyn;                           // [C][nJ][N]
    double *scratch = accd + C * nJ * N + w * PAD; // this wave's exchange buffer
    double *spd = dyn + C * nJ * N + next * PAD;   // [N]: the special limb's t of one component
    // the accumulator's rows (every wave helps): its own, or X^rot[k] * tv (the initial
    // rotation of the test vector, mlwe_RNSc_mul_by_xai: coefficient u = (x - a) mod 2N
    // of tv, negated when u >= N)
    for (int r = 0; r < C * nJ; r++) {
        if (tvp) {
            const VH_GLOBAL uint64_t *src = vh_global(tvp[r]);
            const uint32_t a = (uint32_t)(rot[k] & (2 * N - 1));
            const double qr = limbs[P.J[r % nJ]].qd;
            for (int x = threadIdx.x; x < N; x += blockDim.x) {
                const uint32_t u = (uint32_t)(x + 2 * N - a) & (2 * N - 1);
                const double sv = vh_u2d(src[u & (N - 1)]);
                accd[r * N + x] = (u < N || sv == 0.0) ? sv : qr - sv;
            }
        } else {
            const VH_GLOBAL uint64_t *src = vh_global(accp[(size_t)k * C * nJ + r]);
            for (int x = threadIdx.x; x < N; x += blockDim.x) accd[r * N + x] = vh_u2d(src[x]);
        }
    }
    __syncthreads();
    const vh_limb &L = limbs[P.E[w]];
    const double q = L.qd, qinv = L.qinv;
    const VH_GLOBAL double *twf = vh_global(reinterpret_cast<const double *>(L.tw_fwd));
    const VH_GLOBAL double *twi = vh_global(reinterpret_cast<const double *>(L.tw_inv));
    const VH_GLOBAL uint64_t *am = vh_global(amounts + (size_t)k * P.n);
    for (uint64_t i = 0; i < P.n; i++) {
        const uint32_t a = (uint32_t)(am[i] & (2 * N - 1));
        if (a == 0) continue;
        double out[C][16];
#pragma unroll
        for (int c = 0; c < C; c++)
#pragma unroll
            for (int m = 0; m < 16; m++) out[c][m] = 0.0;
        for (int d = 0; d < nd; d++) {
            const int c = d / nJ, jj = d - c * nJ;
            const double qj = limbs[P.J[jj]].qd;
            const double *row = accd + (c * nJ + jj) * N;
            double v[16];
#pragma unroll
            for (int m = 0; m < 16; m++) {
                const int x = R::jA(lane, m);
                const uint32_t u = (uint32_t)(x + 2 * N - a) & (2 * N - 1);
                const double s = row[u & (N - 1)];
                double r = (u < N ? s : -s) - row[x];     // in (-2 q_j, q_j)
                r = r < 0.0 ? r + qj : r;
                v[m] = r < 0.0 ? r + qj : r;
            }
            R::template fwd<true>(v, scratch, lane, twf, q, qinv);
            const uint64_t *const *kp = bkp + ((i * nd + d) * C) * next + w;
#pragma unroll
            for (int c2 = 0; c2 < C; c2++) {
                const VH_GLOBAL uint64_t *kr = vh_global(kp[(size_t)c2 * next]) + 16 * lane;
                uint64_t kw[16];
#pragma unroll
                for (int m = 0; m < 16; m++) kw[m] = kr[m];
#pragma unroll
                for (int m = 0; m < 16; m++) {
                    const double x = vh_center(v[m], q, qinv);
                    const double kk = vh_u2d(kw[m]);
                    const double hi = x * kk;
                    const double lo = __fma_rn(x, kk, -hi);
                    const double qt = __fma_rn(x, kk * qinv, VH_MAGIC) - VH_MAGIC;
                    const double r = __fma_rn(-qt, q, hi) + lo;   // |r| <= q + 2^47
                    out[c2][m] = (d & 3) == 3 ? vh_center(out[c2][m] + r, q, qinv) : out[c2][m] + r;
                }
            }
        }
        __syncthreads();   // every wave has read the accumulator rows of this index
#pragma unroll
        for (int c2 = 0; c2 < C; c2++) {
            double v[16];
#pragma unroll
            for (int m = 0; m < 16; m++) v[m] = vh_center(out[c2][m], q, qinv);
            R::template inv<true>(v, scratch, lane, twi, q, qinv);
#pragma unroll
            for (int m = 0; m < 16; m++) v[m] = vh_canon(vh_mulmod_tw(v[m], L.ninv_d, L.ninv_q, q), q, qinv);
            if (w == P.sp) {
#pragma unroll
                for (int m = 0; m < 16; m++) spd[R::jA(lane, m)] = vh_canon(v[m] + P.hp, q, qinv);
            }
            __syncthreads();
            if (w != P.sp) {
                double *row = accd + (c2 * nJ + P.jw[w]) * N;
                const double h = P.hdiv[w], iv = P.inv[w], ivq = P.invq[w];
#pragma unroll
                for (int m = 0; m < 16; m++) {
                    const int x = R::jA(lane, m);
                    const double y = vh_mulmod_tw(v[m] + h - spd[x], iv, ivq, q);
                    row[x] = vh_canon(row[x] + y, q, qinv);
                }
            }
            __syncthreads();
        }
    }
    for (int r = 0; r < C * nJ; r++) {
        VH_GLOBAL uint64_t *dst = vh_global(accp[(size_t)k * C * nJ + r]);
        for (int x = threadIdx.x; x < N; x += blockDim.x) dst[x] = vh_d2u(accd[r * N + x]);
    }
}

hipError_t vh_launch_blind_rotate(hipStream_t st, const vh_br_params &P, const vh_limb *limbs, uint64_t *const *accp,
                                  const uint64_t *const *bkp, const uint64_t *amounts, unsigned count,
                                  const uint64_t *const *tvp, const uint64_t *rot)
{
    if (P.C < 1 || P.C > 2 || P.nJ < 1 || P.nJ > VH_BR_MAXJ || P.next != P.nJ + 1 || P.next > VH_BR_MAXE ||
        P.sp < 0 || P.sp >= P.next || count == 0 || (tvp && !rot))
        return hipErrorInvalidValue;
    const dim3 grid(count), block(64 * P.next);
#define VH_BR(CC, JJ)                                                                                      \
    if (P.C == CC && P.nJ == JJ) VH_KL((k_blind_rotate<CC, JJ>), grid, block, 0, st, P, limbs, accp, bkp, amounts, tvp, rot);
    VH_BR(1, 1) VH_BR(1, 2) VH_BR(1, 3) VH_BR(2, 1) VH_BR(2, 2) VH_BR(2, 3)
#undef VH_BR
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched sample extraction: one thread per output word (grid-stride).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_extract_lwe(const uint64_t *const *__restrict__ src,
                                                     const uint64_t *__restrict__ qs, uint64_t *__restrict__ out,
                                                     uint32_t N, uint32_t r, uint32_t na, uint32_t idx, uint64_t total)
{
    const uint64_t width = (uint64_t)r * N + 1;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < total; g += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t rowi = g / width, e = g - rowi * width;   // rowi = k na + j
        const uint32_t j = (uint32_t)(rowi % na);
        const uint64_t *const *rows = src + rowi * (r + 1);
        uint64_t v;
        if (e == (uint64_t)r * N) {
            v = rows[r][idx];
        } else {
            const uint32_t c = (uint32_t)(e / N), i = (uint32_t)(e - (uint64_t)c * N);
            if (i <= idx) {
                v = rows[c][idx - i];
            } else {
                const uint64_t x = rows[c][N + idx - i];
                v = x ? qs[j] - x : 0;
            }
        }
        out[g] = v;
    }
}

hipError_t vh_launch_extract_lwe(hipStream_t st, const uint64_t *const *src, const uint64_t *qs, uint64_t *out,
                                 uint32_t N, uint32_t r, uint32_t na, uint32_t idx, uint64_t count)
{
    const uint64_t total = count * na * ((uint64_t)r * N + 1);
    if (!total) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<uint64_t>((total + 255) / 256, 65536);
    VH_KL(k_extract_lwe, dim3(blocks), dim3(256), 0, st, src, qs, out, N, r, na, idx, total);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ALU rate probes: each thread runs 8 independent chains of `iters` operations of one
// kind (the chains keep the pipeline full), then stores a value that depends on all of
// them so nothing is optimised away.  kind: 0 FP64 FMA, 1 32-bit multiply (low half) +
// add, 2 32-bit multiply high half, 3 32x32->64 multiply-add, 4 FP64 modular product
// (vh_mulmod_tw), 5 64-bit Shoup modular product (vh_shoup).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_alu_rate(int kind, int iters, uint64_t *out)
{
    const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t acc = 0;
    if (kind == 0) {
        double x[8];
        for (int k = 0; k < 8; k++) x[k] = 1.0 + 1e-9 * (id + k);
        const double a = 0.999999999, b = 1e-12;
        for (int i = 0; i < iters; i++)
#pragma unroll
            for (int k = 0; k < 8; k++) x[k] = __fma_rn(x[k], a, b);
        for (int k = 0; k < 8; k++) acc += (uint64_t)__double_as_longlong(x[k]);
    } else if (kind == 1 || kind == 2) {
        uint32_t x[8];
        for (int k = 0; k < 8; k++) x[k] = 0x9E3779B9u * (id + k) + 7u;
        const uint32_t a = 0x85EBCA6Bu, b = 0xC2B2AE35u;
        for (int i = 0; i < iters; i++)
#pragma unroll
            for (int k = 0; k < 8; k++) x[k] = kind == 1 ? x[k] * a + b : __umulhi(x[k], a) ^ b;
        for (int k = 0; k < 8; k++) acc += x[k];
    } else if (kind == 3) {
        uint64_t x[8];
        for (int k = 0; k < 8; k++) x[k] = 0x9E3779B97F4A7C15ull * (id + k);
        const uint32_t a = 0x85EBCA6Bu;
        for (int i = 0; i < iters; i++)
#pragma unroll
            for (int k = 0; k < 8; k++) x[k] = (uint64_t)(uint32_t)x[k] * a + (x[k] >> 32);
        for (int k = 0; k < 8; k++) acc += x[k];
    } else if (kind == 4) {
        const double q = 1125899906842597.0, w = 392817365281.0, wq = w / q;   // a prime below 2^50
        double x[8];
        for (int k = 0; k < 8; k++) x[k] = (double)((id * 8 + k) % 1000003);
        for (int i = 0; i < iters; i++)
#pragma unroll
            for (int k = 0; k < 8; k++) x[k] = vh_mulmod_tw(x[k], w, wq, q);
        for (int k = 0; k < 8; k++) acc += (uint64_t)__double_as_longlong(x[k]);
    } else {
        const uint64_t q = 1125899906842597ull, w = 392817365281ull;
        const uint64_t wsh = (uint64_t)(((unsigned __int128)w << 64) / q);
        uint64_t x[8];
        for (int k = 0; k < 8; k++) x[k] = (id * 8ull + k) % q;
        for (int i = 0; i < iters; i++)
#pragma unroll
            for (int k = 0; k < 8; k++) x[k] = vh_shoup(x[k], w, wsh, q);
        for (int k = 0; k < 8; k++) acc += x[k];
    }
    out[id] = acc;
}

hipError_t vh_launch_alu_rate(hipStream_t st, int kind, int iters, unsigned blocks, uint64_t *out)
{
    VH_KL(k_alu_rate, dim3(blocks), dim3(256), 0, st, kind, iters, out);
    return hipGetLastError();
}

hipError_t vh_launch_ntt_small(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, int dir,
                               int logn)
{
    if (logn < 9 || logn > 12 || nitems == 0) return hipErrorInvalidValue;
    const unsigned per = 4096u >> logn, nb = (nitems + per - 1) / per;
#define VH_SMALL(LG)                                                                                        \
    if (logn == LG) {                                                                                       \
        if (dir == 0) VH_KL((k_ntt_small<0, LG>), dim3(nb), dim3(256), 0, st, items, nitems, limbs);         \
        else VH_KL((k_ntt_small<1, LG>), dim3(nb), dim3(256), 0, st, items, nitems, limbs);                  \
    }
    VH_SMALL(9) VH_SMALL(10) VH_SMALL(11) VH_SMALL(12)
#undef VH_SMALL
    return hipGetLastError();
}

hipError_t vh_launch_ntt_pass(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                              int fp, int dir, int logn, int s0, int R, int col, int G, int first, int last)
{
    const int S = 1 << R;
    const unsigned nsub = 1u << (logn - R);
    dim3 grid(nsub / G, nitems);
    int thr = G * S / 2;
    thr = thr < 64 ? 64 : (thr > 256 ? 256 : thr);
    const size_t lds = (size_t)G * S * 8;
    if (fp) {
        if (dir == 0) VH_KL((k_ntt_pass<0, true>), grid, dim3(thr), lds, st, items, limbs, logn, s0, R, col, G, first, last);
        else VH_KL((k_ntt_pass<1, true>), grid, dim3(thr), lds, st, items, limbs, logn, s0, R, col, G, first, last);
    } else {
        if (dir == 0) VH_KL((k_ntt_pass<0, false>), grid, dim3(thr), lds, st, items, limbs, logn, s0, R, col, G, first, last);
        else VH_KL((k_ntt_pass<1, false>), grid, dim3(thr), lds, st, items, limbs, logn, s0, R, col, G, first, last);
    }
    return hipGetLastError();
}

// Rank-1 tensor product of one limb row (mlwe.c:665-708): d0 = a1 a2,
// d1 = a1 b2 + b1 a2, d2 = b1 b2, in one pass over the four inputs (the
// reference's slot order and arithmetic; canonical results).
__global__ void __launch_bounds__(256)
k_tensor1(const vh_tensor_job *__restrict__ jobs, const vh_limb *__restrict__ limbs, uint64_t n)
{
    const vh_tensor_job jb = jobs[blockIdx.y];
    const vh_limb L = limbs[jb.limb];
    const uint64_t q = L.q, mu = L.mu64;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t a1 = vh_reduce64(jb.a1[i], q, mu), b1 = vh_reduce64(jb.b1[i], q, mu);
        const uint64_t a2 = vh_reduce64(jb.a2[i], q, mu), b2 = vh_reduce64(jb.b2[i], q, mu);
        jb.d0[i] = vh_mulmod(a1, a2, L);
        jb.d1[i] = vh_addmod(vh_mulmod(a1, b2, L), vh_mulmod(b1, a2, L), q);
        jb.d2[i] = vh_mulmod(b1, b2, L);
    }
}

hipError_t vh_launch_tensor1(hipStream_t st, const vh_tensor_job *jobs, unsigned njobs, const vh_limb *limbs,
                             uint64_t n)
{
    VH_KL(k_tensor1, dim3(vh_blocks(n, 256, 64), njobs), dim3(256), 0, st, jobs, limbs, n);
    return hipGetLastError();
}

hipError_t vh_launch_eltwise(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                             int op, uint64_t n)
{
    VH_KL(k_eltwise, dim3(vh_blocks(n, 256, 64), nitems), dim3(256), 0, st, items, limbs, op, n);
    return hipGetLastError();
}

hipError_t vh_launch_split_mul(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                               int mode, uint64_t n, uint64_t S)
{
    VH_KL(k_split_mul, dim3(vh_blocks(n * S, 256, 64), nitems), dim3(256), 0, st, items, limbs, mode, n);
    return hipGetLastError();
}

hipError_t vh_launch_division(hipStream_t st, const vh_div_job *jobs, unsigned njobs, unsigned nout,
                              const vh_limb *limbs, const void *consts, uint64_t n, int kmax)
{
    (void)nout;
    if (kmax <= 4)
        VH_KL(k_division<4>, dim3((unsigned)((n + 255) / 256), njobs), dim3(256), 0, st, jobs, limbs,
                           (const vh_divc *)consts, n);
    else
        VH_KL(k_division<VH_MAX_DIV>, dim3((unsigned)((n + 255) / 256), njobs), dim3(256), 0, st, jobs,
                           limbs, (const vh_divc *)consts, n);
    return hipGetLastError();
}

hipError_t vh_launch_base_conversion(hipStream_t st, const vh_bconv_job *jobs, unsigned njobs, unsigned nout,
                                     const vh_limb *limbs, const void *consts, uint64_t n)
{
    VH_KL(k_base_conversion, dim3((unsigned)((n + 255) / 256), njobs, nout), dim3(256), 0, st, jobs,
                       limbs, (const ulonglong2 *)consts, n);
    return hipGetLastError();
}

hipError_t vh_launch_permute(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, uint64_t N)
{
    VH_KL(k_permute, dim3(vh_blocks(N, 256, 64), nitems), dim3(256), 0, st, items, limbs, N);
    return hipGetLastError();
}

hipError_t vh_launch_keyswitch_mac(hipStream_t st, const vh_mac_job *jobs, unsigned njobs, unsigned next,
                                   const vh_limb *limbs, const uint64_t *const *kptr, uint64_t n, int fp)
{
    const dim3 grid(njobs, (unsigned)((n + 255) / 256), next);
    if (fp) VH_KL(k_keyswitch_mac<true>, grid, dim3(256), 0, st, jobs, limbs, kptr, n);
    else VH_KL(k_keyswitch_mac<false>, grid, dim3(256), 0, st, jobs, limbs, kptr, n);
    return hipGetLastError();
}

hipError_t vh_launch_scatter_split(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                                   uint64_t N, int signed_in)
{
    VH_KL(k_scatter_split, dim3(vh_blocks(N, 256, 64), nitems), dim3(256), 0, st, items, limbs, N,
                       signed_in);
    return hipGetLastError();
}

hipError_t vh_launch_ntt16(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                           int dir, int logn, int s0, int R, int col, int first, int last)
{
    if (R < 4 || R > 8 || logn < R) return hipErrorInvalidValue;
    const unsigned G = 256u >> (R - 4);
    const unsigned nsub = 1u << (logn - R);
    if (nsub % G) return hipErrorInvalidValue;
    dim3 grid(nsub / G, nitems);
    // (dir, first, last, col) take the four combinations a two-pass transform uses;
    // first == 2 marks a forward first pass whose source rows are shared by items.
#define VH_NTT16(D, RR, CC, FF, LL, SS) VH_KL((k_ntt16<D, RR, CC, FF, LL, SS>), grid, dim3(256), 0, st, items, limbs, logn, s0)
#define VH_NTT16_R(D, CC, FF, LL, SS) \
    switch (R) { case 4: VH_NTT16(D, 4, CC, FF, LL, SS); break; case 5: VH_NTT16(D, 5, CC, FF, LL, SS); break; \
                 case 6: VH_NTT16(D, 6, CC, FF, LL, SS); break; case 7: VH_NTT16(D, 7, CC, FF, LL, SS); break; \
                 default: VH_NTT16(D, 8, CC, FF, LL, SS); break; }
    if (dir == 0 && col && first == 2 && !last) { VH_NTT16_R(0, true, true, false, true) }
    else if (dir == 0 && col && first && !last) { VH_NTT16_R(0, true, true, false, false) }
    else if (dir == 0 && !col && !first && last) { VH_NTT16_R(0, false, false, true, false) }
    else if (dir == 1 && !col && first == 1 && !last) { VH_NTT16_R(1, false, true, false, false) }
    else if (dir == 1 && col && !first && last) { VH_NTT16_R(1, true, false, true, false) }
    else return hipErrorInvalidValue;
#undef VH_NTT16_R
#undef VH_NTT16
    return hipGetLastError();
}

hipError_t vh_launch_ntt16p(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, int dir,
                            int logn, int s0, int R, int col, int first, int last, unsigned nblocks, int wide)
{
    if (logn < 13 || logn > 16 || R < 5 || R > 8 || (!col && R != 8) || nblocks == 0) return hipErrorInvalidValue;
    if (col ? (s0 != 0 || R != logn - 8) : (s0 != logn - 8)) return hipErrorInvalidValue;
    if (!col) wide = 0;
    const unsigned G = (wide ? 512u : 256u) >> (R - 4);
    const unsigned nsub = 1u << (logn - R);
    if (nsub % G) return hipErrorInvalidValue;
    const unsigned tpi = nsub / G, ntiles = tpi * nitems;
    if (wide && nblocks != UINT_MAX) nblocks = nblocks / 2 ? nblocks / 2 : 1;   // same threads per CU
    const dim3 grid(nblocks < ntiles ? nblocks : ntiles);
#define VH_P(D, RR, CC, FF, LL, SS) VH_KL((k_ntt16p<D, RR, CC, FF, LL, SS, 256>), grid, dim3(256), 0, st, items, limbs, logn, s0, tpi, ntiles)
#define VH_PC(D, RR, FF, LL, SS) \
    do { if (wide) VH_KL((k_ntt16p<D, RR, true, FF, LL, SS, 512>), grid, dim3(512), 0, st, items, limbs, logn, s0, tpi, ntiles); \
         else VH_P(D, RR, true, FF, LL, SS); } while (0)
#define VH_P_R(D, FF, LL, SS) \
    switch (R) { case 5: VH_PC(D, 5, FF, LL, SS); break; case 6: VH_PC(D, 6, FF, LL, SS); break; \
                 case 7: VH_PC(D, 7, FF, LL, SS); break; default: VH_PC(D, 8, FF, LL, SS); break; }
    // first == 2: forward first pass whose source rows are shared by items (ModUp lift)
    if (dir == 0 && col && first == 2 && !last) { VH_P_R(0, true, false, true) }
    else if (dir == 0 && col && first && !last) { VH_P_R(0, true, false, false) }
    else if (dir == 0 && !col && !first && last) { VH_P(0, 8, false, false, true, false); }
    else if (dir == 1 && !col && first && !last) { VH_P(1, 8, false, true, false, false); }
    else if (dir == 1 && col && !first && last) { VH_P_R(1, false, true, false) }
    else return hipErrorInvalidValue;
#undef VH_P_R
#undef VH_PC
#undef VH_P
    return hipGetLastError();
}

// Workgroups of the pipelined pass one CU holds at once (the smallest over its instances).
int vh_ntt16p_occupancy()
{
    int a = 0, b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_ntt16p<0, 8, false, false, true>, 256, 0) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_ntt16p<1, 8, false, true, false>, 256, 0) != hipSuccess) return 0;
    int m = a < b ? a : b;
    int c = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_ntt16p<0, 8, true, true, false>, 256, 0) != hipSuccess) return 0;
    m = m < c ? m : c;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_ntt16p<1, 8, true, false, true>, 256, 0) != hipSuccess) return 0;
    return m < c ? m : c;
}

hipError_t vh_launch_ntt_xcd(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, int dir,
                             int logn, unsigned *ctl, unsigned *err, unsigned lag, unsigned spin_limit, unsigned nblocks,
                             int occ, int l2x)
{
    if (logn < 13 || logn > 16) return hipErrorInvalidValue;
#define VH_XCD(D, RR, OO) \
    do { \
        if (l2x) VH_KL((k_ntt_xcd<D, RR, OO, true>), dim3(nblocks), dim3(256), 0, st, items, nitems, limbs, logn, ctl, err, lag, spin_limit); \
        else VH_KL((k_ntt_xcd<D, RR, OO, false>), dim3(nblocks), dim3(256), 0, st, items, nitems, limbs, logn, ctl, err, lag, spin_limit); \
    } while (0)
#define VH_XCD_R(D, OO) \
    switch (logn - 8) { case 5: VH_XCD(D, 5, OO); break; case 6: VH_XCD(D, 6, OO); break; \
                        case 7: VH_XCD(D, 7, OO); break; default: VH_XCD(D, 8, OO); }
    if (occ >= 4) { if (dir == 0) { VH_XCD_R(0, 4) } else { VH_XCD_R(1, 4) } }
    else          { if (dir == 0) { VH_XCD_R(0, 2) } else { VH_XCD_R(1, 2) } }
#undef VH_XCD_R
#undef VH_XCD
    return hipGetLastError();
}

__global__ void k_ntt_xcd_verify(const unsigned *__restrict__ ctl, unsigned nitems, unsigned T, unsigned *err)
{
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < nitems; i += gridDim.x * blockDim.x)
        if (ctl[16 + i] != T || ctl[16 + nitems + i] != T) atomicOr(err, 2u);
}

hipError_t vh_launch_ntt_xcd_verify(hipStream_t st, const unsigned *ctl, unsigned nitems, int logn, unsigned *err)
{
    VH_KL(k_ntt_xcd_verify, dim3((nitems + 255) / 256), dim3(256), 0, st, ctl, nitems,
                       1u << (logn - 12), err);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// On-device sampling (rns_polynomial.c:242-271 semantics; the reference draws
// from AES-CTR or BLAKE3, this engine from ChaCha20, so streams differ by
// engine exactly as the reference's two engines differ from each other).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t vh_rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

#define VH_QR(a, b, c, d)                                  \
    a += b; d ^= a; d = vh_rotl32(d, 16);                  \
    c += d; b ^= c; b = vh_rotl32(b, 12);                  \
    a += b; d ^= a; d = vh_rotl32(d, 8);                   \
    c += d; b ^= c; b = vh_rotl32(b, 7);

// ChaCha20 block: 16 output words for (key, counter, nonce).
__device__ __forceinline__ void vh_chacha20(const vh_rng_key &key, uint64_t ctr, uint32_t o[16])
{
    uint32_t x[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                      key.k[0], key.k[1], key.k[2], key.k[3], key.k[4], key.k[5], key.k[6], key.k[7],
                      (uint32_t)ctr, (uint32_t)(ctr >> 32), (uint32_t)key.nonce, (uint32_t)(key.nonce >> 32)};
#pragma unroll
    for (int i = 0; i < 16; i++) o[i] = x[i];
#pragma unroll
    for (int r = 0; r < 10; r++) {
        VH_QR(o[0], o[4], o[8], o[12]) VH_QR(o[1], o[5], o[9], o[13])
        VH_QR(o[2], o[6], o[10], o[14]) VH_QR(o[3], o[7], o[11], o[15])
        VH_QR(o[0], o[5], o[10], o[15]) VH_QR(o[1], o[6], o[11], o[12])
        VH_QR(o[2], o[7], o[8], o[13]) VH_QR(o[3], o[4], o[9], o[14])
    }
#pragma unroll
    for (int i = 0; i < 16; i++) o[i] += x[i];
}
#undef VH_QR

// Uniform residues: r * q / 2^64 of a uniform 64-bit r (the reference's
// mod-switch from 2^64, rns_polynomial.c:250).  Row k of the call uses
// counters [k * n/8, (k+1) * n/8); item.aux holds k.
__global__ void __launch_bounds__(256)
k_sample_uniform(const vh_item *__restrict__ items, const vh_limb *__restrict__ limbs, vh_rng_key key, uint64_t n)
{
    const vh_item it = items[blockIdx.y];
    const uint64_t q = limbs[it.limb].q, nb = n / 8;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t o[16];
        vh_chacha20(key, (uint64_t)it.aux * nb + b, o);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint64_t r = (uint64_t)o[2 * k] | ((uint64_t)o[2 * k + 1] << 32);
            it.dst[8 * b + k] = __umul64hi(r, q);
        }
    }
}

// Rounded Box-Muller normal samples (misc_tp.c:26-31 semantics), one noise
// vector per job written to every row of the job.
__global__ void __launch_bounds__(256)
k_sample_gaussian(const vh_gauss_job *__restrict__ jobs, const vh_item *__restrict__ items,
                  const vh_limb *__restrict__ limbs, vh_rng_key key, double sigma, uint64_t n)
{
    const vh_gauss_job jb = jobs[blockIdx.y];
    const uint64_t nb = n / 4;
    const double two_pi = 6.283185307179586476925286766559;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t o[16];
        vh_chacha20(key, (uint64_t)jb.stream * nb + b, o);
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const uint64_t r0 = (uint64_t)o[4 * s] | ((uint64_t)o[4 * s + 1] << 32);
            const uint64_t r1 = (uint64_t)o[4 * s + 2] | ((uint64_t)o[4 * s + 3] << 32);
            const double u0 = ((double)(r0 >> 11) + 0.5) * 0x1p-53, u1 = ((double)(r1 >> 11) + 0.5) * 0x1p-53;
            const double z = cos(two_pi * u0) * sqrt(-2.0 * log(u1)) * sigma;
            const int64_t v = jb.mode ? (int64_t)z : (int64_t)round(z);
            const uint64_t j = 4 * b + s;
            if (jb.mode) {
                items[jb.first].dst[j] = (uint64_t)v;
                continue;
            }
            const uint64_t mag = v < 0 ? (uint64_t)0 - (uint64_t)v : (uint64_t)v;
            for (uint32_t c = 0; c < jb.count; c++) {
                const vh_item &it = items[jb.first + c];
                const uint64_t q = limbs[it.limb].q;
                const uint64_t a = mag < q ? mag : mag % q;
                it.dst[j] = (v < 0 && a) ? q - a : a;
            }
        }
    }
}

// Raw keystream blocks (diagnostics: the RFC 8439 known answer).
__global__ void __launch_bounds__(256)
k_chacha_raw(vh_rng_key key, uint64_t ctr0, uint64_t nblocks, uint32_t *__restrict__ out)
{
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nblocks; b += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t o[16];
        vh_chacha20(key, ctr0 + b, o);
#pragma unroll
        for (int i = 0; i < 16; i++) out[16 * b + i] = o[i];
    }
}

hipError_t vh_launch_chacha_raw(hipStream_t st, vh_rng_key key, uint64_t ctr0, uint64_t nblocks, uint32_t *out)
{
    VH_KL(k_chacha_raw, dim3(vh_blocks(nblocks, 256, 64)), dim3(256), 0, st, key, ctr0, nblocks, out);
    return hipGetLastError();
}

hipError_t vh_launch_sample_uniform(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                                    vh_rng_key key, uint64_t n)
{
    if (n % 8) return hipErrorInvalidValue;
    VH_KL(k_sample_uniform, dim3(vh_blocks(n / 8, 256, 64), nitems), dim3(256), 0, st, items, limbs, key, n);
    return hipGetLastError();
}

hipError_t vh_launch_sample_gaussian(hipStream_t st, const vh_gauss_job *jobs, unsigned njobs, const vh_item *items,
                                     const vh_limb *limbs, vh_rng_key key, double sigma, uint64_t n)
{
    if (n % 4) return hipErrorInvalidValue;
    VH_KL(k_sample_gaussian, dim3(vh_blocks(n / 4, 256, 64), njobs), dim3(256), 0, st, jobs, items, limbs,
                       key, sigma, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Negacyclic monomial products (rns_polynomial.c:1048-1168):
//   out = X^a * in, or (X^a - 1) * in when item.aux == 1, mod X^n + 1, a < 2n.
// Output coefficient i takes u = (i - a) mod 2n: +in[u] if u < n, else
// -in[u - n].  One thread per output coefficient; out != in.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_mul_xai(const vh_item *__restrict__ items, const vh_limb *__restrict__ limbs, uint64_t n,
          const uint64_t *__restrict__ amounts, uint64_t stride, uint64_t col)
{
    const vh_item it = items[blockIdx.y];
    // the monomial's exponent (< 2n): the row's own, or entry (row's scalar, col) of a table
    const uint64_t a = amounts ? amounts[it.scalar * stride + col] : it.scalar;
    const vh_limb L = limbs[it.limb];
    const uint64_t q = L.q, mu = L.mu64, n2 = 2 * n;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t u = (i + n2 - a) % n2;
        const uint64_t x = vh_reduce64(it.src[u < n ? u : u - n], q, mu);
        uint64_t r = u < n ? x : (x ? q - x : 0);
        if (it.aux) r = vh_submod(r, vh_reduce64(it.src[i], q, mu), q);
        it.dst[i] = r;
    }
}

hipError_t vh_launch_mul_xai(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, uint64_t n,
                             const uint64_t *amounts, uint64_t stride, uint64_t col)
{
    VH_KL(k_mul_xai, dim3(vh_blocks(n, 256, 64), nitems), dim3(256), 0, st, items, limbs, n, amounts, stride, col);
    return hipGetLastError();
}

// Records the XCC id of every workgroup (placement probe).
__global__ void k_xcc_probe(unsigned *out)
{
    if (threadIdx.x == 0) out[blockIdx.x] = __builtin_amdgcn_s_getreg((3 << 11) | 20);
}

hipError_t vh_launch_xcc_probe(hipStream_t st, unsigned *out, unsigned nblocks)
{
    VH_KL(k_xcc_probe, dim3(nblocks), dim3(64), 0, st, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point: work lists and the persistent NTT's
// queue counters of vfhe_hip_ntt_batch_noalloc.  Host code never passes the
// array's name to a kernel or a copy: it is a host-side stand-in there.  The
// device address comes from vh_scratch_resolve, called once per device at the
// engine's set-up with that device current.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[VH_SCRATCH_BYTES];

hipError_t vh_scratch_resolve(void **ptr)
{
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

__global__ void k_zero_u32(unsigned *p, unsigned n)
{
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

hipError_t vh_launch_zero_u32(hipStream_t st, unsigned *p, unsigned n)
{
    VH_KL(k_zero_u32, dim3((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Slot operations on NTT-form rows (rns_polynomial.c:1228-1300, 1384-1432).  A
// row of n = split * ps values holds `split` blocks of ps slots.
//   op 0 broadcast: dst[i] = src[block(i) + scalar]
//   op 1 rotate:    dst[i] = src[block(i) + (slot(i) + scalar) mod ps]
//   op 2 copy:      dst[b*ps + aux] = src[b*ps + scalar] for every block b (i < split)
//   op 3 inverse:   dst[i] = src[i]^(q-2) mod q; a zero slot sets *(unsigned *)src2
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_slot(const vh_item *__restrict__ items, const vh_limb *__restrict__ limbs, int op, uint64_t n, uint64_t ps)
{
    const vh_item it = items[blockIdx.y];
    const vh_limb L = limbs[it.limb];
    const uint64_t cnt = op == 2 ? n / ps : n;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * blockDim.x) {
        if (op == 0) {
            it.dst[i] = it.src[(i / ps) * ps + it.scalar];
        } else if (op == 1) {
            const uint64_t b = (i / ps) * ps, s = i - b;
            it.dst[i] = it.src[b + (s + it.scalar) % ps];
        } else if (op == 2) {
            it.dst[i * ps + it.aux] = it.src[i * ps + it.scalar];
        } else {
            const uint64_t a = vh_reduce64(it.src[i], L.q, L.mu64);
            if (a == 0) {
                atomicOr((unsigned *)it.src2, 1u);
                it.dst[i] = 0;
                continue;
            }
            uint64_t r = 1, base = a, e = L.q - 2;
            while (e) {
                if (e & 1) r = vh_mulmod(r, base, L);
                base = vh_mulmod(base, base, L);
                e >>= 1;
            }
            it.dst[i] = r;
        }
    }
}

hipError_t vh_launch_slot(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, int op,
                          uint64_t n, uint64_t ps)
{
    const uint64_t cnt = op == 2 ? n / ps : n;
    VH_KL(k_slot, dim3(vh_blocks(cnt, 256, 64), nitems), dim3(256), 0, st, items, limbs, op, n, ps);
    return hipGetLastError();
}
