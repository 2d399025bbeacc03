// vh_kernels.hip -- HIP kernels of the hip engine (gfx950).
//
// Every kernel works on a list of vh_item rows so that one launch covers any
// number of (polynomial, limb) rows: the reference's per-limb loops
// (modules/arith/c/src/rns_polynomial.c:582-614 and friends) become the y
// dimension of a single grid.
#include <type_traits>
#include "vh_device.h"
#include "vh_kernels.h"

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

__global__ void __launch_bounds__(256)
k_division(const vh_div_job *__restrict__ jobs, const vh_limb *__restrict__ limbs,
           const vh_divc *__restrict__ consts, uint64_t n)
{
    const vh_div_job jb = jobs[blockIdx.y];
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int K = jb.K, nl = jb.nl;
    uint64_t t[VH_MAX_DIV];
    uint64_t v[VH_MAX_DIV];
    for (int k = 0; k < K; k++) v[k] = jb.base[(uint64_t)jb.div_limb[k] * jb.stride + x];
    for (int k = 0; k < K; k++) {
        const vh_limb Lk = limbs[jb.div_limb[k]];
        const vh_divc ck = consts[k * nl + jb.div_limb[k]];
        t[k] = vh_reduce64(vh_reduce64(v[k], Lk.q, Lk.mu64) + ck.h, Lk.q, Lk.mu64);
        for (int k2 = k + 1; k2 < K; k2++) {
            const vh_limb L2 = limbs[jb.div_limb[k2]];
            const vh_divc c2 = consts[k * nl + jb.div_limb[k2]];
            uint64_t a = vh_reduce64(vh_reduce64(v[k2], L2.q, L2.mu64) + c2.h, L2.q, L2.mu64);
            a = vh_submod(a, vh_reduce64(t[k], L2.q, L2.mu64), L2.q);
            v[k2] = vh_shoup(a, c2.inv, c2.inv_sh, L2.q);
        }
    }
    const int li = jb.out_limb[blockIdx.z];
    const vh_limb Li = limbs[li];
    uint64_t *row = jb.base + (uint64_t)li * jb.stride;
    uint64_t a = vh_reduce64(row[x], Li.q, Li.mu64);
    for (int k = 0; k < K; k++) {
        const vh_divc ck = consts[k * nl + li];
        a = vh_addmod(a, ck.h, Li.q);
        a = vh_submod(a, vh_reduce64(t[k], Li.q, Li.mu64), Li.q);
        a = vh_shoup(a, ck.inv, ck.inv_sh, Li.q);
    }
    row[x] = a;
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
    const vh_bconv_job jb = jobs[blockIdx.y];
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int t = blockIdx.z;
    const int P = jb.P[t];
    const vh_limb Lp = limbs[P];
    uint64_t acc = 0;
    for (int j = 0; j < jb.w; j++) {
        const vh_limb Lj = limbs[jb.D[j]];
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
__global__ void __launch_bounds__(256)
k_keyswitch_mac(const vh_mac_job *__restrict__ jobs, const vh_limb *__restrict__ limbs,
                const uint64_t *const *__restrict__ kptr, uint64_t n)
{
    const vh_mac_job jb = jobs[blockIdx.y];
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int ti = blockIdx.z;                 // index into the extended-limb list
    const int limb = jb.limb[ti];
    const vh_limb L = limbs[limb];
    const int C = jb.comps, D = jb.digits;
    uint64_t acc[VH_MAX_COMPS];
    for (int c = 0; c < C; c++) acc[c] = 0;
    for (int d = 0; d < D; d++) {
        const uint64_t m = jb.modup[((uint64_t)d * jb.next + ti) * n + x];
        const uint64_t *const *kp = kptr + jb.kptr_off + ((uint64_t)ti * D + d) * C;
        for (int c = 0; c < C; c++)
            acc[c] = vh_addmod(acc[c], vh_mulmod(vh_reduce64(kp[c][x], L.q, L.mu64), m, L), L.q);
    }
    for (int c = 0; c < C; c++) {
        uint64_t a = acc[c];
        jb.out[c][(uint64_t)limb * n + x] = a ? L.q - a : 0;
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
        if (dir == 0) hipLaunchKernelGGL((k_ntt_pass<0, true>), grid, dim3(thr), lds, st, items, limbs, logn, s0, R, col, G, first, last);
        else hipLaunchKernelGGL((k_ntt_pass<1, true>), grid, dim3(thr), lds, st, items, limbs, logn, s0, R, col, G, first, last);
    } else {
        if (dir == 0) hipLaunchKernelGGL((k_ntt_pass<0, false>), grid, dim3(thr), lds, st, items, limbs, logn, s0, R, col, G, first, last);
        else hipLaunchKernelGGL((k_ntt_pass<1, false>), grid, dim3(thr), lds, st, items, limbs, logn, s0, R, col, G, first, last);
    }
    return hipGetLastError();
}

hipError_t vh_launch_eltwise(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                             int op, uint64_t n)
{
    hipLaunchKernelGGL(k_eltwise, dim3(vh_blocks(n, 256, 64), nitems), dim3(256), 0, st, items, limbs, op, n);
    return hipGetLastError();
}

hipError_t vh_launch_split_mul(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                               int mode, uint64_t n, uint64_t S)
{
    hipLaunchKernelGGL(k_split_mul, dim3(vh_blocks(n * S, 256, 64), nitems), dim3(256), 0, st, items, limbs, mode, n);
    return hipGetLastError();
}

hipError_t vh_launch_division(hipStream_t st, const vh_div_job *jobs, unsigned njobs, unsigned nout,
                              const vh_limb *limbs, const void *consts, uint64_t n)
{
    hipLaunchKernelGGL(k_division, dim3((unsigned)((n + 255) / 256), njobs, nout), dim3(256), 0, st, jobs, limbs,
                       (const vh_divc *)consts, n);
    return hipGetLastError();
}

hipError_t vh_launch_base_conversion(hipStream_t st, const vh_bconv_job *jobs, unsigned njobs, unsigned nout,
                                     const vh_limb *limbs, const void *consts, uint64_t n)
{
    hipLaunchKernelGGL(k_base_conversion, dim3((unsigned)((n + 255) / 256), njobs, nout), dim3(256), 0, st, jobs,
                       limbs, (const ulonglong2 *)consts, n);
    return hipGetLastError();
}

hipError_t vh_launch_permute(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, uint64_t N)
{
    hipLaunchKernelGGL(k_permute, dim3(vh_blocks(N, 256, 64), nitems), dim3(256), 0, st, items, limbs, N);
    return hipGetLastError();
}

hipError_t vh_launch_keyswitch_mac(hipStream_t st, const vh_mac_job *jobs, unsigned njobs, unsigned next,
                                   const vh_limb *limbs, const uint64_t *const *kptr, uint64_t n)
{
    hipLaunchKernelGGL(k_keyswitch_mac, dim3((unsigned)((n + 255) / 256), njobs, next), dim3(256), 0, st, jobs,
                       limbs, kptr, n);
    return hipGetLastError();
}

hipError_t vh_launch_scatter_split(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                                   uint64_t N, int signed_in)
{
    hipLaunchKernelGGL(k_scatter_split, dim3(vh_blocks(N, 256, 64), nitems), dim3(256), 0, st, items, limbs, N,
                       signed_in);
    return hipGetLastError();
}
