// vh_kernels.h -- launch interface between the host runtime and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "vh_device.h"

#define VH_MAX_DIV 16     // divide primes handled by one fused division launch
#define VH_MAX_LIMBS 64   // masks are uint64_t, as in the reference
#define VH_MAX_COMPS 8    // MLWE components (rank + 1) of one key-switch output

// While set, the launch functions below evaluate their arguments and launch nothing:
// the engine rebuilds a captured graph's work lists this way before replaying it.
extern int g_vh_suppress_launch;
#define VH_KL(...) do { if (!g_vh_suppress_launch) hipLaunchKernelGGL(__VA_ARGS__); } while (0)

// Per-device scratch of the allocation-free entry point (vfhe_hip_ntt_batch_noalloc):
// a file-scope device array in vh_kernels.hip, VH_SCRATCH_BYTES on every device the
// library runs on, for the life of the process.
#define VH_SCRATCH_BYTES ((size_t)1 << 20)
// Resolves the calling thread's current device's copy of the scratch array to a
// device address (hipGetSymbolAddress); call once per device with it current.
hipError_t vh_scratch_resolve(void **ptr);

// Kernel families reported by the engine's launch profiler (vfhe_hip_profile_stop).
enum {
    VH_PROF_NTT_FP64 = 0, VH_PROF_NTT_INT, VH_PROF_ELTWISE, VH_PROF_DIVISION, VH_PROF_KSMAC, VH_PROF_OTHER,
    VH_PROF_TAGS
};

enum {
    VH_OP_MUL = 0, VH_OP_MUL_ADDTO, VH_OP_MUL_SUBTO, VH_OP_SCALE, VH_OP_FMA, VH_OP_ADD, VH_OP_SUB,
    VH_OP_NEGATE, VH_OP_REDUCE, VH_OP_REDUCE_SIGNED, VH_OP_ADD_SCALAR, VH_OP_SUB_SCALAR,
    VH_OP_COPY, VH_OP_ZERO, VH_OP_REDUCE_MP, VH_OP_ADD_AT0
};

// One limb row of a rank-1 tensor product: d0 = a1 a2, d1 = a1 b2 + b1 a2, d2 = b1 b2.
struct vh_tensor_job {
    const uint64_t *a1, *b1, *a2, *b2;
    uint64_t *d0, *d1, *d2;
    uint32_t limb;
};

// One polynomial's sequential division job.
struct vh_div_job {
    uint64_t *base;      // device rows of the polynomial, limb-major
    uint64_t stride;     // row length (N)
    int K;               // number of divide primes
    int nl;              // pool size (row count of the constant table)
    int nout;            // number of output limbs
    int pad;
    int div_limb[VH_MAX_DIV];
    int out_limb[VH_MAX_LIMBS];
};

// ModDown by K special primes, pass-through add and rescale by one prime of a
// key-switch output component (k_moddown_rescale).
struct vh_mdr_job {
    uint64_t *base;          // device rows of the component, limb-major
    const uint64_t *add;     // rows of the pass-through polynomial (same layout), or null
    uint64_t stride;         // row length (N)
    int K;                   // special primes
    int nl;                  // pool size (row count of the constant tables)
    int nout;                // output limbs (the rescale limb excluded)
    int resc;                // the rescale limb
    int div_limb[VH_MAX_DIV];
    int out_limb[VH_MAX_LIMBS];
};

struct vh_bconv_job {
    const uint64_t *in;
    uint64_t *out;
    uint64_t stride;
    int w, v;
    int D[VH_MAX_LIMBS];
    int P[VH_MAX_LIMBS];
};

struct vh_mac_job {
    uint64_t *out[VH_MAX_COMPS];  // device base of each output component
    const uint64_t *modup;        // [digits][next][n] lifted, forward-transformed digits
    uint64_t kptr_off;            // offset of this job's key-row table in kptr
    int comps, digits, next;
    int add;                      // 1: out = +sum (external product), 0: out = -sum (key-switch)
    int limb[VH_MAX_LIMBS];       // extended limb list
    // Fused rows only: out_c += P * pass[c] on the limbs of pass_mask (pass[c] NTT-form rows,
    // limb-major; pmod[limb] = {P mod q centred, RN(that / q)}), null pass[c]: nothing added.
    const uint64_t *pass[VH_MAX_COMPS];
    const double2 *pmod;
    uint64_t pass_mask;
};

// Whole FP64 transforms of N = 2^9 .. 2^12 (reads it.src, writes it.dst), registers radix-16.
hipError_t vh_launch_ntt_small(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, int dir,
                               int logn);
// Fused CGGI blind rotation (N = 1024, FP64 limbs, one special prime): workgroup k steps
// accumulator k through every key index with its rows resident in LDS; wave w works modulo
// extended limb E[w].
#define VH_BR_MAXE 4
#define VH_BR_MAXJ 3
struct vh_br_params {
    int C;                      // components (rank + 1), 1 or 2
    int nJ;                     // level limbs (digits per component)
    int next;                   // extended limbs = waves per workgroup (nJ + 1)
    int sp;                     // wave of the special limb
    int E[VH_BR_MAXE];          // extended limbs, ascending
    int J[VH_BR_MAXJ];          // level limbs, ascending
    int jw[VH_BR_MAXE];         // wave -> index in J (-1: the special limb)
    double hdiv[VH_BR_MAXE];    // floor(p/2) mod q_E[w]
    double inv[VH_BR_MAXE];     // p^-1 mod q_E[w], centred
    double invq[VH_BR_MAXE];    // RN(inv / q_E[w])
    double hp;                  // floor(p/2)
    uint64_t n;                 // key indices
};
// accp[(k * C + c) * nJ + jj]: row J[jj] of component c of accumulator k (coefficient form,
// updated in place); bkp[((i * C * nJ + d) * C + c) * next + w]: key index i, digit d, output
// component c, limb E[w] (NTT form); amounts[k * n + i] < 2N.  tvp (optional): the C nJ rows of a
// test vector; accumulator k then starts as X^rot[k] * tv instead of from its own rows.
hipError_t vh_launch_blind_rotate(hipStream_t st, const vh_br_params &P, const vh_limb *limbs, uint64_t *const *accp,
                                  const uint64_t *const *bkp, const uint64_t *amounts, unsigned count,
                                  const uint64_t *const *tvp = nullptr, const uint64_t *rot = nullptr);
// CGGI over the 32-bit torus (vh_torus.hip).  consts: 576 double2 -- [0, 512) pass 1's
// twiddles with the lane's twist factor, zeta^r w^(r k) at k 64 + r (r < 64, k < 8),
// [512, 576) the pass-2 twiddles w^(8 a s) at a 8 + s (a, s < 8); w = e^(-2 pi i / 512),
// zeta = e^(i pi / 1024).  The slot factors zeta^(64 s) are constants in the kernels.
struct vt_br_args {
    const uint32_t *lwe;      // count x (n + 1): a_0 .. a_(n-1), b
    const uint32_t *tv;       // N test-vector coefficients
    const double2 *bsk;       // n x (2 ell) x 2 x [slot][lane], transform domain
    const double2 *consts;
    uint32_t *out;            // count x (N + 1) extracted samples, or null
    uint32_t *acc_out;        // count x 2N accumulators (a then b), or null
    uint32_t n, ell, base_log;
    uint32_t offset;          // floor(2^32 / (4 torus_base)), added to b before its rounding
    uint32_t dec_off;         // rounding and digit offsets of the signed decomposition
};
hipError_t vh_launch_torus_fft(hipStream_t st, const uint32_t *in, double2 *out, const double2 *consts, uint64_t count,
                               int natural);
hipError_t vh_launch_torus_blind_rotate(hipStream_t st, const vt_br_args &A, uint64_t count);
// Batched sample extraction (mlwe_extract_LWE, mlwe.c:40-66, at coefficient idx): src[(k na + j)
// (r + 1) + c] = row of limb j of component c of sample k (coefficient form), qs[j] its prime;
// out[(k na + j) (r N + 1) + e] = the extracted sample's mask words (e < r N) and body (e = r N).
hipError_t vh_launch_extract_lwe(hipStream_t st, const uint64_t *const *src, const uint64_t *qs, uint64_t *out,
                                 uint32_t N, uint32_t r, uint32_t na, uint32_t idx, uint64_t count);
// ALU rate probe (see k_alu_rate): blocks x 256 threads, 8 chains x iters operations each.
hipError_t vh_launch_alu_rate(hipStream_t st, int kind, int iters, unsigned blocks, uint64_t *out);
hipError_t vh_launch_ntt_pass(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                              int fp, int dir, int logn, int s0, int R, int col, int G, int first, int last);
// Register radix-16 pass (FP64 path): 4 <= R <= 8, G = 4096 / 2^R sub-transforms per block.
// first: 1 = the pass reads it.src; 2 = also, and many items share one source row
// (read through the caches instead of streaming).
hipError_t vh_launch_ntt16(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                           int dir, int logn, int s0, int R, int col, int first, int last);
// Pipelined pass (FP64, N = 2^13..2^16): same arguments as vh_launch_ntt16 (first = 2 for the
// forward column pass only), nblocks persistent workgroups of 256 threads (at most one per tile);
// needs vh_limb::rowf_* for row passes.  wide: a column pass runs 512-thread workgroups over
// twice the columns per tile (nblocks is halved).
hipError_t vh_launch_ntt16p(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, int dir,
                            int logn, int s0, int R, int col, int first, int last, unsigned nblocks, int wide = 0);
int vh_ntt16p_occupancy();
// Persistent two-pass NTT with per-XCD queues; ctl holds 16 + nitems zeroed uint32 counters.
hipError_t vh_launch_ntt_xcd(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, int dir,
                             int logn, unsigned *ctl, unsigned *err, unsigned lag, unsigned spin_limit, unsigned nblocks,
                             int occ, int l2x);
hipError_t vh_launch_ntt_xcd_verify(hipStream_t st, const unsigned *ctl, unsigned nitems, int logn, unsigned *err);
hipError_t vh_launch_xcc_probe(hipStream_t st, unsigned *out, unsigned nblocks);
hipError_t vh_launch_zero_u32(hipStream_t st, unsigned *p, unsigned n);
hipError_t vh_launch_tensor1(hipStream_t st, const vh_tensor_job *jobs, unsigned njobs, const vh_limb *limbs,
                             uint64_t n);
hipError_t vh_launch_eltwise(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                             int op, uint64_t n);
hipError_t vh_launch_split_mul(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                               int mode, uint64_t n, uint64_t S);
// cdown: K x nl division constants of the special primes; cresc: 1 x nl of the rescale prime.
hipError_t vh_launch_moddown_rescale(hipStream_t st, const vh_mdr_job *jobs, unsigned njobs, const vh_limb *limbs,
                                     const void *cdown, const void *cresc, uint64_t n, int kmax);
// FP64 form (every limb below 2^50, K <= 4): cs holds the constant rows built by the
// engine (4 special rows, then one row per pool limb, VH_MDR_W double2 each).
#define VH_MDR_W 8
hipError_t vh_launch_moddown_rescale_fp(hipStream_t st, const vh_mdr_job *jobs, unsigned njobs, const vh_limb *limbs,
                                        const void *cs, uint64_t n);
hipError_t vh_launch_division(hipStream_t st, const vh_div_job *jobs, unsigned njobs, unsigned nout,
                              const vh_limb *limbs, const void *consts, uint64_t n, int kmax);
hipError_t vh_launch_base_conversion(hipStream_t st, const vh_bconv_job *jobs, unsigned njobs, unsigned nout,
                                     const vh_limb *limbs, const void *consts, uint64_t n);
hipError_t vh_launch_permute(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, uint64_t N);
// Forward row pass (last 8 stages) of every digit fused with the inner product
// out_c = -sum_d digit_d * key[d][c]; modup holds the column-pass output.
// FP64 path, 13 <= logn <= 16, comps 1 or 2.
hipError_t vh_launch_ntt16_row_mac(hipStream_t st, const vh_mac_job *jobs, unsigned njobs, unsigned next,
                                   const vh_limb *limbs, const uint64_t *const *kptr, int logn, int comps);
// Forward row stages of it.src and it.src2 (column-pass outputs), their pointwise
// product and the inverse row stages, into it.dst (the inverse column pass input).
// Pipelined form of the same step (persistent workgroups, cus x its occupancy, or one tile per
// workgroup with one_tile); needs vh_limb::rowf_*.
hipError_t vh_launch_ntt16p_row_mul(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                                    int logn, unsigned cus, int one_tile);
hipError_t vh_launch_ntt16_row_mul_inv(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                                       int logn);
// fp != 0: every extended limb has q < 2^50 and every operand is canonical (FP64 products).
hipError_t vh_launch_keyswitch_mac(hipStream_t st, const vh_mac_job *jobs, unsigned njobs, unsigned next,
                                   const vh_limb *limbs, const uint64_t *const *kptr, uint64_t n, int fp);
// On-device sampling: ChaCha20 (RFC 8439 block function, 64-bit counter and
// 64-bit nonce words) keyed by the engine's 256-bit key; one nonce per call.
struct vh_rng_key {
    uint32_t k[8];
    uint64_t nonce;
};
// Gaussian job: items[first .. first+count) are the rows of one polynomial that
// receive the same noise vector; stream numbers the job inside the call.
// mode 0: round() and reduce into each row's limb; mode 1: truncate, raw int64
// into the single row (secret-key sampling).
struct vh_gauss_job {
    uint32_t first, count, stream, mode;
};
hipError_t vh_launch_sample_uniform(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                                    vh_rng_key key, uint64_t n);
hipError_t vh_launch_sample_gaussian(hipStream_t st, const vh_gauss_job *jobs, unsigned njobs, const vh_item *items,
                                     const vh_limb *limbs, vh_rng_key key, double sigma, uint64_t n);
hipError_t vh_launch_chacha_raw(hipStream_t st, vh_rng_key key, uint64_t ctr0, uint64_t nblocks, uint32_t *out);
// Complex FP64 FFT (CKKS encoder): rows of 2n doubles (real | imaginary).
// dir 0 = CT_NR, 1 = GS_RN; fma selects the reference engine whose complex
// product is reproduced (1 avx512ifma, 0 portable).
hipError_t vh_launch_fft(hipStream_t st, double *const *rows, unsigned nrows, const double2 *ws, int logn, int dir,
                         int fma);
hipError_t vh_launch_fft_bitrev(hipStream_t st, double *const *rows, double *const *tmp, unsigned nrows, uint64_t n,
                                uint32_t prec);
hipError_t vh_launch_fft_scale(hipStream_t st, double *const *rows, unsigned nrows, double scale, uint64_t count);
hipError_t vh_launch_fft_round(hipStream_t st, double *const *rows, uint64_t *const *outs, unsigned nrows, uint64_t n);
// out = X^a * in (item.aux = 0) or (X^a - 1) * in (aux = 1) mod X^n + 1, a < 2n: a = item.scalar,
// or with a table, a = amounts[item.scalar * stride + col].
hipError_t vh_launch_mul_xai(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, uint64_t n,
                             const uint64_t *amounts = nullptr, uint64_t stride = 0, uint64_t col = 0);
hipError_t vh_launch_scatter_split(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs,
                                   uint64_t N, int signed_in);

// Slot operations on NTT-form rows of n = split * ps values (op 0 broadcast, 1 rotate,
// 2 copy one slot, 3 per-slot inverse with a zero flag at src2).
hipError_t vh_launch_slot(hipStream_t st, const vh_item *items, unsigned nitems, const vh_limb *limbs, int op,
                          uint64_t n, uint64_t ps);

