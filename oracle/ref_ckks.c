/* Test and baseline infrastructure (never linked into the engine): the CKKS
 * multiply + relinearize + rescale of a ciphertext pair composed from the
 * reference engine's own kernels (ntt_forward / ntt_reverse / mod_eltwise_*,
 * modules/arith/c/src, compiled unmodified into the same library by
 * oracle/Makefile), in the call order of the reference's C and Python:
 *
 *   mlwe_multiply (modules/mlwe/c/src/mlwe.c:711-760)
 *     mlwe_tensor_product (mlwe.c:665-709)            rank-1 slots, NTT form
 *     mlwe_RNS_to_RNSc (mlwe.c:522-529)                 INTT of the slots
 *     mlwe_RNSc_GHS_hybrid_keyswitch (mlwe.c:557-598)
 *       gadget_mul_subto_polynomial (gadget.c:22-37)    per digit: lift (rns_polynomial.c:824-843),
 *                                                       NTT over the extended limbs, mul_subto
 *       mlwe_RNS_to_RNSc + mlwe_round_division_RNSc     ModDown by the special primes
 *       pass-through adds                                (rns_polynomial.c:939-976 per prime)
 *     mlwe_RNSc_to_RNS (mlwe.c:756)                     the product's NTT form
 *   CKKS_Scheme.rescale (modules/fhe/python/src/vfhe/fhe/ckks.py:128-148)
 *     MLWE.round_division: INTT, round division by the level's top prime
 *
 * The relinearization key holds L digits (one per level limb) of two
 * components on the L + K extended limbs, NTT form; the second slot of the rlk
 * (the linear component) is the reference's NULL pass-through.  Pairs are
 * spread over `threads` POSIX threads (the reference kernels are reentrant for
 * disjoint buffers).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "arith.h"

typedef struct {
    uint64_t N, L, K, count, ninputs;
    NTT_proc *procs;
    const uint64_t *primes, *ct1, *ct2, *rlk;
    uint64_t *out;
    uint64_t next;
    pthread_mutex_t mu;
} job_t;

static uint64_t *abuf(size_t words)
{
    void *p = NULL;
    if (posix_memalign(&p, 64, words * 8 ? words * 8 : 64)) abort();
    return (uint64_t *)p;
}

/* polynomial_round_division_RNSc_wo_free (rns_polynomial.c:939-976) on limb rows
 * x[i] (i < l, those with active[i]), dividing by limb idx. */
static void round_divide(uint64_t **x, int *active, uint64_t l, uint64_t idx, const job_t *J, uint64_t *tmp)
{
    const uint64_t N = J->N, p = J->primes[idx], half = p / 2;
    mod_eltwise_add_scalar(x[idx], x[idx], half, N, J->procs[idx]);
    for (uint64_t i = 0; i < l; i++) {
        if (!active[i] || i == idx) continue;
        const uint64_t q = J->primes[i];
        const uint64_t inv = inverse_mod(p, q);
        mod_eltwise_reduce(tmp, x[idx], N, J->procs[i]);
        mod_eltwise_add_scalar(x[i], x[i], half % q, N, J->procs[i]);
        mod_eltwise_sub(x[i], x[i], tmp, N, J->procs[i]);
        mod_eltwise_scale(x[i], x[i], inv, N, J->procs[i]);
    }
    memset(x[idx], 0, N * 8);
    active[idx] = 0;
}

static void multiply_one(const job_t *J, uint64_t k, uint64_t *ws)
{
    const uint64_t N = J->N, L = J->L, E = J->L + J->K, in = k % J->ninputs;
    const uint64_t *a1 = J->ct1 + in * 2 * L * N, *b1 = a1 + L * N;
    const uint64_t *a2 = J->ct2 + in * 2 * L * N, *b2 = a2 + L * N;
    uint64_t *d0 = ws, *d1 = d0 + L * N, *d2 = d1 + L * N;      /* tensor slots, L limbs */
    uint64_t *oa = d2 + L * N, *ob = oa + E * N;                 /* key-switch output, E limbs */
    uint64_t *t = ob + E * N;                                    /* lifted digit, E limbs */
    uint64_t *tmp = t + E * N;                                   /* one row */
    /* tensor product (NTT form) */
    for (uint64_t i = 0; i < L; i++) {
        const uint64_t o = i * N;
        mod_eltwise_mul(d0 + o, (uint64_t *)a1 + o, (uint64_t *)a2 + o, N, J->procs[i]);
        mod_eltwise_mul(d1 + o, (uint64_t *)a1 + o, (uint64_t *)b2 + o, N, J->procs[i]);
        mod_eltwise_mul_addto(d1 + o, (uint64_t *)b1 + o, (uint64_t *)a2 + o, N, J->procs[i]);
        mod_eltwise_mul(d2 + o, (uint64_t *)b1 + o, (uint64_t *)b2 + o, N, J->procs[i]);
    }
    /* to coefficient form */
    for (uint64_t i = 0; i < L; i++) {
        ntt_reverse(d0 + i * N, d0 + i * N, J->procs[i]);
        ntt_reverse(d1 + i * N, d1 + i * N, J->procs[i]);
        ntt_reverse(d2 + i * N, d2 + i * N, J->procs[i]);
    }
    /* GHS key-switch of the quadratic slot: out = -sum_j NTT(lift_j(d0)) * rlk[j] */
    memset(oa, 0, E * N * 8);
    memset(ob, 0, E * N * 8);
    for (uint64_t j = 0; j < L; j++) {
        for (uint64_t i = 0; i < E; i++) {
            if (i == j) memcpy(t + i * N, d0 + j * N, N * 8);
            else mod_eltwise_reduce(t + i * N, d0 + j * N, N, J->procs[i]);
            ntt_forward(t + i * N, t + i * N, J->procs[i]);
        }
        const uint64_t *ka = J->rlk + j * 2 * E * N, *kb = ka + E * N;
        for (uint64_t i = 0; i < E; i++) {
            mod_eltwise_mul_subto(oa + i * N, (uint64_t *)ka + i * N, t + i * N, N, J->procs[i]);
            mod_eltwise_mul_subto(ob + i * N, (uint64_t *)kb + i * N, t + i * N, N, J->procs[i]);
        }
    }
    for (uint64_t i = 0; i < E; i++) {
        ntt_reverse(oa + i * N, oa + i * N, J->procs[i]);
        ntt_reverse(ob + i * N, ob + i * N, J->procs[i]);
    }
    /* ModDown: round division by the special primes, ascending, on each component */
    uint64_t *ra[64], *rb[64];
    int act_a[64], act_b[64];
    for (uint64_t i = 0; i < E; i++) { ra[i] = oa + i * N; rb[i] = ob + i * N; act_a[i] = act_b[i] = 1; }
    for (uint64_t s = L; s < E; s++) round_divide(ra, act_a, E, s, J, tmp);
    for (uint64_t s = L; s < E; s++) round_divide(rb, act_b, E, s, J, tmp);
    /* pass-through: the linear slot and the constant slot keep the key */
    for (uint64_t i = 0; i < L; i++) {
        mod_eltwise_add(oa + i * N, oa + i * N, d1 + i * N, N, J->procs[i]);
        mod_eltwise_add(ob + i * N, ob + i * N, d2 + i * N, N, J->procs[i]);
    }
    /* mlwe_multiply's NTT form, then the rescale's INTT and division by the top prime */
    for (uint64_t i = 0; i < L; i++) {
        ntt_forward(oa + i * N, oa + i * N, J->procs[i]);
        ntt_forward(ob + i * N, ob + i * N, J->procs[i]);
    }
    for (uint64_t i = 0; i < L; i++) {
        ntt_reverse(oa + i * N, oa + i * N, J->procs[i]);
        ntt_reverse(ob + i * N, ob + i * N, J->procs[i]);
    }
    round_divide(ra, act_a, L, L - 1, J, tmp);
    round_divide(rb, act_b, L, L - 1, J, tmp);
    if (k >= J->ninputs) return;   /* a repeated input: the result is not kept */
    uint64_t *o = J->out + k * 2 * (L - 1) * N;
    memcpy(o, oa, (L - 1) * N * 8);
    memcpy(o + (L - 1) * N, ob, (L - 1) * N * 8);
}

static void *worker(void *arg)
{
    job_t *J = (job_t *)arg;
    const uint64_t E = J->L + J->K;
    uint64_t *ws = abuf((3 * J->L + 3 * E + 1) * J->N);
    for (;;) {
        pthread_mutex_lock(&J->mu);
        const uint64_t k = J->next++;
        pthread_mutex_unlock(&J->mu);
        if (k >= J->count) break;
        multiply_one(J, k, ws);
    }
    free(ws);
    return NULL;
}

/* ct1, ct2: ninputs x 2 x L x N (a then b, NTT form); rlk: L x 2 x (L + K) x N (NTT form);
 * out: ninputs x 2 x (L - 1) x N (coefficient form).  Multiplies `count` pairs, pair k
 * being input k mod ninputs (a timing sample larger than its inputs); the results of
 * the first ninputs pairs are kept.  Returns 0, or -1 on bad sizes. */
int ref_ckks_multiply(uint64_t N, uint64_t L, uint64_t K, const uint64_t *primes, const uint64_t *ct1,
                      const uint64_t *ct2, const uint64_t *rlk, uint64_t *out, uint64_t ninputs, uint64_t count,
                      int threads)
{
    if (L < 2 || L + K > 64 || N == 0 || threads < 1 || ninputs == 0) return -1;
    job_t J = {N, L, K, count, ninputs, NULL, primes, ct1, ct2, rlk, out, 0, PTHREAD_MUTEX_INITIALIZER};
    J.procs = (NTT_proc *)malloc(sizeof(NTT_proc) * (L + K));
    for (uint64_t i = 0; i < L + K; i++) J.procs[i] = ntt_new_proc(N, primes[i]);
    pthread_t th[512];
    const int nt = threads > 512 ? 512 : threads;
    for (int i = 0; i < nt; i++) pthread_create(&th[i], NULL, worker, &J);
    for (int i = 0; i < nt; i++) pthread_join(th[i], NULL);
    for (uint64_t i = 0; i < L + K; i++) ntt_free_proc(J.procs[i]);
    free(J.procs);
    return 0;
}
