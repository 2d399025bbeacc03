/*
 * vfhe_oracle.c -- CPU restatement of vfhe's RNS polynomial hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
 * library built from it.  Nothing in the shipped engine links or calls it.
 *
 * Every function restates the reference's algorithm in plain scalar C with
 * unsigned __int128 arithmetic (exact `%`), citing the reference file:line it
 * follows.  Polynomials are flat limb-major arrays p[l][N] (limb i at p + i*N)
 * with an explicit RNS mask, exactly the reference's `coeffs[i]` rows
 * (modules/arith/c/include/arith.h:72-78).
 *
 * How each part is pinned (tests/test_oracle.py):
 *  - number theory: the known answers of modules/arith/c/test/test_number_theory.c
 *    and the psi values of the reference's ntt_new_proc at n = 4096; Python big-int primality cross-check;
 *  - NTT forward: brute-force closed form NTT(x)[i] = sum_j x_j psi^{j(2 brv(i)+1)}
 *    and bit-exact agreement with the reference's own ntt_forward (oracle/_ref,
 *    both the portable and the avx512ifma builds) on every size/prime tested;
 *  - NTT inverse: roundtrip + agreement with the reference ntt_reverse;
 *  - elementwise: Python big-int `%` and the reference mod_eltwise_* (oracle/_ref);
 *  - lift / round / floor division / base conversion / scaled lift: big-integer
 *    CRT reconstruction (closed forms stated in tests/test_oracle.py);
 *  - permute: composition gen * gen^-1 = identity and big-int X -> X^gen;
 *  - tensor product: the reference's slot-layout identity (test_mlwe.py:183-207);
 *  - NTT and mod_eltwise_{mul,add,sub}: also the committed golden vectors
 *    tests/golden/ref_ntt_mod.npz, written from the reference kernels;
 *  - key-switch / multiply / automorphism: decryption with valid keys built by
 *    the mlwe.py:156-176 key layout (phase preserved to |e| < 2^16 out of
 *    Q ~ 2^150; reversed digits or a wrong source key fail), on top of the
 *    pinned lift, NTT, product and division steps they are composed of.  No
 *    reference-produced vector exists for the bit pattern of a key-switched
 *    ciphertext: the reference's C for it (rns_polynomial.c, mlwe.c) includes
 *    blake3.h, which the image lacks, so it is not buildable here.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <assert.h>

typedef unsigned __int128 u128;

#define OAPI __attribute__((visibility("default")))

static inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }
static inline uint64_t addmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a + b) % q); }
static inline uint64_t submod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a + q - (b % q)) % q); }

/* ------------------------------------------------------------------------- */
/* Number theory: modules/arith/c/src/number_theory.c:6-217,                  */
/* modules/arith/python/src/vfhe/arith/number_theory.py:3-46,                 */
/* modules/arith/python/src/vfhe/arith/polynomial.py:151-170                  */
/* ------------------------------------------------------------------------- */

OAPI uint64_t o_power_mod(uint64_t base, uint64_t exp, uint64_t mod)
{ /* number_theory.c:6-20 (square and multiply) */
    uint64_t r = 1 % mod;
    base %= mod;
    while (exp) {
        if (exp & 1) r = mulmod(r, base, mod);
        base = mulmod(base, base, mod);
        exp >>= 1;
    }
    return r;
}

OAPI uint64_t o_inverse_mod(uint64_t a, uint64_t m)
{ /* number_theory.c:103-109 semantics: reduce, 0 maps to 0, else the inverse (here by signed EEA) */
    a %= m;
    if (a == 0) return 0;
    __int128 t = 0, nt = 1, r = m, nr = a;
    while (nr != 0) {
        __int128 qq = r / nr, tmp;
        tmp = t - qq * nt; t = nt; nt = tmp;
        tmp = r - qq * nr; r = nr; nr = tmp;
    }
    if (t < 0) t += m;
    return (uint64_t)t;
}

OAPI int o_is_prime(uint64_t n)
{ /* Miller-Rabin with the first 12 prime bases: deterministic below 2^64
     (number_theory.c:111-152 and number_theory.py:3-46 agree on every n they
     are asked about here). */
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; i++) {
        if (n == small[i]) return 1;
        if (n % small[i] == 0) return 0;
    }
    uint64_t d = n - 1; int s = 0;
    while ((d & 1) == 0) { d >>= 1; s++; }
    for (int i = 0; i < 12; i++) {
        uint64_t x = o_power_mod(small[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int r = 1; r < s; r++) {
            x = mulmod(x, x, n);
            if (x == n - 1) { comp = 0; break; }
        }
        if (comp) return 0;
    }
    return 1;
}

OAPI uint64_t o_next_special_prime(uint64_t x, uint64_t n, int primitive)
{ /* number_theory.c:179-217: smallest prime > x with q = 1 mod 2n (and q != 1 mod 4n if primitive) */
    uint64_t ord = 2 * n, a = x / ord;
    if (a * ord + 1 <= x) a++;
    if (primitive && (a % 2 == 0)) a++;
    for (;;) {
        uint64_t c = a * ord + 1;
        if (o_is_prime(c) && (!primitive || c % (4 * n) != 1)) return c;
        a += primitive ? 2 : 1;
    }
}

OAPI uint64_t o_gen_prime(uint64_t rou_order, uint64_t prime_size, const uint64_t *excl, uint64_t nexcl)
{ /* polynomial.py:151-160 (Ring.gen_prime): a = ((2^b-1)//order)|1, then a -= 2 before the
     first candidate, descending by 2 until a prime not in the exclusion list. */
    uint64_t top = (prime_size >= 64) ? UINT64_MAX : ((1ULL << prime_size) - 1);
    uint64_t a = (top / rou_order) | 1ULL;
    a -= 2;
    for (;;) {
        uint64_t c = a * rou_order + 1;
        int ex = 0;
        for (uint64_t i = 0; i < nexcl; i++) if (excl[i] == c) { ex = 1; break; }
        if (!ex && o_is_prime(c)) return c;
        a -= 2;
    }
}

OAPI uint64_t o_find_psi(uint64_t n, uint64_t q)
{ /* ntt.c:271-281 / ntt_portable.c:141-151: first g = 2,3,... with g^((q-1)/2n) of order 2n */
    for (uint64_t g = 2; g < q; g++) {
        uint64_t c = o_power_mod(g, (q - 1) / (2 * n), q);
        if (o_power_mod(c, n, q) == q - 1) return c;
    }
    return 0;
}

static uint64_t brv(uint64_t x, int bits)
{
    uint64_t r = 0;
    for (int i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

/* ------------------------------------------------------------------------- */
/* Single-limb NTT: ntt_portable.c:18-62 (tables), :76-135 (transforms)        */
/* ------------------------------------------------------------------------- */

static void make_table(uint64_t *tab, uint64_t n, uint64_t q, uint64_t root)
{ /* ntt_portable.c:25-33: tab[brv(i)] = root^i */
    int logn = 0;
    while ((1ULL << logn) < n) logn++;
    uint64_t p = 1;
    for (uint64_t i = 0; i < n; i++) {
        tab[brv(i, logn)] = p;
        p = mulmod(p, root, q);
    }
}

static void ntt_fwd_tab(uint64_t *a, uint64_t n, uint64_t q, const uint64_t *F)
{ /* ntt_portable.c:76-100: merged Cooley-Tukey, natural in, bit-reversed out */
    uint64_t t = n;
    for (uint64_t m = 1; m < n; m <<= 1) {
        t >>= 1;
        for (uint64_t i = 0; i < m; i++) {
            uint64_t w = F[m + i], j1 = 2 * i * t;
            for (uint64_t j = j1; j < j1 + t; j++) {
                uint64_t u = a[j], v = mulmod(a[j + t], w, q);
                a[j] = addmod(u, v, q);
                a[j + t] = submod(u, v, q);
            }
        }
    }
}

static void ntt_inv_tab(uint64_t *a, uint64_t n, uint64_t q, const uint64_t *I)
{ /* ntt_portable.c:102-135: Gentleman-Sande, bit-reversed in, natural out, then n^-1 */
    uint64_t t = 1;
    for (uint64_t m = n; m > 1; m >>= 1) {
        uint64_t h = m >> 1;
        for (uint64_t i = 0; i < h; i++) {
            uint64_t w = I[h + i], j1 = 2 * i * t;
            for (uint64_t j = j1; j < j1 + t; j++) {
                uint64_t u = a[j], v = a[j + t];
                a[j] = addmod(u, v, q);
                a[j + t] = mulmod(submod(u, v, q), w, q);
            }
        }
        t <<= 1;
    }
    uint64_t ninv = o_inverse_mod(n, q);
    for (uint64_t i = 0; i < n; i++) a[i] = mulmod(a[i], ninv, q);
}

OAPI void o_ntt_forward(uint64_t *a, uint64_t n, uint64_t q)
{
    uint64_t *F = malloc(n * sizeof(uint64_t));
    make_table(F, n, q, o_find_psi(n, q));
    ntt_fwd_tab(a, n, q, F);
    free(F);
}

OAPI void o_ntt_inverse(uint64_t *a, uint64_t n, uint64_t q)
{
    uint64_t *I = malloc(n * sizeof(uint64_t));
    make_table(I, n, q, o_inverse_mod(o_find_psi(n, q), q));
    ntt_inv_tab(a, n, q, I);
    free(I);
}

/* ------------------------------------------------------------------------- */
/* Elementwise ops: mod_portable.c:55-161 (with test_mod.c:21-85's oracles)    */
/* ------------------------------------------------------------------------- */

enum {
    O_MUL = 0, O_MUL_ADDTO, O_MUL_SUBTO, O_SCALE, O_FMA, O_ADD, O_SUB, O_NEGATE,
    O_REDUCE, O_REDUCE_SIGNED, O_ADD_SCALAR, O_SUB_SCALAR
};

OAPI void o_eltwise(int op, uint64_t *out, const uint64_t *in1, const uint64_t *in2, uint64_t s,
                    uint64_t n, uint64_t q)
{
    uint64_t sr = s % q;
    for (uint64_t i = 0; i < n; i++) {
        switch (op) {
        case O_MUL: out[i] = mulmod(in1[i], in2[i], q); break;                         /* :55-59 */
        case O_MUL_ADDTO: out[i] = addmod(out[i], mulmod(in1[i], in2[i], q), q); break; /* :61-68 */
        case O_MUL_SUBTO: out[i] = submod(out[i], mulmod(in1[i], in2[i], q), q); break; /* :70-77 */
        case O_SCALE: out[i] = mulmod(in1[i], sr, q); break;                            /* :79-84 */
        case O_FMA: out[i] = addmod(out[i] % q, mulmod(in1[i], sr, q), q); break;       /* :86-94 */
        case O_ADD: out[i] = addmod(in1[i], in2[i], q); break;                          /* :96-100 */
        case O_SUB: out[i] = submod(in1[i], in2[i], q); break;                          /* :102-106 */
        case O_NEGATE: out[i] = (q - in1[i] % q) % q; break;                            /* :108-112 */
        case O_REDUCE: out[i] = in1[i] % q; break;                                      /* :114-118 */
        case O_REDUCE_SIGNED: {                                                         /* :120-137 */
            int64_t v = (int64_t)in1[i];
            uint64_t r = ((v < 0) ? -(uint64_t)v : (uint64_t)v) % q;
            out[i] = (v < 0) ? (r == 0 ? 0 : q - r) : r;
            break;
        }
        case O_ADD_SCALAR: out[i] = addmod(in1[i], sr, q); break;                       /* :139-144 */
        case O_SUB_SCALAR: out[i] = submod(in1[i], sr, q); break;                       /* :146-151 */
        default: abort();
        }
    }
}

OAPI void o_reduce_mp(uint64_t *out, const uint64_t *hi, const uint64_t *lo, uint64_t n, uint64_t q)
{ /* mod_portable.c:153-161 */
    for (uint64_t i = 0; i < n; i++) out[i] = (uint64_t)((((u128)hi[i] << 64) | lo[i]) % q);
}

/* ------------------------------------------------------------------------- */
/* RNS context: the incNTT of rns_polynomial.c:7-61 plus one NTT_proc per limb */
/* ------------------------------------------------------------------------- */

typedef struct o_ctx {
    uint64_t N, split, n, l;
    uint64_t *q, *psi;
    uint64_t **F, **I, **w;
} o_ctx;

OAPI o_ctx *o_ctx_new(uint64_t N, uint64_t split, const uint64_t *primes, uint64_t l)
{
    o_ctx *c = calloc(1, sizeof(*c));
    c->N = N; c->split = split; c->n = N / split; c->l = l;
    c->q = malloc(l * sizeof(uint64_t));
    c->psi = malloc(l * sizeof(uint64_t));
    c->F = malloc(l * sizeof(uint64_t *));
    c->I = malloc(l * sizeof(uint64_t *));
    c->w = malloc(l * sizeof(uint64_t *));
    int logn = 0;
    while ((1ULL << logn) < c->n) logn++;
    for (uint64_t i = 0; i < l; i++) {
        uint64_t q = primes[i], n = c->n;
        c->q[i] = q;
        c->psi[i] = o_find_psi(n, q);
        c->F[i] = malloc(n * sizeof(uint64_t));
        c->I[i] = malloc(n * sizeof(uint64_t));
        make_table(c->F[i], n, q, c->psi[i]);
        make_table(c->I[i], n, q, o_inverse_mod(c->psi[i], q));
        /* rns_polynomial.c:21-27: w[k] = psi^(2 brv(k) + 1), the X^split wrap per slot */
        c->w[i] = malloc(n * sizeof(uint64_t));
        for (uint64_t k = 0; k < n; k++)
            c->w[i][k] = o_power_mod(c->psi[i], 2 * brv(k, logn) + 1, q);
    }
    return c;
}

OAPI void o_ctx_free(o_ctx *c)
{
    for (uint64_t i = 0; i < c->l; i++) { free(c->F[i]); free(c->I[i]); free(c->w[i]); }
    free(c->F); free(c->I); free(c->w); free(c->q); free(c->psi); free(c);
}

OAPI uint64_t o_ctx_psi(const o_ctx *c, uint64_t i) { return c->psi[i]; }
OAPI void o_ctx_w(const o_ctx *c, uint64_t i, uint64_t *out) { memcpy(out, c->w[i], c->n * sizeof(uint64_t)); }

#define ROW(p, i) ((p) + (i) * c->N)
#define ACTIVE(mask, i) (((mask) >> (i)) & 1ULL)

/* polynomial_RNSc_to_RNS / polynomial_RNS_to_RNSc: rns_polynomial.c:582-614 */
OAPI uint64_t o_to_ntt(const o_ctx *c, uint64_t *out, const uint64_t *in, uint64_t mask)
{
    for (uint64_t i = 0; i < c->l; i++) {
        if (!ACTIVE(mask, i)) continue;
        if (out != in) memcpy(ROW(out, i), ROW(in, i), c->N * sizeof(uint64_t));
        for (uint64_t k = 0; k < c->split; k++) ntt_fwd_tab(ROW(out, i) + k * c->n, c->n, c->q[i], c->F[i]);
    }
    return mask;
}

OAPI uint64_t o_to_coeff(const o_ctx *c, uint64_t *out, const uint64_t *in, uint64_t mask)
{
    for (uint64_t i = 0; i < c->l; i++) {
        if (!ACTIVE(mask, i)) continue;
        if (out != in) memcpy(ROW(out, i), ROW(in, i), c->N * sizeof(uint64_t));
        for (uint64_t k = 0; k < c->split; k++) ntt_inv_tab(ROW(out, i) + k * c->n, c->n, c->q[i], c->I[i]);
    }
    return mask;
}

/* Pointwise product of split-degree block polynomials with the X^split wrap
   through w: polynomial_mul_RNS_polynomial / _mul_addto_ / _mul_subto_
   (rns_polynomial.c:311-435).  mode 0: out = in1*in2, 1: out += , 2: out -= . */
static void blockmul(const o_ctx *c, uint64_t i, uint64_t *o, const uint64_t *a, const uint64_t *b, int mode)
{
    uint64_t n = c->n, S = c->split, q = c->q[i];
    uint64_t *acc = calloc(c->N, sizeof(uint64_t));
    for (uint64_t j = 0; j < S; j++)
        for (uint64_t k = 0; k < S; k++) {
            uint64_t dst = (j + k) % S;
            int wrap = (j + k) >= S;
            for (uint64_t x = 0; x < n; x++) {
                uint64_t p = mulmod(a[j * n + x], b[k * n + x], q);
                if (wrap) p = mulmod(p, c->w[i][x], q);
                acc[dst * n + x] = addmod(acc[dst * n + x], p, q);
            }
        }
    for (uint64_t x = 0; x < c->N; x++) {
        if (mode == 0) o[x] = acc[x];
        else if (mode == 1) o[x] = addmod(o[x], acc[x], q);
        else o[x] = submod(o[x], acc[x], q);
    }
    free(acc);
}

OAPI uint64_t o_mul(const o_ctx *c, uint64_t *out, const uint64_t *a, uint64_t ma, const uint64_t *b, uint64_t mb, int mode)
{
    uint64_t m = ma & mb; /* :316, :363, :403 */
    for (uint64_t i = 0; i < c->l; i++)
        if (ACTIVE(m, i)) blockmul(c, i, ROW(out, i), ROW(a, i), ROW(b, i), mode);
    return m;
}

/* polynomial_add/sub_RNS(c)_polynomial (rns_polynomial.c:437-472), negate (:565-580),
   scale (:516-526), scale_RNS (:552-563), scale_addto (:528-545). */
OAPI uint64_t o_addsub(const o_ctx *c, uint64_t *out, const uint64_t *a, uint64_t ma, const uint64_t *b, uint64_t mb, int sub)
{
    uint64_t m = ma & mb;
    for (uint64_t i = 0; i < c->l; i++)
        if (ACTIVE(m, i)) o_eltwise(sub ? O_SUB : O_ADD, ROW(out, i), ROW(a, i), ROW(b, i), 0, c->N, c->q[i]);
    return m;
}

OAPI uint64_t o_negate(const o_ctx *c, uint64_t *out, const uint64_t *a, uint64_t m)
{
    for (uint64_t i = 0; i < c->l; i++)
        if (ACTIVE(m, i)) o_eltwise(O_NEGATE, ROW(out, i), ROW(a, i), NULL, 0, c->N, c->q[i]);
    return m;
}

OAPI uint64_t o_scale(const o_ctx *c, uint64_t *out, const uint64_t *a, uint64_t m, uint64_t s, int addto)
{
    for (uint64_t i = 0; i < c->l; i++)
        if (ACTIVE(m, i)) o_eltwise(addto ? O_FMA : O_SCALE, ROW(out, i), ROW(a, i), NULL, s, c->N, c->q[i]);
    return m;
}

OAPI uint64_t o_scale_rns(const o_ctx *c, uint64_t *out, const uint64_t *a, uint64_t m, const uint64_t *s)
{
    for (uint64_t i = 0; i < c->l; i++)
        if (ACTIVE(m, i)) o_eltwise(O_SCALE, ROW(out, i), ROW(a, i), NULL, s[i], c->N, c->q[i]);
    return m;
}

/* polynomial_RNSc_add_integer (rns_polynomial.c:474-488): add to coefficient 0.
   polynomial_RNS_add_integer (:490-514): add to every slot of the first block. */
OAPI uint64_t o_add_integer(const o_ctx *c, uint64_t *out, const uint64_t *a, uint64_t m, uint64_t v, int ntt_form)
{
    for (uint64_t i = 0; i < c->l; i++) {
        if (!ACTIVE(m, i)) continue;
        uint64_t q = c->q[i];
        uint64_t vm = (v >> 63) ? (q - ((0 - v) % q)) % q : v % q;
        if (out != a) memcpy(ROW(out, i), ROW(a, i), c->N * sizeof(uint64_t));
        if (ntt_form) {
            for (uint64_t x = 0; x < c->n; x++) ROW(out, i)[x] = addmod(ROW(out, i)[x], vm, q);
        } else {
            ROW(out, i)[0] = addmod(ROW(out, i)[0], vm, q);
        }
    }
    return m;
}

/* polynomial_RNSc_mod_reduce_lifted (rns_polynomial.c:824-843): unsigned lift of limb idx */
OAPI void o_mod_reduce_lifted(const o_ctx *c, uint64_t *out, uint64_t out_mask, const uint64_t *in, uint64_t idx)
{
    for (uint64_t i = 0; i < c->l; i++) {
        if (!ACTIVE(out_mask, i)) continue;
        for (uint64_t x = 0; x < c->N; x++) ROW(out, i)[x] = ROW(in, idx)[x] % c->q[i];
    }
}

/* polynomial_round_division_RNSc_wo_free / _floor_ (rns_polynomial.c:905-976):
   for each idx of divide_mask & mask in ascending order, x_idx += floor(p/2)
   (round only); every other active limb i: x_i = (x_i [+ floor(p/2)] - x_idx) * p^-1;
   zero limb idx and drop it from the mask. */
OAPI uint64_t o_division(const o_ctx *c, uint64_t *p, uint64_t mask, uint64_t divide_mask, int round)
{
    uint64_t dm = divide_mask & mask;
    for (uint64_t idx = 0; idx < c->l; idx++) {
        if (!ACTIVE(dm, idx)) continue;
        uint64_t pr = c->q[idx], half = round ? pr / 2 : 0;
        uint64_t *xs = ROW(p, idx);
        if (round) for (uint64_t x = 0; x < c->N; x++) xs[x] = addmod(xs[x], half, pr);
        for (uint64_t i = 0; i < c->l; i++) {
            if (!ACTIVE(mask, i) || i == idx) continue;
            uint64_t q = c->q[i], inv = o_inverse_mod(pr, q), hq = half % q;
            uint64_t *xi = ROW(p, i);
            for (uint64_t x = 0; x < c->N; x++)
                xi[x] = mulmod(submod(addmod(xi[x], hq, q), xs[x] % q, q), inv, q);
        }
        memset(xs, 0, c->N * sizeof(uint64_t));
        mask &= ~(1ULL << idx);
    }
    return mask;
}

/* polynomial_base_conversion_RNSc (rns_polynomial.c:738-822) with
   init_base_conversion_params (:638-715): fast base extension without overflow
   correction, out_p = sum_j [x_j * Dhat_j]_{q_j} * (prod_{k != j} q_k mod p). */
OAPI uint64_t o_base_conversion(const o_ctx *c, uint64_t *out, uint64_t out_mask, const uint64_t *in, uint64_t in_mask)
{
    for (uint64_t i = 0; i < c->l; i++)
        if (ACTIVE(in_mask, i) && ACTIVE(out_mask, i) && out != in)
            memcpy(ROW(out, i), ROW(in, i), c->N * sizeof(uint64_t));
    uint64_t D[64], P[64], w = 0, v = 0;
    for (uint64_t i = 0; i < c->l; i++) if (ACTIVE(in_mask, i)) D[w++] = i;
    for (uint64_t i = 0; i < c->l; i++) if (ACTIVE(out_mask, i) && !ACTIVE(in_mask, i)) P[v++] = i;
    if (v == 0) return out_mask;
    uint64_t *tmp = malloc(c->N * sizeof(uint64_t));
    for (uint64_t t = 0; t < v; t++) memset(ROW(out, P[t]), 0, c->N * sizeof(uint64_t));
    for (uint64_t j = 0; j < w; j++) {
        uint64_t qj = c->q[D[j]], prod = 1;
        for (uint64_t k = 0; k < w; k++) if (k != j) prod = mulmod(prod, c->q[D[k]] % qj, qj);
        uint64_t dhat = o_inverse_mod(prod, qj);
        for (uint64_t x = 0; x < c->N; x++) tmp[x] = mulmod(ROW(in, D[j])[x], dhat, qj);
        for (uint64_t t = 0; t < v; t++) {
            uint64_t pi = c->q[P[t]], dm = 1;
            for (uint64_t k = 0; k < w; k++) if (k != j) dm = mulmod(dm, c->q[D[k]] % pi, pi);
            uint64_t *o = ROW(out, P[t]);
            for (uint64_t x = 0; x < c->N; x++) o[x] = addmod(o[x], mulmod(tmp[x] % pi, dm, pi), pi);
        }
    }
    free(tmp);
    return out_mask;
}

/* rns_compute_scaling_factors (rns_polynomial.c:1463-1494) */
OAPI void o_scaling_factors(const o_ctx *c, uint64_t *delta, uint64_t in_mask, uint64_t out_mask)
{
    uint64_t diff = out_mask & ~in_mask;
    for (uint64_t i = 0; i < c->l; i++) {
        delta[i] = 0;
        if (!ACTIVE(out_mask, i) || !ACTIVE(in_mask, i)) continue;
        uint64_t d = 1 % c->q[i];
        for (uint64_t j = 0; j < c->l; j++) if (ACTIVE(diff, j)) d = mulmod(d, c->q[j] % c->q[i], c->q[i]);
        delta[i] = d;
    }
}

/* polynomial_RNSc_scaled_lift (rns_polynomial.c:1497-1525): out mask unchanged */
OAPI void o_scaled_lift(const o_ctx *c, uint64_t *out, uint64_t out_mask, const uint64_t *in, uint64_t in_mask, const uint64_t *delta)
{
    uint64_t local[64];
    if (!delta) { o_scaling_factors(c, local, in_mask, out_mask); delta = local; }
    for (uint64_t i = 0; i < c->l; i++) {
        if (!ACTIVE(out_mask, i)) continue;
        if (ACTIVE(in_mask, i)) o_eltwise(O_SCALE, ROW(out, i), ROW(in, i), NULL, delta[i], c->N, c->q[i]);
        else memset(ROW(out, i), 0, c->N * sizeof(uint64_t));
    }
}

/* polynomial_RNSc_permute (rns_polynomial.c:996-1035): X -> X^gen in the split
   block layout, coefficient at block i position i2 is X^(i + i2*split). */
OAPI uint64_t o_permute(const o_ctx *c, uint64_t *out, const uint64_t *in, uint64_t m, uint64_t gen)
{
    uint64_t N = c->N, S = c->split, n = c->n;
    int slog = 0;
    while ((1ULL << slog) < S) slog++;
    for (uint64_t j = 0; j < c->l; j++) {
        if (!ACTIVE(m, j)) { memset(ROW(out, j), 0, N * sizeof(uint64_t)); continue; }
        uint64_t q = c->q[j];
        for (uint64_t i = 0; i < S; i++)
            for (uint64_t i2 = 0; i2 < n; i2++) {
                uint64_t e = (i + (i2 << slog)) * gen;
                uint64_t dst = (e & (S - 1)) * n + ((e & (N - 1)) >> slog);
                uint64_t v = ROW(in, j)[i * n + i2] % q;
                ROW(out, j)[dst] = (e & N) ? (q - v) % q : v;
            }
    }
    return m;
}

/* ------------------------------------------------------------------------- */
/* MLWE layer: ciphertext = (r+1) consecutive polys [a_0 .. a_{r-1}, b]         */
/* ------------------------------------------------------------------------- */

#define POLY(ct, k) ((ct) + (k) * c->l * c->N)

OAPI uint64_t o_extended_rank(uint64_t r) { return r * (r + 3) / 2; } /* mlwe.c:659-663 */

/* mlwe_tensor_product (mlwe.c:665-709), NTT domain */
OAPI uint64_t o_tensor_product(const o_ctx *c, uint64_t *out, const uint64_t *in1, const uint64_t *in2, uint64_t r, uint64_t mask)
{
    uint64_t R = o_extended_rank(r), k = 0, m = mask;
    for (uint64_t i = 0; i < r; i++)
        for (uint64_t j = i; j < r; j++) {
            m = o_mul(c, POLY(out, k), POLY(in1, i), mask, POLY(in2, j), mask, 0);
            if (i != j) o_mul(c, POLY(out, k), POLY(in1, j), mask, POLY(in2, i), mask, 1);
            k++;
        }
    for (uint64_t i = 0; i < r; i++) {
        o_mul(c, POLY(out, k), POLY(in1, i), mask, POLY(in2, r), mask, 0);
        o_mul(c, POLY(out, k), POLY(in1, r), mask, POLY(in2, i), mask, 1);
        k++;
    }
    o_mul(c, POLY(out, R), POLY(in1, r), mask, POLY(in2, r), mask, 0);
    return m;
}

/* mlwe_RNSc_GHS_hybrid_keyswitch (mlwe.c:557-598) with gadget_mul_subto_polynomial
   (gadget.c:22-37).  in: Rin+1 polys in coefficient form on in_mask.  ksk[i] is
   NULL (pass-through) or a digit-major array: for the k-th active limb of the
   input (ascending), entry k holds (r+1) polys in NTT form on ext_mask.
   out: r+1 polys, coefficient form; returns the out mask. */
OAPI uint64_t o_ghs_keyswitch(const o_ctx *c, uint64_t *out, uint64_t r, const uint64_t *in, uint64_t Rin,
                              uint64_t in_mask, const uint64_t *const *ksk, uint64_t ext_mask)
{
    uint64_t PL = c->l * c->N;
    uint64_t divide_mask = ext_mask & ~in_mask;
    memset(out, 0, (r + 1) * PL * sizeof(uint64_t)); /* mlwe_RNS_trivial_sample_of_zero */
    uint64_t *tmp = malloc(PL * sizeof(uint64_t));
    for (uint64_t i = 0; i < Rin; i++) {
        if (!ksk[i]) continue;
        uint64_t kidx = 0;
        for (uint64_t j = 0; j < c->l; j++) {
            if (!ACTIVE(in_mask, j)) continue;
            o_mod_reduce_lifted(c, tmp, ext_mask, POLY(in, i), j);
            o_to_ntt(c, tmp, tmp, ext_mask);
            const uint64_t *entry = ksk[i] + kidx * (r + 1) * PL;
            for (uint64_t comp = 0; comp <= r; comp++)
                o_mul(c, POLY(out, comp), entry + comp * PL, ext_mask, tmp, ext_mask, 2);
            kidx++;
        }
    }
    free(tmp);
    uint64_t m = ext_mask;
    for (uint64_t comp = 0; comp <= r; comp++) {
        o_to_coeff(c, POLY(out, comp), POLY(out, comp), ext_mask);
        m = o_division(c, POLY(out, comp), ext_mask, divide_mask, 1);
    }
    uint64_t keep = 0, mout = m;
    for (uint64_t i = 0; i < Rin; i++) {
        if (ksk[i]) continue;
        mout = o_addsub(c, POLY(out, keep), POLY(out, keep), m, POLY(in, i), in_mask, 0);
        keep++;
    }
    mout = o_addsub(c, POLY(out, r), POLY(out, r), m, POLY(in, Rin), in_mask, 0);
    return mout;
}

/* mlwe_multiply (mlwe.c:711-760): tensor, INTT, key-switch with rlk, NTT.
   in1/in2: r+1 polys in NTT form on mask; out: r+1 polys in NTT form. */
OAPI uint64_t o_multiply(const o_ctx *c, uint64_t *out, const uint64_t *in1, const uint64_t *in2, uint64_t r,
                         uint64_t mask, const uint64_t *const *rlk, uint64_t ext_mask)
{
    uint64_t R = o_extended_rank(r), PL = c->l * c->N;
    uint64_t *ext = calloc((R + 1) * PL, sizeof(uint64_t));
    uint64_t m = o_tensor_product(c, ext, in1, in2, r, mask);
    for (uint64_t k = 0; k <= R; k++) o_to_coeff(c, POLY(ext, k), POLY(ext, k), m);
    uint64_t mo = o_ghs_keyswitch(c, out, r, ext, R, m, rlk, ext_mask);
    for (uint64_t k = 0; k <= r; k++) o_to_ntt(c, POLY(out, k), POLY(out, k), mo);
    free(ext);
    return mo;
}

/* mlwe_automorphism_RNSc_GHS (mlwe.c:314-326): permute every component, then key-switch */
OAPI uint64_t o_automorphism(const o_ctx *c, uint64_t *out, const uint64_t *in, uint64_t r, uint64_t mask,
                             uint64_t gen, const uint64_t *const *ksk, uint64_t ext_mask)
{
    uint64_t PL = c->l * c->N;
    uint64_t *tmp = calloc((r + 1) * PL, sizeof(uint64_t));
    for (uint64_t k = 0; k <= r; k++) o_permute(c, POLY(tmp, k), POLY(in, k), mask, gen);
    uint64_t m = o_ghs_keyswitch(c, out, r, tmp, r, mask, ksk, ext_mask);
    free(tmp);
    return m;
}

/* mlwe_round_division_RNSc (mlwe.c:762-772) on every component: the CKKS rescale
   of ckks.py:128-148 is INTT + this with the dropped top prime. */
OAPI uint64_t o_mlwe_round_division(const o_ctx *c, uint64_t *ct, uint64_t r, uint64_t mask, uint64_t divide_mask)
{
    uint64_t m = mask;
    for (uint64_t k = 0; k <= r; k++) m = o_division(c, POLY(ct, k), mask, divide_mask, 1);
    return m;
}

/* ------------------------------------------------------------------------- */
/* Monomial products (rns_polynomial.c:1048-1168) and MGSW (mgsw.c:6-88).      */
/* ------------------------------------------------------------------------- */

/* polynomial_RNSc_mul_by_xai / _minus1: out = X^a in (minus1: (X^a - 1) in) mod X^N + 1,
   coefficient form, a taken mod 2N; a == 0 copies (minus1: zeroes) the rows. */
OAPI uint64_t o_mul_by_xai(const o_ctx *c, uint64_t *out, const uint64_t *in, uint64_t m, uint64_t a, int minus1)
{
    const uint64_t N = c->N;
    a &= 2 * N - 1;
    for (uint64_t j = 0; j < c->l; j++) {
        if (!ACTIVE(m, j)) continue;
        const uint64_t q = c->q[j];
        const uint64_t *x = ROW(in, j);
        uint64_t *y = ROW(out, j);
        for (uint64_t i = 0; i < N; i++) {
            /* coefficient i of X^a * x: sum over e = k + a with e = i or e = i + N (mod 2N) */
            uint64_t v;
            if (a < N) v = i >= a ? x[i - a] % q : (q - x[i + N - a] % q) % q;
            else {
                const uint64_t b = a - N; /* X^a = -X^b */
                v = i >= b ? (q - x[i - b] % q) % q : x[i + N - b] % q;
            }
            y[i] = minus1 ? submod(v, x[i] % q, q) : v;
        }
    }
    return m;
}

/* mgsw_external_product (mgsw.c:6-49), gadget_mul_addto_polynomial (gadget.c:5-20).
   in: r+1 coefficient-form polys on in_mask; mgsw: (r+1)*ell entries of [r+1][l][N]
   (NTT form, extended mask ext_mask); out: r+1 polys, NTT form.  Returns out's mask. */
OAPI uint64_t o_external_product(const o_ctx *c, uint64_t *out, uint64_t r, const uint64_t *in, uint64_t in_mask,
                                 const uint64_t *mgsw, uint64_t ell, uint64_t special_primes, uint64_t ext_mask)
{
    const uint64_t PL = c->l * c->N;
    memset(out, 0, (r + 1) * PL * sizeof(uint64_t));
    uint64_t *tmp = malloc(PL * sizeof(uint64_t));
    for (uint64_t comp = 0; comp <= r; comp++) {
        uint64_t kidx = 0;
        for (uint64_t j = 0; j < c->l; j++) {
            if (!ACTIVE(in_mask, j)) continue;
            o_mod_reduce_lifted(c, tmp, ext_mask, POLY(in, comp), j);
            o_to_ntt(c, tmp, tmp, ext_mask);
            const uint64_t *entry = mgsw + (comp * ell + kidx) * (r + 1) * PL;
            for (uint64_t oc = 0; oc <= r; oc++) o_mul(c, POLY(out, oc), POLY(entry, oc), ext_mask, tmp, ext_mask, 1);
            kidx++;
        }
    }
    free(tmp);
    uint64_t m = ext_mask;
    if (special_primes > 0) {
        uint64_t dm = 0, cnt = 0;
        for (int idx = 63; idx >= 0 && cnt < special_primes; idx--)
            if (ACTIVE(m, (uint64_t)idx)) { dm |= 1ULL << idx; cnt++; }
        uint64_t mo = m;
        for (uint64_t oc = 0; oc <= r; oc++) {
            o_to_coeff(c, POLY(out, oc), POLY(out, oc), m);
            if (dm) mo = o_division(c, POLY(out, oc), m, dm, 1);
            o_to_ntt(c, POLY(out, oc), POLY(out, oc), mo);
        }
        m = mo;
    }
    return m;
}

/* mgsw_CMUX (mgsw.c:51-71): out = ExtProd(mgsw, in2 - in1) + NTT(in1); in1, in2 coefficient form. */
OAPI uint64_t o_cmux(const o_ctx *c, uint64_t *out, uint64_t r, const uint64_t *in1, const uint64_t *in2, uint64_t mask,
                     const uint64_t *mgsw, uint64_t ell, uint64_t special_primes, uint64_t ext_mask)
{
    const uint64_t PL = c->l * c->N;
    uint64_t *diff = calloc((r + 1) * PL, sizeof(uint64_t)), *a = calloc((r + 1) * PL, sizeof(uint64_t));
    for (uint64_t k = 0; k <= r; k++) {
        o_addsub(c, POLY(diff, k), POLY(in2, k), mask, POLY(in1, k), mask, 1);
        o_to_ntt(c, POLY(a, k), POLY(in1, k), mask);
    }
    uint64_t m = o_external_product(c, out, r, diff, mask, mgsw, ell, special_primes, ext_mask);
    uint64_t mo = m;
    for (uint64_t k = 0; k <= r; k++) mo = o_addsub(c, POLY(out, k), POLY(out, k), m, POLY(a, k), mask, 0);
    free(diff);
    free(a);
    return mo;
}
