/*
 * vfhe_hip.h -- C ABI of the `hip` engine for vfhe's RNS polynomial hot path.
 *
 * The engine is a drop-in for the reference's arith + mlwe native surface on
 * the hot path: the functions below keep the reference's names, argument
 * order, argument meaning and error behaviour (failed preconditions abort, as
 * the reference's live asserts do; allocation failure exits), and the structs
 * keep the reference's public field layout so that Python code that casts a
 * handle and reads `rns_mask`, `coeffs`, `l` or `w` keeps working.
 *
 * Memory model.  Every RNS polynomial owns a device buffer of allocated_l
 * limb rows of N uint64 each (limb-major, 256-byte aligned rows) in HBM.  The
 * `coeffs[i]` rows of the public struct are host mirror rows.  All compute
 * entry points run on the device on one in-order HIP stream and return without
 * waiting; a polynomial's host rows are refreshed only by
 * polynomial_sync_to_host(), and host writes into `coeffs` become visible to
 * the device after polynomial_mark_host_dirty().  Those two functions, the
 * vfhe_hip_* functions and the batched entry points have no counterpart in the
 * reference; its Python side needs the two residency calls around its direct
 * reads and writes of `coeffs` (modules/arith/python/src/vfhe/arith/polynomial.py:
 * 437-469): polynomial_sync_to_host before reading, polynomial_mark_host_dirty
 * after writing.
 *
 * Standalone single-limb kernels (ntt_forward, ntt_reverse, mod_eltwise_*)
 * accept host pointers (staged through the device, synchronous, as the
 * reference's C tests call them) or device pointers (asynchronous).
 */
#ifndef VFHE_HIP_H
#define VFHE_HIP_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Types (public prefix identical to the reference)                         */
/* ------------------------------------------------------------------------ */

/* Per-(n, q) NTT context.  Opaque to Python (modules/arith/python/cdef/arith.cdef:8);
   replaces struct _NTT_proc of modules/arith/c/include/arith.h:29-46.  The
   engine keeps n, q and the roots first, then its own device tables. */
typedef struct _NTT_proc *NTT_proc;

/* Prime pool for one (N, split degree): modules/arith/c/include/arith.h:48-54,
   cdef modules/arith/python/cdef/arith.cdef:10-16.  w[i][k] = psi_i^(2 brv(k)+1). */
typedef struct _incNTT {
    NTT_proc *ntt;
    uint64_t split_degree;
    uint64_t **w;
    uint64_t N, l;
} *incNTT;

/* RNS polynomial: modules/arith/c/include/arith.h:72-87, cdef arith.cdef:18-23.
   coeffs[i] are host mirror rows (see the memory model above). */
typedef struct _RNS_Polynomial {
    uint64_t **coeffs;
    incNTT ntt;
    uint64_t rns_mask;
    uint64_t allocated_l;
} *RNS_Polynomial;
typedef RNS_Polynomial RNSc_Polynomial;

/* MLWE sample: modules/mlwe/c/include/mlwe.h:45-55, cdef mlwe.cdef:27-31 */
typedef struct _RNS_MLWE {
    RNS_Polynomial *a, b;
    uint64_t r;
} *RNS_MLWE;
typedef RNS_MLWE RNSc_MLWE;

/* MLWE key: modules/mlwe/c/include/mlwe.h:37-43, cdef mlwe.cdef:33-40 */
typedef struct _IntPolynomial {
    uint64_t *coeffs;
    uint64_t N;
} *IntPolynomial;
typedef struct _RNS_MLWE_Key {
    IntPolynomial *s;
    RNS_Polynomial *s_RNS;
    uint64_t N, l, r;
    double sigma;
} *RNS_MLWE_Key;

/* LWE key and sample: modules/mlwe/c/include/mlwe.h, cdef mlwe.cdef:11-25.
   Host-resident arrays (a[limb][n], b[limb], s[limb][n]). */
typedef struct _LWE_Key {
    uint64_t **s;
    uint64_t n;
    uint64_t l;
    incNTT ntt;
    double sigma;
} *LWE_Key;
typedef struct _LWE {
    uint64_t **a;
    uint64_t *b;
    uint64_t n;
    uint64_t l;
    incNTT ntt;
} *LWE;

/* Fast base conversion parameters: modules/arith/c/include/arith.h:211-221 */
typedef struct _RNS_BaseConversionParams {
    uint64_t in_mask;
    uint64_t out_mask;
    uint32_t w;
    uint32_t v;
    uint32_t *D;
    uint32_t *P;
    uint64_t *Dhat;
    uint64_t **D_mod_p;
} *RNS_BaseConversionParams;

/* ------------------------------------------------------------------------ */
/* Number theory (host): modules/arith/c/src/number_theory.c:6-217           */
/* ------------------------------------------------------------------------ */
uint64_t power_mod(uint64_t base, uint64_t exp, uint64_t mod);       /* number_theory.c:6-20 */
uint64_t inverse_mod(uint64_t a, uint64_t m);                        /* number_theory.c:103-109 */
uint64_t inverse_mod_eea(uint64_t a, uint64_t p);                    /* number_theory.c:22-101 */
bool is_prime(uint64_t n);                                           /* number_theory.c:111-152 */
uint64_t generate_Nth_root_of_unity(uint64_t q, uint64_t n);         /* number_theory.c:154-177 */
uint64_t next_special_prime(uint64_t x, uint64_t n, bool primitive); /* number_theory.c:179-217 */
void bit_rev(uint64_t *out, uint64_t *in, uint64_t n, uint64_t log_n); /* misc.c:186-192 */

/* ------------------------------------------------------------------------ */
/* Single-limb kernels: modules/arith/c/include/arith.h:123-156              */
/* ------------------------------------------------------------------------ */
NTT_proc ntt_new_proc(uint64_t n, uint64_t q);                       /* ntt.c:265-313 */
void ntt_free_proc(NTT_proc proc);                                   /* ntt.c:347-352 */
void ntt_forward(uint64_t *out, uint64_t *in, NTT_proc proc);        /* ntt.c:315-329 */
void ntt_reverse(uint64_t *out, uint64_t *in, NTT_proc proc);        /* ntt.c:331-345 */
NTT_proc *new_ntt_list(uint64_t *primes, uint64_t N, uint64_t l);    /* misc.c:118-126 */

uint64_t add_modq(uint64_t a, uint64_t b, uint64_t q);               /* mod.c:54-58 */
uint64_t sub_modq(uint64_t a, uint64_t b, uint64_t q);               /* mod.c:60 */
uint64_t negate_modq(uint64_t a, uint64_t q);                        /* mod.c:62 */
uint64_t mul_modq(uint64_t a, uint64_t b, NTT_proc proc);            /* mod.c:64-67 */

void mod_eltwise_mul(uint64_t *out, uint64_t *in1, uint64_t *in2, uint64_t n, NTT_proc proc);       /* mod.c:69-83 */
void mod_eltwise_mul_addto(uint64_t *out, uint64_t *in1, uint64_t *in2, uint64_t n, NTT_proc proc); /* mod.c:85-99 */
void mod_eltwise_mul_subto(uint64_t *out, uint64_t *in1, uint64_t *in2, uint64_t n, NTT_proc proc); /* mod.c:101-115 */
void mod_eltwise_scale(uint64_t *out, uint64_t *in, uint64_t scale, uint64_t n, NTT_proc proc);     /* mod.c:117-131 */
void mod_eltwise_fma(uint64_t *out, uint64_t *in, uint64_t scale, uint64_t n, NTT_proc proc);       /* mod.c:133-147 */
void mod_eltwise_add_scalar(uint64_t *out, uint64_t *in, uint64_t scalar, uint64_t n, NTT_proc proc); /* mod.c:149-163 */
void mod_eltwise_sub_scalar(uint64_t *out, uint64_t *in, uint64_t scalar, uint64_t n, NTT_proc proc); /* mod.c:165-179 */
void mod_eltwise_negate(uint64_t *out, uint64_t *in, uint64_t n, NTT_proc proc);                    /* mod.c:181-195 */
void mod_eltwise_add(uint64_t *out, uint64_t *in1, uint64_t *in2, uint64_t n, NTT_proc proc);       /* mod.c:197-211 */
void mod_eltwise_sub(uint64_t *out, uint64_t *in1, uint64_t *in2, uint64_t n, NTT_proc proc);       /* mod.c:213-227 */
void mod_eltwise_reduce(uint64_t *out, uint64_t *in, uint64_t n, NTT_proc proc);                    /* mod.c:229-243 */
void mod_eltwise_reduce_signed(uint64_t *out, int64_t *in, uint64_t n, NTT_proc proc);              /* mod.c:245-259 */
void mod_reduce_array_mp(uint64_t *out, uint64_t *in_high, uint64_t *in_low, uint64_t n, NTT_proc proc); /* mod.c:261-276 */

/* ------------------------------------------------------------------------ */
/* RNS pool and polynomials: modules/arith/c/src/rns_polynomial.c            */
/* ------------------------------------------------------------------------ */
incNTT new_incomplete_ntt_list(uint64_t *primes, uint64_t split_degree, uint64_t N, uint64_t l); /* :7-31 */
void incNTT_extend_with_primes(incNTT ntt, uint64_t *new_primes, uint64_t count);               /* :33-61 */
uint64_t **incNTT_get_rou_matrix(incNTT ntt);                                                  /* :63 */

RNS_Polynomial polynomial_new_RNS_polynomial(uint64_t N, uint64_t rns_mask, incNTT ntt);        /* :65-78 */
RNS_Polynomial *polynomial_new_RNS_polynomial_array(uint64_t size, uint64_t N, uint64_t rns_mask, incNTT ntt); /* :80-90 */
RNS_Polynomial *polynomial_new_array_of_RNS_polynomials(uint64_t N, uint64_t rns_mask, uint64_t size, incNTT ntt); /* :174-181 */
/* Frees are ordered on the calling thread's lane (see vfhe_hip_set_lane for work on other lanes). */
void free_RNS_polynomial(void *p);                                                              /* :154-163 */
void free_RNS_polynomial_array(uint64_t size, RNS_Polynomial *p);                               /* :165-172 */
void polynomial_copy_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in);                     /* :126-136 */
void polynomial_copy_RNSc_polynomial(RNSc_Polynomial out, RNSc_Polynomial in);                  /* :138-141 */
void polynomial_RNS_zero(RNS_Polynomial p);                                                     /* :143-152 */
bool polynomial_eq(RNS_Polynomial a, RNS_Polynomial b);                                         /* :92-123 */
void int_array_to_RNS(RNS_Polynomial out, uint64_t *in);                                        /* :204-223 */
void array_to_RNS(RNS_Polynomial out, uint64_t **in);                                           /* :225-240 */

void polynomial_RNSc_to_RNS(RNS_Polynomial out, RNSc_Polynomial in);                            /* :582-597 */
void polynomial_RNS_to_RNSc(RNSc_Polynomial out, RNS_Polynomial in);                            /* :599-614 */

void polynomial_mul_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, RNS_Polynomial in2);       /* :311-356 */
void polynomial_mul_addto_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, RNS_Polynomial in2); /* :358-396 */
void polynomial_mul_subto_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, RNS_Polynomial in2); /* :398-435 */
void polynomial_multo_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in);                          /* :273-309 */
void polynomial_add_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, RNS_Polynomial in2);       /* :468-472 */
void polynomial_add_RNSc_polynomial(RNSc_Polynomial out, RNSc_Polynomial in1, RNSc_Polynomial in2);   /* :455-466 */
void polynomial_sub_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, RNS_Polynomial in2);       /* :437-448 */
void polynomial_sub_RNSc_polynomial(RNSc_Polynomial out, RNSc_Polynomial in1, RNSc_Polynomial in2);   /* :450-453 */
void polynomial_RNSc_negate(RNSc_Polynomial out, RNSc_Polynomial in);                                 /* :565-575 */
void polynomial_RNS_negate(RNS_Polynomial out, RNS_Polynomial in);                                    /* :577-580 */
void polynomial_scale_RNSc_polynomial(RNSc_Polynomial out, RNSc_Polynomial in1, uint64_t scale);      /* :516-526 */
void polynomial_scale_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, uint64_t scale);         /* :547-550 */
void polynomial_scale_RNS_polynomial_RNS(RNS_Polynomial out, RNS_Polynomial in1, uint64_t *scale);    /* :552-563 */
void polynomial_scale_addto_RNSc_polynomial(RNSc_Polynomial out, RNSc_Polynomial in1, uint64_t scale); /* :528-540 */
void polynomial_scale_addto_RNS_polynomial(RNS_Polynomial out, RNS_Polynomial in1, uint64_t scale);   /* :542-545 */
void polynomial_RNSc_add_integer(RNSc_Polynomial out, RNSc_Polynomial in1, uint64_t in2);              /* :474-488 */
void polynomial_RNS_add_integer(RNS_Polynomial out, RNS_Polynomial in1, uint64_t in2);                /* :490-514 */

void polynomial_RNSc_mod_reduce_lifted(RNSc_Polynomial out, RNSc_Polynomial in, uint64_t idx);        /* :824-843 (ModUp digit) */
void polynomial_RNSc_mod_reduce(RNSc_Polynomial out, RNSc_Polynomial in);                             /* :845-854 */
void polynomial_round_division_RNSc_wo_free(RNSc_Polynomial out, uint64_t divide_mask);               /* :939-976 (rescale, ModDown) */
void polynomial_floor_division_RNSc_wo_free(RNSc_Polynomial out, uint64_t divide_mask);               /* :905-937 */
void polynomial_round_division_RNSc(RNSc_Polynomial out);                                             /* :987-994 */
void polynomial_floor_division_RNSc(RNSc_Polynomial out);                                             /* :978-985 */

RNS_BaseConversionParams init_base_conversion_params(incNTT ntt, uint64_t in_mask, uint64_t out_mask); /* :638-715 */
void free_base_conversion_params(RNS_BaseConversionParams params);                                    /* :717-736 */
void polynomial_base_conversion_RNSc(RNSc_Polynomial out, RNSc_Polynomial in, RNS_BaseConversionParams params); /* :738-822 */
void rns_compute_scaling_factors(uint64_t *delta_out, incNTT ntt, uint64_t in_mask, uint64_t out_mask); /* :1463-1494 */
void polynomial_RNSc_scaled_lift(RNSc_Polynomial out, RNSc_Polynomial in, uint64_t *delta);           /* :1497-1525 */
void polynomial_RNSc_permute(RNSc_Polynomial out, RNSc_Polynomial in, uint64_t gen);                  /* :996-1035 */
/* Slots of NTT-form polynomials (device): */
void polynomial_RNS_broadcast_slot(RNS_Polynomial out, RNS_Polynomial in, uint64_t slot_idx);           /* :1228-1244 */
void polynomial_RNS_rotate_slot(RNS_Polynomial out, RNS_Polynomial in, uint64_t rot);                   /* :1262-1285 */
void polynomial_RNS_copy_slot(RNS_Polynomial out, uint64_t dst, RNS_Polynomial in, uint64_t src);       /* :1287-1300 */
/* Slot-wise inverse (:1384-1432): 0, -2 for a zero slot; -1 for split degree > 1, whose
   extension-field slots (field.c) this engine does not implement. */
int polynomial_RNS_inverse(RNS_Polynomial out, RNS_Polynomial in);

/* ------------------------------------------------------------------------ */
/* MLWE: modules/mlwe/c/src/mlwe.c, gadget.c                                 */
/* ------------------------------------------------------------------------ */
RNS_MLWE mlwe_alloc_RNS_sample(uint64_t N, uint64_t r, uint64_t mask, incNTT ntt);                    /* mlwe.c:98-106 */
RNSc_MLWE mlwe_alloc_RNSc_sample(uint64_t N, uint64_t r, uint64_t mask, incNTT ntt);                  /* mlwe.c:108-111 */
RNS_MLWE *mlwe_alloc_RNS_sample_array(uint64_t size, uint64_t N, uint64_t r, uint64_t mask, incNTT ntt); /* mlwe.c:129-139 */
RNS_MLWE *mlwe_create_copy_array(RNS_MLWE *in, uint64_t size);                                        /* mlwe.c:121-127 */
void free_mlwe_RNS_sample(void *p);                                                                   /* mlwe.c:182-188 */
void free_RNS_mlwe_array(uint64_t size, RNS_MLWE *v);                                                 /* mlwe.c:152-159 */
/* Engine extension: free_mlwe_RNS_sample (mlwe.c:182-188) over an array, one lock acquisition;
   the array itself is not freed. */
void vfhe_hip_free_mlwe_samples(RNS_MLWE *samples, uint64_t count);
void mlwe_copy_RNS_sample(RNS_MLWE out, RNS_MLWE in);                                                 /* mlwe.c:168-175 */
void mlwe_RNS_trivial_sample_of_zero(RNS_MLWE out);                                                   /* mlwe.c:302-312 */
void mlwe_RNSc_to_RNS(RNS_MLWE out, RNSc_MLWE in);                                                    /* mlwe.c:513-520 */
void mlwe_RNS_to_RNSc(RNSc_MLWE out, RNS_MLWE in);                                                    /* mlwe.c:522-529 */
void mlwe_RNS_mul_by_poly(RNS_MLWE out, RNS_MLWE in, RNS_Polynomial poly);                            /* mlwe.c:254-261 */
void mlwe_RNS_mul_addto_by_poly(RNS_MLWE out, RNS_MLWE in, RNS_Polynomial poly);                      /* mlwe.c:263-270 */
void mlwe_RNS_mul_subto_by_poly(RNS_MLWE out, RNS_MLWE in, RNS_Polynomial poly);                      /* mlwe.c:272-279 */
void mlwe_add_RNSc_sample(RNSc_MLWE out, RNSc_MLWE in1, RNSc_MLWE in2);                               /* mlwe.c:446-453 */
void mlwe_sub_RNSc_sample(RNSc_MLWE out, RNSc_MLWE in1, RNSc_MLWE in2);                               /* mlwe.c:484-491 */
void mlwe_add_RNSc_polynomial(RNSc_MLWE out, RNSc_MLWE in1, RNSc_Polynomial in2);                     /* mlwe.c:464-467 */
void mlwe_sub_RNSc_polynomial(RNSc_MLWE out, RNSc_MLWE in1, RNSc_Polynomial in2);                     /* mlwe.c:469-472 */
void mlwe_RNS_add_polynomial(RNS_MLWE out, RNS_MLWE in1, RNS_Polynomial in2);                          /* mlwe.c:474-477 */
void mlwe_RNS_sub_polynomial(RNS_MLWE out, RNS_MLWE in1, RNS_Polynomial in2);                          /* mlwe.c:479-482 */
void mlwe_scale_RNS_mlwe_RNS(RNS_MLWE c, uint64_t *scale);                                            /* mlwe.c:217-225 */
void mlwe_round_division_RNSc(RNSc_MLWE out, uint64_t divide_mask);                                   /* mlwe.c:762-772 */
uint64_t mlwe_extended_rank(uint64_t r);
/* Sample extraction (mlwe.c:40-68): coefficient idx of a coefficient-form sample as an
   LWE sample of dimension r*N over the active limbs of a[0] (host arrays). */
LWE mlwe_extract_LWE(RNSc_MLWE in, uint64_t idx);
/* Trace and packing (mlwe.c:328-445, 596-657), on the device. */
void mlwe_partial_trace(RNSc_MLWE out, RNSc_MLWE in, uint64_t *gens, RNS_MLWE ***ksks, uint64_t size,
                        uint64_t lvl);                                                                   /* mlwe.c:328-341 */
void mlwe_trace(RNSc_MLWE out, RNSc_MLWE in, RNS_MLWE ***ksks, uint64_t lvl);                           /* mlwe.c:436-445 */
void mlwe_full_packing_keyswitch(RNS_MLWE out, LWE *in, uint64_t size, RNS_MLWE **ksk, uint64_t lvl);    /* mlwe.c:343-434 */
void mlwe_full_packing_keyswitch_scaled(RNSc_MLWE *vec, uint64_t ell, RNS_MLWE ***ksks, uint64_t lvl);  /* mlwe.c:600-657 */

/* LWE (modules/mlwe/c/src/lwe.c), host arrays; randomness from the engine's ChaCha20 key. */
LWE_Key lwe_alloc_key(uint64_t n, uint64_t l, incNTT ntt);                                      /* lwe.c:6-18 */
LWE lwe_alloc_sample(uint64_t n, uint64_t l, incNTT ntt);                                       /* lwe.c:20-33 */
void free_lwe_sample(LWE c);                                                                    /* lwe.c:35-44 */
void free_lwe_key(LWE_Key key);                                                                 /* lwe.c:46-54 */
LWE_Key lwe_new_key(uint64_t n, uint64_t l, incNTT ntt, double sec_sigma, double err_sigma);    /* lwe.c:56-70 */
LWE_Key lwe_new_sparse_ternary_key(uint64_t n, uint64_t l, incNTT ntt, uint64_t h, double err_sigma); /* lwe.c:72-95 */
void lwe_sample(LWE c, uint64_t *m, LWE_Key key);                                               /* lwe.c:97-123 */
LWE lwe_new_sample(uint64_t *m, LWE_Key key);                                                   /* lwe.c:125-130 */
LWE lwe_new_trivial_sample(uint64_t *m, uint64_t n, uint64_t l, incNTT ntt);                    /* lwe.c:132-142 */
void lwe_phase(uint64_t *out, LWE c, LWE_Key key);                                              /* lwe.c:144-158 */
void lwe_subto(LWE out, LWE in);                                                                /* lwe.c:160-169 */                                                              /* mlwe.c:659-663 */
void mlwe_tensor_product(RNS_Polynomial *out, RNS_MLWE in1, RNS_MLWE in2);                            /* mlwe.c:665-709 */
void mlwe_RNSc_GHS_hybrid_keyswitch(RNSc_MLWE out, RNSc_MLWE in, RNS_MLWE **ksk, uint64_t lvl);       /* mlwe.c:557-598 + gadget.c:22-37 */
void mlwe_multiply(RNS_MLWE out, RNS_MLWE in1, RNS_MLWE in2, RNS_MLWE **ksk);                         /* mlwe.c:711-760 */
void mlwe_automorphism_RNSc_GHS(RNSc_MLWE out, RNSc_MLWE in, uint64_t gen, RNS_MLWE **ksk, uint64_t lvl); /* mlwe.c:314-326 */
RNS_MLWE_Key mlwe_new_RNS_key_from_array(uint64_t *array, uint64_t N, uint64_t r, uint64_t l, incNTT ntt, double sigma); /* mlwe.c:28-38 */
void free_mlwe_RNS_key(RNS_MLWE_Key key);                                                             /* mlwe.c:21-26 */
void mlwe_RNS_phase(RNS_Polynomial out, RNS_MLWE in, RNS_MLWE_Key key);                               /* mlwe.c:243-252 */

/* ------------------------------------------------------------------------ */
/* Engine identity: modules/misc/python/cdef/misc.cdef:8, vfhe_cpu.cdef:6     */
/* ------------------------------------------------------------------------ */
const char *vfhe_engine_active(void);        /* returns "hip" */
void *safe_aligned_malloc(uint64_t size);    /* misc_tp.c:48-62 (arith.cdef:26): 64-byte aligned host memory, freed with libc free */
int vfhe_cpu_supports(const char *capability); /* "hip": a gfx950 device is visible */

/* ------------------------------------------------------------------------ */
/* Sampling (on the device; ChaCha20 keystream, so the streams differ from     */
/* the reference engines' AES-CTR / BLAKE3 streams, as those differ from each   */
/* other; distributions and masks follow the reference)                         */
/* ------------------------------------------------------------------------ */

/* modules/misc/c/include/misc.h:46-47 (prng.c:132-143): deterministic key from
   a seed (tests), or back to OS entropy. */
void vfhe_prng_set_deterministic_seed(uint64_t seed);
void vfhe_prng_clear_deterministic_seed(void);
/* modules/arith/c/include/arith.h:200 (rns_polynomial.c:242-253): uniform residues
   on every active limb. */
void polynomial_gen_random_RNSc_polynomial(RNSc_Polynomial out);
/* modules/arith/c/include/arith.h:267 (rns_polynomial.c:255-271): one rounded
   N(0, sigma) vector reduced into every active limb. */
void polynomial_gen_gaussian_RNSc_polynomial(RNSc_Polynomial out, double sigma);
/* modules/mlwe/c/include/mlwe.h:74-91 (mlwe.c:70-83, 190-241, 281-300) */
RNS_MLWE_Key mlwe_new_RNS_gaussian_key(uint64_t N, uint64_t r, uint64_t l, double key_sigma, incNTT ntt,
                                       double sigma);
void mlwe_RNS_sample_of_zero(RNS_MLWE out, RNS_MLWE_Key key);
void mlwe_RNSc_sample_of_zero(RNSc_MLWE out, RNS_MLWE_Key key);
RNS_MLWE mlwe_new_RNS_sample_of_zero(RNS_MLWE_Key key);
RNSc_MLWE mlwe_new_RNSc_sample_of_zero(RNS_MLWE_Key key);
RNS_MLWE mlwe_new_RNS_trivial_sample_of_zero(uint64_t N, uint64_t r, uint64_t mask, incNTT ntt);
void mlwe_RNSc_sample(RNSc_MLWE out, RNS_MLWE_Key key, RNSc_Polynomial m);

/* ------------------------------------------------------------------------ */
/* CKKS encoder FFT: modules/arith/python/cdef/arith.cdef:83-93                */
/* (modules/arith/c/src/complex.c).  Rows are 2n doubles, real then imaginary; */
/* tables are the caller's roots, ws[m + i] for group i of the stage with m    */
/* groups.  Results are bit-identical to the avx512ifma engine (option          */
/* "fft.fma" = 1, default) or to the portable engine ("fft.fma" = 0).           */
/* ------------------------------------------------------------------------ */
void *load_rous_CT(double *rous_real, double *rous_imag, uint64_t size);
void *load_rous_GS(double *rous_real, double *rous_imag, uint64_t size);
void CT_NR(double *x, void *ws, uint64_t n);
void GS_RN(double *x, void *ws, uint64_t n);
void bit_reverse_array(double *v, uint64_t N, uint32_t prec);
void complex_poly_scale_double(double *v, double scale, uint64_t N);
void complex_poly_round_to_RNS(RNS_Polynomial out, double *in, uint64_t N);
void complex_polys_ifft_scale_round_to_RNS_batch(void **rows_in, void **outs_rns, uint64_t count, uint64_t n_complex,
                                                 uint32_t log_prec, void *gs_ws, double temp_delta);

/* ------------------------------------------------------------------------ */
/* Monomial products and MGSW (CGGI building blocks)                          */
/* ------------------------------------------------------------------------ */

/* modules/arith/c/include/arith.h:240-241 (rns_polynomial.c:1048-1168):
   out = X^a * in and out = (X^a - 1) * in mod X^N + 1, coefficient form, out != in. */
void polynomial_RNSc_mul_by_xai(RNSc_Polynomial out, RNSc_Polynomial in, uint64_t a);
void polynomial_RNSc_mul_by_xai_minus1(RNSc_Polynomial out, RNSc_Polynomial in, uint64_t a);
/* modules/mlwe/c/include/mlwe.h:96-97 (mlwe.c:493-509): the same on every component. */
void mlwe_RNSc_mul_by_xai(RNSc_MLWE out, RNSc_MLWE in, uint64_t a);
void mlwe_RNSc_mul_by_xai_minus1(RNSc_MLWE out, RNSc_MLWE in, uint64_t a);
/* modules/mlwe/c/include/mlwe.h:143-148 (mgsw.c:6-88): external product of an
   MGSW sample (ell gadget rows per input component, NTT form) with a
   coefficient-form MLWE sample, output in NTT form after the round division by
   the top `special_primes` limbs; CMUX = ExtProd(in2 - in1) + in1; NCMUX =
   CMUX(in1, automorphism_{2N-1}(in2)). */
void mgsw_external_product(RNS_MLWE out, RNS_MLWE *mgsw, RNSc_MLWE in, uint64_t ell, uint64_t special_primes);
void mgsw_CMUX(RNS_MLWE out, RNSc_MLWE in1, RNSc_MLWE in2, RNS_MLWE *mgsw, uint64_t ell, uint64_t special_primes);
void mgsw_NCMUX(RNS_MLWE out, RNSc_MLWE in1, RNSc_MLWE in2, RNS_MLWE *mgsw, RNS_MLWE **ksk, uint64_t ell,
                uint64_t special_primes);

/* ------------------------------------------------------------------------ */
/* Engine extensions (no counterpart in the reference)                       */
/* ------------------------------------------------------------------------ */

/* Residency: refresh the host rows from the device (waits), or declare that the
   host rows were written and must be uploaded before the next device use. */
void polynomial_sync_to_host(RNS_Polynomial p);
void polynomial_mark_host_dirty(RNS_Polynomial p);
/* Raw device address of limb row `limb` (for zero-copy interop, e.g. DLPack). */
uint64_t *polynomial_device_row(RNS_Polynomial p, uint64_t limb);

/* Runtime: wait for all queued device work; return the last HIP error code
   (0 when none) and clear it; device name. */
void vfhe_hip_synchronize(void);
/* Select the calling thread's lane (0..7): a HIP stream with its own staging
   ring; lane 0 is the default.  Work on different lanes may overlap on the
   device, so a caller that uses several lanes orders dependent work itself
   (vfhe_hip_synchronize waits for every lane).  This includes frees: a
   polynomial or sample freed on one lane returns its device block to that
   lane, whose next allocation may reuse it at once, ordered only by that
   lane's stream; work still reading or writing it on another lane must have
   completed (vfhe_hip_synchronize) before the free.  Returns the previous
   lane, or -1 for an invalid one. */
int vfhe_hip_set_lane(int lane);
int vfhe_hip_last_error(void);
const char *vfhe_hip_device_name(void);

/* Tunables, also readable from the environment at first use (VFHE_HIP_NTT_MODE,
   VFHE_HIP_NTT_CHUNK_MB, VFHE_HIP_NTT_TILE_GRID, VFHE_HIP_NTT_COL_WIDE, VFHE_HIP_XCD_LAG,
   VFHE_HIP_XCD_OCC, VFHE_HIP_KS_FUSED, VFHE_HIP_FFT_FMA):
     "ntt.mode"          2: two launches per direction of the pipelined pass kernel (default);
                         0: two launches of the plain pass kernel; 1: persistent per-XCD kernel
     "ntt.chunk_mb"      two-launch modes: MiB of rows per sub-batch, 0 = whole batch
     "ntt.tile_grid"     mode 2: passes launched one tile per workgroup instead of persistent,
                         bit 0 forward column, 1 forward row, 2 inverse row, 3 inverse column,
                         4 the fused product's row kernel (default 15: the four passes)
     "ntt.col_wide"      mode 2: 1 = column passes in 512-thread workgroups over twice the
                         columns per tile when that leaves at least two workgroups per CU
                         (default), 0 = always 256 threads
     "ntt.xcd_lag"       persistent mode: queue slots between the two passes of a transform
     "ntt.xcd_occupancy" persistent mode: workgroups per CU, 2 or 4
     "ntt.spin_limit"    persistent mode: bound on a wait; exceeding it aborts at the next sync
     "ks.fused"          1: key-switch forward row pass fused with the key inner product (default)
     "fft.fma"           1: CKKS encoder FFT with the avx512ifma engine's fused complex product
                         (default), 0: the portable engine's
   get also answers "ntt.xcd_available" (1 when the placement probe found XCC ids 0..7)
   and "probe.xcd0".."probe.xcd7" (workgroups of the probe placed on each XCC).
   set returns 0, or -1 for an unknown key; get returns -1 for an unknown key. */
int vfhe_hip_set_option(const char *key, long value);
long vfhe_hip_get_option(const char *key);

/* Raw ChaCha20 keystream blocks (16 words each) for (key[8], nonce, counters
   ctr0 .. ctr0+nblocks-1), computed on the device: the sampler's known-answer
   check against RFC 8439. */
void vfhe_hip_chacha20_blocks(const uint32_t *key, uint64_t nonce, uint64_t ctr0, uint64_t nblocks, uint32_t *out);

/* NTT schedule autotuner: times round trips of `rows` scratch rows of `ntt`'s
   transform size under each candidate schedule (two launches of the pipelined
   pass kernel: every pass one tile per workgroup with 512-thread column passes,
   row passes one tile per workgroup, every pass persistent; of the plain one over
   the whole batch or 256 / 128 MiB sub-batches; the persistent per-XCD kernel with queue lags 3, 5, 8 where
   available), keeps the fastest in the options above and returns its index;
   us_out (9 entries, may be NULL) receives each candidate's microseconds per
   round trip, -1 when not applicable. */
long vfhe_hip_autotune_ntt(incNTT ntt, uint64_t rows, int reps, double *us_out);
/* Measured device-wide operation rates (operations per second): [0] FP64 FMA, [1] 32-bit
   multiply-low + add, [2] 32-bit multiply-high, [3] 32x32->64 multiply-add, [4] FP64 modular
   products (q < 2^50), [5] 64-bit Shoup modular products. */
void vfhe_hip_probe_alu_rates(double *rates);
/* The same for the fused product (vfhe_hip_ntt_mul_intt_batch): its row kernel one tile per
   workgroup or persistent x its column passes at 256 / 512-by-rule / 512 threads, 6
   candidates timed on the caller's operands (z is overwritten); the fastest is adopted and
   its index returned, us_out[k] (if not null) the microseconds per call of candidate k. */
long vfhe_hip_autotune_product(RNS_Polynomial *z, RNS_Polynomial *x, RNS_Polynomial *y, uint64_t count, int reps,
                               double *us_out);
/* The same for the CKKS multiply batch (vfhe_hip_mlwe_multiply_batch): column-pass width x
   inverse passes one tile per workgroup or persistent x the pass-through components folded
   into the key inner product or added after the ModDown, 12 candidates timed on the
   caller's operands (out is overwritten).  Every candidate is bit-exact. */
long vfhe_hip_autotune_multiply(RNS_MLWE *out, RNS_MLWE *in1, RNS_MLWE *in2, uint64_t count, RNS_MLWE **rlk,
                                uint64_t divide_mask, int reps, double *us_out);

/* Launch profiler: between start and stop every kernel launch is bracketed by
   HIP events on the engine stream.  stop() waits, then writes the summed
   milliseconds and launch counts of each kernel family into ms[6] and
   launches[6]: 0 NTT pass (FP64 path), 1 NTT pass (integer path),
   2 elementwise, 3 sequential division, 4 key-switch accumulation, 5 other. */
void vfhe_hip_profile_start(void);
void vfhe_hip_profile_stop(double *ms, uint64_t *launches);

/* Batched hot path: one launch sequence for many polynomials.
   direction 0 = forward (RNSc -> RNS), 1 = inverse. polys[i] is transformed in place. */
void vfhe_hip_ntt_batch(RNS_Polynomial *polys, uint64_t count, int direction);
/* vfhe_hip_ntt_batch through a HIP graph: the launches of a batch shape (ring, count,
   direction, every polynomial's limb mask and rows, lane, schedule options) are captured
   on its first call and replayed afterwards with the polynomials' work lists rebuilt.
   Results equal vfhe_hip_ntt_batch's. */
void vfhe_hip_ntt_batch_graph(RNS_Polynomial *polys, uint64_t count, int direction);
/* z[i] = INTT(NTT(x[i]) * NTT(y[i])) on the active limbs of x[i] & y[i]
   (coefficient form in, coefficient form out); x and y are left unchanged. */
void vfhe_hip_ntt_mul_intt_batch(RNS_Polynomial *z, RNS_Polynomial *x, RNS_Polynomial *y, uint64_t count);
/* vfhe_hip_ntt_mul_intt_batch through a HIP graph: the first call of a batch shape (ring,
   count, every operand's limb mask and rows, lane, schedule options) runs eagerly and
   captures its launches; later calls of the shape rebuild the operands' work lists and
   replay the graph (one host launch).  Results equal the eager entry point's. */
void vfhe_hip_ntt_mul_intt_batch_graph(RNS_Polynomial *z, RNS_Polynomial *x, RNS_Polynomial *y, uint64_t count);
/* CKKS product for `count` ciphertext pairs sharing one relinearization key:
   out[i] = rescale(relinearize(in1[i] * in2[i])) when divide_mask != 0 (outputs
   in coefficient form on the rescaled mask), or mlwe_multiply semantics (NTT form)
   when divide_mask == 0. */
void vfhe_hip_mlwe_multiply_batch(RNS_MLWE *out, RNS_MLWE *in1, RNS_MLWE *in2, uint64_t count,
                                  RNS_MLWE **rlk, uint64_t divide_mask);
/* Graph form of vfhe_hip_mlwe_multiply_batch (same arguments, same results bit for
   bit).  The first call of a batch shape (ring, count, rank, every operand's and key
   entry's limb mask, divide mask, lane, schedule options) runs eagerly and captures
   the batch's launch sequence -- tensor product, INTT, ModUp column pass, key-switch
   row pass + inner product, INTT, ModDown + rescale -- into a HIP graph owned by the
   engine; later calls of that shape rebuild only the work lists for their operands
   and replay the graph (one launch, no device allocation: the temporaries are the
   plan's own).  Operands written from the host must be marked dirty as usual; they
   are uploaded before the replay.  vfhe_hip_graph_clear frees every plan. */
void vfhe_hip_mlwe_multiply_batch_graph(RNS_MLWE *out, RNS_MLWE *in1, RNS_MLWE *in2, uint64_t count,
                                        RNS_MLWE **rlk, uint64_t divide_mask);
void vfhe_hip_graph_clear(void);

/* mlwe_RNSc_GHS_hybrid_keyswitch (mlwe.c:557-598) of count coefficient-form inputs that share one
   key, one launch sequence: out[k] is in[k] switched by ksk (outputs allocated on the extended ring,
   left on the level limbs in coefficient form). */
void vfhe_hip_mlwe_keyswitch_batch(RNS_MLWE *out, RNS_MLWE *in, uint64_t count, RNS_MLWE **ksk);
/* No counterpart in the reference: the allocation-free form of vfhe_hip_ntt_batch, for
   callers that cannot allocate (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc).  Same transforms and results; the persistent
   per-XCD kernel's work list and queue counters (and any other path's work lists)
   live in a per-device scratch array of 1 MiB (VH_SCRATCH_BYTES) held for the life
   of the process, so no device or host memory is allocated.  Returns 0, or -1 with
   nothing done when the batch needs more than the scratch (about 18,000 limb rows).
   Calls that overlap on one device (several lanes or threads) share the scratch and
   race: order them. */
int vfhe_hip_ntt_batch_noalloc(RNS_Polynomial *polys, uint64_t count, int direction);
/* mgsw_CMUX for `count` independent (out, in1, in2, mgsw) quadruples in one
   launch sequence (a batch of blind rotations stepping through one key index). */
void vfhe_hip_mgsw_cmux_batch(RNS_MLWE *out, RNSc_MLWE *in1, RNSc_MLWE *in2, RNS_MLWE **mgsw, uint64_t count,
                              uint64_t ell, uint64_t special_primes);
/* CGGI16 blind rotation of `count` accumulators at once (cggi16.py:70-90): acc[k] enters
   as X^-rot_b * tv in coefficient form; for i = 0..n-1 every accumulator with a non-zero
   exponent a = steps[k*n + i] becomes CMUX(acc, X^a * acc, bk[i]) (bk[i] the (r+1)*ell
   NTT-form samples of one MGSW encryption), then back to coefficient form.  Results in acc. */
void vfhe_hip_blind_rotate_batch(RNS_MLWE *acc, uint64_t count, RNS_MLWE **bk, uint64_t n, const uint64_t *steps,
                                 uint64_t ell, uint64_t special_primes);
/* The same, each accumulator starting from one test vector (cggi16.py:70-76): acc[k] =
   X^rot_tv[k] * tv (mlwe_RNSc_mul_by_xai; tv in coefficient form), then the blind rotation.
   At N = 1024 on FP64 limbs with one special prime, every key index runs in one launch
   with each accumulator resident in its workgroup's LDS. */
void vfhe_hip_blind_rotate_tv_batch(RNS_MLWE *acc, uint64_t count, RNS_MLWE tv, const uint64_t *rot_tv,
                                    RNS_MLWE **bk, uint64_t n, const uint64_t *steps, uint64_t ell,
                                    uint64_t special_primes);
/* mlwe_extract_LWE (mlwe.c:40-66, mlwe.cdef:87) for `count` coefficient-form samples of one
   ring, rank and mask: out[k] = the LWE sample of coefficient idx of in[k] (free each with
   free_lwe_sample); the rows are gathered on the device and read back with one copy. */
void vfhe_hip_extract_lwe_batch(RNSc_MLWE *in, uint64_t count, uint64_t idx, LWE *out);
/* The rotation amounts of a CGGI16 blind rotation for `count` LWE samples, replacing the
   per-sample CRT and rounding of cggi16.py:54-86 (host arithmetic): rot_tv[k] = 2N - rot_b
   mod 2N, steps[k*n + i] = round(a_i * 2N / q) mod 2N, q the product of the sample's primes.
   The reference rounds a double quotient; an entry whose exact quotient lies within that
   double's rounding error of a half is set to UINT64_MAX for the caller to recompute.
   Returns the number of such entries, or -1 when q * 2N does not fit 127 bits. */
int64_t vfhe_hip_cggi_rotations(LWE *in, uint64_t count, uint64_t N, uint64_t torus_base, uint64_t *rot_tv,
                                uint64_t *steps);
/* mlwe_automorphism_RNSc_GHS (mlwe.c:314-326) for `count` inputs sharing gen and ksk. */
void vfhe_hip_mlwe_automorphism_batch(RNS_MLWE *out, RNS_MLWE *in, uint64_t count, uint64_t gen, RNS_MLWE **ksk);
/* The same two batches with a Galois element / key (and rescale mask) per job:
   jobs may differ in level and key, and share the ring and the rank. */
void vfhe_hip_mlwe_multiply_batch_multi(RNS_MLWE *out, RNS_MLWE *in1, RNS_MLWE *in2, uint64_t count,
                                        RNS_MLWE ***rlks, const uint64_t *divide_masks);
void vfhe_hip_mlwe_automorphism_batch_multi(RNS_MLWE *out, RNS_MLWE *in, uint64_t count, const uint64_t *gens,
                                            RNS_MLWE ***ksks);
/* Views of `count` samples restricted to `mask`: new handles sharing the device
   rows and host mirror of in[k] (no copy).  A key stored once for the top level
   serves every lower level through a view with mask = level limbs | special
   limbs.  Free with free_RNS_mlwe_array before the viewed samples. */
RNS_MLWE *vfhe_hip_mlwe_view_array(RNS_MLWE *in, uint64_t count, uint64_t mask);
/* mlwe_alloc_RNS_sample without clearing the device rows (the reference's own
   allocation leaves them uninitialised too): for outputs that are written
   before they are read.  The host rows behind coeffs[] are created by the
   first polynomial_sync_to_host of a component (until then coeffs[i] is
   NULL), so an output that is only passed on to other engine calls never
   costs a host block. */
RNS_MLWE vfhe_hip_mlwe_alloc_RNS_sample_uninit(uint64_t N, uint64_t r, uint64_t mask, incNTT ntt);
/* The same for `count` samples in one call: out[k] receives each (free each with
   free_mlwe_RNS_sample or all with vfhe_hip_free_mlwe_samples). */
void vfhe_hip_mlwe_alloc_RNS_samples_uninit(RNS_MLWE *out, uint64_t count, uint64_t N, uint64_t r, uint64_t mask,
                                            incNTT ntt);
/* The calling thread's current lane stream (a hipStream_t). */
void *vfhe_hip_current_stream(void);

/* ------------------------------------------------------------------------ */
/* Ciphertext and key format (no counterpart in the reference, which keeps    */
/* keys and ciphertexts in memory only)                                      */
/* ------------------------------------------------------------------------ */
/* `count` samples of one ring, rank and limb mask as one little-endian block:
   "VFHEHIP1", version 1, form (0 unstated, 1 NTT, 2 coefficient), count, N, rank,
   split degree, limb mask, the mask's primes, the rows (sample, component, limb), and a
   64-bit check word over everything before it.  Size in bytes of that block: */
uint64_t vfhe_hip_mlwe_serialized_size(RNS_MLWE *in, uint64_t count);
/* Writes the block into buf (8-byte aligned, size >= the size above): the rows are
   gathered on the device and read back with one copy.  Returns the bytes written, or -1
   (samples of different rings, ranks or masks; buffer too small or unaligned). */
int64_t vfhe_hip_mlwe_serialize(RNS_MLWE *in, uint64_t count, uint32_t form, void *buf, uint64_t size);
/* Reads a block onto the ring `ntt`: allocates out[0 .. count-1] (free each with
   free_mlwe_RNS_sample), uploads the rows with one copy, sets *form (if not null) and
   returns count; or -1 (not such a block, other version, check word wrong), -2 (other
   ring: N, split degree or a prime differs), -3 (more than max_count samples), -4
   (truncated or unaligned).  Nothing is allocated on an error. */
int64_t vfhe_hip_mlwe_deserialize(const void *buf, uint64_t size, incNTT ntt, RNS_MLWE *out, uint64_t max_count,
                                  uint32_t *form);

/* ------------------------------------------------------------------------ */
/* Serving runtime: continuous batching of homomorphic requests              */
/* ------------------------------------------------------------------------ */
#define VFHE_HIP_REQ_MULTIPLY 0      /* in1*in2, relinearize with ksk (rlk), rescale by divide_mask if != 0 */
#define VFHE_HIP_REQ_AUTOMORPHISM 1  /* X -> X^gen on in1, key-switch with ksk (Galois key) */
/* A server with a worker thread.  Requests are admitted while the device bytes
   they hold in flight (output, temporaries, ModUp buffer) fit memory_budget
   (the oldest request is admitted alone when nothing is in flight); requests
   of one class (op, ring, rank; levels, keys, gens and rescale masks may
   differ) are batched, up to max_batch per launch, and batches run on `lanes`
   (1..7) engine lanes; lanes beyond the fourth take a batch only while at most 8
   requests wait (heavy load grows the batches on the first four instead). */
void *vfhe_hip_server_new(uint64_t memory_budget, uint32_t lanes, uint32_t max_batch);
/* Returns every device block parked in the engine's block cache to the allocator (each
   freed in order after its lane's last use) and the bytes released.  The cache keeps at
   most the option `cache.max_mb` (default 32 GiB), evicting the least recently parked
   blocks first, so blocks of sizes an earlier workload used give way to the current
   one's; this call empties it at once and trims the device's stream-ordered pool (it waits
   for every lane), e.g. before a long-lived request mix. */
uint64_t vfhe_hip_release_cached(void);
/* Thread-safe; returns the request id (>= 1), or 0 for an invalid request.
   Inputs and keys must stay alive and unchanged until the request completes.
   Inputs may still be being written by work the caller enqueued on its current
   lane: the request's batch waits for that lane's work submitted before this call. */
uint64_t vfhe_hip_server_submit(void *srv, int op, RNS_MLWE in1, RNS_MLWE in2, RNS_MLWE **ksk, uint64_t gen,
                                uint64_t divide_mask);
/* Up to `max` completed requests: id, output sample (the caller owns it) and
   latency from submit to completion in seconds.  wait != 0 blocks until one
   completes or none is pending.  Returns the number written. */
uint64_t vfhe_hip_server_poll(void *srv, uint64_t *ids, RNS_MLWE *outs, double *latency_s, uint64_t max, int wait);
/* vfhe_hip_server_poll plus each request's latency split in three (parts[3k..3k+2], seconds):
   submit -> admitted into a batch, admitted -> its batch enqueued, enqueued -> completion seen. */
uint64_t vfhe_hip_server_poll_timed(void *srv, uint64_t *ids, RNS_MLWE *outs, double *latency_s, double *parts,
                                    uint64_t max, int wait);
/* stats[10]: batches, requests launched, mean batch size, peak in-flight bytes,
   most lanes busy at once, requests pending, seconds the worker spent
   enqueueing batches, summed device seconds of the batches (each from its
   lane's events), seconds since the server started, longest single batch enqueue (s). */
void vfhe_hip_server_stats(void *srv, double *stats);
/* Drains every submitted request, then stops the worker. */
void vfhe_hip_server_free(void *srv);

/* ---- CGGI over the 32-bit torus with an FP64 FFT ----
   No counterpart in the reference, whose CGGI16 (modules/fhe/python/src/vfhe/fhe/cggi16.py:
   46-140) bootstraps over RNS primes with the NTT.  The same functional bootstrap (rotation
   amounts of cggi16.py:54-86 with q = 2^32, the test vector rotated by -rot_b, one CMUX per
   key index with a non-zero amount, extraction of coefficient 0) on the discretised torus
   Z/2^32: LWE samples of dimension n, one GLWE ring of N = 1024 (rank 1), signed gadget
   decomposition in base 2^base_log with ell levels, external products through a negacyclic
   FP64 FFT of N/2 points.  Torus values are uint32; samples are stored a_0 .. a_(dim-2), b
   with phase b - <a, s> mod 2^32.  Entry points return NULL / a negative code on bad
   arguments instead of aborting. */
typedef struct vfhe_torus_key vfhe_torus_key;
typedef struct vfhe_torus_batch vfhe_torus_batch;
/* A device-resident batch of `count` rows of `dim` torus values (samples, accumulators
   or a test vector), copied from `samples` (count x dim) or zeroed when it is NULL. */
vfhe_torus_batch *vfhe_hip_torus_batch_new(const uint32_t *samples, uint64_t count, uint64_t dim);
/* Copies the batch to `samples` (count x dim) after the work queued on the calling lane. */
int vfhe_hip_torus_batch_read(const vfhe_torus_batch *b, uint32_t *samples);
void vfhe_hip_torus_batch_free(vfhe_torus_batch *b);
/* The bootstrapping key: n GGSW encryptions of the LWE key bits, bsk laid out
   [n][2 ell rows][2 components][N] (row c * ell + l encrypts s_i * 2^(32 - base_log (l+1))
   added to component c: 0 the mask, 1 the body), moved to the transform domain on the
   device.  N must be 1024; ell <= 4, base_log <= 16, ell * base_log <= 32. */
vfhe_torus_key *vfhe_hip_torus_key_new(const uint32_t *bsk, uint64_t n, uint64_t N, uint64_t ell, uint32_t base_log);
void vfhe_hip_torus_key_free(vfhe_torus_key *key);
/* Functional bootstrap of every sample of `in` (dim n + 1) with test vector `tv` (1 x N):
   out (count x (N + 1)) receives the extracted samples (under the GLWE key's coefficients),
   acc_out (count x 2N, may be NULL) the accumulators (mask then body).  torus_base as
   cggi16.py:60: b is offset by floor(2^32 / (4 torus_base)) before its rounding.  One
   launch, queued on the calling lane.  Returns 0, or -1 / -2 / -3 for bad arguments,
   shapes of `in` / `tv`, shapes of the outputs. */
int vfhe_hip_torus_bootstrap(vfhe_torus_batch *out, vfhe_torus_batch *acc_out, const vfhe_torus_key *key,
                             const vfhe_torus_batch *in, const vfhe_torus_batch *tv, uint64_t torus_base);
/* The negacyclic transform of `count` polynomials (N = 1024 coefficients read as signed
   32-bit integers): out[k] = p(zeta w^k), zeta = e^(i pi / N), w = e^(-2 pi i / (N/2)),
   k < N/2, as (re, im) pairs, natural order.  Returns 0 or -1. */
int vfhe_hip_torus_fft(const uint32_t *in, uint64_t count, uint64_t N, double *out);

#ifdef __cplusplus
}
#endif
#endif /* VFHE_HIP_H */
