"""ctypes binding of the hip engine's C ABI (include/vfhe_hip.h).

This plays the role of the reference's cffi extension `_vfhe_native_<engine>`
(modules/misc/python/src/_vfhe_native.py:29-56): it loads the engine library
and declares every entry point with the argument types of the cdefs
(modules/arith/python/cdef/arith.cdef, modules/mlwe/python/cdef/mlwe.cdef).
Handles are passed as `void *`, as there.

The library must have been built (`__graft_entry__.build()` or
`make -C vfhe_amd/csrc`); loading fails loudly otherwise, there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvfhe_hip.so")

U64 = ctypes.c_uint64
U64P = ctypes.POINTER(ctypes.c_uint64)
VP = ctypes.c_void_p
VPP = ctypes.POINTER(ctypes.c_void_p)
DP = ctypes.POINTER(ctypes.c_double)
I32 = ctypes.c_int


class IncNTT(ctypes.Structure):
    """struct _incNTT (modules/arith/python/cdef/arith.cdef:10-16)."""
    _fields_ = [("ntt", VPP), ("split_degree", U64), ("w", ctypes.POINTER(U64P)), ("N", U64), ("l", U64)]


class RNSPolynomial(ctypes.Structure):
    """struct _RNS_Polynomial (modules/arith/python/cdef/arith.cdef:18-23)."""
    _fields_ = [("coeffs", ctypes.POINTER(U64P)), ("ntt", ctypes.POINTER(IncNTT)), ("rns_mask", U64),
                ("allocated_l", U64)]


class RNSMLWE(ctypes.Structure):
    """struct _RNS_MLWE (modules/mlwe/python/cdef/mlwe.cdef:27-31)."""
    _fields_ = [("a", VPP), ("b", VP), ("r", U64)]


class IntPolynomial(ctypes.Structure):
    """struct _IntPolynomial (modules/arith/c/include/arith.h:96-100)."""
    _fields_ = [("coeffs", U64P), ("N", U64)]


class RNSMLWEKey(ctypes.Structure):
    """struct _RNS_MLWE_Key (modules/mlwe/python/cdef/mlwe.cdef:33-40)."""
    _fields_ = [("s", ctypes.POINTER(ctypes.POINTER(IntPolynomial))), ("s_RNS", VPP), ("N", U64), ("l", U64),
                ("r", U64), ("sigma", ctypes.c_double)]


# name: (restype, argtypes)
_SIGNATURES = {
    # number theory / misc
    "power_mod": (U64, [U64, U64, U64]),
    "inverse_mod": (U64, [U64, U64]),
    "inverse_mod_eea": (U64, [U64, U64]),
    "is_prime": (ctypes.c_bool, [U64]),
    "generate_Nth_root_of_unity": (U64, [U64, U64]),
    "next_special_prime": (U64, [U64, U64, ctypes.c_bool]),
    "bit_rev": (None, [U64P, U64P, U64, U64]),
    "vfhe_engine_active": (ctypes.c_char_p, []),
    "safe_aligned_malloc": (VP, [U64]),
    "vfhe_cpu_supports": (I32, [ctypes.c_char_p]),
    # single-limb kernels
    "ntt_new_proc": (VP, [U64, U64]),
    "ntt_free_proc": (None, [VP]),
    "ntt_forward": (None, [VP, VP, VP]),
    "ntt_reverse": (None, [VP, VP, VP]),
    "new_ntt_list": (VPP, [U64P, U64, U64]),
    "add_modq": (U64, [U64, U64, U64]),
    "sub_modq": (U64, [U64, U64, U64]),
    "negate_modq": (U64, [U64, U64]),
    "mul_modq": (U64, [U64, U64, VP]),
    "mod_eltwise_mul": (None, [VP, VP, VP, U64, VP]),
    "mod_eltwise_mul_addto": (None, [VP, VP, VP, U64, VP]),
    "mod_eltwise_mul_subto": (None, [VP, VP, VP, U64, VP]),
    "mod_eltwise_scale": (None, [VP, VP, U64, U64, VP]),
    "mod_eltwise_fma": (None, [VP, VP, U64, U64, VP]),
    "mod_eltwise_add_scalar": (None, [VP, VP, U64, U64, VP]),
    "mod_eltwise_sub_scalar": (None, [VP, VP, U64, U64, VP]),
    "mod_eltwise_negate": (None, [VP, VP, U64, VP]),
    "mod_eltwise_add": (None, [VP, VP, VP, U64, VP]),
    "mod_eltwise_sub": (None, [VP, VP, VP, U64, VP]),
    "mod_eltwise_reduce": (None, [VP, VP, U64, VP]),
    "mod_eltwise_reduce_signed": (None, [VP, VP, U64, VP]),
    "mod_reduce_array_mp": (None, [VP, VP, VP, U64, VP]),
    # pool and polynomials
    "new_incomplete_ntt_list": (VP, [U64P, U64, U64, U64]),
    "incNTT_extend_with_primes": (None, [VP, U64P, U64]),
    "incNTT_get_rou_matrix": (ctypes.POINTER(U64P), [VP]),
    "polynomial_new_RNS_polynomial": (VP, [U64, U64, VP]),
    "polynomial_new_RNS_polynomial_array": (VPP, [U64, U64, U64, VP]),
    "polynomial_new_array_of_RNS_polynomials": (VPP, [U64, U64, U64, VP]),
    "free_RNS_polynomial": (None, [VP]),
    "free_RNS_polynomial_array": (None, [U64, VPP]),
    "polynomial_copy_RNS_polynomial": (None, [VP, VP]),
    "polynomial_copy_RNSc_polynomial": (None, [VP, VP]),
    "polynomial_RNS_zero": (None, [VP]),
    "polynomial_eq": (ctypes.c_bool, [VP, VP]),
    "int_array_to_RNS": (None, [VP, U64P]),
    "array_to_RNS": (None, [VP, ctypes.POINTER(U64P)]),
    "polynomial_RNSc_to_RNS": (None, [VP, VP]),
    "polynomial_RNS_to_RNSc": (None, [VP, VP]),
    "polynomial_mul_RNS_polynomial": (None, [VP, VP, VP]),
    "polynomial_mul_addto_RNS_polynomial": (None, [VP, VP, VP]),
    "polynomial_mul_subto_RNS_polynomial": (None, [VP, VP, VP]),
    "polynomial_multo_RNS_polynomial": (None, [VP, VP]),
    "polynomial_add_RNS_polynomial": (None, [VP, VP, VP]),
    "polynomial_add_RNSc_polynomial": (None, [VP, VP, VP]),
    "polynomial_sub_RNS_polynomial": (None, [VP, VP, VP]),
    "polynomial_sub_RNSc_polynomial": (None, [VP, VP, VP]),
    "polynomial_RNSc_negate": (None, [VP, VP]),
    "polynomial_RNS_negate": (None, [VP, VP]),
    "polynomial_scale_RNSc_polynomial": (None, [VP, VP, U64]),
    "polynomial_scale_RNS_polynomial": (None, [VP, VP, U64]),
    "polynomial_scale_RNS_polynomial_RNS": (None, [VP, VP, U64P]),
    "polynomial_scale_addto_RNSc_polynomial": (None, [VP, VP, U64]),
    "polynomial_scale_addto_RNS_polynomial": (None, [VP, VP, U64]),
    "polynomial_RNSc_add_integer": (None, [VP, VP, U64]),
    "polynomial_RNS_add_integer": (None, [VP, VP, U64]),
    "polynomial_RNSc_mod_reduce_lifted": (None, [VP, VP, U64]),
    "polynomial_RNSc_mod_reduce": (None, [VP, VP]),
    "polynomial_round_division_RNSc_wo_free": (None, [VP, U64]),
    "polynomial_floor_division_RNSc_wo_free": (None, [VP, U64]),
    "polynomial_round_division_RNSc": (None, [VP]),
    "polynomial_floor_division_RNSc": (None, [VP]),
    "init_base_conversion_params": (VP, [VP, U64, U64]),
    "free_base_conversion_params": (None, [VP]),
    "polynomial_base_conversion_RNSc": (None, [VP, VP, VP]),
    "rns_compute_scaling_factors": (None, [U64P, VP, U64, U64]),
    "polynomial_RNSc_scaled_lift": (None, [VP, VP, U64P]),
    "polynomial_RNSc_permute": (None, [VP, VP, U64]),
    # mlwe
    "mlwe_alloc_RNS_sample": (VP, [U64, U64, U64, VP]),
    "mlwe_alloc_RNSc_sample": (VP, [U64, U64, U64, VP]),
    "mlwe_alloc_RNS_sample_array": (VPP, [U64, U64, U64, U64, VP]),
    "mlwe_create_copy_array": (VPP, [VPP, U64]),
    "free_mlwe_RNS_sample": (None, [VP]),
    "free_RNS_mlwe_array": (None, [U64, VPP]),
    "vfhe_hip_free_mlwe_samples": (None, [VPP, U64]),
    "mlwe_copy_RNS_sample": (None, [VP, VP]),
    "mlwe_RNS_trivial_sample_of_zero": (None, [VP]),
    "mlwe_RNSc_to_RNS": (None, [VP, VP]),
    "mlwe_RNS_to_RNSc": (None, [VP, VP]),
    "mlwe_RNS_mul_by_poly": (None, [VP, VP, VP]),
    "mlwe_RNS_mul_addto_by_poly": (None, [VP, VP, VP]),
    "mlwe_RNS_mul_subto_by_poly": (None, [VP, VP, VP]),
    "mlwe_add_RNSc_sample": (None, [VP, VP, VP]),
    "mlwe_sub_RNSc_sample": (None, [VP, VP, VP]),
    "mlwe_add_RNSc_polynomial": (None, [VP, VP, VP]),
    "mlwe_sub_RNSc_polynomial": (None, [VP, VP, VP]),
    "mlwe_RNS_add_polynomial": (None, [VP, VP, VP]),
    "mlwe_RNS_sub_polynomial": (None, [VP, VP, VP]),
    "mlwe_scale_RNS_mlwe_RNS": (None, [VP, U64P]),
    "mlwe_round_division_RNSc": (None, [VP, U64]),
    "mlwe_extended_rank": (U64, [U64]),
    "mlwe_tensor_product": (None, [VPP, VP, VP]),
    "mlwe_RNSc_GHS_hybrid_keyswitch": (None, [VP, VP, VPP, U64]),
    "mlwe_multiply": (None, [VP, VP, VP, VPP]),
    "mlwe_automorphism_RNSc_GHS": (None, [VP, VP, U64, VPP, U64]),
    "mlwe_new_RNS_key_from_array": (VP, [U64P, U64, U64, U64, VP, ctypes.c_double]),
    "free_mlwe_RNS_key": (None, [VP]),
    "mlwe_RNS_phase": (None, [VP, VP, VP]),
    # engine extensions
    "polynomial_sync_to_host": (None, [VP]),
    "polynomial_mark_host_dirty": (None, [VP]),
    "polynomial_device_row": (U64P, [VP, U64]),
    "vfhe_hip_synchronize": (None, []),
    "vfhe_hip_set_lane": (ctypes.c_int, [ctypes.c_int]),
    "vfhe_hip_autotune_ntt": (ctypes.c_long, [VP, U64, I32, ctypes.POINTER(ctypes.c_double)]),
    "vfhe_hip_probe_alu_rates": (None, [ctypes.POINTER(ctypes.c_double)]),
    "vfhe_hip_mlwe_alloc_RNS_samples_uninit": (None, [VPP, U64, U64, U64, U64, VP]),
    "vfhe_hip_torus_batch_new": (VP, [VP, U64, U64]),
    "vfhe_hip_torus_batch_read": (I32, [VP, VP]),
    "vfhe_hip_torus_batch_free": (None, [VP]),
    "vfhe_hip_torus_key_new": (VP, [VP, U64, U64, U64, ctypes.c_uint32]),
    "vfhe_hip_torus_key_free": (None, [VP]),
    "vfhe_hip_torus_bootstrap": (I32, [VP, VP, VP, VP, VP, U64]),
    "vfhe_hip_torus_fft": (I32, [VP, U64, U64, VP]),
    "vfhe_hip_autotune_product": (ctypes.c_long, [VPP, VPP, VPP, U64, I32, ctypes.POINTER(ctypes.c_double)]),
    "vfhe_hip_autotune_multiply": (ctypes.c_long, [VPP, VPP, VPP, U64, VPP, U64, I32,
                                                   ctypes.POINTER(ctypes.c_double)]),
    "vfhe_prng_set_deterministic_seed": (None, [U64]),
    "vfhe_prng_clear_deterministic_seed": (None, []),
    "polynomial_gen_random_RNSc_polynomial": (None, [VP]),
    "polynomial_gen_gaussian_RNSc_polynomial": (None, [VP, ctypes.c_double]),
    "mlwe_new_RNS_gaussian_key": (VP, [U64, U64, U64, ctypes.c_double, VP, ctypes.c_double]),
    "mlwe_RNS_sample_of_zero": (None, [VP, VP]),
    "mlwe_RNSc_sample_of_zero": (None, [VP, VP]),
    "mlwe_new_RNS_sample_of_zero": (VP, [VP]),
    "mlwe_new_RNSc_sample_of_zero": (VP, [VP]),
    "mlwe_new_RNS_trivial_sample_of_zero": (VP, [U64, U64, U64, VP]),
    "mlwe_RNSc_sample": (None, [VP, VP, VP]),
    "vfhe_hip_chacha20_blocks": (None, [VP, U64, U64, U64, VP]),
    "vfhe_hip_mlwe_automorphism_batch": (None, [VPP, VPP, U64, U64, VPP]),
    "vfhe_hip_mlwe_keyswitch_batch": (None, [VPP, VPP, U64, VPP]),
    "vfhe_hip_blind_rotate_batch": (None, [VPP, U64, VPP, U64, U64P, U64, U64]),
    "vfhe_hip_cggi_rotations": (ctypes.c_int64, [VPP, U64, U64, U64, U64P, U64P]),
    "vfhe_hip_blind_rotate_tv_batch": (None, [VPP, U64, VP, U64P, VPP, U64, U64P, U64, U64]),
    "vfhe_hip_extract_lwe_batch": (None, [VPP, U64, U64, VPP]),
    "vfhe_hip_release_cached": (U64, []),
    "vfhe_hip_mlwe_serialized_size": (U64, [VPP, U64]),
    "vfhe_hip_mlwe_serialize": (ctypes.c_int64, [VPP, U64, ctypes.c_uint32, VP, U64]),
    "vfhe_hip_mlwe_deserialize": (ctypes.c_int64, [VP, U64, VP, VPP, U64, ctypes.POINTER(ctypes.c_uint32)]),
    "vfhe_hip_ntt_batch_noalloc": (I32, [VPP, U64, I32]),
    "vfhe_hip_mlwe_multiply_batch_multi": (None, [VPP, VPP, VPP, U64, VPP, U64P]),
    "polynomial_RNSc_mul_by_xai": (None, [VP, VP, U64]),
    "polynomial_RNSc_mul_by_xai_minus1": (None, [VP, VP, U64]),
    "mlwe_RNSc_mul_by_xai": (None, [VP, VP, U64]),
    "mlwe_RNSc_mul_by_xai_minus1": (None, [VP, VP, U64]),
    "mgsw_external_product": (None, [VP, VPP, VP, U64, U64]),
    "mgsw_CMUX": (None, [VP, VP, VP, VPP, U64, U64]),
    "mgsw_NCMUX": (None, [VP, VP, VP, VPP, VPP, U64, U64]),
    "vfhe_hip_mgsw_cmux_batch": (None, [VPP, VPP, VPP, VPP, U64, U64, U64]),
    "vfhe_hip_mlwe_automorphism_batch_multi": (None, [VPP, VPP, U64, U64P, VPP]),
    "vfhe_hip_mlwe_view_array": (VPP, [VPP, U64, U64]),
    "vfhe_hip_current_stream": (VP, []),
    "vfhe_hip_mlwe_alloc_RNS_sample_uninit": (VP, [U64, U64, U64, VP]),
    "vfhe_hip_server_new": (VP, [U64, ctypes.c_uint32, ctypes.c_uint32]),
    "vfhe_hip_server_submit": (U64, [VP, I32, VP, VP, VPP, U64, U64]),
    "vfhe_hip_server_poll": (U64, [VP, U64P, VPP, ctypes.POINTER(ctypes.c_double), U64, I32]),
    "vfhe_hip_server_poll_timed": (U64, [VP, U64P, VPP, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double), U64, I32]),
    "vfhe_hip_server_stats": (None, [VP, ctypes.POINTER(ctypes.c_double)]),
    "vfhe_hip_server_free": (None, [VP]),
    "load_rous_CT": (VP, [DP, DP, U64]),
    "load_rous_GS": (VP, [DP, DP, U64]),
    "CT_NR": (None, [DP, VP, U64]),
    "GS_RN": (None, [DP, VP, U64]),
    "bit_reverse_array": (None, [DP, U64, ctypes.c_uint32]),
    "complex_poly_scale_double": (None, [DP, ctypes.c_double, U64]),
    "complex_poly_round_to_RNS": (None, [VP, DP, U64]),
    "complex_polys_ifft_scale_round_to_RNS_batch": (None, [VPP, VPP, U64, U64, ctypes.c_uint32, VP, ctypes.c_double]),
    "vfhe_hip_last_error": (I32, []),
    "vfhe_hip_device_name": (ctypes.c_char_p, []),
    "vfhe_hip_set_option": (I32, [ctypes.c_char_p, ctypes.c_long]),
    "vfhe_hip_get_option": (ctypes.c_long, [ctypes.c_char_p]),
    "vfhe_hip_profile_start": (None, []),
    "vfhe_hip_profile_stop": (None, [ctypes.POINTER(ctypes.c_double), U64P]),
    # slots, LWE, extraction, trace and packing
    "polynomial_RNS_broadcast_slot": (None, [VP, VP, U64]),
    "polynomial_RNS_rotate_slot": (None, [VP, VP, U64]),
    "polynomial_RNS_copy_slot": (None, [VP, U64, VP, U64]),
    "polynomial_RNS_inverse": (I32, [VP, VP]),
    "mlwe_extract_LWE": (VP, [VP, U64]),
    "mlwe_partial_trace": (None, [VP, VP, U64P, VP, U64, U64]),
    "mlwe_trace": (None, [VP, VP, VP, U64]),
    "mlwe_full_packing_keyswitch": (None, [VP, VPP, U64, VPP, U64]),
    "mlwe_full_packing_keyswitch_scaled": (None, [VPP, U64, VP, U64]),
    "lwe_alloc_key": (VP, [U64, U64, VP]),
    "lwe_alloc_sample": (VP, [U64, U64, VP]),
    "free_lwe_sample": (None, [VP]),
    "free_lwe_key": (None, [VP]),
    "lwe_new_key": (VP, [U64, U64, VP, ctypes.c_double, ctypes.c_double]),
    "lwe_new_sparse_ternary_key": (VP, [U64, U64, VP, U64, ctypes.c_double]),
    "lwe_sample": (None, [VP, U64P, VP]),
    "lwe_new_sample": (VP, [U64P, VP]),
    "lwe_new_trivial_sample": (VP, [U64P, U64, U64, VP]),
    "lwe_phase": (None, [U64P, VP, VP]),
    "lwe_subto": (None, [VP, VP]),
    "vfhe_hip_ntt_batch": (None, [VPP, U64, I32]),
    "vfhe_hip_ntt_batch_graph": (None, [VPP, U64, I32]),
    "vfhe_hip_ntt_mul_intt_batch": (None, [VPP, VPP, VPP, U64]),
    "vfhe_hip_ntt_mul_intt_batch_graph": (None, [VPP, VPP, VPP, U64]),
    "vfhe_hip_mlwe_multiply_batch": (None, [VPP, VPP, VPP, U64, VPP, U64]),
    "vfhe_hip_mlwe_multiply_batch_graph": (None, [VPP, VPP, VPP, U64, VPP, U64]),
    "vfhe_hip_graph_clear": (None, []),
}

_lib = None


def symbols() -> list[str]:
    """Every entry point this binding declares (all are exported by the library)."""
    return list(_SIGNATURES)


def load():
    """Load the engine library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"hip engine not built: {LIB_PATH} is missing (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def voidp_array(items) -> ctypes.Array:
    """A `void *[]` of handles (ints, c_void_p, or nested ctypes arrays); None -> NULL."""
    arr = (VP * max(1, len(items)))()
    for i, x in enumerate(items):
        if isinstance(x, (ctypes.Array, ctypes._Pointer)):
            x = ctypes.cast(x, VP).value
        arr[i] = x
    return arr
