"""CGGI16 functional bootstrap on the hip engine through the host mirror.

Ports modules/fhe/python/test/test_cggi16.py:20-62 (same ring, keys, LUT and
message precision): bootstrap-key generation (MGSW encryptions of the LWE key),
LUT packing, blind rotation (mlwe_RNSc_mul_by_xai + mgsw_CMUX) and sample
extraction, checked by decrypting the extracted LWE sample and reading the
lookup table's entry for the message.  Keys and encryption randomness come from
the seeded host sampler, so the run is deterministic.  The batched blind
rotation (one batched CMUX per key index over several inputs) must give the
same ciphertexts as the one-at-a-time form, bit for bit.
"""
import random

import numpy as np
import pytest

from vfhe_amd import arith
from vfhe_amd.arith import Ring, repr
from vfhe_amd.cggi import CGGI16, _extract_lwe_key
from vfhe_amd.lwe import LWE, LWE_Key, extract_lwe
from vfhe_amd.mlwe import MLWE, MLWE_Scheme

pytestmark = pytest.mark.gpu

MSG_PREC = 5


def mod_switch(v, q, p):
    return round((v * p) / q) % p


@pytest.fixture(scope="module")
def setup():
    arith.NTT_processor_instance.reset()
    arith.set_seed(0xB007C0DE)
    rnd = random.Random(0xB007C0DE)
    Rq = Ring(256, prime_size=[50, 50, 50], split_degree=1)
    in_ring = Rq.quotient_ring(ell=2)
    in_scheme = MLWE_Scheme(in_ring, special_primes=0, module_rank=1)
    out_scheme = MLWE_Scheme(Rq, special_primes=1, module_rank=1, max_lvl=1)
    input_key = in_scheme.key_gen_sparse(17, 3.2, ternary=False)
    input_lwe_key = _extract_lwe_key(input_key)
    output_key = out_scheme.key_gen_sparse(64, 3.2, ternary=False)
    cg = CGGI16(out_scheme)
    print("[cggi] bootstrap key: 256 MGSW encryptions, N=256", flush=True)
    bk = cg.generate_bootstrap_key(input_lwe_key, output_key)
    lut_size = 1 << (MSG_PREC - 1)
    lut = [rnd.randint(0, (1 << (MSG_PREC - 1)) - 1) for _ in range(lut_size)]
    tv = cg.LUT_packing(lut, lut_size, MSG_PREC)
    out_lwe_key = _extract_lwe_key(output_key)
    out_key_in_ring = LWE_Key(ring=in_ring, key=out_lwe_key.get_s(), n=out_lwe_key.n)
    return dict(in_ring=in_ring, key=input_lwe_key, cg=cg, bk=bk, lut=lut, tv=tv, out_key=out_key_in_ring,
                out_scheme=out_scheme, output_key=output_key, rnd=rnd)


def encrypt_msg(s, msg_val):
    ring = s["in_ring"]
    m_scaled = mod_switch(msg_val, 1 << MSG_PREC, ring.q_l)
    return LWE(ring=ring, m=[m_scaled % q for q in ring.primes], key=s["key"])


def test_lwe_mirror_phase_and_extraction(setup):
    """lwe.c:95-156: phase(sample(m)) == m (error-free key); mlwe.c:40-66: the extracted sample's
    phase is coefficient idx of the MLWE phase."""
    s = setup
    ring = s["in_ring"]
    for v in (0, 5, 15):
        m = mod_switch(v, 1 << MSG_PREC, ring.q_l)
        c = LWE(ring=ring, m=[m % q for q in ring.primes], key=s["key"])
        assert c.phase(s["key"], recompose=True) == m % ring.q_l
    sch, key = s["out_scheme"], s["output_key"]
    msg = arith.Polynomial(sch.rings[0]).from_array([(7 * i) % 97 for i in range(256)])
    ct = sch.sample(msg, key)
    ph = sch.phase(ct, key)
    coeffs = ph.get_polynomial()
    lwe_key = _extract_lwe_key(key)
    k = LWE_Key(ring=sch.rings[0], key=lwe_key.get_s(), n=lwe_key.n)
    for idx in (0, 1, 100, 255):
        assert extract_lwe(ct, idx).phase(k, recompose=True) == coeffs[idx]


def test_functional_bootstrap(setup):
    """test_cggi16.py:20-62: the bootstrapped sample decrypts to lut[msg] for several messages."""
    s = setup
    cg, bk, lut, tv = s["cg"], s["bk"], s["lut"], s["tv"]
    msgs = [0, 3, 7, 12, 15]
    for v in msgs:
        print(f"[cggi] bootstrap msg={v}", flush=True)
        lw = encrypt_msg(s, v)
        out = cg.functional_bootstrap(None, tv, lw, bk, torus_base=1 << (MSG_PREC - 1))
        phase = out.phase(s["out_key"], recompose=True)
        assert mod_switch(phase, s["in_ring"].q_l, 1 << MSG_PREC) == lut[v], v


def test_blind_rotate_batch_matches_sequential(setup):
    """The batched blind rotation equals functional_bootstrap_wo_extract input by input."""
    s = setup
    cg, bk, tv = s["cg"], s["bk"], s["tv"]
    lwes = [encrypt_msg(s, v) for v in (1, 9, 14)]
    print("[cggi] batched blind rotation of 3 inputs", flush=True)
    got = cg.blind_rotate_batch([tv] * 3, lwes, bk, torus_base=1 << (MSG_PREC - 1))
    for lw, g in zip(lwes, got):
        ref = MLWE(cg.scheme)
        cg.functional_bootstrap_wo_extract(ref, tv, lw, bk, torus_base=1 << (MSG_PREC - 1))
        ref.to_coeff()
        g.to_coeff()
        for j in range(ref.r):
            assert ref.get_a_poly(j).get_coeff_matrix(repr.coeff) == g.get_a_poly(j).get_coeff_matrix(repr.coeff)
        assert ref.get_b_poly().get_coeff_matrix(repr.coeff) == g.get_b_poly().get_coeff_matrix(repr.coeff)
