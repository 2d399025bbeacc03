"""The CKKS encoder FFT of the reference (oracle/_ref complex builds), pinned on the CPU.

With the special roots of complex.py:91-106 (restated in vfhe_amd.complex),
FFT = CT_NR then bit reversal evaluates c(X) = sum_k c_k X^k at zeta^(5^j),
zeta = exp(2 pi i / 4n) (the CKKS canonical embedding), and IFFT = bit reversal,
GS_RN with the inverse roots, scaling by 1/n inverts it.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vfhe_amd.complex import gen_special_rous_hp


def tables(L, n):
    r = gen_special_rous_hp(2 * n)
    re, im = np.array([v.real for v in r]), np.array([v.imag for v in r])
    ire, iim = np.array([(1 / v).real for v in r]), np.array([(1 / v).imag for v in r])
    return (L.load_rous_CT(O.dptr(re), O.dptr(im), 2 * n), L.load_rous_GS(O.dptr(ire), O.dptr(iim), 2 * n))


@pytest.mark.parametrize("engine", O.ref_engines())
@pytest.mark.parametrize("n", [16, 64])
def test_reference_fft_closed_form_and_inverse(engine, n):
    L = O.ref_complex(engine)
    ct, gs = tables(L, n)
    rng = np.random.default_rng(n)
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = O.aligned_f64(2 * n)
    x[:n], x[n:] = c.real, c.imag
    L.CT_NR(O.dptr(x), ct, n)
    L.bit_reverse_array(O.dptr(x), n, int(np.log2(n)))
    got = x[:n] + 1j * x[n:]
    zeta = np.exp(2j * np.pi / (4 * n))
    exp = np.array([sum(c[k] * zeta ** (k * pow(5, j, 4 * n)) for k in range(n)) for j in range(n)])
    assert np.max(np.abs(got - exp)) < 1e-10 * n
    # a planted mistake (roots 5^j read as 3^j) fails the same check
    bad = np.array([sum(c[k] * zeta ** (k * pow(3, j, 4 * n)) for k in range(n)) for j in range(n)])
    assert np.max(np.abs(got - bad)) > 1e-3
    L.bit_reverse_array(O.dptr(x), n, int(np.log2(n)))
    L.GS_RN(O.dptr(x), gs, n)
    L.complex_poly_scale_double(O.dptr(x), 1.0 / n, n)
    assert np.max(np.abs((x[:n] + 1j * x[n:]) - c)) < 1e-12 * n


def test_reference_engines_agree_to_2_40():
    engines = O.ref_engines()
    if len(engines) < 2:
        pytest.skip("one reference engine on this host")
    n = 4096
    rng = np.random.default_rng(3)
    c = rng.standard_normal(2 * n)
    outs = []
    for e in engines:
        L = O.ref_complex(e)
        ct, _ = tables(L, n)
        x = O.aligned_f64(2 * n)
        x[:] = c
        L.CT_NR(O.dptr(x), ct, n)
        outs.append(x.copy())
    assert np.max(np.abs(outs[0] - outs[1])) <= 2 ** -40 * np.max(np.abs(outs[1]))
