"""The prolate spheroidal taper and the W-kernel generator (host code of librg)."""
import math

import numpy as np
import pytest
import rgamd
import scipy.special as sp


@pytest.mark.parametrize("c", [1.0, 4.0, 4 * math.pi, 20.0])
def test_pswf_matches_scipy(c):
    x = np.linspace(-0.99, 0.99, 23)
    ref = np.array([sp.pro_ang1(0, 0, c, xi)[0] for xi in x]) / sp.pro_ang1(0, 0, c, 0.0)[0]
    got = np.array([rgamd.pswf(c, xi) for xi in x])
    assert np.allclose(got, ref, rtol=1e-7, atol=1e-9)


def test_w0_kernel_is_real_even_and_normalised():
    S, R = 16, 8
    k = rgamd.make_wkernels(1, R, S, 0.1, 0.0, 1.0)[0]  # [R][R][S][S], w = 0
    assert np.abs(k.imag).max() < 1e-12 * np.abs(k.real).max()
    # xf = 0 taps sit at d = x - S/2: K(d) = K(-d)
    k00 = k[0, 0].real
    assert np.allclose(k00[S // 2 + 1:, S // 2], k00[S // 2 - 1:0:-1, S // 2], rtol=1e-9)
    # sum over integer taps ~ A(0)^2 = 1
    assert abs(k00.sum() - 1.0) < 1e-3


def test_w_sign_symmetry():
    """A(l) is real and even, so K_{-w}(d) = conj(K_w(-d)); with the tap offsets
    d = x - S/2 - xf/R this pairs xf = 0 taps (x, y) with (S - x, S - y)."""
    S, R = 16, 8
    k = rgamd.make_wkernels(3, R, S, 0.1, -30.0, 30.0)
    km, kp = k[0, 0, 0], k[2, 0, 0]
    for y in range(1, S):
        for x in range(1, S):
            assert abs(km[y, x] - np.conj(kp[S - y, S - x])) < 1e-12


def test_fp32_table_is_rounded_fp64_table():
    a = rgamd.make_wkernels(2, 4, 8, 0.1, -10.0, 10.0)
    b = rgamd.make_wkernels(2, 4, 8, 0.1, -10.0, 10.0, dtype=np.complex64)
    assert np.array_equal(a.astype(np.complex64), b)
