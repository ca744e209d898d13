"""W-stacking predict in the oracle (src/wstack/predict.cpp,
wstack_common.h): the stack of a single point source against its closed form,
and the whole predict with the reference's Sze-Tan kernels against the direct
sum over the quantised sky."""
import os

import numpy as np
import pytest

import pyoracle as orc
import rgamd

HERE = os.path.dirname(os.path.abspath(__file__))
SZE = os.path.join(HERE, "golden", "sze")
needs_hdf5 = pytest.mark.skipif(not rgamd.lib().rg_hdf5_available(), reason="libhdf5 not loadable")


def sze_kernels():
    uv = rgamd.load_sep_kern(os.path.join(SZE, "sepkern_uv_transpose.hdf5"), True)
    w = rgamd.load_sep_kern(os.path.join(SZE, "sepkern_w_transpose.hdf5"), True)
    lm = rgamd.load_sep_kern(os.path.join(SZE, "sepkern_lm.hdf5"), False)
    n = rgamd.load_sep_kern(os.path.join(SZE, "sepkern_n.hdf5"), False)
    return uv, w, lm, n


def card_points(theta):
    """generate_testcard_dataset (predict.cpp:152-164): 8 points at
    (+-0.95, +-0.95), (0, +-0.5), (+-0.5, 0) times theta / 2."""
    p = np.array([0.95, 0.95, -0.95, -0.95, 0.95, -0.95, -0.95, 0.95, 0.0, 0.5, 0.0, -0.5, 0.5, 0.0,
                  -0.5, 0.0]) * (theta / 2)
    return p[0::2].copy(), p[1::2].copy()


def test_geometry_matches_the_reference_formulas():
    # theta 0.1, lambda 20480: g 2048, og 4096, ceil(144.8 / 49.94) + 4 + 1 planes
    og, wp = orc.wstack_geometry(0.1, 20480.0, 49.93742167907413, 4, 0.25)
    assert og == 4096 and wp == 3 + 4 + 1


def test_single_point_stack_is_its_closed_form():
    """One source of flux f at pixel (row mc, column lc) with unit grid
    correction: plane k at (row r, column c) is
    f exp(-2 pi i ((c - og/2)(lco - og/2) + (r - og/2)(mco - og/2)) / og)
      x exp(2 pi i (t0 + k) dw (1 - sqrt(1 - l^2 - m^2))),
    lco = gd + lc, t0 = -(w_planes / 2), l = theta (lc - g/2) / g."""
    theta, lam, dw, x0, aa_w = 0.1, 160.0, 7.0, 0.25, 2
    g = int(np.floor(theta * lam))
    og, wp = orc.wstack_geometry(theta, lam, dw, aa_w, x0)
    gd = (og - g) // 2
    lc, mc, f = 11, 3, 0.75
    l = (lc / g - 0.5) * theta + 1e-9
    m = (mc / g - 0.5) * theta + 1e-9
    ones = np.ones(11)
    stack, dropped = orc.wstack_build(theta, lam, [l], [m], [f], 5.0, dw, aa_w, x0, ones, ones)
    assert dropped == 0 and stack.shape == (wp, og, og)
    lq, mq = theta * (lc - g // 2) / g, theta * (mc - g // 2) / g
    ph = dw * (1 - np.sqrt(1 - lq * lq - mq * mq))
    r, c = np.meshgrid(np.arange(og), np.arange(og), indexing="ij")
    base = np.exp(-2j * np.pi * ((c - og // 2) * (gd + lc - og // 2) + (r - og // 2) * (gd + mc - og // 2)) / og)
    for k in range(wp):
        ref = f * base * np.exp(2j * np.pi * (k - wp // 2) * ph)
        assert np.abs(stack[k] - ref).max() < 1e-11, k


def test_points_outside_the_sky_are_skipped():
    ones = np.ones(11)
    stack, dropped = orc.wstack_build(0.1, 160.0, [0.2, 0.0], [0.0, 0.0], None, 5.0, 7.0, 2, 0.25,
                                      ones, ones)
    assert dropped == 1
    assert np.abs(stack).max() > 0


def sky(kind, theta):
    if kind == "card":
        return card_points(theta)
    rng = np.random.default_rng(5)
    return rng.uniform(-0.45 * theta, 0.45 * theta, 6), rng.uniform(-0.45 * theta, 0.45 * theta, 6)


@needs_hdf5
@pytest.mark.parametrize("lam,kind", [(640.0, "card"), (2560.0, "card"), (2560.0, "random")])
def test_predict_with_sze_tan_kernels_matches_the_quantised_sum(lam, kind):
    """The reference's test (test/test_wstack.cpp:56-141) at a reduced grid:
    test-card (or random, asymmetric) sky, random u, v in
    [-theta lambda/2, theta lambda/2), w in [-dw, dw); the W-stack predict
    against the direct sum over the quantised points.  The method's own error
    is about 5e-4 per unit source (the kernel is taken at the oversampling
    step below the exact offset, 1/4096 of a cell); a wrong sign, axis or
    index gives errors of order one per source."""
    theta = 0.1
    uv, w, lm, n = sze_kernels()
    assert (uv.size, w.size, lm.size, n.size) == (8, 4, 10001, 10001)
    l, m = sky(kind, theta)
    stack, dropped = orc.wstack_build(theta, lam, l, m, None, uv.du, w.dw, w.size, lm.x0,
                                      lm.data, n.data)
    assert dropped == 0
    rng = np.random.default_rng(11)
    nv = 4000
    u = rng.uniform(-theta * lam / 2, theta * lam / 2, nv)
    v = rng.uniform(-theta * lam / 2, theta * lam / 2, nv)
    ww = rng.uniform(-w.dw, w.dw, nv)
    g = int(np.floor(theta * lam))
    got, nd = orc.wstack_degrid(stack, g, u, v, ww, uv.du, w.dw, uv.data, uv.size, uv.oversampling,
                                w.data, w.size, w.oversampling)
    assert nd == 0
    ref = orc.predict_quantized(l, m, None, theta, lam, u, v, ww)
    err = np.abs(got - ref)
    assert err.max() < 1e-3 * len(l), err.max()
    assert np.mean(err) < 4e-4 * len(l), np.mean(err)


@needs_hdf5
def test_degrid_skips_visibilities_outside_the_stack():
    theta, lam = 0.1, 640.0
    uv, w, lm, n = sze_kernels()
    l, m = card_points(theta)
    stack, _ = orc.wstack_build(theta, lam, l, m, None, uv.du, w.dw, w.size, lm.x0, lm.data, n.data)
    g = int(np.floor(theta * lam))
    u = np.array([0.0, 1e6, 0.0, np.nan])
    v = np.array([0.0, 0.0, 0.0, 0.0])
    ww = np.array([0.0, 0.0, 1e4, 0.0])
    got, nd = orc.wstack_degrid(stack, g, u, v, ww, uv.du, w.dw, uv.data, uv.size, uv.oversampling,
                                w.data, w.size, w.oversampling)
    assert nd == 3
    assert got[0] != 0 and np.all(got[1:] == 0)
