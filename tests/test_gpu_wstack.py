"""W-stacking predict on the GPU against the oracle restatement (reduced
grid, FP64 to 1e-10) and, at the reference test's full geometry (theta 0.1,
lambda 20480: 4096^2 x 8 planes), against the direct sum over the quantised
sky within the method's error."""
import os
import re
import subprocess

import numpy as np
import pytest
import torch

import pyoracle as orc
import rgamd
from gpu_util import dev

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not rgamd.lib().rg_hdf5_available(), reason="libhdf5 not loadable")]

SZE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sze")


@pytest.fixture(scope="module")
def kernels():
    uv = rgamd.load_sep_kern(os.path.join(SZE, "sepkern_uv_transpose.hdf5"), True)
    w = rgamd.load_sep_kern(os.path.join(SZE, "sepkern_w_transpose.hdf5"), True)
    lm = rgamd.load_sep_kern(os.path.join(SZE, "sepkern_lm.hdf5"), False)
    n = rgamd.load_sep_kern(os.path.join(SZE, "sepkern_n.hdf5"), False)
    d = {k: dev(x.data) for k, x in dict(uv=uv, w=w, lm=lm, n=n).items()}
    desc = {k: rgamd.sep_kern_desc(d[k], x.size, x.oversampling)
            for k, x in dict(uv=uv, w=w, lm=lm, n=n).items()}
    return dict(host=dict(uv=uv, w=w, lm=lm, n=n), dev=d, desc=desc)


def card(theta):
    p = np.array([0.95, 0.95, -0.95, -0.95, 0.95, -0.95, -0.95, 0.95, 0.0, 0.5, 0.0, -0.5, 0.5, 0.0,
                  -0.5, 0.0]) * (theta / 2)
    return p[0::2].copy(), p[1::2].copy()


def uvw(rng, theta, lam, dw, nv):
    return (rng.uniform(-theta * lam / 2, theta * lam / 2, nv),
            rng.uniform(-theta * lam / 2, theta * lam / 2, nv), rng.uniform(-dw, dw, nv))


def params(k, theta, lam):
    h = k["host"]
    return rgamd.wstack_params(theta, lam, h["uv"].du, h["w"].dw, h["lm"].x0, h["uv"].size,
                               h["w"].size)


@pytest.mark.parametrize("lam,sky", [(640.0, "card"), (2560.0, "random"), (1280.0, "many")])
def test_predict_matches_oracle(ctx, kernels, lam, sky):
    """Skies of up to 64 sources transform their source rows only and sum
    the columns directly; larger ones (here 100 sources) run the 2-D FFT."""
    theta = 0.1
    h, d = kernels["host"], kernels["desc"]
    rng = np.random.default_rng(21)
    if sky == "card":
        l, m = card(theta)
        flux = None
    else:
        ns = 40 if sky == "random" else 100
        l = rng.uniform(-0.45 * theta, 0.45 * theta, ns)
        m = rng.uniform(-0.45 * theta, 0.45 * theta, ns)
        flux = rng.uniform(0.1, 1.0, ns)
    u, v, w = uvw(rng, theta, lam, h["w"].dw, 20000)
    stack, skipped = orc.wstack_build(theta, lam, l, m, flux, h["uv"].du, h["w"].dw, h["w"].size,
                                      h["lm"].x0, h["lm"].data, h["n"].data)
    g = int(np.floor(theta * lam))
    ref, nd = orc.wstack_degrid(stack, g, u, v, w, h["uv"].du, h["w"].dw, h["uv"].data,
                                h["uv"].size, h["uv"].oversampling, h["w"].data, h["w"].size,
                                h["w"].oversampling)
    assert skipped == 0 and nd == 0
    p = params(kernels, theta, lam)
    out = torch.empty(len(u), dtype=torch.complex128, device="cuda")
    og, wp, sk = ctx.wstack_build(p, dev(l), dev(m), None if flux is None else dev(flux), d["lm"],
                                  d["n"])
    assert (og, wp, sk) == (stack.shape[1], stack.shape[0], 0)
    gs = ctx.wstack_stack(og, wp)
    assert np.abs(gs - stack).max() < 1e-11 * np.abs(stack).max()
    ctx.wstack_degrid(p, dev(u), dev(v), dev(w), d["uv"], d["w"], out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max()
    st = ctx.stats()
    assert st.n_vis == len(u) and st.n_dropped == 0 and st.n_gridded == len(u)


def test_reference_geometry_matches_the_quantised_sum(ctx, kernels):
    """test/test_wstack.cpp:56-141 at its own geometry (4096^2 x 8 planes),
    1 M random visibilities; checked against the direct sum on 20 000 of
    them (the method's error is about 5e-4 per unit source)."""
    theta, lam = 0.1, 20480.0
    h, d = kernels["host"], kernels["desc"]
    l, m = card(theta)
    rng = np.random.default_rng(22)
    nv = 1_000_000
    u, v, w = uvw(rng, theta, lam, h["w"].dw, nv)
    out = torch.empty(nv, dtype=torch.complex128, device="cuda")
    p = params(kernels, theta, lam)
    ctx.wstack_predict(p, dev(l), dev(m), None, dev(u), dev(v), dev(w), d["uv"], d["w"], d["lm"],
                       d["n"], out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    sel = rng.choice(nv, 20000, replace=False)
    ref = orc.predict_quantized(l, m, None, theta, lam, u[sel], v[sel], w[sel])
    err = np.abs(got[sel] - ref)
    assert err.max() < 1e-3 * len(l), err.max()
    assert err.mean() < 4e-4 * len(l), err.mean()
    assert ctx.stats().n_dropped == 0


def test_outside_visibilities_and_points_are_counted(ctx, kernels):
    theta, lam = 0.1, 640.0
    d = kernels["desc"]
    p = params(kernels, theta, lam)
    l = np.array([0.0, 0.2, 0.01])
    m = np.array([0.0, 0.0, np.nan])
    og, wp, sk = ctx.wstack_build(p, dev(l), dev(m), None, d["lm"], d["n"])
    assert sk == 2
    u = np.array([0.0, 1e6, 0.0, np.nan, 1.0])
    v = np.zeros(5)
    w = np.array([0.0, 0.0, 1e4, 0.0, -3.0])
    out = torch.full((5,), 7.0, dtype=torch.complex128, device="cuda")
    ctx.wstack_degrid(p, dev(u), dev(v), dev(w), d["uv"], d["w"], out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert got[0] != 0 and got[4] != 0 and np.all(got[1:4] == 0)
    st = ctx.stats()
    assert st.n_dropped == 3 and st.n_gridded == 2


def test_invalid_parameters_are_rejected(ctx, kernels):
    d = kernels["desc"]
    p = params(kernels, 0.1, 640.0)
    p.aa_support_uv = 7
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(rgamd.RGError):
        ctx.wstack_build(p, e, e, None, d["lm"], d["n"])
    p = params(kernels, 0.1, 600.0)  # og = 120 is no power of two
    with pytest.raises(rgamd.RGError):
        ctx.wstack_build(p, e, e, None, d["lm"], d["n"])


def test_cli_reports_the_aggregate_error():
    """wstack.out (src/wstack/wstack.cpp) in predict mode with the reference's
    kernel files: random visibilities (seeded) and a fixed (pu, pv, pw)."""
    exe = os.path.join(os.path.dirname(SZE), "..", "..", "radio-gridders_amd", "wstack.out")
    base = [exe, "-theta=0.1", "-lambda=20480", "-mode=0", "-cuda=1",
            f"-sepkern_uv={SZE}/sepkern_uv_transpose.hdf5", f"-sepkern_w={SZE}/sepkern_w_transpose.hdf5",
            f"-sepkern_lm={SZE}/sepkern_lm.hdf5", f"-sepkern_n={SZE}/sepkern_n.hdf5"]
    r = subprocess.run(base + ["-npts=200000", "-random_points=1", "-seed=4"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    per = float(re.search(r"Aggregate Error / Npts: (\S+)", r.stdout).group(1))
    assert 0 < per < 8 * 4e-4, r.stdout
    r = subprocess.run(base + ["-npts=16", "-pu=100.3", "-pv=-250.7", "-pw=12.5"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    per = float(re.search(r"Aggregate Error / Npts: (\S+)", r.stdout).group(1))
    assert per < 8 * 1e-3, r.stdout


def test_other_supports_match_oracle(ctx, kernels):
    """Supports other than the reference's 8 x 4 (here 4 x 2 and 16 x 4, with
    synthetic positive kernel tables) run the generic degridding kernel."""
    theta, lam = 0.1, 1280.0
    h, d = kernels["host"], kernels["desc"]
    rng = np.random.default_rng(31)
    l, m = card(theta)
    for auv, aw in ((4, 2), (16, 4)):
        ov = 64
        cuv = rng.uniform(0.1, 1.0, auv * ov)
        cw = rng.uniform(0.1, 1.0, aw * ov)
        u, v, w = uvw(rng, theta, lam, h["w"].dw, 5000)
        stack, _ = orc.wstack_build(theta, lam, l, m, None, h["uv"].du, h["w"].dw, aw, h["lm"].x0,
                                    h["lm"].data, h["n"].data)
        g = int(np.floor(theta * lam))
        ref, nd = orc.wstack_degrid(stack, g, u, v, w, h["uv"].du, h["w"].dw, cuv, auv, ov, cw,
                                    aw, ov)
        p = rgamd.wstack_params(theta, lam, h["uv"].du, h["w"].dw, h["lm"].x0, auv, aw)
        ctx.wstack_build(p, dev(l), dev(m), None, d["lm"], d["n"])
        duv, dw_ = dev(cuv), dev(cw)
        out = torch.empty(len(u), dtype=torch.complex128, device="cuda")
        ctx.wstack_degrid(p, dev(u), dev(v), dev(w), rgamd.sep_kern_desc(duv, auv, ov),
                          rgamd.sep_kern_desc(dw_, aw, ov), out)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max(), (auv, aw)
        assert ctx.stats().n_dropped == nd


@pytest.mark.parametrize("layout", ["sparse", "concentrated"])
def test_sort_handles_sparse_and_concentrated_streams(ctx, kernels, layout):
    """The degridding sort at a 1024^2 stack (24 576 (tile, plane) keys):
    a few visibilities spread over the whole stack (chunks of the bucketed
    records span more keys than the LDS window and rank with device atomics),
    and one (u, v, w) repeated 50 000 times (every chunk reserves its share of
    the same key)."""
    theta, lam = 0.1, 5120.0
    h, d = kernels["host"], kernels["desc"]
    rng = np.random.default_rng(41)
    l, m = card(theta)
    g = int(np.floor(theta * lam))
    reach = (g - h["uv"].size // 2 - 1) * h["uv"].du
    if layout == "sparse":
        nv = 6000
        u = rng.uniform(-reach, reach, nv)
        v = rng.uniform(-reach, reach, nv)
        w = rng.uniform(-h["w"].dw, h["w"].dw, nv)
    else:
        nv = 50000
        u, v, w = np.full(nv, 123.4), np.full(nv, -77.7), np.full(nv, 12.5)
    stack, skipped = orc.wstack_build(theta, lam, l, m, None, h["uv"].du, h["w"].dw, h["w"].size,
                                      h["lm"].x0, h["lm"].data, h["n"].data)
    sel = slice(None) if layout == "sparse" else slice(0, 1)
    ref, nd = orc.wstack_degrid(stack, g, u[sel], v[sel], w[sel], h["uv"].du, h["w"].dw,
                                h["uv"].data, h["uv"].size, h["uv"].oversampling, h["w"].data,
                                h["w"].size, h["w"].oversampling)
    assert skipped == 0 and nd == 0
    p = params(kernels, theta, lam)
    ctx.wstack_build(p, dev(l), dev(m), None, d["lm"], d["n"])
    out = torch.empty(nv, dtype=torch.complex128, device="cuda")
    ctx.wstack_degrid(p, dev(u), dev(v), dev(w), d["uv"], d["w"], out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    if layout == "concentrated":
        ref = np.full(nv, ref[0])
    assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max()
    st = ctx.stats()
    assert st.n_vis == nv and st.n_dropped == 0 and st.n_gridded == nv


def test_empty_and_all_dropped_streams(ctx, kernels):
    """No visibilities, and visibilities that all leave the stack: the sort
    sees no records (empty buckets, no work items) and every output is 0."""
    theta, lam = 0.1, 640.0
    d = kernels["desc"]
    p = params(kernels, theta, lam)
    l, m = card(theta)
    ctx.wstack_build(p, dev(l), dev(m), None, d["lm"], d["n"])
    e = torch.zeros(0, dtype=torch.float64, device="cuda")
    out0 = torch.zeros(0, dtype=torch.complex128, device="cuda")
    ctx.wstack_degrid(p, e, e, e, d["uv"], d["w"], out0)
    torch.cuda.synchronize()
    assert ctx.stats().n_vis == 0 and ctx.stats().n_gridded == 0
    nv = 3000
    u = np.full(nv, 1e7)
    out = torch.full((nv,), 5.0, dtype=torch.complex128, device="cuda")
    ctx.wstack_degrid(p, dev(u), dev(np.zeros(nv)), dev(np.zeros(nv)), d["uv"], d["w"], out)
    torch.cuda.synchronize()
    assert torch.all(out == 0)
    st = ctx.stats()
    assert st.n_vis == nv and st.n_dropped == nv and st.n_gridded == 0
