"""File-level entry points (rg_wproj_flat, rg_wtowers_flat,
rg_image_dft_host_flat) and the grid.out / dft.out CLIs on HDF5 files in the
reference layouts, against the oracle."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest
import torch

import pyoracle as orc
import rgamd
import synth
from gpu_util import rel_max

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not rgamd.lib().rg_hdf5_available(), reason="libhdf5 not loadable")]

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "radio-gridders_amd")
THETA, LAM = 0.1, 1280.0
N = int(math.floor(THETA * LAM))


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("h5")
    pos = synth.layout_y(9, 450.0, np.random.default_rng(1))
    a1, a2, enu = synth.baselines(pos)
    T = 12
    uvw = synth.uvw_tracks(enu, synth.hour_angles(T, 4.0), math.radians(-30.0))
    freq = np.array([synth.FREQ_HZ])
    u, v, w = synth.flatten(uvw, freq)
    l, m, flux = synth.sources(3, np.random.default_rng(2))
    vis = synth.predict_numpy(u, v, w, l, m, flux).reshape(len(a1), T, 1)
    visfile = str(d / "vis.h5")
    rgamd.write_vis_hdf5(visfile, a1, a2, np.arange(T, dtype=float), freq, uvw, vis)
    # W-projection table over the data, W-towers table over [-5, 5]
    w_min, P = synth.kernel_range(w, 10.0)
    tab = rgamd.make_wkernels(P, 8, 16, THETA, w_min, 10.0)
    wkp = str(d / "wk_proj.h5")
    rgamd.write_wkern_hdf5(wkp, THETA, tab, w_min, 10.0)
    tabt = rgamd.make_wkernels(11, 8, 16, THETA, -5.0, 1.0)
    wkt = str(d / "wk_towers.h5")
    rgamd.write_wkern_hdf5(wkt, THETA, tabt, -5.0, 1.0)
    return dict(visfile=visfile, wkp=wkp, wkt=wkt, u=u, v=v, w=w, vis=vis.ravel(), tab=tab,
                w_min=w_min, tabt=tabt, dir=d)


def test_wproj_flat_matches_oracle(files):
    g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    h = np.empty((N, N), np.complex128)
    for mode in (rgamd.RG_MODE_PLAIN, rgamd.RG_MODE_BATCHED):
        rgamd.check(rgamd.lib().rg_set_file_mode(mode))
        rgamd.check(rgamd.lib().rg_wproj_flat(files["visfile"].encode(), files["wkp"].encode(),
                                              g.data_ptr(), h.ctypes.data, N, THETA, LAM, 0.0,
                                              1e300, 256))
        torch.cuda.synchronize()
        ref, _, bad = orc.grid_wprojection(N, THETA, files["u"], files["v"], files["w"],
                                           files["vis"], files["tab"], files["w_min"], 10.0)
        assert bad == 0
        assert rel_max(g.cpu().numpy(), orc.grid_to_image(ref)) < 1e-12
    rgamd.check(rgamd.lib().rg_set_file_mode(rgamd.RG_MODE_BATCHED))


def test_wtowers_flat_matches_oracle(files):
    g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    h = np.empty((N, N), np.complex128)
    rgamd.check(rgamd.lib().rg_wtowers_flat(files["visfile"].encode(), files["wkt"].encode(),
                                            g.data_ptr(), h.ctypes.data, N, THETA, LAM, 0.0, 1e300,
                                            64, 16, 10.0))
    torch.cuda.synchronize()
    ref, _, _ = orc.grid_wtowers(N, THETA, files["u"], files["v"], files["w"], files["vis"],
                                 files["tabt"], -5.0, 1.0, 64, 16, 10.0)
    assert rel_max(g.cpu().numpy(), orc.grid_to_image(ref)) < 1e-11


def test_image_dft_host_flat_matches_oracle(files):
    g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    h = np.empty((N, N), np.complex128)
    rgamd.check(rgamd.lib().rg_image_dft_host_flat(files["visfile"].encode(), h.ctypes.data,
                                                   g.data_ptr(), N, THETA, LAM, 0.0, 1e300,
                                                   N * N // 128, 128))
    ys, xs = np.divmod(np.arange(N * N), N)
    ref = orc.dft(xs, ys, N, LAM, files["u"], files["v"], files["w"], files["vis"]).reshape(N, N)
    assert rel_max(h, ref) < 1e-11
    bad = rgamd.lib().rg_image_dft_host_flat(files["visfile"].encode(), h.ctypes.data,
                                             g.data_ptr(), N, THETA, LAM, 0.0, 1e300, 3, 128)
    assert bad == rgamd.RG_ERR_PARAM


def test_missing_file_returns_error(files):
    g = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    st = rgamd.lib().rg_wproj_flat(b"/nonexistent.h5", files["wkp"].encode(), g.data_ptr(), None,
                                   N, THETA, LAM, 0.0, 1e300, 256)
    assert st == rgamd.RG_ERR_IO


def test_clis_write_the_real_image(files):
    d = files["dir"]
    img = str(d / "grid.img")
    r = subprocess.run([os.path.join(PKG, "grid.out"), f"-vis={files['visfile']}",
                        f"-wkernel={files['wkp']}", f"-theta={THETA}", f"-lambda={LAM}", "-wproj",
                        "-flat", f"-image={img}"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(img, dtype=np.float64).reshape(N, N)
    ref, _, _ = orc.grid_wprojection(N, THETA, files["u"], files["v"], files["w"], files["vis"],
                                     files["tab"], files["w_min"], 10.0)
    assert rel_max(got, orc.grid_to_image(ref).real) < 1e-12
    dimg = str(d / "dft.img")
    r = subprocess.run([os.path.join(PKG, "dft.out"), f"-vis={files['visfile']}",
                        f"-theta={THETA}", f"-lambda={LAM}", f"-blocks={N * N // 64}",
                        "-threadblock=64", "-flat", f"-image={dimg}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(dimg, dtype=np.float64).reshape(N, N)
    ys, xs = np.divmod(np.arange(N * N), N)
    ref = orc.dft(xs, ys, N, LAM, files["u"], files["v"], files["w"], files["vis"]).reshape(N, N)
    assert rel_max(got, ref.real) < 1e-11


@pytest.mark.parametrize("wproj", [True, False])
def test_cli_cpu_image_matches_oracle_pipeline(files, wproj):
    """grid.out -cpu_image: the CPU reference's image (uniform weighting,
    hermitian completion, shift / forward FFT / shift; src/reference/main.c)."""
    d = files["dir"]
    img = str(d / f"cpu_{wproj}.img")
    args = [os.path.join(PKG, "grid.out"), f"-vis={files['visfile']}", f"-theta={THETA}",
            f"-lambda={LAM}", "-flat", f"-image={img}", "-cpu_image"]
    args += ([f"-wkernel={files['wkp']}", "-wproj"] if wproj else
             [f"-wkernel={files['wkt']}", "-subgrid=64", "-margin=16", "-winc=10"])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(img, dtype=np.float64).reshape(N, N)
    wvis, _ = orc.weight(N, THETA, files["u"], files["v"], files["vis"])
    if wproj:
        ref, _, _ = orc.grid_wprojection(N, THETA, files["u"], files["v"], files["w"], wvis,
                                         files["tab"], files["w_min"], 10.0)
    else:
        ref, _, _ = orc.grid_wtowers(N, THETA, files["u"], files["v"], files["w"], wvis,
                                     files["tabt"], -5.0, 1.0, 64, 16, 10.0)
    ref = orc.fft_shift(orc.fft2d(orc.fft_shift(orc.make_hermitian(ref)), -1))
    assert rel_max(got, ref.real) < 1e-11
