"""HDF5 readers and writers of the reference layouts (host code of librg)."""
import math

import numpy as np
import pytest

import rgamd

C_LIGHT = 299792458.0
pytestmark = pytest.mark.skipif(not rgamd.lib().rg_hdf5_available(), reason="libhdf5 not loadable")


def make_file(path, rng, nbl=5, T=4, F=3):
    a1 = np.array([0, 0, 1, 1, 2][:nbl], np.int32)
    a2 = np.array([1, 2, 2, 3, 3][:nbl], np.int32)
    uvw = rng.uniform(-3000, 3000, (nbl, T, 3))
    # baseline lengths (first sample) chosen for the filter test
    uvw[:, 0, 0] = [100.0, 500.0, 1000.0, 2000.0, 4000.0][:nbl]
    uvw[:, 0, 1] = 0.0
    time = np.arange(T, dtype=np.float64)
    freq = np.array([1.0e8, 1.5e8, 2.0e8][:F])
    vis = rng.standard_normal((nbl, T, F)) + 1j * rng.standard_normal((nbl, T, F))
    rgamd.write_vis_hdf5(path, a1, a2, time, freq, uvw, vis)
    return uvw, freq, vis


def test_visibility_roundtrip_flatten_and_filter(tmp_path):
    rng = np.random.default_rng(1)
    path = str(tmp_path / "vis.h5")
    uvw, freq, vis = make_file(path, rng)
    u, v, w, fv = rgamd.load_flat_vis(path)
    nbl, T, F = vis.shape
    assert len(u) == nbl * T * F
    k = 0
    for b in range(nbl):
        for t in range(T):
            for f in range(F):
                assert u[k] == uvw[b, t, 0] * freq[f] / C_LIGHT
                assert v[k] == uvw[b, t, 1] * freq[f] / C_LIGHT
                assert w[k] == uvw[b, t, 2] * freq[f] / C_LIGHT
                assert fv[k] == vis[b, t, f]
                k += 1
    # [bl_min, bl_max) in metres on the first sample: keeps 500 and 1000
    u2, _, _, _ = rgamd.load_flat_vis(path, 500.0, 2000.0)
    assert len(u2) == 2 * T * F
    assert u2[0] == uvw[1, 0, 0] * freq[0] / C_LIGHT


def test_wkernel_roundtrip_and_plane_index(tmp_path):
    path = str(tmp_path / "wk.h5")
    rng = np.random.default_rng(2)
    P, R, S = 5, 4, 8
    tab = rng.standard_normal((P, R, R, S, S)) + 1j * rng.standard_normal((P, R, R, S, S))
    rgamd.write_wkern_hdf5(path, 0.1, tab, -20.0, 10.0)
    k = rgamd.load_wkern(path, 0.1)
    assert k.data.shape == tab.shape and np.array_equal(k.data, tab)
    assert k.w_min == -20.0 and k.w_step == 10.0
    # the group is wkern/<theta %g>: a different theta is not found
    with pytest.raises(rgamd.RGError) as e:
        rgamd.load_wkern(path, 0.2)
    assert e.value.code == rgamd.RG_ERR_IO


def test_missing_and_foreign_files(tmp_path):
    with pytest.raises(rgamd.RGError) as e:
        rgamd.load_flat_vis(str(tmp_path / "none.h5"))
    assert e.value.code == rgamd.RG_ERR_IO
    p = tmp_path / "text.h5"
    p.write_text("not hdf5")
    with pytest.raises(rgamd.RGError) as e:
        rgamd.load_flat_vis(str(p))
    assert e.value.code == rgamd.RG_ERR_IO


def test_empty_baseline_groups_are_skipped(tmp_path):
    """Groups with no time samples hold nothing to grid: they are skipped
    instead of having their first uvw sample read (which the reference reads
    out of bounds, hdf5.c:104-106); the other baselines of a file still load."""
    path = str(tmp_path / "empty.h5")
    a = np.array([0, 0], np.int32)
    b = np.array([1, 2], np.int32)
    rgamd.write_vis_hdf5(path, a, b, np.zeros(0), np.array([1.5e8]), np.zeros((2, 0, 3)),
                         np.zeros((2, 0, 1), np.complex128))
    u, v, w, vis = rgamd.load_flat_vis(path)
    assert len(u) == 0 and len(vis) == 0
    u, v, w, vis = rgamd.load_flat_vis(path, 10.0, 1000.0)  # the filter reads no sample either
    assert len(u) == 0
