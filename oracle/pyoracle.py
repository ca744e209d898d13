"""ctypes wrapper of the oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is the checker: only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py use it.  Arrays are numpy, complex values
complex128.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}


class C128(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]

_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_L = C.c_int64


def lib(fast: bool = False):
    """liboracle.so (checker) or liboracle_fast.so (CPU-baseline timing build)."""
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run make -C oracle")
        L = C.CDLL(path)
        L.orc_frac_coord.argtypes = [_I, _I, _I, _D, _D, _D, _D, _D, C.POINTER(_I), C.POINTER(_I),
                                     C.POINTER(_I), C.POINTER(_I)]
        L.orc_frac_coord.restype = None
        L.orc_frac_coord_packed.argtypes = [_I, _I, _I, _D, _D, _D, _D, _D, C.POINTER(_I), C.POINTER(_I)]
        L.orc_frac_coord_packed.restype = None
        L.orc_w_plane.argtypes = [_D, _D, _D]
        L.orc_w_plane.restype = _I
        L.orc_tower_index.argtypes = [_D, _D, _I, _I]
        L.orc_tower_index.restype = _I
        L.orc_floor_index.argtypes = [_D, _D]
        L.orc_floor_index.restype = _I
        L.orc_grid_wprojection_flat.argtypes = [_P, _I, _D, _P, _P, _P, _P, _L, _P, _I, _I, _I, _D, _D,
                                                C.POINTER(_L)]
        L.orc_grid_wprojection_flat.restype = C.c_uint64
        L.orc_grid_wprojection_flat_omp.argtypes = [_P, _I, _D, _P, _P, _P, _P, _L, _P, _I, _I, _I, _D,
                                                    _D, _I, C.POINTER(_L)]
        L.orc_grid_wprojection_flat_omp.restype = C.c_uint64
        L.orc_grid_wtowers_flat.argtypes = [_P, _I, _D, _P, _P, _P, _P, _L, _P, _I, _I, _I, _D, _D,
                                            _I, _I, _D, _I, C.POINTER(_L)]
        L.orc_grid_wtowers_flat.restype = C.c_uint64
        L.orc_wtransfer.argtypes = [_P, _I, _D, _D]
        L.orc_wtransfer.restype = None
        L.orc_cipow.argtypes = [C128, _I]
        L.orc_cipow.restype = C128
        L.orc_fft2d.argtypes = [_P, _I, _I]
        L.orc_fft2d.restype = None
        L.orc_fft1d.argtypes = [_P, _I, _I]
        L.orc_fft1d.restype = None
        L.orc_fft_shift.argtypes = [_P, _I]
        L.orc_fft_shift.restype = None
        L.orc_make_hermitian.argtypes = [_P, _I]
        L.orc_make_hermitian.restype = None
        L.orc_weight_flat.argtypes = [_P, _I, _D, _P, _P, _P, _L]
        L.orc_weight_flat.restype = None
        L.orc_grid_to_image.argtypes = [_P, _I]
        L.orc_grid_to_image.restype = None
        L.orc_dft_flat.argtypes = [_P, _P, _P, _L, _I, _D, _P, _P, _P, _P, _L, _I]
        L.orc_dft_flat.restype = None
        L.orc_wstack_geometry.argtypes = [_D, _D, _D, _I, _D, C.POINTER(_I), C.POINTER(_I)]
        L.orc_wstack_geometry.restype = _I
        L.orc_wstack_build.argtypes = [_P, _D, _D, _P, _P, _P, _I, _D, _D, _I, _D, _P, _I, _P, _I]
        L.orc_wstack_build.restype = _I
        L.orc_wstack_degrid.argtypes = [_P, _P, _P, _P, _L, _P, _I, _I, _I, _D, _D, _I, _I, _P, _I,
                                        _P, _I, _I]
        L.orc_wstack_degrid.restype = _L
        L.orc_predict_quantized.argtypes = [_P, _P, _P, _P, _I, _D, _D, _P, _P, _P, _L, _I]
        L.orc_predict_quantized.restype = None
        L.orc_epic_grid.argtypes = [_P, _P, _P, _P, _I, _I, _P, _I, _I]
        L.orc_epic_grid.restype = None
        _libs[name] = L
    return _libs[name]


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def frac_coord(N, S, R, theta, u, v, du=0.0, dv=0.0):
    gx, gy, xf, yf = _I(), _I(), _I(), _I()
    lib().orc_frac_coord(N, S, R, theta, u, v, du, dv, C.byref(gx), C.byref(gy), C.byref(xf),
                         C.byref(yf))
    return gx.value, gy.value, xf.value, yf.value


def frac_coord_packed(N, S, R, theta, u, v, du=0.0, dv=0.0):
    go, so = _I(), _I()
    lib().orc_frac_coord_packed(N, S, R, theta, u, v, du, dv, C.byref(go), C.byref(so))
    return go.value, so.value


def w_plane(w, w_min, w_step):
    return lib().orc_w_plane(w, w_min, w_step)


def tower_index(u, theta, N, chunk):
    return lib().orc_tower_index(u, theta, N, chunk)


def floor_index(w, winc):
    return lib().orc_floor_index(w, winc)


def grid_wprojection(N, theta, u, v, w, vis, kern, w_min, w_step, nthreads=0, fast=False):
    """Returns (grid [N][N] complex128, flops, n_bad)."""
    u, v, w = _c(u, np.float64), _c(v, np.float64), _c(w, np.float64)
    vis = _c(vis, np.complex128)
    kern = _c(kern, np.complex128)
    P, R, _, S, _ = kern.shape
    g = np.zeros((N, N), np.complex128)
    bad = _L()
    L = lib(fast)
    if nthreads and nthreads > 1:
        fl = L.orc_grid_wprojection_flat_omp(g.ctypes.data, N, theta, u.ctypes.data, v.ctypes.data,
                                             w.ctypes.data, vis.ctypes.data, len(u), kern.ctypes.data,
                                             P, R, S, w_min, w_step, nthreads, C.byref(bad))
    else:
        fl = L.orc_grid_wprojection_flat(g.ctypes.data, N, theta, u.ctypes.data, v.ctypes.data,
                                         w.ctypes.data, vis.ctypes.data, len(u), kern.ctypes.data,
                                         P, R, S, w_min, w_step, C.byref(bad))
    return g, int(fl), bad.value


def grid_wtowers(N, theta, u, v, w, vis, kern, w_min, w_step, subgrid, margin, winc, nthreads=1,
                 fast=False):
    u, v, w = _c(u, np.float64), _c(v, np.float64), _c(w, np.float64)
    vis = _c(vis, np.complex128)
    kern = _c(kern, np.complex128)
    P, R, _, S, _ = kern.shape
    g = np.zeros((N, N), np.complex128)
    bad = _L()
    fl = lib(fast).orc_grid_wtowers_flat(g.ctypes.data, N, theta, u.ctypes.data, v.ctypes.data,
                                         w.ctypes.data, vis.ctypes.data, len(u), kern.ctypes.data,
                                         P, R, S, w_min, w_step, subgrid, margin, winc, nthreads,
                                         C.byref(bad))
    return g, int(fl), bad.value


def cipow(z: complex, e: int) -> complex:
    r = lib().orc_cipow(C128(z.real, z.imag), int(e))
    return complex(r.re, r.im)


def wtransfer(n, theta, winc):
    out = np.zeros((n, n), np.complex128)
    lib().orc_wtransfer(out.ctypes.data, n, theta, winc)
    return out


def fft2d(a, sign):
    a = np.array(a, dtype=np.complex128, order="C", copy=True)
    lib().orc_fft2d(a.ctypes.data, a.shape[0], sign)
    return a


def fft1d(a, sign):
    a = np.array(a, dtype=np.complex128, order="C", copy=True)
    lib().orc_fft1d(a.ctypes.data, a.shape[0], sign)
    return a


def fft_shift(a):
    a = np.array(a, dtype=np.complex128, order="C", copy=True)
    lib().orc_fft_shift(a.ctypes.data, a.shape[0])
    return a


def make_hermitian(a):
    a = np.array(a, dtype=np.complex128, order="C", copy=True)
    lib().orc_make_hermitian(a.ctypes.data, a.shape[0])
    return a


def weight(N, theta, u, v, vis):
    u, v = _c(u, np.float64), _c(v, np.float64)
    vis = np.array(vis, dtype=np.complex128, copy=True)
    wg = np.zeros((N, N), np.uint32)
    lib().orc_weight_flat(wg.ctypes.data, N, theta, u.ctypes.data, v.ctypes.data, vis.ctypes.data,
                          len(u))
    return vis, wg


def grid_to_image(g):
    g = np.array(g, dtype=np.complex128, order="C", copy=True)
    lib().orc_grid_to_image(g.ctypes.data, g.shape[0])
    return g


def dft(px, py, N, lam, u, v, w, vis, nthreads=0):
    px, py = _c(px, np.int32), _c(py, np.int32)
    u, v, w = _c(u, np.float64), _c(v, np.float64), _c(w, np.float64)
    vis = _c(vis, np.complex128)
    out = np.zeros(len(px), np.complex128)
    lib().orc_dft_flat(out.ctypes.data, px.ctypes.data, py.ctypes.data, len(px), N, lam,
                       u.ctypes.data, v.ctypes.data, w.ctypes.data, vis.ctypes.data, len(u),
                       nthreads or (os.cpu_count() or 1))
    return out


def wstack_geometry(theta, lam, dw, aa_w, x0):
    """(oversampled grid size, number of w-planes) of wstack_predict."""
    og, wp = _I(), _I()
    if lib().orc_wstack_geometry(theta, lam, dw, aa_w, x0, C.byref(og), C.byref(wp)) != 0:
        raise ValueError("inconsistent W-stack geometry")
    return og.value, wp.value


def wstack_build(theta, lam, l, m, flux, du, dw, aa_w, x0, corr_lm, corr_n, fast=False):
    """W-stack [w_planes][og][og] complex128 and the number of skipped points."""
    og, wp = wstack_geometry(theta, lam, dw, aa_w, x0)
    l, m = _c(l, np.float64), _c(m, np.float64)
    fl = None if flux is None else _c(flux, np.float64)
    clm, cn = _c(corr_lm, np.float64), _c(corr_n, np.float64)
    stack = np.zeros((wp, og, og), np.complex128)
    dropped = lib(fast).orc_wstack_build(stack.ctypes.data, theta, lam, l.ctypes.data, m.ctypes.data,
                                         None if fl is None else fl.ctypes.data, len(l), du, dw,
                                         aa_w, x0, clm.ctypes.data, len(clm), cn.ctypes.data, len(cn))
    return stack, dropped


def wstack_degrid(stack, g, u, v, w, du, dw, conv_uv, aa_uv, ov_uv, conv_w, aa_w, ov_w,
                  nthreads=0, fast=False):
    """Predicted visibilities (complex128) and the number not predicted."""
    u, v, w = _c(u, np.float64), _c(v, np.float64), _c(w, np.float64)
    cuv, cw = _c(conv_uv, np.float64), _c(conv_w, np.float64)
    stack = np.ascontiguousarray(stack, dtype=np.complex128)
    out = np.zeros(len(u), np.complex128)
    dropped = lib(fast).orc_wstack_degrid(out.ctypes.data, u.ctypes.data, v.ctypes.data,
                                          w.ctypes.data, len(u), stack.ctypes.data, stack.shape[1],
                                          stack.shape[0], g, du, dw, aa_uv, aa_w, cuv.ctypes.data,
                                          ov_uv, cw.ctypes.data, ov_w, nthreads)
    return out, dropped


def predict_quantized(l, m, flux, theta, lam, u, v, w, nthreads=0):
    """Direct sum over the quantised points (predict_visibility_quantized_vec)."""
    l, m = _c(l, np.float64), _c(m, np.float64)
    fl = None if flux is None else _c(flux, np.float64)
    u, v, w = _c(u, np.float64), _c(v, np.float64), _c(w, np.float64)
    out = np.zeros(len(u), np.complex128)
    lib().orc_predict_quantized(out.ctypes.data, l.ctypes.data, m.ctypes.data,
                                None if fl is None else fl.ctypes.data, len(l), theta, lam,
                                u.ctypes.data, v.ctypes.data, w.ctypes.data, len(u), nthreads)
    return out


def epic_grid(f, x, y, illum, g):
    """E-field grids [nbatch][g][g] complex128; f [nbatch][n_ant], x, y int, illum [S][S]."""
    f = np.ascontiguousarray(f, dtype=np.complex128)
    nbatch, n_ant = f.shape
    x, y = _c(x, np.int32), _c(y, np.int32)
    il = np.ascontiguousarray(illum, dtype=np.complex128)
    out = np.empty((nbatch, g, g), np.complex128)
    lib().orc_epic_grid(out.ctypes.data, f.ctypes.data, x.ctypes.data, y.ctypes.data, n_ant, nbatch,
                        il.ctypes.data, il.shape[0], g)
    return out
