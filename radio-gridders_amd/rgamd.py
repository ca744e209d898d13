"""ctypes binding of librg (include/rg.h) for tests, benchmarks and scripts.

Device memory comes from torch tensors (torch is plumbing here: allocation and
streams); every computation runs in librg's HIP kernels.  Loading fails loudly
when librg.so has not been built -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librg.so")

RG_OK = 0
RG_ERR_PARAM = -1
RG_ERR_IO = -2
RG_ERR_SHAPE = -3
RG_ERR_HIP = -4
RG_ERR_RANGE = -5
RG_ERR_NOMEM = -6
RG_ERR_UNSUPPORTED = -7

RG_C128 = 0
RG_C64 = 1
RG_MODE_PLAIN = 0
RG_MODE_BATCHED = 1
RG_DFT_FP64 = 0
RG_DFT_MIXED = 1
RG_IMAGE_GPU = 0
RG_IMAGE_CPU = 1

TIMERS = {"wproj_tiles": 0, "wproj_plain": 1, "binning": 2, "tower_tiles": 3, "tower_reduce": 4,
          "tower_cols": 5, "tower_finish": 6, "tower_combine": 7}


class RGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"librg error {code}: {msg}")
        self.code = code


class FlatVis(C.Structure):
    _fields_ = [("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p), ("vis", C.c_void_p),
                ("n", C.c_int64), ("vis_type", C.c_int32), ("reserved", C.c_int32)]


class WKern(C.Structure):
    _fields_ = [("data", C.c_void_p), ("type", C.c_int32), ("plane_count", C.c_int32),
                ("oversampling", C.c_int32), ("size", C.c_int32), ("w_min", C.c_double),
                ("w_step", C.c_double)]


class GridParams(C.Structure):
    _fields_ = [("grid_size", C.c_int32), ("mode", C.c_int32), ("theta", C.c_double),
                ("subgrid_size", C.c_int32), ("subgrid_margin", C.c_int32), ("winc", C.c_double),
                ("tile", C.c_int32), ("vis_per_item", C.c_int32), ("compensated", C.c_int32),
                ("floor_batch", C.c_int32), ("fft_radix", C.c_int32)]


class SepKernDesc(C.Structure):
    _fields_ = [("data", C.c_void_p), ("size", C.c_int32), ("oversampling", C.c_int32)]


class WStackParams(C.Structure):
    _fields_ = [("theta", C.c_double), ("lambda_", C.c_double), ("du", C.c_double),
                ("dw", C.c_double), ("x0", C.c_double), ("aa_support_uv", C.c_int32),
                ("aa_support_w", C.c_int32)]


class GridStats(C.Structure):
    _fields_ = [("n_vis", C.c_int64), ("n_gridded", C.c_int64), ("n_dropped", C.c_int64),
                ("n_bad_plane", C.c_int64), ("n_items", C.c_int64), ("n_tasks", C.c_int64),
                ("n_towers", C.c_int64), ("n_records", C.c_int64), ("n_bands", C.c_int64),
                ("n_chunks", C.c_int64)]


# (name, restype, argtypes) of every exported symbol of include/rg.h
_P = C.c_void_p
_I = C.c_int32
_D = C.c_double
_L = C.c_int64
SIGNATURES = [
    ("rg_last_error", C.c_char_p, []),
    ("rg_abi_version", C.c_int, []),
    ("rg_context_create", C.c_int, [C.POINTER(_P), C.c_int]),
    ("rg_context_destroy", C.c_int, [_P]),
    ("rg_context_stats", C.c_int, [_P, C.POINTER(GridStats)]),
    ("rg_context_thaw", C.c_int, [_P]),
    ("rg_context_set_timing", C.c_int, [_P, _I]),
    ("rg_context_timer", C.c_int, [_P, _I, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    ("rg_grid", C.c_int, [_P, C.POINTER(GridParams), C.POINTER(FlatVis), C.POINTER(WKern), _P, _I, _P]),
    ("rg_grid_channels", C.c_int, [_P, C.POINTER(GridParams), C.POINTER(FlatVis), C.POINTER(C.c_int64),
                                   _I, C.POINTER(WKern), _P, _I, _P]),
    ("rg_grid_noalloc", C.c_int, [C.POINTER(GridParams), C.POINTER(FlatVis), C.POINTER(WKern), _P, _I,
                                  _I, _P]),
    ("rg_grid_to_image", C.c_int, [_P, _P, _I, _I, _P]),
    ("rg_fft2d", C.c_int, [_P, _P, _I, _I, _I, _I, _P]),
    ("rg_fft_shift", C.c_int, [_P, _I, _I, _P]),
    ("rg_weight", C.c_int, [_P, _P, _P, _L, _P, _I, _I, _D, _P, _P]),
    ("rg_make_hermitian", C.c_int, [_P, _I, _I, _P]),
    ("rg_gcf_correct", C.c_int, [_P, _I, _I, _P, _P]),
    ("rg_set_image_convention", C.c_int, [_I]),
    ("rg_dft", C.c_int, [_P, C.POINTER(FlatVis), _P, _P, _L, _I, _D, _P, _I, _P]),
    ("rg_dft_image", C.c_int, [_P, C.POINTER(FlatVis), _I, _D, _P, _I, _I, _I, _P]),
    ("rg_predict_points", C.c_int, [_P, _P, _P, _L, _P, _P, _P, _I, _P, _I, _P]),
    ("rg_wstack_build", C.c_int, [_P, C.POINTER(WStackParams), _P, _P, _P, _I,
                                  C.POINTER(SepKernDesc), C.POINTER(SepKernDesc), C.POINTER(_I),
                                  C.POINTER(_I), C.POINTER(C.c_int64), _P]),
    ("rg_wstack_degrid", C.c_int, [_P, C.POINTER(WStackParams), _P, _P, _P, _L,
                                   C.POINTER(SepKernDesc), C.POINTER(SepKernDesc), _P, _P]),
    ("rg_wstack_copy", C.c_int, [_P, _P, _L, _P]),
    ("rg_epic_grid", C.c_int, [_P, _P, _P, _P, _P, _I, _I, _P, _I, _I, _P, _P]),
    ("rg_wstack_predict", C.c_int, [_P, C.POINTER(WStackParams), _P, _P, _P, _I, _P, _P, _P, _L,
                                    C.POINTER(SepKernDesc), C.POINTER(SepKernDesc),
                                    C.POINTER(SepKernDesc), C.POINTER(SepKernDesc), _P, _P]),
    ("rg_pswf", C.c_double, [_D, _D]),
    ("rg_pswf_eval", C.c_int, [_D, _P, _L, _P]),
    ("rg_wkern_generate", C.c_int, [_P, _I, _I, _I, _I, _D, _D, _D, _I, _D, _I]),
    ("rg_wtowers_flat", C.c_int, [C.c_char_p, C.c_char_p, _P, _P, C.c_int, _D, _D, _D, _D, C.c_int, C.c_int, _D]),
    ("rg_wtowers", C.c_int, [C.c_char_p, C.c_char_p, _P, _P, C.c_int, _D, _D, _D, _D, C.c_int, C.c_int, _D]),
    ("rg_wproj_flat", C.c_int, [C.c_char_p, C.c_char_p, _P, _P, C.c_int, _D, _D, _D, _D, C.c_int]),
    ("rg_wproj", C.c_int, [C.c_char_p, C.c_char_p, _P, _P, C.c_int, _D, _D, _D, _D, C.c_int]),
    ("rg_image_dft_host_flat", C.c_int, [C.c_char_p, _P, _P, C.c_int, _D, _D, _D, _D, C.c_int, C.c_int]),
    ("rg_image_dft_host", C.c_int, [C.c_char_p, _P, _P, C.c_int, _D, _D, _D, _D, C.c_int, C.c_int]),
    ("rg_set_file_mode", C.c_int, [_I]),
    ("rg_hdf5_available", C.c_int, []),
    ("rg_load_flat_vis", C.c_int, [C.c_char_p, _D, _D, C.POINTER(C.c_int64), C.POINTER(_P), C.POINTER(_P),
                                   C.POINTER(_P), C.POINTER(_P)]),
    ("rg_load_wkern", C.c_int, [C.c_char_p, _D, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I),
                                C.POINTER(_D), C.POINTER(_D), C.POINTER(_P)]),
    ("rg_free_host", None, [_P]),
    ("rg_load_sep_kern", C.c_int, [C.c_char_p, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_D),
                                   C.POINTER(_D), C.POINTER(_D), C.POINTER(_P)]),
    ("rg_write_vis_hdf5", C.c_int, [C.c_char_p, _I, _P, _P, _I, _I, _P, _P, _P, _P]),
    ("rg_write_wkern_hdf5", C.c_int, [C.c_char_p, _D, _I, _I, _I, _D, _D, _P]),
]

_lib = None


def lib():
    """Loads librg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (__graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code: int):
    if code != RG_OK:
        raise RGError(code, lib().rg_last_error().decode())


def grid_noalloc(p: "GridParams", fv: "FlatVis", wk: "WKern", grid, accumulate=False, stream=None):
    """rg_grid_noalloc: batched W-projection on the current device through
    librg's static scratch (no context, no allocation)."""
    check(lib().rg_grid_noalloc(C.byref(p), C.byref(fv), C.byref(wk), _ptr(grid),
                                grid_type_of(grid), 1 if accumulate else 0, _stream(stream)))


def _ptr(t) -> int:
    """Device/host address of a torch tensor or numpy array (0 for None)."""
    if t is None:
        return 0
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


def _stream(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream


def flat_vis(u, v, w, vis) -> FlatVis:
    """Descriptor of device tensors: u, v, w float64; vis complex128 or complex64.
    The descriptor keeps the tensors alive: it holds their raw addresses."""
    import torch
    assert u.dtype == torch.float64 and v.dtype == torch.float64 and w.dtype == torch.float64
    assert vis.dtype in (torch.complex128, torch.complex64)
    n = u.numel()
    assert v.numel() == n and w.numel() == n and vis.numel() == n
    for t in (u, v, w, vis):
        assert t.is_contiguous() and t.is_cuda
    d = FlatVis(_ptr(u), _ptr(v), _ptr(w), _ptr(vis), n,
                RG_C128 if vis.dtype == torch.complex128 else RG_C64, 0)
    d._keep = (u, v, w, vis)
    return d


def wkern_desc(table, w_min: float, w_step: float) -> WKern:
    """Descriptor of a device table [P][R][R][S][S] (complex128 or complex64);
    keeps the table alive."""
    import torch
    assert table.dim() == 5 and table.is_contiguous() and table.is_cuda
    P, R, R2, S, S2 = table.shape
    assert R == R2 and S == S2
    typ = RG_C128 if table.dtype == torch.complex128 else RG_C64
    d = WKern(_ptr(table), typ, P, R, S, float(w_min), float(w_step))
    d._keep = table
    return d


def params(grid_size, theta, mode=RG_MODE_BATCHED, subgrid=0, margin=0, winc=0.0, tile=0,
           vis_per_item=0, compensated=0, floor_batch=0, fft_radix=0) -> GridParams:
    return GridParams(int(grid_size), int(mode), float(theta), int(subgrid), int(margin),
                      float(winc), int(tile), int(vis_per_item), int(compensated), int(floor_batch),
                      int(fft_radix))


def grid_type_of(t) -> int:
    import torch
    return RG_C128 if t.dtype == torch.complex128 else RG_C64


class Context:
    """An rg_context on one device."""

    def __init__(self, device: int = 0):
        self.device = int(device)
        self.h = C.c_void_p()
        check(lib().rg_context_create(C.byref(self.h), int(device)))

    def close(self):
        if self.h:
            lib().rg_context_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid(self, p: GridParams, fv: FlatVis, wk: WKern, grid, stream=None):
        check(lib().rg_grid(self.h, C.byref(p), C.byref(fv), C.byref(wk), _ptr(grid),
                            grid_type_of(grid), _stream(stream)))

    def grid_channels(self, p: GridParams, fv: FlatVis, offsets, wk: WKern, grids, stream=None):
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        check(lib().rg_grid_channels(self.h, C.byref(p), C.byref(fv),
                                     off.ctypes.data_as(C.POINTER(C.c_int64)), len(off) - 1,
                                     C.byref(wk), _ptr(grids), grid_type_of(grids), _stream(stream)))

    def thaw(self):
        """rg_context_thaw: no graph captured from this context is replayed again."""
        check(lib().rg_context_thaw(self.h))

    def set_timing(self, enable: bool):
        check(lib().rg_context_set_timing(self.h, 1 if enable else 0))

    def timer(self, name: str):
        """(total ms, launches) of a kernel timer since its last read."""
        ms, cnt = C.c_double(), C.c_int64()
        check(lib().rg_context_timer(self.h, TIMERS[name], C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def stats(self) -> GridStats:
        s = GridStats()
        check(lib().rg_context_stats(self.h, C.byref(s)))
        return s

    def grid_to_image(self, grid, stream=None):
        check(lib().rg_grid_to_image(self.h, _ptr(grid), grid_type_of(grid), grid.shape[-1],
                                     _stream(stream)))

    def fft2d(self, data, sign: int, stream=None):
        n = data.shape[-1]
        batch = data.numel() // (n * n)
        check(lib().rg_fft2d(self.h, _ptr(data), grid_type_of(data), n, batch, int(sign),
                             _stream(stream)))

    def weight(self, u, v, vis, grid_size: int, theta: float, counts=None, stream=None):
        """Uniform weighting of the device visibilities `vis` in place
        (weight, src/reference/grid.c:59-81); `counts` (int32/uint32 device
        tensor of grid_size^2) receives the per-cell counts when given."""
        check(lib().rg_weight(self.h, _ptr(u), _ptr(v), u.numel(), _ptr(vis), grid_type_of(vis),
                              int(grid_size), float(theta), _ptr(counts), _stream(stream)))

    def wstack_build(self, p: WStackParams, l, m, flux, corr_lm: SepKernDesc, corr_n: SepKernDesc,
                     stream=None):
        """Builds the context's W-stack; returns (og, w_planes, skipped points)."""
        og, wp, sk = C.c_int32(), C.c_int32(), C.c_int64()
        check(lib().rg_wstack_build(self.h, C.byref(p), _ptr(l), _ptr(m), _ptr(flux), l.numel(),
                                    C.byref(corr_lm), C.byref(corr_n), C.byref(og), C.byref(wp),
                                    C.byref(sk), _stream(stream)))
        return og.value, wp.value, sk.value

    def wstack_degrid(self, p: WStackParams, u, v, w, conv_uv: SepKernDesc, conv_w: SepKernDesc,
                      out, stream=None):
        check(lib().rg_wstack_degrid(self.h, C.byref(p), _ptr(u), _ptr(v), _ptr(w), u.numel(),
                                     C.byref(conv_uv), C.byref(conv_w), _ptr(out), _stream(stream)))

    def wstack_stack(self, og: int, planes: int, stream=None) -> np.ndarray:
        """The context's W-stack as a host array [planes][og][og] complex128."""
        out = np.empty((planes, og, og), np.complex128)
        check(lib().rg_wstack_copy(self.h, out.ctypes.data, out.nbytes, _stream(stream)))
        return out

    def wstack_predict(self, p: WStackParams, l, m, flux, u, v, w, conv_uv, conv_w, corr_lm,
                       corr_n, out, stream=None):
        check(lib().rg_wstack_predict(self.h, C.byref(p), _ptr(l), _ptr(m), _ptr(flux), l.numel(),
                                      _ptr(u), _ptr(v), _ptr(w), u.numel(), C.byref(conv_uv),
                                      C.byref(conv_w), C.byref(corr_lm), C.byref(corr_n),
                                      _ptr(out), _stream(stream)))

    def epic_grid(self, f, x, y, illum, grids, stream=None):
        """E-field grids: f [nbatch][n_ant] complex64, x, y int32 device tensors,
        illum [S][S] complex64, grids [nbatch][g][g] complex64 (all device)."""
        nbatch, n_ant = f.shape
        check(lib().rg_epic_grid(self.h, _ptr(f), _ptr(x), _ptr(y), None, n_ant, nbatch,
                                 _ptr(illum), illum.shape[0], grids.shape[-1], _ptr(grids),
                                 _stream(stream)))

    def dft(self, fv: FlatVis, px, py, grid_size: int, lam: float, out, mode=RG_DFT_FP64,
            stream=None):
        check(lib().rg_dft(self.h, C.byref(fv), _ptr(px), _ptr(py), px.numel(), int(grid_size),
                           float(lam), _ptr(out), int(mode), _stream(stream)))

    def dft_image(self, fv: FlatVis, grid_size: int, lam: float, out, blocks: int, threads: int,
                  mode=RG_DFT_FP64, stream=None):
        check(lib().rg_dft_image(self.h, C.byref(fv), int(grid_size), float(lam), _ptr(out),
                                 int(blocks), int(threads), int(mode), _stream(stream)))


def fft_shift(data, stream=None):
    check(lib().rg_fft_shift(_ptr(data), grid_type_of(data), data.shape[-1], _stream(stream)))


def make_hermitian(grid, stream=None):
    """Hermitian completion of an n x n device grid in place
    (make_hermitian, src/reference/grid.c:639-662)."""
    check(lib().rg_make_hermitian(_ptr(grid), grid_type_of(grid), grid.shape[-1], _stream(stream)))


def gcf_correct(image, taper, stream=None):
    """image[r][c] /= taper[r] * taper[c] in place (taper: float64 device tensor)."""
    check(lib().rg_gcf_correct(_ptr(image), grid_type_of(image), image.shape[-1], _ptr(taper),
                               _stream(stream)))


def set_image_convention(convention: int):
    """Image convention of the file-level wrappers (RG_IMAGE_GPU / RG_IMAGE_CPU)."""
    check(lib().rg_set_image_convention(int(convention)))


def predict_points(u, v, w, l, m, flux, out, stream=None):
    """Point-source visibilities on the device (all arguments device tensors)."""
    check(lib().rg_predict_points(_ptr(u), _ptr(v), _ptr(w), u.numel(), _ptr(l), _ptr(m),
                                  _ptr(flux), l.numel(), _ptr(out), grid_type_of(out),
                                  _stream(stream)))


def pswf(c: float, x: float) -> float:
    return lib().rg_pswf(float(c), float(x))


def pswf_eval(c: float, x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    check(lib().rg_pswf_eval(float(c), x.ctypes.data, x.size, out.ctypes.data))
    return out


def make_wkernels(planes: int, oversampling: int, size: int, theta: float, w_min: float,
                  w_step: float, dtype=np.complex128, image_samples: int = 128,
                  pswf_c: float = 0.0, nthreads: int = 0) -> np.ndarray:
    """W-kernel table [P][R][R][S][S] generated on the host (FP64)."""
    out = np.empty((planes, oversampling, oversampling, size, size), dtype=dtype)
    typ = RG_C128 if dtype == np.complex128 else RG_C64
    check(lib().rg_wkern_generate(out.ctypes.data, typ, planes, oversampling, size, float(theta),
                                  float(w_min), float(w_step), int(image_samples), float(pswf_c),
                                  int(nthreads)))
    return out


@dataclass
class LoadedKern:
    data: np.ndarray
    w_min: float
    w_step: float


def load_flat_vis(path: str, bl_min: float = 0.0, bl_max: float = 1.7976931348623158e308):
    n = C.c_int64()
    ptrs = [C.c_void_p() for _ in range(4)]
    check(lib().rg_load_flat_vis(path.encode(), bl_min, bl_max, C.byref(n), *[C.byref(p) for p in ptrs]))
    out = []
    for k, p in enumerate(ptrs):
        cnt = n.value * (2 if k == 3 else 1)
        arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(max(cnt, 1),))[:cnt].copy()
        lib().rg_free_host(p)
        out.append(arr)
    u, v, w, vis = out
    return u, v, w, vis.view(np.complex128)


def load_wkern(path: str, theta: float) -> LoadedKern:
    P, R, S = C.c_int32(), C.c_int32(), C.c_int32()
    wmin, wstep = C.c_double(), C.c_double()
    p = C.c_void_p()
    check(lib().rg_load_wkern(path.encode(), theta, C.byref(P), C.byref(R), C.byref(S),
                              C.byref(wmin), C.byref(wstep), C.byref(p)))
    cnt = P.value * R.value * R.value * S.value * S.value * 2
    arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(cnt,)).copy()
    lib().rg_free_host(p)
    return LoadedKern(arr.view(np.complex128).reshape(P.value, R.value, R.value, S.value, S.value),
                      wmin.value, wstep.value)


@dataclass
class SepKern:
    """Separable kernel: data[oversampling][size] in file order, and the
    spacing attributes of the file."""
    data: np.ndarray
    size: int
    oversampling: int
    du: float
    dw: float
    x0: float


def sep_kern_desc(data, size: int, oversampling: int) -> SepKernDesc:
    """Descriptor of a device float64 kernel table."""
    return SepKernDesc(_ptr(data), int(size), int(oversampling))


def wstack_params(theta, lam, du, dw, x0, aa_uv, aa_w) -> WStackParams:
    return WStackParams(float(theta), float(lam), float(du), float(dw), float(x0), int(aa_uv),
                        int(aa_w))


def load_sep_kern(path: str, transposed: bool = False) -> SepKern:
    size, ov = C.c_int32(), C.c_int32()
    du, dw, x0 = C.c_double(), C.c_double(), C.c_double()
    p = C.c_void_p()
    check(lib().rg_load_sep_kern(path.encode(), int(transposed), C.byref(size), C.byref(ov),
                                 C.byref(du), C.byref(dw), C.byref(x0), C.byref(p)))
    cnt = size.value * ov.value
    arr = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(cnt,)).copy()
    lib().rg_free_host(p)
    return SepKern(arr, size.value, ov.value, du.value, dw.value, x0.value)


def write_vis_hdf5(path, a1, a2, time, freq, uvw, vis):
    """uvw [n_bl][T][3] metres, vis [n_bl][T][F] complex128."""
    a1 = np.ascontiguousarray(a1, np.int32)
    a2 = np.ascontiguousarray(a2, np.int32)
    time = np.ascontiguousarray(time, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    uvw = np.ascontiguousarray(uvw, np.float64)
    vis = np.ascontiguousarray(vis, np.complex128)
    nbl, T, F = vis.shape
    check(lib().rg_write_vis_hdf5(path.encode(), nbl, a1.ctypes.data, a2.ctypes.data, T, F,
                                  time.ctypes.data, freq.ctypes.data, uvw.ctypes.data,
                                  vis.ctypes.data))


def write_wkern_hdf5(path, theta, table, w_min, w_step):
    t = np.ascontiguousarray(table, np.complex128)
    P, R, _, S, _ = t.shape
    check(lib().rg_write_wkern_hdf5(path.encode(), float(theta), P, R, S, float(w_min),
                                    float(w_step), t.ctypes.data))


def autotune(ctx, base: GridParams, fv: FlatVis, wk: WKern, grid, candidates, offsets=None,
             reps: int = 3):
    """Times every candidate (a dict of GridParams fields over `base`, e.g.
    {"tile": 48, "vis_per_item": 1024}) on the caller's own data and returns
    (best GridParams, [(candidate, ms)]).  Each candidate runs once to warm up
    (scratch memory, caches), then `reps` times between device events; the
    grid is overwritten."""
    import torch

    timings = []
    best, best_ms = None, float("inf")
    for cand in candidates:
        p = GridParams()
        for name, _ in GridParams._fields_:
            setattr(p, name, getattr(base, name))
        for k, v in cand.items():
            setattr(p, k, v)

        def run():
            if offsets is None:
                ctx.grid(p, fv, wk, grid)
            else:
                ctx.grid_channels(p, fv, offsets, wk, grid)

        run()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        timings.append((dict(cand), ms))
        if ms < best_ms:
            best, best_ms = p, ms
    return best, timings

