"""Shared helpers of the GPU parity tests (device staging, error metrics)."""
import numpy as np
import torch

import rgamd
import synth


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def stage(u, v, w, vis, vis_dtype=torch.complex128):
    du, dv, dw = dev(u, torch.float64), dev(v, torch.float64), dev(w, torch.float64)
    dvis = dev(vis).to(vis_dtype).contiguous()
    return (du, dv, dw, dvis), rgamd.flat_vis(du, dv, dw, dvis)


def kern(table, w_min, w_step, dtype=torch.complex128):
    t = dev(table).to(dtype).contiguous()
    return t, rgamd.wkern_desc(t, w_min, w_step)


def rel_max(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rel_rms(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.sqrt(np.mean(np.abs(a - b) ** 2)) / max(np.sqrt(np.mean(np.abs(b) ** 2)), 1e-300))


def small_obs(n_st=12, n_dumps=24, arm_m=500.0, theta=0.1, lam=1280.0, n_src=3, seed=1):
    obs = synth.config_small(seed=seed, n_st=n_st, n_dumps=n_dumps, arm_m=arm_m, theta=theta,
                             lam=lam, n_src=n_src)
    vis = synth.predict_numpy(obs.u, obs.v, obs.w, obs.l, obs.m, obs.flux)
    return obs, vis
