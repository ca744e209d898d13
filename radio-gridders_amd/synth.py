"""Seeded synthetic observations for the gridding workloads.

Station layouts, Earth-rotation uvw tracks, point-source skies and W-kernel
table ranges.  Every random choice is drawn from numpy Generators with fixed
seeds (layout 1, sources 2, channel counts 3, pixel sample 4) so runs are
reproducible.  Visibilities follow V = sum_s flux_s exp(-2 pi i (u l_s + v m_s
+ w (n_s - 1))) (the predict.cpp:27-45 convention of the reference).

Config names follow the workload table: C1 plain W-projection (27 stations),
C2 W-towers (512 stations), C4 many channels, C5 stress.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

C_LIGHT = 299792458.0
SITE_LAT_DEG = -26.7  # recorded site latitude of every synthetic layout
FREQ_HZ = 150e6


def layout_y(n_st: int, arm_m: float, rng) -> np.ndarray:
    """Random Y: station k on arm k % 3 (arms at 90, 210, 330 deg), radius U(0, arm_m)."""
    arm = np.arange(n_st) % 3
    r = rng.uniform(0.0, arm_m, n_st)
    ang = np.deg2rad(90.0 + 120.0 * arm)
    return np.stack([r * np.cos(ang), r * np.sin(ang), np.zeros(n_st)], axis=1)  # E, N, U


def layout_radial(n_st: int, scale_m: float, rng) -> np.ndarray:
    """Radius ~ Exp(scale), azimuth U(0, 2 pi)."""
    r = rng.exponential(scale_m, n_st)
    ang = rng.uniform(0.0, 2 * math.pi, n_st)
    return np.stack([r * np.cos(ang), r * np.sin(ang), np.zeros(n_st)], axis=1)


def baselines(pos: np.ndarray):
    a1, a2 = np.triu_indices(len(pos), k=1)
    return a1.astype(np.int32), a2.astype(np.int32), pos[a2] - pos[a1]


def uvw_tracks(enu: np.ndarray, ha_rad: np.ndarray, dec_rad: float,
               lat_rad: float = math.radians(SITE_LAT_DEG)) -> np.ndarray:
    """uvw [n_bl][T][3] in metres for local (E, N, U) baselines."""
    E, N, U = enu[:, 0:1], enu[:, 1:2], enu[:, 2:3]
    X = -math.sin(lat_rad) * N + math.cos(lat_rad) * U
    Y = E
    Z = math.cos(lat_rad) * N + math.sin(lat_rad) * U
    sh, ch = np.sin(ha_rad)[None, :], np.cos(ha_rad)[None, :]
    sd, cd = math.sin(dec_rad), math.cos(dec_rad)
    u = sh * X + ch * Y
    v = -sd * ch * X + sd * sh * Y + cd * Z
    w = cd * ch * X - cd * sh * Y + sd * Z
    return np.stack([u, v, w], axis=2)


def hour_angles(n_dumps: int, hours: float) -> np.ndarray:
    """HA symmetric about transit, n_dumps samples over `hours`."""
    h = np.linspace(-hours / 2, hours / 2, n_dumps) if n_dumps > 1 else np.zeros(1)
    return h * (2 * math.pi / 24.0)


def dec_for_zenith_angle(za_deg: float, lat_deg: float = SITE_LAT_DEG) -> float:
    """Declination that transits at zenith angle za_deg (north of the site)."""
    return math.radians(lat_deg + za_deg)


def filter_first_sample(uvw: np.ndarray, bl_min: float, bl_max: float) -> np.ndarray:
    """Baselines whose first uvw sample has length in [bl_min, bl_max) metres (hdf5.c:104-106)."""
    L = np.hypot(uvw[:, 0, 0], uvw[:, 0, 1])
    return (L >= bl_min) & (L < bl_max)


def flatten(uvw: np.ndarray, freqs: np.ndarray):
    """Flat SoA in wavelengths, order baseline -> time -> frequency."""
    f = np.asarray(freqs, np.float64)
    lam = uvw[:, :, None, :] * f[None, None, :, None] / C_LIGHT  # [bl][T][F][3]
    flat = lam.reshape(-1, 3)
    return (np.ascontiguousarray(flat[:, 0]), np.ascontiguousarray(flat[:, 1]),
            np.ascontiguousarray(flat[:, 2]))


def sources(n_src: int, rng, lm_max: float = 0.05):
    l = rng.uniform(-lm_max, lm_max, n_src)
    m = rng.uniform(-lm_max, lm_max, n_src)
    flux = rng.uniform(0.1, 1.0, n_src)
    return l, m, flux


def predict_numpy(u, v, w, l, m, flux) -> np.ndarray:
    """CPU point-source prediction (small inputs)."""
    out = np.zeros(len(u), np.complex128)
    for lk, mk, fk in zip(l, m, flux):
        nk = math.sqrt(1 - lk * lk - mk * mk)
        ph = u * lk + v * mk + w * (nk - 1)
        out += fk * np.exp(-2j * math.pi * ph)
    return out


def kernel_range(w: np.ndarray, w_step: float, pad: int = 1):
    """W-plane table covering the data: (w_min, planes) on a grid of w_step."""
    lo = math.floor(float(np.min(w)) / w_step) - pad
    hi = math.ceil(float(np.max(w)) / w_step) + pad
    return lo * w_step, hi - lo + 1


@dataclass
class Observation:
    name: str
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    l: np.ndarray
    m: np.ndarray
    flux: np.ndarray
    theta: float
    lam: float
    n_bl: int
    n_dumps: int
    freqs: np.ndarray
    chan_offsets: np.ndarray | None = None
    info: dict = field(default_factory=dict)

    @property
    def grid_size(self) -> int:
        return int(math.floor(self.lam * self.theta))

    @property
    def n_vis(self) -> int:
        return len(self.u)


def observe(name, pos, n_dumps, hours, dec_rad, freqs, n_src, src_seed, theta, lam,
            bl_max=10_000.0, bl_min=0.0, lm_max=0.05) -> Observation:
    a1, a2, enu = baselines(pos)
    uvw = uvw_tracks(enu, hour_angles(n_dumps, hours), dec_rad)
    keep = filter_first_sample(uvw, bl_min, bl_max)
    uvw = uvw[keep]
    u, v, w = flatten(uvw, freqs)
    l, m, flux = sources(n_src, np.random.default_rng(src_seed), lm_max)
    return Observation(name, u, v, w, l, m, flux, theta, lam, int(keep.sum()), n_dumps,
                       np.asarray(freqs, np.float64),
                       info={"stations": len(pos), "site_lat_deg": SITE_LAT_DEG,
                             "dec_deg": math.degrees(dec_rad), "hours": hours,
                             "bl_max_m": bl_max, "freq_hz": [float(f) for f in np.atleast_1d(freqs)]})


def config_c1(seed: int = 1) -> Observation:
    """27 stations on a random Y (10 km arms), 256 dumps over 4 h at dec -30 deg,
    150 MHz, bl_max 10 km, 10 sources, theta 0.1, lambda 20480 (2048^2)."""
    pos = layout_y(27, 10_000.0, np.random.default_rng(seed))
    return observe("C1", pos, 256, 4.0, math.radians(-30.0), [FREQ_HZ], 10, 2, 0.1, 20480.0)


def config_c2(seed: int = 1, n_st: int = 512, n_dumps: int = 64) -> Observation:
    """512 stations, radius ~ Exp(3 km), 64 dumps over 4 h at dec -30 deg, 150 MHz,
    bl_max 10 km, 50 sources, theta 0.1, lambda 20480 (2048^2)."""
    pos = layout_radial(n_st, 3_000.0, np.random.default_rng(seed))
    return observe("C2", pos, n_dumps, 4.0, math.radians(-30.0), [FREQ_HZ], 50, 2, 0.1, 20480.0)


def config_c4(seed: int = 1, n_st: int = 256, n_chan: int = 256, n_dumps: int = 64,
              median: float = 40_000.0, sigma: float = 1.0, count_seed: int = 3,
              f_lo: float = 100e6, f_hi: float = 200e6) -> Observation:
    """Many channels with uneven counts: n_st stations (radius ~ Exp(3 km)), n_dumps
    over 4 h at dec -30 deg, bl_max 10 km (first sample), n_chan channels linear in
    [f_lo, f_hi], channel c keeps round(LogNormal(ln median, sigma)) (seed
    count_seed) of its (baseline, dump) samples, drawn without replacement and kept
    in baseline-track order.  The stream is channel-major (chan_offsets).  10
    sources, theta 0.05, lambda 20480 (1024^2 grids)."""
    pos = layout_radial(n_st, 3_000.0, np.random.default_rng(seed))
    a1, a2, enu = baselines(pos)
    uvw = uvw_tracks(enu, hour_angles(n_dumps, 4.0), math.radians(-30.0))
    uvw = uvw[filter_first_sample(uvw, 0.0, 10_000.0)]
    flat = uvw.reshape(-1, 3)  # baseline -> time
    freqs = np.linspace(f_lo, f_hi, n_chan)
    rng = np.random.default_rng(count_seed)
    counts = np.minimum(np.rint(rng.lognormal(math.log(median), sigma, n_chan)).astype(np.int64),
                        len(flat))
    us, vs, ws, offs = [], [], [], [0]
    for c in range(n_chan):
        pick = np.sort(rng.choice(len(flat), size=int(counts[c]), replace=False))
        sc = freqs[c] / C_LIGHT
        us.append(flat[pick, 0] * sc)
        vs.append(flat[pick, 1] * sc)
        ws.append(flat[pick, 2] * sc)
        offs.append(offs[-1] + int(counts[c]))
    l, m, flux = sources(10, np.random.default_rng(2))
    return Observation("C4", np.concatenate(us), np.concatenate(vs), np.concatenate(ws), l, m, flux,
                       0.05, 20480.0, uvw.shape[0], n_dumps, freqs,
                       chan_offsets=np.asarray(offs, np.int64),
                       info={"stations": n_st, "channels": n_chan, "median": median,
                             "sigma": sigma, "count_seed": count_seed})


def config_c5(seed: int = 1, n_st: int = 512, n_chan: int = 16, n_dumps: int = 128,
              za_deg: float = 60.0, chan_hz: float = 1e6) -> Observation:
    """Stress workload: n_st stations (radius ~ Exp(3 km)), n_dumps over 4 h with the
    phase centre transiting at zenith angle za_deg (|w| up to ~5000 wavelengths),
    bl_max 10 km, n_chan channels chan_hz apart centred on 150 MHz, stream order
    baseline -> time -> channel, 50 sources, theta 0.2, lambda 20480 (4096^2)."""
    pos = layout_radial(n_st, 3_000.0, np.random.default_rng(seed))
    freqs = FREQ_HZ + chan_hz * (np.arange(n_chan) - (n_chan - 1) / 2)
    return observe("C5", pos, n_dumps, 4.0, dec_for_zenith_angle(za_deg), freqs, 50, 2, 0.2,
                   20480.0)


def config_small(seed: int = 1, n_st: int = 8, n_dumps: int = 16, arm_m: float = 600.0,
                 theta: float = 0.1, lam: float = 640.0, n_src: int = 3) -> Observation:
    """Small Y-layout observation for parity tests (N = theta * lambda)."""
    pos = layout_y(n_st, arm_m, np.random.default_rng(seed))
    return observe("small", pos, n_dumps, 4.0, math.radians(-30.0), [FREQ_HZ], n_src, 2, theta,
                   lam, bl_max=1e12)
