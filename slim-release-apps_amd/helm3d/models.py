"""Synthetic velocity models and sources of the benchmark configurations.

Inputs only (generated on the host with fixed seeds); the hot path never
depends on how they were made.  Arrays are (nx, ny, nz) in physical-grid
order; flattening for the operator uses MATLAB column-major order (x fastest),
``vec(a) = a.reshape(-1, order='F')``.

  C1  64^3 + 10-pt PML, h 10 m, 5 Hz: Gaussian-smoothed (sigma 8) random field
      scaled to 1500-4500 m/s; unit impulse at (32,32,5).
  C2  200x200x100 + 20-pt PML, h 10 m, 6 Hz: 1500->4000 m/s linear in depth
      times (1 + 0.1 s), s a sigma-8 smoothed field scaled to [-1,1];
      impulse at (100,100,5).
  C3  C2 model, 8 impulses at z=5, y=100, x=round(linspace(20,179,8)).
  C4  8 subproblems, grids drawn in [96^3, 256x256x128] + 20-pt PML, h 10 m,
      frequencies uniform in 3-8 Hz, each with the C2 velocity recipe on its
      own grid and 4-32 impulses on a seeded surface grid (z = 5).
  C5  256x256x128 + 20-pt PML, h 12.5 m: smooth background (C1 recipe with
      sigma 20) plus 12 Gaussian blobs (+-300 m/s, radius 10-25 pts); start
      model = true model smoothed with sigma 20.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Tuple

import numpy as np
from scipy.ndimage import gaussian_filter

SEED_BASE = 20261202


def vec(a: np.ndarray) -> np.ndarray:
    """MATLAB vec(): column-major flatten (x fastest)."""
    return np.asarray(a).reshape(-1, order="F")


def unvec(v: np.ndarray, n) -> np.ndarray:
    return np.asarray(v).reshape(tuple(n), order="F")


@dataclass
class Config:
    name: str
    n_phys: Tuple[int, int, int]
    npml: int
    h: float
    freq: float
    v: np.ndarray                      # physical velocity (m/s), (nx,ny,nz)
    sources: List[Tuple[int, int, int]]  # physical 0-based indices

    @property
    def nt(self):
        return tuple(a + 2 * self.npml for a in self.n_phys)

    @property
    def npml6(self):
        return (self.npml,) * 6


def _scaled_smooth(shape, sigma, rng, lo, hi):
    s = gaussian_filter(rng.standard_normal(shape), sigma=sigma, mode="reflect")
    return lo + (hi - lo) * (s - s.min()) / (s.max() - s.min())


def config_c1(seed: int = SEED_BASE + 1) -> Config:
    rng = np.random.default_rng(seed)
    n = (64, 64, 64)
    v = _scaled_smooth(n, 8, rng, 1500.0, 4500.0)
    return Config("C1", n, 10, 10.0, 5.0, v, [(32, 32, 5)])


def _c2_velocity(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n = (200, 200, 100)
    z = np.arange(n[2], dtype=np.float64)
    vz = 1500.0 + 2500.0 * z / (n[2] - 1)
    s = _scaled_smooth(n, 8, rng, -1.0, 1.0)
    return vz[None, None, :] * (1.0 + 0.1 * s)


def config_c2(seed: int = SEED_BASE + 2) -> Config:
    return Config("C2", (200, 200, 100), 20, 10.0, 6.0, _c2_velocity(seed), [(100, 100, 5)])


def config_c3(seed: int = SEED_BASE + 2) -> Config:
    xs = np.round(np.linspace(20, 179, 8)).astype(int)
    return Config("C3", (200, 200, 100), 20, 10.0, 6.0, _c2_velocity(seed), [(int(x), 100, 5) for x in xs])


def _gradient_velocity(n, rng) -> np.ndarray:
    """C2 recipe on any grid: 1500->4000 m/s linear in depth times
    (1 + 0.1 s), s a sigma-8 smoothed field scaled to [-1,1]."""
    z = np.arange(n[2], dtype=np.float64)
    vz = 1500.0 + 2500.0 * z / max(n[2] - 1, 1)
    s = _scaled_smooth(n, 8, rng, -1.0, 1.0)
    return vz[None, None, :] * (1.0 + 0.1 * s)


def config_c4(seed: int = SEED_BASE + 4, scale: float = 1.0, nsub: int = 8) -> List[Config]:
    """Uneven batched workload: ``nsub`` subproblems with (nx, ny) drawn
    uniformly in [96, 256] and nz in [96, 128] (so every grid lies between
    96^3 and 256x256x128), 20-point PML, h = 10 m, frequency uniform in
    [3, 8] Hz, and 4-32 unit impulses at depth index 5 on distinct nodes of
    a seeded surface grid (spacing 8 points, offset 4).  ``scale`` < 1 shrinks
    every grid, the PML and the source spacing for quick parity runs (the
    number of sources is kept)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nsub):
        n = [int(rng.integers(96, 257)), int(rng.integers(96, 257)), int(rng.integers(96, 129))]
        freq = float(rng.uniform(3.0, 8.0))
        nsrc = int(rng.integers(4, 33))
        sub_seed = int(rng.integers(0, 2 ** 31))
        if scale != 1.0:
            n = [max(12, int(round(a * scale))) for a in n]
        npml = max(4, int(round(20 * min(scale, 1.0))))
        sp = max(2, int(round(8 * min(scale, 1.0))))
        srng = np.random.default_rng(sub_seed)
        v = _gradient_velocity(tuple(n), srng)
        gx = np.arange(sp // 2, n[0] - sp // 2, sp)
        gy = np.arange(sp // 2, n[1] - sp // 2, sp)
        nodes = [(int(x), int(y)) for y in gy for x in gx]
        pick = srng.choice(len(nodes), size=min(nsrc, len(nodes)), replace=False)
        zs = min(5, n[2] - 1)
        srcs = [(nodes[i][0], nodes[i][1], zs) for i in sorted(pick)]
        out.append(Config(f"C4.{k}", tuple(n), npml, 10.0, freq, v, srcs))
    return out


def config_small(n=(32, 32, 24), npml=8, h=10.0, freq=5.0, seed: int = SEED_BASE + 9) -> Config:
    """Reduced C1-style case for quick parity solves."""
    rng = np.random.default_rng(seed)
    v = _scaled_smooth(n, 4, rng, 1500.0, 4500.0)
    return Config("small", tuple(n), npml, h, freq, v, [(n[0] // 2, n[1] // 2, 3)])


def fwi_c5(seed: int = SEED_BASE + 5, n=(256, 256, 128), h: float = 12.5, npml: int = 20,
           freqs=(4.0, 6.0), nblobs: int = 12):
    """FWI workload: 256x256x128 at 12.5 m + 20-point PML; true model = smooth
    1500-4500 m/s background (sigma 20 points) + 12 seeded Gaussian blobs of
    +-300 m/s with radii 10-25 points; start model = true model smoothed with
    sigma 20; 16 sources on a 4x4 surface grid and 64x64 receivers at depth
    index 2; frequencies 4 and 6 Hz; Ricker spectrum with f0 = 10 Hz, t0 = 0.
    Returns (FWIModel in s2/m2, m_true, m_start) with m on the physical grid
    (x fastest)."""
    from .fwi import FWIModel
    rng = np.random.default_rng(seed)
    v = _scaled_smooth(n, 20, rng, 1500.0, 4500.0)
    ax = [np.arange(a, dtype=np.float64) for a in n]
    for _ in range(nblobs):
        c = rng.uniform(0, 1, 3) * np.array(n)
        r = rng.uniform(10.0, 25.0)
        amp = 300.0 * rng.choice([-1.0, 1.0])
        gx = np.exp(-((ax[0] - c[0]) / r) ** 2)
        gy = np.exp(-((ax[1] - c[1]) / r) ** 2)
        gz = np.exp(-((ax[2] - c[2]) / r) ** 2)
        v += amp * gx[:, None, None] * gy[None, :, None] * gz[None, None, :]
    v0 = gaussian_filter(v, sigma=20, mode="reflect")
    src = np.round(np.linspace(n[0] / 8, n[0] * 7 / 8, 4)) * h
    rx = (2 + 4 * np.arange(64)) * h
    model = FWIModel(o=(0.0, 0.0, 0.0), d=(h, h, h), n=n, freq=list(freqs),
                     xsrc=src, ysrc=src, zsrc=np.array([2 * h]),
                     xrec=rx, yrec=rx, zrec=np.array([2 * h]), f0=10.0, t0=0.0, unit="s2/m2")
    return model, vec(1.0 / v ** 2), vec(1.0 / v0 ** 2)


def extend(v: np.ndarray, npml: int) -> np.ndarray:
    """Edge-replicating extension into the PML (opExtension mode 1)."""
    return np.pad(v, npml, mode="edge")


def wavenumber(v_pml: np.ndarray, freq: float, unit: str = "m/s") -> np.ndarray:
    """param2wavenum.m:20-49 (wn only)."""
    c = (2 * np.pi * freq) ** 2
    if unit == "m/s":
        return c * v_pml ** (-2)
    if unit == "s2/m2":
        return c * v_pml
    if unit == "s2/km2":
        return 1e-6 * c * v_pml
    raise ValueError(unit)


def point_source(cfg: Config, src) -> np.ndarray:
    """Unit impulse on the PML grid at physical index src (vec order)."""
    b = np.zeros(cfg.nt, dtype=np.complex128)
    p = cfg.npml
    b[src[0] + p, src[1] + p, src[2] + p] = 1.0
    return vec(b)
