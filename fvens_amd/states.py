"""Deterministic synthetic flow states for parity tests and the benchmark.

``perturbed_freestream`` is a smooth, non-converged state with realistic gradients: density
and pressure modulated by 5% with sin(2 pi x/L) cos(2 pi y/L), velocity rotated by
2 deg * sin(pi x/L), plus seeded uniform noise of +-1e-3 relative on every variable.
"""
from __future__ import annotations

import numpy as np


def freestream(gamma: float, Minf: float, aoa_deg: float) -> np.ndarray:
    a = np.deg2rad(aoa_deg)
    p = 1.0 / (gamma * Minf ** 2)
    return np.array([1.0, np.cos(a), np.sin(a), p / (gamma - 1.0) + 0.5])


def cell_centres(raw) -> np.ndarray:
    cnt = raw.nnode.astype(np.float64)
    idx = np.where(raw.inpoel >= 0, raw.inpoel, 0)
    xs = raw.coords[idx]                               # (nelem, 4, 2)
    mask = (raw.inpoel >= 0)[..., None]
    return (xs * mask).sum(axis=1) / cnt[:, None]


def perturbed_freestream(raw, gamma=1.4, Minf=0.5, aoa_deg=0.0, L=1.0, amp=0.05,
                         noise=1e-3, seed=20261202) -> np.ndarray:
    x = cell_centres(raw)
    rng = np.random.default_rng(seed)
    mod = 1.0 + amp * np.sin(2 * np.pi * x[:, 0] / L) * np.cos(2 * np.pi * x[:, 1] / L)
    rho = mod
    p = mod / (gamma * Minf ** 2)
    ang = np.deg2rad(aoa_deg) + np.deg2rad(2.0) * np.sin(np.pi * x[:, 0] / L)
    vx, vy = np.cos(ang), np.sin(ang)
    w = np.stack([rho, vx, vy, p], axis=1)
    w *= 1.0 + noise * (2.0 * rng.random(w.shape) - 1.0)
    u = np.empty_like(w)
    u[:, 0] = w[:, 0]
    u[:, 1] = w[:, 0] * w[:, 1]
    u[:, 2] = w[:, 0] * w[:, 2]
    u[:, 3] = w[:, 3] / (gamma - 1.0) + 0.5 * w[:, 0] * (w[:, 1] ** 2 + w[:, 2] ** 2)
    return np.ascontiguousarray(u)
