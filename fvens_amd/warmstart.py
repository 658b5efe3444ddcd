"""Learned warm start workflow on top of the library's network kernels (fvx.MPNN) and batched
solver (fvx.Batch): parameter initialisation, training-set generation from converged batched
solves, training, and warm-started solves with the freestream solve as fallback.

No reference counterpart: the reference starts every solve from the freestream or from a
first-order starter solve (src/utilities/casesolvers.cpp:225-386).
"""
from __future__ import annotations

import time

import numpy as np

from . import fvx

WIDTH, F_IN, F_EDGE, NLAYER, F_OUT = 64, 8, 4, 4, 4


def layout():
    """(name, shape) of the flat parameter vector, matrices [in][out] row-major."""
    L = [("Wenc", (F_IN, WIDTH)), ("benc", (WIDTH,))]
    for l in range(NLAYER):
        L += [(f"Wa{l}", (WIDTH, WIDTH)), (f"Wb{l}", (WIDTH, WIDTH)), (f"We{l}", (F_EDGE, WIDTH)),
              (f"bm{l}", (WIDTH,)), (f"Wu{l}", (WIDTH, WIDTH)), (f"Wv{l}", (WIDTH, WIDTH)),
              (f"bu{l}", (WIDTH,))]
    return L + [("Wdec", (WIDTH, F_OUT)), ("bdec", (F_OUT,))]


def init_params(seed: int = 0, dec_scale: float = 0.01) -> np.ndarray:
    """He-scaled normal weights for the relu layers, residual branches halved, zero biases and a
    small decoder so that the untrained network predicts close to the freestream."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in layout():
        if name.startswith("b"):
            parts.append(np.zeros(shape))
        elif name == "Wenc":
            parts.append(rng.standard_normal(shape) * np.sqrt(1.0 / F_IN))
        elif name.startswith(("Wa", "Wb")):
            parts.append(rng.standard_normal(shape) * np.sqrt(1.0 / WIDTH))
        elif name.startswith("We"):
            parts.append(rng.standard_normal(shape) * 0.5)
        elif name.startswith(("Wu", "Wv")):
            parts.append(rng.standard_normal(shape) * 0.5 * np.sqrt(2.0 / (2 * WIDTH)))
        else:
            parts.append(rng.standard_normal(shape) * dec_scale)
    w = np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)
    assert w.size == fvx.MPNN.nparams()
    return w


def cosine_lr(step: int, total: int, lr0: float, lr_min: float, warmup: int = 50) -> float:
    if step < warmup:
        return lr0 * (step + 1) / warmup
    t = (step - warmup) / max(1, total - warmup)
    return lr_min + 0.5 * (lr0 - lr_min) * (1 + np.cos(np.pi * min(1.0, t)))


def train(net: "fvx.MPNN", targets, mach, aoa_deg, steps: int, graphs_per_step: int,
          lr0: float = 2e-3, lr_min: float = 5e-5, seed: int = 0, log_every: int = 0,
          log=print):
    """Adam training on a resident target set [nsamples, ncell, 4] whose sample i was solved
    at (mach[i], aoa_deg[i]); each step draws `graphs_per_step` samples without replacement."""
    rng = np.random.default_rng(seed)
    mach = np.asarray(mach)
    aoa_deg = np.asarray(aoa_deg)
    ns = len(mach)
    G = min(graphs_per_step, ns, net.max_graphs)
    history = []
    order = rng.permutation(ns)
    pos = 0
    t0 = time.perf_counter()
    for it in range(steps):
        if pos + G > ns:
            order, pos = rng.permutation(ns), 0
        ids = order[pos:pos + G]
        pos += G
        net.set_points(mach[ids], aoa_deg[ids])
        want = log_every > 0 and (it % log_every == 0 or it == steps - 1)
        loss = net.train_step(targets, ids, cosine_lr(it, steps, lr0, lr_min), want_loss=want)
        if want:
            history.append((it, loss))
            log(f"step {it:6d} loss {loss:.4e} ({time.perf_counter() - t0:.1f} s)")
    return history
