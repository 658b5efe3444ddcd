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


def usable(stats, max_rel: float = 1e-2):
    """Indices of training fields worth learning: converged solves, and solves that stopped at
    the step limit within max_rel of their normaliser; non-finite solves are dropped."""
    return [i for i, s in enumerate(stats)
            if s["status"] == 0 or (s["status"] == 4 and s["final_rel"] < max_rel)]


def train(net: "fvx.MPNN", targets, mach, aoa_deg, steps: int, graphs_per_step: int,
          lr0: float = 2e-3, lr_min: float = 5e-5, seed: int = 0, log_every: int = 0,
          log=print, valid=None):
    """Adam training on a resident target set [nsamples, ncell, 4] whose sample i was solved
    at (mach[i], aoa_deg[i]); each step draws `graphs_per_step` samples without replacement
    from `valid` (default: all samples)."""
    rng = np.random.default_rng(seed)
    mach = np.asarray(mach)
    aoa_deg = np.asarray(aoa_deg)
    pool = np.arange(len(mach)) if valid is None else np.asarray(valid)
    ns = len(pool)
    G = min(graphs_per_step, ns, net.max_graphs)
    history = []
    order = pool[rng.permutation(ns)]
    pos = 0
    t0 = time.perf_counter()
    for it in range(steps):
        if pos + G > ns:
            order, pos = pool[rng.permutation(ns)], 0
        ids = order[pos:pos + G]
        pos += G
        net.set_points(mach[ids], aoa_deg[ids])
        want = log_every > 0 and (it % log_every == 0 or it == steps - 1)
        loss = net.train_step(targets, ids, cosine_lr(it, steps, lr0, lr_min), want_loss=want)
        if want:
            history.append((it, loss))
            log(f"step {it:6d} loss {loss:.4e} ({time.perf_counter() - t0:.1f} s)")
    return history


def freestream_states(batch: "fvx.Batch"):
    """[ninst, ncell, 4] device tensor of every instance's freestream state."""
    import torch
    fs = torch.tensor(batch.freestream(), device="cuda")
    return fs[:, None, :].expand(batch.ninst, batch.disc.mesh.nelem, 4).contiguous()


def solve_batch(disc2: "fvx.FlowFV", mach, aoa_deg, u=None, starter: "fvx.FlowFV" = None,
                starter_cfg=None, main_cfg=None, norms=None):
    """Steady solves of the points (mach[i], aoa_deg[i]) in one batch: an optional first-order
    starter solve (the reference's case driver, src/utilities/casesolvers.cpp:225-386) then the
    main solve on disc2. u: initial states (default freestream), updated in place and returned
    with per-instance statistics and the wall time of each phase."""
    import torch
    b2 = fvx.Batch(disc2, mach, aoa_deg)
    if u is None:
        u = freestream_states(b2)
    if norms is not None:
        b2.set_norms(norms)
    out = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if starter is not None:
        b1 = fvx.Batch(starter, mach, aoa_deg)
        if norms is not None:
            b1.set_norms(norms)
        out["starter"] = b1.steady_solve(u, **starter_cfg)
        torch.cuda.synchronize()
    t1 = time.perf_counter()
    out["main"] = b2.steady_solve(u, **main_cfg)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out["t_starter"], out["t_main"] = t1 - t0, t2 - t1
    out["functionals"] = b2.functionals(u)
    return u, out


def generate_targets(net: "fvx.MPNN", disc2, starter, mach, aoa_deg, starter_cfg, main_cfg,
                     chunk: int = 64, log=print):
    """Training set: converged fields of the batched solver for every point, turned into network
    targets [npoints, ncell, 4] (fp32, device). Returns the targets and per-point statistics."""
    import torch
    n = len(mach)
    T = torch.zeros((n, net.N, 4), dtype=torch.float32, device="cuda")
    stats = []
    for c0 in range(0, n, chunk):
        sl = slice(c0, min(n, c0 + chunk))
        b = fvx.Batch(disc2, mach[sl], aoa_deg[sl])
        u0 = freestream_states(b)
        norms = b.residual_norms(u0)
        u, st = solve_batch(disc2, mach[sl], aoa_deg[sl], u=u0, starter=starter,
                            starter_cfg=starter_cfg, main_cfg=main_cfg, norms=norms)
        for s0 in range(sl.start, sl.stop, net.max_graphs):
            ss = slice(s0, min(sl.stop, s0 + net.max_graphs))
            net.set_points(mach[ss], aoa_deg[ss])
            net.targets(u[ss.start - c0:ss.stop - c0], T[ss], ngraph=ss.stop - ss.start)
        for i, s in enumerate(st["main"]):
            stats.append(dict(mach=float(mach[c0 + i]), aoa=float(aoa_deg[c0 + i]),
                              steps=s["steps"], status=s["status"], final_rel=s["final_rel"],
                              starter_steps=st["starter"][i]["steps"] if starter else 0))
        log(f"targets {sl.stop}/{n}: starter {st['t_starter']:.1f} s, main {st['t_main']:.1f} s, "
            f"converged {sum(s['status'] == 0 for s in st['main'])}/{sl.stop - sl.start}")
    return T, stats


def warm_states(net: "fvx.MPNN", mach, aoa_deg, floor: float = 0.05):
    """Predicted conserved states [npoints, ncell, 4] (device, float64)."""
    import torch
    u = torch.zeros((len(mach), net.N, 4), dtype=torch.float64, device="cuda")
    for c0 in range(0, len(mach), net.max_graphs):
        sl = slice(c0, min(len(mach), c0 + net.max_graphs))
        net.set_points(mach[sl], aoa_deg[sl])
        net.warm_state(u[sl], ngraph=sl.stop - sl.start, floor=floor)
    return u
