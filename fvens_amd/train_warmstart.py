"""Trains the learned warm start of the inviscid airfoil sweep and measures what it gains.

    python -m fvens_amd.train_warmstart --ntrain 2000 --ntest 200 --out DIR
    python -m fvens_amd.train_warmstart --case bump --ntrain 256 --ntest 64 --tol 1e-10 \
        --maxiter 3000 --steps 1500 --out DIR

The second form trains the same network on the Gaussian bump channel (4,608 cells, Roe,
Green-Gauss, unlimited) over inflow Mach U[0.3, 0.6] and angle U[-2, 2] deg, whose solves reach
1e-10 of the free-stream residual norm.

1. Training set: `ntrain` operating points (Mach ~ U[0.3, 0.75], AoA ~ U[-4, 6] deg from
   std::mt19937_64 seed 2000) on the 98,304-cell NACA 4-digit O-mesh (Roe, weighted least
   squares, Venkatakrishnan K = 5), solved by the batched solver (first-order starter, then the
   second-order solve, FP32 Jacobian and preconditioner) and turned into network targets.
2. Training: Adam on the device kernels, `graphs` fields per step.
3. Held-out A/B on `ntest` points (seed 200): the plain solve (freestream start, first-order
   starter, FP64 Jacobian and preconditioner) against the warm start (network prediction, five
   low-CFL first-order steps, then the main solve, FP32 Jacobian and preconditioner), both to
   `tol` times the freestream residual norm of the point; a warm start that fails falls back to
   the plain solve and is charged for both.
   Batched over all test points, plus single-instance wall times for the first `single`.

Writes DIR/warmstart_k2.npz (parameters) and DIR/warmstart_k2.json (statistics).
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import fvx, meshgen
from . import warmstart as ws
from .points import operating_points

AIRFOIL_BCS = [(2, "slipwall"), (4, "farfield")]
BUMP_BCS = [(2, "slipwall"), (3, "inflowoutflow"), (4, "inflowoutflow")]
# operating-point ranges (Mach, angle of the free stream in degrees) of each case's family
RANGES = {"airfoil": ((0.3, 0.75), (-4.0, 6.0)), "bump": ((0.3, 0.6), (-2.0, 2.0))}


def case_mesh(case):
    """The case's mesh (RCM order for the airfoil, the generator's order for the bump)."""
    if case == "bump":
        return meshgen.mesh_k1()
    raw = meshgen.mesh_k2()
    return raw.permuted(fvx.rcm_permutation(raw))


def case_points(case, seed, count):
    """Seeded operating points of the case's family (std::mt19937_64)."""
    mach, aoa = RANGES[case]
    return operating_points(seed, count, mach=mach, aoa_deg=aoa)


def case_setup(case, gm, sweeps, maxiter, tol):
    """(starter discretisation, main discretisation, FP32 and FP64 preconditioner settings,
    starter settings, main settings) of the case. The bump family is the Gaussian channel of
    tests/inv-gaussianbump (Roe, Green-Gauss, unlimited; starter CFL 25 -> 500, main
    250 -> 2000 from its base.ctrl) over a range of inflow Mach numbers and angles."""
    if case != "bump":
        d1, d2 = discretisations(gm)
        return (d1, d2) + solver_settings(sweeps, maxiter, tol)
    d1 = fvx.FlowFV(gm, fvx.physics_cfg(Minf=0.5, bcs=BUMP_BCS),
                    fvx.numerics_cfg(flux="ROE", gradient="NONE", order2=False))
    d2 = fvx.FlowFV(gm, fvx.physics_cfg(Minf=0.5, bcs=BUMP_BCS),
                    fvx.numerics_cfg(flux="ROE", gradient="GREENGAUSS", recon="NONE"))
    pc32 = dict(pc="sgs", pc_sweeps=sweeps, pc_precision="f32", jac_precision="f32",
                min_relax=0.2)
    pc64 = dict(pc32, pc_precision="f64", jac_precision="f64")
    starter = dict(cflinit=25.0, cflfin=500.0, tol=1e-1, maxiter=300)
    # matrix-free Newton-Krylov with a high CFL ceiling: 15 main steps to 1e-10 where defect
    # correction at the control file's ceiling (2000) needs 40, in the same time
    main = dict(cflinit=250.0, cflfin=1e5, tol=tol, maxiter=maxiter, matrix_free=True)
    return d1, d2, pc32, pc64, starter, main


def case_warm(case):
    """(starting CFL of the warm main solve, first-order pre-steps) of the case: the airfoil's
    predictions need a few low-CFL first-order steps at the trailing edge, the bump's do not."""
    return (1e4, 0) if case == "bump" else (1000.0, WARM_PRE_STEPS)


def discretisations(gm, limiter_param=5.0):
    """First-order starter and second-order operator of the sweep (Roe, weighted least squares,
    Venkatakrishnan with parameter K = limiter_param)."""
    d1 = fvx.FlowFV(gm, fvx.physics_cfg(Minf=0.5, bcs=AIRFOIL_BCS),
                    fvx.numerics_cfg(flux="ROE", gradient="NONE", order2=False))
    d2 = fvx.FlowFV(gm, fvx.physics_cfg(Minf=0.5, bcs=AIRFOIL_BCS),
                    fvx.numerics_cfg(flux="ROE", gradient="LEASTSQUARES",
                                     recon="VENKATAKRISHNAN", limiter_param=limiter_param))
    return d1, d2


def solver_settings(sweeps, maxiter, tol):
    """Preconditioner settings (FP32 / FP64 storage) and the starter / main pseudo-time
    settings. The main solve is matrix-free (the second-order residual as the Krylov operator,
    the first-order Jacobian as the preconditioner): with the Roe flux, defect correction with
    the first-order Jacobian as the operator stalls at 1e-5 .. 1e-1 of the freestream norm on this
    mesh, while the matrix-free Newton-Krylov iteration reaches the limiter's floor (~2e-9)."""
    pc32 = dict(pc="sgs", pc_sweeps=sweeps, pc_precision="f32", jac_precision="f32",
                min_relax=0.2)
    pc64 = dict(pc32, pc_precision="f64", jac_precision="f64")
    # the reference's starter ramp (testcases/naca0012/transonic-sanity-test-muscl.ctrl:
    # CFL 50 -> 1000); a fixed CFL of 1000 from the freestream diverges at high angles of attack
    starter = dict(cflinit=50.0, cflfin=1000.0, tol=1e-5, maxiter=300)
    # the main solve starts at CFL 200 after the starter: from 1000, 8 more of the 64 sweep
    # points diverge or stall (strong shocks / low Mach at high angle of attack)
    main = dict(cflinit=200.0, cflfin=10000.0, tol=tol, maxiter=maxiter, matrix_free=True)
    return pc32, pc64, starter, main


def warm_settings(main, cflinit):
    """Main-solve settings of a warm start: the same solve with its own initial CFL."""
    return dict(main, cflinit=cflinit)


# First-order pseudo-time steps applied to a predicted state before the main solve: the
# prediction is least accurate at the trailing edge, where a full Newton step from it can drive
# the pressure negative; a few low-CFL first-order steps smooth that corner and leave the rest of
# the field as predicted (27 of 64 held-out warm solves failed without them, 7 with them).
WARM_PRE_STEPS, WARM_PRE_CFL = 5, 50.0


def warm_solve(net, d1, d2, mach, aoa, pc32, main, norms, warm_cfl=1000.0,
               pre_steps=WARM_PRE_STEPS):
    """Warm-started solve of the points: prediction, first-order smoothing (pre_steps steps),
    main solve. Returns the states, the main-solve statistics and the time of each phase."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u = ws.warm_states(net, mach, aoa)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pre = [dict(steps=0)]*len(mach)
    if pre_steps > 0:
        b1 = fvx.Batch(d1, mach, aoa)
        b1.set_norms(norms)
        pre = b1.steady_solve(u, cflinit=WARM_PRE_CFL, cflfin=WARM_PRE_CFL, tol=1e-300,
                              maxiter=pre_steps, **pc32)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    _, o = ws.solve_batch(d2, mach, aoa, u=u, main_cfg=dict(warm_settings(main, warm_cfl), **pc32),
                          norms=norms)
    o["t_inference"], o["t_pre"] = t1 - t0, t2 - t1
    o["pre"] = pre
    return u, o


def solve_with_fallback(net, d1, d2, mach, aoa, pc_warm, pc_plain, starter, main, norms,
                        warm_cfl=1000.0):
    """Warm-started solves of the points; every point whose warm solve ends non-finite or
    unconverged is solved again by the plain method (freestream, starter, main solve). Returns
    the final states [npoints, ncell, 4] and per point the path taken and its statistics."""
    u, w = warm_solve(net, d1, d2, mach, aoa, pc_warm, main, norms, warm_cfl)
    out = [dict(path="warm", **w["main"][i]) for i in range(len(mach))]
    failed = [i for i, s in enumerate(w["main"]) if s["status"] != 0]
    if failed:
        uf, fb = ws.solve_batch(d2, mach[failed], aoa[failed], starter=d1,
                                starter_cfg=dict(starter, **pc_plain),
                                main_cfg=dict(main, **pc_plain), norms=norms[failed])
        for k, i in enumerate(failed):
            u[i] = uf[k]
            out[i] = dict(path="fallback", **fb["main"][k])
    return u, out


def generate_targets_warm(net, d1, d2, mach, aoa, pc32, starter, main, chunk, log):
    """Training targets of the points from warm-started solves with the current network (the
    plain solve where a warm start fails): the network speeds up the generation of its own
    next training set."""
    import torch
    n = len(mach)
    T = torch.zeros((n, net.N, 4), dtype=torch.float32, device="cuda")
    stats = []
    for c0 in range(0, n, chunk):
        sl = slice(c0, min(n, c0 + chunk))
        m, a = mach[sl], aoa[sl]
        b = fvx.Batch(d2, m, a)
        norms = b.residual_norms(ws.freestream_states(b))
        t0 = time.perf_counter()
        u, res = solve_with_fallback(net, d1, d2, m, a, pc32, pc32, starter, main, norms)
        torch.cuda.synchronize()
        for s0 in range(0, len(m), net.max_graphs):
            ss = slice(s0, min(len(m), s0 + net.max_graphs))
            net.set_points(m[ss], a[ss])
            net.targets(u[ss], T[sl.start + ss.start:sl.start + ss.stop], ngraph=ss.stop - ss.start)
        for i, r in enumerate(res):
            stats.append(dict(mach=float(m[i]), aoa=float(a[i]), steps=r["steps"],
                              status=r["status"], final_rel=r["final_rel"], path=r["path"]))
        log(f"targets {sl.stop}/{n} (warm-started): {time.perf_counter() - t0:.1f} s, "
            f"fallbacks {sum(r['path'] == 'fallback' for r in res)}, "
            f"converged {sum(r['status'] == 0 for r in res)}/{len(m)}")
    return T, stats


def ab_test(net, d1, d2, mach, aoa, pc32, pc64, starter, main, log, warm_cfl=1000.0,
            chunk=64, pre_steps=WARM_PRE_STEPS):
    """Batched A/B over the points (in batches of `chunk`) plus per-point statistics."""
    pts = []
    tot = dict(t_cold=0.0, t_warm=0.0, t_inference=0.0, t_pre=0.0, t_fallback=0.0, fallbacks=0)
    for c0 in range(0, len(mach), chunk):
        m, a = mach[c0:c0 + chunk], aoa[c0:c0 + chunk]
        b = fvx.Batch(d2, m, a)
        norms = b.residual_norms(ws.freestream_states(b))
        _, cold = ws.solve_batch(d2, m, a, starter=d1, starter_cfg=dict(starter, **pc64),
                                 main_cfg=dict(main, **pc64), norms=norms)
        log(f"A/B points {c0}..{c0 + len(m) - 1}: plain {cold['t_starter'] + cold['t_main']:.2f} s")
        start_ratio = b.residual_norms(ws.warm_states(net, m, a)) / norms
        _, warm = warm_solve(net, d1, d2, m, a, pc32, main, norms, warm_cfl, pre_steps)
        failed = [i for i, s in enumerate(warm["main"]) if s["status"] != 0]
        fb = None
        if failed:
            # fallback: the plain solve of the points whose warm start did not converge
            _, fb = ws.solve_batch(d2, m[failed], a[failed], starter=d1,
                                   starter_cfg=dict(starter, **pc64),
                                   main_cfg=dict(main, **pc64), norms=norms[failed])
        t_w = warm["t_inference"] + warm["t_pre"] + warm["t_main"]
        t_f = (fb["t_starter"] + fb["t_main"]) if fb else 0.0
        log(f"A/B points {c0}..{c0 + len(m) - 1}: warm {t_w:.2f} s + fallback {t_f:.2f} s "
            f"({len(failed)} points)")
        tot["t_cold"] += cold["t_starter"] + cold["t_main"]
        tot["t_warm"] += t_w
        tot["t_inference"] += warm["t_inference"]
        tot["t_pre"] += warm["t_pre"]
        tot["t_fallback"] += t_f
        tot["fallbacks"] += len(failed)
        for i in range(len(m)):
            c, w = cold["main"][i], warm["main"][i]
            fc, fw = cold["functionals"][i], warm["functionals"][i]
            pts.append(dict(mach=float(m[i]), aoa=float(a[i]),
                            cold_steps=cold["starter"][i]["steps"] + c["steps"],
                            cold_status=c["status"],
                            warm_steps=warm["pre"][i]["steps"] + w["steps"],
                            warm_status=w["status"],
                            warm_start_residual_ratio=float(start_ratio[i]),
                            dCL=abs(fw["CL"] - fc["CL"]), dCDp=abs(fw["CDp"] - fc["CDp"]),
                            CL=fc["CL"], CDp=fc["CDp"]))
    out = dict(tot, points=pts)
    out["speedup_batched"] = out["t_cold"] / (out["t_warm"] + out["t_fallback"])
    log(f"A/B batched: plain {out['t_cold']:.2f} s, warm {out['t_warm']:.2f} s "
        f"(+{out['t_fallback']:.2f} s fallback for {out['fallbacks']}), "
        f"speedup {out['speedup_batched']:.2f}")
    return out


def warm_arm(net, d1, d2, mach, aoa, pcw, pc64, starter, main, warm_cfl=1000.0, chunk=64,
             pre_steps=WARM_PRE_STEPS):
    """The warm arm of ab_test alone with preconditioner settings pcw: warm-started solves in
    batches of `chunk`, then the plain FP64 solve of the points whose warm start did not
    converge. Returns the wall time, the fallback count and each point's functionals and
    status (the fallback's where it ran)."""
    t = 0.0
    fallbacks = 0
    pts = []
    for c0 in range(0, len(mach), chunk):
        m, a = mach[c0:c0 + chunk], aoa[c0:c0 + chunk]
        b = fvx.Batch(d2, m, a)
        norms = b.residual_norms(ws.freestream_states(b))
        _, warm = warm_solve(net, d1, d2, m, a, pcw, main, norms, warm_cfl, pre_steps)
        t += warm["t_inference"] + warm["t_pre"] + warm["t_main"]
        res = [dict(status=w["status"], **f) for w, f in zip(warm["main"], warm["functionals"])]
        failed = [i for i, s in enumerate(warm["main"]) if s["status"] != 0]
        if failed:
            _, fb = ws.solve_batch(d2, m[failed], a[failed], starter=d1,
                                   starter_cfg=dict(starter, **pc64),
                                   main_cfg=dict(main, **pc64), norms=norms[failed])
            t += fb["t_starter"] + fb["t_main"]
            for k, i in enumerate(failed):
                res[i] = dict(status=fb["main"][k]["status"], **fb["functionals"][k])
        fallbacks += len(failed)
        pts += res
    return dict(t=t, fallbacks=fallbacks, points=pts)


def single_ab(net, d1, d2, mach, aoa, pc32, pc64, starter, main, log, warm_cfl=1000.0,
              pre_steps=WARM_PRE_STEPS):
    import torch
    rows = []
    for i in range(len(mach)):
        m1, a1 = mach[i:i + 1], aoa[i:i + 1]
        b = fvx.Batch(d2, m1, a1)
        norm = b.residual_norms(ws.freestream_states(b))
        _, c = ws.solve_batch(d2, m1, a1, starter=d1, starter_cfg=dict(starter, **pc64),
                              main_cfg=dict(main, **pc64), norms=norm)
        _, w = warm_solve(net, d1, d2, m1, a1, pc32, main, norm, warm_cfl, pre_steps)
        r = dict(mach=float(m1[0]), aoa=float(a1[0]), t_cold=c["t_starter"] + c["t_main"],
                 t_warm=w["t_inference"] + w["t_pre"] + w["t_main"],
                 t_inference=w["t_inference"],
                 cold_status=c["main"][0]["status"], warm_status=w["main"][0]["status"],
                 cold_steps=c["starter"][0]["steps"] + c["main"][0]["steps"],
                 warm_steps=w["pre"][0]["steps"] + w["main"][0]["steps"],
                 dCL=abs(w["functionals"][0]["CL"] - c["functionals"][0]["CL"]),
                 dCDp=abs(w["functionals"][0]["CDp"] - c["functionals"][0]["CDp"]))
        if r["warm_status"] != 0:
            r["t_warm"] += r["t_cold"]          # fallback: the plain solve after the failure
        rows.append(r)
        log(f"single {i}: M {r['mach']:.3f} aoa {r['aoa']:.2f}: plain {r['t_cold']:.3f} s "
            f"({r['cold_steps']} steps), warm {r['t_warm']:.3f} s ({r['warm_steps']} steps)")
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", choices=["airfoil", "bump"], default="airfoil",
                    help="operating-point family: the 98,304-cell airfoil or the 4,608-cell "
                         "Gaussian bump channel")
    ap.add_argument("--ntrain", type=int, default=2000)
    ap.add_argument("--ntest", type=int, default=200)
    ap.add_argument("--single", type=int, default=8)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--graphs", type=int, default=8)
    ap.add_argument("--label-tol", type=float, default=1e-6)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--maxiter", type=int, default=200)
    ap.add_argument("--sweeps", type=int, default=16)
    ap.add_argument("--chunk", type=int, default=64)
    ap.add_argument("--params", default=None, help="start from these parameters (npz)")
    ap.add_argument("--warm-cfl", type=float, default=None,
                    help="starting CFL of the warm main solve (default: the case's)")
    ap.add_argument("--skip", type=int, default=0,
                    help="use training draws skip .. skip+ntrain-1 of the seed-2000 sequence")
    ap.add_argument("--warm-labels", action="store_true",
                    help="generate the training fields with warm starts from --params")
    ap.add_argument("--from-scratch", action="store_true",
                    help="train from the initial parameters even when --params is given")
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=".")
    a = ap.parse_args()
    import torch

    def log(s):
        print(s, flush=True)

    os.makedirs(a.out, exist_ok=True)
    raw = case_mesh(a.case)
    gm = fvx.Mesh(raw)
    d1, d2, pc32, pc64, starter, main_cfg = case_setup(a.case, gm, a.sweeps, a.maxiter, a.tol)
    name = "warmstart_k2" if a.case == "airfoil" else "warmstart_bump"
    net = fvx.MPNN(gm, 2, a.graphs)
    res = dict(args=vars(a), cells=raw.nelem, nparams=net.nparams())
    if a.ntrain > 0:
        mtr, atr = case_points(a.case, 2000, a.skip + a.ntrain)
        mtr, atr = mtr[a.skip:], atr[a.skip:]
        t0 = time.perf_counter()
        label = dict(main_cfg, tol=a.label_tol)
        if a.warm_labels:
            net.set_params(np.load(a.params)["params"])
            T, st = generate_targets_warm(net, d1, d2, mtr, atr, pc32, starter, label, a.chunk,
                                          log)
        else:
            T, st = ws.generate_targets(net, d2, d1, mtr, atr, dict(starter, **pc32),
                                        dict(label, **pc32), chunk=a.chunk, log=log)
        res["t_data"] = time.perf_counter() - t0
        res["train_points"] = st
        valid = ws.usable(st)
        res["train_fields_used"] = len(valid)
        log(f"training on {len(valid)} of {len(st)} fields (non-finite solves dropped)")
        net.set_params(ws.init_params(0) if a.params is None or a.from_scratch
                       else np.load(a.params)["params"])
        t0 = time.perf_counter()
        res["history"] = ws.train(net, T, mtr, atr, steps=a.steps, graphs_per_step=a.graphs,
                                  lr0=a.lr, log_every=max(1, a.steps // 20), log=log,
                                  valid=valid)
        torch.cuda.synchronize()
        res["t_train"] = time.perf_counter() - t0
        del T
        torch.cuda.empty_cache()
        np.savez(os.path.join(a.out, name + ".npz"), params=net.get_params())
    else:
        net.set_params(np.load(a.params)["params"])
    if a.ntest > 0:
        mte, ate = case_points(a.case, 200, a.ntest)
        wcfl, pre = case_warm(a.case)
        wcfl = a.warm_cfl if a.warm_cfl is not None else wcfl
        res["ab"] = ab_test(net, d1, d2, mte, ate, pc32, pc64, starter, main_cfg, log,
                            wcfl, pre_steps=pre)
        res["single"] = single_ab(net, d1, d2, mte[:a.single], ate[:a.single], pc32, pc64,
                                  starter, main_cfg, log, wcfl, pre_steps=pre)
    with open(os.path.join(a.out, name + ".json"), "w") as fh:
        json.dump(res, fh, default=float)
    log("done")


if __name__ == "__main__":
    main()
