"""Benchmark: second-order FP64 residual throughput (Mcell-residuals/s) on one MI355X.

Workload: inviscid transonic airfoil operator (HLLC, weighted least squares, Venkatakrishnan
K=5, M0.8, alpha 1.25 deg) on the refined 2048x512 NACA 4-digit (t=0.12) O-mesh with the outer
256 rings split into triangles: 1,572,864 cells, ordered along a Hilbert curve of the cell
centres (--order rcm: the reference's RCM order; per-cell results agree to rounding).
A step is one full FlowFV::compute_residual (gradients + limiter pass, face-flux / gather pass
with local time steps) over a synthetic, seeded perturbed free-stream state resident in HBM; the
library picks the flux pass's form (wave-task form on this order). The same workload in the
other ordering is timed after the timed region ("other_order").

Prints one JSON line. It also times pseudo-time solves to a residual tolerance (the second
half of the metric, under "time_to_tol": the bump, the airfoil, the reference's laminar case and
the warm-start A/B), the 64-point batched operating-point sweep ("batch", unless --no-batch) and
the warm-start network's training and inference kernels ("mpnn"). The CPU baseline is this
repository's CPU oracle in its reference-structured OpenMP mode (face loops with atomic
scatters), timed on a bounded sample of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# bytes per cell of the face-flux pass, from the two-pass compulsory-traffic model of the
# residual (408 B/cell for this operator): u read 32 + limited gradient read 64 + residual
# write 32 + time-step write 8 + area 8 + face data 48 B/face x 1.668 faces/cell = 224 B/cell
FLUX_PASS_BYTES_PER_CELL = 224.0
RESIDUAL_BYTES_PER_CELL = 408.0
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="bounded CPU-baseline sample (seconds of CPU work)")
    p.add_argument("--cpu-seconds-1t", type=float, default=8.0,
                   help="bounded one-thread CPU-baseline sample (seconds)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--order", choices=["rcm", "hilbert"], default="hilbert",
                   help="cell ordering of the residual workload")
    p.add_argument("--flux-form", choices=["auto", "quad", "wave"], default="auto",
                   help="flux pass: the library's choice, a quad of lanes per cell, or face "
                        "tasks per wave64 cell group")
    p.add_argument("--no-alt-order", action="store_true",
                   help="skip the residual timing in the other cell ordering")
    p.add_argument("--no-batch", action="store_true",
                   help="skip the batched operating-point sweep")
    p.add_argument("--batch-instances", type=int, default=64)
    p.add_argument("--batch-maxiter", type=int, default=200)
    p.add_argument("--warm-params", default=os.path.join(ROOT, "fvens_amd", "data",
                                                          "warmstart_k2.npz"))
    p.add_argument("--warm-points", type=int, default=32)
    p.add_argument("--warm-params-bump", default=os.path.join(ROOT, "fvens_amd", "data",
                                                               "warmstart_bump.npz"))
    p.add_argument("--no-warm-start", action="store_true",
                   help="skip the warm-start A/B among the time-to-tolerance legs")
    p.add_argument("--only", choices=["residual", "solve", "batch", "mpnn"], default=None,
                   help="run one leg only (for profiling); the JSON line then lacks the others")
    p.add_argument("--no-solve", action="store_true",
                   help="skip the time-to-tolerance pseudo-time solves")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="run only a plain device-to-device copy of known size (the reference "
                        "kernel for scaling HBM counter readings)")
    return p.parse_args()


def calibrate_copy(steps):
    """Plain copy of 2^27 doubles (1 GiB read + 1 GiB write per launch)."""
    import torch
    n = 1 << 27
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8,
                      "copy_GBs": round(2 * n * 8 / dt / 1e9, 1), "launches": steps + 3}))


def residual_leg(raw, pcfg, ncfg, u_host, flux_form, steps, warmup):
    """Residual evaluations of one cell ordering: `steps` timed back to back between two
    device synchronisations, then the median of per-evaluation event times (>= 50) and the
    library's per-kernel event times, each in a loop of its own."""
    import torch
    from fvens_amd import fvx
    mesh = fvx.Mesh(raw)
    disc = fvx.FlowFV(mesh, pcfg, ncfg)
    if flux_form == "wave":
        disc.set_launch("flux_wave", 64, True)
    elif flux_form == "quad":
        disc.set_launch("flux", 256, True)
    form = "wave" if disc.get_launch("flux_wave")[0] == 64 else "quad"
    u = torch.tensor(u_host, device="cuda")
    r = torch.zeros_like(u)
    dt = torch.zeros(raw.nelem, dtype=torch.float64, device="cuda")
    for _ in range(warmup):
        disc.compute_residual(u, r, True, dt)
    # the timed region: exactly `steps` residual evaluations, nothing else recorded
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        disc.compute_residual(u, r, True, dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # median over >= 50 evaluations, each bracketed by events on the launch stream
    nmed = max(50, steps)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(nmed)]
    for e0, e1 in evs:
        e0.record()
        disc.compute_residual(u, r, True, dt)
        e1.record()
    torch.cuda.synchronize()
    med_ms = statistics.median(e0.elapsed_time(e1) for e0, e1 in evs)
    # per-kernel device times from the library's own event pairs, in a loop of their own
    disc.profile_enable(True)
    for _ in range(nmed):
        disc.compute_residual(u, r, True, dt)
    flux_ms, flux_n = disc.profile_read("flux")
    grad_ms, grad_n = disc.profile_read("gradients")
    disc.profile_enable(False)
    assert torch.isfinite(r).all().item() and torch.isfinite(dt).all().item()
    return {"form": form, "faces": int(mesh.naface), "nmed": nmed,
            "ms_per_step": (t1 - t0) * 1e3 / steps,
            "value": raw.nelem * steps / (t1 - t0) / 1e6, "median_ms": med_ms,
            "flux_s": flux_ms / flux_n / 1e3, "grad_s": grad_ms / grad_n / 1e3}


def build_case(order="rcm"):
    import numpy as np
    from fvens_amd import fvx, meshgen, states
    raw = meshgen.mesh_k2l()
    perm = fvx.rcm_permutation(raw) if order == "rcm" else fvx.hilbert_permutation(raw)
    raw = raw.permuted(perm)
    bcs = [(2, "slipwall"), (4, "farfield")]
    pcfg = fvx.physics_cfg(Minf=0.8, aoa_deg=1.25, bcs=bcs)
    ncfg = fvx.numerics_cfg(flux="HLLC", gradient="LEASTSQUARES", recon="VENKATAKRISHNAN",
                            limiter_param=5.0)
    u = states.perturbed_freestream(raw, Minf=0.8, aoa_deg=1.25, L=1.0)
    return raw, bcs, pcfg, ncfg, np.ascontiguousarray(u)


BUMP_BCS = [(2, "slipwall"), (3, "inflowoutflow"), (4, "inflowoutflow")]
AIRFOIL_BCS = [(2, "slipwall"), (4, "farfield")]


def _solve_case(name, order="rcm"):
    """Mesh, physics and (first-order, second-order) numerics of a time-to-tolerance case."""
    from fvens_amd import fvx, meshgen
    if name == "bump":
        raw = meshgen.mesh_k1()
        phys = dict(Minf=0.5, aoa_deg=0.0, bcs=BUMP_BCS)
        n1 = dict(flux="ROE", gradient="NONE", order2=False)
        n2 = dict(flux="ROE", gradient="GREENGAUSS", recon="NONE")
        # tests/inv-gaussianbump/base.ctrl: starter CFL 25..500, main CFL 250..2000
        ramps = ((25.0, 500.0), (250.0, 2000.0))
    else:
        raw = meshgen.mesh_k2()
        phys = dict(Minf=0.5, aoa_deg=1.25, bcs=AIRFOIL_BCS)
        n1 = dict(flux="HLLC", gradient="NONE", order2=False)
        n2 = dict(flux="HLLC", gradient="LEASTSQUARES", recon="VENKATAKRISHNAN",
                  limiter_param=5.0)
        ramps = ((1000.0, 1000.0), (1000.0, 10000.0))
    raw = raw.permuted(fvx.rcm_permutation(raw) if order == "rcm" else fvx.hilbert_permutation(raw))
    return raw, phys, n1, n2, ramps


PC_NAMES = {"sgs": "multicolour SGS", "ilu0": "multicolour block ILU(0)",
            "sgs_level": "natural-order SGS", "ilu0_level": "natural-order block ILU(0)",
            "bjacobi_ilu0": "block Jacobi + ILU(0)"}


def gpu_time_to_tol(name, tol, precision, sweeps=16, pc="sgs", subdomain=256, order="rcm",
                    pc_precision=None):
    """Pseudo-time solve from free stream (first-order starter, then the main solve) until
    sqrt(sum r_E^2 area) <= tol x the same norm at the free-stream state."""
    import numpy as np
    import torch
    from fvens_amd import fvx
    raw, phys, n1, n2, ramps = _solve_case(name, order)
    mesh = fvx.Mesh(raw)
    phys_cfg = fvx.physics_cfg(**phys)
    d1 = fvx.FlowFV(mesh, phys_cfg, fvx.numerics_cfg(**n1))
    d2 = fvx.FlowFV(mesh, phys_cfg, fvx.numerics_cfg(**n2))
    uinf = torch.tensor(np.tile(d1.freestream(), (raw.nelem, 1)), device="cuda")
    r = torch.zeros_like(uinf)
    dt = torch.zeros(raw.nelem, dtype=torch.float64, device="cuda")
    d2.compute_residual(uinf, r, True, dt)
    _, _, area, _ = mesh.faces()
    norm0 = float(np.sqrt((r[:, 3].cpu().numpy() ** 2 * area).sum()))
    kw = dict(min_relax=0.2, pc=pc, pc_sweeps=sweeps, pc_precision=pc_precision or precision,
              jac_precision=precision, pc_subdomain_cells=subdomain, raise_on_failure=False)
    u = uinf.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s1 = d1.steady_solve(u, cflinit=ramps[0][0], cflfin=ramps[0][1], tol=1e-1 if name == "bump"
                         else 1e-6, maxiter=300, **kw)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    s2 = d2.steady_solve(u, cflinit=ramps[1][0], cflfin=ramps[1][1], tol=tol, maxiter=3000,
                         fixed_norm=norm0, **kw)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    f = d2.functionals(u, 2)
    return dict(seconds=round(t2 - t0, 4), main_seconds=round(t2 - t1, 4),
                starter_steps=s1["steps"], steps=s2["steps"], lin_iters=s2["lin_iters"],
                final_rel=s2["final_rel"], converged=bool(s2["converged"]),
                CL=f["CL"], CDp=f["CDp"], precision=precision,
                pc_precision=pc_precision or precision,
                pc=f"{PC_NAMES[pc]} x{sweeps}"
                   + (f" ({subdomain}-cell subdomains)" if pc == "bjacobi_ilu0" else ""))


def viscous_time_to_tol(precision, tol=1e-10, oracle=False):
    """The reference's laminar NACA0012 case (testcases/visc-naca0012/laminar-implicit.ctrl,
    13,156-cell hybrid mesh, Roe + least squares, adiabatic wall, M 0.5, Re 5000): first-order
    starter (CFL 200 -> 1000 to 1e-1) then the main solve (CFL 500 -> 5000) to tol x the first
    main-step norm. oracle=True times the CPU oracle (natural-order ILU(0), one thread)."""
    import numpy as np
    from fvens_amd import gmsh
    mesh_file = os.path.join(ROOT, "tests", "golden", "NACA0012_lam_hybrid_1.msh")
    raw = gmsh.read_gmsh2(mesh_file)
    bcs = [(4, "inflowoutflow"), (2, "adiabaticwall", (0.0, 0.0))]
    phys = dict(Minf=0.5, aoa_deg=0.0, Tinf=288.15, Reinf=5000.0, Pr=0.72, viscous=True, bcs=bcs)
    if oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        m = po.Mesh(raw)
        d1 = po.Disc(m, po.make_config(**phys, flux="ROE", gradient="LEASTSQUARES", order2=False))
        d2 = po.Disc(m, po.make_config(**phys, flux="ROE", gradient="LEASTSQUARES", recon="NONE"))
        u = np.tile(d1.freestream(), (raw.nelem, 1))
        t0 = time.perf_counter()
        u, s1 = d1.steady_solve(u, cflinit=200.0, cflfin=1000.0, tol=1e-1, maxiter=50,
                                fullupdate=1, min_relax=0.2, pc="ilu0")
        u, s2 = d2.steady_solve(u, cflinit=500.0, cflfin=5000.0, tol=tol, maxiter=400,
                                fullupdate=1, min_relax=0.2, pc="ilu0")
        f = d2.functionals(u, 2)
        return dict(seconds=round(time.perf_counter() - t0, 3), steps=s1["steps"] + s2["steps"],
                    converged=bool(s2["converged"]), CL=f["CL"], CDp=f["CDp"], CDsf=f["CDsf"],
                    threads=1, pc="natural-order block ILU(0)")
    import torch
    from fvens_amd import fvx
    gm = fvx.Mesh(raw)
    d1 = fvx.FlowFV(gm, fvx.physics_cfg(**phys),
                    fvx.numerics_cfg(flux="ROE", gradient="LEASTSQUARES", order2=False))
    d2 = fvx.FlowFV(gm, fvx.physics_cfg(**phys),
                    fvx.numerics_cfg(flux="ROE", gradient="LEASTSQUARES", recon="NONE"))
    u = torch.tensor(np.tile(d1.freestream(), (raw.nelem, 1)), device="cuda")
    kw = dict(pc="sgs", pc_sweeps=16, pc_precision=precision, jac_precision=precision,
              min_relax=0.2, fullupdate=True, raise_on_failure=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s1 = d1.steady_solve(u, cflinit=200.0, cflfin=1000.0, tol=1e-1, maxiter=50, **kw)
    s2 = d2.steady_solve(u, cflinit=500.0, cflfin=5000.0, tol=tol, maxiter=400, **kw)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    f = d2.functionals(u, 2)
    return dict(seconds=round(t, 4), starter_steps=s1["steps"], steps=s2["steps"],
                lin_iters=s2["lin_iters"], final_rel=s2["final_rel"],
                converged=bool(s2["converged"]), CL=f["CL"], CDp=f["CDp"], CDsf=f["CDsf"],
                precision=precision, pc="multicolour SGS x16")


def cpu_time_to_tol(name, tol):
    """The oracle's FP64 implicit solve of the same case: natural-order block ILU(0), the
    reference's bjacobi + ILU on one rank (tests/inv-gaussianbump/ibump.solverc)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    raw, phys, n1, n2, ramps = _solve_case(name)
    m = po.Mesh(raw)
    d1 = po.Disc(m, po.make_config(**phys, **n1))
    d2 = po.Disc(m, po.make_config(**phys, **n2))
    u = np.tile(d1.freestream(), (raw.nelem, 1))
    r, _ = d2.residual(u)
    norm0 = float(np.sqrt((r[:, 3] ** 2 * m.area).sum()))
    t0 = time.perf_counter()
    u, s1 = d1.steady_solve(u, cflinit=ramps[0][0], cflfin=ramps[0][1], tol=1e-1, maxiter=300,
                            fullupdate=0, min_relax=0.2, pc="ilu0")
    u, s2 = d2.steady_solve(u, cflinit=ramps[1][0], cflfin=ramps[1][1], tol=tol, maxiter=3000,
                            fullupdate=0, min_relax=0.2, pc="ilu0", fixed_norm=norm0)
    f = d2.functionals(u, 2)
    return dict(seconds=round(time.perf_counter() - t0, 3), steps=s2["steps"],
                lin_iters=s2["lin_iters"], final_rel=s2["rel"], converged=bool(s2["converged"]),
                CL=f["CL"], CDp=f["CDp"], threads=1, pc="natural-order block ILU(0)")


def batch_sweep(ninst, maxiter, sweeps=16, tol=1e-8, pc_precision="f32"):
    """Batched operating-point sweep on the 98,304-cell airfoil: ninst seeded points
    (std::mt19937_64 seed 4242, Mach U[0.3, 0.75], AoA U[-4, 6] deg), Roe + WLS +
    Venkatakrishnan K=5, first-order starter then the second-order matrix-free solve, FP32
    Jacobian and preconditioner, FP64 residual, retirement at tol x each point's freestream
    residual norm."""
    import numpy as np
    import torch
    from fvens_amd import fvx, meshgen, train_warmstart as tw, warmstart as ws
    from fvens_amd.points import operating_points
    raw = meshgen.mesh_k2()
    raw = raw.permuted(fvx.rcm_permutation(raw))
    gm = fvx.Mesh(raw)
    d1, d2 = tw.discretisations(gm)
    pc32, _, starter, main = tw.solver_settings(sweeps, maxiter, tol)
    pc32 = dict(pc32, pc_precision=pc_precision)
    mach, aoa = operating_points(4242, ninst)
    b = fvx.Batch(d2, mach, aoa)
    norms = b.residual_norms(ws.freestream_states(b))
    u, o = ws.solve_batch(d2, mach, aoa, starter=d1, starter_cfg=dict(starter, **pc32),
                          main_cfg=dict(main, **pc32), norms=norms)
    wall = o["t_starter"] + o["t_main"]
    st1 = [s["steps"] for s in o["starter"]]
    st2 = [s["steps"] for s in o["main"]]
    status = [s["status"] for s in o["main"]]
    cell_iters = raw.nelem * (sum(st1) + sum(st2))
    del u
    torch.cuda.empty_cache()
    return {
        "case": f"{ninst} operating points on the 98,304-cell airfoil, Roe, WLS, "
                f"Venkatakrishnan K=5, FP32 Jacobian + multicolour SGS x{sweeps} "
                f"({pc_precision.upper()} factors), FP64 residual",
        "value": round(cell_iters / wall / 1e6, 2), "unit": "Mcell-iterations/s (aggregate)",
        "seconds": round(wall, 3), "starter_seconds": round(o["t_starter"], 3),
        "tol": tol, "maxiter": maxiter,
        "converged": int(sum(1 for x in status if x == 0)),
        "not_converged": int(sum(1 for x in status if x == 4)),
        "nonfinite": int(sum(1 for x in status if x == 3)),
        "starter_steps": st1, "main_steps": st2, "status": status,
        "final_rel": [float("%.3e" % s["final_rel"]) for s in o["main"]],
        "lin_iters": [s["lin_iters"] for s in o["main"]],
    }


def warm_start_ab(params_path, ntest, nsingle, maxiter, tol=1e-8, sweeps=16, case="airfoil"):
    """A/B of the learned warm start on held-out points (seed 200): plain FP64 solve from the
    freestream with the first-order starter against the network's prediction with FP32
    Jacobian + preconditioner, both to tol x the freestream residual norm, with fallback."""
    import numpy as np
    from fvens_amd import fvx, train_warmstart as tw
    raw = tw.case_mesh(case)
    gm = fvx.Mesh(raw)
    d1, d2, pc32, pc64, starter, main = tw.case_setup(case, gm, sweeps, maxiter, tol)
    net = fvx.MPNN(gm, 2, 16)
    net.set_params(np.load(params_path)["params"])
    mach, aoa = tw.case_points(case, 200, ntest)
    quiet = lambda s: None  # noqa: E731
    wcfl, pre = tw.case_warm(case)
    ab = tw.ab_test(net, d1, d2, mach, aoa, pc32, pc64, starter, main, quiet, wcfl,
                    pre_steps=pre)
    single = tw.single_ab(net, d1, d2, mach[:nsingle], aoa[:nsingle], pc32, pc64, starter,
                          main, quiet, wcfl, pre_steps=pre)
    # the same warm arm with bfloat16 preconditioner factors over the FP32 Jacobian
    w16 = tw.warm_arm(net, d1, d2, mach, aoa, dict(pc32, pc_precision="bf16"), pc64, starter,
                      main, wcfl, pre_steps=pre)
    pts = ab["points"]
    ok = [p for p in pts if p["cold_status"] == 0 and p["warm_status"] == 0]
    ok16 = [(p, q) for p, q in zip(pts, w16["points"]) if p["cold_status"] == 0
            and q["status"] == 0]
    return {
        "case": (f"{ntest} held-out points (seed 200) on the 98,304-cell airfoil, Roe, WLS, "
                 f"Venkatakrishnan K=5" if case == "airfoil" else
                 f"{ntest} held-out points (seed 200: inflow Mach U[0.3, 0.6], angle U[-2, 2] "
                 f"deg) of the 4,608-cell Gaussian bump channel, Roe, Green-Gauss, unlimited")
                + f"; tol {tol:g} x freestream residual norm",
        "batched": {"plain_fp64_seconds": round(ab["t_cold"], 3),
                    "warm_fp32_seconds": round(ab["t_warm"] + ab["t_fallback"], 3),
                    "speedup": round(ab["speedup_batched"], 3), "fallbacks": ab["fallbacks"],
                    "warm_phases_seconds": {k: round(ab[k], 3) for k in
                                            ("t_inference", "t_pre", "t_fallback")},
                    "plain_steps": [p["cold_steps"] for p in pts],
                    "warm_steps": [p["warm_steps"] for p in pts],
                    "both_converged": len(ok),
                    "max_dCL": max([p["dCL"] for p in ok], default=None),
                    "max_dCDp": max([p["dCDp"] for p in ok], default=None),
                    "warm_bf16_pc_seconds": round(w16["t"], 3),
                    "speedup_bf16_pc": round(ab["t_cold"] / w16["t"], 3),
                    "fallbacks_bf16_pc": w16["fallbacks"],
                    "max_dCL_bf16_pc": max([abs(q["CL"] - p["CL"]) for p, q in ok16],
                                           default=None)},
        "single": [{k: (float("%.4g" % v) if isinstance(v, float) else v) for k, v in r.items()}
                   for r in single],
        "single_speedup": round(sum(r["t_cold"] for r in single)
                                / max(1e-9, sum(r["t_warm"] for r in single)), 3),
    }


def mpnn_timing(graphs, steps):
    """Device time of one training step (forward, backward, Adam) and of one inference pass of
    the warm-start network over `graphs` fields of the 98,304-cell airfoil graph (synthetic
    targets), measured with HIP events on the launch stream."""
    import numpy as np
    import torch
    from fvens_amd import fvx, meshgen, warmstart as ws
    raw = meshgen.mesh_k2()
    raw = raw.permuted(fvx.rcm_permutation(raw))
    gm = fvx.Mesh(raw)
    net = fvx.MPNN(gm, 2, graphs)
    net.set_params(ws.init_params(0))
    rng = np.random.default_rng(0)
    T = torch.tensor((rng.standard_normal((graphs, raw.nelem, 4)) * 0.1).astype(np.float32),
                     device="cuda")
    net.set_points(np.linspace(0.3, 0.75, graphs), np.linspace(-4, 6, graphs))
    ids = np.arange(graphs)
    y = torch.zeros((graphs, raw.nelem, 4), dtype=torch.float32, device="cuda")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    for _ in range(3):
        net.train_step(T, ids, 1e-4, want_loss=False)
        net.predict(y)
    ev[0].record()
    for _ in range(steps):
        net.train_step(T, ids, 1e-4, want_loss=False)
    ev[1].record()
    ev[2].record()
    for _ in range(steps):
        net.predict(y)
    ev[3].record()
    torch.cuda.synchronize()
    tr = ev[0].elapsed_time(ev[1]) / steps
    inf = ev[2].elapsed_time(ev[3]) / steps
    rows = graphs * raw.nelem
    return {"case": f"4-layer width-64 message passing over {graphs} fields of the "
                    f"98,304-cell airfoil graph ({rows} rows), bf16 MFMA, fp32 accumulation",
            "train_step_ms": round(tr, 4), "inference_ms": round(inf, 4),
            "train_rows_per_s": round(rows / tr * 1e3), "infer_rows_per_s": round(rows / inf * 1e3)}


def host_cpu():
    """CPU model, sockets x cores per socket, threads per core, logical CPUs and the CPUs
    this process may run on."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"),
            "threads_per_core": info.get("Thread(s) per core"), "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0))}


def native_oracle():
    """Builds the CPU oracle with -O3 -march=native for this machine (a few seconds) in a
    temporary directory and loads it; falls back to the portable -O3 build."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    out = tempfile.mkdtemp(prefix="fvo_native_")
    p = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "native",
                        f"NATIVE_DIR={out}"], capture_output=True, text=True)
    if p.returncode == 0:
        po.use_library(os.path.join(out, "liboracle.so"))
        return po, "g++ -O3 -march=native -fopenmp"
    return po, "g++ -O3 -fopenmp (portable build; the native build failed)"


def cpu_baseline(po, flags, raw, bcs, u, budget_s, budget_1t):
    """The oracle's reference-structured OpenMP residual (face loops with atomic scatters) on
    the same operator and state: all the host cores this process may use, then one thread."""
    m = po.Mesh(raw)
    d = po.Disc(m, po.make_config(Minf=0.8, aoa_deg=1.25, flux="HLLC", gradient="LEASTSQUARES",
                                  recon="VENKATAKRISHNAN", limiter_param=5.0, bcs=bcs))
    threads = len(os.sched_getaffinity(0))

    def run(nthreads, budget, nmax):
        po.set_threads(nthreads)
        d.residual(u, parallel=1)                   # warm-up
        times = []
        while sum(times) < budget and len(times) < nmax:
            t0 = time.perf_counter()
            d.residual(u, parallel=1)
            times.append(time.perf_counter() - t0)
        return times

    tall = run(threads, budget_s, 50)
    t1 = run(1, budget_1t, 10)
    po.set_threads(threads)
    med = statistics.median(tall)
    med1 = statistics.median(t1)
    return dict(value=raw.nelem / med / 1e6, unit="Mcell-residuals/s", cores=threads,
                kind="port",
                sample=f"median of {len(tall)} full residual evaluations of the same "
                       f"{raw.nelem:,}-cell operator and state ({sum(tall):.1f} s), oracle "
                       f"OpenMP face loops with atomic scatters, {threads} threads "
                       f"(OMP_PLACES/OMP_PROC_BIND as set), {flags}",
                one_thread={"value": raw.nelem / med1 / 1e6, "evaluations": len(t1),
                            "seconds": round(sum(t1), 2)},
                host=host_cpu())


def time_to_tol(a):
    """Second half of the metric: pseudo-time solves to a residual tolerance on the device,
    with the CPU oracle beside them unless --no-cpu-baseline."""
    # SGS sweeps per application: the fastest of 10, 12, 16, 20 and 24 measured on each case
    bump64 = gpu_time_to_tol("bump", 1e-10, "f64", sweeps=12)
    bump32 = gpu_time_to_tol("bump", 1e-10, "f32", sweeps=12)
    air64 = gpu_time_to_tol("airfoil", 1e-8, "f64", sweeps=24)
    air32 = gpu_time_to_tol("airfoil", 1e-8, "f32", sweeps=24)
    # bfloat16 preconditioner factors over the FP32 Jacobian
    bump16 = gpu_time_to_tol("bump", 1e-10, "f32", sweeps=12, pc_precision="bf16")
    air16 = gpu_time_to_tol("airfoil", 1e-8, "f32", sweeps=24, pc_precision="bf16")
    tt = {
        "criterion": "sqrt(sum r_energy^2 area) <= tol x the same norm at the free stream",
        "bump_1e-10": {
            "case": "Gaussian bump 96x32 hybrid, 4,608 cells, M0.5, Roe, Green-Gauss, "
                    "unlimited; first-order starter then main solve",
            "gpu_fp64": bump64, "gpu_fp32_jacobian_and_pc": bump32,
            "fp32_speedup": round(bump64["seconds"] / bump32["seconds"], 3),
            "gpu_fp32_jacobian_bf16_pc": bump16,
            "bf16_pc_speedup": round(bump64["seconds"] / bump16["seconds"], 3),
            "warm_start": None},
        "airfoil_1e-8": {
            "case": "NACA 4-digit O-mesh 512x128, 98,304 cells, M0.5, alpha 1.25, HLLC, WLS, "
                    "Venkatakrishnan K=5 (the limiter stalls the residual near 3e-9, so "
                    "1e-8 is the airfoil target)",
            "gpu_fp64": air64, "gpu_fp32_jacobian_and_pc": air32,
            "fp32_speedup": round(air64["seconds"] / air32["seconds"], 3),
            "CL_CDp_fp32_vs_fp64": [abs(air32["CL"] / air64["CL"] - 1),
                                    abs(air32["CDp"] / air64["CDp"] - 1)],
            "gpu_fp32_jacobian_bf16_pc": air16,
            "bf16_pc_speedup": round(air64["seconds"] / air16["seconds"], 3),
            "CL_CDp_bf16_pc_vs_fp64": [abs(air16["CL"] / air64["CL"] - 1),
                                       abs(air16["CDp"] / air64["CDp"] - 1)]},
    }
    if not a.no_cpu_baseline:
        tt["bump_1e-10"]["cpu_oracle_fp64"] = cpu_time_to_tol("bump", 1e-10)
        tt["bump_1e-10"]["gpu_vs_cpu"] = round(
            tt["bump_1e-10"]["cpu_oracle_fp64"]["seconds"] / bump64["seconds"], 2)
    v64 = viscous_time_to_tol("f64")
    v32 = viscous_time_to_tol("f32")
    tt["viscous_naca0012_1e-10"] = {
        "case": "the reference's laminar NACA0012 case (13,156-cell hybrid mesh, Roe, least "
                "squares, adiabatic wall, M0.5, Re 5000), starter then main solve to 1e-10",
        "gpu_fp64": v64, "gpu_fp32_jacobian_and_pc": v32,
        "fp32_speedup": round(v64["seconds"] / v32["seconds"], 3),
        "CL_CDp_CDsf_fp32_vs_fp64": [abs(v32["CL"] - v64["CL"]),
                                     abs(v32["CDp"] / v64["CDp"] - 1),
                                     abs(v32["CDsf"] / v64["CDsf"] - 1)]}
    if not a.no_cpu_baseline:
        vo = viscous_time_to_tol("f64", oracle=True)
        tt["viscous_naca0012_1e-10"]["cpu_oracle_fp64"] = vo
        tt["viscous_naca0012_1e-10"]["gpu_vs_cpu"] = round(vo["seconds"] / v64["seconds"], 2)
    if a.no_warm_start:
        tt["warm_start_bump_1e-10"] = {"skipped": "--no-warm-start"}
    elif os.path.exists(a.warm_params_bump):
        # the metric's 1e-10 A/B: the bump family, whose solves reach 1e-10 of the free-stream
        # norm (the airfoil's floor is ~3e-9 of it, the resolution of an FP64 state)
        tt["warm_start_bump_1e-10"] = warm_start_ab(a.warm_params_bump, a.warm_points, 4, 3000,
                                                    tol=1e-10, case="bump")
    else:
        tt["warm_start_bump_1e-10"] = {"skipped": "no trained parameters at "
                                                  + a.warm_params_bump}
    if a.no_warm_start:
        tt["warm_start_airfoil"] = {"skipped": "--no-warm-start"}
    elif os.path.exists(a.warm_params):
        tt["warm_start_airfoil"] = warm_start_ab(a.warm_params, a.warm_points, 4,
                                                 a.batch_maxiter)
    else:
        tt["warm_start_airfoil"] = {"skipped": "no trained parameters at " + a.warm_params}
    return tt


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(the engine runs on a single GPU; domain sharding is not implemented)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from fvens_amd import fvx
    assert torch.cuda.is_available(), "no GPU"
    if a.calibrate_copy:
        calibrate_copy(a.steps)
        return
    if a.only == "batch":
        print(json.dumps({"batch": batch_sweep(a.batch_instances, a.batch_maxiter)}), flush=True)
        return
    if a.only == "mpnn":
        print(json.dumps({"mpnn": mpnn_timing(8, min(a.steps, 50))}), flush=True)
        return
    if a.only == "solve":
        print(json.dumps({"time_to_tol": time_to_tol(a)}), flush=True)
        return
    if a.only == "residual":
        a.no_batch = a.no_solve = True
    raw, bcs, pcfg, ncfg, u_host = build_case(a.order)
    res = residual_leg(raw, pcfg, ncfg, u_host, a.flux_form, a.steps, a.warmup)
    ms_per_step, value, med_ms = res["ms_per_step"], res["value"], res["median_ms"]
    flux_avg_s, grad_avg_s = res["flux_s"], res["grad_s"]
    achieved = FLUX_PASS_BYTES_PER_CELL * raw.nelem / flux_avg_s / 1e9
    form = res["form"]
    traffic = None
    tf = os.path.join(ROOT, "profiles", "k_flux_traffic.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            tj = json.load(fh)
        if (tj.get("order") == a.order and tj.get("cells") == raw.nelem
                and tj.get("form") == form):
            traffic = tj.get("hbm_bytes_per_launch")
    out = {
        "metric": "Mcell-residuals/s; time to 1e-10 residual: warm-start+FP32 PC vs plain FP64",
        "value": round(value, 3),
        "unit": "Mcell-residuals/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "inviscid transonic airfoil residual, NACA 4-digit t=0.12 O-mesh "
                               f"2048x512 (outer 256 rings split), {a.order.upper()} order, "
                               + ("wave-task" if form == "wave" else "quad-per-cell")
                               + " flux pass",
                   "cells": raw.nelem, "faces": res["faces"], "flux": "HLLC",
                   "gradient": "weighted least squares", "limiter": "Venkatakrishnan K=5",
                   "mach": 0.8, "aoa_deg": 1.25, "state": "seeded perturbed free stream",
                   "parallelism": "single GPU"},
        "median_ms_per_evaluation": round(med_ms, 5),
        "median_value": round(raw.nelem / med_ms / 1e3, 3),
        "kernels_ms": {"k_gradients": round(grad_avg_s * 1e3, 5),
                       "k_flux": round(flux_avg_s * 1e3, 5), "evaluations": res["nmed"]},
        "roofline": {"bound": "hbm", "kernel": "k_flux_wave" if form == "wave" else "k_flux",
                     "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_cell": FLUX_PASS_BYTES_PER_CELL},
        "residual_roofline_frac": round(RESIDUAL_BYTES_PER_CELL * raw.nelem
                                        / ((flux_avg_s + grad_avg_s) * 1e9) / HBM_PEAK_GBS, 4),
    }
    if not a.no_alt_order:
        # the same workload in the other cell ordering (RCM is the reference's), same flux
        # form choice, outside the timed region above
        alt = "rcm" if a.order == "hilbert" else "hilbert"
        raw2, _, pcfg2, ncfg2, u2 = build_case(alt)
        r2 = residual_leg(raw2, pcfg2, ncfg2, u2, a.flux_form, a.steps, a.warmup)
        out["other_order"] = {
            "order": alt, "flux_form": r2["form"], "value": round(r2["value"], 3),
            "median_value": round(raw2.nelem / r2["median_ms"] / 1e3, 3),
            "kernels_ms": {"k_gradients": round(r2["grad_s"] * 1e3, 5),
                           "k_flux": round(r2["flux_s"] * 1e3, 5)},
            "k_flux_frac": round(FLUX_PASS_BYTES_PER_CELL * raw2.nelem / r2["flux_s"] / 1e9
                                 / HBM_PEAK_GBS, 4)}
        del raw2, u2, r2
    po = None
    if not a.no_cpu_baseline:
        po, flags = native_oracle()
        out["cpu_baseline"] = cpu_baseline(po, flags, raw, bcs, u_host, a.cpu_seconds,
                                           a.cpu_seconds_1t)
    if not a.no_solve:
        # second half of the metric: time to 1e-10 of the free-stream residual norm
        out["time_to_tol"] = time_to_tol(a)
    if not a.no_batch:
        out["batch"] = batch_sweep(a.batch_instances, a.batch_maxiter)
        out["batch_bf16_pc"] = batch_sweep(a.batch_instances, a.batch_maxiter, pc_precision="bf16")
        out["mpnn"] = mpnn_timing(8, 20)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
