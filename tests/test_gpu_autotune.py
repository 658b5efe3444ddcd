"""Autotuner: every launch shape of the cell kernels gives bitwise the same residual, time
steps and Jacobian (each cell is computed independently of how cells are grouped), the tuner
returns its timing table and keeps a shape from it, and the preconditioner tuner returns a
schedule that reaches the tolerance."""
import numpy as np
import pytest

from fvens_amd import fvx, meshgen, states

pytestmark = pytest.mark.gpu

BCS = [(2, "slipwall"), (4, "farfield")]


def _case():
    import torch
    raw = meshgen.airfoil_omesh(128, 32, 16)
    raw = raw.permuted(fvx.rcm_permutation(raw))
    gm = fvx.Mesh(raw)
    d = fvx.FlowFV(gm, fvx.physics_cfg(Minf=0.7, aoa_deg=2.0, bcs=BCS),
                   fvx.numerics_cfg(flux="HLLC", gradient="LEASTSQUARES", recon="VENKATAKRISHNAN"))
    u = torch.tensor(states.perturbed_freestream(raw, Minf=0.7, aoa_deg=2.0), device="cuda")
    return raw, gm, d, u


def _eval(raw, gm, d, u):
    import torch
    r = torch.zeros_like(u)
    dt = torch.zeros(raw.nelem, dtype=torch.float64, device="cuda")
    d.compute_residual(u, r, True, dt)
    A = fvx.Matrix(gm)
    d.assemble_jacobian(u, A)
    return r.clone(), dt.clone(), A.to_bsr()[2]


def test_results_do_not_depend_on_launch_shape(torch_cuda):
    import torch
    raw, gm, d, u = _case()
    r0, dt0, J0 = _eval(raw, gm, d, u)
    for block in (64, 128, 256, 512):
        for xcd in (False, True):
            d.set_launch("gradients", block, xcd)
            d.set_launch("flux", block, xcd)
            d.set_launch("jacobian", min(block, 256), xcd)
            r, dt, J = _eval(raw, gm, d, u)
            assert torch.equal(r, r0) and torch.equal(dt, dt0), (block, xcd)
            assert np.array_equal(J, J0), (block, xcd)
    # the wave-task form of the flux pass
    for xcd in (False, True):
        d.set_launch("flux_wave", 64, xcd)
        r, dt, J = _eval(raw, gm, d, u)
        assert torch.equal(r, r0) and torch.equal(dt, dt0), ("wave", xcd)


def test_autotuner_keeps_the_fastest_shape(torch_cuda):
    raw, gm, d, u = _case()
    table = d.autotune(u, reps=5)
    # 8 shapes each for the gradient and quad-form flux kernels, the wave-task flux form in
    # both block orders, 6 Jacobian shapes
    assert len(table) == 8 + 8 + 2 + 6
    assert all(t[3] > 0 for t in table)
    for k in ("gradients", "jacobian"):
        best = min((t for t in table if t[0] == k), key=lambda t: t[3])
        assert d.get_launch(k) == (best[1], best[2])
    quad = min((t for t in table if t[0] == "flux"), key=lambda t: t[3])
    best = min((t for t in table if t[0] in ("flux", "flux_wave")), key=lambda t: t[3])
    assert d.get_launch("flux")[0] == quad[1]
    if best[0] == "flux_wave":
        assert d.get_launch("flux_wave") == (64, best[2])
    else:
        assert d.get_launch("flux") == (best[1], best[2]) and d.get_launch("flux_wave")[0] == 0


def test_preconditioner_tuner_returns_a_schedule_that_converges(torch_cuda):
    import torch
    raw, gm, d, u = _case()
    (pc, sweeps), table = d.autotune_pc(u, cfl=200.0, rtol=0.1)
    assert len(table) == 10
    ok = [t for t in table if t[3] > 0]
    assert ok and (pc, sweeps) == min(ok, key=lambda t: t[3])[:2]
    # the chosen schedule solves the system to the tolerance
    A = fvx.Matrix(gm)
    r = torch.zeros_like(u)
    dt = torch.zeros(raw.nelem, dtype=torch.float64, device="cuda")
    dtm = torch.zeros_like(dt)
    d.compute_residual(u, r, True, dt)
    d.assemble_jacobian(u, A)
    d.add_pseudotime(A, dt, 200.0, dtm)
    P = fvx.Preconditioner(gm, pc, sweeps=sweeps)
    P.setup(A)
    x = torch.zeros_like(u)
    assert fvx.linear_solve(A, P, r, x, restart=30, rtol=0.1) < 30
