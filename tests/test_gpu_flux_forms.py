"""The two forms of the face-flux pass give the same residual bits.

The face-task form (each face evaluated once per workgroup, LDS-staged cell data) and the
quad-per-cell form (each cell's lanes evaluate its own faces) use the same face arithmetic and
the same per-cell summation order, so their residuals and time steps must be bitwise equal on
every mesh, ordering, flux and instance set; the face-task form is also checked against the
oracle (per-cell residual within 1e-12) at the orderings the benchmarks use.
"""
import numpy as np
import pytest

import pyoracle as po
from fvens_amd import fvx, meshgen, states

pytestmark = pytest.mark.gpu

BUMP_BCS = [(2, "slipwall"), (3, "inflowoutflow"), (4, "inflowoutflow")]
AIRFOIL_BCS = [(2, "slipwall"), (4, "farfield")]


def _residual(disc, ut, form, nelem):
    import torch
    disc.set_launch("flux_tasks" if form == "tasks" else "flux", 256, True)
    r = torch.full((nelem, 4), 0.25, dtype=torch.float64, device="cuda")
    dt = torch.zeros(nelem, dtype=torch.float64, device="cuda")
    disc.compute_residual(ut, r, True, dt)
    torch.cuda.synchronize()
    return r, dt


def _both(raw, pcfg, ncfg, u):
    import torch
    disc = fvx.FlowFV(fvx.Mesh(raw), pcfg, ncfg)
    ut = torch.tensor(u, device="cuda")
    rq, dq = _residual(disc, ut, "quad", raw.nelem)
    rt, dtt = _residual(disc, ut, "tasks", raw.nelem)
    return rq, dq, rt, dtt


@pytest.mark.parametrize("flux", ["ROE", "HLLC", "HLL", "LLF", "AUSM", "AUSMPLUS", "VANLEER"])
def test_forms_bitwise_equal_on_the_bump(torch_cuda, flux):
    raw = meshgen.mesh_k1()
    u = states.perturbed_freestream(raw, Minf=0.5, L=3.0)
    rq, dq, rt, dtt = _both(raw, fvx.physics_cfg(Minf=0.5, bcs=BUMP_BCS),
                            fvx.numerics_cfg(flux=flux, gradient="LEASTSQUARES",
                                             recon="VENKATAKRISHNAN", limiter_param=5.0), u)
    assert torch_cuda.equal(rq, rt) and torch_cuda.equal(dq, dtt)


@pytest.mark.parametrize("order", ["rcm", "hilbert", "natural"])
def test_forms_bitwise_equal_on_the_airfoil(torch_cuda, order):
    raw = meshgen.mesh_k2()
    if order != "natural":
        raw = raw.permuted(fvx.rcm_permutation(raw) if order == "rcm"
                           else fvx.hilbert_permutation(raw))
    u = states.perturbed_freestream(raw, Minf=0.8, aoa_deg=1.25, L=1.0)
    rq, dq, rt, dtt = _both(raw, fvx.physics_cfg(Minf=0.8, aoa_deg=1.25, bcs=AIRFOIL_BCS),
                            fvx.numerics_cfg(flux="HLLC", gradient="LEASTSQUARES",
                                             recon="VENKATAKRISHNAN", limiter_param=5.0), u)
    assert torch_cuda.equal(rq, rt) and torch_cuda.equal(dq, dtt)


def test_face_task_form_matches_the_oracle_in_hilbert_order(torch_cuda):
    import torch
    raw = meshgen.mesh_k2()
    raw = raw.permuted(fvx.hilbert_permutation(raw))
    phys = dict(Minf=0.8, aoa_deg=1.25, bcs=AIRFOIL_BCS)
    num = dict(flux="HLLC", gradient="LEASTSQUARES", recon="VENKATAKRISHNAN", limiter_param=5.0)
    u = states.perturbed_freestream(raw, Minf=0.8, aoa_deg=1.25, L=1.0)
    od = po.Disc(po.Mesh(raw), po.make_config(**phys, **num))
    r_ref, dt_ref = od.residual(u)
    scale = np.maximum(np.abs(r_ref), od.flux_magnitude(u))
    disc = fvx.FlowFV(fvx.Mesh(raw), fvx.physics_cfg(**phys), fvx.numerics_cfg(**num))
    disc.set_launch("flux_tasks", 256, True)
    ut = torch.tensor(u, device="cuda")
    r = torch.zeros_like(ut)
    dt = torch.zeros(raw.nelem, dtype=torch.float64, device="cuda")
    disc.compute_residual(ut, r, True, dt)
    rel = np.abs(r.cpu().numpy() - r_ref) / scale
    assert rel.max() <= 1e-12, rel.max()
    assert np.allclose(dt.cpu().numpy(), dt_ref, rtol=1e-12, atol=0)


def test_forms_bitwise_equal_for_a_batch(torch_cuda):
    import torch
    raw = meshgen.mesh_k1()
    disc = fvx.FlowFV(fvx.Mesh(raw), fvx.physics_cfg(Minf=0.5, bcs=BUMP_BCS),
                      fvx.numerics_cfg(flux="ROE", gradient="GREENGAUSS", recon="NONE"))
    mach, aoa = np.array([0.3, 0.5, 0.7]), np.array([0.0, 2.0, -1.0])
    b = fvx.Batch(disc, mach, aoa)
    u = np.stack([states.perturbed_freestream(raw, Minf=m, aoa_deg=a, L=3.0, seed=7 + i)
                  for i, (m, a) in enumerate(zip(mach, aoa))])
    ut = torch.tensor(u, device="cuda")
    out = []
    for form in ("flux", "flux_tasks"):
        disc.set_launch(form, 256, True)
        r = torch.zeros_like(ut)
        dt = torch.zeros((3, raw.nelem), dtype=torch.float64, device="cuda")
        b.compute_residual(ut, r, True, dt)
        torch.cuda.synchronize()
        out.append((r, dt))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
