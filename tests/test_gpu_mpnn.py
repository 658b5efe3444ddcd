"""Learned warm start kernels against the numpy restatement (oracle/mpnn_oracle.py).

Tolerances: the kernels and the restatement round the same quantities to bf16 (weights, GEMM
operands), but fp32 sums in a different order can put a value on the other side of a bf16
rounding boundary, which moves it by one bf16 unit (2^-8 relative). Such flips are rare and
diluted by the following sums, so outputs and gradients are compared in relative L2 norm
against 2^-8/4 ~ 1e-3 (outputs) and 1e-2 (gradient tensors, whose small entries are sums of
many cancelling terms), and in the max norm against 4 x 2^-8.
"""
import numpy as np
import pytest

import mpnn_oracle as mo
import pyoracle as po
from fvens_amd import fvx, meshgen, warmstart

pytestmark = pytest.mark.gpu

WALL = 2
MACH = np.array([0.45, 0.62, 0.74])
AOA = np.array([-3.0, 1.25, 5.5])


@pytest.fixture(scope="module")
def setup(torch_cuda):
    raw = meshgen.mesh_k1()          # bump channel: wall marker 2
    gm = fvx.Mesh(raw)
    net = fvx.MPNN(gm, WALL, 4)
    g = mo.graph(po.Mesh(raw), raw, WALL)
    return raw, gm, net, g


def _params(seed):
    w = warmstart.init_params(seed, dec_scale=0.3)
    rng = np.random.default_rng(seed + 100)
    # non-zero biases so that every bias path is exercised
    for name, sl in _slices().items():
        if name.startswith("b"):
            w[sl] = rng.standard_normal(sl.stop - sl.start).astype(np.float32) * 0.1
    return w


def _slices():
    out, o = {}, 0
    for name, shape in mo.layout():
        n = int(np.prod(shape))
        out[name] = slice(o, o + n)
        o += n
    return out


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_layout_matches(setup):
    assert fvx.MPNN.nparams() == mo.nparams() == 67908
    assert warmstart.layout() == mo.layout()


def test_graph_features_match(setup):
    raw, gm, net, g = setup
    f = net.features()
    assert np.array_equal(f["nbr"], g["nbr"])
    assert np.array_equal(f["rev"], g["rev"])
    assert np.array_equal(f["invdeg"], g["inv_deg"])
    np.testing.assert_allclose(f["stat"], g["static"], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(f["edge"], g["edge"], rtol=1e-6, atol=1e-6)
    # the wall-adjacent flag marks exactly the cells with a wall face
    assert f["stat"][:, 5].sum() == np.sum(raw.btag == WALL)


def test_forward_matches_oracle(setup):
    import torch
    raw, gm, net, g = setup
    w = _params(1)
    net.set_params(w)
    net.set_points(MACH, AOA)
    y = torch.zeros((3, raw.nelem, 4), dtype=torch.float32, device="cuda")
    net.predict(y)
    yg = y.cpu().numpy()
    yo, _ = mo.forward(w, g, MACH, AOA, quant=True)
    assert _rel(yg, yo) < 1e-3, _rel(yg, yo)
    assert np.abs(yg - yo).max() <= 4 * 2.0 ** -8 * np.abs(yo).max()
    # a graph's prediction does not depend on the other graphs of the call
    net.set_points(MACH[1:2], AOA[1:2])
    y1 = torch.zeros((1, raw.nelem, 4), dtype=torch.float32, device="cuda")
    net.predict(y1)
    assert torch.equal(y1[0], y[1])


def test_gradient_matches_oracle(setup):
    import torch
    raw, gm, net, g = setup
    w = _params(2)
    net.set_params(w)
    rng = np.random.default_rng(5)
    tgt = (rng.standard_normal((4, raw.nelem, 4)) * 0.2).astype(np.float32)
    ids = np.array([2, 0, 3])
    net.set_points(MACH, AOA)
    tt = torch.tensor(tgt, device="cuda")
    grad = torch.zeros(mo.nparams(), dtype=torch.float32, device="cuda")
    loss = net.gradient(tt, ids, grad)
    gg = grad.cpu().numpy()
    lo, go = mo.loss_and_grad(w, g, MACH, AOA, tgt[ids], quant=True)
    assert abs(loss - lo) <= 1e-3 * lo
    for name, sl in _slices().items():
        assert _rel(gg[sl], go[sl]) < 1e-2, (name, _rel(gg[sl], go[sl]))
    # deterministic: a second evaluation is bitwise identical
    grad2 = torch.zeros_like(grad)
    assert net.gradient(tt, ids, grad2) == loss
    assert torch.equal(grad, grad2)


def test_adam_step_matches_oracle(setup):
    import torch
    raw, gm, net, g = setup
    w = _params(3)
    net.set_params(w)
    rng = np.random.default_rng(6)
    tgt = torch.tensor((rng.standard_normal((2, raw.nelem, 4)) * 0.2).astype(np.float32),
                       device="cuda")
    net.set_points(MACH[:2], AOA[:2])
    grad = torch.zeros(mo.nparams(), dtype=torch.float32, device="cuda")
    net.gradient(tgt, [0, 1], grad)
    g1 = grad.cpu().numpy()
    net.train_step(tgt, [0, 1], 1e-3)
    w1 = net.get_params()
    we, m, v = mo.adam(w, np.zeros_like(w), np.zeros_like(w), g1, 1, 1e-3)
    np.testing.assert_allclose(w1, we, rtol=1e-6, atol=1e-9)
    # second step continues the moments
    net.gradient(tgt, [0, 1], grad)
    g2 = grad.cpu().numpy()
    net.train_step(tgt, [0, 1], 1e-3)
    we2, _, _ = mo.adam(we, m, v, g2, 2, 1e-3)
    np.testing.assert_allclose(net.get_params(), we2, rtol=1e-6, atol=1e-9)


def test_targets_and_warm_state(setup):
    import torch
    raw, gm, net, g = setup
    rng = np.random.default_rng(7)
    G, N = 3, raw.nelem
    rho = 1 + 0.2 * rng.standard_normal((G, N))
    vx = np.cos(np.deg2rad(AOA))[:, None] + 0.1 * rng.standard_normal((G, N))
    vy = np.sin(np.deg2rad(AOA))[:, None] + 0.1 * rng.standard_normal((G, N))
    p = (1 + 0.2 * rng.standard_normal((G, N))) / (1.4 * MACH[:, None] ** 2)
    u = np.stack([rho, rho * vx, rho * vy, p / 0.4 + 0.5 * rho * (vx ** 2 + vy ** 2)], -1)
    net.set_points(MACH, AOA)
    ut = torch.tensor(u, device="cuda")
    t = torch.zeros((G, N, 4), dtype=torch.float32, device="cuda")
    net.targets(ut, t)
    to = mo.targets(u, MACH, AOA)
    np.testing.assert_allclose(t.cpu().numpy(), to, rtol=1e-6, atol=1e-7)
    # the warm state of the targets is the state itself (to fp32 precision of the outputs)
    w = _params(4)
    net.set_params(w)
    y = torch.zeros_like(t)
    net.predict(y)
    uw = torch.zeros_like(ut)
    net.warm_state(uw, floor=0.05)
    uo = mo.warm_state(y.cpu().numpy(), MACH, AOA, floor=0.05)
    np.testing.assert_allclose(uw.cpu().numpy(), uo, rtol=1e-12, atol=1e-12)


def test_training_reduces_loss(setup):
    import torch
    raw, gm, net, g = setup
    # smooth synthetic fields that depend on the operating point and the position
    mach = np.linspace(0.35, 0.7, 8)
    aoa = np.linspace(-3, 5, 8)
    xc = g["static"][:, 0]
    yc = g["static"][:, 1]
    tgt = np.stack([np.stack([0.2 * (m - 0.5) * np.sin(3 * xc), 0.05 * a * yc,
                              0.1 * np.cos(2 * xc) * m, 0.3 * xc * yc], -1)
                    for m, a in zip(mach, aoa)]).astype(np.float32)
    tt = torch.tensor(tgt, device="cuda")
    net.set_params(warmstart.init_params(0))
    hist = warmstart.train(net, tt, mach, aoa, steps=300, graphs_per_step=4, lr0=3e-3,
                           log_every=299, log=lambda s: None)
    assert hist[-1][1] < 0.1 * hist[0][1], hist
