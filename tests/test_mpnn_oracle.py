"""Pins the numpy restatement of the learned warm start network (oracle/mpnn_oracle.py) and the
seeded operating points, without a GPU.

- The backward pass equals central finite differences of the loss in float64 (quant=False).
- bf16 rounding matches hand-worked cases (ties to even, carries into the exponent).
- One Adam step from zero moments moves every weight by lr in the direction of -sign(grad)
  (the closed form of the first bias-corrected step).
- Targets of the freestream are zero and warm_state(targets(u)) == u.
- The point generator reproduces the C++ standard's known answer for std::mt19937_64 and
  draws of libstdc++'s uniform_real_distribution (tests/golden/mt19937_64_points.json).
"""
import json

import numpy as np
import pytest

import mpnn_oracle as mo
import pyoracle as po
from conftest import golden
from fvens_amd import meshgen
from fvens_amd.points import MT19937_64, operating_points


@pytest.fixture(scope="module")
def small_graph():
    raw = meshgen.airfoil_omesh(24, 8, 3)
    return raw, mo.graph(po.Mesh(raw), raw, 2)


def test_finite_difference_gradient(small_graph):
    raw, g = small_graph
    rng = np.random.default_rng(0)
    n = mo.nparams()
    w = rng.standard_normal(n) * 0.15
    mach, aoa = np.array([0.5, 0.7]), np.array([1.0, -2.0])
    t = rng.standard_normal((2, g["N"], 4)) * 0.1
    _, gr = mo.loss_and_grad(w, g, mach, aoa, t, quant=False)
    # a sample from every parameter tensor
    o = 0
    idx = []
    for name, shape in mo.layout():
        k = int(np.prod(shape))
        idx += list(o + rng.choice(k, 3, replace=False))
        o += k
    for i in idx:
        e = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[i] += e
        wm[i] -= e
        fd = (mo.loss_and_grad(wp, g, mach, aoa, t, quant=False)[0]
              - mo.loss_and_grad(wm, g, mach, aoa, t, quant=False)[0]) / (2 * e)
        assert abs(gr[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, gr[i], fd)


def test_planted_backward_mistake_is_caught(small_graph):
    """The finite-difference check fails on a backward pass with a dropped term."""
    raw, g = small_graph
    rng = np.random.default_rng(1)
    w = rng.standard_normal(mo.nparams()) * 0.15
    mach, aoa = np.array([0.6]), np.array([2.0])
    t = rng.standard_normal((1, g["N"], 4)) * 0.1
    g2 = dict(g)
    g2["rev"] = np.where(g["rev"] >= 0, (g["rev"] + 1) % 4, -1)   # wrong reverse slots
    _, gr = mo.loss_and_grad(w, g2, mach, aoa, t, quant=False)
    i = mo.unpack(np.arange(mo.nparams()))["Wb0"].reshape(-1)[5]
    e = 1e-6
    wp, wm = w.copy(), w.copy()
    wp[i] += e
    wm[i] -= e
    fd = (mo.loss_and_grad(wp, g, mach, aoa, t, quant=False)[0]
          - mo.loss_and_grad(wm, g, mach, aoa, t, quant=False)[0]) / (2 * e)
    assert abs(gr[i] - fd) > 1e-4 * max(1.0, abs(fd))


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38,
                  1.0 + 2 ** -8 + 2 ** -20], np.float32)
    big = np.array([0x7F620000], np.uint32).view(np.float32)[0]
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, big, 1.0 + 2 ** -7], np.float32)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-8 is a tie -> even (1+2^-6); 3e38 = 0x7F61B1E6
    # has low half 0xB1E6 > 0x8000 -> 0x7F62; just above a tie rounds up
    got = mo.bf16(x)
    assert np.array_equal(got, want), got


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(2)
    w = rng.standard_normal(100).astype(np.float32)
    g = rng.standard_normal(100).astype(np.float32)
    w1, m, v = mo.adam(w, np.zeros(100), np.zeros(100), g, 1, 1e-3)
    np.testing.assert_allclose(w1, w - 1e-3 * np.sign(g), rtol=0, atol=1e-6)


def test_targets_roundtrip():
    mach, aoa = np.array([0.5, 0.73]), np.array([0.0, 4.0])
    N = 7
    a = np.deg2rad(aoa)
    fs = np.stack([np.ones(2), np.cos(a), np.sin(a),
                   1 / (1.4 * 0.4 * mach ** 2) + 0.5], -1)
    u = np.repeat(fs[:, None, :], N, 1)
    t = mo.targets(u, mach, aoa)
    assert np.abs(t).max() < 1e-6
    rng = np.random.default_rng(3)
    u2 = u * (1 + 0.05 * rng.standard_normal(u.shape))
    back = mo.warm_state(mo.targets(u2, mach, aoa).astype(np.float64), mach, aoa)
    np.testing.assert_allclose(back, u2, rtol=1e-6)


def test_point_generator_known_answers():
    g = MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042       # [rand.predef] of the C++ standard
    ref = json.load(open(golden("mt19937_64_points.json")))
    for case in ref:
        m, a = operating_points(case["seed"], len(case["mach"]))
        assert list(m) == case["mach"] and list(a) == case["aoa_deg"]
