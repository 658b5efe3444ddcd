"""CPU restatement of the learned warm start's message-passing network (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the smoke check as the checker of the HIP
training and inference kernels, never by the product package.

The network has no counterpart in the reference (the reference starts every solve from the
freestream or from a first-order starter solve, `src/utilities/casesolvers.cpp:225-386`), so this
file is the specification the kernels are checked against, and the finite-difference check in
tests/test_mpnn_oracle.py pins its backward pass.

Graph: nodes are cells; the edges of cell i are its interior neighbour slots (the reference's
`esuel`, `src/mesh/mesh.cpp:659-762`, entries below nelem).

Features
  node (8): (xc-0.5)/(1+r), yc/(1+r), log(1+r)/log(51), log10(dw+1e-4)/2, log10(sqrt(area))/2,
            wall-adjacent flag, 4(M-0.5), aoa_deg/5      with r = |centre - (0.5, 0)|, dw the
            distance from the centre to the nearest wall face segment
  edge (4) of slot k of cell i towards j: dx/L, dy/L, log10(L/sqrt(area_i)), log(area_j/area_i)
            with (dx, dy) = centre_j - centre_i, L its length

Network (width 64, 4 message-passing layers, relu = max(0, .)):
  h0 = x Wenc + benc
  layer l:  P = hb Wa,  Q = hb Wb                   (hb = h_l rounded to bf16)
            s_ik = P_i + Q_j + e_ik We + bm,  a_i = (1/deg_i) sum_k relu(s_ik)
            h_{l+1} = h_l + relu(hb Wu + ab Wv + bu)   (ab = a rounded to bf16)
  y  = h4 Wdec + bdec     (4 outputs: rho-1, vx-cos(aoa), vy-sin(aoa), p/p_inf-1)
  loss = mean over graphs, cells and outputs of (y - t)^2
All weight matrices enter the arithmetic rounded to bf16; biases, activations kept between
kernels and accumulations are fp32. In `quant=False` mode everything is float64 (used by the
finite-difference check).
"""
from __future__ import annotations

import numpy as np

F_IN, F_EDGE, WIDTH, NLAYER, F_OUT = 8, 4, 64, 4, 4


def layout():
    """Flat parameter layout: list of (name, shape); matrices are [in][out] row-major."""
    L = [("Wenc", (F_IN, WIDTH)), ("benc", (WIDTH,))]
    for l in range(NLAYER):
        L += [(f"Wa{l}", (WIDTH, WIDTH)), (f"Wb{l}", (WIDTH, WIDTH)), (f"We{l}", (F_EDGE, WIDTH)),
              (f"bm{l}", (WIDTH,)), (f"Wu{l}", (WIDTH, WIDTH)), (f"Wv{l}", (WIDTH, WIDTH)),
              (f"bu{l}", (WIDTH,))]
    L += [("Wdec", (WIDTH, F_OUT)), ("bdec", (F_OUT,))]
    return L


def nparams():
    return sum(int(np.prod(s)) for _, s in layout())


def unpack(flat):
    out, o = {}, 0
    for name, shape in layout():
        n = int(np.prod(shape))
        out[name] = flat[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name]).reshape(-1) for name, _ in layout()])


def bf16(x):
    """Round to the nearest bf16 (ties to even), returned as float32."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def _point_segment(px, py, ax, ay, bx, by):
    dx, dy = bx - ax, by - ay
    ll = dx * dx + dy * dy
    t = np.clip(((px - ax) * dx + (py - ay) * dy) / np.where(ll > 0, ll, 1.0), 0.0, 1.0)
    qx, qy = ax + t * dx - px, ay + t * dy - py
    return np.sqrt(qx * qx + qy * qy)


def graph(mesh, raw, wall_tag):
    """Graph arrays of a mesh: `mesh` is the oracle mesh (pyoracle.Mesh), `raw` the raw mesh."""
    N = mesh.nelem
    esuel = mesh.esuel
    valid = (esuel >= 0) & (esuel < N)
    nbr = np.where(valid, esuel, -1).astype(np.int64)
    deg = valid.sum(1)
    inv_deg = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    # reverse slot: slot of cell i in the list of its neighbour j that points back to i
    rev = np.full((N, 4), -1, np.int64)
    for k in range(4):
        j = nbr[:, k]
        ok = j >= 0
        for kk in range(4):
            hit = ok & (nbr[np.maximum(j, 0), kk] == np.arange(N))
            rev[hit, k] = kk
    assert np.all((rev >= 0) == valid)
    rc, area = mesh.rc, mesh.area
    xc, yc = rc[:, 0], rc[:, 1]
    r = np.hypot(xc - 0.5, yc)
    wall = np.nonzero(raw.btag == wall_tag)[0]
    dw = np.full(N, np.inf)
    for ib in wall:
        a, b = raw.coords[raw.bface[ib, 0]], raw.coords[raw.bface[ib, 1]]
        dw = np.minimum(dw, _point_segment(xc, yc, a[0], a[1], b[0], b[1]))
    if len(wall) == 0:
        dw[:] = 1e4
    walladj = np.zeros(N)
    bslot = esuel >= N
    for k in range(4):
        ib = np.where(bslot[:, k], esuel[:, k] - N, 0)
        walladj = np.maximum(walladj, np.where(bslot[:, k], (raw.btag[ib] == wall_tag) * 1.0, 0.0))
    static = np.stack([(xc - 0.5) / (1 + r), yc / (1 + r), np.log1p(r) / np.log(51.0),
                       np.log10(dw + 1e-4) / 2, np.log10(np.sqrt(area)) / 2, walladj], 1)
    jj = np.maximum(nbr, 0)
    dx = rc[jj, 0] - xc[:, None]
    dy = rc[jj, 1] - yc[:, None]
    L = np.hypot(dx, dy)
    Ls = np.where(valid, L, 1.0)
    edge = np.stack([dx / Ls, dy / Ls, np.log10(Ls / np.sqrt(area)[:, None]),
                     np.log(area[jj] / area[:, None])], 2)
    edge = np.where(valid[:, :, None], edge, 0.0)
    return dict(N=N, nbr=nbr, valid=valid, inv_deg=inv_deg.astype(np.float32),
                rev=rev, static=static.astype(np.float32), edge=edge.astype(np.float32))


def node_features(g, mach, aoa_deg):
    """(G, N, 8) float32 input features for G operating points."""
    G, N = len(mach), g["N"]
    x = np.empty((G, N, F_IN), np.float32)
    x[:, :, :6] = g["static"][None]
    x[:, :, 6] = np.float32(4.0 * (np.asarray(mach) - 0.5))[:, None]
    x[:, :, 7] = np.float32(np.asarray(aoa_deg) / 5.0)[:, None]
    return x


def _cast(quant):
    if quant:
        return (lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float64),
                lambda a: bf16(np.asarray(a, np.float64).astype(np.float32)).astype(np.float64))
    ident = lambda a: np.asarray(a, np.float64)  # noqa: E731
    return ident, ident


def forward(flat, g, mach, aoa_deg, quant=True):
    """Outputs y (G, N, 4) and the cache the backward pass needs."""
    f32, b16 = _cast(quant)
    p = {k: np.asarray(v, np.float64) for k, v in unpack(np.asarray(flat)).items()}
    W = {k: (b16(v) if k.startswith("W") else f32(v)) for k, v in p.items()}
    x = node_features(g, mach, aoa_deg).astype(np.float64)
    nbr, valid, inv_deg = g["nbr"], g["valid"], g["inv_deg"].astype(np.float64)
    e = g["edge"].astype(np.float64)
    jj = np.maximum(nbr, 0)
    h = f32(x @ W["Wenc"] + W["benc"])
    cache = dict(x=x, W=W, layers=[])
    for l in range(NLAYER):
        hb = b16(h)
        P = f32(hb @ W[f"Wa{l}"])
        Q = f32(hb @ W[f"Wb{l}"])
        s = P[:, :, None, :] + Q[:, jj, :] + (e @ W[f"We{l}"])[None] + W[f"bm{l}"]
        s = np.where(valid[None, :, :, None], s, 0.0)
        a = f32(inv_deg[None, :, None] * np.maximum(s, 0.0).sum(2))
        ab = b16(a)
        pre = hb @ W[f"Wu{l}"] + ab @ W[f"Wv{l}"] + W[f"bu{l}"]
        cache["layers"].append(dict(h=h, hb=hb, s=s, ab=ab, pre=pre))
        h = f32(h + np.maximum(pre, 0.0))
    y = f32(h @ W["Wdec"] + W["bdec"])
    cache["h4"] = h
    return y, cache


def loss_and_grad(flat, g, mach, aoa_deg, targets, quant=True):
    """Mean-squared loss and its gradient (flat, float64) with respect to the parameters.

    The bf16 rounding of activations is passed through unchanged in the backward pass
    (straight-through); weight gradients use the rounded weights, as the kernels do."""
    f32, b16 = _cast(quant)
    y, c = forward(flat, g, mach, aoa_deg, quant)
    W = c["W"]
    t = np.asarray(targets, np.float64)
    cnt = y.size
    diff = y - t
    loss = float((diff * diff).sum() / cnt)
    dy = f32(2.0 * diff / cnt)
    gr = {}
    gr["Wdec"] = np.einsum("gnc,gnd->cd", c["h4"], dy)
    gr["bdec"] = dy.sum((0, 1))
    gh = f32(dy @ W["Wdec"].T)
    nbr, valid, inv_deg, rev = g["nbr"], g["valid"], g["inv_deg"].astype(np.float64), g["rev"]
    e = g["edge"].astype(np.float64)
    N = g["N"]
    for l in reversed(range(NLAYER)):
        L = c["layers"][l]
        dpre = np.where(L["pre"] > 0, gh, 0.0)
        dpb = b16(dpre)
        gr[f"Wu{l}"] = np.einsum("gnc,gnd->cd", L["hb"], dpb)
        gr[f"Wv{l}"] = np.einsum("gnc,gnd->cd", L["ab"], dpb)
        gr[f"bu{l}"] = dpre.sum((0, 1))
        gpart = f32(gh + dpb @ W[f"Wu{l}"].T)
        da = f32(dpb @ W[f"Wv{l}"].T)
        ds = np.where((L["s"] > 0) & valid[None, :, :, None],
                      (da * inv_deg[None, :, None])[:, :, None, :], 0.0)   # (G, N, 4, 64)
        dP = ds.sum(2)
        # dQ_j = sum over the slots of j of ds at the neighbour's reverse slot
        jj = np.maximum(nbr, 0)
        rr = np.maximum(rev, 0)
        dQ = np.where(valid[None, :, :, None], ds[:, jj, rr, :], 0.0).sum(2)
        gr[f"We{l}"] = np.einsum("nkf,gnkc->fc", e, ds)
        gr[f"bm{l}"] = ds.sum((0, 1, 2))
        dPb, dQb = b16(dP), b16(dQ)
        gr[f"Wa{l}"] = np.einsum("gnc,gnd->cd", L["hb"], dPb)
        gr[f"Wb{l}"] = np.einsum("gnc,gnd->cd", L["hb"], dQb)
        gh = f32(gpart + dPb @ W[f"Wa{l}"].T + dQb @ W[f"Wb{l}"].T)
    gr["Wenc"] = np.einsum("gnc,gnd->cd", c["x"], gh)
    gr["benc"] = gh.sum((0, 1))
    del N
    return loss, pack(gr)


def adam(w, m, v, grad, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma and Ba, 2015) in float32; step counts from 1."""
    w, m, v, grad = (np.asarray(a, np.float32) for a in (w, m, v, grad))
    m = np.float32(b1) * m + np.float32(1 - b1) * grad
    v = np.float32(b2) * v + np.float32(1 - b2) * grad * grad
    mh = m / np.float32(1 - b1 ** step)
    vh = v / np.float32(1 - b2 ** step)
    return w - np.float32(lr) * mh / (np.sqrt(vh) + np.float32(eps)), m, v


def targets(u, mach, aoa_deg, gamma=1.4):
    """Network targets from conserved fields u (G, N, 4) in the reference's non-dimensional
    form (rho_inf = 1, |v_inf| = 1, p_inf = 1/(gamma M^2), `src/physics/aphysics.cpp:17-58`)."""
    u = np.asarray(u, np.float64)
    a = np.deg2rad(np.asarray(aoa_deg))[:, None]
    pinf = 1.0 / (gamma * np.asarray(mach) ** 2)[:, None]
    rho = u[..., 0]
    vx, vy = u[..., 1] / rho, u[..., 2] / rho
    p = (gamma - 1) * (u[..., 3] - 0.5 * rho * (vx * vx + vy * vy))
    return np.stack([rho - 1, vx - np.cos(a), vy - np.sin(a), p / pinf - 1], -1).astype(np.float32)


def warm_state(y, mach, aoa_deg, gamma=1.4, floor=0.05):
    """Conserved state from network outputs; density and pressure are floored at `floor` times
    their freestream values."""
    y = np.asarray(y, np.float64)
    a = np.deg2rad(np.asarray(aoa_deg))[:, None]
    pinf = 1.0 / (gamma * np.asarray(mach) ** 2)[:, None]
    rho = np.maximum(1 + y[..., 0], floor)
    vx, vy = np.cos(a) + y[..., 1], np.sin(a) + y[..., 2]
    p = np.maximum(1 + y[..., 3], floor) * pinf
    return np.stack([rho, rho * vx, rho * vy, p / (gamma - 1) + 0.5 * rho * (vx * vx + vy * vy)], -1)
