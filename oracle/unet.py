"""FP64 numpy restatement of the learned-guess network (test infrastructure).

Only the tests use this module; the product path runs the network on the
device (fopis_amd/csrc/guess.hip). The network has no counterpart in the
reference, whose solver is SOR from the previous potential
(src/Grid/RootGrid/RootGridField.cpp:41-103); this file states the
network's arithmetic so that the device kernels can be checked:

- conv3d: channels-last cross-correlation over the 27 taps, zero padding,
  tap t = (di+1)*9 + (dj+1)*3 + (dk+1) reads x[i+di, j+dj, k+dk];
  y = sum_t x_shift_t @ W[t] + b;
- up: nearest-neighbour upsampling by 2 (input read at index >> 1);
- pool: mean over 2x2x2 blocks;
- loss: 0.5 * mean over voxels of interior nodes of (out - target/k)^2, or of
  the residual (A(k out) - rho_hat)^2 with A the 7-point operator.

Pins: tests/test_guess_oracle.py checks conv3d against torch.nn.functional
.conv3d on the CPU in float64, and the hand-written backward of the whole
network against central finite differences.
"""
from __future__ import annotations

import numpy as np

LAYERS = [
    (8, 16, 0, 0, 1), (16, 16, 0, 0, 1),
    (16, 32, 1, 0, 1), (32, 32, 1, 0, 1),
    (32, 48, 2, 0, 1), (48, 48, 2, 0, 1),
    (48, 64, 3, 0, 1), (64, 64, 3, 0, 1),
    (64, 48, 2, 1, 1), (48, 48, 2, 0, 1),
    (48, 32, 1, 1, 1), (32, 32, 1, 0, 1),
    (32, 16, 0, 1, 1), (16, 16, 0, 0, 1),
    (16, 16, 0, 0, 0),
]
E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, D2A, D2B, D1A, D1B, D0A, D0B, OUT = range(15)


def bf16_round(a):
    """Round float64 values to the nearest bf16 (ties to even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unpack(params):
    """Flat parameter vector -> [(W [27,cin,cout], b [cout])] in layer order."""
    out, off = [], 0
    for ci, co, _, _, _ in LAYERS:
        nw = 27 * ci * co
        out.append((np.asarray(params[off:off + nw], np.float64).reshape(27, ci, co),
                    np.asarray(params[off + nw:off + nw + co], np.float64)))
        off += nw + co
    return out


def pack(grads):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in grads])


def upsample(x):
    return x.repeat(2, 0).repeat(2, 1).repeat(2, 2)


def conv3d(x, W, b, up=False):
    if up:
        x = upsample(x)
    n0, n1, n2, cin = x.shape
    xp = np.zeros((n0 + 2, n1 + 2, n2 + 2, cin))
    xp[1:-1, 1:-1, 1:-1] = x
    y = np.zeros((n0, n1, n2, W.shape[2]))
    for t in range(27):
        a, bb, c = t // 9, (t // 3) % 3, t % 3
        y += xp[a:a + n0, bb:bb + n1, c:c + n2] @ W[t]
    return y + b


def conv3d_backward(x, W, gy, up=False):
    """dW, db and dx (at the input's own resolution) of conv3d."""
    xu = upsample(x) if up else x
    n0, n1, n2, cin = xu.shape
    cout = W.shape[2]
    xp = np.zeros((n0 + 2, n1 + 2, n2 + 2, cin))
    xp[1:-1, 1:-1, 1:-1] = xu
    g2 = gy.reshape(-1, cout)
    dW = np.zeros_like(W)
    dxp = np.zeros_like(xp)
    for t in range(27):
        a, bb, c = t // 9, (t // 3) % 3, t % 3
        dW[t] = xp[a:a + n0, bb:bb + n1, c:c + n2].reshape(-1, cin).T @ g2
        dxp[a:a + n0, bb:bb + n1, c:c + n2] += gy @ W[t].T
    dx = dxp[1:-1, 1:-1, 1:-1]
    if up:
        dx = dx.reshape(n0 // 2, 2, n1 // 2, 2, n2 // 2, 2, cin).sum(axis=(1, 3, 5))
    return dW, gy.sum(axis=(0, 1, 2)), dx


def pool(x):
    n0, n1, n2, c = x.shape
    return x.reshape(n0 // 2, 2, n1 // 2, 2, n2 // 2, 2, c).mean(axis=(1, 3, 5))


def pool_backward(gc):
    return upsample(gc) / 8.0


def forward(x0, params, round_act=False):
    """x0 [n,n,n,8] -> (out [n,n,n] channel 0 of the last layer, saved activations)."""
    P = unpack(params)
    rnd = bf16_round if round_act else (lambda a: a)

    def layer(li, x, add=None):
        ci, co, lvl, up, relu = LAYERS[li]
        W, b = P[li]
        y = conv3d(x, W, b, up=bool(up))
        if add is not None:
            y = y + add
        if relu:
            y = np.maximum(y, 0.0)
        return rnd(y)

    A = {}
    A["X0"] = x0
    A["A0"] = layer(E0A, x0)
    A["S0"] = layer(E0B, A["A0"])
    A["P0"] = rnd(pool(A["S0"]))
    A["A1"] = layer(E1A, A["P0"])
    A["S1"] = layer(E1B, A["A1"])
    A["P1"] = rnd(pool(A["S1"]))
    A["A2"] = layer(E2A, A["P1"])
    A["S2"] = layer(E2B, A["A2"])
    A["P2"] = rnd(pool(A["S2"]))
    A["A3"] = layer(E3A, A["P2"])
    A["B3"] = layer(E3B, A["A3"])
    A["D2"] = layer(D2A, A["B3"], A["S2"])
    A["E2"] = layer(D2B, A["D2"])
    A["D1"] = layer(D1A, A["E2"], A["S1"])
    A["E1"] = layer(D1B, A["D1"])
    A["D0"] = layer(D0A, A["E1"], A["S0"])
    A["E0"] = layer(D0B, A["D0"])
    W, b = P[OUT]
    out = conv3d(A["E0"], W, b)[..., 0]
    return out, A


def laplace7(u, m):
    """6 u - sum of the 6 neighbours on the voxels m, with u taken as 0 outside m."""
    u = np.where(m, u, 0.0)
    p = np.pad(u, 1)
    nb = (p[:-2, 1:-1, 1:-1] + p[2:, 1:-1, 1:-1] + p[1:-1, :-2, 1:-1] + p[1:-1, 2:, 1:-1] +
          p[1:-1, 1:-1, :-2] + p[1:-1, 1:-1, 2:])
    return np.where(m, 6.0 * u - nb, 0.0)


def loss_grad(x0, params, target, k, valid, round_act=False, kind=0):
    """Loss and flat gradient; valid = (vn0, vn1, vn2) voxels of interior nodes.
    kind 0: 0.5 mean((out - target/k)^2); kind 1: 0.5 mean((A(k out) - rho_hat)^2)
    with rho_hat = x0[..., 0] and A the 7-point operator (laplace7)."""
    out, A = forward(x0, params, round_act)
    P = unpack(params)
    m = np.zeros(out.shape, bool)
    m[: valid[0], : valid[1], : valid[2]] = True
    n = m.sum()
    G = [None] * len(LAYERS)
    gout = np.zeros(out.shape + (16,))
    if kind == 1:
        q = np.where(m, laplace7(k * out, m) - x0[..., 0], 0.0)
        loss = 0.5 * (q * q).sum() / n
        gout[..., 0] = k / n * laplace7(q, m)
    else:
        d = np.where(m, out - target / k, 0.0)
        loss = 0.5 * (d * d).sum() / n
        gout[..., 0] = d / n

    def back(li, x, gy, y=None):
        ci, co, lvl, up, relu = LAYERS[li]
        if relu:
            gy = gy * (y > 0)
        dW, db, dx = conv3d_backward(x, P[li][0], gy, up=bool(up))
        G[li] = (dW, db)
        return gy, dx

    _, gE0 = back(OUT, A["E0"], gout)
    _, gD0 = back(D0B, A["D0"], gE0, A["E0"])
    gS0skip, gE1 = back(D0A, A["E1"], gD0, A["D0"])
    _, gD1 = back(D1B, A["D1"], gE1, A["E1"])
    gS1skip, gE2 = back(D1A, A["E2"], gD1, A["D1"])
    _, gD2 = back(D2B, A["D2"], gE2, A["E2"])
    gS2skip, gB3 = back(D2A, A["B3"], gD2, A["D2"])
    _, gA3 = back(E3B, A["A3"], gB3, A["B3"])
    _, gP2 = back(E3A, A["P2"], gA3, A["A3"])
    gS2 = pool_backward(gP2) + gS2skip
    _, gA2 = back(E2B, A["A2"], gS2, A["S2"])
    _, gP1 = back(E2A, A["P1"], gA2, A["A2"])
    gS1 = pool_backward(gP1) + gS1skip
    _, gA1 = back(E1B, A["A1"], gS1, A["S1"])
    _, gP0 = back(E1A, A["P0"], gA1, A["A1"])
    gS0 = pool_backward(gP0) + gS0skip
    _, gA0 = back(E0B, A["A0"], gS0, A["S0"])
    back(E0A, A["X0"], gA0, A["A0"])
    return loss, pack(G)
