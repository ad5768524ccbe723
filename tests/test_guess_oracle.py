"""Pins of the FP64 network oracle (oracle/unet.py) on the CPU: its conv3d and
conv3d backward against torch's CPU float64 conv3d and autograd, its pool and
upsample against their definitions, and the hand-written backward of the
whole network against central finite differences."""
import numpy as np
import pytest
import torch

from oracle import unet


def _torch_conv(x, W, b, up):
    """x [n0,n1,n2,cin], W [27,cin,cout] through torch (NCDHW, cross-correlation)."""
    xt = torch.tensor(x).permute(3, 0, 1, 2)[None]
    if up:
        xt = torch.nn.functional.interpolate(xt, scale_factor=2, mode="nearest")
    wt = torch.tensor(W).reshape(3, 3, 3, W.shape[1], W.shape[2]).permute(4, 3, 0, 1, 2)
    y = torch.nn.functional.conv3d(xt, wt, torch.tensor(b), padding=1)
    return y[0].permute(1, 2, 3, 0).numpy()


@pytest.mark.parametrize("shape,cin,cout,up", [((5, 6, 7), 3, 4, False), ((4, 6, 8), 8, 16, True),
                                                ((3, 3, 3), 16, 2, False)])
def test_conv3d_matches_torch(shape, cin, cout, up):
    rng = np.random.default_rng(3)
    xs = tuple(s // 2 for s in shape) if up else shape
    x = rng.standard_normal(xs + (cin,))
    W = rng.standard_normal((27, cin, cout))
    b = rng.standard_normal(cout)
    np.testing.assert_allclose(unet.conv3d(x, W, b, up), _torch_conv(x, W, b, up), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("up", [False, True])
def test_conv3d_backward_matches_torch_autograd(up):
    rng = np.random.default_rng(4)
    shape, cin, cout = (4, 6, 4), 5, 3
    xs = tuple(s // 2 for s in shape) if up else shape
    x = rng.standard_normal(xs + (cin,))
    W = rng.standard_normal((27, cin, cout))
    b = rng.standard_normal(cout)
    gy = rng.standard_normal(shape + (cout,))
    xt = torch.tensor(x, requires_grad=True)
    Wt = torch.tensor(W, requires_grad=True)
    bt = torch.tensor(b, requires_grad=True)
    xi = xt.permute(3, 0, 1, 2)[None]
    if up:
        xi = torch.nn.functional.interpolate(xi, scale_factor=2, mode="nearest")
    wk = Wt.reshape(3, 3, 3, cin, cout).permute(4, 3, 0, 1, 2)
    y = torch.nn.functional.conv3d(xi, wk, bt, padding=1)[0].permute(1, 2, 3, 0)
    (y * torch.tensor(gy)).sum().backward()
    dW, db, dx = unet.conv3d_backward(x, W, gy, up)
    np.testing.assert_allclose(dW, Wt.grad.numpy(), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(dx, xt.grad.numpy(), rtol=1e-10, atol=1e-10)


def test_pool_and_upsample_definitions():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((4, 6, 2, 3))
    p = unet.pool(x)
    assert p.shape == (2, 3, 1, 3)
    assert np.isclose(p[1, 2, 0, 1], x[2:4, 4:6, 0:2, 1].mean())
    u = unet.upsample(p)
    assert u.shape == x.shape and u[3, 5, 1, 2] == p[1, 2, 0, 2]
    # pool_backward is the adjoint of pool: <pool(x), g> = <x, pool_backward(g)>
    g = rng.standard_normal(p.shape)
    assert np.isclose((p * g).sum(), (x * unet.pool_backward(g)).sum())


@pytest.mark.parametrize("kind", [0, 1])
def test_network_backward_matches_finite_differences(kind):
    rng = np.random.default_rng(6)
    n = 8
    total = sum(27 * ci * co + co for ci, co, *_ in unet.LAYERS)
    params = np.zeros(total)
    off = 0
    for ci, co, *_ in unet.LAYERS:
        params[off:off + 27 * ci * co] = rng.standard_normal(27 * ci * co) * np.sqrt(2.0 / (27 * ci))
        params[off + 27 * ci * co:off + 27 * ci * co + co] = 0.05 * rng.standard_normal(co)
        off += 27 * ci * co + co
    x0 = np.zeros((n, n, n, 8))
    x0[..., 0] = rng.standard_normal((n, n, n))
    x0[2:5, 2:6, 3:5, 1] = 1.0
    target = rng.standard_normal((n, n, n))
    valid = (7, 7, 7)
    loss, grad = unet.loss_grad(x0, params, target, 1.3, valid, kind=kind)
    idx = rng.choice(total, 24, replace=False)
    # include parameters of the first, a skip-fed and the output layer
    idx = np.concatenate([idx, [5, total - 3, 27 * 8 * 16 + 16 + 100]])
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (unet.loss_grad(x0, p1, target, 1.3, valid, kind=kind)[0] -
              unet.loss_grad(x0, p2, target, 1.3, valid, kind=kind)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-7, (i, fd, grad[i])


def test_laplace7_is_the_poisson_operator():
    """laplace7 against the 7-point operator written out, and its symmetry."""
    rng = np.random.default_rng(8)
    u = rng.standard_normal((5, 6, 7))
    m = np.zeros(u.shape, bool)
    m[:4, :5, :6] = True
    L = unet.laplace7(u, m)
    i, j, k = 2, 3, 4
    want = 6 * u[i, j, k] - (u[i - 1, j, k] + u[i + 1, j, k] + u[i, j - 1, k] + u[i, j + 1, k] + u[i, j, k - 1] +
                             u[i, j, k + 1])
    assert np.isclose(L[i, j, k], want)
    assert L[3, 4, 5] == 6 * u[3, 4, 5] - (u[2, 4, 5] + u[3, 3, 5] + u[3, 4, 4])  # neighbours outside m count as 0
    v = rng.standard_normal(u.shape)
    assert np.isclose((unet.laplace7(u, m) * np.where(m, v, 0)).sum(), (np.where(m, u, 0) * unet.laplace7(v, m)).sum())
