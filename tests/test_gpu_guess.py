"""The learned initial guess on the device against the FP64 network oracle.

Bars: a convolution's BF16 output agrees with the FP64 result on the same
BF16 inputs and BF16-rounded weights to one BF16 rounding (2^-8 relative, the
FP32 accumulation adds ~1e-6); weight and bias gradients (FP32 outputs) to
1e-4 relative; the whole network's parameter gradient, whose activations are
BF16 on the device, to cosine similarity 0.999 and 5 % relative L2 per layer.
The trained network must lower the multigrid cycles to 1e-8 against a start
from zero on held-out pairs."""
import os

import numpy as np
import pytest

from oracle import unet

pytestmark = pytest.mark.gpu

F = pytest.importorskip("fopis_amd.lib")
from fopis_amd import guess as GU  # noqa: E402


def _bf16_input(rng, shape, scale=1.0):
    return GU.to_bf16(rng.uniform(-scale, scale, shape).astype(np.float32))


def _close_bf16(got_bf16, ref):
    got = GU.from_bf16(got_bf16).astype(np.float64)
    tol = 2.0 ** -7 * np.abs(ref) + 1e-5 * np.max(np.abs(ref))
    bad = np.abs(got - ref) > tol
    assert not bad.any(), (np.argwhere(bad)[:5], got[bad][:5], ref[bad][:5])


@pytest.mark.parametrize("shape,cin,cout,up,relu,add", [
    ((8, 8, 16), 8, 16, False, True, False),
    ((6, 10, 12), 16, 48, False, False, True),
    ((8, 8, 8), 64, 48, True, True, True),
    ((4, 4, 4), 48, 64, False, True, False),
    ((5, 7, 9), 32, 32, False, True, True),
])
def test_conv_forward_matches_oracle(shape, cin, cout, up, relu, add):
    rng = np.random.default_rng(11)
    xs = tuple(s // 2 for s in shape) if up else shape
    x = _bf16_input(rng, xs + (cin,))
    W = (rng.standard_normal((27, cin, cout)) * np.sqrt(2.0 / (27 * cin))).astype(np.float32)
    b = rng.uniform(-0.2, 0.2, cout).astype(np.float32)
    a = _bf16_input(rng, shape + (cout,), 0.5) if add else None
    y = GU.conv3d(x, W, b, up=up, relu=relu, add_bf16=a)
    ref = unet.conv3d(GU.from_bf16(x).astype(np.float64), unet.bf16_round(W), b.astype(np.float64), up)
    if add:
        ref = ref + GU.from_bf16(a)
    if relu:
        ref = np.maximum(ref, 0.0)
    _close_bf16(y, ref)


@pytest.mark.parametrize("shape,cin,cout,up", [((8, 8, 8), 16, 32, False), ((8, 8, 8), 48, 16, True),
                                                ((6, 6, 10), 8, 16, False), ((4, 4, 4), 64, 64, False)])
def test_conv_backward_matches_oracle(shape, cin, cout, up):
    rng = np.random.default_rng(12)
    xs = tuple(s // 2 for s in shape) if up else shape
    x = _bf16_input(rng, xs + (cin,))
    W = (rng.standard_normal((27, cin, cout)) * 0.1).astype(np.float32)
    gy = _bf16_input(rng, shape + (cout,))
    want_dx = cin % 16 == 0
    dw, db, dx = GU.conv3d_backward(x, W, gy, up=up, want_dx=want_dx)
    rW, rb, rdx = unet.conv3d_backward(GU.from_bf16(x).astype(np.float64), unet.bf16_round(W),
                                       GU.from_bf16(gy).astype(np.float64), up)
    np.testing.assert_allclose(dw, rW, rtol=1e-4, atol=1e-4 * np.max(np.abs(rW)))
    np.testing.assert_allclose(db, rb, rtol=1e-4, atol=1e-4 * np.max(np.abs(rb)))
    if want_dx:
        # the device returns the gradient at the conv's output grid; fold it for an upsampled input
        d = GU.from_bf16(dx).astype(np.float64)
        if up:
            n0, n1, n2 = shape
            d = d.reshape(n0 // 2, 2, n1 // 2, 2, n2 // 2, 2, cin).sum(axis=(1, 3, 5))
            tol = 8 * 2.0 ** -8 * np.abs(rdx).max()
            assert np.max(np.abs(d - rdx)) <= tol
        else:
            _close_bf16(dx, rdx)


def _desc(n):
    return F.make_desc(n, n, n, [(-1.0, 1.0, 1.0)], eps0=1.0, mg_precision=F.FP64)


def _smooth(rng, shape, k=3):
    f = rng.standard_normal(shape)
    for ax in range(3):
        for _ in range(k):
            f = (np.roll(f, 1, ax) + 2 * f + np.roll(f, -1, ax)) / 4
    return f


@pytest.mark.parametrize("kind", [0, 1])
def test_network_gradient_matches_oracle(kind):
    n = 17  # network grid 16^3: levels 16, 8, 4, 2
    ctx = F.Context(_desc(n))
    g = GU.Guess(ctx, seed=5)
    g.set_loss(kind)
    assert g.grid == (16, 16, 16)
    g.dataset_alloc(1)
    rng = np.random.default_rng(13)
    rho = GU.to_bf16(rng.standard_normal(g.grid).astype(np.float32))
    mask = np.zeros(g.grid, np.uint8)
    mask[4:9, 5:11, 6:10] = 1
    phi = GU.to_bf16(_smooth(rng, g.grid).astype(np.float32) * 3)
    g.dataset_set(0, rho, mask, phi)
    loss, grad = g.grad(0)
    params, k = g.get_params()
    assert k == 1.0
    # the device convolves with BF16 weights and FP32 biases
    P = params.astype(np.float64).copy()
    for name, ws, _, _ in GU.param_slices()[0]:
        P[ws] = unet.bf16_round(params[ws])
    x0 = np.zeros(g.grid + (8,))
    x0[..., 0] = GU.from_bf16(rho)
    x0[..., 1] = mask
    rl, rg = unet.loss_grad(x0, P, GU.from_bf16(phi).astype(np.float64), 1.0, (n - 2,) * 3, round_act=True,
                            kind=kind)
    assert abs(loss - rl) <= 2e-2 * abs(rl), (loss, rl)
    for name, ws, _, bs in GU.param_slices()[0]:
        for sl in (ws, bs):
            a, b = grad[sl].astype(np.float64), rg[sl]
            if name == "e0a" and sl is ws:
                a, b = a.reshape(27, 8, 16)[:, :2], b.reshape(27, 8, 16)[:, :2]  # live input channels
            if name == "out":
                a, b = a.reshape(-1, 16)[:, 0] if sl is ws else a[:1], b.reshape(-1, 16)[:, 0] if sl is ws else b[:1]
            cos = (a * b).sum() / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300)
            rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-300)
            assert cos > 0.999 and rel < 0.05, (name, sl, cos, rel)
    ctx.close()


def test_synth_boxes_charge_sits_on_box_surfaces():
    """The synthetic training problems put each box's charge on the nodes
    whose eight cells are partly inside the box (cells lo <= c < hi of the
    (n+1)^3 cell map): the support, and the sign of each box's charge, match
    a numpy count of the inside cells around every interior node."""
    n = 33
    ctx = F.Context(_desc(n))
    g = GU.Guess(ctx, seed=1)
    shp = ctx.field_shape(F.RHO)
    ctx.set_field(F.RHO, np.zeros(shp))
    boxes = [[5, 6, 7, 14, 12, 20], [18, 20, 3, 26, 29, 11]]
    charges = [0.7, -0.4]
    g.synth_boxes(boxes, charges, 3)
    rho = ctx.get_field(F.RHO)
    cm = np.zeros((n + 1,) * 3, np.int32)
    for b, bx in enumerate(boxes):
        cm[bx[0]:bx[3], bx[1]:bx[4], bx[2]:bx[5]] = b + 1
    cx = n + 2
    count = np.zeros((cx,) * 3, np.int32)
    owner = np.zeros((cx,) * 3, np.int32)
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                sl = cm[a:a + cx - 2, b:b + cx - 2, c:c + cx - 2]  # cells (i-1+a, ...) of nodes i = 1..cx-2
                count[1:cx - 1, 1:cx - 1, 1:cx - 1] += sl > 0
                owner[1:cx - 1, 1:cx - 1, 1:cx - 1] = np.maximum(owner[1:cx - 1, 1:cx - 1, 1:cx - 1], sl)
    interior = np.zeros((cx,) * 3, bool)
    interior[2:cx - 2, 2:cx - 2, 2:cx - 2] = True
    support = interior & (count > 0) & (count < 8)
    assert support.sum() > 0
    np.testing.assert_array_equal(rho != 0, support)
    assert np.all(rho[support & (owner == 1)] > 0) and np.all(rho[support & (owner == 2)] < 0)
    ctx.close()


def _make_pair(ctx, g, rng, seed, slot=None):
    """rho = plasma-like noise + 1-3 boxes with surface charge; phi from multigrid to 1e-8 from zero.
    Returns the cycles of that solve."""
    n = ctx.nx
    rho = np.zeros(ctx.field_shape(F.RHO))
    rho[2:n, 2:n, 2:n] = rng.standard_normal((n - 2,) * 3) * 0.05
    ctx.set_field(F.RHO, rho)
    boxes, charges = [], []
    for _ in range(rng.integers(1, 4)):
        s = rng.integers(4, 10, 3)
        lo = [int(rng.integers(2, n - 1 - s[a])) for a in range(3)]
        boxes.append(lo + [lo[a] + int(s[a]) for a in range(3)])
        charges.append(float(rng.uniform(-1.0, 1.0)))
    g.synth_boxes(boxes, charges, seed)
    ctx.set_field(F.PHI, np.zeros(ctx.field_shape(F.PHI)))
    cyc, rr = ctx.solve_poisson()
    assert rr <= 1e-8
    if slot is not None:
        g.dataset_store(slot)
    return cyc


def test_trained_guess_lowers_multigrid_cycles():
    n = 33  # network grid 32^3
    ctx = F.Context(_desc(n))
    g = GU.Guess(ctx, seed=9)
    g.set_loss(1)  # train on the multigrid's starting residual
    ntrain, nheld = 48, 8
    g.dataset_alloc(ntrain + nheld)
    rng = np.random.default_rng(21)
    for i in range(ntrain + nheld):
        _make_pair(ctx, g, rng, 1000 + i, slot=i)
    before = g.eval(ntrain, ntrain + nheld)
    losses = g.train(ntrain, epochs=40, lr=2e-3, seed=3)
    after = g.eval(ntrain, ntrain + nheld)
    print(f"[guess] train loss {losses[0]:.4g} -> {losses[-1]:.4g}; held-out {before:.4g} -> {after:.4g}")
    assert losses[-1] < 0.5 * losses[0]
    assert after < 0.5 * before
    # multigrid from zero vs from the network's guess on fresh held-out problems
    rng = np.random.default_rng(77)
    zero, guessed = [], []
    for i in range(6):
        state = rng.bit_generator.state
        zero.append(_make_pair(ctx, g, rng, 5000 + i))
        rng.bit_generator.state = state
        _make_pair(ctx, g, rng, 5000 + i)  # same problem again
        ctx.set_field(F.PHI, np.zeros(ctx.field_shape(F.PHI)))
        g.apply()
        cyc, rr = ctx.solve_poisson()
        assert rr <= 1e-8
        guessed.append(cyc)
    print(f"[guess] multigrid cycles to 1e-8: from zero {zero}, from the network {guessed}")
    assert np.mean(guessed) < np.mean(zero)
    ctx.close()


def test_step_with_guess_matches_step_without():
    """main_loop_es with the guess on (mode 1) starts the step's first solve
    from the network; multigrid still iterates to 1e-8, so phi agrees with the
    warm-started step to the solver's tolerance even for an untrained network."""
    from oracle import refcpu
    from tests.cases import device_desc, load_config, oracle_params

    gr = refcpu.Grid(oracle_params(load_config("c1"), nx=33, ppc=4))
    gr.call("rc_load_initial_particles")
    a, b = F.Context(device_desc(gr)), F.Context(device_desc(gr))
    for ctx in (a, b):
        for s in range(gr.S):
            ctx.set_particles(s, gr.get_particles(s))
        ctx.initialize()
    for q in range(3):
        b.set_field(F.EXREF + q, a.get_field(F.EXREF + q))
    b.set_field(F.PHI, a.get_field(F.PHI))
    g = GU.Guess(b, seed=4)
    g.set_mode(1)
    a.main_loop_es()
    b.main_loop_es()
    assert F.solver_stats(b)["last_relres"] <= 1e-8
    pa, pb = a.get_field(F.PHI), b.get_field(F.PHI)
    assert np.max(np.abs(pa - pb)) <= 1e-6 * np.max(np.abs(pa))
    a.close()
    b.close()


def test_training_pairs_hdf5_round_trip(tmp_path):
    """Guess.export_pairs writes the stored (rho_hat, mask, phi_hat) pairs to
    HDF5 and import_pairs reads them into another context's dataset bit for
    bit: the pairs, and the loss the network gets on them, are the same."""
    from fopis_amd import h5

    if not os.path.exists(h5.LIB_PATH):
        pytest.skip("libfopis_io.so not built (no HDF5)")
    n = 33
    a = F.Context(_desc(n))
    ga = GU.Guess(a, seed=1)
    ga.dataset_alloc(4)
    rng = np.random.default_rng(31)
    for i in range(4):
        _make_pair(a, ga, rng, 500 + i, slot=i)
    path = tmp_path / "pairs.h5"
    assert ga.export_pairs(str(path), 1, 4) == 3
    with h5.H5File(path) as f:
        assert tuple(f.read("grid", np.int32)) == tuple(ga.grid)
        assert int(f.read("count", np.int32)) == 3
        assert f.shape("pairs/000002/phi_hat") == tuple(ga.grid)
    b = F.Context(_desc(n))
    gb = GU.Guess(b, seed=1)
    gb.dataset_alloc(5)
    assert gb.import_pairs(str(path), slot0=2) == 3
    for i in range(3):
        for x, y in zip(ga.dataset_get(1 + i), gb.dataset_get(2 + i)):
            np.testing.assert_array_equal(x, y)
    p, k = ga.get_params()
    gb.set_params(p, k)
    assert ga.eval(1, 4) == gb.eval(2, 5)
    a.close()
    b.close()


def _lap7(u):
    """6 u - sum of the six neighbours, zero outside (the solver's operator, dx = 1)."""
    p = np.pad(u, 1)
    return (6 * u - p[:-2, 1:-1, 1:-1] - p[2:, 1:-1, 1:-1] - p[1:-1, :-2, 1:-1] - p[1:-1, 2:, 1:-1]
            - p[1:-1, 1:-1, :-2] - p[1:-1, 1:-1, 2:])


def test_warm_start_correction_records_trains_and_lowers_the_start():
    """Guess mode 3 records, from real PIC steps (33^3 nodes, 2 species, 16
    ppc), pairs (residual charge of the warm start, potential the solve adds
    to it); each pair satisfies A delta = r to the BF16 storage of its values
    (checked here in numpy with the 7-point operator). The network trained on
    them (mode 2: warm start plus its correction) starts the first solve of
    later steps from a smaller residual than the warm start, and the solve
    still ends at 1e-8. The reference warm-starts from the previous phi
    (RootGridField.cpp:41-103)."""
    n = 33
    # 16 ppc of size 4 and eps0 = 6400: w_pe dt = 0.1 and a Debye length of one cell for v_th = 0.1
    ctx = F.Context(F.make_desc(n, n, n, [(-1.0, 1.0, 4.0), (1.0, 100.0, 4.0)], eps0=6400.0, precision=F.FP32))
    lo, hi = (0.0, 0.0, 0.0), (n - 1.0,) * 3
    ctx.load_maxwellian(0, (n - 1) ** 3 * 16, 11, lo, hi, (0.0, 0.0, 0.0), 0.1)
    ctx.load_maxwellian(1, (n - 1) ** 3 * 16, 12, lo, hi, (0.0, 0.0, 0.0), 0.01)
    ctx.initialize()
    g = GU.Guess(ctx, seed=5)
    g.set_loss(1)
    ntrain, nheld = 40, 8
    g.dataset_alloc(ntrain + nheld)
    for _ in range(2):
        ctx.main_loop_es()
    g.set_mode(3)
    while g.recorded() < ntrain + nheld:
        ctx.main_loop_es()
    g.set_mode(0)
    assert g.recorded() == ntrain + nheld
    vn = n - 2  # voxels of interior nodes
    for i in (0, ntrain + nheld - 1):
        r, mask, d = g.dataset_get(i)
        r = GU.from_bf16(r).astype(np.float64)[:vn, :vn, :vn]
        d = GU.from_bf16(d).astype(np.float64)[:vn, :vn, :vn]
        assert not mask.any()
        assert abs(np.sqrt(np.mean(r ** 2)) - 1.0) < 1e-2  # normalised by its rms
        err = np.linalg.norm(_lap7(d) - r) / np.linalg.norm(r)
        assert err < 3e-2, err
    before = g.eval(ntrain, ntrain + nheld)
    g.train(ntrain, epochs=30, lr=2e-3, seed=3)
    after = g.eval(ntrain, ntrain + nheld)
    print(f"[guess] warm-start pairs: held-out loss {before:.4g} -> {after:.4g}")
    assert after < 0.5 * before
    starts = {}
    for mode in (0, 2):
        g.set_mode(mode)
        ctx.reset_timers()
        for _ in range(4):
            ctx.main_loop_es()
            assert F.solver_stats(ctx)["last_relres"] <= 1e-8
        starts[mode] = F.solver_stats(ctx)["start_relres_first"]
    print(f"[guess] first solve starts at relres {starts[0]:.3g} (warm) and {starts[2]:.3g} (warm + network)")
    assert starts[2] < 0.5 * starts[0]
    ctx.close()
