"""ctypes binding of the learned Poisson initial guess (fopis_guess_* and
fopis_nn_* in include/fopis.h).

The network and its training run in libfopis_hip.so; this module only moves
host buffers. BF16 tensors cross the boundary as uint16 bit patterns.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as F

# (cin, cout, level, upsampled input, relu) of the 15 convolutions, in the
# order of the flat parameter vector: per layer W[27][cin][cout] then b[cout].
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
NAMES = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "d2a", "d2b", "d1a", "d1b", "d0a", "d0b", "out"]


def param_slices():
    """(name, W slice, W shape, b slice) of each layer in the flat vector."""
    out, off = [], 0
    for name, (ci, co, _, _, _) in zip(NAMES, LAYERS):
        nw = 27 * ci * co
        out.append((name, slice(off, off + nw), (27, ci, co), slice(off + nw, off + nw + co)))
        off += nw + co
    return out, off


def to_bf16(a) -> np.ndarray:
    """float32 -> bf16 bit patterns, round to nearest even (as the device does)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def from_bf16(h) -> np.ndarray:
    return (np.ascontiguousarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def _bind(L):
    if getattr(L, "_guess_bound", False):
        return
    vp, i32, i64, u64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
    fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
    u16p, u8p, i32p = C.POINTER(C.c_uint16), C.POINTER(C.c_uint8), C.POINTER(C.c_int32)
    L.fopis_guess_init.argtypes = [vp, u64]
    L.fopis_guess_num_params.argtypes = [vp, C.POINTER(i64)]
    L.fopis_guess_get_params.argtypes = [vp, fp, dp]
    L.fopis_guess_set_params.argtypes = [vp, fp, d]
    L.fopis_guess_set_mode.argtypes = [vp, C.c_int]
    L.fopis_guess_recorded.argtypes = [vp, C.POINTER(i32)]
    L.fopis_guess_set_loss.argtypes = [vp, C.c_int]
    L.fopis_guess_apply.argtypes = [vp]
    L.fopis_guess_dataset_alloc.argtypes = [vp, i32]
    L.fopis_guess_dataset_store.argtypes = [vp, i32, dp]
    L.fopis_guess_dataset_get.argtypes = [vp, i32, u16p, u8p, u16p]
    L.fopis_guess_dataset_set.argtypes = [vp, i32, u16p, u8p, u16p]
    L.fopis_guess_grad.argtypes = [vp, i32, dp, fp]
    L.fopis_guess_train.argtypes = [vp, i32, i32, d, d, u64, dp]
    L.fopis_guess_eval.argtypes = [vp, i32, i32, dp]
    L.fopis_guess_synth_boxes.argtypes = [vp, i32, i32p, dp, u64]
    L.fopis_nn_conv3d.argtypes = [i32] * 7 + [u16p, fp, fp, u16p, u16p, i32, fp]
    L.fopis_nn_conv3d_backward.argtypes = [i32] * 6 + [u16p, fp, u16p, fp, fp, u16p]
    L._guess_bound = True


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _check(L, rc):
    if rc != 0:
        raise F.FopisError(rc, L.fopis_last_error().decode())


def net_grid(nx: int, ny: int, nz: int):
    """Voxels per axis of the network grid: interior nodes rounded up to 8."""
    return tuple((n - 2 + 7) // 8 * 8 for n in (nx, ny, nz))


class Guess:
    """The learned initial guess attached to one Context."""

    def __init__(self, ctx: "F.Context", seed: int = 1):
        self.ctx = ctx
        self.L = ctx.L
        _bind(self.L)
        _check(self.L, self.L.fopis_guess_init(ctx.h, seed))
        n = C.c_int64()
        _check(self.L, self.L.fopis_guess_num_params(ctx.h, C.byref(n)))
        self.nparams = n.value
        self.grid = net_grid(ctx.nx, ctx.ny, ctx.nz)
        self.nvox = int(np.prod(self.grid))

    def get_params(self):
        p = np.empty(self.nparams, np.float32)
        k = C.c_double()
        _check(self.L, self.L.fopis_guess_get_params(self.ctx.h, _p(p, C.c_float), C.byref(k)))
        return p, k.value

    def set_params(self, p, out_scale: float):
        a = np.ascontiguousarray(p, dtype=np.float32)
        _check(self.L, self.L.fopis_guess_set_params(self.ctx.h, _p(a, C.c_float), out_scale))

    def set_mode(self, mode: int):
        _check(self.L, self.L.fopis_guess_set_mode(self.ctx.h, mode))

    def recorded(self) -> int:
        """Warm-start pairs recorded since mode 3 was set."""
        n = C.c_int32()
        _check(self.L, self.L.fopis_guess_recorded(self.ctx.h, C.byref(n)))
        return n.value

    def set_loss(self, kind: int):
        """0: potential MSE; 1: the multigrid's starting residual ||A(k out) - rho_hat||^2."""
        _check(self.L, self.L.fopis_guess_set_loss(self.ctx.h, kind))

    def apply(self):
        _check(self.L, self.L.fopis_guess_apply(self.ctx.h))

    def dataset_alloc(self, capacity: int):
        _check(self.L, self.L.fopis_guess_dataset_alloc(self.ctx.h, capacity))

    def dataset_store(self, slot: int) -> float:
        r = C.c_double()
        _check(self.L, self.L.fopis_guess_dataset_store(self.ctx.h, slot, C.byref(r)))
        return r.value

    def dataset_get(self, slot: int):
        rho = np.empty(self.nvox, np.uint16)
        mask = np.empty(self.nvox, np.uint8)
        phi = np.empty(self.nvox, np.uint16)
        _check(self.L, self.L.fopis_guess_dataset_get(self.ctx.h, slot, _p(rho, C.c_uint16), _p(mask, C.c_uint8),
                                                      _p(phi, C.c_uint16)))
        shp = self.grid
        return rho.reshape(shp), mask.reshape(shp), phi.reshape(shp)

    def dataset_set(self, slot: int, rho_hat_bf16, mask, phi_hat_bf16):
        r = np.ascontiguousarray(rho_hat_bf16, dtype=np.uint16).ravel()
        m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
        p = np.ascontiguousarray(phi_hat_bf16, dtype=np.uint16).ravel()
        _check(self.L, self.L.fopis_guess_dataset_set(self.ctx.h, slot, _p(r, C.c_uint16), _p(m, C.c_uint8),
                                                      _p(p, C.c_uint16)))

    def export_pairs(self, path: str, begin: int, end: int) -> int:
        """Writes dataset slots [begin, end) to an HDF5 file laid out as the
        reference's resume files are (groups of plain datasets,
        RootGridIO.cpp:25-290): grid (int32, 3) and count (int32 scalar), then per pair i
        pairs/<i:06d>/rho_hat and phi_hat (float32, the network grid; the
        bf16 values widened exactly) and mask (uint8)."""
        from fopis_amd.h5 import H5File

        with H5File(path, "w") as f:
            f.write("grid", np.array(self.grid, np.int32))
            f.write("count", np.int32(end - begin))
            for i in range(begin, end):
                rho, mask, phi = self.dataset_get(i)
                k = f"pairs/{i - begin:06d}"
                f.write(k + "/rho_hat", from_bf16(rho).reshape(self.grid))
                f.write(k + "/mask", mask)
                f.write(k + "/phi_hat", from_bf16(phi).reshape(self.grid))
        return end - begin

    def import_pairs(self, path: str, slot0: int = 0) -> int:
        """Reads a file written by export_pairs into slots slot0, slot0 + 1, ...
        (bit for bit: the stored values are bf16 values). The network grid must
        match; the dataset must have room."""
        from fopis_amd.h5 import H5File

        with H5File(path, "r") as f:
            grid = tuple(int(v) for v in f.read("grid", np.int32))
            if grid != tuple(self.grid):
                raise ValueError(f"pairs for grid {grid}, network grid {tuple(self.grid)}")
            n = int(f.read("count", np.int32))
            for i in range(n):
                k = f"pairs/{i:06d}"
                self.dataset_set(slot0 + i, to_bf16(f.read(k + "/rho_hat", np.float32)), f.read(k + "/mask", np.uint8),
                                 to_bf16(f.read(k + "/phi_hat", np.float32)))
        return n

    def grad(self, slot: int):
        g = np.empty(self.nparams, np.float32)
        loss = C.c_double()
        _check(self.L, self.L.fopis_guess_grad(self.ctx.h, slot, C.byref(loss), _p(g, C.c_float)))
        return loss.value, g

    def train(self, npairs: int, epochs: int, lr: float = 1e-3, lr_end: float | None = None, seed: int = 7):
        out = np.zeros(epochs)
        lr_end = lr / 20 if lr_end is None else lr_end
        _check(self.L, self.L.fopis_guess_train(self.ctx.h, npairs, epochs, lr, lr_end, seed, _p(out, C.c_double)))
        return out

    def eval(self, begin: int, end: int) -> float:
        r = C.c_double()
        _check(self.L, self.L.fopis_guess_eval(self.ctx.h, begin, end, C.byref(r)))
        return r.value

    def synth_boxes(self, boxes, charges, seed: int = 0):
        """boxes: [(lo0, lo1, lo2, hi0, hi1, hi2)] glue cell ranges; charges: surface charge per node."""
        b = np.ascontiguousarray(boxes, dtype=np.int32).reshape(-1, 6)
        q = np.ascontiguousarray(charges, dtype=np.float64).ravel()
        _check(self.L, self.L.fopis_guess_synth_boxes(self.ctx.h, len(b), _p(b, C.c_int32), _p(q, C.c_double),
                                                      seed))


def conv3d(x_bf16, w, b, up=False, relu=False, add_bf16=None, reps=0):
    """One convolution on the device. x_bf16 [n0',n1',n2',cin] (half-size when up), w [27,cin,cout]
    float32, b [cout]; returns y [n0,n1,n2,cout] as bf16 bit patterns (and ms per launch if reps)."""
    L = F.load()
    _bind(L)
    x = np.ascontiguousarray(x_bf16, dtype=np.uint16)
    cin = x.shape[-1]
    n = tuple(s * 2 for s in x.shape[:3]) if up else x.shape[:3]
    wv = np.ascontiguousarray(w, dtype=np.float32)
    cout = wv.shape[-1]
    bv = np.ascontiguousarray(b, dtype=np.float32)
    y = np.empty(n + (cout,), np.uint16)
    ad = None if add_bf16 is None else np.ascontiguousarray(add_bf16, dtype=np.uint16)
    ms = C.c_float(0.0)
    _check(L, L.fopis_nn_conv3d(n[0], n[1], n[2], cin, cout, int(up), int(relu), _p(x, C.c_uint16),
                                _p(wv, C.c_float), _p(bv, C.c_float), _p(ad, C.c_uint16), _p(y, C.c_uint16), reps,
                                C.byref(ms)))
    return (y, ms.value) if reps else y


def conv3d_backward(x_bf16, w, gy_bf16, up=False, want_dx=True):
    """Weight, bias and (optionally) data gradient of conv3d for output gradient gy."""
    L = F.load()
    _bind(L)
    x = np.ascontiguousarray(x_bf16, dtype=np.uint16)
    gy = np.ascontiguousarray(gy_bf16, dtype=np.uint16)
    cin, cout = x.shape[-1], gy.shape[-1]
    n = gy.shape[:3]
    wv = np.ascontiguousarray(w, dtype=np.float32)
    dw = np.empty((27, cin, cout), np.float32)
    db = np.empty(cout, np.float32)
    dx = np.empty(n + (cin,), np.uint16) if want_dx else None
    _check(L, L.fopis_nn_conv3d_backward(n[0], n[1], n[2], cin, cout, int(up), _p(x, C.c_uint16), _p(wv, C.c_float),
                                         _p(gy, C.c_uint16), _p(dw, C.c_float), _p(db, C.c_float),
                                         _p(dx, C.c_uint16)))
    return dw, db, dx


def bench_layers(n: int = 128, reps: int = 20, seed: int = 1):
    """Per-layer time of the network's convolutions at an n^3 grid (random BF16
    input and weights), and of one whole guess (apply) on an (n+1)^3-node context."""
    rng = np.random.default_rng(seed)
    out = []
    for name, (ci, co, lvl, up, relu) in zip(NAMES, LAYERS):
        s = n >> lvl
        xs = (s // 2,) * 3 if up else (s,) * 3
        x = to_bf16(rng.uniform(-1, 1, xs + (ci,)).astype(np.float32))
        w = (rng.standard_normal((27, ci, co)) * 0.05).astype(np.float32)
        _, ms = conv3d(x, w, np.zeros(co, np.float32), up=bool(up), relu=bool(relu), reps=reps)
        flop = 2.0 * 27 * ci * co * s ** 3
        byt = 2.0 * (ci * (s // 2 if up else s) ** 3 + co * s ** 3 * (2 if name in ("d2a", "d1a", "d0a") else 1))
        out.append({"layer": name, "cin": ci, "cout": co, "grid": s, "up": bool(up), "us": ms * 1e3,
                    "tflops": flop / ms / 1e9, "gbs": byt / ms / 1e6})
    nn = n + 1
    ctx = F.Context(F.make_desc(nn, nn, nn, [(-1.0, 1.0, 1.0)], eps0=1.0, mg_precision=F.FP32))
    g = Guess(ctx, seed=seed)
    rho = np.zeros(ctx.field_shape(F.RHO))
    rho[2:nn, 2:nn, 2:nn] = rng.standard_normal((nn - 2,) * 3)
    ctx.set_field(F.RHO, rho)
    for _ in range(3):
        g.apply()
    ctx.synchronize()
    ctx.reset_timers()
    for _ in range(reps):
        g.apply()
    ctx.synchronize()
    apply_ms = ctx.phase_time("guess")[0]
    ctx.close()
    return out, apply_ms


if __name__ == "__main__":
    import json
    import sys

    n = int(sys.argv[1]) if len(sys.argv) > 1 else 128
    layers, apply_ms = bench_layers(n)
    for r in layers:
        print(json.dumps(r))
    print(json.dumps({"apply_ms": apply_ms, "grid": n}))
