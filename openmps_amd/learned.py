"""Learned CG warm start: model and snapshot-dataset wrappers, snapshot
generation from the FP64 solver, training on the GPU, and weight files.

Everything numeric runs in libmps_hip.so (mps_mlp.hip): the feature kernel,
the BF16-MFMA forward and backward passes and Adam.  This module only drives
them.  Weight files are .npz with the host layout of mps_model_get_params:
w1 [256][132], b1 [256], w2 [256][256], b2 [256], w3 [256], b3 [1].
"""
from __future__ import annotations

import ctypes
import os
import sys
import time

import numpy as np

from . import _lib as L
from .computer import Computer, _raise, make_environment_from  # noqa: F401

FEAT = 132
HID = 256
PARAMS = ("w1", "b1", "w2", "b2", "w3", "b3")
SHAPES = {"w1": (HID, FEAT), "b1": (HID,), "w2": (HID, HID), "b2": (HID,), "w3": (HID,), "b3": (1,)}
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models",
                               "warmstart_mlp.npz")


class Model:
    """The warm-start MLP on one GPU."""

    def __init__(self, seed: int = 0, device: int = 0):
        h = ctypes.c_void_p()
        _raise(L.lib().mps_model_create(device, seed, ctypes.byref(h)), None)
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            L.lib().mps_model_destroy(self._h)
            self._h = None

    def params(self) -> dict:
        out = {k: np.empty(SHAPES[k], dtype=np.float32) for k in PARAMS}
        _raise(L.lib().mps_model_get_params(self._h, *[out[k].ctypes.data for k in PARAMS]), None)
        return out

    def set_params(self, p: dict) -> None:
        arrs = [np.ascontiguousarray(p[k], dtype=np.float32).reshape(SHAPES[k]) for k in PARAMS]
        _raise(L.lib().mps_model_set_params(self._h, *[a.ctypes.data for a in arrs]), None)

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty(x.shape[0], dtype=np.float32)
        _raise(L.lib().mps_model_forward(self._h, x.shape[0], x.ctypes.data, y.ctypes.data), None)
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_float()
        g = {k: np.empty(SHAPES[k], dtype=np.float32) for k in PARAMS}
        _raise(L.lib().mps_model_gradients(self._h, x.shape[0], x.ctypes.data, t.ctypes.data,
                                           ctypes.byref(loss), *[g[k].ctypes.data for k in PARAMS]),
               None)
        return float(loss.value), g

    def train(self, ds: "Dataset", epochs: int, batch: int = 8192, lr: float = 1e-3,
              seed: int = 0) -> np.ndarray:
        losses = np.zeros(epochs)
        _raise(L.lib().mps_model_train(self._h, ds._h, epochs, batch, lr, seed,
                                       losses.ctypes.data), None)
        return losses

    def save(self, path: str) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, **self.params())

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS, device: int = 0) -> "Model":
        m = cls(0, device)
        with np.load(path, allow_pickle=False) as f:
            m.set_params({k: f[k] for k in PARAMS})
        return m


class Dataset:
    """Device-resident (features, target) snapshot rows."""

    def __init__(self, capacity: int, device: int = 0):
        h = ctypes.c_void_p()
        _raise(L.lib().mps_dataset_create(device, capacity, ctypes.byref(h)), None)
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            L.lib().mps_dataset_destroy(self._h)
            self._h = None

    def __len__(self) -> int:
        n = ctypes.c_size_t()
        _raise(L.lib().mps_dataset_size(self._h, ctypes.byref(n)), None)
        return n.value

    def rows(self, first: int = 0, count: int | None = None):
        count = len(self) - first if count is None else count
        x = np.empty((count, FEAT), dtype=np.float32)
        t = np.empty(count, dtype=np.float32)
        _raise(L.lib().mps_dataset_get(self._h, first, count, x.ctypes.data, t.ctypes.data), None)
        return x, t


def attach_model(c: Computer, m: Model | None) -> None:
    _raise(L.lib().mps_attach_model(c._h, m._h if m is not None else None), c._h)
    c._model_ref = m


def attach_dataset(c: Computer, ds: Dataset | None, fraction: float, seed: int) -> None:
    _raise(L.lib().mps_attach_dataset(c._h, ds._h if ds is not None else None, fraction, seed),
           c._h)
    c._dataset_ref = ds


def features(c: Computer) -> np.ndarray:
    out = np.empty((c.size(), FEAT), dtype=np.float32)
    _raise(L.lib().mps_get_features(c._h, out.ctypes.data), c._h)
    return out


def snapshot_scenes(count: int, seed: int):
    """Seeded jittered dam-break scenes at l0 = 0.7 mm (the 1M workload's
    resolution) with columns of varying size and aspect ratio."""
    from . import scenes
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        l0 = 0.0007
        col_w = rng.uniform(0.06, 0.14)
        col_h = col_w * rng.uniform(1.0, 2.5)
        tank_w, tank_h = 2.5 * col_w, 1.2 * col_h
        parts, dom = scenes.dam_break(tank_w, tank_h, col_w, col_h, l0, jitter=0.02,
                                      seed=int(rng.integers(1 << 31)))
        env = dict(max_dt=1e-4, courant=scenes.COURANT, g=scenes.G, rho=scenes.RHO, nu=scenes.NU,
                   r_e_by_l0=2.1, l0=l0, r_lap_by_l0=3.1, **dom)
        out.append((parts, env))
    return out


def generate(ds: Dataset, runs: int, steps: int, fraction: float, seed: int,
             log=print) -> None:
    """FP64-solver snapshots: `runs` seeded scenes x `steps` steps each; at
    every step a `fraction` of the alive rows is appended to ds."""
    for k, (parts, env) in enumerate(snapshot_scenes(runs, seed)):
        c = Computer(1e-10, make_environment_from(env))
        c.set_solver(kind=L.SOLVER_CG_FP64, stop_rule=L.STOP_RHS_NORM, tol=1e-10)
        c.add_particles(parts)
        attach_dataset(c, ds, fraction, seed * 1000 + k)
        t0 = time.time()
        for s in range(steps):
            c.forward_time(env["max_dt"])
            if (s + 1) % 100 == 0:
                log(f"  snapshot run {k + 1}/{runs}: {len(parts['x'])} particles, step {s + 1}/"
                    f"{steps}, {time.time() - t0:.1f} s, dataset {len(ds)} rows")
        attach_dataset(c, None, 0.0, 0)
        del c


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description="train the warm-start MLP on FP64 solver snapshots")
    ap.add_argument("--runs", type=int, default=10)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--fraction", type=float, default=0.05)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--capacity", type=int, default=12_000_000)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    ds = Dataset(args.capacity)
    log(f"generating {args.runs} x {args.steps} snapshots ...")
    generate(ds, args.runs, args.steps, args.fraction, args.seed, log)
    m = Model(args.seed)
    log(f"training on {len(ds)} rows ...")
    t0 = time.time()
    for e in range(args.epochs):
        lr = args.lr * (0.5 ** (e // 4))
        loss = m.train(ds, 1, args.batch, lr, args.seed + e)[0]
        log(f"  epoch {e + 1}/{args.epochs}: lr {lr:.2e}, mean loss {loss:.4e} "
            f"({time.time() - t0:.1f} s)")
    m.save(args.out)
    log(f"saved {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
