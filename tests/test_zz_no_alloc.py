"""The allocation-free FP64 CG entry point (mps_cg_solve_csr_noalloc) against
the staged solve it varies (mps_set_ppe + mps_stage_ppe_solve) on the same
systems: the pressure systems of the 131.6k- and 2.2k-particle dam breaks
after one step, solved twice in a row (the second call smaller), each from a
zero guess with the scene's allowable residual.  Device buffers for the
caller's side come from the HIP runtime the library links (the entry point
itself allocates nothing)."""
import ctypes
import sys

import numpy as np
import pytest
import scipy.sparse

import cases
from openmps_amd import _lib as L, scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libmps_hip.so links."""

    def __init__(self):
        L.lib()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        self.bufs.append(p)
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def pressure_system(name):
    cfg = scenes.config(name)
    c = cases.make("gpu", cfg["env"], cfg["eps"])
    c.add_particles(cfg["particles"])
    c.forward_time(cfg["dt"])
    rp, col, val, b, _ = c.ppe()
    return cfg, rp, col, val, b


def staged(cfg, rp, col, val, b):
    c = cases.make("gpu", cfg["env"], cfg["eps"])
    c.set_ppe(rp, col, val, b, np.zeros(len(b)))
    c.stage("ppe_solve")
    return c.ppe()[4], c.cg_stats()["iterations"]


def noalloc(hip, rp, col, val, b, eps):
    n = len(b)
    x0 = np.zeros(n)
    d = [hip.put(a) for a in (rp.astype(np.uint64), col.astype(np.uint32), val, b, x0)]
    it = ctypes.c_int64(-1)
    err = ctypes.create_string_buffer(256)
    rc = L.lib().mps_cg_solve_csr_noalloc(0, n, d[0], d[1], d[2], d[3], d[4], eps, 0, None,
                                          ctypes.byref(it), err, 256)
    assert rc == L.MPS_OK, err.value.decode()
    x = hip.get(d[4], x0)
    hip.free()
    return x, it.value


def test_noalloc_cg_matches_staged_solve_twice():
    hip = Hip()
    for name in ("dambreak_125k", "dambreak_1k"):  # the second call is the smaller one
        cfg, rp, col, val, b = pressure_system(name)
        say(f"no-alloc CG on the {name} pressure system: {len(b)} rows, {len(val)} entries")
        x_na, it_na = noalloc(hip, rp, col, val, b, cfg["eps"])
        x_st, it_st = staged(cfg, rp, col, val, b)
        A = scipy.sparse.csr_matrix((val, col.astype(np.int64), rp.astype(np.int64)), shape=(len(b), len(b)))
        rel = np.linalg.norm(x_na - x_st) / np.linalg.norm(x_st)
        res = np.linalg.norm(b - A @ x_na) / np.linalg.norm(b)
        say(f"  iterations no-alloc {it_na}, staged {it_st}; |x_na - x_st|/|x_st| {rel:.2e}; "
            f"|b - A x_na|/|b| {res:.2e}")
        assert np.all(np.isfinite(x_na))
        assert rel <= 1e-6
        assert abs(it_na - it_st) <= max(2, 0.02 * it_st)
        assert res <= 10 * cfg["eps"]


def test_noalloc_rejects_more_rows_than_its_work_area():
    """n above MPS_NOALLOC_MAX_ROWS is refused before anything runs (no
    pointer is read)."""
    err = ctypes.create_string_buffer(256)
    rc = L.lib().mps_cg_solve_csr_noalloc(0, (1 << 20) + 1, None, None, None, None, None, 1e-10, 0,
                                          None, None, err, 256)
    assert rc == L.MPS_ERR_BAD_ARG
    assert b"exceed" in err.value
