"""lfd_stokes_solve_noalloc (the allocation-free variant of the drop-in's
lfd_stokes_fill_done + lfd_stokes_solve) through the C ABI, against the
existing entry points on the same matrix: two calls in a row, the second
with a smaller system, each equal to the drop-in's solve with the same
solver (6x6 block-Jacobi) to the solve tolerance and to the oracle's FP64
direct solve within 1e-8; sizes beyond the buffer and an unsorted block
list are refused with an error.  The device arrays are the test's own,
from the HIP runtime the library itself loaded.
"""
import ctypes

import numpy as np
import pytest

from tests.conftest import conf

pytestmark = pytest.mark.gpu


def _matrix(oracle, pos, rad, L, sd):
    R = oracle.System(pos, rad, L, sd)
    R.first_update()
    db, cp, rows, od = R.blocks()
    p0 = np.repeat(np.arange(len(rad)), np.diff(cp)).astype(np.int32)
    return db, cp, p0, rows.astype(np.int32), od, R.rhs()


class Dev:
    """Device copies of host arrays (hipMalloc / hipMemcpy of the HIP runtime
    liblfdem_amd.so links), freed together."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p.value)
        assert self.hip.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0      # host to device
        return p.value

    def get(self, p, n, dtype=np.float64):
        out = np.empty(n, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0        # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc(capi, D, db, p0, p1, od, b):
    n = len(db)
    t = [D.put(a) for a in (db, p0, p1, od, b)]
    x = D.put(np.full(6 * n, np.nan))
    st = capi.solve_noalloc(n, len(p0), t[0], t[1], t[2], t[3], t[4], x)
    return D.get(x, 6 * n), st


def _existing(capi, db, cp, p1, od, b):
    n = len(db)
    ss = capi.StokesSolver(n, opts=capi.default_opts(precond=0))
    ss.resetResistanceMatrix(len(p1), db, True)
    zero = np.zeros(15)
    for i in range(n - 1):
        ss.startNewColumn()
        for k in range(cp[i], cp[i + 1]):
            ss.addResistanceBlocks(i, p1[k], zero, zero, od[k])
    ss.matrixFillingDone()
    ss.resetRHS()
    ss.setRHS(b)
    u, w = ss.solve()
    return np.concatenate([u, w], axis=1).ravel()


def test_noalloc_matches_the_drop_in_on_two_calls(capi, oracle, golden):
    D = Dev()
    cases = [("N=500 phi=0.5", conf(500, 0.5, 1) + (np.zeros(3),)),
             ("N=200 reference fixture", (golden["dat_pos"], golden["dat_radius"], golden["dat_L"],
                                          golden["dat_disp"]))]
    for name, (pos, rad, L, sd) in cases:
        print("no-alloc solve:", name, flush=True)
        db, cp, p0, p1, od, b = _matrix(oracle, pos, rad, L, sd)
        x, st = _noalloc(capi, D, db, p0, p1, od, b)
        assert st["res_norm"] <= 1e-12 * st["rhs_norm"] * 1.0000001
        y = _existing(capi, db, cp, p1, od, b)
        diff = np.abs(x - y).max() / np.abs(y).max()
        print("  vs lfd_stokes_solve: max rel diff", diff, "bit-identical", bool(np.array_equal(x, y)), flush=True)
        assert diff <= 1e-9
        x_ref = oracle.Chol(len(rad), db, cp, p1, od).solve(b)
        assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) <= 1e-8
    D.free()


def test_noalloc_refuses_what_it_cannot_take(capi, oracle, golden):
    D = Dev()
    db, cp, p0, p1, od, b = _matrix(oracle, golden["dat_pos"], golden["dat_radius"], golden["dat_L"],
                                    golden["dat_disp"])
    n = len(db)
    ptrs = [D.put(a) for a in (db, p0, p1, od, b)]
    x = D.put(np.zeros(6 * n))
    print("no-alloc: size beyond the buffer", flush=True)
    with pytest.raises(capi.LfdError) as e:      # refused on the host before any device work
        capi.solve_noalloc(16385, 0, ptrs[0], None, None, None, ptrs[4], x)
    assert e.value.code == -5
    print("no-alloc: unsorted block list", flush=True)
    q0, q1 = p0.copy(), p1.copy()
    q0[[3, 4]], q1[[3, 4]] = q0[[4, 3]], q1[[4, 3]]
    assert (q0[3], q1[3]) > (q0[4], q1[4])
    tq = [D.put(a) for a in (q0, q1)]
    with pytest.raises(capi.LfdError) as e:
        capi.solve_noalloc(n, len(p0), ptrs[0], tq[0], tq[1], ptrs[3], ptrs[4], x)
    assert e.value.code == -1
    # the buffer is usable again after a refused call
    x2, st = _noalloc(capi, D, db, p0, p1, od, b)
    assert st["res_norm"] <= 1e-12 * st["rhs_norm"] * 1.0000001
    D.free()
