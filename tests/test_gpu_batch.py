"""Batched and served instances (lfd_batch_*) through the C ABI.

Instances in a batch share no arithmetic: each one's cell grid, pair list,
matrix rows, Krylov sums and time step are computed from its own data in the
same order as when it runs alone.  The bar is therefore bit-identical
trajectories (tolerance 0) between an instance alone and the same instance in
a mix, including instances admitted and retired between steps (continuous
batching).  Parity of a lone instance with the oracle is covered by
test_gpu_parity.py.
"""
import numpy as np
import pytest

from tests.conftest import conf

pytestmark = pytest.mark.gpu

# (n, phi, seed, repulsion): uneven sizes, densities and rate (repulsion = 1/rate)
MIX = [(300, 0.50, 11, 5.0), (700, 0.55, 12, 0.5), (1100, 0.45, 13, 20.0)]


def _inst(capi, n, phi, seed, rep):
    pos, rad, L = conf(n, phi, seed)
    return pos, rad, L, capi.default_params(repulsion=rep)


def _alone(capi, spec, nsteps):
    pos, rad, L, p = _inst(capi, *spec)
    S = capi.System(pos, rad, L, np.zeros(3), params=p)
    S.first_update()
    S.step(nsteps)
    return S.state(), S.pairs()


def _same(a, b):
    for k in ("pos", "vel", "ang_vel", "na_vel", "na_ang_vel", "shear_disp"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("dt", "time", "strain"):
        assert a[k] == b[k], k


def _caps(capi, specs):
    slots = cells = 0
    for s in specs:
        pos, rad, L, p = _inst(capi, *s)
        sl, ce, _ = capi.footprint(len(rad), rad, L, p)
        slots += sl
        cells += ce
    return slots, cells


def test_batch_equals_alone(capi):
    nsteps = 4
    ref = [_alone(capi, s, nsteps) for s in MIX]
    slots, cells = _caps(capi, MIX)
    B = capi.Batch(slots, cells, len(MIX))
    ids = []
    for s in MIX:
        pos, rad, L, p = _inst(capi, *s)
        ids.append(B.add(pos, rad, L, params=p))
    print("batch", B.info(), flush=True)
    B.step(nsteps)
    B.sync()
    for iid, (st, pr) in zip(ids, ref):
        _same(B.state(iid), st)
        bp = B.pairs(iid)
        assert np.array_equal(bp["p0"], pr["p0"]) and np.array_equal(bp["p1"], pr["p1"])
        assert np.array_equal(bp["disp_tan"], pr["disp_tan"])


def test_continuous_batching(capi):
    """A runs 2 steps alone in the batch, B joins for 3 steps, A leaves, B runs 2
    more, C joins in A's freed slots: every trajectory equals its lone run."""
    a, b, c = MIX[0], MIX[1], (260, 0.52, 14, 2.0)
    ra = _alone(capi, a, 5)[0]
    rb = _alone(capi, b, 5)[0]
    rc = _alone(capi, c, 2)[0]
    slots, cells = _caps(capi, [a, b])
    B = capi.Batch(slots, cells, 2)
    ia = B.add(*_inst(capi, *a)[:3], params=_inst(capi, *a)[3])
    B.step(2)
    ib = B.add(*_inst(capi, *b)[:3], params=_inst(capi, *b)[3])
    B.step(3)
    _same(B.state(ia), ra)
    B.remove(ia)
    assert B.info()["live"] == 1
    B.step(2)
    _same(B.state(ib), rb)
    # C fits in A's slots and cells; a third instance of B's size does not
    sl, ce, _ = capi.footprint(c[0], _inst(capi, *c)[1], _inst(capi, *c)[2])
    assert B.room(sl, ce)
    ic = B.add(*_inst(capi, *c)[:3], params=_inst(capi, *c)[3])
    assert not B.room(*capi.footprint(b[0], _inst(capi, *b)[1], _inst(capi, *b)[2])[:2])
    B.step(2)
    _same(B.state(ic), rc)


def test_admission_capacity(capi):
    a = MIX[0]
    pos, rad, L, p = _inst(capi, *a)
    sl, ce, by = capi.footprint(len(rad), rad, L, p)
    assert sl % 128 == 0 and sl >= len(rad) and by > 0
    B = capi.Batch(sl, ce, 2)
    B.add(pos, rad, L, params=p)
    with pytest.raises(capi.LfdError) as e:
        B.add(pos, rad, L, params=p)
    assert e.value.code == -5          # LFD_ERR_CAPACITY
    assert B.info()["live"] == 1
