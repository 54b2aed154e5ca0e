"""The HIP step loop against the reference's multi-step outputs (GPU).

lfd_system_time_evolution (System::timeEvolution, System.cpp:1203-1279) on
the reference's two regression runs, read through the library's parameter
and configuration readers: every row of LF_DEM's data_ files (11 per run,
1e-3 strain apart) is reproduced to the printed 6 digits in every column
the current Simulation::outputData shares with the file (see
tests/golden_rows.py), with the oracle's step counts, and within 1e-9 of
the oracle's viscosity in every row.
"""
import numpy as np
import pytest

from tests import golden_rows as G

pytestmark = pytest.mark.gpu


def run(capi, oracle, tag, strains):
    c, S = G.setup(tag)
    out = []
    for mk in (lambda: capi.System(c.pos, c.radius, c.L, shear_disp=c.shear_disp, params=capi.default_params(**S.params)),
               lambda: oracle.System(c.pos, c.radius, c.L, shear_disp=c.shear_disp,
                                     params=oracle.default_params(**S.params))):
        X = mk()
        if S.control == "stress":
            X.set_control("stress", S.target_stress)
        X.first_update()
        out.append(G.rows(X, c, S, strains))
    return out


@pytest.mark.parametrize("tag,first", [("stress5r", 1), ("rate02r", 0)])
def test_golden_rows_on_gpu(capi, oracle, tag, first):
    _, g = G.golden(tag)
    print("golden rows on the GPU:", tag, flush=True)
    gpu, orc = run(capi, oracle, tag, g[first:, 1])
    assert [p[2] for p in gpu] == [p[2] for p in orc]
    assert G.mismatches(tag, gpu, first) == []
    for (rg, _, _), (ro, _, _) in zip(gpu, orc):
        assert rg["number of interaction"] == ro["number of interaction"]
        assert rg["viscosity"] == pytest.approx(ro["viscosity"], rel=1e-9)
        assert rg["shear rate"] == pytest.approx(ro["shear rate"], rel=1e-9)
