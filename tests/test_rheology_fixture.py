"""The committed rheology series (tests/golden/rheology_n200.npz) is the
oracle's: its first two outputs are recomputed here (CPU, a few seconds)."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def test_rheology_fixture_is_the_oracles(golden):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_rheology", os.path.join(HERE, "golden", "make_rheology.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    ref = np.load(os.path.join(HERE, "golden", "rheology_n200.npz"))
    visc, steps, strain = mod.series(golden["dat_pos"], golden["dat_radius"], golden["dat_L"], golden["dat_disp"],
                                     outputs=2)
    assert steps.tolist() == ref["steps"][:2].tolist()
    assert np.array_equal(visc, ref["viscosity"][:2])
    assert np.allclose(strain, [0.05, 0.1], rtol=0, atol=1e-12)
