"""Checkpoint clock files (chk_<name>.dat, States.cpp:9-79) on the CPU."""
import struct

import pytest

from lf_dem_amd import checkpoint, formats


def test_clock_file_layout(tmp_path):
    """int format 1, then cumulated strain, time, time in simulation units
    (State::outputStateBinary writes them in this order, unpadded)."""
    p = str(tmp_path / "chk_x.dat")
    checkpoint.write_state(p, checkpoint.Clock(time=2.5, time_in_simulation_units=0.5, cumulated_strain=1.25))
    data = open(p, "rb").read()
    assert data == struct.pack("<i", 1) + struct.pack("<3d", 1.25, 2.5, 0.5)
    c = checkpoint.read_state(p)
    assert (c.time, c.time_in_simulation_units, c.cumulated_strain) == (2.5, 0.5, 1.25)
    assert not c.is_zero() and checkpoint.Clock(0, 0, 0).is_zero()


def test_clock_file_rejects_other_formats(tmp_path):
    p = str(tmp_path / "chk_bad.dat")
    open(p, "wb").write(struct.pack("<i3d", 2, 0, 0, 0))
    with pytest.raises(formats.FormatError, match="incorrect binary format"):
        checkpoint.read_state(p)


def test_checkpoint_names():
    assert checkpoint.names("run1", "/d") == ("/d/conf_run1.dat", "/d/chk_run1.dat")
