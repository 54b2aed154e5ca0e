"""LF_DEM checkpoints: the clock file chk_<name>.dat beside the configuration
conf_<name>.dat, and the restart from them.

Simulation::checkpoint (Simulation.cpp:604-612) writes the binary
configuration with the active contacts' states (Simulation.cpp:513-600; here
lf_dem_amd.formats.write_bin) and the clock (State::outputStateBinary,
States.cpp:9-29), little-endian:
    int format (1 = basic), double cumulated strain, double time,
    double time in simulation (output) units.
A restart (main.cpp:214-220, System.cpp:505-520) builds the system from the
configuration, finds its pairs, restores the contacts' states
(System::setContacts) and takes the clock from the checkpoint.  Velocities,
the time step and every other pair quantity are recomputed from the
positions, as LF_DEM does.
"""
import os
import struct

import numpy as np

from . import capi, formats

BASIC = 1                       # State::StateFileFormat::basic (States.h:10-12)


class Clock:
    """State::Clock (States.h:13-17)."""

    def __init__(self, time, time_in_simulation_units, cumulated_strain):
        self.time = float(time)
        self.time_in_simulation_units = float(time_in_simulation_units)
        self.cumulated_strain = float(cumulated_strain)

    def is_zero(self):
        """State::isZeroTimeChkp (States.cpp:74-79)."""
        return self.time == 0 and self.time_in_simulation_units == 0 and self.cumulated_strain == 0


def write_state(path, clock):
    """State::outputStateBinary (States.cpp:9-29)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<i3d", BASIC, clock.cumulated_strain, clock.time, clock.time_in_simulation_units))


def read_state(path):
    """State::readBasicCheckpoint (States.cpp:43-67)."""
    data = open(path, "rb").read()
    if len(data) < 4 + 24:
        raise formats.FormatError("checkpoint file too short: %s" % path)
    fmt, = struct.unpack_from("<i", data, 0)
    if fmt != BASIC:
        raise formats.FormatError("readBasicCheckpoint(): got incorrect binary format.")
    strain, time, time_sim = struct.unpack_from("<3d", data, 4)
    return Clock(time, time_sim, strain)


def names(simu_name, directory="."):
    """(configuration, clock) file names of Simulation::checkpoint."""
    return (os.path.join(directory, "conf_%s.dat" % simu_name), os.path.join(directory, "chk_%s.dat" % simu_name))


def configuration_of(system):
    """System::getConfiguration (System.cpp:650-670) of a device system:
    positions, radii, box, volume fraction, Lees-Edwards displacement and the
    active contacts' states."""
    st = system.state()
    c = system.contacts()
    vf = float(np.sum(4.0 / 3 * np.pi * system.radius ** 3) / np.prod(system.L))
    contacts = [(int(a), int(b), c["disp_tan"][k].copy(), c["disp_rolling"][k].copy())
                for k, (a, b) in enumerate(zip(c["p0"], c["p1"]))]
    return formats.Configuration(st["pos"], system.radius, system.L, vf, st["shear_disp"], contacts,
                                 fmt=formats.BIN_BASE_NEW)


def checkpoint(system, simu_name, directory=".", time_in_simulation_units=None):
    """Simulation::checkpoint: conf_<name>.dat and chk_<name>.dat; returns
    their paths.  The time in simulation units defaults to the internal time."""
    conf_path, chk_path = names(simu_name, directory)
    st = system.state()
    formats.write_bin(conf_path, configuration_of(system))
    tsim = st["time"] if time_in_simulation_units is None else time_in_simulation_units
    write_state(chk_path, Clock(st["time"], tsim, st["strain"]))
    return conf_path, chk_path


def resume(conf_path, chk_path, params=None, opts=None):
    """A device system restarted from a checkpoint: (system, clock)."""
    c = formats.read(conf_path)
    clock = read_state(chk_path)
    S = capi.System(c.pos, c.radius, c.L, c.shear_disp, params=params, opts=opts)
    S.first_update()
    if c.contacts:
        S.set_contacts([k[0] for k in c.contacts], [k[1] for k in c.contacts],
                       np.array([k[2] for k in c.contacts]), np.array([k[3] for k in c.contacts]))
    S.set_clock(clock.time, clock.cumulated_strain)
    return S, clock
