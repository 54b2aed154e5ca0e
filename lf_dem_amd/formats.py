"""LF_DEM configuration files: text and binary readers and writers.

Text configuration (GenerateInitConfig.cpp:215-240 writes it,
Configuration.h:300-345 reads it): two header lines, the first naming the
fields (`# np1 np2 vf lx ly lz vf1 vf2 disp`), the second giving their values
(`# 100 100 0.45 ...`); then one particle per line, `x y z radius`.  The
reader takes lx, ly, lz and vf from the header by name, and the Lees-Edwards
displacement from `dispx` / `dispy` (0 when absent: the `disp` field the
generator writes is not read back).  A `format` field or a `radius_in` field
marks the fixed-velocity and circular-Couette layouts (Configuration.h:
206-224), which this library does not run.

Binary configuration (Simulation.cpp:513-600 writes it, Configuration.h:
179-250 reads it), little-endian:
  v2: int -1, int 2, int np, double vf, lx, ly, lz, disp_x, disp_y,
      np x (double x, y, z, radius), unsigned ncontact,
      ncontact x (unsigned p0, p1, double disp_tan[3], disp_rolling[3]);
  v1: as v2 without the two leading ints (and possibly 16-bit p0, p1, told
      apart by p0 > np as LF_DEM does, Configuration.h:60-110);
  v3: the fixed-velocity layout (-1, 3, np, np_fixed, ...), not run here.
"""
import struct

import numpy as np

BIN_BASE_OLD, BIN_BASE_NEW, BIN_FIXED_VEL = 1, 2, 3
TXT_BASE_OLD, TXT_BASE_NEW, TXT_FIXED_VEL, TXT_COUETTE = 4, 5, 6, 7


class FormatError(ValueError):
    pass


class Configuration:
    """Positions (np, 3), radii (np,), box L (3,), volume fraction, Lees-Edwards
    displacement (3,), contact states (list of (p0, p1, disp_tan[3], disp_rolling[3]))."""

    def __init__(self, pos, radius, L, vf, shear_disp=(0.0, 0.0, 0.0), contacts=None, fmt=None):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        self.radius = np.ascontiguousarray(radius, dtype=np.float64)
        self.L = np.asarray(L, dtype=np.float64)
        self.vf = float(vf)
        self.shear_disp = np.asarray(shear_disp, dtype=np.float64)
        self.contacts = list(contacts or [])
        self.format = fmt

    @property
    def np(self):
        return len(self.radius)


def _split(line):
    return line.split()


def _meta(l1, l2):
    """getConfMetaData (Configuration.h:160-172): field names of the first
    header line against the values of the second, skipping the leading '#'."""
    a, b = _split(l1), _split(l2)
    if len(a) != len(b):
        raise FormatError("Ill-formed header in the configuration file")
    return dict(zip(a[1:], b[1:]))


def txt_format(path):
    """getTxtConfigurationFileFormat (Configuration.h:206-224)."""
    with open(path) as f:
        meta = _meta(f.readline(), f.readline())
    if "format" in meta:
        return int(meta["format"])
    if "radius_in" in meta:
        return TXT_COUETTE
    return TXT_BASE_OLD


def read_txt(path):
    """readTxtBaseConfiguration (Configuration.h:300-345)."""
    fmt = txt_format(path)
    if fmt not in (TXT_BASE_OLD, TXT_BASE_NEW):
        raise FormatError("text configuration format %d (fixed-velocity / Couette) is not supported" % fmt)
    with open(path) as f:
        meta = _meta(f.readline(), f.readline())
        vals = []
        for line in f:
            vals.extend(line.split())
    for key in ("lx", "ly", "lz", "vf"):
        if key not in meta:
            raise FormatError("parameter '%s' not found in the header of the configuration file" % key)
    n = len(vals) // 4
    a = np.array([float(v) for v in vals[:4 * n]], dtype=np.float64).reshape(n, 4)
    L = [float(meta["lx"]), float(meta["ly"]), float(meta["lz"])]
    disp = [float(meta.get("dispx", "0")), float(meta.get("dispy", "0")), 0.0]
    return Configuration(a[:, :3], a[:, 3], L, float(meta["vf"]), disp, fmt=fmt)


def write_txt(path, conf, np1=None, vf1=None, vf2=None):
    """The generator's layout (GenerateInitConfig.cpp:222-236): header
    `# np1 np2 vf lx ly lz vf1 vf2 disp`, 15 significant digits."""
    n = conf.np
    if np1 is None:
        np1 = int(np.sum(conf.radius == conf.radius.min()))
    vol = 4.0 / 3 * np.pi * conf.radius ** 3
    box = float(np.prod(conf.L))
    small = conf.radius == conf.radius.min()
    vf1 = float(vol[small].sum() / box) if vf1 is None else vf1
    vf2 = float(vol[~small].sum() / box) if vf2 is None else vf2
    g = lambda x: "%.15g" % x
    with open(path, "w") as f:
        f.write("# np1 np2 vf lx ly lz vf1 vf2 disp\n")
        f.write("# %d %d %s %s %s %s %s %s %s\n" % (np1, n - np1, g(conf.vf), g(conf.L[0]), g(conf.L[1]),
                                                     g(conf.L[2]), g(vf1), g(vf2), g(0)))
        for (x, y, z), a in zip(conf.pos, conf.radius):
            f.write("%s %s %s %s \n" % (g(x), g(y), g(z), g(a)))


def bin_format(data):
    """getBinaryConfigurationFileFormat (Configuration.h:188-204)."""
    sw, = struct.unpack_from("<i", data, 0)
    if sw == -1:
        return struct.unpack_from("<i", data, 4)[0]
    return BIN_BASE_OLD


def _contacts(data, off, n):
    """readContactStatesBStream (Configuration.h:60-110)."""
    ncont, = struct.unpack_from("<I", data, off)
    off += 4
    out = []
    if ncont:
        p0, = struct.unpack_from("<I", data, off)
        ushort = p0 > n
        rec = "<HH6d" if ushort else "<II6d"
        size = struct.calcsize(rec)
        for k in range(ncont):
            v = struct.unpack_from(rec, data, off + k * size)
            out.append((int(v[0]), int(v[1]), np.array(v[2:5]), np.array(v[5:8])))
    return out


def read_bin(path):
    """readBinaryBaseConfiguration (Configuration.h:226-250)."""
    data = open(path, "rb").read()
    fmt = bin_format(data)
    if fmt not in (BIN_BASE_OLD, BIN_BASE_NEW):
        raise FormatError("binary configuration format %d (fixed-velocity) is not supported" % fmt)
    off = 8 if fmt == BIN_BASE_NEW else 0
    n, = struct.unpack_from("<i", data, off)
    vf, lx, ly, lz, dx, dy = struct.unpack_from("<6d", data, off + 4)
    off += 4 + 48
    a = np.frombuffer(data, dtype="<f8", count=4 * n, offset=off).reshape(n, 4)
    off += 32 * n
    contacts = _contacts(data, off, n)
    return Configuration(a[:, :3].copy(), a[:, 3].copy(), [lx, ly, lz], vf, [dx, dy, 0.0], contacts, fmt=fmt)


def write_bin(path, conf):
    """outputConfigurationBinary (Simulation.cpp:513-600), format v2."""
    parts = [struct.pack("<iii", -1, BIN_BASE_NEW, conf.np),
             struct.pack("<6d", conf.vf, conf.L[0], conf.L[1], conf.L[2], conf.shear_disp[0], conf.shear_disp[1]),
             np.ascontiguousarray(np.column_stack([conf.pos, conf.radius]), dtype="<f8").tobytes(),
             struct.pack("<I", len(conf.contacts))]
    for p0, p1, dt, dr in conf.contacts:
        parts.append(struct.pack("<II6d", p0, p1, *np.asarray(dt, dtype=float), *np.asarray(dr, dtype=float)))
    with open(path, "wb") as f:
        f.write(b"".join(parts))


def read(path):
    """Text or binary by content (binary when the first bytes are not a '#' header)."""
    with open(path, "rb") as f:
        head = f.read(1)
    return read_txt(path) if head == b"#" else read_bin(path)
