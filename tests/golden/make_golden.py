"""Writes tests/golden/check_n200.npz from LF_DEM's own regression fixtures.

Reads (as data) LF_DEM/check/confs/D3N200VF0.45Bidi1.4_0.5Cubic_1.{dat,bin} and
the t=0 blocks of LF_DEM/check/reference/{int,par}_*_repulsive_{rate0.2r,
stress_stress5r}.dat.  Run once in a container where the reference tree is
mounted; the npz it writes is committed, so tests never read the reference.
"""
import os
import struct
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/LF_DEM/check"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "check_n200.npz")
STEM = "D3N200VF0.45Bidi1.4_0.5Cubic_1"


def read_dat(path):
    with open(path) as f:
        f.readline()
        hdr = f.readline().split()[1:]
        rows = [list(map(float, l.split())) for l in f if l.strip() and not l.startswith("#")]
    np1, np2 = int(hdr[0]), int(hdr[1])
    vf, lx, ly, lz = map(float, hdr[2:6])
    disp = float(hdr[8])
    a = np.array(rows)
    assert a.shape[0] == np1 + np2
    return a[:, :3], a[:, 3], np.array([lx, ly, lz]), np.array([disp, 0.0, 0.0]), vf


def read_bin(path):
    b = open(path, "rb").read()
    sw, fmt, n = struct.unpack_from("<iii", b, 0)
    assert sw == -1 and fmt == 2
    vf, lx, ly, lz, dx, dy = struct.unpack_from("<6d", b, 12)
    a = np.frombuffer(b, dtype="<f8", count=4 * n, offset=60).reshape(n, 4)
    ncont = struct.unpack_from("<I", b, 60 + 32 * n)[0]
    assert ncont == 0
    return a[:, :3].copy(), a[:, 3].copy(), np.array([lx, ly, lz]), np.array([dx, dy, 0.0]), vf


def first_block(path):
    """Rows of the first snapshot (after the first '# t ...' line) of an int_/par_ file."""
    rows, seen = [], 0
    header = None
    with open(path) as f:
        for l in f:
            if l.startswith("#"):
                parts = l[1:].split()
                if len(parts) >= 3 and all(_isnum(p) for p in parts):
                    seen += 1
                    if seen == 1:
                        header = list(map(float, parts))
                    if seen == 2:
                        break
                continue
            if seen == 1 and l.strip():
                rows.append(list(map(float, l.split())))
    return np.array(rows), np.array(header)


def _isnum(s):
    try:
        float(s)
        return True
    except ValueError:
        return False


def main():
    d = {}
    pos, rad, L, disp, vf = read_dat(os.path.join(REF, "confs", STEM + ".dat"))
    d.update(dat_pos=pos, dat_radius=rad, dat_L=L, dat_disp=disp, dat_vf=vf)
    pos, rad, L, disp, vf = read_bin(os.path.join(REF, "confs", STEM + ".bin"))
    d.update(bin_pos=pos, bin_radius=rad, bin_L=L, bin_disp=disp, bin_vf=vf)
    for tag, run in (("stress", "repulsive_stress_stress5r"), ("rate", "repulsive_rate0.2r")):
        ia, ih = first_block(os.path.join(REF, "reference", "int_%s_%s.dat" % (STEM, run)))
        pa, ph = first_block(os.path.join(REF, "reference", "par_%s_%s.dat" % (STEM, run)))
        d["%s_int" % tag] = ia
        d["%s_int_header" % tag] = ih
        d["%s_par" % tag] = pa
        d["%s_par_header" % tag] = ph
    np.savez_compressed(OUT, **d)
    print("wrote", OUT, {k: getattr(v, "shape", v) for k, v in d.items()})


if __name__ == "__main__":
    main()
