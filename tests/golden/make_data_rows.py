"""Writes tests/golden/data_rows.npz: the rows and column names of LF_DEM's
two multi-step regression outputs,

  check/reference/data_D3N200VF0.45Bidi1.4_0.5Cubic_1_repulsive_stress_stress5r.dat
      (`LF_DEM -s 5r confs/D3N200VF0.45Bidi1.4_0.5Cubic_1.dat repulsive_stress.txt`)
  check/reference/data_D3N200VF0.45Bidi1.4_0.5Cubic_1_repulsive_rate0.2r.dat
      (`LF_DEM -r 0.2r confs/D3N200VF0.45Bidi1.4_0.5Cubic_1.bin repulsive.txt`)

read as data.  Run once where the reference tree is mounted; the npz is
committed, so the tests never read the reference.
"""
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/LF_DEM/check/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data_rows.npz")
STEM = "D3N200VF0.45Bidi1.4_0.5Cubic_1"


def read(path):
    names, rows = [], []
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                head = line[1:].split(":", 1)
                if len(head) == 2 and head[0].strip().isdigit():
                    names.append(head[1].strip())
                continue
            if line.strip():
                rows.append([float(x) for x in line.split()])
    return np.array(names), np.array(rows)


def main():
    d = {}
    for tag, run in (("stress5r", "repulsive_stress_stress5r"), ("rate02r", "repulsive_rate0.2r")):
        names, rows = read(os.path.join(REF, "data_%s_%s.dat" % (STEM, run)))
        assert rows.shape[1] == len(names)
        d[tag + "_names"] = names
        d[tag + "_rows"] = rows
    np.savez_compressed(OUT, **d)
    print("wrote", OUT, {k: v.shape for k, v in d.items()})


if __name__ == "__main__":
    main()
