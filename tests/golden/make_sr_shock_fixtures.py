"""Converts the reference's SR MHD shock-tube solutions
(tst/regression/data/sr_mhd_shock{1,2,4}_hlld.vtk: legacy VTK, big-endian float32,
10,000 cells, conserved variables) into compact .npz fixtures.  Run once in a
container that has the reference tree; the fixtures are committed."""
import os
import re
import struct
import sys

import numpy as np

SRC = "/root/reference/tst/regression/data"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_vtk(path):
    raw = open(path, "rb").read()
    head = raw[:200].decode("ascii", "replace")
    t = float(re.search(r"time=([0-9.eE+-]+)", head).group(1))
    out = {"time": t}
    m = re.search(rb"X_COORDINATES (\d+) float\n", raw)
    nf = int(m.group(1))
    out["x1f"] = np.array(struct.unpack(">%df" % nf, raw[m.end():m.end() + 4 * nf]), dtype=np.float32)
    nc = int(re.search(rb"CELL_DATA (\d+)\n", raw).group(1))
    for m in re.finditer(rb"SCALARS (\S+) float\nLOOKUP_TABLE default\n", raw):
        name = m.group(1).decode()
        out[name] = np.array(struct.unpack(">%df" % nc, raw[m.end():m.end() + 4 * nc]),
                             dtype=np.float32)
    for m in re.finditer(rb"VECTORS (\S+) float\n", raw):
        name = m.group(1).decode().replace("-", "")
        v = np.array(struct.unpack(">%df" % (3 * nc), raw[m.end():m.end() + 12 * nc]),
                     dtype=np.float32).reshape(nc, 3)
        out[name] = v.T.copy()
    return out


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    for i in (1, 2, 4):
        d = read_vtk(os.path.join(src, "sr_mhd_shock%d_hlld.vtk" % i))
        np.savez_compressed(os.path.join(HERE, "sr_mhd_shock%d_hlld.npz" % i), **d)
        print(i, d["time"], {k: np.shape(v) for k, v in d.items()})
