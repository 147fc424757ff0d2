"""MHD blast wave (src/pgen/blast.cpp, Cartesian branch) and the seeded ensemble of
blast waves of uneven sizes.

ProblemGenerator (blast.cpp:34-211): uniform density `damb`, pressure `pamb` outside
radius `radius` and `prat*pamb` inside (`ramp = 0`, `drat = 1`), zero velocity, uniform
face field, and total energy p/(gamma-1) + B^2/2.  The reference orients the field in
the x1-x2 plane by `angle` (blast.cpp:149-210); here the field direction is any unit
vector, which the ensemble draws isotropically.
"""
from __future__ import annotations

import numpy as np

from ..mesh import MeshDesc, block_coords, new_block_arrays, IDN, IEN

EDGES = (32, 48, 64, 96, 128)


def blast_mesh(n, bn, **kw):
    """inputs/mhd/athinput.blast: [-1,1]^3, periodic, gamma 5/3, CFL 0.3, VL2 + PLM."""
    kw.setdefault("gamma", 5.0 / 3.0)
    kw.setdefault("cfl", 0.3)
    return MeshDesc(nx=(n, n, n), bnx=(bn, bn, bn), xmin=(-1.0, -1.0, -1.0),
                    xmax=(1.0, 1.0, 1.0), **kw)


class Blast:
    def __init__(self, m: MeshDesc, prat=100.0, pamb=0.1, damb=1.0, radius=0.1, b0=1.0,
                 bdir=(np.cos(np.pi / 6), np.sin(np.pi / 6), 0.0), center=(0.0, 0.0, 0.0)):
        self.m = m
        self.prat, self.pamb, self.damb, self.radius = prat, pamb, damb, radius
        bdir = np.asarray(bdir, dtype=np.float64)
        self.b = b0 * bdir / np.linalg.norm(bdir)
        self.center = center

    def block(self, gid):
        m = self.m
        a = new_block_arrays(m)
        c = block_coords(m, gid)
        ng = m.ng
        n1, n2, n3 = m.bnx
        x = c.x1v[ng:ng + n1][None, None, :]
        y = c.x2v[ng:ng + n2][None, :, None]
        z = c.x3v[ng:ng + n3][:, None, None]
        x0, y0, z0 = self.center
        rad = np.sqrt((x - x0) ** 2 + (y - y0) ** 2 + (z - z0) ** 2)
        pres = np.where(rad < self.radius, self.prat * self.pamb, self.pamb)
        sl = (slice(ng, ng + n3), slice(ng, ng + n2), slice(ng, ng + n1))
        u = a["u"]
        u[(IDN,) + sl] = self.damb
        bx, by, bz = self.b
        u[(IEN,) + sl] = pres / (m.gamma - 1.0) + 0.5 * (bx * bx + by * by + bz * bz)
        a["b1"][ng:ng + n3, ng:ng + n2, ng:ng + n1 + 1] = bx
        a["b2"][ng:ng + n3, ng:ng + n2 + 1, ng:ng + n1] = by
        a["b3"][ng:ng + n3 + 1, ng:ng + n2, ng:ng + n1] = bz
        return a


def blast_ensemble(nmembers=48, seed=20261204, edges=EDGES, block_sizes=(16, 32), **kw):
    """Seeded ensemble: edge uniform in `edges`, MeshBlocks of 16^3 or 32^3 (16^3 where
    the edge is not a multiple of 32), P_in/P_out log-uniform in [10, 1000], |B0|
    uniform in [0.1, 1] with an isotropic direction.  Returns [(MeshDesc, Blast)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nmembers):
        n = int(rng.choice(edges))
        bn = int(rng.choice([b for b in block_sizes if n % b == 0]))
        prat = float(np.exp(rng.uniform(np.log(10.0), np.log(1000.0))))
        b0 = float(rng.uniform(0.1, 1.0))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        m = blast_mesh(n, bn, **kw)
        out.append((m, Blast(m, prat=prat, b0=b0, bdir=d)))
    return out
