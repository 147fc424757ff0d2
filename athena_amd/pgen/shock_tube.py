"""MHD shock tube (pgen/shock_tube.cpp) with the Ryu & Jones 2a states of
inputs/mhd/athinput.rj2a, and its L1 error against the Dai & Woodward solution
(shock_tube.cpp:50-261, Mesh::UserWorkAfterLoop).

The reference's regression test (tst/regression/scripts/tests/mhd/rj2a_shock.py) runs
the tube along x1, x2 and x3 at 256 and 512 cells with cfl_number 0.3, outflow faces
along the tube and periodic transverse faces; along x3 the mesh is 4 x 4 x nx3, a 3-D
run.  Here every orientation is a 3-D mesh with 4 transverse cells per direction.
"""
from __future__ import annotations

import math

import numpy as np

from ..mesh import IB1, IB2, IB3, IDN, IEN, IM1, IM2, IM3, MeshDesc, block_coords, \
    new_block_arrays

PI = 3.1415926535897932          # defs.hpp.in:140

# <problem> of inputs/mhd/athinput.rj2a: (d, p, u, v, w, bx, by, bz)
RJ2A_LEFT = (1.08, 0.95, 1.2, 0.01, 0.5, 0.5641895835477562, 1.0155412503859613,
             0.5641895835477562)
RJ2A_RIGHT = (1.0, 1.0, 0.0, 0.0, 0.0, 0.5641895835477562, 1.1283791670955125,
              0.5641895835477562)
RJ2A_GAMMA = float("1.666666666667")      # <hydro> gamma as written in the input file
RJ2A_TLIM = 0.2
RJ2A_CFL = 0.3                            # rj2a_shock.py overrides time/cfl_number


def rj2a_mesh(nx, shock_dir=3, ntrans=4, bnx=None, **kw) -> MeshDesc:
    """The mesh of rj2a_shock.py for one direction and resolution: [-0.5, 0.5]^3 with
    nx cells along the tube, ntrans across it, one MeshBlock unless bnx is given."""
    d = shock_dir - 1
    n = [ntrans] * 3
    n[d] = nx
    bc = ["periodic"] * 6
    bc[2 * d] = bc[2 * d + 1] = "outflow"
    return MeshDesc(nx=tuple(n), bnx=tuple(bnx or n), xmin=(-0.5,) * 3, xmax=(0.5,) * 3,
                    gamma=kw.pop("gamma", RJ2A_GAMMA), cfl=kw.pop("cfl", RJ2A_CFL),
                    xorder=kw.pop("xorder", 2), rsolver=kw.pop("rsolver", "hlld"),
                    bc=tuple(bc), **kw)


# (momentum slot of the tube's x, y, z velocity) and (face component carrying bx, by,
# bz) per shock direction (shock_tube.cpp:395-490, 503-532)
_MOM = {1: (IM1, IM2, IM3), 2: (IM2, IM3, IM1), 3: (IM3, IM1, IM2)}
_FACE = {1: (0, 1, 2), 2: (1, 2, 0), 3: (2, 0, 1)}


class ShockTube:
    """MeshBlock::ProblemGenerator of shock_tube.cpp for an MHD Riemann problem."""

    def __init__(self, m: MeshDesc, shock_dir=3, left=RJ2A_LEFT, right=RJ2A_RIGHT,
                 xshock=0.0):
        if shock_dir not in (1, 2, 3):
            raise ValueError("shock_dir must be 1, 2 or 3")
        self.m, self.dir, self.xshock = m, shock_dir, xshock
        self.left, self.right = left, right

    def block(self, gid):
        m = self.m
        c = block_coords(m, gid)
        a = new_block_arrays(m)
        ng = m.ng
        n1, n2, n3 = m.bnx
        s1, s2, s3 = slice(ng, ng + n1), slice(ng, ng + n2), slice(ng, ng + n3)
        xv = {1: c.x1v, 2: c.x2v, 3: c.x3v}[self.dir]
        gm1 = m.gamma - 1.0
        shape = (n3, n2, n1)
        # per active cell: True on the left state (x < xshock)
        along = xv[ng:ng + m.bnx[self.dir - 1]] < self.xshock
        if self.dir == 1:
            lft = np.broadcast_to(along[None, None, :], shape)
        elif self.dir == 2:
            lft = np.broadcast_to(along[None, :, None], shape)
        else:
            lft = np.broadcast_to(along[:, None, None], shape)
        L, R = self.left, self.right

        def pick(q):
            return np.where(lft, L[q], R[q])

        d, p = pick(0), pick(1)
        vx, vy, vz = pick(2), pick(3), pick(4)
        bx, by, bz = pick(5), pick(6), pick(7)
        mx, my, mz = _MOM[self.dir]
        u = a["u"]
        u[IDN, s3, s2, s1] = d
        u[mx, s3, s2, s1] = vx * d
        u[my, s3, s2, s1] = vy * d
        u[mz, s3, s2, s1] = vz * d
        e = p / gm1
        e = e + (0.5 * d) * (vx * vx + vy * vy + vz * vz)
        faces = [None, None, None]
        fx, fy, fz = _FACE[self.dir]
        faces[fx], faces[fy], faces[fz] = bx, by, bz
        a["b1"][s3, s2, s1] = faces[0]
        a["b2"][s3, s2, s1] = faces[1]
        a["b3"][s3, s2, s1] = faces[2]
        e = e + 0.5 * (faces[0] * faces[0] + faces[1] * faces[1] + faces[2] * faces[2])
        u[IEN, s3, s2, s1] = e
        # the faces at ie+1, je+1, ke+1 copy ie, je, ke (shock_tube.cpp:541-556)
        a["b1"][s3, s2, ng + n1] = a["b1"][s3, s2, ng + n1 - 1]
        a["b2"][s3, ng + n2, s1] = a["b2"][s3, ng + n2 - 1, s1]
        a["b3"][ng + n3, s2, s1] = a["b3"][ng + n3 - 1, s2, s1]
        return a


def rj2a_exact(r, tlim, gamma):
    """Dai & Woodward 1994 tables Ia and Ib at position r along the tube: conserved
    (d, mx, my, mz, e) and (bx, by, bz) as shock_tube.cpp:74-200 evaluates them."""
    xfp = 2.2638 * tlim
    xrp = (0.53432 + 1.0 / math.sqrt(PI * 1.309)) * tlim
    xsp = (0.53432 + 0.48144 / 1.309) * tlim
    xc = 0.57538 * tlim
    xsm = (0.60588 - 0.51594 / 1.4903) * tlim
    xrm = (0.60588 - 1.0 / math.sqrt(PI * 1.4903)) * tlim
    xfm = (1.2 - 2.3305 / 1.08) * tlim
    gm1 = gamma - 1.0
    s4 = math.sqrt(4.0 * PI)
    bx = 2.0 / s4
    if r > xfp:
        d0, mx, my, mz, by, bz, p = 1.0, 0.0, 0.0, 0.0, 4.0 / s4, 2.0 / s4, 1.0
    elif r > xrp:
        d0 = 1.3090
        mx, my, mz = 0.53432 * d0, -0.094572 * d0, -0.047286 * d0
        by, bz, p = 5.3452 / s4, 2.6726 / s4, 1.5844
    elif r > xsp:
        d0 = 1.3090
        mx, my, mz = 0.53432 * d0, -0.18411 * d0, 0.17554 * d0
        by, bz, p = 5.7083 / s4, 1.7689 / s4, 1.5844
    elif r > xc:
        d0 = 1.4735
        mx, my, mz = 0.57538 * d0, 0.047601 * d0, 0.24734 * d0
        by, bz, p = 5.0074 / s4, 1.5517 / s4, 1.9317
    elif r > xsm:
        d0 = 1.6343
        mx, my, mz = 0.57538 * d0, 0.047601 * d0, 0.24734 * d0
        by, bz, p = 5.0074 / s4, 1.5517 / s4, 1.9317
    elif r > xrm:
        d0 = 1.4903
        mx, my, mz = 0.60588 * d0, 0.22157 * d0, 0.30125 * d0
        by, bz, p = 5.5713 / s4, 1.7264 / s4, 1.6558
    elif r > xfm:
        d0 = 1.4903
        mx, my, mz = 0.60588 * d0, 0.11235 * d0, 0.55686 * d0
        by, bz, p = 5.0987 / s4, 2.8326 / s4, 1.6558
    else:
        d0 = 1.08
        mx, my, mz = 1.2 * d0, 0.01 * d0, 0.5 * d0
        by, bz, p = 3.6 / s4, 2.0 / s4, 0.95
    e0 = p / gm1
    e0 += 0.5 * ((mx * mx + my * my + mz * mz) / d0 + (bx * bx + by * by + bz * bz))
    return (d0, mx, my, mz, e0), (bx, by, bz)


def rj2a_errors(m: MeshDesc, u, bcc, shock_dir=3, tlim=RJ2A_TLIM):
    """L1 errors of MeshBlock 0 (u: (5, nc3, nc2, nc1) conserved, bcc: (3, ...)) in the
    order of shock-errors.dat (d, M1, M2, M3, E, B1c, B2c, B3c) and their RMS,
    normalised by the mesh's total cell count (shock_tube.cpp:177-261)."""
    ng = m.ng
    n1, n2, n3 = m.bnx
    c = block_coords(m, 0)
    im = _MOM[shock_dir]
    ib = {1: (IB1, IB2, IB3), 2: (IB2, IB3, IB1), 3: (IB3, IB1, IB2)}[shock_dir]
    xv = {1: c.x1v, 2: c.x2v, 3: c.x3v}[shock_dir]
    err = [0.0] * 8
    cache = {}
    for k in range(ng, ng + n3):
        for j in range(ng, ng + n2):
            for i in range(ng, ng + n1):
                r = xv[(i, j, k)[shock_dir - 1]]
                if r not in cache:
                    cache[r] = rj2a_exact(r, tlim, m.gamma)
                (d0, mx, my, mz, e0), (bx, by, bz) = cache[r]
                err[IDN] += abs(d0 - u[IDN, k, j, i])
                err[im[0]] += abs(mx - u[im[0], k, j, i])
                err[im[1]] += abs(my - u[im[1], k, j, i])
                err[im[2]] += abs(mz - u[im[2], k, j, i])
                err[IEN] += abs(e0 - u[IEN, k, j, i])
                err[5 + ib[0]] += abs(bx - bcc[ib[0], k, j, i])
                err[5 + ib[1]] += abs(by - bcc[ib[1], k, j, i])
                err[5 + ib[2]] += abs(bz - bcc[ib[2], k, j, i])
    total = float(m.nx[0] * m.nx[1] * m.nx[2])
    err = [e / total for e in err]
    rms = 0.0
    for e in err:
        rms += e * e
    return math.sqrt(rms), err
