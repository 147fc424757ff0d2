"""Special-relativistic MHD problem generators.

* ``prim2cons`` restates ``EquationOfState::PrimitiveToConserved`` for SR MHD
  (adiabatic_mhd_sr.cpp:284-332).  Primitives are (rho, u^1, u^2, u^3, p_gas) with
  u^i the spatial components of the 4-velocity.
* ``ShockTube`` restates ``gr_shock_tube.cpp`` (shock_dir = 1, Minkowski/SR branch,
  :55-330) for the Mignone, Ugliano & Bodo (2009) tubes of inputs/mhd_sr/athinput.mub_*.
  The library models periodic meshes only, so the tube is embedded in a periodic
  domain three times as long: the second interface created by the periodic wrap is
  1.5 away from the compared region [-0.5, 0.5], farther than any signal (< c) travels
  before tlim <= 0.55.  Transverse directions have two cells; the data are uniform
  in them, so every transverse flux difference vanishes exactly.
* ``ShearLayer``: the relativistic shear-layer configuration, v_x = +-0.5 tanh
  profiles of width 0.01 (two layers for periodicity), rho = P = 1, B_x = 0.1, seeded
  v_y perturbations.  kh.cpp iprob=3 writes Newtonian-form conserved variables even in
  SR (kh.cpp:221-283), so the conserved state here is built with prim2cons instead.
Both write w and w1 (the SR con2prim seed register) like the reference pgens
(gr_shock_tube.cpp:200-214).
"""
from __future__ import annotations

import math

import numpy as np

from ..mesh import MeshDesc, block_coords, new_block_arrays


def prim2cons(prim, bc, gamma):
    """prim (5, ...) = rho, u1, u2, u3, pgas; bc (3, ...) cell-centred B.  Returns (5, ...)."""
    gamma_adi_red = gamma / (gamma - 1.0)
    rho, u1, u2, u3, pgas = prim
    bb1, bb2, bb3 = bc
    u0 = np.sqrt(1.0 + u1 * u1 + u2 * u2 + u3 * u3)
    b0 = bb1 * u1 + bb2 * u2 + bb3 * u3
    b1 = (bb1 + b0 * u1) / u0
    b2 = (bb2 + b0 * u2) / u0
    b3 = (bb3 + b0 * u3) / u0
    b_sq = -(b0 * b0) + b1 * b1 + b2 * b2 + b3 * b3
    wtot_u02 = (rho + gamma_adi_red * pgas + b_sq) * u0 * u0
    return np.array([rho * u0, wtot_u02 * u1 / u0 - b0 * b1, wtot_u02 * u2 / u0 - b0 * b2,
                     wtot_u02 * u3 / u0 - b0 * b3,
                     wtot_u02 - b0 * b0 - (pgas + 0.5 * b_sq)])


# problem parameters of inputs/mhd_sr/athinput.mub_{1,2,4} (tlim, nx1 per unit length,
# gamma, left state, right state: rho, pgas, vx, vy, vz, Bx, By, Bz); tols are the
# reference tests' relative-L1 tolerances for HLLE (mhd_shocks_hlle.py) and tols_hlld
# for HLLD (mhd_shocks_hlld.py)
MUB = {
    1: dict(tlim=0.4, nx1=400, gamma=2.0,
            left=(1.0, 1.0, 0.0, 0.0, 0.0, 0.5, 1.0, 0.0),
            right=(0.125, 0.1, 0.0, 0.0, 0.0, 0.5, -1.0, 0.0),
            tols=[0.02, 0.02, 0.03, 0.08, 0.0, 0.0, 0.02, 0.0],
            tols_hlld=[0.02, 0.01, 0.02, 0.04, 0.0, 0.0, 0.01, 0.0]),
    2: dict(tlim=0.55, nx1=800, gamma=1.6666666666666667,
            left=(1.08, 0.95, 0.4, 0.3, 0.2, 2.0, 0.3, 0.3),
            right=(1.0, 1.0, -0.45, -0.2, 0.2, 2.0, -0.7, 0.5),
            tols=[0.003, 0.002, 0.007, 0.01, 0.007, 0.0, 0.005, 0.007],
            tols_hlld=[0.003, 0.002, 0.007, 0.01, 0.005, 0.0, 0.004, 0.007]),
    4: dict(tlim=0.5, nx1=800, gamma=1.6666666666666667,
            left=(1.0, 5.0, 0.0, 0.3, 0.4, 1.0, 6.0, 2.0),
            right=(0.9, 5.3, 0.0, 0.0, 0.0, 1.0, 5.0, 2.0),
            tols=[0.003, 0.002, 0.03, 0.04, 0.008, 0.0, 0.002, 0.005],
            tols_hlld=[0.002, 0.001, 0.02, 0.03, 0.004, 0.0, 0.001, 0.003]),
}


def shock_tube_mesh(problem, nx1=None, xorder=2, rsolver="hlle", outflow=False, **kw):
    """The tube on a 3-D mesh with two periodic transverse cells.  outflow=True: the
    reference's own [-0.5, 0.5] between outflow faces; otherwise a periodic box three
    times as long (the wrap-around waves do not reach [-0.5, 0.5] by tlim)."""
    p = MUB[problem]
    if outflow:
        n = nx1 or p["nx1"]
        lo, hi = -0.5, 0.5
        kw.setdefault("bc", ("outflow", "outflow") + ("periodic",) * 4)
    else:
        n = (nx1 or p["nx1"]) * 3
        lo, hi = -1.5, 1.5
    return MeshDesc(nx=(n, 2, 2), bnx=kw.pop("bnx", (n, 2, 2)), xmin=(lo, -0.5, -0.5),
                    xmax=(hi, 0.5, 0.5), gamma=p["gamma"], cfl=0.4, xorder=xorder,
                    rsolver=rsolver, relativistic=True, **kw)


def _state(s):
    rho, pgas, vx, vy, vz, bbx, bby, bbz = s
    ut = math.sqrt(1.0 / (1.0 - (vx * vx + vy * vy + vz * vz)))
    ux, uy, uz = ut * vx, ut * vy, ut * vz
    bt = bbx * ux + bby * uy + bbz * uz
    bx = (bbx + bt * ux) / ut
    by = (bby + bt * uy) / ut
    bz = (bbz + bt * uz) / ut
    # cell-centred lab-frame field used by PrimitiveToConserved (gr_shock_tube.cpp:219-221)
    bb = (bx * ut - bt * ux, by * ut - bt * uy, bz * ut - bt * uz)
    return (rho, ux, uy, uz, pgas), bb, (bbx, bby, bbz)


class ShockTube:
    def __init__(self, m: MeshDesc, problem):
        self.m, self.p = m, MUB[problem]
        self.L = _state(self.p["left"])
        self.R = _state(self.p["right"])

    def _left(self, x):
        # left state on [-1.5, 0), right state on [0, 1.5); xshock = 0
        return x < 0.0

    def block(self, gid):
        m = self.m
        c = block_coords(m, gid)
        a = new_block_arrays(m)
        ng = m.ng
        n1, n2, n3 = m.bnx
        is_, ie = ng, ng + n1 - 1
        left = self._left(c.x1v)                      # per i (shock_dir = 1)
        prim = np.zeros((5, m.nc[0]))
        bb = np.zeros((3, m.nc[0]))
        bf = np.zeros((3, m.nc[0]))
        for i in range(m.nc[0]):
            s = self.L if left[i] else self.R
            prim[:, i] = s[0]
            bb[:, i] = s[1]
            bf[:, i] = s[2]
        sl = (slice(ng, ng + n3), slice(ng, ng + n2), slice(is_, ie + 1))
        for n in range(5):
            a["w"][(n,) + sl] = prim[n, is_:ie + 1]
            a["w1"][(n,) + sl] = prim[n, is_:ie + 1]
        cons = prim2cons(prim[:, is_:ie + 1], bb[:, is_:ie + 1], m.gamma)
        for n in range(5):
            a["u"][(n,) + sl] = cons[n]
        # face fields: face i takes the state of cell i (gr_shock_tube.cpp:235-318)
        a["b1"][ng:ng + n3, ng:ng + n2, is_:ie + 2] = bf[0, is_:ie + 2]
        a["b2"][ng:ng + n3, ng:ng + n2 + 1, is_:ie + 1] = bf[1, is_:ie + 1]
        a["b3"][ng:ng + n3 + 1, ng:ng + n2, is_:ie + 1] = bf[2, is_:ie + 1]
        # The reference tube is 1-D: it has no transverse ghost cells, so every NH seed
        # (w1) comes from a filled cell.  In the 3-D embedding the transverse ghosts of
        # w/w1 are therefore filled with their periodic images; left at zero they would
        # seed the first primitive recovery differently from their images and break the
        # exact transverse uniformity at the 1e-13 level.
        if m.nblocks == 1:
            for key in ("w", "w1"):
                a[key] = _wrap_ghosts(a[key], ng, (n1, n2, n3))
        return a

    def compare(self, get_block, ref):
        """Reference comparison of mhd_shocks_hlle.py:53-85 on [-0.5, 0.5]: relative L1
        of dens, Etot, mom(3), Bcc(3) against the committed solution (piecewise-constant
        union-grid L1 of scripts/utils/comparison.py:14-30)."""
        m = self.m
        assert m.nblocks == 1
        b = get_block(0)
        c = block_coords(m, 0)
        ng = m.ng
        n1 = m.bnx[0]
        xf = c.x1f[ng:ng + n1 + 1]
        lo = int(np.searchsorted(xf, -0.5 - 1e-12))
        hi = int(np.searchsorted(xf, 0.5 + 1e-12))
        xf = xf[lo:hi]
        j, k = ng, ng
        cell = slice(ng + lo, ng + hi - 1)
        new = [b["u"][0, k, j, cell], b["u"][4, k, j, cell], b["u"][1, k, j, cell],
               b["u"][2, k, j, cell], b["u"][3, k, j, cell], b["bcc"][0, k, j, cell],
               b["bcc"][1, k, j, cell], b["bcc"][2, k, j, cell]]
        refs = [ref["dens"], ref["Etot"], ref["mom"][0], ref["mom"][1], ref["mom"][2],
                ref["ccB"][0], ref["ccB"][1], ref["ccB"][2]]
        out = []
        for arr_new, arr_ref in zip(new, refs):
            eps = l1_diff(ref["x1f"].astype(np.float64), arr_ref.astype(np.float64), xf,
                          np.asarray(arr_new, dtype=np.float32).astype(np.float64))
            norm = l1_norm(ref["x1f"].astype(np.float64), arr_ref.astype(np.float64))
            out.append((eps, norm))
        return out


def _wrap_ghosts(arr, ng, n):
    n1, n2, n3 = n
    core = arr[:, ng:ng + n3, ng:ng + n2, ng:ng + n1]
    idx3 = (np.arange(n3 + 2 * ng) - ng) % n3
    idx2 = (np.arange(n2 + 2 * ng) - ng) % n2
    idx1 = (np.arange(n1 + 2 * ng) - ng) % n1
    return core[:, idx3][:, :, idx2][:, :, :, idx1].copy()


def l1_norm(faces, vals):
    return math.fsum(np.abs(vals) * np.diff(faces))


def l1_diff(f1, v1, f2, v2):
    fr = np.union1d(f1, f2)
    mid = 0.5 * (fr[:-1] + fr[1:])

    def fill(f, v):
        idx = np.searchsorted(f, mid, side="right") - 1
        ok = (idx >= 0) & (idx < len(v))
        out = np.zeros(len(mid))
        out[ok] = v[idx[ok]]
        return out
    return math.fsum(np.abs(fill(f1, v1) - fill(f2, v2)) * np.diff(fr))


def shear_mesh(n=256, bn=32, **kw):
    kw.setdefault("gamma", 4.0 / 3.0)
    kw.setdefault("cfl", 0.3)
    kw.setdefault("xorder", 3)
    kw.setdefault("ng", 3)
    return MeshDesc(nx=(n, n, n), bnx=(bn, bn, bn), xmin=(-0.5, -0.5, -0.5),
                    xmax=(0.5, 0.5, 0.5), rsolver="hlle", relativistic=True, **kw)


class ShearLayer:
    def __init__(self, m: MeshDesc, vshear=0.5, width=0.01, amp=1.0e-3, bx=0.1,
                 seed=20261202):
        self.m, self.vshear, self.width, self.amp, self.bx = m, vshear, width, amp, bx
        self.rng = np.random.default_rng(seed)
        self.noise = self.rng.uniform(-1.0, 1.0, (m.nx[2], m.nx[1], m.nx[0]))

    def block(self, gid):
        m = self.m
        c = block_coords(m, gid)
        a = new_block_arrays(m)
        ng = m.ng
        n1, n2, n3 = m.bnx
        l1, l2, l3 = m.loc(gid)
        y = c.x2v[ng:ng + n2][None, :, None]
        vx = self.vshear * (np.tanh((y + 0.25) / self.width) - np.tanh((y - 0.25) / self.width)
                            - 1.0) * np.ones((n3, n2, n1))
        vy = self.amp * self.noise[l3 * n3:(l3 + 1) * n3, l2 * n2:(l2 + 1) * n2,
                                   l1 * n1:(l1 + 1) * n1]
        ut = 1.0 / np.sqrt(1.0 - (vx * vx + vy * vy))
        prim = np.array([np.ones_like(vx), ut * vx, ut * vy, np.zeros_like(vx),
                         np.ones_like(vx)])
        bc = np.array([np.full_like(vx, self.bx), np.zeros_like(vx), np.zeros_like(vx)])
        cons = prim2cons(prim, bc, m.gamma)
        sl = (slice(ng, ng + n3), slice(ng, ng + n2), slice(ng, ng + n1))
        for n in range(5):
            a["u"][(n,) + sl] = cons[n]
            a["w"][(n,) + sl] = prim[n]
            a["w1"][(n,) + sl] = prim[n]
        a["b1"][ng:ng + n3, ng:ng + n2, ng:ng + n1 + 1] = self.bx
        return a
