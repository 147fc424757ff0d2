"""Special-relativistic MHD linear wave (pgen/gr_linear_wave.cpp, the SR branch) and the
regression test built on it (tst/regression/scripts/tests/sr/sr_mhd_linwave.py).

The background state and the eigenvector of each of the 7 waves follow Anton et al.
2010 as gr_linear_wave.cpp:76-440 evaluates them (equation numbers in the comments are
the paper's); the wave runs along x1 over one wavelength, [-0.5, 0.5], and after
t = 1/|lambda| it is compared with the initial conserved state and cell-centred field
(gr_linear_wave.cpp:441-558).  The test runs res x res/2 x res/2 cells in MeshBlocks of
(res/2)^3, res = 16 and 32, HLLD, PLM, VL2, cfl 0.3, gamma 4/3.
"""
from __future__ import annotations

import math

import numpy as np

from ..mesh import IDN, IEN, IM1, IM2, IM3, IPR, IVX, IVY, IVZ, MeshDesc, block_coords, \
    new_block_arrays

ONE_OVER_SQRT2 = 0.70710678118654752440     # defs.hpp.in
TWO_PI = 6.2831853071795862                  # defs.hpp.in
# sr_mhd_linwave.py: background and per-wave high-resolution bounds, ratio bound
BACKGROUND = dict(rho=1.0, pgas=0.5, vx=0.1, vy=0.15, vz=0.05, bx=1.0, by=2.0 / 3.0,
                  bz=1.0 / 3.0)
GAMMA = 4.0 / 3.0
HIGH_RES_ERRORS = (4.0e-8, 3.0e-8, 3.0e-8, 2.0e-8, 4.0e-8, 3.0e-8, 3.0e-8)
ERROR_RATIO = 0.4
NAMES = ("leftgoing fast", "leftgoing Alfven", "leftgoing slow", "entropy",
         "rightgoing slow", "rightgoing Alfven", "rightgoing fast")


def sr_linear_wave_mesh(res, **kw) -> MeshDesc:
    """res x res/2 x res/2 on [-0.5, 0.5]^3 in MeshBlocks of (res/2)^3."""
    h = res // 2
    return MeshDesc(nx=(res, h, h), bnx=(h, h, h), xmin=(-0.5,) * 3, xmax=(0.5,) * 3,
                    gamma=kw.pop("gamma", GAMMA), cfl=kw.pop("cfl", 0.3),
                    xorder=kw.pop("xorder", 2), rsolver=kw.pop("rsolver", "hlld"),
                    relativistic=True, **kw)


def _background(p, gamma):
    vx, vy, vz = p["vx"], p["vy"], p["vz"]
    bx, by, bz = p["bx"], p["by"], p["bz"]
    v_sq = vx * vx + vy * vy + vz * vz
    u = [0.0] * 4
    u[0] = 1.0 / math.sqrt(1.0 - v_sq)
    u[1], u[2], u[3] = u[0] * vx, u[0] * vy, u[0] * vz
    b = [0.0] * 4
    b[0] = bx * u[1] + by * u[2] + bz * u[3]
    b[1] = 1.0 / u[0] * (bx + b[0] * u[1])
    b[2] = 1.0 / u[0] * (by + b[0] * u[2])
    b[3] = 1.0 / u[0] * (bz + b[0] * u[3])
    return u, b


def wavespeed(wave_flag, p=BACKGROUND, gamma=GAMMA):
    """Lab-frame speed of wave_flag as sr_mhd_linwave.py's calculate_wavespeed computes
    it (numpy's polynomial roots for the magnetosonic quartic); the run ends at
    1/|speed|."""
    if wave_flag == 3:
        return p["vx"]
    u, b = _background(p, gamma)
    rho, pgas, vx = p["rho"], p["pgas"], p["vx"]
    red = gamma / (gamma - 1.0)
    b_sq = -b[0] ** 2 + b[1] ** 2 + b[2] ** 2 + b[3] ** 2
    wgas = rho + red * pgas
    wtot = wgas + b_sq
    cs_sq = gamma * pgas / wgas
    lap = (b[1] + wtot ** 0.5 * u[1]) / (b[0] + wtot ** 0.5 * u[0])
    lam = (b[1] - wtot ** 0.5 * u[1]) / (b[0] - wtot ** 0.5 * u[0])
    if wave_flag == 1:
        return min(lap, lam)
    if wave_flag == 5:
        return max(lap, lam)
    roots = sorted(_quartic(p, gamma).real)
    return roots[{0: 0, 2: 1, 4: 2, 6: 3}[wave_flag]]


def _quartic(p, gamma):
    u, b = _background(p, gamma)
    rho, pgas, vx = p["rho"], p["pgas"], p["vx"]
    red = gamma / (gamma - 1.0)
    b_sq = -b[0] ** 2 + b[1] ** 2 + b[2] ** 2 + b[3] ** 2
    wgas = rho + red * pgas
    cs_sq = gamma * pgas / wgas
    fa = wgas * (1.0 / cs_sq - 1.0)
    fb = -(wgas + b_sq / cs_sq)
    g2 = u[0] * u[0]
    g4 = g2 * g2
    a4 = fa * g4 - fb * g2 - b[0] * b[0]
    a3 = -fa * 4.0 * g4 * vx + fb * 2.0 * g2 * vx + 2.0 * b[0] * b[1]
    a2 = fa * 6.0 * g4 * vx * vx + fb * g2 * (1.0 - vx * vx) + b[0] * b[0] - b[1] * b[1]
    a1 = -fa * 4.0 * g4 * vx * vx * vx - fb * 2.0 * g2 * vx - 2.0 * b[0] * b[1]
    a0 = fa * g4 * (vx * vx) * (vx * vx) + fb * g2 * vx * vx + b[1] * b[1]
    return np.roots([a4, a3, a2, a1, a0])


def eigen(wave_flag, p=BACKGROUND, gamma=GAMMA):
    """(lambda, delta_rho, delta_pgas, delta_u[4], delta_b[4]) of gr_linear_wave.cpp:
    116-335 (MHD branch), scaled to unit L2 norm over (delta_rho, delta_pgas,
    delta_u[0..3], delta_b[0..3]) as gr_linear_wave.cpp:400-411 does before the
    amplitude is applied."""
    lam, drho, dp, du, db = _eigen_raw(wave_flag, p, gamma)
    size = drho * drho + dp * dp
    for mu in range(4):
        size += du[mu] * du[mu] + db[mu] * db[mu]
    size = math.sqrt(size)
    return (lam, drho / size, dp / size, [x / size for x in du], [x / size for x in db])


def _eigen_raw(wave_flag, p=BACKGROUND, gamma=GAMMA):
    """The un-normalised eigenvector of gr_linear_wave.cpp:116-335 (MHD branch)."""
    if not 0 <= wave_flag <= 6:
        raise ValueError("wave_flag must be 0..6")
    u, b = _background(p, gamma)
    rho, pgas = p["rho"], p["pgas"]
    vx, vy, vz, bx, by, bz = p["vx"], p["vy"], p["vz"], p["bx"], p["by"], p["bz"]
    red = gamma / (gamma - 1.0)
    b_sq = -b[0] * b[0] + b[1] * b[1] + b[2] * b[2] + b[3] * b[3]
    wgas = rho + red * pgas
    wtot = wgas + b_sq
    cs_sq = gamma * pgas / wgas
    sq = math.sqrt
    lambda_ap = (b[1] + sq(wtot) * u[1]) / (b[0] + sq(wtot) * u[0])            # (A 38)
    lambda_am = (b[1] - sq(wtot) * u[1]) / (b[0] - sq(wtot) * u[0])            # (A 38)
    if wave_flag == 3:                                                           # (A 46)
        return vx, 1.0, 0.0, [0.0] * 4, [0.0] * 4

    def alphas(lam):
        a1 = [u[3], lam * u[3], 0.0, u[0] - lam * u[1]]                          # (A 58)
        a2 = [-u[2], -lam * u[2], lam * u[1] - u[0], 0.0]                        # (A 59)
        g_1 = 1.0 / u[0] * (by + lam * vy / (1.0 - lam * vx) * bx)              # (A 60)
        g_2 = 1.0 / u[0] * (bz + lam * vz / (1.0 - lam * vx) * bx)              # (A 61)
        if g_1 == 0.0 and g_2 == 0.0:
            f_1 = f_2 = ONE_OVER_SQRT2                                           # (A 67)
        else:
            f_1 = g_1 / sq(g_1 * g_1 + g_2 * g_2)                               # (A 66)
            f_2 = g_2 / sq(g_1 * g_1 + g_2 * g_2)
        return a1, a2, g_1, g_2, f_1, f_2

    if wave_flag in (1, 5):                                                      # (A 65)
        sign = 1.0
        if lambda_ap > lambda_am:
            if wave_flag == 1:
                sign = -1.0
        elif wave_flag == 5:
            sign = -1.0
        lam = lambda_ap if sign > 0 else lambda_am
        a1, a2, _, _, f_1, f_2 = alphas(lam)
        du = [f_1 * a1[m] + f_2 * a2[m] for m in range(4)]
        db = [-sign * sq(wtot) * du[m] for m in range(4)]
        return lam, 0.0, 0.0, du, db

    # magnetosonic (A 71)
    fa = wgas * (1.0 / cs_sq - 1.0)
    fb = -(wgas + b_sq / cs_sq)
    roots = sorted(_quartic(p, gamma).real)
    l_fl, l_sl, l_sr, l_fr = roots
    lam, other = {0: (l_fl, l_sl), 2: (l_sl, l_fl), 4: (l_sr, l_fr), 6: (l_fr, l_sr)}[wave_flag]
    lambda_a, sign = lambda_ap, 1.0
    if lambda_ap > lambda_am:
        if wave_flag < 3:
            lambda_a, sign = lambda_am, -1.0
    elif wave_flag > 3:
        lambda_a, sign = lambda_am, -1.0
    a = u[0] * (vx - lam)                                                        # (A 39)
    g = 1.0 - lam * lam                                                          # (A 41)
    b_over_a = -sign * sq(-fb - fa * a * a / g)                                  # (A 68)
    a1, a2, g_1, g_2, f_1, f_2 = alphas(lam)
    a11 = -a1[0] * a1[0]
    a12 = -a1[0] * a2[0]
    a22 = -a2[0] * a2[0]
    for i in range(1, 4):
        a11 += a1[i] * a1[i]
        a12 += a1[i] * a2[i]
        a22 += a2[i] * a2[i]
    det = a11 * a22 - a12 * a12
    c_1 = (g_1 * a12 + g_2 * a22) / det * u[0] * (1.0 - lam * vx)               # (A 63)
    c_2 = -(g_1 * a11 + g_2 * a12) / det * u[0] * (1.0 - lam * vx)
    b_t = [c_1 * a1[m] + c_2 * a2[m] for m in range(4)]                          # (A 62)
    phi = [a * u[m] for m in range(4)]
    phi[0] += lam
    phi[1] += 1.0
    if abs(lam - lambda_a) <= abs(other - lambda_a):
        denom = sq(det * (f_1 * f_1 * a11 + 2.0 * f_1 * f_2 * a12 + f_2 * f_2 * a22))
        btn = [((f_1 * a12 + f_2 * a22) * a1[m] - (f_1 * a11 + f_2 * a12) * a2[m]) / denom
               for m in range(4)]                                                # (A 75)
        bt_norm = -b_t[0] * b_t[0]
        for i in range(1, 4):
            bt_norm += b_t[i] * b_t[i]
        bt_norm = sq(bt_norm)
        denom = a * a - (g + a * a) * cs_sq
        dp = 0.0 if denom == 0.0 else -(g + a * a) * cs_sq / denom * bt_norm    # (A 74)
        drho = rho / (gamma * pgas) * dp
        du = [-a * dp / (wgas * cs_sq * (g + a * a)) * phi[m] - b_over_a / wgas * btn[m]
              for m in range(4)]                                                 # (A 72)
        db = [-b_over_a * dp / wgas * u[m] - (1.0 + a * a / g) * btn[m]
              for m in range(4)]                                                 # (A 73)
    else:
        dp = -1.0                                                                # (A 78)
        drho = rho / (gamma * pgas) * dp
        denom = wgas * a * a - b_sq * g
        btr = [b_t[m] / denom for m in range(4)] if denom != 0.0 else [0.0] * 4  # (A 79)
        du = [a / (wgas * cs_sq * (g + a * a)) * phi[m] - b_over_a * g / wgas * btr[m]
              for m in range(4)]                                                 # (A 76)
        db = [b_over_a / wgas * u[m] - (1.0 + a * a / g) * g * btr[m]
              for m in range(4)]                                                 # (A 77)
    return lam, drho, dp, du, db


class SRLinearWave:
    """MeshBlock::ProblemGenerator of gr_linear_wave.cpp, SR branch (:560-850)."""

    def __init__(self, m: MeshDesc, wave_flag, amp=1.0e-6, p=BACKGROUND):
        if not m.relativistic:
            raise ValueError("the SR linear wave needs a relativistic mesh")
        self.m, self.wave_flag, self.amp, self.p = m, wave_flag, amp, p
        self.u, self.b = _background(p, m.gamma)
        self.lam, self.drho, self.dp, self.du, self.db = eigen(wave_flag, p, m.gamma)
        self.k = TWO_PI / (m.xmax[0] - m.xmin[0])
        self.red = m.gamma / (m.gamma - 1.0)

    def tlim(self):
        return 1.0 / abs(wavespeed(self.wave_flag, self.p, self.m.gamma))

    def _states(self, x):
        """Perturbed u^mu, b^mu, rho, pgas at positions x (arrays)."""
        la = self.amp * np.sin(self.k * x)
        rho = self.p["rho"] + la * self.drho
        pg = self.p["pgas"] + la * self.dp
        um = [self.u[m] + la * self.du[m] for m in range(4)]
        bm = [self.b[m] + la * self.db[m] for m in range(4)]
        return rho, pg, um, bm

    def block(self, gid):
        m = self.m
        c = block_coords(m, gid)
        a = new_block_arrays(m)
        x = c.x1v                                   # every cell, ghosts included
        rho, pg, um, bm = self._states(x)
        # b_mu b^mu with the Minkowski lowering (-, +, +, +)
        b_sq = -bm[0] * bm[0] + bm[1] * bm[1] + bm[2] * bm[2] + bm[3] * bm[3]
        wtot = rho + self.red * pg + b_sq
        ptot = pg + 0.5 * b_sq
        w = np.stack([rho, um[1], um[2], um[3], pg])
        a["w"][:] = w[:, None, None, :]
        a["w1"][:] = w[:, None, None, :]
        uu = np.stack([um[0] * rho,
                       wtot * um[0] * um[1] - bm[0] * bm[1],
                       wtot * um[0] * um[2] - bm[0] * bm[2],
                       wtot * um[0] * um[3] - bm[0] * bm[3],
                       wtot * um[0] * um[0] - bm[0] * bm[0] - ptot])
        a["u"][:] = uu[:, None, None, :]
        # faces (gr_linear_wave.cpp:792-815): B^x constant, B^y and B^z from the cell
        # centre x1v(i) of their column
        by = bm[2] * um[0] - bm[0] * um[2]
        bz = bm[3] * um[0] - bm[0] * um[3]
        a["b1"][:] = self.p["bx"]
        a["b2"][:] = by[None, None, :]
        a["b3"][:] = bz[None, None, :]
        return a

    def initial(self, gid):
        """Active conserved variables and cell-centred field recorded at set-up."""
        blk = self.block(gid)
        ng = self.m.ng
        n1, n2, n3 = self.m.bnx
        sl = (slice(ng, ng + n3), slice(ng, ng + n2), slice(ng, ng + n1))
        u0 = blk["u"][(slice(None),) + sl]
        b1, b2, b3 = blk["b1"], blk["b2"], blk["b3"]
        bcc = np.stack([
            0.5 * b1[ng:ng + n3, ng:ng + n2, ng:ng + n1] + 0.5 * b1[ng:ng + n3, ng:ng + n2, ng + 1:ng + n1 + 1],
            0.5 * b2[ng:ng + n3, ng:ng + n2, ng:ng + n1] + 0.5 * b2[ng:ng + n3, ng + 1:ng + n2 + 1, ng:ng + n1],
            0.5 * b3[ng:ng + n3, ng:ng + n2, ng:ng + n1] + 0.5 * b3[ng + 1:ng + n3 + 1, ng:ng + n2, ng:ng + n1]])
        return u0, bcc

    def errors(self, get):
        """RMS and per-variable L1 errors of gr_linear_wave.cpp:441-558 (volume-weighted,
        D, M1, M2, M3, E, B1c, B2c, B3c; RMS over the 8).  get(gid, name) returns a
        block register."""
        m = self.m
        ng = m.ng
        n1, n2, n3 = m.bnx
        err = np.zeros(8)
        vol_tot = 0.0
        for g in range(m.nblocks):
            c = block_coords(m, g)
            u0, b0 = self.initial(g)
            u = get(g, "u")[:, ng:ng + n3, ng:ng + n2, ng:ng + n1]
            bcc = get(g, "bcc")[:, ng:ng + n3, ng:ng + n2, ng:ng + n1]
            vol = (c.dx1f[ng:ng + n1][None, None, :] * c.dx2f[ng:ng + n2][None, :, None]
                   * c.dx3f[ng:ng + n3][:, None, None])
            for n in range(5):
                err[n] += float(np.sum(np.abs(u[n] - u0[n]) * vol))
            for n in range(3):
                err[5 + n] += float(np.sum(np.abs(bcc[n] - b0[n]) * vol))
            vol_tot += float(np.sum(vol))
        err /= vol_tot
        return math.sqrt(float(np.sum(err * err)) / 8.0), err
