"""MHD linear-wave problem: initial conditions and error analysis.

Restates ``src/pgen/linear_wave.cpp`` for adiabatic MHD:
* ``Mesh::InitUserMeshData`` wave geometry and eigensystem (``:78-184``),
* ``Eigensystem`` adiabatic-MHD right eigenvectors (``:626-773``),
* ``MeshBlock::ProblemGenerator`` face fields from the vector potential and cell
  conserved variables (``:436-580``, with the averaged vector potential on edges next to
  finer MeshBlocks of a statically refined mesh),
* ``Mesh::UserWorkAfterLoop`` volume-weighted L1 errors (``:190-430``).
"""
from __future__ import annotations

import math

import numpy as np

from ..mesh import MeshDesc, block_coords, new_block_arrays, IDN, IM1, IM2, IM3, IEN

TINY_NUMBER = 1.0e-20
PI = 3.1415926535897932


def _sign(x):
    return -1.0 if x < 0.0 else 1.0


def eigensystem(d, v1, v2, v3, h, b1, b2, b3, x, y, gamma):
    """Adiabatic MHD eigenvalues and right eigenvectors (linear_wave.cpp:626-773).
    Returns (ev[7], rem[7][7]) with rem[row][wave]."""
    gm1 = gamma - 1.0
    vsq = v1 * v1 + v2 * v2 + v3 * v3
    btsq = b2 * b2 + b3 * b3
    bt_starsq = (gm1 - (gm1 - 1.0) * y) * btsq
    vaxsq = b1 * b1 / d
    hp = h - (vaxsq + btsq / d)
    twid_asq = max((gm1 * (hp - 0.5 * vsq) - (gm1 - 1.0) * x), TINY_NUMBER)
    ct2 = bt_starsq / d
    tsum = vaxsq + ct2 + twid_asq
    tdif = vaxsq + ct2 - twid_asq
    cf2_cs2 = math.sqrt(tdif * tdif + 4.0 * twid_asq * ct2)
    cfsq = 0.5 * (tsum + cf2_cs2)
    cf = math.sqrt(cfsq)
    cssq = twid_asq * vaxsq / cfsq
    cs = math.sqrt(cssq)
    bt = math.sqrt(btsq)
    bt_star = math.sqrt(bt_starsq)
    if bt == 0.0:
        bet2, bet3 = 1.0, 0.0
    else:
        bet2, bet3 = b2 / bt, b3 / bt
    bet2_star = bet2 / math.sqrt(gm1 - (gm1 - 1.0) * y)
    bet3_star = bet3 / math.sqrt(gm1 - (gm1 - 1.0) * y)
    bet_starsq = bet2_star * bet2_star + bet3_star * bet3_star
    vbet = v2 * bet2_star + v3 * bet3_star
    if (cfsq - cssq) == 0.0:
        alpha_f, alpha_s = 1.0, 0.0
    elif (twid_asq - cssq) <= 0.0:
        alpha_f, alpha_s = 0.0, 1.0
    elif (cfsq - twid_asq) <= 0.0:
        alpha_f, alpha_s = 1.0, 0.0
    else:
        alpha_f = math.sqrt((twid_asq - cssq) / (cfsq - cssq))
        alpha_s = math.sqrt((cfsq - twid_asq) / (cfsq - cssq))
    sqrtd = math.sqrt(d)
    isqrtd = 1.0 / sqrtd
    s = _sign(b1)
    twid_a = math.sqrt(twid_asq)
    qf = cf * alpha_f * s
    qs = cs * alpha_s * s
    af_prime = twid_a * alpha_f * isqrtd
    as_prime = twid_a * alpha_s * isqrtd
    afpbb = af_prime * bt_star * bet_starsq
    aspbb = as_prime * bt_star * bet_starsq
    vax = math.sqrt(vaxsq)
    ev = [v1 - cf, v1 - vax, v1 - cs, v1, v1 + cs, v1 + vax, v1 + cf]
    r = [[0.0] * 7 for _ in range(7)]
    r[0] = [alpha_f, 0.0, alpha_s, 1.0, alpha_s, 0.0, alpha_f]
    r[1] = [alpha_f * ev[0], 0.0, alpha_s * ev[2], v1, alpha_s * ev[4], 0.0, alpha_f * ev[6]]
    qa, qb, qc, qd = alpha_f * v2, alpha_s * v2, qs * bet2_star, qf * bet2_star
    r[2] = [qa + qc, -bet3, qb - qd, v2, qb + qd, bet3, qa - qc]
    qa, qb, qc, qd = alpha_f * v3, alpha_s * v3, qs * bet3_star, qf * bet3_star
    r[3] = [qa + qc, bet2, qb - qd, v3, qb + qd, -bet2, qa - qc]
    r41 = -(v2 * bet3 - v3 * bet2)
    r[4] = [alpha_f * (hp - v1 * cf) + qs * vbet + aspbb, r41,
            alpha_s * (hp - v1 * cs) - qf * vbet - afpbb,
            0.5 * vsq + (gm1 - 1.0) * x / gm1,
            alpha_s * (hp + v1 * cs) + qf * vbet - afpbb, -r41,
            alpha_f * (hp + v1 * cf) - qs * vbet + aspbb]
    r50, r51, r52 = as_prime * bet2_star, -bet3 * s * isqrtd, -af_prime * bet2_star
    r[5] = [r50, r51, r52, 0.0, r52, r51, r50]
    r60, r61, r62 = as_prime * bet3_star, bet2 * s * isqrtd, -af_prime * bet3_star
    r[6] = [r60, r61, r62, 0.0, r62, r61, r60]
    return ev, r


class LinearWave:
    """Global wave parameters (InitUserMeshData) for a cubic-or-not periodic mesh."""

    def __init__(self, m: MeshDesc, wave_flag=0, amp=1.0e-6, vflow=0.0):
        self.m = m
        self.wave_flag = wave_flag
        self.amp = amp
        self.vflow = vflow
        gam = m.gamma
        self.gm1 = gam - 1.0
        x1size = m.xmax[0] - m.xmin[0]
        x2size = m.xmax[1] - m.xmin[1]
        x3size = m.xmax[2] - m.xmin[2]
        ang_3 = math.atan(x1size / x2size)
        self.sin_a3, self.cos_a3 = math.sin(ang_3), math.cos(ang_3)
        ang_2 = math.atan(0.5 * (x1size * self.cos_a3 + x2size * self.sin_a3) / x3size)
        self.sin_a2, self.cos_a2 = math.sin(ang_2), math.cos(ang_2)
        x1 = x1size * self.cos_a2 * self.cos_a3
        x2 = x2size * self.cos_a2 * self.sin_a3
        x3 = x3size * self.sin_a2
        lam = x1
        if ang_3 != 0.0:
            lam = min(lam, x2)
        if ang_2 != 0.0:
            lam = min(lam, x3)
        self.lam = lam
        self.k_par = 2.0 * PI / lam
        self.d0 = 1.0
        self.u0 = vflow
        self.bx0, self.by0, self.bz0 = 1.0, math.sqrt(2.0), 0.5
        self.p0 = 1.0 / gam
        h0 = ((self.p0 / self.gm1 + 0.5 * self.d0 * (self.u0 * self.u0 + 0.0 + 0.0))
              + self.p0) / self.d0
        h0 += (self.bx0 * self.bx0 + self.by0 * self.by0 + self.bz0 * self.bz0) / self.d0
        self.ev, self.rem = eigensystem(self.d0, self.u0, 0.0, 0.0, h0, self.bx0,
                                        self.by0, self.bz0, 0.0, 1.0, gam)
        self.dby = amp * self.rem[5][wave_flag]
        self.dbz = amp * self.rem[6][wave_flag]

    # vector potential components (linear_wave.cpp:584-620)
    def _xy(self, x1, x2, x3):
        x = x1 * self.cos_a2 * self.cos_a3 + x2 * self.cos_a2 * self.sin_a3 + x3 * self.sin_a2
        y = -x1 * self.sin_a3 + x2 * self.cos_a3
        return x, y

    def A1(self, x1, x2, x3):
        x, y = self._xy(x1, x2, x3)
        Ay = self.bz0 * x - (self.dbz / self.k_par) * np.cos(self.k_par * (x))
        Az = -self.by0 * x + (self.dby / self.k_par) * np.cos(self.k_par * (x)) + self.bx0 * y
        return -Ay * self.sin_a3 - Az * self.sin_a2 * self.cos_a3

    def A2(self, x1, x2, x3):
        x, y = self._xy(x1, x2, x3)
        Ay = self.bz0 * x - (self.dbz / self.k_par) * np.cos(self.k_par * (x))
        Az = -self.by0 * x + (self.dby / self.k_par) * np.cos(self.k_par * (x)) + self.bx0 * y
        return Ay * self.cos_a3 - Az * self.sin_a2 * self.sin_a3

    def A3(self, x1, x2, x3):
        x, y = self._xy(x1, x2, x3)
        Az = -self.by0 * x + (self.dby / self.k_par) * np.cos(self.k_par * (x)) + self.bx0 * y
        return Az * self.cos_a2

    def analytic_cons(self, x1v, x2v, x3v):
        """Analytic conserved state + cell-centred B at the given centres
        (UserWorkAfterLoop :218-255).  Inputs broadcast as (k, j, i)."""
        wf, amp, rem = self.wave_flag, self.amp, self.rem
        x = self.cos_a2 * (x1v * self.cos_a3 + x2v * self.sin_a3) + x3v * self.sin_a2
        sn = np.sin(self.k_par * x)
        ca2, sa2, ca3, sa3 = self.cos_a2, self.sin_a2, self.cos_a3, self.sin_a3
        d1 = self.d0 + amp * sn * rem[0][wf]
        mx = self.d0 * self.vflow + amp * sn * rem[1][wf]
        my = amp * sn * rem[2][wf]
        mz = amp * sn * rem[3][wf]
        m1 = mx * ca2 * ca3 - my * sa3 - mz * sa2 * ca3
        m2 = mx * ca2 * sa3 + my * ca3 - mz * sa2 * sa3
        m3 = mx * sa2 + mz * ca2
        e0 = self.p0 / self.gm1 + 0.5 * self.d0 * self.u0 * self.u0 + amp * sn * rem[4][wf]
        e0 = e0 + 0.5 * (self.bx0 * self.bx0 + self.by0 * self.by0 + self.bz0 * self.bz0)
        bx = self.bx0
        by = self.by0 + amp * sn * rem[5][wf]
        bz = self.bz0 + amp * sn * rem[6][wf]
        b1 = bx * ca2 * ca3 - by * sa3 - bz * sa2 * ca3
        b2 = bx * ca2 * sa3 + by * ca3 - bz * sa2 * sa3
        b3 = bx * sa2 + bz * ca2
        return [d1, m1, m2, m3, e0], [b1, b2, b3]

    def block(self, gid):
        """Initial registers of one block (ProblemGenerator :436-580)."""
        m = self.m
        c = block_coords(m, gid)
        a = new_block_arrays(m)
        ng = m.ng
        n1, n2, n3 = m.bnx
        is_, js, ks = ng, ng, ng
        ie, je, ke = is_ + n1 - 1, js + n2 - 1, ks + n3 - 1
        # vector potential on [ks..ke+1][js..je+1][is..ie+1]
        K = np.arange(ks, ke + 2)[:, None, None]
        J = np.arange(js, je + 2)[None, :, None]
        I = np.arange(is_, ie + 2)[None, None, :]
        a1 = self.A1(c.x1v[I], c.x2f[J], c.x3f[K])
        a2 = self.A2(c.x1f[I], c.x2v[J], c.x3f[K])
        a3 = self.A3(c.x1f[I], c.x2f[J], c.x3v[K])
        if m.nblevel is not None:
            self._refined_edges(m, gid, c, a1, a2, a3, (is_, ie, js, je, ks, ke))
        # shapes (n3+1, n2+1, n1+1); local index (k-ks, j-js, i-is)
        dx1 = c.dx1f[is_:ie + 2][None, None, :]
        dx2 = c.dx2f[js:je + 2][None, :, None]
        dx3 = c.dx3f[ks:ke + 2][:, None, None]
        b1 = (a3[:-1, 1:, :] - a3[:-1, :-1, :]) / dx2[:, :-1, :] - \
             (a2[1:, :-1, :] - a2[:-1, :-1, :]) / dx3[:-1]
        a["b1"][ks:ke + 1, js:je + 1, is_:ie + 2] = b1
        b2 = (a1[1:, :, :-1] - a1[:-1, :, :-1]) / dx3[:-1] - \
             (a3[:-1, :, 1:] - a3[:-1, :, :-1]) / dx1[:, :, :-1]
        a["b2"][ks:ke + 1, js:je + 2, is_:ie + 1] = b2
        b3 = (a2[:, :-1, 1:] - a2[:, :-1, :-1]) / dx1[:, :, :-1] - \
             (a1[:, 1:, :-1] - a1[:, :-1, :-1]) / dx2[:, :-1, :]
        a["b3"][ks:ke + 2, js:je + 1, is_:ie + 1] = b3
        # conserved variables
        Kc = np.arange(ks, ke + 1)[:, None, None]
        Jc = np.arange(js, je + 1)[None, :, None]
        Ic = np.arange(is_, ie + 1)[None, None, :]
        x = self.cos_a2 * (c.x1v[Ic] * self.cos_a3 + c.x2v[Jc] * self.sin_a3) + \
            c.x3v[Kc] * self.sin_a2
        sn = np.sin(self.k_par * x)
        wf, amp, rem = self.wave_flag, self.amp, self.rem
        u = a["u"]
        sl = (slice(ks, ke + 1), slice(js, je + 1), slice(is_, ie + 1))
        u[(IDN,) + sl] = self.d0 + amp * sn * rem[0][wf]
        mx = self.d0 * self.vflow + amp * sn * rem[1][wf]
        my = amp * sn * rem[2][wf]
        mz = amp * sn * rem[3][wf]
        ca2, sa2, ca3, sa3 = self.cos_a2, self.sin_a2, self.cos_a3, self.sin_a3
        u[(IM1,) + sl] = mx * ca2 * ca3 - my * sa3 - mz * sa2 * ca3
        u[(IM2,) + sl] = mx * ca2 * sa3 + my * ca3 - mz * sa2 * sa3
        u[(IM3,) + sl] = mx * sa2 + mz * ca2
        e = self.p0 / self.gm1 + 0.5 * self.d0 * self.u0 * self.u0 + amp * sn * rem[4][wf]
        e = e + 0.5 * (self.bx0 * self.bx0 + self.by0 * self.by0 + self.bz0 * self.bz0)
        u[(IEN,) + sl] = e
        return a

    def _refined_edges(self, m, gid, c, a1, a2, a3, rng):
        """On a statically refined mesh, a block edge that touches a finer neighbour
        takes the mean of the vector potential at the two fine edge centres, so that the
        coarse faces there equal the restriction of the fine ones (linear_wave.cpp:
        454-510).  a1..a3 are indexed (k-ks, j-js, i-is) over [ks..ke+1] x [js..je+1] x
        [is..ie+1]."""
        is_, ie, js, je, ks, ke = rng
        nbl = m.nblevel[gid]
        lev = int(m.locs[gid][0])
        fin = lambda o3, o2, o1: bool(nbl[o3 + 1, o2 + 1, o1 + 1] > lev)  # noqa: E731
        K = np.arange(ks, ke + 2)[:, None, None]
        J = np.arange(js, je + 2)[None, :, None]
        I = np.arange(is_, ie + 2)[None, None, :]
        lo_i, hi_i, lo_j, hi_j = I == is_, I == ie + 1, J == js, J == je + 1
        lo_k, hi_k = K == ks, K == ke + 1
        shp = a1.shape
        m1 = np.zeros(shp, dtype=bool)
        for flag, cond in ((fin(0, -1, 0), lo_j), (fin(0, 1, 0), hi_j), (fin(-1, 0, 0), lo_k),
                           (fin(1, 0, 0), hi_k), (fin(-1, -1, 0), lo_j & lo_k),
                           (fin(-1, 1, 0), hi_j & lo_k), (fin(1, -1, 0), lo_j & hi_k),
                           (fin(1, 1, 0), hi_j & hi_k)):
            if flag:
                m1 |= np.broadcast_to(cond, shp)
        m2 = np.zeros(shp, dtype=bool)
        for flag, cond in ((fin(0, 0, -1), lo_i), (fin(0, 0, 1), hi_i), (fin(-1, 0, 0), lo_k),
                           (fin(1, 0, 0), hi_k), (fin(-1, 0, -1), lo_i & lo_k),
                           (fin(-1, 0, 1), hi_i & lo_k), (fin(1, 0, -1), lo_i & hi_k),
                           (fin(1, 0, 1), hi_i & hi_k)):
            if flag:
                m2 |= np.broadcast_to(cond, shp)
        m3 = np.zeros(shp, dtype=bool)
        for flag, cond in ((fin(0, 0, -1), lo_i), (fin(0, 0, 1), hi_i), (fin(0, -1, 0), lo_j),
                           (fin(0, 1, 0), hi_j), (fin(0, -1, -1), lo_i & lo_j),
                           (fin(0, -1, 1), hi_i & lo_j), (fin(0, 1, -1), lo_i & hi_j),
                           (fin(0, 1, 1), hi_i & hi_j)):
            if flag:
                m3 |= np.broadcast_to(cond, shp)
        kk, jj, ii = np.nonzero(m1)
        if len(kk):
            i, j, k = ii + is_, jj + js, kk + ks
            x1l = c.x1f[i] + 0.25 * c.dx1f[i]
            x1r = c.x1f[i] + 0.75 * c.dx1f[i]
            a1[kk, jj, ii] = 0.5 * (self.A1(x1l, c.x2f[j], c.x3f[k]) +
                                    self.A1(x1r, c.x2f[j], c.x3f[k]))
        kk, jj, ii = np.nonzero(m2)
        if len(kk):
            i, j, k = ii + is_, jj + js, kk + ks
            x2l = c.x2f[j] + 0.25 * c.dx2f[j]
            x2r = c.x2f[j] + 0.75 * c.dx2f[j]
            a2[kk, jj, ii] = 0.5 * (self.A2(c.x1f[i], x2l, c.x3f[k]) +
                                    self.A2(c.x1f[i], x2r, c.x3f[k]))
        kk, jj, ii = np.nonzero(m3)
        if len(kk):
            i, j, k = ii + is_, jj + js, kk + ks
            x3l = c.x3f[k] + 0.25 * c.dx3f[k]
            x3r = c.x3f[k] + 0.75 * c.dx3f[k]
            a3[kk, jj, ii] = 0.5 * (self.A3(c.x1f[i], c.x2f[j], x3l) +
                                    self.A3(c.x1f[i], c.x2f[j], x3r))

    def errors(self, get_block):
        """Volume-weighted L1 errors (UserWorkAfterLoop :190-370).  ``get_block(gid)``
        returns a dict with ``u`` and ``bcc`` arrays.  Returns (rms, l1[8], maxerr[8])."""
        m = self.m
        ng = m.ng
        l1 = np.zeros(8)
        mx = np.zeros(8)
        for gid in range(m.nblocks):
            c = block_coords(m, gid)
            b = get_block(gid)
            n1, n2, n3 = m.bnx
            sl = (slice(ng, ng + n3), slice(ng, ng + n2), slice(ng, ng + n1))
            K = c.x3v[ng:ng + n3][:, None, None]
            J = c.x2v[ng:ng + n2][None, :, None]
            I = c.x1v[ng:ng + n1][None, None, :]
            cons, bc = self.analytic_cons(I, J, K)
            vol = (c.dx1f[ng:ng + n1][None, None, :] * c.dx2f[ng:ng + n2][None, :, None]) * \
                c.dx3f[ng:ng + n3][:, None, None]
            for n in range(5):
                err = np.abs(cons[n] - b["u"][(n,) + sl])
                l1[n] += np.sum(err * vol)
                mx[n] = max(mx[n], err.max())
            for n in range(3):
                err = np.abs(bc[n] - b["bcc"][(n,) + sl])
                l1[5 + n] += np.sum(err * vol)
                mx[5 + n] = max(mx[5 + n], err.max())
        volt = (m.xmax[0] - m.xmin[0]) * (m.xmax[1] - m.xmin[1]) * (m.xmax[2] - m.xmin[2])
        l1 = l1 / volt
        rms = math.sqrt(float(np.sum(l1 * l1)))
        return rms, l1, mx


def linear_wave_mesh(nx1=32, wave_flag=0, **kw):
    """The athinput.linear_wave3d configuration on a uniform grid: domain 3 x 1.5 x
    1.5, nx2 = nx3 = nx1/2, gamma=1.666666666666667, CFL 0.3, one MeshBlock unless
    ``bnx`` is given."""
    nx = (nx1, nx1 // 2, nx1 // 2)
    kw.setdefault("bnx", nx)
    kw.setdefault("gamma", 1.666666666666667)
    kw.setdefault("cfl", 0.3)
    return MeshDesc(nx=nx, xmin=(0.0, 0.0, 0.0), xmax=(3.0, 1.5, 1.5), **kw)


# tlim of the reference's regression runs (mhd_linwave.py:43-80)
WAVE_TLIM = {0: 0.5, 1: 1.0, 2: 2.0, 3: 1.0}
