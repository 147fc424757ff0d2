"""Decaying-turbulence initial condition: uniform density, pressure and field with a
seeded Gaussian random velocity field.

The reference drives/initialises turbulence in ``src/fft/turbulence.cpp:244-330`` with
FFTW (amplitude ~ k^{-(expo+2)/2} on a shell of wavenumbers); FFTW is not available
here, so the same spectrum is synthesised with numpy's FFT on the global grid and then
split into MeshBlocks.  Conserved variables follow
``EquationOfState::PrimitiveToConserved`` (adiabatic_mhd.cpp:99-137).
"""
from __future__ import annotations

import numpy as np

from ..mesh import MeshDesc, new_block_arrays, IDN, IM1, IM2, IM3, IEN


def turbulence_mesh(n=256, bn=32, **kw):
    """256^3 on [-0.5,0.5]^3 split into (n/bn)^3 MeshBlocks, gamma = 5/3."""
    kw.setdefault("gamma", 5.0 / 3.0)
    kw.setdefault("cfl", 0.3)
    return MeshDesc(nx=(n, n, n), bnx=(bn, bn, bn), xmin=(-0.5, -0.5, -0.5),
                    xmax=(0.5, 0.5, 0.5), **kw)


def random_velocity(n, seed=20261202, kmin=2.0, kmax=8.0, expo=2.0, mach=1.0,
                    cs=np.sqrt(5.0 / 3.0)):
    """(3, n, n, n) velocity, zero mean, rms |v| = mach * cs, |v_k| ~ k^{-(expo+2)/2}."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kk = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + k1[None, None, :] ** 2)
    shell = (kk >= kmin) & (kk <= kmax)
    amp = np.zeros_like(kk)
    amp[shell] = kk[shell] ** (-(expo + 2.0) / 2.0)
    v = np.empty((3, n, n, n))
    for d in range(3):
        c = (rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))) * amp
        v[d] = np.fft.ifftn(c).real
        v[d] -= v[d].mean()
    rms = np.sqrt(np.mean(np.sum(v * v, axis=0)))
    v *= mach * cs / rms
    return v


class Turbulence:
    def __init__(self, m: MeshDesc, rho=1.0, pres=1.0, b0=(0.1, 0.0, 0.0), seed=20261202,
                 mach=1.0):
        self.m = m
        self.rho, self.pres, self.b0 = rho, pres, b0
        n = m.nx[0]
        assert m.nx[0] == m.nx[1] == m.nx[2]
        cs = np.sqrt(m.gamma * pres / rho)
        self.v = random_velocity(n, seed=seed, mach=mach, cs=cs)

    def block(self, gid):
        m = self.m
        a = new_block_arrays(m)
        ng = m.ng
        l1, l2, l3 = m.loc(gid)
        n1, n2, n3 = m.bnx
        v = self.v[:, l3 * n3:(l3 + 1) * n3, l2 * n2:(l2 + 1) * n2, l1 * n1:(l1 + 1) * n1]
        sl = (slice(ng, ng + n3), slice(ng, ng + n2), slice(ng, ng + n1))
        bx, by, bz = self.b0
        a["b1"][ng:ng + n3, ng:ng + n2, ng:ng + n1 + 1] = bx
        a["b2"][ng:ng + n3, ng:ng + n2 + 1, ng:ng + n1] = by
        a["b3"][ng:ng + n3 + 1, ng:ng + n2, ng:ng + n1] = bz
        u = a["u"]
        d = self.rho
        u[(IDN,) + sl] = d
        u[(IM1,) + sl] = v[0] * d
        u[(IM2,) + sl] = v[1] * d
        u[(IM3,) + sl] = v[2] * d
        igm1 = 1.0 / (m.gamma - 1.0)
        u[(IEN,) + sl] = self.pres * igm1 + 0.5 * (d * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
                                                     + (bx * bx + by * by + bz * bz))
        return a
