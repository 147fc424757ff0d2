"""Full-size device runs checked through properties that hold at any size: the 256^3
turbulence mesh of the throughput benchmark (512 MeshBlocks of 32^3, VL2 + PLM + HLLD
+ CT) and the 256^3 SR shear layer (PPM + HLLE, NGHOST 3), both periodic.

* Conservation: the device history sums (mass, three momenta, total energy) are
  unchanged to round-off after several cycles (finite-volume update with shared face
  fluxes; outputs/history.cpp sums).
* div B stays at round-off (CT on shared, averaged edge EMFs).
* The same 256^3 turbulence run split into 64 MeshBlocks of 64^3 instead of 512 of 32^3
  gives the same state and time steps bit for bit: a block boundary only makes the EMF
  of a shared face or edge the average of 2 or 4 equal values, and (a + a) * 0.5 and
  ((a + a) + a + a) * 0.25 round back to a exactly (the rounding error of 3a is below
  half an ulp of 4a).
These sizes are beyond what the oracle runs in seconds; the oracle comparisons are in
tests/test_gpu_parity.py and the other GPU tests at smaller sizes.
"""
import numpy as np
import pytest

from athena_amd.gpu import GpuMesh
from athena_amd.pgen.sr import ShearLayer, shear_mesh
from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh
from harness import div_b

pytestmark = pytest.mark.gpu
NCYC = 4


def run(m, ic, ncycles=NCYC):
    d = GpuMesh(m)
    for g in range(m.nblocks):
        b = ic.block(g)
        for k in ("u", "b1", "b2", "b3", "w", "w1"):
            if k in b:
                d.set(g, k, b[k])
    d.initialize()
    h0 = d.history()
    dts = []
    for _ in range(ncycles):
        d.run(1)
        dts.append(d.state()[1])
    return d, h0, d.history(), dts


def assert_conserved(h0, h1):
    mom = slice(1, 4)
    assert abs(h1[0] - h0[0]) <= 1e-13 * abs(h0[0]), (h0[0], h1[0])
    # momentum totals are near zero: measure against sqrt(2 M KE)
    scale = np.sqrt(2.0 * h0[0] * max(float(np.sum(h0[4:7])), 1e-300))
    assert np.max(np.abs(h1[mom] - h0[mom])) <= 1e-13 * scale, (h0[mom], h1[mom])
    assert abs(h1[7] - h0[7]) <= 1e-13 * abs(h0[7]), (h0[7], h1[7])


def assemble(m, d):
    """The active region of every block placed in one (5, N3, N2, N1) array."""
    ng = m.ng
    n1, n2, n3 = m.bnx
    out = np.empty((5, m.nx[2], m.nx[1], m.nx[0]))
    for g in range(m.nblocks):
        l1, l2, l3 = m.loc(g)
        u = d.get(g, "u")
        out[:, l3 * n3:(l3 + 1) * n3, l2 * n2:(l2 + 1) * n2, l1 * n1:(l1 + 1) * n1] = \
            u[:, ng:ng + n3, ng:ng + n2, ng:ng + n1]
    return out


@pytest.mark.timeout(600)
def test_turbulence_256_invariants_and_decomposition():
    print("case turbulence 256^3 as 512 blocks of 32^3", flush=True)
    m32 = turbulence_mesh(256, 32)
    ic = Turbulence(m32)
    d32, h0, h1, dt32 = run(m32, ic)
    assert_conserved(h0, h1)
    db = div_b(m32, lambda g, k: d32.get(g, k))
    print("  div B * dx max %.3e, dt %s" % (db, dt32), flush=True)
    assert db <= 1e-12
    u32 = assemble(m32, d32)
    del d32
    print("case turbulence 256^3 as 64 blocks of 64^3", flush=True)
    m64 = turbulence_mesh(256, 64)
    d64, g0, g1, dt64 = run(m64, Turbulence(m64))
    for q in (0, 7):                   # same initial mass and energy, summed in another order
        assert abs(g0[q] - h0[q]) <= 1e-13 * abs(h0[q])
    assert dt64 == dt32, (dt64, dt32)
    u64 = assemble(m64, d64)
    for n in range(5):
        ndiff = int(np.count_nonzero(u64[n] != u32[n]))
        print("  u%d values that differ: %d" % (n, ndiff), flush=True)
        assert ndiff == 0, n


@pytest.mark.timeout(600)
def test_sr_shear_256_invariants():
    print("case SR shear layer 256^3 as 512 blocks of 32^3", flush=True)
    m = shear_mesh(256, 32)
    d, h0, h1, dts = run(m, ShearLayer(m), ncycles=2)
    assert all(np.isfinite(dts)) and all(t > 0 for t in dts)
    assert_conserved(h0, h1)
    db = div_b(m, lambda g, k: d.get(g, k))
    print("  div B * dx max %.3e" % db, flush=True)
    assert db <= 1e-12
