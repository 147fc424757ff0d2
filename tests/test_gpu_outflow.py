"""GPU parity with outflow faces: the device library against the oracle, bit for bit
over whole registers (ghost layers included), on the reference's RJ2a shock tube
(tst/regression/scripts/tests/mhd/rj2a_shock.py, the 3-D case along x3) and on
multi-block turbulence meshes whose outflow faces meet at edges and corners; and the
RJ2a criteria of the reference's test met by the device run itself."""
import numpy as np
import pytest

from athena_amd.gpu import GpuMesh
from athena_amd.pgen.shock_tube import RJ2A_TLIM, ShockTube, rj2a_errors, rj2a_mesh
from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh
from oracle_bind import OracleMesh

pytestmark = pytest.mark.gpu
KEYS = ("u", "w", "bcc", "b1", "b2", "b3")


def run_both(m, blocks, ncycles=1 << 30, tlim=1e300):
    o = OracleMesh(m)
    o.load(blocks)
    d = GpuMesh(m)
    d.load(blocks)
    o.initialize(tlim)
    d.initialize(tlim)
    no = o.run(ncycles, tlim)
    nd = d.run(ncycles, tlim)
    assert no == nd
    return o, d, nd


def assert_identical(m, o, d):
    for g in range(m.nblocks):
        for k in KEYS:
            a, b = d.get(g, k), o.get(g, k)
            if not np.array_equal(a, b):
                bad = np.argwhere(a != b)
                raise AssertionError("block %d %s: %d values differ, first at %s (%r vs %r)"
                                     % (g, k, len(bad), tuple(bad[0]), a[tuple(bad[0])],
                                        b[tuple(bad[0])]))
    to, tg = o.state(), d.state()
    assert to[2] == tg[2] and to[0] == tg[0] and to[1] == tg[1], (to, tg)


def test_rj2a_x3_bitwise_and_reference_criteria():
    errs = {}
    for nx in (256, 512):
        print("case rj2a x3", nx, flush=True)
        m = rj2a_mesh(nx, shock_dir=3)
        o, d, n = run_both(m, [ShockTube(m, 3).block(0)], tlim=RJ2A_TLIM)
        assert_identical(m, o, d)
        errs[nx] = rj2a_errors(m, d.get(0, "u"), d.get(0, "bcc"), 3)[0]
        print("  cycles %d rms %e" % (n, errs[nx]), flush=True)
    assert errs[512] <= 0.01
    assert errs[512] / errs[256] <= 0.6


BCS = {
    "x1x3": ("outflow", "outflow", "periodic", "periodic", "outflow", "outflow"),
    "x2": ("periodic", "periodic", "outflow", "outflow", "periodic", "periodic"),
    "all": ("outflow",) * 6,
}


@pytest.mark.parametrize("bcname,rsolver,xorder,ng", [
    ("x1x3", "hlld", 2, 2), ("all", "hlld", 2, 2), ("x2", "hlle", 2, 2),
    ("all", "hlld", 3, 3), ("all", "hlle", 1, 2)])
def test_turbulence_outflow_multiblock(bcname, rsolver, xorder, ng):
    print("case turbulence 16^3/8^3 outflow", bcname, rsolver, xorder, ng, flush=True)
    m = turbulence_mesh(16, 8, rsolver=rsolver, xorder=xorder, ng=ng, bc=BCS[bcname])
    t = Turbulence(m)
    o, d, _ = run_both(m, [t.block(g) for g in range(m.nblocks)], ncycles=12)
    assert_identical(m, o, d)


def test_outflow_noncubic_blocks():
    print("case turbulence 16^3 in 8x4x16 blocks, all outflow", flush=True)
    from athena_amd.mesh import MeshDesc
    m = MeshDesc(nx=(16, 16, 16), bnx=(8, 4, 16), xmin=(-0.5,) * 3, xmax=(0.5,) * 3,
                 gamma=5.0 / 3.0, cfl=0.3, bc=BCS["all"])
    t = Turbulence(m)
    o, d, _ = run_both(m, [t.block(g) for g in range(m.nblocks)], ncycles=8)
    assert_identical(m, o, d)


def test_outflow_mesh_refuses_ensemble_and_autotune():
    from athena_amd.gpu import AthGpuError, Ensemble
    m = rj2a_mesh(64, shock_dir=3)
    d = GpuMesh(m)
    d.load([ShockTube(m, 3).block(0)])
    d.initialize(RJ2A_TLIM)
    with pytest.raises(AthGpuError, match="periodic"):
        Ensemble([d])
    with pytest.raises(AthGpuError, match="outflow"):
        d.autotune()
