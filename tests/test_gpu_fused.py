"""The fused flux + flux-divergence kernel (k_flux_div, the production stage) against the
separate flux and update kernels, bit for bit over whole registers (ghosts included) and
the clock, on meshes covering every reconstruction, Riemann solver and dynamics, outflow
faces, unequal and tiny MeshBlocks and several band heights; the separate-kernel path is
itself pinned to the oracle by the other GPU parity tests, and the default (fused) path
is compared with the oracle here too."""
import numpy as np
import pytest

from athena_amd.gpu import GpuMesh
from athena_amd.mesh import MeshDesc
from athena_amd.pgen.linear_wave import LinearWave, linear_wave_mesh
from athena_amd.pgen.shock_tube import ShockTube, rj2a_mesh
from athena_amd.pgen.sr import ShearLayer, shear_mesh
from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh
from oracle_bind import OracleMesh

pytestmark = pytest.mark.gpu
KEYS = ("u", "w", "bcc", "b1", "b2", "b3", "e1", "e2", "e3")


def cases():
    yield "lw-hlld-plm-16blocks", linear_wave_mesh(32, bnx=(8, 8, 8)), LinearWave, {}
    yield "turb-hlle-dc", turbulence_mesh(16, 8, rsolver="hlle", xorder=1), Turbulence, {}
    yield "turb-hlld-ppm-ng3", turbulence_mesh(16, 8, xorder=3, ng=3), Turbulence, {}
    yield ("turb-unequal-ng4", MeshDesc(nx=(24, 24, 24), bnx=(12, 8, 4), xmin=(-0.5,) * 3,
                                        xmax=(0.5,) * 3, xorder=3, ng=4), Turbulence, {})
    yield ("turb-2cube", MeshDesc(nx=(8, 8, 8), bnx=(2, 2, 2), xmin=(-0.5,) * 3,
                                  xmax=(0.5,) * 3, ng=2), Turbulence, {})
    yield "sr-shear-ppm-hlle", shear_mesh(16, 8), ShearLayer, {}
    yield "rj2a-outflow-x3", rj2a_mesh(64, shock_dir=3, bnx=(4, 4, 16)), ShockTube, {}
    yield "rj2a-outflow-x1", rj2a_mesh(64, shock_dir=1, bnx=(16, 4, 4)), ShockTube, \
        {"shock_dir": 1}


CASES = list(cases())


def run(m, blocks, rows, ncycles):
    d = GpuMesh(m)
    d.load(blocks)
    d.set_fused(rows)
    d.initialize()
    d.run(ncycles)
    return d


def same_state(m, a, b, what):
    for g in range(m.nblocks):
        for k in KEYS:
            x, y = a.get(g, k), b.get(g, k)
            if not np.array_equal(x, y):
                bad = np.argwhere(x != y)
                raise AssertionError("%s: block %d %s: %d values differ, first at %s"
                                     % (what, g, k, len(bad), tuple(bad[0])))
    assert a.state() == b.state(), what


@pytest.mark.parametrize("name,m,ic,kw", CASES, ids=[c[0] for c in CASES])
def test_fused_equals_separate_kernels(name, m, ic, kw):
    print("case fused", name, m.nx, m.bnx, flush=True)
    p = ic(m, **kw)
    blocks = [p.block(g) for g in range(m.nblocks)]
    ref = run(m, blocks, 0, 4)
    assert ref.fused() == 0
    dflt = GpuMesh(m)
    assert dflt.fused() == 0                  # separate kernels unless asked for
    dflt.close()
    rows_all = sorted({-1, 1, m.bnx[1]} | ({2} if m.bnx[1] > 2 else set()))
    for rows in rows_all:
        if rows > 0 and (rows + 2) * (m.bnx[0] + 2) > 256:
            continue
        d = run(m, blocks, rows, 4)
        assert d.fused() > 0
        print("  rows %d (set %d) bitwise" % (d.fused(), rows), flush=True)
        same_state(m, d, ref, "%s rows %d" % (name, rows))


def test_fused_default_matches_oracle():
    m = turbulence_mesh(32, 16)
    print("case fused default vs oracle", m.nx, m.bnx, flush=True)
    p = Turbulence(m)
    blocks = [p.block(g) for g in range(m.nblocks)]
    d = run(m, blocks, -1, 3)
    assert d.fused() == min(256 // 18 - 2, 16)
    o = OracleMesh(m)
    o.load(blocks)
    o.initialize()
    o.run(3)
    for g in range(m.nblocks):
        for k in ("u", "w", "bcc", "b1", "b2", "b3"):
            assert np.array_equal(d.get(g, k), o.get(g, k)), (g, k)
    assert d.state()[:2] == o.state()[:2]
