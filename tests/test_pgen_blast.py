"""CPU checks of the blast-wave problem generator and the ensemble draws
(src/pgen/blast.cpp:34-211 semantics, inputs/mhd/athinput.blast defaults)."""
import numpy as np

from athena_amd.mesh import block_coords
from athena_amd.pgen.blast import EDGES, Blast, blast_ensemble, blast_mesh


def test_blast_state_follows_the_reference_generator():
    m = blast_mesh(32, 16)
    b = Blast(m, prat=100.0, pamb=0.1, damb=1.0, radius=0.3, b0=1.0, bdir=(1.0, 2.0, 2.0))
    ng = m.ng
    gm1 = m.gamma - 1.0
    inside = outside = 0
    for g in range(m.nblocks):
        a = b.block(g)
        c = block_coords(m, g)
        x, y, z = np.meshgrid(c.x3v[ng:-ng], c.x2v[ng:-ng], c.x1v[ng:-ng], indexing="ij")
        r = np.sqrt(x * x + y * y + z * z)      # symmetric in the three axes
        e = a["u"][4, ng:-ng, ng:-ng, ng:-ng]
        assert np.all(a["u"][0, ng:-ng, ng:-ng, ng:-ng] == 1.0)
        assert np.all(a["u"][1:4] == 0.0)
        want = np.where(r < 0.3, 10.0, 0.1) / gm1 + 0.5
        np.testing.assert_allclose(e, want, rtol=1e-15)
        inside += np.count_nonzero(r < 0.3)
        outside += np.count_nonzero(r >= 0.3)
        np.testing.assert_allclose(a["b1"][ng:-ng, ng:-ng, ng:-ng], 1.0 / 3.0, rtol=1e-15)
        np.testing.assert_allclose(a["b2"][ng:-ng, ng:-ng, ng:-ng], 2.0 / 3.0, rtol=1e-15)
        np.testing.assert_allclose(a["b3"][ng:-ng, ng:-ng, ng:-ng], 2.0 / 3.0, rtol=1e-15)
    assert inside > 0 and outside > 0
    # the sphere of radius 0.3 on dx = 1/16 holds about 4/3 pi (0.3*16)^3 cells
    assert abs(inside - 4.0 / 3.0 * np.pi * (0.3 * 16) ** 3) < 0.15 * inside


def test_ensemble_draws():
    e = blast_ensemble(48, seed=20261204)
    assert len(e) == 48
    for m, b in e:
        assert m.nx[0] in EDGES and m.nx[0] == m.nx[1] == m.nx[2]
        assert m.bnx[0] in (16, 32) and m.nx[0] % m.bnx[0] == 0
        assert 10.0 <= b.prat <= 1000.0
        assert 0.1 <= np.linalg.norm(b.b) <= 1.0
    assert {m.nx[0] for m, _ in e} == set(EDGES)
    # reproducible
    e2 = blast_ensemble(48, seed=20261204)
    assert [(m.nx, m.bnx, b.prat) for m, b in e] == [(m.nx, m.bnx, b.prat) for m, b in e2]
