"""The static-mesh-refinement restatement (oracle/smr_oracle.cpp) pinned without the GPU,
and the library's host-side SMR planner checked against it.

* Known answer: the reference's SMR linear-wave test (tst/regression/scripts/tests/mhd/
  mhd_linwave.py:42-97, athinput.linear_wave3d with its five <refinementN> regions,
  MeshBlocks of nx/4) for the fast wave: RMS L1 error at nx1 = 64 under 4.5e-8 and
  64/32 ratio under 0.4 (:122-135).  The uniform grid misses that threshold
  (tests/golden/linear_wave_errors.json: 5.973017e-8), so the refined levels carry it.
* Reduction: a refined mesh description without refined regions steps like the uniform
  restatement (athena_oracle.cpp) to round-off in every active cell and face.
* Invariants with planted omissions: over both levels, volume-weighted mass and energy stay
  conserved, and the two sides of every MeshBlock face hold the same magnetic flux (a
  coarse face the mean of the four fine faces on it); leaving out the flux correction
  breaks conservation and leaving out the EMF correction breaks the face consistency, so
  each check sees its correction.
* Planner: the library's neighbour lists (ath_gpu_smr_plan, host only) equal the
  restatement's on several trees, and the planner's range checks pass.
"""
import numpy as np
import pytest

from athena_amd.mesh import refined
from athena_amd.pgen.linear_wave import LinearWave, linear_wave_mesh
from oracle_bind import OracleMesh, OracleSmrMesh
from smr_common import face_mismatch, totals

REGIONS = [((0.3, 0.3, 0.3), (0.4, 0.4, 0.4), 1), ((0.3, 1.0, 1.0), (0.4, 1.1, 1.1), 1),
           ((1.0, 1.0, 0.3), (1.1, 1.1, 0.4), 1), ((1.0, 0.3, 1.0), (1.1, 0.4, 1.1), 1),
           ((1.9, 0.3, 0.3), (2.0, 1.2, 1.2), 1)]


def smr_wave(nx1, wave_flag=0, regions=REGIONS, **kw):
    m = refined(linear_wave_mesh(nx1, wave_flag, bnx=(nx1 // 4,) * 3, **kw), regions)
    return m, LinearWave(m, wave_flag=wave_flag, vflow=1.0 if wave_flag == 3 else 0.0)


def loaded(m, lw):
    o = OracleSmrMesh(m)
    o.load([lw.block(g) for g in range(m.nblocks)])
    return o


def test_tree_of_the_regression_mesh():
    m, _ = smr_wave(32)
    levels = np.bincount(m.locs[:, 0] - m.root_level)
    # 16 root blocks of 8^3; the five regions refine 8 of them into 64 children
    assert list(levels) == [8, 64]


def test_fast_wave_known_answer():
    rms = {}
    for nx1 in (32, 64):
        m, lw = smr_wave(nx1, 0)
        o = loaded(m, lw)
        o.initialize(0.5)
        o.run(10 ** 6, 0.5)
        assert o.state()[0] == pytest.approx(0.5, abs=1e-12)
        rms[nx1] = lw.errors(o.block)[0]
    assert rms[64] < 4.5e-8
    assert rms[64] / rms[32] < 0.4
    assert rms[64] < 5.973017e-08   # better than the uniform grid at the same root size


def test_unrefined_matches_uniform_restatement():
    m0 = linear_wave_mesh(32, 1, bnx=(8, 8, 8))
    m = refined(m0, [])
    assert m.nblocks == m0.nblocks
    lw0, lw = LinearWave(m0, wave_flag=1), LinearWave(m, wave_flag=1)
    u = OracleMesh(m0)
    u.load([lw0.block(g) for g in range(m0.nblocks)])
    s = loaded(m, lw)
    # the tree orders blocks in Z order, the uniform mesh row-major: match by location
    by_loc = {tuple(int(v) for v in m.locs[g][1:]): g for g in range(m.nblocks)}
    u.initialize(1.0)
    s.initialize(1.0)
    u.run(6, 1.0)
    s.run(6, 1.0)
    assert u.state()[:2] == pytest.approx(s.state()[:2], rel=1e-14)
    # active cells and faces; ghost faces differ at round-off already at the start,
    # since the multilevel exchange fills edge and corner ghosts from other neighbours
    ng = m.ng
    act = (slice(ng, -ng),) * 3
    for g0 in range(m0.nblocks):
        g = by_loc[m0.loc(g0)]
        for k in ("u", "b1", "b2", "b3"):
            a, b = u.get(g0, k), s.get(g, k)
            sel = (slice(None),) + act if k == "u" else act
            scale = np.abs(a).max()
            assert np.abs(a[sel] - b[sel]).max() <= 1e-14 * scale, (g0, k)


@pytest.mark.parametrize("skip", [0, 1, 2])
def test_conservation_and_face_consistency_see_the_corrections(skip):
    # a larger amplitude makes the coarse/fine mismatches visible above round-off
    m = refined(linear_wave_mesh(32, 0, bnx=(8, 8, 8)), REGIONS)
    lw = LinearWave(m, wave_flag=0, amp=1e-2)
    o = loaded(m, lw)
    o.skip(skip)
    o.initialize(0.5)
    t0, f0 = totals(o, m), face_mismatch(o, m)
    o.run(20, 0.5)
    t1, f1 = totals(o, m), face_mismatch(o, m)
    mass_ok = abs(t1[0] - t0[0]) <= 1e-14 * abs(t0[0])
    energy_ok = abs(t1[4] - t0[4]) <= 1e-14 * abs(t0[4])
    faces_ok = f1 < 1e-13
    assert f0 < 1e-13
    if skip == 0:
        assert mass_ok and energy_ok and faces_ok
    elif skip == 1:
        assert not (mass_ok and energy_ok)
        assert faces_ok
    else:
        assert not faces_ok


@pytest.mark.parametrize("regions", [
    REGIONS,
    [((0.6, 0.5, 0.5), (0.9, 0.8, 0.8), 2)],               # three levels
    [((0.0, 0.0, 0.0), (0.2, 0.2, 0.2), 1)],               # a corner wrapping periodically
    [((2.9, 0.0, 0.6), (3.0, 1.5, 0.7), 1)],               # a slab across the x1 boundary
])
def test_planner_neighbours_equal_the_restatement(regions):
    from athena_amd import gpu
    m, _ = smr_wave(32, 0, regions=regions)
    o = OracleSmrMesh(m)
    for g in range(m.nblocks):
        mine, steps = gpu.smr_plan(m, g)
        assert np.array_equal(mine, o.neighbors(g)), g
    assert steps[2] > 0
