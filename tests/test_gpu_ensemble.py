"""GPU checks of ensembles: every member advanced inside an ensemble (one launch per
stage kernel for all members, persistent work-groups, HIP-graph replay, any tile shape)
evolves bit for bit as it does on its own, which the single-mesh parity tests tie to the
CPU oracle; a blast-wave member is also compared with the oracle directly."""
import numpy as np
import pytest

from athena_amd.gpu import AthGpuError, Ensemble, GpuMesh
from athena_amd.pgen.blast import Blast, blast_ensemble, blast_mesh
from harness import compare, div_b
from oracle_bind import OracleMesh

pytestmark = pytest.mark.gpu
KEYS = ("u", "w", "bcc", "b1", "b2", "b3")

# small members of uneven sizes and block shapes, with the ensemble's random draws
SMALL = [(16, 8), (24, 8), (32, 16), (16, 16), (24, 12)]


def small_members(seed=5):
    rng = np.random.default_rng(seed)
    out = []
    for n, bn in SMALL:
        m = blast_mesh(n, bn)
        d = rng.normal(size=3)
        b = Blast(m, prat=float(np.exp(rng.uniform(np.log(10), np.log(1000)))),
                  b0=float(rng.uniform(0.1, 1.0)), bdir=d, radius=0.3)
        out.append((m, b))
    return out


def fresh(members):
    meshes = []
    for m, ic in members:
        d = GpuMesh(m)
        d.load([ic.block(g) for g in range(m.nblocks)])
        d.initialize()
        meshes.append(d)
    return meshes


def snapshot(d):
    return {(g, k): d.get(g, k) for g in range(d.m.nblocks) for k in KEYS}, d.state()


def assert_same(a, b):
    sa, ca = a
    sb, cb = b
    assert ca == cb, (ca, cb)
    for key in sa:
        assert np.array_equal(sa[key], sb[key]), key


@pytest.fixture(scope="module")
def solo():
    """Each member run on its own; snapshots after every cycle 1..4."""
    members = small_members()
    meshes = fresh(members)
    snaps = []
    for _ in range(4):
        for d in meshes:
            d.run(1)
        snaps.append([snapshot(d) for d in meshes])
    return members, snaps


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_matches_members_run_alone(solo, graph):
    # every replay of the captured cycle is checked, with host copies (eager work and
    # synchronisation) between the replays
    print("case ensemble graph=%s" % graph, flush=True)
    members, snaps = solo
    meshes = fresh(members)
    ens = Ensemble(meshes)
    for c in range(4):
        ens.run(1, graph=graph)
        for d, r in zip(meshes, snaps[c]):
            assert_same(snapshot(d), r)


@pytest.mark.parametrize("tiles,persist", [([1, 2, 3, 4], [0, 0, 0, 0]),
                                           ([5, 4, 2, 1], [64, 96, 128, 200]),
                                           ([3, 5, 1, 2], [1, 1, 1, 1]),
                                           ([0, 6, 0, 0], [0, 0, 0, 0]),
                                           ([2, 7, 4, 0], [0, 77, 0, 0])])
def test_ensemble_tiles_and_persistent_grids(solo, tiles, persist):
    print("case ensemble tiles=%s persist=%s" % (tiles, persist), flush=True)
    members, snaps = solo
    meshes = fresh(members)
    ens = Ensemble(meshes)
    ens.set_tiles(tiles, persist)
    ens.run(4, graph=True)
    for d, r in zip(meshes, snaps[3]):
        assert_same(snapshot(d), r)


def test_autotune_keeps_state_and_result(solo):
    print("case ensemble autotune", flush=True)
    members, snaps = solo
    meshes = fresh(members)
    ens = Ensemble(meshes)
    before = [snapshot(d) for d in meshes]
    best = ens.autotune(reps=2)
    t, p = ens.tiles()
    print("  chosen tiles %s persist %s best ms %s" % (t, p, best), flush=True)
    for d, b in zip(meshes, before):
        assert_same(snapshot(d), b)
    ens.run(4)
    for d, r in zip(meshes, snaps[3]):
        assert_same(snapshot(d), r)


def test_single_stage_on_member_is_detected(solo):
    members, _ = solo
    meshes = fresh(members)
    ens = Ensemble(meshes)
    meshes[1].stage(1)
    with pytest.raises(AthGpuError):
        ens.run(1)


def test_blast_member_matches_oracle():
    print("case blast vs oracle", flush=True)
    m = blast_mesh(16, 8)
    ic = Blast(m, prat=300.0, b0=0.7, bdir=(0.3, -0.5, 0.8), radius=0.3)
    blocks = [ic.block(g) for g in range(m.nblocks)]
    o = OracleMesh(m)
    o.load(blocks)
    o.initialize(1e300)
    o.run(5, 1e300)
    d = GpuMesh(m)
    d.load(blocks)
    d.initialize()
    ens = Ensemble([d])
    ens.run(5)
    res = compare(m, lambda g, k: d.get(g, k), lambda g, k: o.get(g, k))
    assert all(eq for _, eq in res.values()), res
    assert o.state()[2] == d.state()[2] == 5


def test_full_ensemble_builds_and_steps():
    """The benchmark's 48 blast waves (31.6 M cells): every member advances, conserves
    mass, momentum and energy to round-off (periodic boxes; device history sums) and
    keeps div B at round-off."""
    print("case 48-member ensemble, 2 cycles", flush=True)
    members = blast_ensemble()
    meshes = fresh(members)
    h0 = [d.history() for d in meshes]
    ens = Ensemble(meshes)
    n, cells = ens.size()
    assert n == 48 and cells == sum(m.nx[0] ** 3 for m, _ in members)
    ens.run(2)
    for q, d in enumerate(meshes):
        t, dt, nc = d.state()
        assert nc == 2 and t > 0 and np.isfinite(dt)
        u = d.get(0, "u")
        assert np.all(np.isfinite(u))
        h1 = d.history()
        assert abs(h1[0] - h0[q][0]) <= 1e-13 * h0[q][0], q
        assert abs(h1[7] - h0[q][7]) <= 1e-13 * h0[q][7], q
        scale = np.sqrt(2.0 * h0[q][0] * h1[7])
        assert np.max(np.abs(h1[1:4] - h0[q][1:4])) <= 1e-13 * scale, q
    for q in range(3):
        m = members[q][0]
        assert div_b(m, lambda g, k: meshes[q].get(g, k)) <= 1e-12, q


def test_single_mesh_tiles_and_autotune(solo):
    print("case single-mesh tiles and autotune", flush=True)
    members, snaps = solo
    m, ic = members[2]
    d = fresh([(m, ic)])[0]
    before = snapshot(d)
    d.autotune(reps=2)
    print("  chosen tiles %s" % d.tiles(), flush=True)
    assert_same(snapshot(d), before)
    d.run(4)
    assert_same(snapshot(d), snaps[3][2])
    d2 = fresh([(m, ic)])[0]
    d2.set_tiles([4, 5, 3, 2])
    assert list(d2.tiles()) == [4, 5, 3, 2]
    d2.run(4)
    assert_same(snapshot(d2), snaps[3][2])
    for tiles in ([0, 6, 0, 0], [0, 7, 0, 0]):
        # the corner kernel's z-march, 8 and 16 planes per segment
        d3 = fresh([(m, ic)])[0]
        d3.set_tiles(tiles)
        assert list(d3.tiles()) == tiles
        d3.run(4)
        assert_same(snapshot(d3), snaps[3][2])
    for bad in ([0, 0, 7, 0], [0, 0, 0, 6]):   # the z-march is for fluxes and corners only
        with pytest.raises(AthGpuError):
            fresh([(m, ic)])[0].set_tiles(bad)


@pytest.mark.parametrize("mode", ["learned", "newton"])
def test_sr_members_with_recovery_modes_match_solo(mode):
    """Relativistic members using the learned seed (a seed launch over all members, then
    the learned recovery) or Newton recovery inside an ensemble, eager and replayed as a
    HIP graph, evolve bit for bit as on their own in the same mode."""
    from athena_amd import ml
    from athena_amd.gpu import C2P_MODES
    from athena_amd.pgen.sr import ShearLayer, shear_mesh
    print("case SR ensemble recovery mode %s" % mode, flush=True)
    net = ml.load_mlp() if mode == "learned" else None
    specs = [(shear_mesh(16, 8), {}), (shear_mesh(24, 8), {"amp": 3.0e-3}),
             (shear_mesh(16, 16), {"vshear": 0.3})]

    def make():
        out = []
        for m, kw in specs:
            ic = ShearLayer(m, **kw)
            d = GpuMesh(m)
            d.load([ic.block(g) for g in range(m.nblocks)])
            d.set_c2p(C2P_MODES[mode], net)
            d.initialize()
            out.append(d)
        return out

    solo = make()
    snaps = []
    for _ in range(3):
        for d in solo:
            d.run(1)
        snaps.append([snapshot(d) for d in solo])
    for graph in (False, True):
        meshes = make()
        ens = Ensemble(meshes)
        for c in range(3):
            ens.run(1, graph=graph)
            for d, r in zip(meshes, snaps[c]):
                assert_same(snapshot(d), r)
        ens.close()


def test_members_with_different_recovery_modes_are_refused():
    from athena_amd.gpu import C2P_MODES
    from athena_amd.pgen.sr import ShearLayer, shear_mesh
    meshes = []
    for mode in ("reference", "newton"):
        m = shear_mesh(16, 8)
        ic = ShearLayer(m)
        d = GpuMesh(m)
        d.load([ic.block(g) for g in range(m.nblocks)])
        d.set_c2p(C2P_MODES[mode], None)
        d.initialize()
        meshes.append(d)
    with pytest.raises(AthGpuError, match="primitive-recovery mode"):
        Ensemble(meshes)
