"""CPU tests of the MeshBlock tree of a statically refined mesh (ath_gpu_tree_build; the
reference's meshblock_tree.cpp and the <refinementN> set-up of mesh.cpp:306,351-482).
Host code only: the library is loaded, no device call is made.

* Without refinement the block list is the root grid in the Z order that the restart
  layout uses (athena_amd/restart.py), for power-of-two and other root grids.
* Known answers counted by hand on a 4^3 root grid of [0,1]^3: one octant refined once
  gives 63 + 8 = 71 blocks; one corner block refined twice forces its level-1 neighbours
  to exist, so with periodic faces the 8 root blocks around the corner are refined
  (56 + 8 * 8 - 1 + 8 = 127 blocks) and with outflow faces only the corner one
  (63 + 7 + 8 = 78).  A missing periodic wrap or a missing balance step changes them.
* On random root grids, boundary flags and refinement regions, checked by brute force
  from the leaf set alone: the leaves tile the mesh (no leaf inside another, volumes sum
  to the root grid); every one of the 26 neighbour entries equals what the geometry
  gives (the same-level leaf at that position, else the coarser leaf containing it, else
  -2 with the touching finer leaves one level down, -1 across a non-periodic face), so
  adjacent blocks differ by at most one level; each region's interior is at its level.
"""
import itertools

import numpy as np
import pytest

from athena_amd.gpu import AthGpuError, tree_build
from athena_amd.restart import root_level, ztree_order


@pytest.mark.parametrize("nrb", [(4, 4, 4), (3, 5, 2), (1, 1, 1), (8, 2, 1)])
def test_uniform_root_grid_is_the_restart_z_order(nrb):
    locs, nbr, rl = tree_build(nrb)
    assert rl == root_level(nrb)
    assert np.all(locs[:, 0] == rl)
    assert [tuple(int(v) for v in l[1:]) for l in locs] == [tuple(t) for t in ztree_order(nrb)]


def test_known_block_counts():
    octant = [((0.0,) * 3, (0.25,) * 3, 1)]
    corner = [((0.0,) * 3, (0.125,) * 3, 2)]
    assert len(tree_build((4, 4, 4), regions=octant)[0]) == 71
    assert len(tree_build((4, 4, 4), regions=corner)[0]) == 127
    assert len(tree_build((4, 4, 4), bc=(1,) * 6, regions=corner)[0]) == 78


def _check_tree(nrb, bc, xmin, xmax, regions):
    locs, nbr, rl = tree_build(nrb, bc, xmin, xmax, regions)
    leaf = {tuple(int(v) for v in l): g for g, l in enumerate(locs)}
    n = len(locs)
    # tiling: no leaf has a leaf ancestor, volumes add up
    for (lev, a, b, c) in leaf:
        for up in range(1, lev + 1):
            assert (lev - up, a >> up, b >> up, c >> up) not in leaf
    vol = sum(8.0 ** (rl - l[0]) for l in locs)
    assert vol == float(np.prod(nrb))
    periodic = [bc[2 * q] == 0 for q in range(3)]
    for g in range(n):
        lev = int(locs[g, 0])
        nmax = [nrb[q] << (lev - rl) for q in range(3)]
        for o in itertools.product((-1, 0, 1), repeat=3):   # (o3, o2, o1)
            off = (o[2], o[1], o[0])
            p = [int(locs[g, 1 + q]) + off[q] for q in range(3)]
            outside = False
            for q in range(3):
                if p[q] < 0 or p[q] >= nmax[q]:
                    if not periodic[q]:
                        outside = True
                    p[q] %= nmax[q]
            got = int(nbr[g, o[0] + 1, o[1] + 1, o[2] + 1])
            if outside:
                assert got == -1
                continue
            if (lev, *p) in leaf:
                assert got == leaf[(lev, *p)]
            elif lev > 0 and (lev - 1, *(v >> 1 for v in p)) in leaf:
                assert got == leaf[(lev - 1, *(v >> 1 for v in p))]
            else:
                assert got == -2, (g, off)
                # the children of p that touch this block are leaves one level down
                sides = [(0, 1) if off[q] == 0 else ((0,) if off[q] > 0 else (1,))
                         for q in range(3)]
                for c in itertools.product(*sides):
                    ch = (lev + 1, *(2 * p[q] + c[q] for q in range(3)))
                    assert ch in leaf, (g, off, ch)
    # the interior of each region reaches its level
    for (a, b, rlev) in regions:
        centre = [0.5 * (a[q] + b[q]) for q in range(3)]
        frac = [(centre[q] - xmin[q]) / (xmax[q] - xmin[q]) for q in range(3)]
        deepest = 0
        for (lev, *l) in leaf:
            cells = [nrb[q] << (lev - rl) for q in range(3)]
            if all(l[q] <= frac[q] * cells[q] < l[q] + 1 for q in range(3)):
                deepest = lev
        assert deepest >= rl + rlev
    return n


@pytest.mark.parametrize("seed", range(6))
def test_random_trees_by_brute_force(seed):
    rng = np.random.default_rng(seed)
    nrb = tuple(int(v) for v in rng.integers(1, 5, size=3))
    bc = []
    for q in range(3):
        f = int(rng.integers(0, 2))
        bc += [f, f]
    xmin = tuple(float(v) for v in rng.uniform(-1.0, 0.0, size=3))
    xmax = tuple(x + float(w) for x, w in zip(xmin, rng.uniform(0.5, 2.0, size=3)))
    regions = []
    for _ in range(int(rng.integers(1, 4))):
        a = [xmin[q] + (xmax[q] - xmin[q]) * float(rng.uniform(0.0, 0.8)) for q in range(3)]
        b = [a[q] + (xmax[q] - a[q]) * float(rng.uniform(0.05, 0.4)) for q in range(3)]
        regions.append((a, b, int(rng.integers(1, 4))))
    n = _check_tree(nrb, tuple(bc), xmin, xmax, regions)
    print("seed %d: root %s bc %s, %d regions -> %d MeshBlocks" % (seed, nrb, bc, len(regions), n))
    assert n > int(np.prod(nrb))


def test_region_on_block_edges():
    # a region whose faces lie on block faces covers exactly the blocks inside it
    # (mesh.cpp:403-416: strict '>' at the low edge, '>=' at the high edge)
    locs, _, rl = tree_build((4, 4, 4), bc=(1,) * 6,
                             regions=[((0.25, 0.25, 0.25), (0.75, 0.75, 0.75), 1)])
    fine = [tuple(int(v) for v in l[1:]) for l in locs if l[0] == rl + 1]
    assert len(fine) == 64 and len(locs) == 64 - 8 + 64
    assert min(min(t) for t in fine) == 2 and max(max(t) for t in fine) == 5


@pytest.mark.parametrize("kw, msg", [
    (dict(bc=(0, 1, 0, 0, 0, 0)), "bc flags"),
    (dict(regions=[((0.0,) * 3, (1.5, 0.5, 0.5), 1)]), "inside the mesh"),
    (dict(regions=[((0.0,) * 3, (0.5,) * 3, 0)]), "region level"),
    (dict(regions=[((0.6, 0.0, 0.0), (0.5, 0.5, 0.5), 1)]), "invalid refinement region"),
])
def test_rejects_bad_input(kw, msg):
    with pytest.raises(AthGpuError, match=msg):
        tree_build((2, 2, 2), **kw)


def test_deep_region_matches_an_ordered_scan():
    # the region's logical range (mesh.cpp:400-416, an ordered scan in the reference) at
    # level 18: the deepest leaves are the sibling-widened cells containing the region
    from athena_amd.restart import root_level as rlev
    nrb, lev = (2, 2, 2), 18
    a, b = (0.3, 0.41, 0.77), (0.3000001, 0.4100001, 0.7700001)
    locs, nbr, rl = tree_build(nrb, regions=[(a, b, lev)])
    assert rl == rlev(nrb)
    deep = locs[locs[:, 0] == rl + lev]
    n = nrb[0] << lev
    for q in range(3):
        faces = 0.5 + (np.arange(1, n + 1) / n - 0.5)          # right faces, [0, 1] mesh
        lo = int(np.argmax(faces > a[q]))
        hi = int(np.argmax(faces >= b[q]))
        lo -= lo % 2
        hi += 1 - hi % 2            # odd: the pair (hi - 1, hi) is the last one added
        assert deep[:, 1 + q].min() == lo and deep[:, 1 + q].max() == hi, q
    assert len(deep) % 8 == 0


def test_oversized_region_fails_before_building():
    # level 9 over the whole of an 8^3 root grid asks for 2^36 blocks: refused from the
    # region ranges alone, without growing the node vector first
    import time
    t0 = time.time()
    with pytest.raises(AthGpuError, match="tree nodes"):
        tree_build((8, 8, 8), regions=[((0.0,) * 3, (1.0,) * 3, 9)])
    assert time.time() - t0 < 1.0
    # a zero-width region is accepted, as the reference accepts it
    locs, _, _ = tree_build((4, 4, 4), regions=[((0.3, 0.3, 0.3), (0.3, 0.3, 0.3), 1)])
    assert len(locs) >= 64
