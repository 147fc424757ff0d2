"""CPU bounds checks of the stage kernels' launch geometry (no device calls): the
position box each kernel covers, combined with the stencil it reads, stays inside the
MeshBlock's arrays for every reconstruction order and ghost width, and covers exactly
the reference's loop ranges.  Stencils: donor cell reads cells i-1, i for face i
(dc.cpp:24-103), PLM i-2..i+1 (plm.cpp:27-167), PPM i-3..i+2 (ppm.cpp:44-334); corner
EMFs read the four cells around an edge (calculate_corner_e.cpp:187-229); cell-centred
fields read faces i and i+1 (field.cpp:112-180)."""
import itertools

import pytest

from athena_amd.gpu import kernel_box

STENCIL = {1: (1, 0), 2: (2, 1), 3: (3, 2)}   # cells below / above a face


@pytest.mark.parametrize("nx,ng,order", [(8, 2, 1), (8, 2, 2), (16, 3, 3), (32, 2, 2),
                                         (32, 3, 3), (12, 4, 3), (16, 2, 1), (2, 2, 2),
                                         (3, 3, 3)])
def test_flux_box_reads_inside_the_block(nx, ng, order):
    if order == 3 and ng < 3:
        pytest.skip("PPM needs NGHOST >= 3")
    lo, n = kernel_box(0, nx, ng)
    nc = nx + 2 * ng
    is_, ie = ng, ng + nx - 1
    below, above = STENCIL[order]
    faces = set()
    for d in range(3):
        for c in range(lo[d], lo[d] + n[d]):
            if c < is_:                   # the cell computes its low face only from is on
                continue
            assert c - below >= 0 and c + above <= nc - 1, (d, c)
            faces.add(c)
    # faces is..ie+1 along every direction, cells is-1..ie+1 transverse
    assert faces == set(range(is_, ie + 2))
    assert all(lo[d] == is_ - 1 and lo[d] + n[d] - 1 == ie + 1 for d in range(3))


@pytest.mark.parametrize("nx,ng", [(8, 2), (32, 2), (16, 3), (2, 2), (3, 3)])
def test_corner_update_and_recovery_boxes(nx, ng):
    nc = nx + 2 * ng
    is_, ie = ng, ng + nx - 1
    for kind in (1, 2):                   # edges / CT faces is..ie+1
        lo, n = kernel_box(kind, nx, ng)
        for d in range(3):
            assert lo[d] == is_ and lo[d] + n[d] - 1 == ie + 1
            # corner EMFs read cells c-1 and c; face arrays have nc+1 entries
            assert lo[d] - 1 >= 0 and lo[d] + n[d] - 1 <= nc
    lo, n = kernel_box(3, nx, ng)         # every cell; reads faces c and c+1 <= nc
    assert list(lo) == [0, 0, 0] and list(n) == [nc] * 3


def test_boxes_fit_every_tile_shape():
    # the tile shapes mask positions beyond the box; only the work-group count changes
    tiles = [(64, 4, 1), (32, 8, 1), (16, 16, 1), (16, 4, 4), (8, 8, 4)]
    for (tx, ty, tz), kind in itertools.product(tiles, range(4)):
        assert tx * ty * tz == 256
        lo, n = kernel_box(kind, 32, 2)
        chunks = -(-n[0] // tx) * -(-n[1] // ty) * -(-n[2] // tz)
        assert chunks * 256 >= n[0] * n[1] * n[2]


@pytest.mark.parametrize("nx,ng", [(8, 2), (6, 3), (4, 4), (5, 2), (2, 2), (3, 3)])
def test_ghost_exchange_covers_every_ghost_once_from_an_active_source(nx, ng):
    # the exchange kernel's numbering, read back through the library: for the cell array
    # and each face component, ghost numbers 0..count-1 hit every ghost position exactly
    # once, in memory order, and read an active position of the neighbour that the
    # reference's same-level exchange sends there: cells is-ng..is-1 <- ie-ng+1..ie and
    # ie+1..ie+ng <- is..is+ng-1 (bvals_cc.cpp:287-427); faces along the normal
    # s-ng..s-1 <- e-ng+1..e and e+2..e+ng+1 <- s+1..s+ng, faces s and e+1 staying with
    # their owner (bvals_fc.cpp:344-688)
    from athena_amd.gpu import ghost_position
    nc = nx + 2 * ng
    is_, ie = ng, ng + nx - 1
    for array in range(4):
        n = [nc + (array == d + 1) for d in range(3)]
        hi = [ie + (array == d + 1) for d in range(3)]      # last active index
        ghosts = {(i, j, k) for k in range(n[2]) for j in range(n[1]) for i in range(n[0])
                  if not all(is_ <= x <= h for x, h in zip((i, j, k), hi))}
        seen, last = set(), -1
        for t in range(len(ghosts)):
            i, j, k, si, sj, sk, p1, p2, p3 = ghost_position(nx, ng, array, t)
            assert (i, j, k) in ghosts and (i, j, k) not in seen
            seen.add((i, j, k))
            lin = (k * n[1] + j) * n[0] + i
            assert lin > last                                 # memory order
            last = lin
            for x, s, p, h in zip((i, j, k), (si, sj, sk), (p1, p2, p3), hi):
                assert is_ <= s <= h                          # an active source
                if p == -1:
                    assert x < is_ and s == x + nx
                elif p == 1:
                    assert x > h and s == x - nx
                else:
                    assert s == x and is_ <= x <= h
        assert seen == ghosts
        with pytest.raises(RuntimeError):
            ghost_position(nx, ng, array, len(ghosts))


def _refine_touch_ok(op, d, comp, lo, hi, nx, ng):
    """Whether every index the refinement kernel `op` touches along direction d, for
    coarse indices lo..hi, lies inside its arrays (refine.hip's stencils): fine index
    f = (c - ng) * 2 + ng; cells have nc = nx + 2 ng entries, faces along their normal
    nc + 1; coarse arrays ncc = nx/2 + 2 ng (+1 for faces along the normal)."""
    nc, ncc = nx + 2 * ng, nx // 2 + 2 * ng
    if lo > hi:
        return False
    cs = range(lo, hi + 1)
    fs = [(c - ng) * 2 + ng for c in cs]
    cells = lambda idx, n: all(0 <= v <= n - 1 for v in idx)
    if op == "restrict_cc" or (op in ("restrict_fc", "internal") and d != comp):
        return cells(cs, ncc) and cells(fs + [f + 1 for f in fs], nc)
    if op == "emf":                    # edges along comp: cells along comp, faces across
        if d == comp:
            return cells(cs, ncc) and cells(fs + [f + 1 for f in fs], nc)
        return cells(cs, ncc + 1) and cells(fs, nc + 1)
    if op == "restrict_fc" or (op == "shared" and d == comp):
        return cells(cs, ncc + 1) and cells(fs, nc + 1)
    if op in ("prolongate_cc", "shared"):
        return (cells([c - 1 for c in cs] + [c + 1 for c in cs], ncc)
                and cells(fs + [f + 1 for f in fs], nc))
    # internal, along the face component's own normal: faces f, f+1, f+2
    return cells(cs, ncc) and cells(fs + [f + 2 for f in fs], nc + 1)


@pytest.mark.parametrize("ng", [2, 3])
def test_refinement_ranges_are_checked_before_any_launch(ng):
    # every (lo, hi) along every direction: the library accepts exactly the ranges whose
    # accesses stay inside the arrays, and rejects the others on the host before any
    # device call (accepted ones reach the launch, which has no device here)
    from athena_amd.gpu import AthGpuError, Refinement
    nx = (8, 6, 4)
    R = Refinement(nx, ng)
    p = 4096
    calls = {"restrict_cc": lambda r: R.restrict_cc(1, 1, p, p, p, r),
             "prolongate_cc": lambda r: R.prolongate_cc(1, 1, p, p, p, p, r)}
    for comp in range(3):
        calls["restrict_fc%d" % comp] = lambda r, c=comp: R.restrict_fc(1, c, p, p, p, r)
        calls["shared%d" % comp] = (lambda r, c=comp:
                                    R.prolongate_fc_shared(1, c, p, p, p, p, r))
    calls["internal"] = lambda r: R.prolongate_fc_internal(1, p, p, p, p, r)
    for comp in range(3):
        calls["emf%d" % comp] = lambda r, c=comp: R.restrict_emf(1, c, p, p, p, r)

    def accepted(fn, r):
        try:
            fn(r)
        except AthGpuError as e:
            return "range outside" not in str(e) and "bad argument" not in str(e)
        return True

    base = [[ng, ng + n // 2 - 1] for n in nx]          # a valid range per direction
    for name, fn in calls.items():
        op = name.rstrip("012")
        comp = int(name[-1]) if name[-1].isdigit() else (3 if op == "internal" else -1)
        for d in range(3):
            ncc = nx[d] // 2 + 2 * ng
            for lo in range(-2, ncc + 3):
                for hi in range(lo - 1, ncc + 3):
                    r = [v for pair in base for v in pair]
                    r[2 * d], r[2 * d + 1] = lo, hi
                    if op == "internal":
                        want = all(_refine_touch_ok("internal", dd, dd if dd == d else -1,
                                                    r[2 * dd], r[2 * dd + 1], nx[dd], ng)
                                   and _refine_touch_ok("internal", dd, -1, r[2 * dd],
                                                        r[2 * dd + 1], nx[dd], ng)
                                   for dd in range(3))
                    else:
                        want = all(_refine_touch_ok(op, dd, comp, r[2 * dd], r[2 * dd + 1],
                                                    nx[dd], ng) for dd in range(3))
                    assert accepted(fn, r) == want, (name, d, lo, hi)
    with pytest.raises(AthGpuError, match="bad argument"):
        Refinement((7, 8, 8), 2).restrict_cc(1, 1, p, p, p, [2, 4, 2, 5, 2, 5])   # odd nx
