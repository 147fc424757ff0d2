"""Restart files: the reference's binary layout (outputs/restart.cpp:36-240), block
order of the reference's tree (meshblock_tree.cpp:87-110, 336-352), the file read back
by a restatement of the reference's own reader (oracle/restart_oracle.cpp: parameter
parsing, packed header, ID list, tree reconstruction and block offsets as mesh.cpp and
meshblock.cpp read them), which also catches planted writer mistakes (block order,
header padding), and, on the GPU, checkpoint/resume that continues bit for bit, on the
device and on the oracle from the file read the reference's way."""
import os

import numpy as np
import pytest

import oracle_bind
from athena_amd import restart
from athena_amd.mesh import MeshDesc


def tree_order_by_traversal(nbx):
    """Independent construction: build the root grid recursively as CreateRootGrid does
    (a child exists where its lowest block index is inside the grid) and list the leaves
    depth first, children n = i + 2j + 4k in increasing n."""
    lev = restart.root_level(nbx)
    out = []

    def visit(l1, l2, l3, level):
        if level == lev:
            out.append((l1, l2, l3))
            return
        levfac = 1 << (lev - level - 1)
        for n in range(8):
            i, j, k = n & 1, (n >> 1) & 1, (n >> 2) & 1
            if ((l3 * 2 + k) * levfac < nbx[2] and (l2 * 2 + j) * levfac < nbx[1]
                    and (l1 * 2 + i) * levfac < nbx[0]):
                visit(l1 * 2 + i, l2 * 2 + j, l3 * 2 + k, level + 1)

    visit(0, 0, 0, 0)
    return out


@pytest.mark.parametrize("nbx", [(1, 1, 1), (2, 2, 2), (8, 8, 8), (3, 5, 2), (4, 1, 3),
                                 (6, 6, 6)])
def test_block_order_is_the_tree_order(nbx):
    assert restart.ztree_order(nbx) == tree_order_by_traversal(nbx)


def test_param_dump_format():
    txt = restart.param_dump({"mesh": {"nx1": "64", "x1min": "-1.0"},
                              "time": {"cfl_number": "0.3"}})
    assert txt == ("#------------------------- PAR_DUMP -------------------------\n"
                   "<mesh>\n"
                   "nx1   = 64   \n"
                   "x1min = -1.0 \n"
                   "<time>\n"
                   "cfl_number = 0.3 \n"
                   "#------------------------- PAR_DUMP -------------------------\n"
                   "<par_end>\n")


def test_round_trip_and_layout(tmp_path):
    m = MeshDesc(nx=(24, 16, 8), bnx=(8, 8, 8), xmin=(-1.0, 0.0, 0.0), xmax=(1.0, 1.0, 0.5))
    rng = np.random.default_rng(0)
    data = {(g, k): rng.normal(size=m.shape(k)) for g in range(m.nblocks)
            for k in ("u", "b1", "b2", "b3")}
    path = str(tmp_path / "a.rst")
    restart.write(path, m, lambda g, k: data[(g, k)], 0.25, 1e-3, 17)
    head, blocks, params = restart.read(path, m)
    assert head["nbtotal"] == 6 and head["root_level"] == 2
    assert head["time"] == 0.25 and head["dt"] == 1e-3 and head["ncycle"] == 17
    assert head["region"][:6] == (-1.0, 0.0, 0.0, 1.0, 1.0, 0.5)
    assert head["region"][9:] == (24, 16, 8)
    for g in range(m.nblocks):
        for k in ("u", "b1", "b2", "b3"):
            assert np.array_equal(blocks[g][k], data[(g, k)])
    datasize = 8 * sum(int(np.prod(m.shape(k))) for k in ("u", "b1", "b2", "b3"))
    assert head["datasize"] == datasize
    assert os.path.getsize(path) == (len(params) + 8 + 88 + 16 + 4 + 8
                                     + 40 * m.nblocks + m.nblocks * datasize)


def _sample(m, seed=0):
    rng = np.random.default_rng(seed)
    return {(g, k): rng.normal(size=m.shape(k)) for g in range(m.nblocks)
            for k in ("u", "b1", "b2", "b3")}


def _per_block(m):
    return sum(int(np.prod(m.shape(k))) for k in ("u", "b1", "b2", "b3"))


def _split(m, flat):
    out, p = {}, 0
    for k in ("u", "b1", "b2", "b3"):
        n = int(np.prod(m.shape(k)))
        out[k] = flat[p:p + n].reshape(m.shape(k))
        p += n
    return out


def read_as_reference(path, m):
    """The file read by the restatement of the reference's reader: the header, the block
    size it takes from <meshblock>, and each block's registers under the location the
    reference would assign them to."""
    per = _per_block(m)
    ints, reals, locs, data = oracle_bind.restart_read(path, m.ng, m.nblocks,
                                                       per * m.nblocks)
    blocks = {}
    for i, (level, l1, l2, l3) in enumerate(locs[:ints[0]]):
        blocks[m.gid(int(l1), int(l2), int(l3))] = _split(m, data[per * i:per * (i + 1)])
    return ints, reals, blocks


@pytest.mark.parametrize("nx,bnx,ng", [((24, 16, 8), (8, 8, 8), 2), ((32, 32, 32), (8, 16, 16), 3),
                                       ((16, 16, 16), (16, 16, 16), 2),
                                       ((48, 24, 40), (8, 8, 8), 4)])
def test_reference_reader_reads_the_written_file(tmp_path, nx, bnx, ng):
    m = MeshDesc(nx=nx, bnx=bnx, xmin=(-1.0, 0.0, 0.0), xmax=(1.0, 1.0, 0.5), ng=ng)
    data = _sample(m)
    path = str(tmp_path / "r.rst")
    restart.write(path, m, lambda g, k: data[(g, k)], 0.375, 2.5e-3, 41)
    ints, reals, blocks = read_as_reference(path, m)
    assert ints[:3] == [m.nblocks, restart.root_level(m.nbx), 41]
    assert tuple(ints[3:6]) == m.bnx and tuple(ints[6:9]) == m.nx
    assert ints[9] == 1                       # the file lists its blocks in tree order
    assert tuple(reals[:2]) == (0.375, 2.5e-3)
    assert tuple(reals[2:8]) == (-1.0, 0.0, 0.0, 1.0, 1.0, 0.5)
    assert tuple(reals[8:11]) == (1.0, 1.0, 1.0)
    assert sorted(blocks) == list(range(m.nblocks))
    for g in range(m.nblocks):
        for k in ("u", "b1", "b2", "b3"):
            assert np.array_equal(blocks[g][k], data[(g, k)]), (g, k)


def test_reference_reader_catches_planted_writer_mistakes(tmp_path, monkeypatch):
    m = MeshDesc(nx=(32, 16, 16), bnx=(8, 8, 8), xmin=(0.0,) * 3, xmax=(1.0,) * 3)
    data = _sample(m, 1)
    get = lambda g, k: data[(g, k)]    # noqa: E731
    # blocks listed row-major instead of in the tree's order: the reference would put
    # one block's data under another's location
    rowmajor = [(l1, l2, l3) for l3 in range(m.nbx[2]) for l2 in range(m.nbx[1])
                for l1 in range(m.nbx[0])]
    assert rowmajor != restart.ztree_order(m.nbx)
    path = str(tmp_path / "order.rst")
    with monkeypatch.context() as mp:
        mp.setattr(restart, "ztree_order", lambda nbx: rowmajor)
        restart.write(path, m, get, 0.0, 1e-3, 0)
    ints, _, blocks = read_as_reference(path, m)
    assert ints[9] == 0
    assert any(not np.array_equal(blocks[g]["u"], data[(g, "u")]) for g in range(m.nblocks))
    # RegionSize written without its 4 bytes of tail padding: the reader's packed header
    # is shifted, so the time, dt and block sizes no longer come out
    path = str(tmp_path / "pad.rst")
    import struct
    with monkeypatch.context() as mp:
        mp.setattr(restart, "REGION", struct.Struct("<9d3i"))
        restart.write(path, m, get, 0.5, 1e-3, 3)
    try:
        ints, reals, _ = read_as_reference(path, m)
        assert not (reals[0] == 0.5 and reals[1] == 1e-3 and ints[2] == 3)
    except RuntimeError:
        pass                                  # refused as broken: also caught


@pytest.mark.gpu
def test_checkpoint_resume_is_bitwise(tmp_path):
    from athena_amd.gpu import GpuMesh
    from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh
    print("case restart resume 32^3/16^3", flush=True)
    m = turbulence_mesh(32, 16)
    ic = Turbulence(m)
    blocks = [ic.block(g) for g in range(m.nblocks)]
    a = GpuMesh(m)
    a.load(blocks)
    a.initialize()
    a.run(3)
    t, dt, nc = a.state()
    path = str(tmp_path / "c.rst")
    restart.write(path, m, a.get, t, dt, nc)
    a.run(3)
    head, rb, _ = restart.read(path, m)
    b = GpuMesh(m)
    b.load([rb[g] for g in range(m.nblocks)])
    b.initialize()
    b.set_clock(head["time"], head["dt"], head["ncycle"])
    b.run(3)
    assert a.state() == b.state()
    for g in range(m.nblocks):
        for k in ("u", "w", "bcc", "b1", "b2", "b3"):
            assert np.array_equal(a.get(g, k), b.get(g, k)), (g, k)
    # the same file read by the restatement of the reference's reader and continued by
    # the oracle ends where the device run ends
    ints, reals, rblocks = read_as_reference(path, m)
    o = oracle_bind.OracleMesh(m)
    o.load([rblocks[g] for g in range(m.nblocks)])
    o.initialize()
    o.set_clock(reals[0], reals[1], ints[2])
    o.run(3)
    assert o.state()[:2] == a.state()[:2]
    for g in range(m.nblocks):
        for k in ("u", "w", "bcc", "b1", "b2", "b3"):
            assert np.array_equal(a.get(g, k), o.get(g, k)), ("oracle", g, k)


REGIONS = [((0.3, 0.3, 0.3), (0.4, 0.4, 0.4), 1), ((0.3, 1.0, 1.0), (0.4, 1.1, 1.1), 1),
           ((1.0, 1.0, 0.3), (1.1, 1.1, 0.4), 1), ((1.0, 0.3, 1.0), (1.1, 0.4, 1.1), 1),
           ((1.9, 0.3, 0.3), (2.0, 1.2, 1.2), 1)]


def _refined_mesh(regions=REGIONS):
    from athena_amd.mesh import refined
    from athena_amd.pgen.linear_wave import linear_wave_mesh
    return refined(linear_wave_mesh(32, 0, bnx=(8, 8, 8)), regions)


def read_refined_as_reference(path, m):
    """As read_as_reference, for a refined mesh: blocks keyed by the gid of the location
    (level included) the reference assigns each one's data to."""
    per = _per_block(m)
    ints, reals, locs, data = oracle_bind.restart_read(path, m.ng, m.nblocks,
                                                       per * m.nblocks)
    gid_of = {tuple(int(v) for v in r): g for g, r in enumerate(m.locs)}
    blocks = {}
    for i, loc in enumerate(locs[:ints[0]]):
        blocks[gid_of[tuple(int(v) for v in loc)]] = _split(m, data[per * i:per * (i + 1)])
    return ints, reals, [tuple(int(v) for v in r) for r in locs[:ints[0]]], blocks


@pytest.mark.parametrize("regions", [REGIONS, [((0.6, 0.5, 0.5), (0.9, 0.8, 0.8), 2)]],
                         ids=["two-levels", "three-levels"])
def test_refined_mesh_file_read_by_the_reference_reader(tmp_path, regions):
    # the reference rebuilds the tree from the listed locations (CreateRootGrid, then
    # AddMeshBlockWithoutRefine along each path) and assigns block data in the order of
    # its own leaf list: the written list must be that order, levels included
    m = _refined_mesh(regions)
    data = _sample(m, 2)
    path = str(tmp_path / "smr.rst")
    restart.write(path, m, lambda g, k: data[(g, k)], 0.125, 2e-3, 9)
    assert "refinement = static" in open(path, "rb").read().split(b"<par_end>")[0].decode()
    ints, reals, locs, blocks = read_refined_as_reference(path, m)
    assert ints[:3] == [m.nblocks, m.root_level, 9]
    assert ints[9] == 1
    assert locs == [tuple(int(v) for v in r) for r in m.locs]
    assert tuple(reals[:2]) == (0.125, 2e-3)
    for g in range(m.nblocks):
        for k in ("u", "b1", "b2", "b3"):
            assert np.array_equal(blocks[g][k], data[(g, k)]), (g, k)
    head, back, _ = restart.read(path, m)
    assert head["root_level"] == m.root_level and len(head["locations"]) == m.nblocks
    for g in range(m.nblocks):
        for k in ("u", "b1", "b2", "b3"):
            assert np.array_equal(back[g][k], data[(g, k)]), (g, k)


def test_refined_reader_catches_a_block_list_out_of_tree_order(tmp_path, monkeypatch):
    m = _refined_mesh()
    data = _sample(m, 3)
    order, gids, lev = restart.block_order(m)
    # swap two fine blocks of one parent: still a valid tree, a different list order
    q = next(i for i in range(len(order) - 1) if order[i][0] > lev and order[i + 1][0] > lev)
    order2, gids2 = list(order), list(gids)
    order2[q], order2[q + 1] = order2[q + 1], order2[q]
    gids2[q], gids2[q + 1] = gids2[q + 1], gids2[q]
    path = str(tmp_path / "swap.rst")
    with monkeypatch.context() as mp:
        mp.setattr(restart, "block_order", lambda mm: (order2, gids2, lev))
        restart.write(path, m, lambda g, k: data[(g, k)], 0.0, 1e-3, 0)
    ints, _, locs, blocks = read_refined_as_reference(path, m)
    assert ints[9] == 0
    # the reference puts each swapped block's data under the other's location
    a, b = gids[q], gids[q + 1]
    assert np.array_equal(blocks[a]["u"], data[(b, "u")])
    assert np.array_equal(blocks[b]["u"], data[(a, "u")])


@pytest.mark.gpu
def test_refined_checkpoint_resume_is_bitwise(tmp_path):
    from athena_amd.gpu import GpuMesh
    from athena_amd.pgen.linear_wave import LinearWave
    print("case restart resume on the refined linear-wave mesh", flush=True)
    m = _refined_mesh()
    lw = LinearWave(m, wave_flag=0)
    blocks = [lw.block(g) for g in range(m.nblocks)]
    a = GpuMesh(m)
    a.load(blocks)
    a.initialize()
    a.run(3)
    t, dt, nc = a.state()
    path = str(tmp_path / "c.rst")
    restart.write(path, m, a.get, t, dt, nc)
    a.run(3)
    head, rb, _ = restart.read(path, m)
    b = GpuMesh(m)
    b.load([rb[g] for g in range(m.nblocks)])
    b.initialize()
    b.set_clock(head["time"], head["dt"], head["ncycle"])
    b.run(3)
    assert a.state() == b.state()
    for g in range(m.nblocks):
        for k in ("u", "w", "bcc", "b1", "b2", "b3"):
            assert np.array_equal(a.get(g, k), b.get(g, k)), (g, k)
    # read the way the reference reads it and continued by the multilevel restatement
    ints, reals, _, rblocks = read_refined_as_reference(path, m)
    o = oracle_bind.OracleSmrMesh(m)
    o.load([rblocks[g] for g in range(m.nblocks)])
    o.initialize()
    o.set_clock(reals[0], reals[1], ints[2])
    o.run(3)
    assert o.state()[:2] == a.state()[:2]
    for g in range(m.nblocks):
        for k in ("u", "w", "bcc", "b1", "b2", "b3"):
            assert np.array_equal(a.get(g, k), o.get(g, k)), ("oracle", g, k)
