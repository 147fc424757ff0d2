"""Restart files: the reference's binary layout (outputs/restart.cpp:36-240), block
order of the reference's tree (meshblock_tree.cpp:87-110, 336-352) and, on the GPU,
checkpoint/resume that continues bit for bit."""
import os

import numpy as np
import pytest

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
