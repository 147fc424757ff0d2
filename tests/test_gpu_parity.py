"""GPU parity: the device library against the CPU oracle on identical inputs.

Bar: relative L1 <= 1e-12 per variable after the run (the stated FP64 tolerance); the
Newtonian path is expected to be bit-identical (same operation order, no FMA
contraction, IEEE division/sqrt), which is reported and, for the single-block linear
wave, asserted.  Also: the reference's own linear-wave errors reproduced on the GPU,
conservation to round-off and div(B) at round-off.
"""
import json
import os

import numpy as np
import pytest

from athena_amd.gpu import GpuMesh
from athena_amd.pgen.linear_wave import LinearWave, linear_wave_mesh
from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh
from harness import compare, div_b, totals
from oracle_bind import OracleMesh

pytestmark = pytest.mark.gpu
TOL = 1e-12
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "linear_wave_errors.json")))


def both(m, blocks, ncycles=None, tlim=float("inf")):
    o = OracleMesh(m)
    o.load(blocks)
    d = GpuMesh(m)
    d.load(blocks)
    o.initialize(tlim if np.isfinite(tlim) else 1e300)
    d.initialize(tlim)
    n = ncycles if ncycles is not None else (1 << 30)
    no = o.run(n, tlim if np.isfinite(tlim) else 1e300)
    nd = d.run(n, tlim)
    assert no == nd
    return o, d


def check(m, o, d, bitwise=False):
    res = compare(m, lambda g, k: d.get(g, k), lambda g, k: o.get(g, k))
    bad = {k: v for k, v in res.items() if not v[0] <= TOL}
    print(flush=True)
    for k, (r, eq) in sorted(res.items()):
        print("  %-5s relL1=%.3e bitwise=%s" % (k, r, eq), flush=True)
    assert not bad, bad
    if bitwise:
        assert all(eq for _, eq in res.values()), res
    to, tg = o.state(), d.state()
    assert to[2] == tg[2]
    assert abs(to[0] - tg[0]) <= 1e-14 * max(1.0, abs(to[0]))
    assert abs(to[1] - tg[1]) <= 1e-13 * to[1]


@pytest.mark.parametrize("case", [c for c in GOLD["cases"] if c["nx1"] == 32],
                         ids=lambda c: c["name"])
def test_linear_wave_single_block(case):
    print("case linear_wave", case["name"], flush=True)
    m = linear_wave_mesh(32)
    lw = LinearWave(m, wave_flag=case["wave_flag"], amp=1e-6, vflow=case["vflow"])
    o, d = both(m, [lw.block(g) for g in range(m.nblocks)], tlim=case["tlim"])
    check(m, o, d, bitwise=True)
    rms = lw.errors(d.block)[0]
    assert float("%e" % rms) == case["rms_l1"]


def test_linear_wave_multiblock():
    print("case linear_wave fast 4x2x2 blocks", flush=True)
    m = linear_wave_mesh(32, bnx=(8, 8, 8))
    lw = LinearWave(m, wave_flag=0, amp=1e-6)
    o, d = both(m, [lw.block(g) for g in range(m.nblocks)], tlim=0.25)
    check(m, o, d)


@pytest.mark.parametrize("rsolver,xorder,ng", [("hlld", 2, 2), ("hlle", 2, 2),
                                               ("hlld", 3, 3), ("hlle", 1, 2)])
def test_turbulence_multiblock(rsolver, xorder, ng):
    print("case turbulence 32^3/16^3", rsolver, xorder, ng, flush=True)
    m = turbulence_mesh(32, 16, rsolver=rsolver, xorder=xorder, ng=ng)
    t = Turbulence(m)
    o, d = both(m, [t.block(g) for g in range(m.nblocks)], ncycles=10)
    check(m, o, d)


@pytest.mark.parametrize("bnx,rsolver,xorder,ng", [((16, 8, 4), "hlld", 2, 2),
                                                   ((4, 16, 8), "hlle", 3, 3),
                                                   ((8, 4, 16), "hlld", 3, 4)])
def test_linear_wave_noncubic_blocks(bnx, rsolver, xorder, ng):
    # MeshBlocks of unequal sides and NGHOST up to 4: every work map, the corner
    # kernel's z-march tile width and the ghost exchange's numbering per direction
    print("case linear_wave 32x16x16 blocks", bnx, rsolver, xorder, ng, flush=True)
    m = linear_wave_mesh(32, bnx=bnx, rsolver=rsolver, xorder=xorder, ng=ng)
    lw = LinearWave(m, wave_flag=1, amp=1e-6)
    o, d = both(m, [lw.block(g) for g in range(m.nblocks)], ncycles=6)
    check(m, o, d, bitwise=True)


@pytest.mark.parametrize("nx1,bnx,rsolver,xorder,ng", [(32, (2, 2, 2), "hlld", 2, 2),
                                                       (24, (3, 3, 3), "hlle", 3, 3)])
def test_linear_wave_smallest_blocks(nx1, bnx, rsolver, xorder, ng):
    # MeshBlocks as small as the ghost width allows: each ghost layer is a neighbour's
    # whole active region and the corner z-march covers fewer planes than its length
    print("case linear_wave smallest blocks", nx1, bnx, rsolver, xorder, ng, flush=True)
    m = linear_wave_mesh(nx1, bnx=bnx, rsolver=rsolver, xorder=xorder, ng=ng)
    lw = LinearWave(m, wave_flag=0, amp=1e-6)
    o, d = both(m, [lw.block(g) for g in range(m.nblocks)], ncycles=4)
    check(m, o, d, bitwise=True)


def test_conservation_and_divb_gpu():
    print("case conservation 32^3/16^3", flush=True)
    m = turbulence_mesh(32, 16)
    t = Turbulence(m)
    d = GpuMesh(m)
    d.load([t.block(g) for g in range(m.nblocks)])
    d.initialize()
    tot0 = totals(m, d.get)
    d.run(10)
    tot1 = totals(m, d.get)
    scale = np.array([tot0[0], tot0[0], tot0[0], tot0[0], tot0[4]])
    assert np.all(np.abs(tot1 - tot0) <= 1e-13 * scale), tot1 - tot0
    assert div_b(m, d.get) < 1e-13
