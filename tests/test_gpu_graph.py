"""Single-mesh cycles replayed as one captured HIP graph (ath_gpu_mesh_set_graph) against
the eager cycles, bit for bit over whole registers and the clock after every cycle, with
the program's own eager work (register copies to the host, history sums) between the
replays; the graph is recaptured when the mesh's configuration changes (recovery mode,
fused band height, tiles), and kernel timing falls back to eager launches."""
import numpy as np
import pytest

from athena_amd.gpu import C2P_MODES, GpuMesh
from athena_amd.pgen.shock_tube import ShockTube, rj2a_mesh
from athena_amd.pgen.sr import ShearLayer, shear_mesh
from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh

pytestmark = pytest.mark.gpu
KEYS = ("u", "w", "bcc", "b1", "b2", "b3")


def snap(d):
    return {(g, k): d.get(g, k) for g in range(d.m.nblocks) for k in KEYS}, d.state()


def same(a, b, what):
    (sa, ca), (sb, cb) = a, b
    assert ca == cb, (what, ca, cb)
    for key in sa:
        assert np.array_equal(sa[key], sb[key]), (what, key)


def mesh(m, ic, mode=None, net=None):
    d = GpuMesh(m)
    d.load([ic.block(g) for g in range(m.nblocks)])
    if mode is not None:
        d.set_c2p(C2P_MODES[mode], net)
    d.initialize()
    return d


@pytest.mark.parametrize("case", ["turbulence-hlld-plm", "sr-shear-learned", "rj2a-outflow"])
def test_graph_cycles_equal_eager_cycles(case):
    print("case graph", case, flush=True)
    net = None
    if case == "turbulence-hlld-plm":
        m = turbulence_mesh(32, 16)
        ic, mode = Turbulence(m), None
    elif case == "sr-shear-learned":
        from athena_amd import ml
        m = shear_mesh(16, 8)
        ic, mode, net = ShearLayer(m), "learned", ml.load_mlp()
    else:
        m = rj2a_mesh(64, shock_dir=3, bnx=(4, 4, 16))
        ic, mode = ShockTube(m), None
    eager, graph = mesh(m, ic, mode, net), mesh(m, ic, mode, net)
    graph.set_graph(True)
    for c in range(4):
        eager.run(1)
        graph.run(1)
        same(snap(graph), snap(eager), "cycle %d" % c)
        graph.history()                       # eager device work between replays
    # several cycles per call, replayed back to back
    eager.run(3)
    graph.run(3)
    same(snap(graph), snap(eager), "three cycles")


def test_graph_is_recaptured_when_the_configuration_changes():
    print("case graph recapture", flush=True)
    m = turbulence_mesh(32, 16)
    ic = Turbulence(m)
    eager, graph = mesh(m, ic), mesh(m, ic)
    graph.set_graph(True)
    eager.run(1)
    graph.run(1)
    for d in (eager, graph):
        d.set_fused(-1)                       # another stage kernel sequence
    eager.run(2)
    graph.run(2)
    same(snap(graph), snap(eager), "after set_fused")
    for d in (eager, graph):
        d.set_fused(0)
        d.set_tiles([4, 0, 3, 2])             # other launch grids
    eager.run(2)
    graph.run(2)
    same(snap(graph), snap(eager), "after set_tiles")
    for d in (eager, graph):
        d.set_tiles([6, 6, 0, 0])             # the flux z-march: another kernel
    eager.run(2)
    graph.run(2)
    same(snap(graph), snap(eager), "after the flux z-march tile")
    graph.kernel_timing(True)                 # timed cycles run eagerly
    eager.run(1)
    graph.run(1)
    assert graph.kernel_stats()
    graph.kernel_timing(False)
    same(snap(graph), snap(eager), "timed cycle")
    graph.set_graph(False)
    eager.run(1)
    graph.run(1)
    same(snap(graph), snap(eager), "graph off")
