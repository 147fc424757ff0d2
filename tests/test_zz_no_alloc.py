"""ath_gpu_block_new_dt_noalloc (the allocation-free NewBlockTimeStep over a fixed device
scratch array) against ath_gpu_block_new_dt (which allocates its scratch) through the C
ABI: bit for bit on two calls in a row, the second on a smaller MeshBlock, so the second
call reads a scratch array still holding the first call's larger set of partial minima.
Both entries are also checked against the oracle's NewBlockTimeStep, and a MeshBlock
beyond the scratch limit is refused with an error."""
import numpy as np
import pytest

from athena_amd.gpu import AthGpuError, Block
from athena_amd.mesh import MeshDesc
from athena_amd.pgen.sr import ShearLayer
from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh
from oracle_bind import OracleMesh

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def state(n, relativistic=False):
    if relativistic:
        m = MeshDesc(nx=(n, n, n), bnx=(n, n, n), xmin=(-0.5,) * 3, xmax=(0.5,) * 3,
                     gamma=4.0 / 3.0, cfl=0.3, xorder=2, ng=2, rsolver="hlle",
                     relativistic=True)
        ic = ShearLayer(m)
    else:
        m = turbulence_mesh(n, n, rsolver="hlld", ng=2)
        ic = Turbulence(m)
    o = OracleMesh(m)
    o.load([ic.block(0)])
    o.initialize()
    o.run(2)
    return m, o


def dev(o, n):
    return torch.from_numpy(np.ascontiguousarray(o.get(0, n)).ravel().copy()).cuda()


@pytest.mark.parametrize("relativistic", [False, True], ids=["newtonian", "sr"])
def test_noalloc_new_dt_matches_allocating_entry(relativistic):
    # 48^3 active cells = 432 partial minima, then 16^3 = 16: the second call leaves
    # stale partials of the first call in scratch slots 16 .. 431, which must not be read
    for n in (48, 16):
        print("case no_alloc new_dt relativistic=%d n=%d" % (relativistic, n), flush=True)
        m, o = state(n, relativistic)
        blk = Block(m)
        t = {k: dev(o, k) for k in ("w", "bcc", "b1", "b2", "b3")}
        args = [t[k].data_ptr() for k in ("w", "bcc", "b1", "b2", "b3")]
        dt_a = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
        dt_b = torch.full((1,), -2.0, dtype=torch.float64, device="cuda")
        blk.new_dt(*args, dt_a.data_ptr())
        blk.new_dt_noalloc(*args, dt_b.data_ptr())
        torch.cuda.synchronize()
        a, b = float(dt_a.cpu()[0]), float(dt_b.cpu()[0])
        want = o.call(0, 5)[0]
        print("  new_dt %.17g noalloc %.17g oracle %.17g" % (a, b, want), flush=True)
        assert b == a
        assert a == want
        blk.close()


def test_noalloc_new_dt_refuses_oversized_block():
    # 260^3 active cells need 68,657 partial minima, beyond the 65,536 the scratch holds;
    # the check runs on the host before any launch, so the arrays are never read
    m = MeshDesc(nx=(260, 260, 260), bnx=(260, 260, 260), xmin=(0.0,) * 3, xmax=(1.0,) * 3,
                 gamma=5.0 / 3.0, cfl=0.3, xorder=2, ng=2, rsolver="hlld")
    blk = Block(m)
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda")
    p = dummy.data_ptr()
    with pytest.raises(AthGpuError, match="scratch limit"):
        blk.new_dt_noalloc(p, p, p, p, p, p)
    blk.close()
