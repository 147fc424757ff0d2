"""GPU parity of the per-MeshBlock entry points over caller-owned device arrays
(ath_gpu_block_*), each against the oracle's restatement of the same reference function
on the same block state, bit for bit:

  Reconstruction::{DonorCell,PiecewiseLinear,PiecewiseParabolic}X1/X2/X3 on pencils,
  Hydro::RiemannSolver on pencils, Hydro::CalculateFluxes, Field::ComputeCornerE,
  Field::CT, Hydro::AddFluxDivergence, EquationOfState::ConservedToPrimitive over whole
  and partial ranges, Hydro::NewBlockTimeStep.

The block state is a decaying-turbulence (Newtonian) or shear-layer (SR) MeshBlock after
two VL2 cycles of the oracle, so every branch sees non-trivial data; device arrays start
as copies of the oracle's, so positions outside a function's range must also agree.
"""
import numpy as np
import pytest

from athena_amd.gpu import AthGpuError, Block
from athena_amd.mesh import MeshDesc
from athena_amd.pgen.sr import ShearLayer
from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh
from oracle_bind import OracleMesh

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

REGS = ("flux1", "flux2", "flux3", "e3_x1f", "e2_x1f", "e1_x2f", "e3_x2f", "e2_x3f",
        "e1_x3f", "wght1", "wght2", "wght3", "e1", "e2", "e3")
CASES = [("newtonian-hlld-ng2", False, "hlld", 2), ("newtonian-hlle-ng3", False, "hlle", 3),
         ("sr-hlle-ng3", True, "hlle", 3), ("sr-hlld-ng3", True, "hlld", 3)]


def make_state(relativistic, rsolver, ng):
    if relativistic:
        m = MeshDesc(nx=(16, 16, 16), bnx=(16, 16, 16), xmin=(-0.5,) * 3, xmax=(0.5,) * 3,
                     gamma=4.0 / 3.0, cfl=0.3, xorder=2, ng=ng, rsolver=rsolver,
                     relativistic=True)
        ic = ShearLayer(m)
    else:
        m = turbulence_mesh(16, 16, rsolver=rsolver, ng=ng)
        ic = Turbulence(m)
    o = OracleMesh(m)
    o.load([ic.block(0)])
    o.initialize()
    o.run(2)
    return m, o


class Dev:
    """Device copies of the oracle block's registers (torch tensors)."""

    def __init__(self, o, names):
        self.t = {n: torch.from_numpy(np.ascontiguousarray(o.get(0, n)).ravel().copy()).cuda()
                  for n in names}

    def __getitem__(self, n):
        return self.t[n].data_ptr()

    def get(self, n):
        torch.cuda.synchronize()
        return self.t[n].cpu().numpy()


def same(a, b, what):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    assert bad.size == 0, (what, bad[:5], a[bad[:5]], b[bad[:5]])


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request):
    name, sr, rs, ng = request.param
    print("case block api", name, flush=True)
    m, o = make_state(sr, rs, ng)
    return m, o


def test_calculate_fluxes_corner_ct_divergence(case):
    m, o = case
    blk = Block(m)
    dt = o.state()[1]
    for order in ((1, 2, 3) if m.ng >= 3 else (1, 2)):
        d = Dev(o, ("w", "bcc", "b1", "b2", "b3", "u") + REGS)
        regs = {n: d[n] for n in REGS}
        blk.calculate_fluxes(d["w"], d["b1"], d["b2"], d["b3"], d["bcc"], order, dt, regs)
        o.call(0, 0, order=order)
        for n in REGS[:12]:
            same(d.get(n), o.get(0, n), "order %d %s" % (order, n))
        blk.compute_corner_e(d["w"], d["bcc"], regs)
        o.call(0, 1)
        for n in ("e1", "e2", "e3"):
            same(d.get(n), o.get(0, n), "order %d %s" % (order, n))
        wght = 0.37 * dt
        blk.ct(wght, regs, d["b1"], d["b2"], d["b3"])
        blk.add_flux_divergence(wght, regs, d["u"])
        o.call(0, 2, wght=wght)
        o.call(0, 3, wght=wght)
        for n in ("b1", "b2", "b3", "u"):
            same(d.get(n), o.get(0, n), "order %d %s" % (order, n))
    blk.close()


def test_cons2prim_ranges_and_new_dt(case):
    m, o = case
    blk = Block(m)
    nc = m.nc
    for rng in ((0, nc[0] - 1, 0, nc[1] - 1, 0, nc[2] - 1),
                (m.ng, nc[0] - m.ng - 1, 1, nc[1] - 3, 2, nc[2] - 2)):
        d = Dev(o, ("u", "w", "w1", "bcc", "b1", "b2", "b3"))
        blk.cons2prim(d["u"], d["w1"], d["b1"], d["b2"], d["b3"], d["w"], d["bcc"], rng)
        o.call(0, 4, rng=rng)
        for n in ("u", "w", "bcc"):
            same(d.get(n), o.get(0, n), "range %s %s" % (rng, n))
    d = Dev(o, ("w", "bcc", "b1", "b2", "b3"))
    dt = torch.zeros(1, dtype=torch.float64, device="cuda")
    blk.new_dt(d["w"], d["bcc"], d["b1"], d["b2"], d["b3"], dt.data_ptr())
    want = o.call(0, 5)[0]
    got = float(dt.cpu()[0])
    print("  new_block_dt %.17g (oracle %.17g)" % (got, want), flush=True)
    assert got == want
    blk.close()


def test_reconstruct_and_riemann_pencils(case):
    m, o = case
    blk = Block(m)
    nc1 = m.nc[0]
    ng = m.ng
    dt = o.state()[1]
    for order in ((1, 2, 3) if ng >= 3 else (1, 2)):
        e = order - 1
        for d_ in range(3):
            # a pencil inside the stencil's reach: cells is-1 .. ie+1 (x1), cell j / k
            k, j = ng + 1, ng + 2
            if d_ == 0:
                il, iu = max(ng - 1, e), min(nc1 - ng, nc1 - 1 - max(e, 1))
            else:
                il, iu = ng - 1, nc1 - ng
            dv = Dev(o, ("w", "bcc"))
            wl = torch.zeros(7 * nc1, dtype=torch.float64, device="cuda")
            wr = torch.zeros(7 * nc1, dtype=torch.float64, device="cuda")
            blk.reconstruct(d_, order, k, j, il, iu, dv["w"], dv["bcc"], wl.data_ptr(),
                            wr.data_ptr())
            torch.cuda.synchronize()
            out = o.call(0, 6, order=order, rng=(il, iu, j, 0, k, d_),
                         out_size=2 * 7 * (nc1 + 1))
            owl = out[:7 * (nc1 + 1)].reshape(7, nc1 + 1)[:, :nc1]
            owr = out[7 * (nc1 + 1):].reshape(7, nc1 + 1)[:, :nc1]
            same(wl.cpu().numpy().reshape(7, nc1), owl, "wl dir %d order %d" % (d_, order))
            same(wr.cpu().numpy().reshape(7, nc1), owr, "wr dir %d order %d" % (d_, order))
            # the Riemann solve of these states on faces of the pencil
            fl, fr = (il + 1, iu) if d_ == 0 else (il, iu)
            fk, fj = (k, j)
            names = {0: ("b1", "flux1", "e3_x1f", "e2_x1f", "wght1"),
                     1: ("b2", "flux2", "e1_x2f", "e3_x2f", "wght2"),
                     2: ("b3", "flux3", "e2_x3f", "e1_x3f", "wght3")}[d_]
            dr = Dev(o, names)
            coords = {0: "dx1f", 1: "dx2f", 2: "dx3f"}
            if d_ == 0:
                dxw = o.get(0, "dx1f")
            else:   # CenterWidth along x2 / x3 is the same for every i of the pencil
                dxw = np.full(nc1, o.get(0, coords[d_])[j if d_ == 1 else k])
            dxw_t = torch.from_numpy(np.ascontiguousarray(dxw)).cuda()
            blk.riemann(fk, fj, fl, fr, d_ + 1, dr[names[0]], wl.data_ptr(), wr.data_ptr(),
                        dr[names[1]], dr[names[2]], dr[names[3]], dr[names[4]],
                        dxw_t.data_ptr(), dt)
            wlo = np.zeros((7, nc1 + 1))
            wro = np.zeros((7, nc1 + 1))
            wlo[:, :nc1] = wl.cpu().numpy().reshape(7, nc1)
            wro[:, :nc1] = wr.cpu().numpy().reshape(7, nc1)
            o.riemann_pencil(0, d_, fk, fj, fl, fr, wlo, wro)
            for n in names[1:]:
                same(dr.get(n), o.get(0, n), "riemann dir %d order %d %s" % (d_, order, n))
    blk.close()


def test_block_entries_reject_bad_ranges():
    m = turbulence_mesh(8, 8)
    blk = Block(m)
    nc1 = m.nc[0]
    with pytest.raises(AthGpuError, match="stencil"):
        blk.reconstruct(0, 2, 3, 3, 0, nc1 - 2, 1, 1, 1, 1)     # PLM reads cell -1
    with pytest.raises(AthGpuError, match="PPM"):
        blk.reconstruct(1, 3, 3, 3, 1, 5, 1, 1, 1, 1)           # nghost 2
    with pytest.raises(AthGpuError, match="outside"):
        blk.cons2prim(1, 1, 1, 1, 1, 1, 1, (0, nc1, 0, 1, 0, 1))
    with pytest.raises(AthGpuError, match="ivx"):
        blk.riemann(2, 2, 2, 4, 0, 1, 1, 1, 1, 1, 1, 1, 1, 0.1)
    blk.close()
