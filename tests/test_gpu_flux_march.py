"""GPU checks of the flux kernel's z-march (flux tile options 6 and 7, k_flux_zm: each
lane walks a segment of about 8 or 16 planes of one column, reconstructs its cell once
along x1, passing the high-face states one lane up, and once along x3, carrying them to
the next plane).  Every face state is the value k_flux3 computes for that face, so whole
runs are bit for bit the default runs, which the parity tests tie to the CPU oracle; one
case is also compared with the oracle directly.  Cases cover PLM and PPM (donor-cell
launches keep linear chunks under these tiles), every Riemann solver, NGHOST 2-4,
MeshBlocks of unequal sides and the smallest ones, outflow faces, a statically refined
mesh, and segments taller than the box."""
import numpy as np
import pytest

from athena_amd.gpu import GpuMesh
from athena_amd.mesh import MeshDesc
from athena_amd.pgen.linear_wave import LinearWave, linear_wave_mesh
from athena_amd.pgen.sr import ShearLayer, shear_mesh
from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh
from harness import compare
from oracle_bind import OracleMesh

pytestmark = pytest.mark.gpu
KEYS = ("u", "w", "bcc", "b1", "b2", "b3")
OUTFLOW = ("outflow",) * 6


def _lw(bnx, rsolver, xorder, ng, nx1=32, wave_flag=1):
    m = linear_wave_mesh(nx1, bnx=bnx, rsolver=rsolver, xorder=xorder, ng=ng)
    return m, LinearWave(m, wave_flag=wave_flag, amp=1e-6)


def _turb(n, bn, **kw):
    m = turbulence_mesh(n, bn, **kw)
    return m, Turbulence(m)


def _sr_hlld():
    m = MeshDesc(nx=(32, 32, 32), bnx=(16, 16, 16), xmin=(-0.5,) * 3, xmax=(0.5,) * 3,
                 gamma=4.0 / 3.0, cfl=0.3, xorder=2, ng=2, rsolver="hlld", relativistic=True)
    return m, ShearLayer(m)


def _smr():
    # the reference's refined linear-wave mesh (two levels, 72 MeshBlocks of 8^3)
    from athena_amd.mesh import refined
    regions = [((0.3, 0.3, 0.3), (0.4, 0.4, 0.4), 1), ((0.3, 1.0, 1.0), (0.4, 1.1, 1.1), 1),
               ((1.0, 1.0, 0.3), (1.1, 1.1, 0.4), 1), ((1.0, 0.3, 1.0), (1.1, 0.4, 1.1), 1),
               ((1.9, 0.3, 0.3), (2.0, 1.2, 1.2), 1)]
    m = refined(linear_wave_mesh(32, 0, bnx=(8, 8, 8)), regions)
    return m, LinearWave(m, wave_flag=0)


def _sr_ppm():
    m = shear_mesh(32, 16, xorder=3, ng=3)
    return m, ShearLayer(m)


CASES = {
    "lw_plm_hlld_16x8x4": lambda: _lw((16, 8, 4), "hlld", 2, 2),
    "lw_ppm_hlle_4x16x8_ng3": lambda: _lw((4, 16, 8), "hlle", 3, 3),
    "lw_ppm_hlld_8x4x16_ng4": lambda: _lw((8, 4, 16), "hlld", 3, 4),
    "lw_plm_hlld_2cubed": lambda: _lw((2, 2, 2), "hlld", 2, 2, wave_flag=0),
    "lw_ppm_hlle_3cubed": lambda: _lw((3, 3, 3), "hlle", 3, 3, nx1=24, wave_flag=0),
    "turb_plm_hlld": lambda: _turb(32, 16),
    "turb_dc_hlle": lambda: _turb(32, 16, rsolver="hlle", xorder=1),
    "turb_ppm_hlld_outflow": lambda: _turb(16, 8, xorder=3, ng=3, bc=OUTFLOW),
    "sr_shear_ppm_hlle": _sr_ppm,
    "sr_shear_plm_hlld": _sr_hlld,
    "smr_linwave_plm_hlld": _smr,
}


def run(m, ic, ncycles, tiles=None):
    d = GpuMesh(m)
    d.load([ic.block(g) for g in range(m.nblocks)])
    if tiles is not None:
        d.set_tiles(tiles)
    d.initialize()
    d.run(ncycles)
    return d


def snapshot(d):
    return {(g, k): d.get(g, k) for g in range(d.m.nblocks) for k in KEYS}, d.state()


@pytest.mark.parametrize("name", list(CASES))
def test_flux_march_runs_bit_for_bit_as_the_default(name):
    m, ic = CASES[name]()
    print("case flux march", name, flush=True)
    ref = snapshot(run(m, ic, 5))
    for flux_tile in (6, 7):
        d = run(m, ic, 5, tiles=[flux_tile, 6, 0, 0])
        assert list(d.tiles()) == [flux_tile, 6, 0, 0]
        s, c = snapshot(d)
        assert c == ref[1], (flux_tile, c, ref[1])
        for key in s:
            assert np.array_equal(s[key], ref[0][key]), (flux_tile, key)


def test_flux_march_against_the_oracle():
    m, ic = _turb(32, 16)
    print("case flux march vs oracle, turbulence 32^3/16^3 PLM HLLD", flush=True)
    blocks = [ic.block(g) for g in range(m.nblocks)]
    o = OracleMesh(m)
    o.load(blocks)
    o.initialize(1e300)
    o.run(6, 1e300)
    d = run(m, ic, 6, tiles=[6, 6, 0, 0])
    res = compare(m, lambda g, k: d.get(g, k), lambda g, k: o.get(g, k))
    assert all(eq for _, eq in res.values()), res
    assert o.state()[2] == d.state()[2]


def test_ensembles_refuse_the_flux_march():
    from athena_amd.gpu import AthGpuError, Ensemble
    m, ic = _turb(32, 16)
    d = GpuMesh(m)
    d.load([ic.block(g) for g in range(m.nblocks)])
    d.initialize()
    e = Ensemble([d])
    with pytest.raises(AthGpuError, match="corner"):
        e.set_tiles([6, 6, 0, 0], [0, 0, 0, 0])


def test_single_mesh_autotune_times_the_flux_march_and_keeps_the_bits():
    # PPM, where the z-march can win; the tuned mesh's state is unchanged by tuning and
    # its run is bit for bit the default run whatever flux tile was kept
    m, ic = _turb(32, 16, xorder=3, ng=3)
    print("case autotune with the flux march, turbulence 32^3/16^3 PPM HLLD", flush=True)
    ref = snapshot(run(m, ic, 5))
    d = GpuMesh(m)
    d.load([ic.block(g) for g in range(m.nblocks)])
    d.initialize()
    before = snapshot(d)
    d.autotune(reps=2)
    print("  chosen tiles %s" % list(d.tiles()), flush=True)
    after = snapshot(d)
    assert after[1] == before[1]
    for key in before[0]:
        assert np.array_equal(after[0][key], before[0][key]), key
    d.run(5)
    s, c = snapshot(d)
    assert c == ref[1]
    for key in s:
        assert np.array_equal(s[key], ref[0][key]), key
