"""SR MHD linear waves on the device (tst/regression/scripts/tests/sr/sr_mhd_linwave.py:
7 waves, res x res/2 x res/2 in (res/2)^3 MeshBlocks, res = 16 and 32, HLLD, PLM, VL2,
cfl 0.3), and device-vs-oracle parity on one of them.

The test's convergence criterion (error ratio <= 0.4 from 16 to 32 cells) holds for
every wave.  Its absolute bounds at 32 cells (2-4e-8) are not met: the errors here are
1.7-3x larger, identical on the oracle, dominated by the PLM reconstruction (a 10x
smaller CFL number leaves them unchanged, PPM cuts them 3x) although the initial states
are exact eigenvectors and the SR HLLD resolves isolated Alfven and contact jumps to
1e-7.  That test is kept as a strict expected failure so the gap stays visible.
"""
import numpy as np
import pytest

from athena_amd.gpu import GpuMesh
from athena_amd.pgen.sr_linear_wave import (ERROR_RATIO, HIGH_RES_ERRORS, NAMES,
                                            SRLinearWave, sr_linear_wave_mesh)
from harness import compare
from oracle_bind import OracleMesh

pytestmark = pytest.mark.gpu


def device_rms(wave_flag, res):
    m = sr_linear_wave_mesh(res)
    lw = SRLinearWave(m, wave_flag)
    d = GpuMesh(m)
    d.load([lw.block(g) for g in range(m.nblocks)])
    tl = lw.tlim()
    d.initialize(tl)
    n = d.run(1 << 30, tl)
    return lw.errors(lambda g, k: d.get(g, k))[0], n


@pytest.fixture(scope="module")
def errors():
    out = {}
    for w in range(7):
        for res in (16, 32):
            print("case sr linear wave", NAMES[w], res, flush=True)
            out[w, res] = device_rms(w, res)
            print("  cycles %d rms %e" % (out[w, res][1], out[w, res][0]), flush=True)
    return out


def test_sr_linwave_convergence(errors):
    for w in range(7):
        ratio = errors[w, 32][0] / errors[w, 16][0]
        print("%-18s rms %e -> %e  ratio %.3f  (bound at 32: %.1e)"
              % (NAMES[w], errors[w, 16][0], errors[w, 32][0], ratio, HIGH_RES_ERRORS[w]))
        assert ratio <= ERROR_RATIO, (w, ratio)


@pytest.mark.xfail(strict=True, reason="PLM errors 1.7-3x the reference test's absolute "
                                       "bounds, on the oracle as on the device")
def test_sr_linwave_reference_bounds(errors):
    for w in range(7):
        assert errors[w, 32][0] <= HIGH_RES_ERRORS[w], (NAMES[w], errors[w, 32][0])


@pytest.mark.parametrize("wave_flag", [5, 2])
def test_sr_linwave_parity(wave_flag):
    print("case sr linear wave parity", NAMES[wave_flag], flush=True)
    m = sr_linear_wave_mesh(16)
    lw = SRLinearWave(m, wave_flag)
    blocks = [lw.block(g) for g in range(m.nblocks)]
    o = OracleMesh(m)
    o.load(blocks)
    d = GpuMesh(m)
    d.load(blocks)
    tl = lw.tlim()
    o.initialize(tl)
    d.initialize(tl)
    assert o.run(1 << 30, tl) == d.run(1 << 30, tl)
    res = compare(m, lambda g, k: d.get(g, k), lambda g, k: o.get(g, k), floor=1e-9)
    for k, (r, eq) in sorted(res.items()):
        print("  %-5s relL1=%.3e bitwise=%s" % (k, r, eq))
    assert all(r <= 1e-12 for r, _ in res.values()), res
    ro = lw.errors(lambda g, k: o.get(g, k))[0]
    rd = lw.errors(lambda g, k: d.get(g, k))[0]
    assert abs(ro - rd) <= 1e-6 * ro
