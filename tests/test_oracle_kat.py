"""CPU tests: the oracle against the reference's own known answers.

* RMS L1 linear-wave errors of the compiled reference (tests/golden/) -- pins the
  whole VL2 + PLM + HLLD + CT + con2prim restatement, including the ghost exchange,
  the EMF averaging and the CFL time step.
* mhd_linwave.py thresholds and convergence ratio.
* invariants: conservation under periodic BCs, div B at round-off, MeshBlock
  decomposition independence (mpi_linwave.py / omp_linwave.py analogue).
"""
import json
import os

import numpy as np
import pytest

from athena_amd.pgen.linear_wave import LinearWave, linear_wave_mesh
from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh
from harness import div_b, totals
from oracle_bind import OracleMesh

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "linear_wave_errors.json")))


def run_linwave(nx1, wave_flag, vflow, tlim, **kw):
    m = linear_wave_mesh(nx1, **kw)
    lw = LinearWave(m, wave_flag=wave_flag, amp=1e-6, vflow=vflow)
    o = OracleMesh(m)
    o.load([lw.block(g) for g in range(m.nblocks)])
    o.initialize(tlim)
    o.run(1 << 30, tlim)
    rms, l1, mx = lw.errors(o.block)
    return rms, o, m


@pytest.mark.parametrize("case", [c for c in GOLD["cases"] if c["nx1"] == 32],
                         ids=lambda c: "%s%d" % (c["name"], c["nx1"]))
def test_oracle_matches_reference_linear_wave_errors(case):
    rms, _, _ = run_linwave(case["nx1"], case["wave_flag"], case["vflow"], case["tlim"])
    # the reference prints %e (7 significant digits)
    assert float("%e" % rms) == case["rms_l1"], (rms, case)


@pytest.mark.slow
@pytest.mark.parametrize("case", [c for c in GOLD["cases"] if c["nx1"] == 64],
                         ids=lambda c: "%s%d" % (c["name"], c["nx1"]))
def test_oracle_matches_reference_linear_wave_errors_64(case):
    if not os.environ.get("ATH_SLOW"):
        pytest.skip("set ATH_SLOW=1 for the 64^3 known-answer runs (~1 min each)")
    rms, _, _ = run_linwave(case["nx1"], case["wave_flag"], case["vflow"], case["tlim"])
    # the printed digits of the reference's own uniform-grid run; mhd_linwave.py's
    # absolute thresholds (thresholds_at_64) belong to its static-refinement runs,
    # which the uniform errors exceed, so they are not asserted here
    assert float("%e" % rms) == case["rms_l1"], (rms, case)


def test_convergence_ratio_from_golden():
    by = {}
    for c in GOLD["cases"]:
        by.setdefault(c["name"], {})[c["nx1"]] = c["rms_l1"]
    for name, v in by.items():
        assert v[64] / v[32] <= GOLD["max_convergence_ratio"], name


def test_oracle_decomposition_independent():
    """Same problem, 1 MeshBlock vs 16 MeshBlocks: EMF averaging and ghost exchange
    reproduce the single-block run to round-off (mpi_linwave.py:66-74 analogue)."""
    r1, o1, m1 = run_linwave(32, 0, 0.0, 0.25)
    r2, o2, m2 = run_linwave(32, 0, 0.0, 0.25, bnx=(8, 8, 8))
    assert m2.nblocks == 16
    assert abs(r1 - r2) <= 1e-12 * r1


def test_oracle_conservation_and_divb():
    m = turbulence_mesh(16, 8)
    t = Turbulence(m)
    o = OracleMesh(m)
    o.load([t.block(g) for g in range(m.nblocks)])
    o.initialize()
    tot0 = totals(m, o.get)
    o.run(5)
    tot1 = totals(m, o.get)
    scale = np.array([tot0[0], abs(tot0[0]), abs(tot0[0]), abs(tot0[0]), tot0[4]])
    assert np.all(np.abs(tot1 - tot0) <= 1e-13 * scale), (tot1 - tot0)
    assert div_b(m, o.get) < 1e-14


def test_oracle_lr_symmetry():
    """mhd_linwave.py:159: the L- and R-going fast waves give identical errors."""
    m = linear_wave_mesh(16)
    errs = []
    for wf in (0, 6):
        lw = LinearWave(m, wave_flag=wf, amp=1e-6)
        o = OracleMesh(m)
        o.load([lw.block(g) for g in range(m.nblocks)])
        o.initialize(0.1)
        o.run(1 << 30, 0.1)
        errs.append(lw.errors(o.block)[0])
    assert errs[0] == pytest.approx(errs[1], rel=1e-12)
