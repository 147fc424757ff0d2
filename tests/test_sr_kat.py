"""SR MHD known answers from the reference's own data: the Mignone-Ugliano-Bodo shock
tubes (tst/regression/scripts/tests/sr/mhd_shocks_hlle.py:53-85, solutions
tst/regression/data/sr_mhd_shock{1,2,4}_hlld.vtk committed as tests/golden/*.npz).
Run with HLLE + PLM + VL2 + CT, compared with the reference test's tolerances
(relative L1; 0.0 means the component must stay exactly equal)."""
import os

import numpy as np
import pytest

from athena_amd.pgen.sr import MUB, ShockTube, shock_tube_mesh
from oracle_bind import OracleMesh

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
NAMES = ["dens", "Etot", "mom1", "mom2", "mom3", "Bcc1", "Bcc2", "Bcc3"]


def ref(problem):
    return dict(np.load(os.path.join(GOLDEN, "sr_mhd_shock%d_hlld.npz" % problem)))


def assert_tube(res, tols):
    for name, (eps, norm), tol in zip(NAMES, res, tols):
        if tol == 0.0:
            assert eps == 0.0, (name, eps)
        else:
            assert eps / norm <= tol, (name, eps / norm, tol)


def run_oracle(problem):
    m = shock_tube_mesh(problem)
    st = ShockTube(m, problem)
    o = OracleMesh(m)
    o.load([st.block(0)])
    tl = MUB[problem]["tlim"]
    o.initialize(tl)
    o.run(1 << 30, tl)
    return st.compare(o.block, ref(problem))


def test_oracle_sr_shock_tube_1():
    assert_tube(run_oracle(1), MUB[1]["tols"])


@pytest.mark.parametrize("problem", [2, 4])
def test_oracle_sr_shock_tubes_slow(problem):
    if not os.environ.get("ATH_SLOW"):
        pytest.skip("set ATH_SLOW=1 (about 30 s each on one core)")
    assert_tube(run_oracle(problem), MUB[problem]["tols"])
