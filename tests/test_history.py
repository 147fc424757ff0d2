"""History output: the format of outputs/history.cpp:355-452 (CPU) and the device sums
against the same volume-weighted sums over the active cells formed on the host (GPU), on
a uniform and on a statically refined mesh, plus their conservation over cycles on a
periodic mesh."""
import numpy as np
import pytest

from athena_amd.history import NAMES, HistoryWriter, header, read


def test_header_and_line_format(tmp_path):
    h = header()
    assert h.startswith("# Athena++ history data\n# [1]=time     [2]=dt       [3]=mass     ")
    assert "[10]=tot-E   [11]=1-ME    [12]=2-ME    [13]=3-ME    \n" in h
    w = HistoryWriter(str(tmp_path / "a.hst"))
    ln = w.line(0.5, 0.01, list(range(11)))
    # each value: a blank then %12.5e (outputs.cpp:320-321), e.g. "  5.00000e-01"
    assert ln.startswith("  5.00000e-01  1.00000e-02  0.00000e+00  1.00000e+00")
    assert ln == "".join(" %12.5e" % v for v in [0.5, 0.01] + list(range(11))) + "\n"
    assert len(NAMES) == 13


def _refined_wave():
    from athena_amd.mesh import refined
    from athena_amd.pgen.linear_wave import LinearWave, linear_wave_mesh
    regions = [((0.3, 0.3, 0.3), (0.4, 0.4, 0.4), 1), ((0.3, 1.0, 1.0), (0.4, 1.1, 1.1), 1),
               ((1.0, 1.0, 0.3), (1.1, 1.1, 0.4), 1), ((1.0, 0.3, 1.0), (1.1, 0.4, 1.1), 1),
               ((1.9, 0.3, 0.3), (2.0, 1.2, 1.2), 1)]
    m = refined(linear_wave_mesh(32, 0, bnx=(8, 8, 8)), regions)
    return m, LinearWave(m, wave_flag=0, amp=1e-2)


def _turbulence():
    from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh
    m = turbulence_mesh(32, 16)
    return m, Turbulence(m)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["turbulence", "refined"])
def test_device_history_sums(case):
    # on the refined mesh each block's cell volumes are those of its level, so the sums
    # weight fine cells by 1/8 of a coarse one, as the reference's history does
    from athena_amd.gpu import GpuMesh
    from athena_amd.mesh import block_coords
    print("case history sums", case, flush=True)
    m, ic = _turbulence() if case == "turbulence" else _refined_wave()
    d = GpuMesh(m)
    d.load([ic.block(g) for g in range(m.nblocks)])
    d.initialize()
    h0 = d.history()
    ng = m.ng
    want = np.zeros(11)
    for g in range(m.nblocks):
        c = block_coords(m, g)
        sl = (slice(ng, -ng),) * 3
        vol = (c.dx1f[ng:-ng][None, None, :] * c.dx2f[ng:-ng][None, :, None]
               * c.dx3f[ng:-ng][:, None, None])
        u = d.get(g, "u")[(slice(None),) + sl]
        b = d.get(g, "bcc")[(slice(None),) + sl]
        want += [np.sum(vol * u[0]), np.sum(vol * u[1]), np.sum(vol * u[2]),
                 np.sum(vol * u[3]), np.sum(vol * 0.5 * u[1] ** 2 / u[0]),
                 np.sum(vol * 0.5 * u[2] ** 2 / u[0]), np.sum(vol * 0.5 * u[3] ** 2 / u[0]),
                 np.sum(vol * u[4]), np.sum(vol * 0.5 * b[0] ** 2),
                 np.sum(vol * 0.5 * b[1] ** 2), np.sum(vol * 0.5 * b[2] ** 2)]
    scale = np.maximum(np.abs(want), 1e-300)
    assert np.all(np.abs(h0 - want) <= 1e-12 * np.maximum(scale, abs(want[0])))
    # periodic: mass, momenta and total energy are conserved to round-off
    d.run(5)
    h1 = d.history()
    for q in (0, 1, 2, 3, 7):
        assert abs(h1[q] - h0[q]) <= 1e-13 * abs(h0[0] if q in (1, 2, 3) else h0[q]), q
    assert np.array_equal(d.history(), h1)      # deterministic
