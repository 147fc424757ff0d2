"""CPU checks of the flux kernel's z-march launch map (ath_gpu_flux_march_position, the
map k_flux_zm itself evaluates; no device call).  The kernel computes, for every owned
lane and every plane k0..k1 of its segment, the faces k_flux3 computes for cell
(k, j, i): the flux box [is-1, ie+1]^3 of each MeshBlock (calculate_fluxes.cpp:68-76,
176-182, 299-302).  Its x1 faces take their left states from the lane before, so that
lane must hold cell (k, j, i-1) at the same planes whenever the face exists (i >= is);
lane 0 of a wavefront repeats the previous wavefront's last position and lanes past the
end compute a clamped copy, so their stencils must stay inside the box too."""
import pytest

from athena_amd.gpu import flux_march_position


@pytest.mark.parametrize("shape,seg", [((8, 8, 8, 2, 3), 8), ((8, 8, 8, 2, 3), 16),
                                       ((16, 8, 4, 2, 3), 8), ((4, 16, 8, 3, 2), 16),
                                       ((2, 2, 2, 2, 5), 8), ((3, 3, 3, 3, 4), 8),
                                       ((32, 32, 32, 2, 1), 8), ((6, 5, 7, 2, 2), 1),
                                       ((6, 5, 7, 2, 2), 3), ((6, 5, 7, 2, 2), 100)])
def test_every_flux_position_once_with_its_x1_neighbour_in_the_lane_before(shape, seg):
    nx1, nx2, nx3, ng, nb = shape
    lo = ng - 1
    box = (nx1 + 2, nx2 + 2, nx3 + 2)
    _, nwave = flux_march_position(shape, seg, 0, 0)
    seen = {}
    prev = None
    for wv in range(nwave):
        lanes = [flux_march_position(shape, seg, wv, ln)[0] for ln in range(64)]
        for ln, p in enumerate(lanes):
            inside, g, j, i, k0, k1, hmax = (int(v) for v in p)
            # clamped or not, every lane's cell and planes lie in its block's box
            assert 0 <= g < nb
            assert lo <= i < lo + box[0] and lo <= j < lo + box[1], (wv, ln)
            assert lo <= k0 <= k1 < lo + box[2] and k1 - k0 + 1 <= hmax, (wv, ln)
            if ln == 0 and prev is not None:
                assert tuple(p) == tuple(prev), wv     # repeats the last lane before
            if not inside or ln == 0:
                continue
            for k in range(k0, k1 + 1):
                key = (g, k, j, i)
                assert key not in seen, (key, wv, ln, seen.get(key))
                seen[key] = (wv, ln)
            if i >= ng:                                # the x1 face (k, j, i) exists
                q = lanes[ln - 1]
                assert (int(q[1]), int(q[2]), int(q[3]), int(q[4]), int(q[5])) == \
                    (g, j, i - 1, k0, k1), (wv, ln)
        prev = lanes[63]
    want = {(g, k, j, i) for g in range(nb) for k in range(lo, lo + box[2])
            for j in range(lo, lo + box[1]) for i in range(lo, lo + box[0])}
    assert set(seen) == want


def test_bad_arguments_fail():
    from athena_amd.gpu import AthGpuError
    with pytest.raises(AthGpuError):
        flux_march_position((8, 8, 8, 2, 1), 0, 0, 0)
    _, nw = flux_march_position((8, 8, 8, 2, 1), 8, 0, 0)
    with pytest.raises(AthGpuError):
        flux_march_position((8, 8, 8, 2, 1), 8, nw, 0)
    with pytest.raises(AthGpuError):
        flux_march_position((8, 8, 8, 2, 1), 8, 0, 64)
