"""CPU checks of the fused flux + flux-divergence kernel's plan (k_flux_div), through the
library's host enumeration, which runs the kernel's own index helpers (fdiv_band,
fdiv_thread) over every band and plane:

* every band fits one work-group, and every flux an owned cell reads from LDS is the
  face it needs, computed by the neighbouring thread in the same plane (the library
  fails otherwise);
* the faces whose EMFs and CT weights are written are exactly the reference's MHD flux
  ranges (calculate_fluxes.cpp:68-76, 176-182, 299-302), each once;
* the cells updated are exactly the active cells, each once;
* one work-group's LDS fits three work-groups on a 160 KiB compute unit, and bands that
  would need more threads than a work-group has are refused.
A planted mistake in the range logic (a band that skips the extra x2 row, say) fails the
coverage check."""
import itertools

import numpy as np
import pytest

from athena_amd.gpu import AthGpuError, flux_div_plan

SHAPES = [((32, 32, 32), 2, 5), ((32, 32, 32), 2, 4), ((32, 32, 32), 2, 1),
          ((16, 8, 4), 3, 8), ((16, 8, 4), 3, 3), ((64, 16, 8), 4, 1), ((2, 2, 2), 2, 2),
          ((3, 3, 3), 3, 1), ((8, 16, 4), 2, 16), ((5, 7, 3), 2, 4), ((80, 4, 4), 2, 1),
          ((32, 30, 8), 2, 5)]


def reference_faces(nx, ng):
    n1, n2, n3 = nx
    s = ng
    ie, je, ke = s + n1 - 1, s + n2 - 1, s + n3 - 1
    want = set()
    for k in range(s - 1, ke + 2):            # x1: k in [ks-1, ke+1], j in [js-1, je+1]
        for j in range(s - 1, je + 2):
            for i in range(s, ie + 2):        #     i in [is, ie+1]
                want.add((0, k, j, i))
    for k in range(s - 1, ke + 2):            # x2: j in [js, je+1], i in [is-1, ie+1]
        for j in range(s, je + 2):
            for i in range(s - 1, ie + 2):
                want.add((1, k, j, i))
    for k in range(s, ke + 2):                # x3: k in [ks, ke+1]
        for j in range(s - 1, je + 2):
            for i in range(s - 1, ie + 2):
                want.add((2, k, j, i))
    return want


@pytest.mark.parametrize("nx,ng,rows", SHAPES)
def test_plan_covers_reference_ranges_once(nx, ng, rows):
    faces, cells, lds = flux_div_plan(nx, ng, rows)
    got = [tuple(int(v) for v in f) for f in faces]
    want = reference_faces(nx, ng)
    assert len(got) == len(set(got)), "a face's EMFs are written twice"
    assert set(got) == want
    cl = [tuple(int(v) for v in c) for c in cells]
    assert len(cl) == len(set(cl)), "a cell is updated twice"
    act = set(itertools.product(range(ng, ng + nx[2]), range(ng, ng + nx[1]),
                                range(ng, ng + nx[0])))
    assert set(cl) == act
    assert lds * 8 <= 64 * 1024


def test_default_band_fits_three_work_groups_per_cu():
    _, _, lds = flux_div_plan((32, 32, 32), 2, 5)
    assert 3 * lds * 8 <= 160 * 1024


@pytest.mark.parametrize("nx,ng,rows", [((32, 32, 32), 2, 6), ((32, 32, 32), 2, 33),
                                        ((300, 4, 4), 2, 1), ((32, 32, 32), 2, 0),
                                        ((128, 4, 4), 2, 1)])
def test_plan_refuses_bands_beyond_one_cell_per_thread(nx, ng, rows):
    with pytest.raises(AthGpuError):
        flux_div_plan(nx, ng, rows)
