"""CPU pins of the CPU restatement of the MeshRefinement operators
(oracle/refine_oracle.cpp, mesh_refinement.cpp) by what the mathematics fixes:
restriction is exact on constant and linear fields (cells: volume averages; faces: area
averages); prolongation reproduces linear fields (minmod of equal one-sided slopes is
the slope); restriction after prolongation is the identity for cells and shared faces
(the offsets of the eight / four fine values cancel); the divergence-preserving internal
prolongation leaves every fine cell divergence-free when the coarse field is the
discrete curl of an edge potential, and reproduces linear divergence-free fields."""
import numpy as np
import pytest

from refine_common import (active_coarse_range, block_coords, dims, edge_shape,
                           face_shape, ora_prolongate_cc, ora_prolongate_fc_internal,
                           ora_prolongate_fc_shared, ora_restrict_cc, ora_restrict_emf,
                           ora_restrict_fc, shape_of, tables)

SHAPES = [((8, 8, 8), 2), ((8, 4, 6), 3), ((4, 6, 8), 4)]
COEF = np.array([0.7, 1.3, -0.4, 2.1])            # f = c0 + c1 x + c2 y + c3 z


def lin(x, y, z, c=COEF):
    return c[0] + c[1] * x + c[2] * y + c[3] * z


def cell_field(t, n, nvar=1, c=COEF):
    z, y, x = np.meshgrid(t["x3v"], t["x2v"], t["x1v"], indexing="ij")
    return np.array([lin(x, y, z, c) + v for v in range(nvar)])


def face_field(t, comp, c=COEF):
    xs = [t["x1v"], t["x2v"], t["x3v"]]
    xs[comp] = t["x%df" % (comp + 1)]
    z, y, x = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    return lin(x, y, z, c)


def sel(a, r):
    return a[..., r[4]:r[5] + 1, r[2]:r[3] + 1, r[0]:r[1] + 1]


@pytest.mark.parametrize("nx,ng", SHAPES)
def test_cell_restriction_exact_on_linear_fields(nx, ng):
    s = shape_of(nx, ng)
    fn, cn = dims(s)
    fc, cc = block_coords(s)
    ft, ct = tables(fc, fn), tables(cc, cn)
    fine = cell_field(ft, fn, nvar=3)
    coarse = np.zeros((3, cn[2], cn[1], cn[0]))
    r = active_coarse_range(s)
    ora_restrict_cc(s, fc, fine, coarse, r)
    want = cell_field(ct, cn, nvar=3)
    np.testing.assert_allclose(sel(coarse, r), sel(want, r), rtol=1e-14, atol=1e-14)
    const = np.full_like(fine, 3.25)
    ora_restrict_cc(s, fc, const, coarse, r)
    np.testing.assert_allclose(sel(coarse, r), 3.25, rtol=1e-15)


@pytest.mark.parametrize("nx,ng", SHAPES)
def test_cell_prolongation_reproduces_linear_fields(nx, ng):
    s = shape_of(nx, ng)
    fn, cn = dims(s)
    fc, cc = block_coords(s)
    ft, ct = tables(fc, fn), tables(cc, cn)
    coarse = cell_field(ct, cn, nvar=2)
    fine = np.zeros((2, fn[2], fn[1], fn[0]))
    r = active_coarse_range(s)
    ora_prolongate_cc(s, fc, cc, coarse, fine, r)
    want = cell_field(ft, fn, nvar=2)
    a = [ng, ng + nx[0] - 1, ng, ng + nx[1] - 1, ng, ng + nx[2] - 1]
    np.testing.assert_allclose(sel(fine, a), sel(want, a), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("nx,ng", SHAPES)
def test_cell_restriction_undoes_prolongation(nx, ng):
    rng = np.random.default_rng(7)
    s = shape_of(nx, ng)
    fn, cn = dims(s)
    fc, cc = block_coords(s)
    coarse = rng.uniform(0.5, 2.0, size=(2, cn[2], cn[1], cn[0]))
    fine = np.zeros((2, fn[2], fn[1], fn[0]))
    r = active_coarse_range(s)
    ora_prolongate_cc(s, fc, cc, coarse, fine, r)
    back = np.zeros_like(coarse)
    ora_restrict_cc(s, fc, fine, back, r)
    np.testing.assert_allclose(sel(back, r), sel(coarse, r), rtol=2e-15)


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("nx,ng", SHAPES)
def test_face_restriction_and_shared_prolongation(nx, ng, comp):
    rng = np.random.default_rng(11 + comp)
    s = shape_of(nx, ng)
    fn, cn = dims(s)
    fc, cc = block_coords(s)
    ft, ct = tables(fc, fn), tables(cc, cn)
    r = active_coarse_range(s, comp)
    # restriction of the area averages of a linear field
    coarse = np.zeros(face_shape(cn, comp))
    ora_restrict_fc(s, comp, fc, np.ascontiguousarray(face_field(ft, comp)), coarse, r)
    np.testing.assert_allclose(sel(coarse, r), sel(face_field(ct, comp), r), rtol=1e-14,
                               atol=1e-14)
    # shared-face prolongation reproduces the linear field on the fine faces
    fine = np.zeros(face_shape(fn, comp))
    ora_prolongate_fc_shared(s, comp, fc, cc, np.ascontiguousarray(face_field(ct, comp)),
                             fine, r)
    a = [ng, ng + nx[0] - 1, ng, ng + nx[1] - 1, ng, ng + nx[2] - 1]
    a[2 * comp + 1] += 1
    want = face_field(ft, comp)
    # only fine faces that lie on coarse faces are written: every other one along comp
    stride = [slice(None)] * 3
    stride[2 - comp] = slice(0, None, 2)
    np.testing.assert_allclose(sel(fine, a)[tuple(stride)], sel(want, a)[tuple(stride)],
                               rtol=1e-14, atol=1e-14)
    # and restriction undoes it
    crand = rng.uniform(-1.0, 1.0, size=face_shape(cn, comp))
    fine[:] = 0.0
    ora_prolongate_fc_shared(s, comp, fc, cc, crand, fine, r)
    back = np.zeros_like(crand)
    ora_restrict_fc(s, comp, fc, fine, back, r)
    np.testing.assert_allclose(sel(back, r), sel(crand, r), rtol=1e-14, atol=1e-15)


def _curl_field(ct, cn, rng):
    """Face fields of the discrete curl of a random edge potential on the coarse grid:
    div B = 0 in every coarse cell to round-off."""
    n1, n2, n3 = cn
    a1 = rng.normal(size=(n3 + 1, n2 + 1, n1))
    a2 = rng.normal(size=(n3 + 1, n2, n1 + 1))
    a3 = rng.normal(size=(n3, n2 + 1, n1 + 1))
    dx1, dx2, dx3 = ct["dx1f"], ct["dx2f"], ct["dx3f"]
    b1 = ((a3[:, 1:, :] - a3[:, :-1, :]) / dx2[None, :, None]
          - (a2[1:, :, :] - a2[:-1, :, :]) / dx3[:, None, None])
    b2 = ((a1[1:, :, :] - a1[:-1, :, :]) / dx3[:, None, None]
          - (a3[:, :, 1:] - a3[:, :, :-1]) / dx1[None, None, :])
    b3 = ((a2[:, :, 1:] - a2[:, :, :-1]) / dx1[None, None, :]
          - (a1[:, 1:, :] - a1[:, :-1, :]) / dx2[None, :, None])
    return [np.ascontiguousarray(b) for b in (b1, b2, b3)]


def _prolongate_all(s, fc, cc, cb):
    fn, cn = dims(s)
    fb = [np.zeros(face_shape(fn, c)) for c in range(3)]
    for comp in range(3):
        ora_prolongate_fc_shared(s, comp, fc, cc, cb[comp], fb[comp],
                                 active_coarse_range(s, comp))
    ora_prolongate_fc_internal(s, fc, fb[0], fb[1], fb[2], active_coarse_range(s))
    return fb


def _fine_div(s, ft, fb):
    nx1, nx2, nx3, ng = (int(v) for v in s)
    i, j, k = (slice(ng, ng + n) for n in (nx1, nx2, nx3))
    ip, jp, kp = (slice(ng + 1, ng + n + 1) for n in (nx1, nx2, nx3))
    dx1, dx2, dx3 = ft["dx1f"][i], ft["dx2f"][j], ft["dx3f"][k]
    return ((fb[0][k, j, ip] - fb[0][k, j, i]) / dx1[None, None, :]
            + (fb[1][k, jp, i] - fb[1][k, j, i]) / dx2[None, :, None]
            + (fb[2][kp, j, i] - fb[2][k, j, i]) / dx3[:, None, None])


@pytest.mark.parametrize("nx,ng", SHAPES)
def test_internal_prolongation_keeps_the_field_divergence_free(nx, ng):
    rng = np.random.default_rng(23)
    s = shape_of(nx, ng)
    fn, cn = dims(s)
    fc, cc = block_coords(s)
    ft, ct = tables(fc, fn), tables(cc, cn)
    cb = _curl_field(ct, cn, rng)
    fb = _prolongate_all(s, fc, cc, cb)
    div = _fine_div(s, ft, fb)
    scale = max(np.abs(b).max() for b in cb) / ft["dx1f"][ng]
    print("  max |div B| / (|B|/dx) = %.2e" % (np.abs(div).max() / scale))
    assert np.abs(div).max() < 1e-12 * scale
    # without the internal step the fine cells are not divergence-free
    fb0 = [np.zeros(face_shape(fn, c)) for c in range(3)]
    for comp in range(3):
        ora_prolongate_fc_shared(s, comp, fc, cc, cb[comp], fb0[comp],
                                 active_coarse_range(s, comp))
    assert np.abs(_fine_div(s, ft, fb0)).max() > 1e-3 * scale


@pytest.mark.parametrize("nx,ng", SHAPES)
def test_internal_prolongation_reproduces_linear_divergence_free_fields(nx, ng):
    # B = (x2 + 2 x3, 3 x3 - x1, 0.5 x1 + x2): every component is linear in the two
    # transverse coordinates only, so div B = 0 and face averages are centre values
    s = shape_of(nx, ng)
    fn, cn = dims(s)
    fc, cc = block_coords(s)
    ft, ct = tables(fc, fn), tables(cc, cn)
    cs = [np.array([0.0, 0.0, 1.0, 2.0]), np.array([0.0, -1.0, 0.0, 3.0]),
          np.array([0.0, 0.5, 1.0, 0.0])]
    cb = [np.ascontiguousarray(face_field(ct, c, cs[c])) for c in range(3)]
    fb = _prolongate_all(s, fc, cc, cb)
    nx1, nx2, nx3, _ = (int(v) for v in s)
    for comp in range(3):
        want = face_field(ft, comp, cs[comp])
        a = [ng, ng + nx1 - 1, ng, ng + nx2 - 1, ng, ng + nx3 - 1]
        a[2 * comp + 1] += 1
        np.testing.assert_allclose(sel(fb[comp], a), sel(want, a), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("nx,ng", SHAPES)
def test_emf_restriction_exact_on_linear_edge_fields(nx, ng, comp):
    # the length-weighted average of the two fine edges of a coarse edge is the line
    # average over the coarse edge: exact for a field linear along the edge (and any
    # transverse dependence, since both fine edges sit on the same line)
    s = shape_of(nx, ng)
    fn, cn = dims(s)
    fc, cc = block_coords(s)
    ft, ct = tables(fc, fn), tables(cc, cn)

    def edge_field(t):
        xs = [t["x1f"], t["x2f"], t["x3f"]]
        xs[comp] = t["x%dv" % (comp + 1)]
        z, y, x = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
        transverse = (y * z, x * z, x * y)[comp]      # both fine edges share these
        return lin(x, y, z) + transverse

    fine = np.ascontiguousarray(edge_field(ft))
    coarse = np.zeros(edge_shape(cn, comp))
    r = active_coarse_range(s)
    for d in range(3):
        if d != comp:
            r[2 * d + 1] += 1
    ora_restrict_emf(s, comp, fc, fine, coarse, r)
    np.testing.assert_allclose(sel(coarse, r), sel(edge_field(ct), r), rtol=1e-14, atol=1e-14)
