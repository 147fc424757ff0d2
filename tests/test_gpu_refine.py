"""Device MeshRefinement operators (athena_amd/csrc/refine.hip) against the CPU
restatement (oracle/refine_oracle.cpp), bit for bit: several blocks per launch, each with
its own stretched coordinate tables, random data, ranges that include ghost zones, and
one call per operator and component."""
import numpy as np
import pytest
import torch

from athena_amd.gpu import Refinement
from refine_common import (coords, dims, edge_shape, face_shape, ora_prolongate_cc,
                           ora_prolongate_fc_internal, ora_prolongate_fc_shared,
                           ora_restrict_cc, ora_restrict_emf, ora_restrict_fc, shape_of)

pytestmark = pytest.mark.gpu

SHAPES = [((8, 8, 8), 2), ((8, 4, 6), 3), ((4, 6, 8), 4)]
NB = 3


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _setup(nx, ng, seed):
    rng = np.random.default_rng(seed)
    s = shape_of(nx, ng)
    fn, cn = dims(s)
    fcs, ccs = [], []
    for b in range(NB):
        x0 = tuple(rng.uniform(-1.0, 1.0, 3))
        h = rng.uniform(0.05, 0.2)
        fcs.append(coords(fn, x0, (h, 1.1 * h, 0.9 * h), ng, stretch=0.2, rng=rng))
        ccs.append(coords(cn, x0, (2 * h, 2.2 * h, 1.8 * h), ng, stretch=0.2, rng=rng))
    return rng, s, fn, cn, fcs, ccs


def _ranges(s, comp=None):
    """Coarse ranges the reference uses (its active region) and the widest the
    operators accept (into the ghost zones)."""
    nx1, nx2, nx3, ng = (int(v) for v in s)
    act = [ng, ng + nx1 // 2 - 1, ng, ng + nx2 // 2 - 1, ng, ng + nx3 // 2 - 1]
    if comp is not None:
        act[2 * comp + 1] += 1
    return act


@pytest.mark.parametrize("nx,ng", SHAPES)
def test_restrict_and_prolongate_cells(nx, ng):
    print("case refine cells", nx, ng, flush=True)
    rng, s, fn, cn, fcs, ccs = _setup(nx, ng, 3)
    R = Refinement(nx, ng)
    nvar = 5
    fine = rng.normal(size=(NB, nvar, fn[2], fn[1], fn[0]))
    coarse0 = rng.normal(size=(NB, nvar, cn[2], cn[1], cn[0]))
    # restriction over the widest range whose fine cells exist: coarse 1 .. ncc-2
    r = [ng - ng // 2, ng + nx[0] // 2 - 1 + ng // 2, ng - ng // 2,
         ng + nx[1] // 2 - 1 + ng // 2, ng - ng // 2, ng + nx[2] // 2 - 1 + ng // 2]
    want = coarse0.copy()
    for b in range(NB):
        ora_restrict_cc(s, fcs[b], np.ascontiguousarray(fine[b]), want[b], r)
    fc_d, cc_d = _dev(np.concatenate(fcs)), _dev(np.concatenate(ccs))
    f_d, c_d = _dev(fine), _dev(coarse0)
    R.restrict_cc(NB, nvar, fc_d.data_ptr(), f_d.data_ptr(), c_d.data_ptr(), r)
    torch.cuda.synchronize()
    assert np.array_equal(c_d.cpu().numpy(), want)
    # prolongation over the active coarse cells
    r = _ranges(s)
    wantf = fine.copy()
    for b in range(NB):
        ora_prolongate_cc(s, fcs[b], ccs[b], np.ascontiguousarray(coarse0[b]), wantf[b], r)
    f_d = _dev(fine)
    c_d = _dev(coarse0)
    R.prolongate_cc(NB, nvar, fc_d.data_ptr(), cc_d.data_ptr(), c_d.data_ptr(),
                    f_d.data_ptr(), r)
    torch.cuda.synchronize()
    got = f_d.cpu().numpy()
    assert np.array_equal(got, wantf)
    assert not np.array_equal(got, fine)          # something was written


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("nx,ng", SHAPES)
def test_restrict_and_prolongate_faces(nx, ng, comp):
    print("case refine faces", nx, ng, comp, flush=True)
    rng, s, fn, cn, fcs, ccs = _setup(nx, ng, 5 + comp)
    R = Refinement(nx, ng)
    fshape, cshape = face_shape(fn, comp), face_shape(cn, comp)
    fine = rng.normal(size=(NB,) + fshape)
    coarse0 = rng.normal(size=(NB,) + cshape)
    fc_d, cc_d = _dev(np.concatenate(fcs)), _dev(np.concatenate(ccs))
    r = _ranges(s, comp)
    want = coarse0.copy()
    for b in range(NB):
        ora_restrict_fc(s, comp, fcs[b], np.ascontiguousarray(fine[b]), want[b], r)
    f_d, c_d = _dev(fine), _dev(coarse0)
    R.restrict_fc(NB, comp, fc_d.data_ptr(), f_d.data_ptr(), c_d.data_ptr(), r)
    torch.cuda.synchronize()
    assert np.array_equal(c_d.cpu().numpy(), want)
    wantf = fine.copy()
    for b in range(NB):
        ora_prolongate_fc_shared(s, comp, fcs[b], ccs[b], np.ascontiguousarray(coarse0[b]),
                                 wantf[b], r)
    f_d, c_d = _dev(fine), _dev(coarse0)
    R.prolongate_fc_shared(NB, comp, fc_d.data_ptr(), cc_d.data_ptr(), c_d.data_ptr(),
                           f_d.data_ptr(), r)
    torch.cuda.synchronize()
    assert np.array_equal(f_d.cpu().numpy(), wantf)


@pytest.mark.parametrize("nx,ng", SHAPES)
def test_internal_face_prolongation(nx, ng):
    print("case refine internal faces", nx, ng, flush=True)
    rng, s, fn, cn, fcs, ccs = _setup(nx, ng, 9)
    R = Refinement(nx, ng)
    b = [rng.normal(size=(NB,) + face_shape(fn, c)) for c in range(3)]
    r = _ranges(s)
    want = [x.copy() for x in b]
    for q in range(NB):
        ora_prolongate_fc_internal(s, fcs[q], want[0][q], want[1][q], want[2][q], r)
    fc_d = _dev(np.concatenate(fcs))
    d = [_dev(x) for x in b]
    R.prolongate_fc_internal(NB, fc_d.data_ptr(), d[0].data_ptr(), d[1].data_ptr(),
                             d[2].data_ptr(), r)
    torch.cuda.synchronize()
    for c in range(3):
        assert np.array_equal(d[c].cpu().numpy(), want[c]), c


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("nx,ng", SHAPES)
def test_emf_restriction(nx, ng, comp):
    print("case refine emf", nx, ng, comp, flush=True)
    rng, s, fn, cn, fcs, ccs = _setup(nx, ng, 13 + comp)
    R = Refinement(nx, ng)
    fine = rng.normal(size=(NB,) + edge_shape(fn, comp))
    coarse0 = rng.normal(size=(NB,) + edge_shape(cn, comp))
    r = _ranges(s)
    for d in range(3):
        if d != comp:
            r[2 * d + 1] += 1
    want = coarse0.copy()
    for b in range(NB):
        ora_restrict_emf(s, comp, fcs[b], np.ascontiguousarray(fine[b]), want[b], r)
    fc_d = _dev(np.concatenate(fcs))
    f_d, c_d = _dev(fine), _dev(coarse0)
    R.restrict_emf(NB, comp, fc_d.data_ptr(), f_d.data_ptr(), c_d.data_ptr(), r)
    torch.cuda.synchronize()
    assert np.array_equal(c_d.cpu().numpy(), want)
