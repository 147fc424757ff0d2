"""Shared set-up for the MeshRefinement operator tests: block shapes, coordinate tables
(the packed x1f, x1v, dx1f, x2f, ... layout of the operators), sample fields and the
ctypes binding of the CPU restatement (oracle/refine_oracle.cpp)."""
import ctypes

import numpy as np

import oracle_bind as ob


def shape_of(nx, ng):
    nx1, nx2, nx3 = nx
    return np.array([nx1, nx2, nx3, ng], dtype=np.int32)


def dims(shape):
    nx1, nx2, nx3, ng = (int(v) for v in shape)
    fine = (nx1 + 2 * ng, nx2 + 2 * ng, nx3 + 2 * ng)
    coarse = (nx1 // 2 + 2 * ng, nx2 // 2 + 2 * ng, nx3 // 2 + 2 * ng)
    return fine, coarse


def coords(n, x0, dx, ng, stretch=0.0, rng=None):
    """Packed tables of a grid with n cells per direction (tuple), first active face at
    x0[d] and spacing dx[d]; stretch > 0 perturbs the interior faces."""
    out = []
    for d in range(3):
        xf = x0[d] + (np.arange(n[d] + 1) - ng) * dx[d]
        if stretch > 0.0:
            xf = xf + stretch * dx[d] * rng.uniform(-1.0, 1.0, size=xf.size)
        xv = 0.5 * (xf[1:] + xf[:-1])
        out += [xf, xv, xf[1:] - xf[:-1]]
    return np.concatenate(out)


def tables(c, n):
    """Unpack coords() output into {'x1f': ..., 'x1v': ..., 'dx1f': ...}."""
    t, o = {}, 0
    for d, nd in zip("123", n):
        for name, ln in (("x%sf", nd + 1), ("x%sv", nd), ("dx%sf", nd)):
            t[name % d] = c[o:o + ln]
            o += ln
    return t


def block_coords(shape, x0=(0.0, 0.0, 0.0), h=0.1):
    """Fine and coarse tables of one block whose active region starts at x0 with fine
    spacing h; the coarse buffer covers the same region with spacing 2h."""
    (fn, cn) = dims(shape)
    ng = int(shape[3])
    fc = coords(fn, x0, (h, h, h), ng)
    cc = coords(cn, x0, (2 * h, 2 * h, 2 * h), ng)
    return fc, cc


def lib():
    L = ob.lib()
    P = ctypes.c_void_p
    I = ctypes.c_int
    L.ora_restrict_cc.argtypes = [P, I, P, P, P, P]
    L.ora_restrict_fc.argtypes = [P, I, P, P, P, P]
    L.ora_prolongate_cc.argtypes = [P, I, P, P, P, P, P]
    L.ora_prolongate_fc_shared.argtypes = [P, I, P, P, P, P, P]
    L.ora_prolongate_fc_internal.argtypes = [P, P, P, P, P, P]
    L.ora_restrict_emf.argtypes = [P, I, P, P, P, P]
    return L


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def ora_restrict_cc(shape, fc, fine, coarse, r):
    lib().ora_restrict_cc(_p(shape), fine.shape[0], _p(fc), _p(fine), _p(coarse),
                          _p(np.asarray(r, dtype=np.int32)))


def ora_restrict_fc(shape, comp, fc, fine, coarse, r):
    lib().ora_restrict_fc(_p(shape), comp, _p(fc), _p(fine), _p(coarse),
                          _p(np.asarray(r, dtype=np.int32)))


def ora_prolongate_cc(shape, fc, cc, coarse, fine, r):
    lib().ora_prolongate_cc(_p(shape), coarse.shape[0], _p(fc), _p(cc), _p(coarse),
                            _p(fine), _p(np.asarray(r, dtype=np.int32)))


def ora_prolongate_fc_shared(shape, comp, fc, cc, coarse, fine, r):
    lib().ora_prolongate_fc_shared(_p(shape), comp, _p(fc), _p(cc), _p(coarse), _p(fine),
                                   _p(np.asarray(r, dtype=np.int32)))


def ora_prolongate_fc_internal(shape, fc, b1, b2, b3, r):
    lib().ora_prolongate_fc_internal(_p(shape), _p(fc), _p(b1), _p(b2), _p(b3),
                                     _p(np.asarray(r, dtype=np.int32)))


def ora_restrict_emf(shape, comp, fc, fine, coarse, r):
    lib().ora_restrict_emf(_p(shape), comp, _p(fc), _p(fine), _p(coarse),
                           _p(np.asarray(r, dtype=np.int32)))


def edge_shape(n, comp):
    """(n3, n2, n1) of the edges along comp of a grid with n = (n1, n2, n3) cells."""
    n1, n2, n3 = n
    return (n3 + (comp != 2), n2 + (comp != 1), n1 + (comp != 0))


def face_shape(n, comp):
    """(n3, n2, n1) of face component comp of a grid with n = (n1, n2, n3) cells."""
    n1, n2, n3 = n
    return (n3 + (comp == 2), n2 + (comp == 1), n1 + (comp == 0))


def active_coarse_range(shape, comp=None):
    """The coarse cells (or faces of component comp) over a block's active region."""
    nx1, nx2, nx3, ng = (int(v) for v in shape)
    r = [ng, ng + nx1 // 2 - 1, ng, ng + nx2 // 2 - 1, ng, ng + nx3 // 2 - 1]
    if comp is not None:
        r[2 * comp + 1] += 1
    return r
