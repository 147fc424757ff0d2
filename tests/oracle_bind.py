"""ctypes binding of the CPU restatement in oracle/ (test infrastructure only).

The oracle is the checker: tests feed it the same initial registers as the device
library and compare the results.  It is built by ``make -C oracle``."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from athena_amd.mesh import BOUNDARY, MeshDesc, RSOLVER

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")

WHAT = {"u": 0, "w": 1, "bcc": 2, "b1": 3, "b2": 4, "b3": 5, "w1": 6,
        "x1f": 7, "x2f": 8, "x3f": 9, "x1v": 10, "x2v": 11, "x3v": 12,
        "dx1f": 13, "dx2f": 14, "dx3f": 15, "flux1": 16, "flux2": 17, "flux3": 18,
        "e1": 19, "e2": 20, "e3": 21, "e3_x1f": 22, "e2_x1f": 23, "e1_x2f": 24,
        "e3_x2f": 25, "e2_x3f": 26, "e1_x3f": 27, "wght1": 28, "wght2": 29, "wght3": 30}


class OraCfg(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int * 3), ("bnx", ctypes.c_int * 3), ("ng", ctypes.c_int),
                ("xorder", ctypes.c_int), ("rsolver", ctypes.c_int),
                ("relativistic", ctypes.c_int),
                ("xmin", ctypes.c_double * 3), ("xmax", ctypes.c_double * 3),
                ("gamma", ctypes.c_double), ("dfloor", ctypes.c_double),
                ("pfloor", ctypes.c_double), ("cfl", ctypes.c_double),
                ("sigma_max", ctypes.c_double), ("beta_min", ctypes.c_double),
                ("gamma_max", ctypes.c_double), ("bc", ctypes.c_int * 6)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        D = ctypes.POINTER(ctypes.c_double)
        L.ora_new.restype = P
        L.ora_restart_read.argtypes = [ctypes.c_char_p, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), D,
                                       ctypes.POINTER(ctypes.c_longlong), ctypes.c_int, D,
                                       ctypes.c_long]
        L.ora_restart_error.restype = ctypes.c_char_p
        L.ora_new.argtypes = [ctypes.POINTER(OraCfg)]
        L.ora_free.argtypes = [P]
        L.ora_nblocks.argtypes = [P]
        L.ora_size.restype = ctypes.c_long
        L.ora_size.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.ora_set.argtypes = [P, ctypes.c_int, ctypes.c_int, D]
        L.ora_get.argtypes = [P, ctypes.c_int, ctypes.c_int, D]
        L.ora_initialize.argtypes = [P, ctypes.c_double]
        L.ora_run.argtypes = [P, ctypes.c_int, ctypes.c_double]
        L.ora_run.restype = ctypes.c_int
        L.ora_state.argtypes = [P, D, D, ctypes.POINTER(ctypes.c_int)]
        L.ora_set_dt.argtypes = [P, ctypes.c_double]
        L.ora_set_clock.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.ora_threads.restype = ctypes.c_int
        L.ora_block_call.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, ctypes.POINTER(ctypes.c_int), D]
        L.ora_block_riemann.argtypes = [P] + [ctypes.c_int] * 6 + [D, D]
        L.ora_riemann.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  D, D, D, D]
        L.ora_sr_con2prim.argtypes = [ctypes.c_int] + [ctypes.c_double] * 6 + \
            [D, D, D, ctypes.POINTER(ctypes.c_int)]
        L.ora_sr_prim2cons.argtypes = [ctypes.c_int, ctypes.c_double, D, D, D]
        L.ora_mlp_nw.restype = ctypes.c_int
        L.ora_mlp_forward.argtypes = [D, ctypes.c_int, D, D]
        L.ora_mlp_grad.argtypes = [D, ctypes.c_int, D, D, D, D]
        L.ora_mlp_features.argtypes = [ctypes.c_int, D, D, D]
        L.ora_smr_new.restype = P
        L.ora_smr_new.argtypes = [ctypes.POINTER(OraCfg), ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
        L.ora_smr_error.restype = ctypes.c_char_p
        L.ora_smr_free.argtypes = [P]
        L.ora_smr_initialize.argtypes = [P, ctypes.c_double]
        L.ora_smr_run.argtypes = [P, ctypes.c_int, ctypes.c_double]
        L.ora_smr_neighbors.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                        ctypes.c_int]
        L.ora_smr_coords.restype = ctypes.c_long
        L.ora_smr_skip.argtypes = [P, ctypes.c_int]
        L.ora_smr_coords.argtypes = [P, ctypes.c_int, ctypes.c_int, D, ctypes.c_long]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def cfg_of(m: MeshDesc) -> OraCfg:
    c = OraCfg()
    c.nx[:] = list(m.nx)
    c.bnx[:] = list(m.bnx)
    c.ng = m.ng
    c.xorder = m.xorder
    c.rsolver = RSOLVER[m.rsolver]
    c.relativistic = int(m.relativistic)
    c.xmin[:] = list(m.xmin)
    c.xmax[:] = list(m.xmax)
    c.gamma = m.gamma
    c.dfloor, c.pfloor = m.dfloor, m.pfloor
    c.cfl = m.cfl
    c.sigma_max, c.beta_min, c.gamma_max = m.sigma_max, m.beta_min, m.gamma_max
    c.bc[:] = [BOUNDARY[f] for f in m.bc]
    return c


class OracleMesh:
    """CPU restatement of one uniform mesh (periodic or outflow faces)."""

    def __init__(self, m: MeshDesc):
        self.m = m
        self._cfg = cfg_of(m)
        self.h = lib().ora_new(ctypes.byref(self._cfg))
        if not self.h:
            raise ValueError("the oracle refused this mesh description")

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_free(self.h)
            self.h = None

    def set(self, gid, what, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        n = lib().ora_size(self.h, gid, WHAT[what])
        assert a.size == n, (what, a.size, n)
        assert lib().ora_set(self.h, gid, WHAT[what], _dp(a)) == 0

    def get(self, gid, what):
        n = lib().ora_size(self.h, gid, WHAT[what])
        a = np.empty(n)
        assert lib().ora_get(self.h, gid, WHAT[what], _dp(a)) == 0
        if what in self.m.shape.__code__.co_consts or what in ("u", "w", "w1", "bcc", "b1", "b2", "b3"):
            return a.reshape(self.m.shape(what))
        return a

    def load(self, blocks):
        for gid, b in enumerate(blocks):
            for k in ("u", "b1", "b2", "b3", "w", "w1"):
                if k in b:
                    self.set(gid, k, b[k])

    def initialize(self, tlim=1e300):
        lib().ora_initialize(self.h, tlim)

    def run(self, ncycles, tlim=1e300):
        return lib().ora_run(self.h, ncycles, tlim)

    def set_clock(self, time, dt, ncycle):
        lib().ora_set_clock(self.h, float(time), float(dt), int(ncycle))

    def state(self):
        t, dt, nc = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        lib().ora_state(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(nc))
        return t.value, dt.value, nc.value

    def block(self, gid):
        return {k: self.get(gid, k) for k in ("u", "w", "bcc", "b1", "b2", "b3")}

    def call(self, gid, fn, order=0, wght=0.0, rng=(0, 0, 0, 0, 0, 0), out_size=1):
        """One per-MeshBlock function on block gid (ora_block_call's fn codes)."""
        r = np.ascontiguousarray(rng, dtype=np.int32)
        out = np.zeros(out_size)
        rc = lib().ora_block_call(self.h, gid, fn, order, wght,
                                  r.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(out))
        assert rc == 0
        return out

    def riemann_pencil(self, gid, d, k, j, il, iu, wl, wr):
        wl = np.ascontiguousarray(wl, dtype=np.float64)
        wr = np.ascontiguousarray(wr, dtype=np.float64)
        lib().ora_block_riemann(self.h, gid, d, k, j, il, iu, _dp(wl), _dp(wr))


class OracleSmrMesh(OracleMesh):
    """CPU restatement of a statically refined mesh (``athena_amd.mesh.refined``):
    periodic faces, Newtonian MHD."""

    def __init__(self, m: MeshDesc):
        assert m.smr, "OracleSmrMesh needs a refined MeshDesc"
        self.m = m
        self._cfg = cfg_of(m)
        self._locs = np.ascontiguousarray(m.locs, dtype=np.int64)
        self.h = lib().ora_smr_new(ctypes.byref(self._cfg), len(self._locs),
                                   self._locs.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                   int(m.root_level))
        if not self.h:
            raise ValueError(lib().ora_smr_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_smr_free(self.h)
            self.h = None

    def initialize(self, tlim=1e300):
        if lib().ora_smr_initialize(self.h, tlim) != 0:
            raise RuntimeError(lib().ora_smr_error().decode())

    def run(self, ncycles, tlim=1e300):
        n = lib().ora_smr_run(self.h, ncycles, tlim)
        if n < 0:
            raise RuntimeError(lib().ora_smr_error().decode())
        return n

    def neighbors(self, gid):
        """SearchAndSetNeighbors order: rows (gid, level, ox1, ox2, ox3, type, fi1, fi2,
        eid or fid), type 1 face, 2 edge, 3 corner."""
        out = np.zeros(9 * 56, dtype=np.int32)
        n = lib().ora_smr_neighbors(self.h, gid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                    56)
        return out[:9 * n].reshape(n, 9)

    def skip(self, mask):
        """Planted omissions (test knob): 1 the flux correction, 2 the EMF correction."""
        lib().ora_smr_skip(self.h, mask)

    def coords(self, gid, coarse=False):
        n = lib().ora_smr_coords(self.h, gid, int(coarse), None, 0)
        a = np.empty(n)
        lib().ora_smr_coords(self.h, gid, int(coarse), _dp(a), n)
        return a


def riemann(solver, relativistic, gamma, wl, wr, bx):
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    bx = np.ascontiguousarray(bx, dtype=np.float64)
    out = np.empty_like(wl)
    lib().ora_riemann(solver, int(relativistic), gamma, wl.shape[0], _dp(wl), _dp(wr),
                      _dp(bx), _dp(out))
    return out


def sr_con2prim(cons, prim_seed, bcc, gamma, dfloor, pfloor, sigma_max=0.0, beta_min=0.0,
                gamma_max=1000.0):
    cons = np.array(cons, dtype=np.float64, order="C")
    prim = np.array(prim_seed, dtype=np.float64, order="C")
    bcc = np.ascontiguousarray(bcc, dtype=np.float64)
    iters = np.empty(cons.shape[0], dtype=np.int32)
    lib().ora_sr_con2prim(cons.shape[0], gamma, dfloor, pfloor, sigma_max, beta_min,
                          gamma_max, _dp(cons), _dp(prim), _dp(bcc),
                          iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return cons, prim, iters


def sr_prim2cons(prim, bcc, gamma):
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    bcc = np.ascontiguousarray(bcc, dtype=np.float64)
    cons = np.empty_like(prim)
    lib().ora_sr_prim2cons(prim.shape[0], gamma, _dp(prim), _dp(bcc), _dp(cons))
    return cons


def mlp_forward(params, x):
    """De-normalised network outputs for normalised inputs x (n,6), in double."""
    P = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(x.shape[0])
    lib().ora_mlp_forward(_dp(P), x.shape[0], _dp(x), _dp(y))
    return y


def mlp_grad(params, x, t):
    """(gradient (4673,), loss) of the mean squared error in normalised units."""
    P = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    g = np.empty(lib().ora_mlp_nw())
    loss = np.empty(1)
    lib().ora_mlp_grad(_dp(P), t.size, _dp(x), _dp(t), _dp(g), _dp(loss))
    return g, float(loss[0])


def mlp_features(cons, bcc):
    cons = np.ascontiguousarray(cons, dtype=np.float64)
    bcc = np.ascontiguousarray(bcc, dtype=np.float64)
    x = np.empty((cons.shape[0], 6))
    lib().ora_mlp_features(cons.shape[0], _dp(cons), _dp(bcc), _dp(x))
    return x


def restart_read(path, ng, nblocks_cap, doubles_cap):
    """The restatement of the reference's restart reader (oracle/restart_oracle.cpp):
    (ints, reals, locations (n, 4) as level, lx1, lx2, lx3 in the order the data is
    assigned, data (doubles of every block's u, b.x1f, b.x2f, b.x3f))."""
    L = lib()
    ints = (ctypes.c_int * 11)()
    reals = np.zeros(11)
    locs = np.zeros(4 * nblocks_cap, dtype=np.int64)
    data = np.zeros(doubles_cap)
    rc = L.ora_restart_read(path.encode(), ng, ints,
                            reals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            locs.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                            nblocks_cap, data.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            doubles_cap)
    if rc != 0:
        raise RuntimeError(L.ora_restart_error().decode())
    return list(ints), reals, locs.reshape(-1, 4), data
