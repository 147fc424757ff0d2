"""Python mirror of the C ABI in include/athena_gpu.h (ctypes).

This is the host-side handle the problem generators, tests and bench use.  It loads
the in-tree ``athena_amd/libathena_gpu.so`` and raises if it is missing or if a call
fails: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .mesh import BOUNDARY, MeshDesc, RSOLVER

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libathena_gpu.so")

ATH = {"u": 0, "w": 1, "bcc": 2, "b1": 3, "b2": 4, "b3": 5, "w1": 6,
       "flux1": 20, "flux2": 21, "flux3": 22, "e1": 23, "e2": 24, "e3": 25,
       "wght1": 26, "wght2": 27, "wght3": 28}

EXPORTS = ["ath_gpu_mesh_create", "ath_gpu_mesh_destroy", "ath_gpu_nblocks",
           "ath_gpu_block_size", "ath_gpu_set_block", "ath_gpu_get_block",
           "ath_gpu_initialize", "ath_gpu_vl2_stage", "ath_gpu_cycle",
           "ath_gpu_get_clock", "ath_gpu_set_clock", "ath_gpu_calculate_fluxes",
           "ath_gpu_compute_corner_e", "ath_gpu_stream", "ath_gpu_synchronize",
           "ath_gpu_riemann_batch", "ath_gpu_sr_con2prim_batch", "ath_gpu_last_error",
           "ath_gpu_kernel_timing", "ath_gpu_kernel_stats",
           "ath_ml_param_count", "ath_ml_create", "ath_ml_destroy", "ath_ml_set_params",
           "ath_ml_get_params", "ath_ml_grad", "ath_ml_train", "ath_ml_infer",
           "ath_ml_features", "ath_ml_last_error", "ath_gpu_sr_con2prim_modes",
           "ath_gpu_sr_con2prim_device", "ath_gpu_mesh_set_c2p", "ath_gpu_mesh_c2p_stats",
           "ath_gpu_ensemble_create", "ath_gpu_ensemble_destroy", "ath_gpu_ensemble_size",
           "ath_gpu_ensemble_cycle", "ath_gpu_ensemble_get_tiles",
           "ath_gpu_ensemble_set_tiles", "ath_gpu_ensemble_autotune",
           "ath_gpu_ensemble_stream", "ath_gpu_mesh_set_tiles", "ath_gpu_mesh_get_tiles",
           "ath_gpu_mesh_autotune", "ath_gpu_kernel_box", "ath_gpu_history",
           "ath_gpu_ghost_position", "ath_gpu_restrict_cc", "ath_gpu_restrict_fc",
           "ath_gpu_prolongate_cc", "ath_gpu_prolongate_fc_shared",
           "ath_gpu_prolongate_fc_internal", "ath_gpu_restrict_emf", "ath_gpu_tree_build",
           "ath_gpu_block_create", "ath_gpu_block_destroy", "ath_gpu_block_reconstruct",
           "ath_gpu_block_riemann", "ath_gpu_block_calculate_fluxes",
           "ath_gpu_block_compute_corner_e", "ath_gpu_block_ct",
           "ath_gpu_block_add_flux_divergence", "ath_gpu_block_cons2prim",
           "ath_gpu_block_new_dt", "ath_gpu_block_new_dt_noalloc",
           "ath_gpu_mesh_set_fused", "ath_gpu_mesh_get_fused", "ath_gpu_flux_div_plan",
           "ath_gpu_mesh_set_graph", "ath_gpu_smr_mesh_create", "ath_gpu_smr_block_level",
           "ath_gpu_smr_plan", "ath_gpu_flux_march_position"]
NKERNELS = 26


class AthGpuError(RuntimeError):
    pass


class RefineRegionC(ctypes.Structure):
    _fields_ = [("xmin", ctypes.c_double * 3), ("xmax", ctypes.c_double * 3),
                ("level", ctypes.c_int)]


class MeshDescC(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int * 3), ("bnx", ctypes.c_int * 3), ("ng", ctypes.c_int),
                ("xorder", ctypes.c_int), ("rsolver", ctypes.c_int),
                ("relativistic", ctypes.c_int),
                ("xmin", ctypes.c_double * 3), ("xmax", ctypes.c_double * 3),
                ("gamma", ctypes.c_double), ("dfloor", ctypes.c_double),
                ("pfloor", ctypes.c_double), ("cfl", ctypes.c_double),
                ("sigma_max", ctypes.c_double), ("beta_min", ctypes.c_double),
                ("gamma_max", ctypes.c_double), ("bc", ctypes.c_int * 6)]


class BlockDescC(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int * 3), ("ng", ctypes.c_int)] + \
        [(n, ctypes.POINTER(ctypes.c_double)) for n in
         ("x1f", "x2f", "x3f", "x1v", "x2v", "x3v", "dx1f", "dx2f", "dx3f")] + \
        [("rsolver", ctypes.c_int), ("relativistic", ctypes.c_int)] + \
        [(n, ctypes.c_double) for n in
         ("gamma", "dfloor", "pfloor", "sigma_max", "beta_min", "gamma_max", "cfl")]


class BlockRegsC(ctypes.Structure):
    _fields_ = [("flux", ctypes.c_void_p * 3)] + \
        [(n, ctypes.c_void_p) for n in
         ("e3_x1f", "e2_x1f", "e1_x2f", "e3_x2f", "e2_x3f", "e1_x3f")] + \
        [("wght", ctypes.c_void_p * 3), ("e", ctypes.c_void_p * 3)]


_lib = None


def load():
    """Load the library (no device call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise AthGpuError("libathena_gpu.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIBPATH)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.c_int
    for name in EXPORTS:
        getattr(L, name).restype = ctypes.c_int
    L.ath_gpu_last_error.restype = ctypes.c_char_p
    L.ath_gpu_mesh_create.argtypes = [ctypes.POINTER(MeshDescC), ctypes.POINTER(P)]
    L.ath_gpu_mesh_destroy.argtypes = [P]
    L.ath_gpu_nblocks.argtypes = [P, ctypes.POINTER(I)]
    L.ath_gpu_block_size.argtypes = [P, I, ctypes.POINTER(ctypes.c_long)]
    L.ath_gpu_set_block.argtypes = [P, I, I, D]
    L.ath_gpu_get_block.argtypes = [P, I, I, D]
    L.ath_gpu_initialize.argtypes = [P, ctypes.c_double]
    L.ath_gpu_vl2_stage.argtypes = [P, I]
    L.ath_gpu_cycle.argtypes = [P, I, ctypes.c_double, ctypes.POINTER(I)]
    L.ath_gpu_get_clock.argtypes = [P, D, D, ctypes.POINTER(I)]
    L.ath_gpu_set_clock.argtypes = [P, ctypes.c_double, ctypes.c_double, I]
    L.ath_gpu_calculate_fluxes.argtypes = [P, I]
    L.ath_gpu_compute_corner_e.argtypes = [P]
    L.ath_gpu_stream.argtypes = [P, ctypes.POINTER(P)]
    L.ath_gpu_synchronize.argtypes = [P]
    L.ath_gpu_kernel_timing.argtypes = [P, I]
    L.ath_gpu_kernel_stats.argtypes = [P, I, ctypes.c_char_p, I, D, ctypes.POINTER(ctypes.c_long)]
    L.ath_gpu_riemann_batch.argtypes = [I, ctypes.c_double, I, D, D, D, D]
    L.ath_gpu_sr_con2prim_batch.argtypes = [I] + [ctypes.c_double] * 6 + \
        [D, D, D, ctypes.POINTER(I)]
    F = ctypes.POINTER(ctypes.c_float)
    IP = ctypes.POINTER(I)
    L.ath_ml_last_error.restype = ctypes.c_char_p
    L.ath_ml_param_count.argtypes = []
    L.ath_ml_create.argtypes = [ctypes.POINTER(P)]
    L.ath_ml_destroy.argtypes = [P]
    L.ath_ml_set_params.argtypes = [P, F]
    L.ath_ml_get_params.argtypes = [P, F]
    L.ath_ml_grad.argtypes = [P, I, F, F, F, F]
    L.ath_ml_train.argtypes = [P, I, F, F, I, I, IP, ctypes.c_float, F]
    L.ath_ml_infer.argtypes = [P, I, F, F, I]
    L.ath_ml_features.argtypes = [I, D, D, F]
    L.ath_gpu_sr_con2prim_modes.argtypes = [P, I, I] + [ctypes.c_double] * 6 + \
        [D, D, D, IP, IP, IP]
    L.ath_gpu_sr_con2prim_device.argtypes = [P, I, I] + [ctypes.c_double] * 6 + \
        [P, P, P, P, P, P, P]
    L.ath_gpu_mesh_set_c2p.argtypes = [P, P, I]
    LP = ctypes.POINTER(ctypes.c_long)
    L.ath_gpu_mesh_set_fused.argtypes = [P, I]
    L.ath_gpu_mesh_set_graph.argtypes = [P, I]
    L.ath_gpu_mesh_get_fused.argtypes = [P, IP]
    L.ath_gpu_flux_div_plan.argtypes = [IP, I, I, IP, ctypes.c_long, LP, IP, ctypes.c_long,
                                        LP, LP]
    L.ath_gpu_mesh_c2p_stats.argtypes = [P, I, ctypes.POINTER(ctypes.c_ulonglong)]
    L.ath_gpu_ensemble_create.argtypes = [ctypes.POINTER(P), I, ctypes.POINTER(P)]
    L.ath_gpu_ensemble_destroy.argtypes = [P]
    L.ath_gpu_ensemble_size.argtypes = [P, IP, ctypes.POINTER(ctypes.c_long)]
    L.ath_gpu_ensemble_cycle.argtypes = [P, I, I]
    L.ath_gpu_ensemble_get_tiles.argtypes = [P, IP, IP]
    L.ath_gpu_ensemble_set_tiles.argtypes = [P, IP, IP]
    L.ath_gpu_ensemble_autotune.argtypes = [P, I, D]
    L.ath_gpu_ensemble_stream.argtypes = [P, ctypes.POINTER(P)]
    L.ath_gpu_mesh_set_tiles.argtypes = [P, IP]
    L.ath_gpu_mesh_get_tiles.argtypes = [P, IP]
    L.ath_gpu_mesh_autotune.argtypes = [P, I, D]
    L.ath_gpu_kernel_box.argtypes = [I, I, I, IP, IP]
    L.ath_gpu_ghost_position.argtypes = [I, I, I, I, IP]
    L.ath_gpu_flux_march_position.argtypes = [IP, I, ctypes.c_long, I, IP,
                                              ctypes.POINTER(ctypes.c_long)]
    L.ath_gpu_restrict_cc.argtypes = [IP, I, I, P, P, P, IP, P]
    L.ath_gpu_restrict_fc.argtypes = [IP, I, I, P, P, P, IP, P]
    L.ath_gpu_prolongate_cc.argtypes = [IP, I, I, P, P, P, P, IP, P]
    L.ath_gpu_prolongate_fc_shared.argtypes = [IP, I, I, P, P, P, P, IP, P]
    L.ath_gpu_prolongate_fc_internal.argtypes = [IP, I, P, P, P, P, IP, P]
    L.ath_gpu_restrict_emf.argtypes = [IP, I, I, P, P, P, IP, P]
    L.ath_gpu_tree_build.argtypes = [IP, IP, D, D, ctypes.POINTER(RefineRegionC), I,
                                     ctypes.POINTER(ctypes.c_longlong), IP, I, IP, IP]
    L.ath_gpu_history.argtypes = [P, D]
    BR = ctypes.POINTER(BlockRegsC)
    Dd = ctypes.c_double
    L.ath_gpu_block_create.argtypes = [ctypes.POINTER(BlockDescC), ctypes.POINTER(P)]
    L.ath_gpu_block_destroy.argtypes = [P]
    L.ath_gpu_block_reconstruct.argtypes = [P, I, I, I, I, I, I, P, P, P, P, P]
    L.ath_gpu_block_riemann.argtypes = [P, I, I, I, I, I, P, P, P, P, P, P, P, P, Dd, P]
    L.ath_gpu_block_calculate_fluxes.argtypes = [P, P, P, P, P, P, I, Dd, BR, P]
    L.ath_gpu_block_compute_corner_e.argtypes = [P, P, P, BR, P]
    L.ath_gpu_block_ct.argtypes = [P, Dd, BR, P, P, P, P]
    L.ath_gpu_block_add_flux_divergence.argtypes = [P, Dd, BR, P, P]
    L.ath_gpu_block_cons2prim.argtypes = [P, P, P, P, P, P, P, P, I, I, I, I, I, I, P]
    L.ath_gpu_block_new_dt.argtypes = [P, P, P, P, P, P, P, P]
    L.ath_gpu_block_new_dt_noalloc.argtypes = [P, P, P, P, P, P, P, P]
    L.ath_gpu_smr_mesh_create.argtypes = [ctypes.POINTER(MeshDescC),
                                          ctypes.POINTER(ctypes.c_longlong), I, I,
                                          ctypes.POINTER(P)]
    L.ath_gpu_smr_block_level.argtypes = [P, I, IP]
    L.ath_gpu_smr_plan.argtypes = [ctypes.POINTER(MeshDescC), ctypes.POINTER(ctypes.c_longlong),
                                   I, I, I, IP, I, IP, LP]
    _lib = L
    return L


def tree_build(nrb, bc=(0,) * 6, xmin=(0.0,) * 3, xmax=(1.0,) * 3, regions=()):
    """MeshBlock tree of a statically refined mesh (ath_gpu_tree_build; host only).
    regions: (xmin[3], xmax[3], level) triples.  Returns (locs, nbr, root_level): locs an
    (n, 4) int64 array of (level, lx1, lx2, lx3) in Z order, nbr an (n, 3, 3, 3) int32
    array indexed [o3+1, o2+1, o1+1] (leaf id, -2 finer, -1 none)."""
    L = load()
    nrb_a = np.ascontiguousarray(nrb, dtype=np.int32)
    bc_a = np.ascontiguousarray(bc, dtype=np.int32)
    lo = np.ascontiguousarray(xmin, dtype=np.float64)
    hi = np.ascontiguousarray(xmax, dtype=np.float64)
    reg = (RefineRegionC * max(len(regions), 1))()
    for r, (a, b, lev) in enumerate(regions):
        reg[r].xmin[:] = [float(v) for v in a]
        reg[r].xmax[:] = [float(v) for v in b]
        reg[r].level = int(lev)
    nb, rl = ctypes.c_int(0), ctypes.c_int(0)
    # first call sizes the output (it fails only on cap when the arguments are valid)
    rc = L.ath_gpu_tree_build(_ip(nrb_a), _ip(bc_a), _dp(lo), _dp(hi), reg, len(regions), None,
                          None, 0, ctypes.byref(nb), ctypes.byref(rl))
    if rc != 0 and nb.value == 0:
        _check(rc)
    n = nb.value
    locs = np.zeros((n, 4), dtype=np.int64)
    nbr = np.zeros((n, 27), dtype=np.int32)
    _check(L.ath_gpu_tree_build(_ip(nrb_a), _ip(bc_a), _dp(lo), _dp(hi), reg, len(regions),
                            locs.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                            _ip(nbr), n, ctypes.byref(nb), ctypes.byref(rl)))
    return locs, nbr.reshape(n, 3, 3, 3), rl.value


def _check(rc):
    if rc != 0:
        raise AthGpuError(load().ath_gpu_last_error().decode())


def _check_ml(rc):
    if rc != 0:
        raise AthGpuError(load().ath_ml_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def desc_of(m: MeshDesc) -> MeshDescC:
    c = MeshDescC()
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


class GpuMesh:
    """Device mirror of one mesh (all registers resident in HBM): uniform, or statically
    refined when ``m`` comes from ``athena_amd.mesh.refined``."""

    def __init__(self, m: MeshDesc):
        L = load()
        self.m = m
        self._desc = desc_of(m)
        h = ctypes.c_void_p()
        if m.smr:
            self._locs = np.ascontiguousarray(m.locs, dtype=np.int64)
            _check(L.ath_gpu_smr_mesh_create(
                ctypes.byref(self._desc),
                self._locs.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                len(self._locs), int(m.root_level), ctypes.byref(h)))
        else:
            _check(L.ath_gpu_mesh_create(ctypes.byref(self._desc), ctypes.byref(h)))
        self.h = h

    def level(self, gid):
        """Logical level of MeshBlock gid (0 on a uniform mesh)."""
        lv = ctypes.c_int()
        _check(load().ath_gpu_smr_block_level(self.h, gid, ctypes.byref(lv)))
        return lv.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().ath_gpu_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self, what):
        n = ctypes.c_long()
        _check(load().ath_gpu_block_size(self.h, ATH[what], ctypes.byref(n)))
        return n.value

    def set(self, gid, what, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.size != self.size(what):
            raise AthGpuError("register %s of block %d: got %d doubles, need %d"
                              % (what, gid, a.size, self.size(what)))
        _check(load().ath_gpu_set_block(self.h, gid, ATH[what], _dp(a)))

    def get(self, gid, what):
        a = np.empty(self.size(what))
        _check(load().ath_gpu_get_block(self.h, gid, ATH[what], _dp(a)))
        if what in ("u", "w", "w1", "bcc", "b1", "b2", "b3"):
            return a.reshape(self.m.shape(what))
        return a

    def load(self, blocks):
        for gid, b in enumerate(blocks):
            for k in ("u", "b1", "b2", "b3", "w", "w1"):
                if k in b:
                    self.set(gid, k, b[k])

    def initialize(self, tlim=float("inf")):
        _check(load().ath_gpu_initialize(self.h, tlim))

    def run(self, ncycles, tlim=float("inf")):
        done = ctypes.c_int()
        _check(load().ath_gpu_cycle(self.h, ncycles, tlim, ctypes.byref(done)))
        return done.value

    def stage(self, s):
        _check(load().ath_gpu_vl2_stage(self.h, s))

    def state(self):
        t, dt, nc = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _check(load().ath_gpu_get_clock(self.h, ctypes.byref(t), ctypes.byref(dt),
                                        ctypes.byref(nc)))
        return t.value, dt.value, nc.value

    def set_clock(self, time, dt, ncycle):
        _check(load().ath_gpu_set_clock(self.h, time, dt, ncycle))

    def synchronize(self):
        _check(load().ath_gpu_synchronize(self.h))

    def stream(self):
        s = ctypes.c_void_p()
        _check(load().ath_gpu_stream(self.h, ctypes.byref(s)))
        return s.value

    def block(self, gid):
        return {k: self.get(gid, k) for k in ("u", "w", "bcc", "b1", "b2", "b3")}

    def set_c2p(self, mode, mlp=None):
        """First solver of the SR ConservedToPrimitive: 0 reference iteration, 1 Newton
        from the previous pressure, 2 Newton from the learned guess of `mlp`."""
        _check(load().ath_gpu_mesh_set_c2p(self.h, mlp.h if mlp is not None else None,
                                           int(mode)))
        self._mlp = mlp   # the mesh keeps a device reference to the network

    def c2p_stats(self, enable=-1):
        """(evaluations, fallbacks, cells) counted before this call; enable 1 clears and
        starts counting, 0 stops, -1 only reads."""
        t = (ctypes.c_ulonglong * 3)()
        _check(load().ath_gpu_mesh_c2p_stats(self.h, int(enable), t))
        return int(t[0]), int(t[1]), int(t[2])

    def history(self):
        """The 11 history sums (mass, 1..3-mom, 1..3-KE, tot-E, 1..3-ME)."""
        h = np.zeros(11)
        _check(load().ath_gpu_history(self.h, _dp(h)))
        return h

    def set_tiles(self, tile_idx):
        t = np.ascontiguousarray(tile_idx, dtype=np.int32)
        _check(load().ath_gpu_mesh_set_tiles(self.h, _ip(t)))

    def set_fused(self, rows=-1):
        """Rows per band of the fused flux + flux-divergence kernel (0: separate flux and
        update kernels, -1: the default)."""
        _check(load().ath_gpu_mesh_set_fused(self.h, int(rows)))

    def set_graph(self, on=True):
        """Replay each cycle of run() as one captured HIP graph (same bits as eager)."""
        _check(load().ath_gpu_mesh_set_graph(self.h, int(bool(on))))

    def fused(self):
        r = ctypes.c_int()
        _check(load().ath_gpu_mesh_get_fused(self.h, ctypes.byref(r)))
        return r.value

    def tiles(self):
        t = np.zeros(4, dtype=np.int32)
        _check(load().ath_gpu_mesh_get_tiles(self.h, _ip(t)))
        return t

    def autotune(self, reps=3):
        """Choose each kernel kind's tile shape by timing; the state is unchanged."""
        best = np.zeros(4)
        _check(load().ath_gpu_mesh_autotune(self.h, int(reps), _dp(best)))
        return best

    def kernel_timing(self, enable=True):
        _check(load().ath_gpu_kernel_timing(self.h, int(enable)))

    def kernel_stats(self):
        """{name: (total_ms, launches)} for kernels launched since timing was enabled."""
        out = {}
        buf = ctypes.create_string_buffer(64)
        for kid in range(NKERNELS):
            ms, cnt = ctypes.c_double(), ctypes.c_long()
            _check(load().ath_gpu_kernel_stats(self.h, kid, buf, 64, ctypes.byref(ms),
                                               ctypes.byref(cnt)))
            if cnt.value:
                out[buf.value.decode()] = (ms.value, cnt.value)
        return out


def riemann_batch(solver, gamma, wl, wr, bx):
    """solver: 0 HLLD, 1 HLLE, 2 SR HLLE, 3 SR HLLD; rotated-frame states (n,7)."""
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    bx = np.ascontiguousarray(bx, dtype=np.float64)
    out = np.empty_like(wl)
    _check(load().ath_gpu_riemann_batch(solver, gamma, wl.shape[0], _dp(wl), _dp(wr), _dp(bx),
                                        _dp(out)))
    return out


def sr_con2prim_batch(cons, prim_seed, bcc, gamma, dfloor, pfloor, sigma_max=0.0,
                      beta_min=0.0, gamma_max=1000.0):
    cons = np.array(cons, dtype=np.float64, order="C")
    prim = np.array(prim_seed, dtype=np.float64, order="C")
    bcc = np.ascontiguousarray(bcc, dtype=np.float64)
    iters = np.empty(cons.shape[0], dtype=np.int32)
    _check(load().ath_gpu_sr_con2prim_batch(
        cons.shape[0], gamma, dfloor, pfloor, sigma_max, beta_min, gamma_max, _dp(cons),
        _dp(prim), _dp(bcc), iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    return cons, prim, iters


class Ensemble:
    """Independent meshes advanced together (ath_gpu_ensemble_*).  The members stay
    usable for get/set of registers between calls."""

    KINDS = ("flux", "corner", "update", "cons2prim")
    TILES = ("linear", "64x4x1", "32x8x1", "16x16x1", "16x4x4", "8x8x4", "zmarch8",
             "zmarch16")   # zmarch: corner kernel only

    def __init__(self, meshes):
        L = load()
        self.meshes = list(meshes)
        arr = (ctypes.c_void_p * len(self.meshes))(*[m.h.value for m in self.meshes])
        h = ctypes.c_void_p()
        _check(L.ath_gpu_ensemble_create(arr, len(self.meshes), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().ath_gpu_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self):
        n, c = ctypes.c_int(), ctypes.c_long()
        _check(load().ath_gpu_ensemble_size(self.h, ctypes.byref(n), ctypes.byref(c)))
        return n.value, c.value

    def run(self, ncycles, graph=True):
        _check(load().ath_gpu_ensemble_cycle(self.h, int(ncycles), int(bool(graph))))

    def tiles(self):
        t = np.zeros(4, dtype=np.int32)
        p = np.zeros(4, dtype=np.int32)
        _check(load().ath_gpu_ensemble_get_tiles(self.h, _ip(t), _ip(p)))
        return t, p

    def set_tiles(self, tile_idx, persist):
        t = np.ascontiguousarray(tile_idx, dtype=np.int32)
        p = np.ascontiguousarray(persist, dtype=np.int32)
        _check(load().ath_gpu_ensemble_set_tiles(self.h, _ip(t), _ip(p)))

    def autotune(self, reps=3):
        best = np.zeros(4)
        _check(load().ath_gpu_ensemble_autotune(self.h, int(reps), _dp(best)))
        return best

    def stream(self):
        s = ctypes.c_void_p()
        _check(load().ath_gpu_ensemble_stream(self.h, ctypes.byref(s)))
        return s.value


C2P_MODES = {"reference": 0, "newton": 1, "learned": 2}


class Mlp:
    """Device copy of the learned-guess network (ath_ml_* entry points)."""

    def __init__(self, params=None):
        L = load()
        self.np = L.ath_ml_param_count()
        h = ctypes.c_void_p()
        _check_ml(L.ath_ml_create(ctypes.byref(h)))
        self.h = h
        if params is not None:
            self.set_params(params)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().ath_ml_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != self.np:
            raise AthGpuError("parameter vector has %d floats, need %d" % (p.size, self.np))
        _check_ml(load().ath_ml_set_params(self.h, _fp(p)))

    def params(self):
        p = np.empty(self.np, dtype=np.float32)
        _check_ml(load().ath_ml_get_params(self.h, _fp(p)))
        return p

    def grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(NW, dtype=np.float32)
        loss = np.empty(1, dtype=np.float32)
        _check_ml(load().ath_ml_grad(self.h, t.size, _fp(x), _fp(t), _fp(g), _fp(loss)))
        return g, float(loss[0])

    def train(self, x, t, order, batch, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        order = np.ascontiguousarray(order, dtype=np.int32)
        steps = order.size // batch
        hist = np.empty(steps, dtype=np.float32)
        _check_ml(load().ath_ml_train(self.h, t.size, _fp(x), _fp(t), steps, batch,
                                      _ip(order), lr, _fp(hist)))
        return hist

    def infer(self, x, bf16=True):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty(x.shape[0], dtype=np.float32)
        _check_ml(load().ath_ml_infer(self.h, x.shape[0], _fp(x), _fp(y), int(bool(bf16))))
        return y


NW = 4673          # trained parameters W1..b3 (the rest is normalisation)


def ml_features(cons, bcc):
    """The solver's raw inputs (n,6) float32 of cons (n,5) and bcc (n,3)."""
    cons = np.ascontiguousarray(cons, dtype=np.float64)
    bcc = np.ascontiguousarray(bcc, dtype=np.float64)
    x = np.empty((cons.shape[0], 6), dtype=np.float32)
    _check_ml(load().ath_ml_features(cons.shape[0], _dp(cons), _dp(bcc), _fp(x)))
    return x


def sr_con2prim_modes(mode, cons, prim_seed, bcc, gamma, dfloor, pfloor, mlp=None,
                      sigma_max=0.0, beta_min=0.0, gamma_max=1000.0):
    """Batched SR recovery with the first solver `mode` (0/1/2 or a C2P_MODES key).
    Returns cons, prim, iters, evals, fallback."""
    mode = C2P_MODES.get(mode, mode)
    cons = np.array(cons, dtype=np.float64, order="C")
    prim = np.array(prim_seed, dtype=np.float64, order="C")
    bcc = np.ascontiguousarray(bcc, dtype=np.float64)
    n = cons.shape[0]
    it, ev, fb = (np.empty(n, dtype=np.int32) for _ in range(3))
    _check_ml(load().ath_gpu_sr_con2prim_modes(
        mlp.h if mlp is not None else None, int(mode), n, gamma, dfloor, pfloor, sigma_max,
        beta_min, gamma_max, _dp(cons), _dp(prim), _dp(bcc), _ip(it), _ip(ev), _ip(fb)))
    return cons, prim, it, ev, fb


def sr_con2prim_device(mode, n, cons, prim, bcc, gamma, dfloor, pfloor, mlp=None,
                       iters=None, evals=None, fallback=None, stream=None, sigma_max=0.0,
                       beta_min=0.0, gamma_max=1000.0):
    """Asynchronous form on device addresses (ints, e.g. a torch tensor's data_ptr()):
    cons (n,5), prim (n,5), bcc (n,3) float64; iters/evals/fallback int32 or None."""
    mode = C2P_MODES.get(mode, mode)
    _check_ml(load().ath_gpu_sr_con2prim_device(
        mlp.h if mlp is not None else None, int(mode), int(n), gamma, dfloor, pfloor,
        sigma_max, beta_min, gamma_max, cons, prim, bcc, iters, evals, fallback, stream))


def kernel_box(kind, nx, ng):
    """(lo, n) of the position box of kernel kind 0 flux, 1 corner, 2 update,
    3 cons2prim in a MeshBlock of nx^3 cells and ng ghosts (no device call)."""
    lo = np.zeros(3, dtype=np.int32)
    n = np.zeros(3, dtype=np.int32)
    _check(load().ath_gpu_kernel_box(int(kind), int(nx), int(ng), _ip(lo), _ip(n)))
    return lo, n


class Refinement:
    """MeshRefinement operators on device arrays (mesh_refinement.cpp, 3-D uniform
    Cartesian): ``shape`` = (nx1, nx2, nx3, ng) of the fine blocks, nb blocks per call
    stored back to back; arguments are device addresses (ints, e.g. a torch tensor's
    data_ptr()), ``r`` the coarse index range (si, ei, sj, ej, sk, ek)."""

    def __init__(self, nx, ng, stream=None):
        self.shape = np.array(list(nx) + [ng], dtype=np.int32)
        self.stream = stream

    def _r(self, r):
        return np.ascontiguousarray(r, dtype=np.int32)

    def restrict_cc(self, nb, nvar, fcoord, fine, coarse, r):
        r = self._r(r)
        _check(load().ath_gpu_restrict_cc(_ip(self.shape), nb, nvar, fcoord, fine, coarse,
                                          _ip(r), self.stream))

    def restrict_fc(self, nb, comp, fcoord, fine, coarse, r):
        r = self._r(r)
        _check(load().ath_gpu_restrict_fc(_ip(self.shape), nb, comp, fcoord, fine, coarse,
                                          _ip(r), self.stream))

    def prolongate_cc(self, nb, nvar, fcoord, ccoord, coarse, fine, r):
        r = self._r(r)
        _check(load().ath_gpu_prolongate_cc(_ip(self.shape), nb, nvar, fcoord, ccoord,
                                            coarse, fine, _ip(r), self.stream))

    def prolongate_fc_shared(self, nb, comp, fcoord, ccoord, coarse, fine, r):
        r = self._r(r)
        _check(load().ath_gpu_prolongate_fc_shared(_ip(self.shape), nb, comp, fcoord,
                                                   ccoord, coarse, fine, _ip(r),
                                                   self.stream))

    def restrict_emf(self, nb, comp, fcoord, fine, coarse, r):
        r = self._r(r)
        _check(load().ath_gpu_restrict_emf(_ip(self.shape), nb, comp, fcoord, fine, coarse,
                                           _ip(r), self.stream))

    def prolongate_fc_internal(self, nb, fcoord, b1, b2, b3, r):
        r = self._r(r)
        _check(load().ath_gpu_prolongate_fc_internal(_ip(self.shape), nb, fcoord, b1, b2,
                                                     b3, _ip(r), self.stream))


def ghost_position(nx, ng, array, t):
    """Ghost position number t of the exchange kernel (array 0 cells, 1..3 face
    components): (i, j, k, source i, source j, source k, p1, p2, p3); no device call."""
    pos = np.zeros(9, dtype=np.int32)
    _check(load().ath_gpu_ghost_position(int(nx), int(ng), int(array), int(t), _ip(pos)))
    return pos


def flux_march_position(shape, seg, wave, lane):
    """Lane ``lane`` of wavefront ``wave`` of the flux kernel's z-march on MeshBlocks of
    shape (nx1, nx2, nx3, ng, nblocks) with segments of about ``seg`` planes:
    ((in, block, j, i, k0, k1, tallest segment), wavefronts per launch); no device call."""
    pos = np.zeros(7, dtype=np.int32)
    nw = ctypes.c_long(0)
    sh = np.asarray(shape, dtype=np.int32)
    _check(load().ath_gpu_flux_march_position(_ip(sh), int(seg), int(wave), int(lane),
                                              _ip(pos), ctypes.byref(nw)))
    return pos, nw.value


class Block:
    """One MeshBlock whose registers the caller keeps in device memory
    (ath_gpu_block_*): the reference's per-MeshBlock functions with its own signatures.
    Array arguments are device addresses (ints, e.g. a torch tensor's data_ptr());
    ``regs`` maps the names of ath_block_regs (flux1..3, e3_x1f, e2_x1f, e1_x2f, e3_x2f,
    e2_x3f, e1_x3f, wght1..3, e1..e3) to device addresses.  Calls are asynchronous on
    ``stream`` (None: the default stream)."""

    def __init__(self, m: MeshDesc, gid=0, stream=None):
        from .mesh import block_coords
        c = block_coords(m, gid)
        self._keep = [np.ascontiguousarray(getattr(c, n), dtype=np.float64) for n in
                      ("x1f", "x2f", "x3f", "x1v", "x2v", "x3v", "dx1f", "dx2f", "dx3f")]
        d = BlockDescC()
        d.nx[:] = list(m.bnx)
        d.ng = m.ng
        for name, a in zip(("x1f", "x2f", "x3f", "x1v", "x2v", "x3v", "dx1f", "dx2f", "dx3f"),
                           self._keep):
            setattr(d, name, _dp(a))
        d.rsolver = RSOLVER[m.rsolver]
        d.relativistic = int(m.relativistic)
        d.gamma, d.dfloor, d.pfloor = m.gamma, m.dfloor, m.pfloor
        d.sigma_max, d.beta_min, d.gamma_max = m.sigma_max, m.beta_min, m.gamma_max
        d.cfl = m.cfl
        self._desc = d
        h = ctypes.c_void_p()
        _check(load().ath_gpu_block_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.stream = stream

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().ath_gpu_block_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def regs(r):
        c = BlockRegsC()
        for q in range(3):
            c.flux[q] = r.get("flux%d" % (q + 1))
            c.wght[q] = r.get("wght%d" % (q + 1))
            c.e[q] = r.get("e%d" % (q + 1))
        for n in ("e3_x1f", "e2_x1f", "e1_x2f", "e3_x2f", "e2_x3f", "e1_x3f"):
            setattr(c, n, r.get(n))
        return c

    def reconstruct(self, dir, order, k, j, il, iu, w, bcc, wl, wr):
        _check(load().ath_gpu_block_reconstruct(self.h, dir, order, k, j, il, iu, w, bcc, wl,
                                                wr, self.stream))

    def riemann(self, k, j, il, iu, ivx, bx, wl, wr, flx, ey, ez, wct, dxw, dt):
        _check(load().ath_gpu_block_riemann(self.h, k, j, il, iu, ivx, bx, wl, wr, flx, ey, ez,
                                            wct, dxw, dt, self.stream))

    def calculate_fluxes(self, w, b1, b2, b3, bcc, order, dt, regs):
        r = self.regs(regs)
        _check(load().ath_gpu_block_calculate_fluxes(self.h, w, b1, b2, b3, bcc, order, dt,
                                                     ctypes.byref(r), self.stream))

    def compute_corner_e(self, w, bcc, regs):
        r = self.regs(regs)
        _check(load().ath_gpu_block_compute_corner_e(self.h, w, bcc, ctypes.byref(r),
                                                     self.stream))

    def ct(self, wght, regs, b1, b2, b3):
        r = self.regs(regs)
        _check(load().ath_gpu_block_ct(self.h, wght, ctypes.byref(r), b1, b2, b3, self.stream))

    def add_flux_divergence(self, wght, regs, u):
        r = self.regs(regs)
        _check(load().ath_gpu_block_add_flux_divergence(self.h, wght, ctypes.byref(r), u,
                                                        self.stream))

    def cons2prim(self, cons, prim_old, b1, b2, b3, prim, bcc, rng):
        il, iu, jl, ju, kl, ku = (int(v) for v in rng)
        _check(load().ath_gpu_block_cons2prim(self.h, cons, prim_old, b1, b2, b3, prim, bcc,
                                              il, iu, jl, ju, kl, ku, self.stream))

    def new_dt(self, w, bcc, b1, b2, b3, dt_out):
        _check(load().ath_gpu_block_new_dt(self.h, w, bcc, b1, b2, b3, dt_out, self.stream))

    def new_dt_noalloc(self, w, bcc, b1, b2, b3, dt_out):
        """ath_gpu_block_new_dt_noalloc: the same result through the library's fixed
        device scratch array, without allocating."""
        _check(load().ath_gpu_block_new_dt_noalloc(self.h, w, bcc, b1, b2, b3, dt_out,
                                                   self.stream))


def smr_plan(m: MeshDesc, gid=0):
    """Host-only check of a refined mesh (ath_gpu_smr_plan): returns (neighbour list of
    block gid as (n, 9) int32 rows, [step counts of the flux, EMF and boundary programs,
    then their numbers of phases])."""
    L = load()
    desc = desc_of(m)
    locs = np.ascontiguousarray(m.locs, dtype=np.int64)
    out = np.zeros(9 * 64, dtype=np.int32)
    nn = ctypes.c_int()
    steps = np.zeros(6, dtype=np.int64)
    _check(L.ath_gpu_smr_plan(ctypes.byref(desc),
                              locs.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), len(locs),
                              int(m.root_level), gid, _ip(out), 64, ctypes.byref(nn),
                              steps.ctypes.data_as(ctypes.POINTER(ctypes.c_long))))
    return out[:9 * nn.value].reshape(nn.value, 9), steps


def flux_div_plan(nx, ng, rows):
    """The fused flux kernel's plan for one MeshBlock shape, checked on the host by the
    library with the kernel's own index helpers: (faces (n, 4) as dir, k, j, i whose EMFs
    and weights it writes; cells (n, 3) as k, j, i that it updates; LDS doubles per
    work-group).  Raises AthGpuError where an LDS access would leave the array or read the
    wrong face."""
    nxa = np.ascontiguousarray(nx, dtype=np.int32)
    nf, nc, nl = ctypes.c_long(), ctypes.c_long(), ctypes.c_long()
    _check(load().ath_gpu_flux_div_plan(_ip(nxa), ng, rows, None, 0, ctypes.byref(nf), None, 0,
                                        ctypes.byref(nc), ctypes.byref(nl)))
    faces = np.zeros((max(nf.value, 1), 4), dtype=np.int32)
    cells = np.zeros((max(nc.value, 1), 3), dtype=np.int32)
    _check(load().ath_gpu_flux_div_plan(_ip(nxa), ng, rows, _ip(faces), nf.value,
                                        ctypes.byref(nf), _ip(cells), nc.value,
                                        ctypes.byref(nc), ctypes.byref(nl)))
    return faces[:nf.value], cells[:nc.value], nl.value
