#!/usr/bin/env python3
"""Throughput benchmark of the MI355X per-MeshBlock MHD update.

Workload (the throughput configuration): 3-D ideal MHD on a 256^3 periodic mesh split
into 512 MeshBlocks of 32^3 (NGHOST=2), rho=1, P=1, B=(0.1,0,0), gamma=5/3, velocity a
seeded Gaussian random field with k^-2 power on 2<=|k|<=8 scaled to rms Mach 1;
VL2 + PLM + HLLD + CT in FP64.  A "step" is one VL2 cycle (two stages + new dt).

Prints one JSON line: zone-cycles/s (Athena++'s own formula, main.cpp:641-649:
cycles x active cells / wall time) with inputs resident in HBM, the roofline of the
dominant kernel and of the whole step, and the CPU restatement timed on a bounded
sample on the host cores.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12          # MI355X HBM3E, B/s
TILE_NAMES = ("linear", "64x4x1", "32x8x1", "16x16x1", "16x4x4", "8x8x4", "zmarch8",
              "zmarch16")
BYTES_PER_ZONE_CYCLE = 536  # compulsory FP64 traffic of one VL2 MHD zone-cycle
METRIC = "zone-cycles/s on 1 MI355X (3D MHD VL2+PLM+HLLD FP64) + % peak HBM GB/s"


class HipEvents:
    """hipEventRecord on an explicit stream (the library's), via the HIP runtime."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                               ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def new(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def destroy(self, e):
        assert self.h.hipEventDestroy(e) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value


# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on the
# default workload (256^3 as 512 MeshBlocks of 32^3; tiles linear, zmarch8, linear,
# linear), fetched KiB x 2.0 + written KiB x 1.0 (the factors a 512 MiB device copy
# gives): profiles/vl2_hlld_plm_turbulence256_hbm_traffic.md.
MEASURED_TRAFFIC = {
    "flux_xyz_plm": 11396524816,
    "flux_xyz_dc": 8299554176,
    "update_u_b_ct": 7940971036,
    "corner_emf": 3773546268,
    "cons2prim": 3309664999,
    "ghost_exchange": 2299010489,
}
TRAFFIC_TILES = ["linear", "zmarch8", "linear", "linear"]
# FP64 FLOP per launch (64 x (ADD + MUL + 2 FMA) of the SQ_INSTS_VALU_*_F64 counters) on
# the same workload: profiles/vl2_hlld_plm_turbulence256_fp64_valu.md.
MEASURED_FP64_FLOP = {
    "flux_xyz_plm": 47786607216,
    "flux_xyz_dc": 31422754608,
    "update_u_b_ct": 4173332480,
    "corner_emf": 2422603776,
    "cons2prim": 3371907300,
}
FP64_PEAK = 78.6e12
TRAFFIC_SOURCE = "profiles/vl2_hlld_plm_turbulence256_hbm_traffic.md"
FP64_SOURCE = "profiles/vl2_hlld_plm_turbulence256_fp64_valu.md"

# The oracle against the reference's own CPU build on identical inputs (a cubic 64^3
# MeshBlock, diagonal fast wave, VL2+PLM+HLLD, 10 cycles, one thread, same container):
# 6.96e5 zone-cycles per CPU-second for the oracle against 5.48e5 for the reference
# compiled from its sources with its own configure line (g++ -O3).  The oracle is this
# much faster than the reference; the bench divides its CPU figure by it.
ORACLE_OVER_REFERENCE = 6.96e5 / 5.48e5
CALIBRATION_BASIS = ("cubic 64^3 MeshBlock, diagonal fast wave, VL2+PLM+HLLD, 10 cycles, 1 "
                     "thread: oracle 6.96e5, reference built from source 5.48e5 "
                     "zone-cycles per CPU-second")


def kernel_algorithmic_bytes(name, m):
    """Compulsory HBM bytes of one launch of a production kernel over all MeshBlocks of
    mesh m: every array element the kernel must read or write, counted once (no
    re-reads, no cache effects).  Per MeshBlock with n active cells per direction and
    nc = n + 2*NGHOST:
      flux_xyz_*:  w, bcc (8 doubles) on the stencil box (n+2e)^3 (e = 1 DC, 2 PLM,
                   3 PPM), normal face fields (3), and per face flux(5)+EMF(2)+weight(1)
                   on 3 (n+1)(n+2)^2 faces;
      flux_div_*:  the same reads and per face EMF(2)+weight(1) writes, plus U^n (5) read
                   and U (5) written on the active cells (the fluxes stay on chip);
      update_b_ct: 3 face fields read and written and 3 edge EMF components read on the
                   active faces;
      corner_emf:  v, bcc (6) on (n+2)^3 cells, 6 face EMFs + 3 weights on the faces,
                   3 edge EMFs written;
      update_u_b_ct: U^n (5) and 3 face fields on the active region, 15 flux and 3 edge
                   components read, active U (5) and B (3) written;
      ghost_exchange: every ghost position of U (5 components) and of the 3 face
                   components read from its source and written;
      cons2prim:   U (5) and faces (3) read, W (5) and bcc (3) written on nc^3 cells."""
    n1, n2, n3 = m.bnx
    assert n1 == n2 == n3
    n = n1
    nc = n + 2 * m.ng
    nb = m.nblocks
    faces = 3 * (n + 1) * (n + 2) ** 2
    if name.startswith("flux_xyz_"):
        e = {"dc": 1, "plm": 2, "ppm": 3}[name.split("_")[2]]
        per = 8 * (n + 2 * e) ** 3 + faces + 8 * faces
    elif name.startswith("flux_div_"):
        e = {"dc": 1, "plm": 2, "ppm": 3}[name.split("_")[2]]
        per = 8 * (n + 2 * e) ** 3 + faces + 3 * faces + 10 * n ** 3
    elif name == "update_b_ct":
        per = 3 * (n + 1) * n * n + 3 * (n + 1) ** 2 * n + 3 * (n + 1) * n * n
    elif name == "corner_emf":
        per = 6 * (n + 2) ** 3 + 9 * (n + 1) * (n + 2) ** 2 + 3 * (n + 1) ** 2 * (n + 2)
    elif name == "update_u_b_ct":
        per = (5 * n ** 3 + 15 * (n + 1) * n * n + 3 * (n + 1) ** 2 * n + 3 * (n + 1) * n * n
               + 5 * n ** 3 + 3 * (n + 1) * n * n)
    elif name == "ghost_exchange":
        per = 2 * (5 * (nc ** 3 - n ** 3) + 3 * ((nc + 1) * nc * nc - (n + 1) * n * n))
    elif name == "cons2prim":
        per = 5 * nc ** 3 + 3 * (nc + 1) * nc * nc + 8 * nc ** 3
    else:
        return None
    return 8 * nb * per


def cpu_baseline(sample_blocks=(4, 4, 4), ncycles=2, min_seconds=10.0):
    """The oracle (CPU restatement, OpenMP over MeshBlocks) on a bounded sample of the
    same workload: 128^3 cells in 64 MeshBlocks of 32^3, ``ncycles`` VL2 cycles."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import OracleMesh, lib as oralib
    from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh
    n = 32 * sample_blocks[0]
    m = turbulence_mesh(n, 32)
    t = Turbulence(m)
    o = OracleMesh(m)
    for g in range(m.nblocks):
        b = t.block(g)
        for k in ("u", "b1", "b2", "b3"):
            o.set(g, k, b[k])
    o.initialize()
    t0 = time.perf_counter()
    done = 0
    while True:                          # at least `ncycles`, then until ~10 s of work
        o.run(1)
        done += 1
        dt = time.perf_counter() - t0
        if done >= ncycles and dt >= min_seconds:
            break
    zc = done * m.nx[0] * m.nx[1] * m.nx[2]
    return {"value": zc / dt, "unit": "zone-cycles/s", "cores": int(oralib().ora_threads()),
            "kind": "port",
            "sample": "%d^3 cells as %d MeshBlocks of 32^3, %d VL2 cycles (%.2f s)"
                      % (n, m.nblocks, done, dt),
            "calibration": {"oracle_over_reference": ORACLE_OVER_REFERENCE,
                            "basis": CALIBRATION_BASIS,
                            "reference_equivalent_value": zc / dt / ORACLE_OVER_REFERENCE}}


C2P_BYTES_PER_SOLVE = 112   # read D, M, E, Bcc and a seed; write 5 primitives
C2P_METRIC = ("SR con2prim cell-solves/s on 1 MI355X at the 1e-12 tolerance, learned "
              "guess (BF16 MFMA MLP + FP64 Newton, plain fallback)")


def bench_sr_c2p(args):
    """Batched SR primitive recovery of 128^3 states from the learned-guess distribution,
    cold (no usable previous pressure), in the three solver modes."""
    import torch
    from athena_amd import gpu, ml
    from athena_amd.mesh import DEFAULT_FLOOR as FL
    n = args.c2p_n ** 3
    net = ml.load_mlp()
    cons, bcc, _ = ml.conserved_batch(n, 20261203)
    dev = torch.device("cuda:0")
    C0 = torch.from_numpy(cons).to(dev)
    B = torch.from_numpy(bcc).to(dev)
    C, P = torch.empty_like(C0), torch.empty_like(C0)
    it = torch.empty(n, dtype=torch.int32, device=dev)
    evs, fbs = torch.empty_like(it), torch.empty_like(it)
    ev = HipEvents()
    res = {}
    # A batch launches K calls back to back, each on its own copy of the states (filled
    # beforehand, cold seed), so no host latency sits between them and no call finds its
    # input in the caches.  Untimed batches over every mode run first for a fixed wall
    # time: the GPU clocks need ~150 calls of ~0.2 ms to settle (a cold call runs ~25 %
    # slower), which a few warm-up calls per mode would not cover.
    Cs = [torch.empty_like(C0) for _ in range(args.steps)]
    Ps = [torch.empty_like(C0) for _ in range(args.steps)]

    def batch(mode):
        for Ci, Pi in zip(Cs, Ps):
            Ci.copy_(C0)
            Pi.zero_()                      # cold: the seed register holds nothing
        torch.cuda.synchronize()
        e0, e1 = ev.new(), ev.new()
        ev.record(e0, 0)
        for Ci, Pi in zip(Cs, Ps):
            gpu.sr_con2prim_device(mode, n, Ci.data_ptr(), Pi.data_ptr(), B.data_ptr(),
                                   ml.GAMMA_AD, FL, FL, mlp=net)
        ev.record(e1, 0)
        ms = ev.elapsed_ms(e0, e1)
        ev.destroy(e0)
        ev.destroy(e1)
        return ms

    t_end = time.perf_counter() + args.c2p_warm_s
    while time.perf_counter() < t_end:
        for mode in (0, 1, 2):
            batch(mode)
    for mode, name in ((0, "reference"), (1, "newton"), (2, "learned")):
        # untimed call with the per-cell counters
        C.copy_(C0)
        P.zero_()
        gpu.sr_con2prim_device(mode, n, C.data_ptr(), P.data_ptr(), B.data_ptr(),
                               ml.GAMMA_AD, FL, FL, mlp=net, iters=it.data_ptr(),
                               evals=evs.data_ptr(), fallback=fbs.data_ptr())
        for r in range(1 + args.warmup):    # warm-up batches, then the timed one
            total = batch(mode)
        e = evs.cpu().numpy().astype(np.float64)
        res[name] = {"cell_solves_per_s": n * args.steps / (total * 1e-3),
                     "ms_per_call": total / args.steps,
                     "mean_evaluations": float(e.mean()),
                     "p99_evaluations": float(np.percentile(e, 99)),
                     "fallback_rate": float(fbs.cpu().numpy().mean()),
                     "failed_rate": float((it.cpu().numpy() < 0).mean())}
        print("  %-9s %.4f ms/call  %.3e solves/s  mean evals %.3f p99 %d"
              % (name, res[name]["ms_per_call"], res[name]["cell_solves_per_s"],
                 e.mean(), np.percentile(e, 99)), file=sys.stderr, flush=True)
    lr = res["learned"]
    achieved = C2P_BYTES_PER_SOLVE * lr["cell_solves_per_s"]
    roof = {"kernel": "k_mlp_seed + k_sr_c2p_modes<2> (one call)", "bound": "hbm",
            "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": None,
            "avg_launch_ms": lr["ms_per_call"],
            "algorithmic_bytes_per_launch": C2P_BYTES_PER_SOLVE * n,
            "note": "FP64-VALU bound (transcendentals and iteration count), not HBM"}
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_bind import sr_con2prim as ora_c2p, lib as oralib
        t0 = time.perf_counter()
        reps = 0
        while True:                         # about 10 s of host work
            ora_c2p(cons, np.zeros((n, 5)), bcc, ml.GAMMA_AD, FL, FL)
            reps += 1
            dt = time.perf_counter() - t0
            if dt > 10.0 or reps >= 200:
                break
        cpu = {"value": reps * n / dt, "unit": "cell-solves/s",
               "cores": int(oralib().ora_threads()), "kind": "port",
               "sample": "the same %d states x %d passes, reference iteration from a "
                         "cold seed (%.1f s)" % (n, reps, dt)}
    out = {"metric": C2P_METRIC, "value": lr["cell_solves_per_s"], "unit": "cell-solves/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": lr["ms_per_call"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 (solve), bf16 MFMA (guess)",
           "data": "synthetic",
           "config": {"workload": "SR MHD con2prim, %d^3 states: Lorentz factor log-uniform "
                                  "[1,10], sigma=b^2/rho log-uniform [0.01,10], rho "
                                  "log-uniform [1e-3,1], p/rho log-uniform [1e-2,10], "
                                  "gamma 4/3, isotropic directions, cold seed"
                                  % args.c2p_n, "states": n},
           "modes": res,
           "speedup_vs_reference": lr["cell_solves_per_s"] / res["reference"]["cell_solves_per_s"],
           "speedup_vs_newton": lr["cell_solves_per_s"] / res["newton"]["cell_solves_per_s"],
           "evaluations_saved_vs_reference": res["reference"]["mean_evaluations"] - lr["mean_evaluations"],
           "evaluations_saved_vs_newton": res["newton"]["mean_evaluations"] - lr["mean_evaluations"],
           "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


ENS_METRIC = ("zone-cycles/s on 1 MI355X, ensemble of 48 MHD blast waves of uneven sizes "
              "(VL2+PLM+HLLD FP64, one persistent launch per stage kernel, HIP graph)")
KIND_KERNELS = ("flux_xyz_dc", "corner_emf", "update_u_b_ct", "cons2prim")


def bench_ensemble(args):
    from athena_amd.gpu import Ensemble, GpuMesh
    from athena_amd.pgen.blast import blast_ensemble
    t0 = time.perf_counter()
    members = blast_ensemble(args.members)
    meshes = []
    for m, ic in members:
        d = GpuMesh(m)
        d.load([ic.block(g) for g in range(m.nblocks)])
        d.initialize()
        meshes.append(d)
    ens = Ensemble(meshes)
    nmem, cells = ens.size()
    print("setup %.1f s: %d members, %d active cells" % (time.perf_counter() - t0, nmem,
                                                         cells), file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    best = None if args.no_autotune else ens.autotune(reps=3)
    tiles, persist = ens.tiles()
    print("autotune %.1f s: tiles %s persist %s" % (time.perf_counter() - t1,
                                                     [Ensemble.TILES[t] for t in tiles],
                                                     list(persist)), file=sys.stderr, flush=True)
    ens.run(args.warmup)
    ev = HipEvents()
    stream = ens.stream()
    e0, e1 = ev.new(), ev.new()
    ev.record(e0, stream)
    ens.run(args.steps)
    ev.record(e1, stream)
    ms = ev.elapsed_ms(e0, e1)
    zcs = args.steps * cells / (ms * 1e-3)
    print("timed %d cycles: %.3f ms/cycle, %.3e zone-cycles/s" % (args.steps, ms / args.steps,
                                                                   zcs), file=sys.stderr,
          flush=True)
    # per-kind launch times measured by the autotuner (stage-1 launches over all members;
    # the update kind is the update followed by the ghost exchange)
    kinds, roof = {}, None
    for k, name in enumerate(KIND_KERNELS if best is not None else ()):
        ab = sum(kernel_algorithmic_bytes(name, m) for m, _ in members)
        if name == "update_u_b_ct":
            name = "update_u_b_ct + ghost_exchange"
            ab += sum(kernel_algorithmic_bytes("ghost_exchange", m) for m, _ in members)
        kinds[name] = {"avg_launch_ms": best[k], "algorithmic_bytes_per_launch": ab,
                       "achieved_GBps": ab / (best[k] * 1e-3) / 1e9}
    if kinds:
        dom = max(kinds, key=lambda k: kinds[k]["avg_launch_ms"])
        roof = {"kernel": dom + " (ensemble launch)", "bound": "hbm",
                "achieved": kinds[dom]["achieved_GBps"], "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": kinds[dom]["achieved_GBps"] * 1e9 / HBM_PEAK,
                "traffic": None, "avg_launch_ms": kinds[dom]["avg_launch_ms"],
                "algorithmic_bytes_per_launch": kinds[dom]["algorithmic_bytes_per_launch"]}
    out = {"metric": ENS_METRIC, "value": zcs, "unit": "zone-cycles/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "%d 3D MHD blast waves, edges from {32,48,64,96,128}, "
                                  "MeshBlocks 16^3 or 32^3, P_in/P_out log-uniform "
                                  "[10,1000], |B0| uniform [0.1,1] isotropic, VL2+PLM+HLLD+CT"
                                  % nmem, "members": nmem, "cells": cells,
                      "tiles": [Ensemble.TILES[t] for t in tiles],
                      "persistent_grid": [int(p) for p in persist]},
           "kernels": kinds, "roofline": roof,
           "step_roofline": {"bytes_per_zone_cycle": BYTES_PER_ZONE_CYCLE,
                             "achieved": zcs * BYTES_PER_ZONE_CYCLE / 1e9, "unit": "GB/s",
                             "frac": zcs * BYTES_PER_ZONE_CYCLE / HBM_PEAK},
           "cpu_baseline": None}
    print(json.dumps(out), flush=True)


SHEAR_METRIC = ("zone-cycles/s on 1 MI355X, 3D SR MHD shear layer (VL2+PPM+HLLE+CT FP64, "
                "learned con2prim guess)")


def bench_sr_shear(args):
    """The SR shear layer at 256^3 in 32^3 MeshBlocks (NGHOST 3, PPM + HLLE + CT), run in
    each primitive-recovery mode from the same initial state for the same cycles."""
    from athena_amd.gpu import GpuMesh, C2P_MODES
    from athena_amd.pgen.sr import ShearLayer, shear_mesh
    from athena_amd import ml
    m = shear_mesh(args.n, args.bn)
    ic = ShearLayer(m)
    blocks = [ic.block(g) for g in range(m.nblocks)]
    net = ml.load_mlp()
    cells = m.nx[0] * m.nx[1] * m.nx[2]
    ev = HipEvents()
    runs, finals = {}, {}
    for name in ("learned", "reference", "newton"):
        d = GpuMesh(m)
        d.load(blocks)
        d.set_c2p(C2P_MODES[name], net if name == "learned" else None)
        d.initialize()
        d.set_fused(args.fused_rows)
        d.run(args.warmup)
        stream = d.stream()
        e0, e1 = ev.new(), ev.new()
        d.synchronize()
        ev.record(e0, stream)
        d.run(args.steps)
        ev.record(e1, stream)
        ms = ev.elapsed_ms(e0, e1)
        # per-kernel events, then the recovery counters (atomics), each in cycles of its
        # own outside the timed ones
        d.kernel_timing(True)
        d.run(2)
        stats = d.kernel_stats()
        d.kernel_timing(False)
        d.c2p_stats(1)
        d.run(2)
        evals, fb, nsolve = d.c2p_stats(0)
        zcs = args.steps * cells / (ms * 1e-3)
        runs[name] = {"zone_cycles_per_s": zcs, "ms_per_step": ms / args.steps,
                      "evaluations_per_solve": evals / max(nsolve, 1),
                      "fallback_rate": fb / max(nsolve, 1), "solves": nsolve,
                      "kernels_ms_per_cycle": {k: v[0] / 2 for k, v in stats.items()},
                      "kernel_launches": {k: v[1] for k, v in stats.items()}}
        finals[name] = [d.get(g, "u") for g in range(m.nblocks)]
        print("  %-9s %.3f ms/cycle %.3e zc/s, %.3f evaluations/solve, fallback %.2e"
              % (name, ms / args.steps, zcs, evals / max(nsolve, 1), fb / max(nsolve, 1)),
              file=sys.stderr, flush=True)
        d.close()
    # config 5 as stated: the learned guess inside persistent ensemble kernels, each cycle
    # replayed as one HIP graph (a one-member ensemble), from the same state for the same
    # cycles as the eager learned run, whose final state it must equal bit for bit
    from athena_amd.gpu import Ensemble
    d = GpuMesh(m)
    d.load(blocks)
    d.set_c2p(C2P_MODES["learned"], net)
    d.initialize()
    ens = Ensemble([d])
    ens.run(args.warmup, graph=True)
    stream = ens.stream()
    e0, e1 = ev.new(), ev.new()
    d.synchronize()
    ev.record(e0, stream)
    ens.run(args.steps, graph=True)
    ev.record(e1, stream)
    ms = ev.elapsed_ms(e0, e1)
    ens.run(4, graph=True)                  # the eager runs' timing and counter cycles
    same = all(np.array_equal(d.get(g, "u"), finals["learned"][g]) for g in range(m.nblocks))
    zcs = args.steps * cells / (ms * 1e-3)
    runs["learned_ensemble_graph"] = {"zone_cycles_per_s": zcs, "ms_per_step": ms / args.steps,
                                      "persistent_kernels": True, "hip_graph": True,
                                      "bitwise_equal_to_eager_learned": same}
    print("  %-9s %.3f ms/cycle %.3e zc/s, learned guess in ensemble kernels replayed as a "
          "HIP graph, bitwise equal to the eager run: %s" % ("graph", ms / args.steps, zcs,
                                                             same),
          file=sys.stderr, flush=True)
    ens.close()
    d.close()
    ng = m.ng
    diffs = {}
    for name in ("learned", "newton"):
        num = den = 0.0
        for a, b in zip(finals[name], finals["reference"]):
            sa = a[:, ng:-ng, ng:-ng, ng:-ng]
            sb = b[:, ng:-ng, ng:-ng, ng:-ng]
            num += np.abs(sa - sb).sum(axis=(1, 2, 3))
            den += np.abs(sb).sum(axis=(1, 2, 3))
        diffs[name] = [float(x) for x in num / den]
    lr = runs["learned"]
    kms = lr["kernels_ms_per_cycle"]
    dom = max(kms, key=kms.get)
    ab = kernel_algorithmic_bytes(dom, m)
    avg = kms[dom] * 2 / lr["kernel_launches"][dom]
    roof = {"kernel": dom, "bound": "hbm",
            "achieved": ab / (avg * 1e-3) / 1e9 if ab else None, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": ab / (avg * 1e-3) / HBM_PEAK if ab else None,
            "traffic": None, "avg_launch_ms": avg, "algorithmic_bytes_per_launch": ab}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_sr()
    gr = runs["learned_ensemble_graph"]
    out = {"metric": SHEAR_METRIC, "value": gr["zone_cycles_per_s"], "unit": "zone-cycles/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": gr["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 (bf16 MFMA guess)", "data": "synthetic",
           "config": {"workload": "3D SR MHD shear layer %d^3 as %d MeshBlocks of %d^3, "
                                  "NGHOST 3, v_x=+-0.5 tanh width 0.01, v_y noise 1e-3, "
                                  "rho=1, P=1, B_x=0.1, VL2+PPM+HLLE+CT, learned "
                                  "recovery guess, persistent ensemble kernels, HIP-graph "
                                  "cycle" % (args.n, m.nblocks, args.bn), "cells": cells},
           "modes": runs, "rel_l1_vs_reference_mode": diffs,
           "speedup_learned_vs_reference": lr["zone_cycles_per_s"]
           / runs["reference"]["zone_cycles_per_s"],
           "roofline": roof,
           "step_roofline": {"bytes_per_zone_cycle": BYTES_PER_ZONE_CYCLE,
                             "achieved": lr["zone_cycles_per_s"] * BYTES_PER_ZONE_CYCLE / 1e9,
                             "unit": "GB/s",
                             "frac": lr["zone_cycles_per_s"] * BYTES_PER_ZONE_CYCLE / HBM_PEAK},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


SMR_METRIC = ("zone-cycles/s on 1 MI355X, 3D MHD on a statically refined mesh "
              "(VL2+PLM+HLLD+CT FP64, two levels, flux and EMF correction)")
# <refinement1..5> of inputs/mhd/athinput.linear_wave3d (the reference's SMR regression mesh)
SMR_REGIONS = [((0.3, 0.3, 0.3), (0.4, 0.4, 0.4), 1), ((0.3, 1.0, 1.0), (0.4, 1.1, 1.1), 1),
               ((1.0, 1.0, 0.3), (1.1, 1.1, 0.4), 1), ((1.0, 0.3, 1.0), (1.1, 0.4, 1.1), 1),
               ((1.9, 0.3, 0.3), (2.0, 1.2, 1.2), 1)]


def bench_smr(args):
    """The reference's SMR linear-wave mesh (athinput.linear_wave3d with its five refined
    regions) on a root grid of n x n/2 x n/2 cells in MeshBlocks of bn^3: zone-cycles over
    the active cells of every level, and the share of the cycle spent in the refinement
    programs (flux correction, EMF correction, boundary exchange with prolongation)."""
    from athena_amd.gpu import GpuMesh
    from athena_amd.mesh import refined
    from athena_amd.pgen.linear_wave import LinearWave, linear_wave_mesh
    t0 = time.perf_counter()
    m = refined(linear_wave_mesh(args.n, 0, bnx=(args.bn,) * 3), SMR_REGIONS)
    lw = LinearWave(m, wave_flag=0)
    d = GpuMesh(m)
    for g in range(m.nblocks):
        b = lw.block(g)
        for k in ("u", "b1", "b2", "b3"):
            d.set(g, k, b[k])
    d.initialize()
    levels = np.bincount(np.asarray(m.locs)[:, 0] - m.root_level).tolist()
    cells = m.nblocks * args.bn ** 3
    print("setup %.1f s: %d MeshBlocks of %d^3 (per level from the root: %s), %d cells"
          % (time.perf_counter() - t0, m.nblocks, args.bn, levels, cells), file=sys.stderr,
          flush=True)
    d.set_graph(not args.no_graph)
    d.run(args.warmup)
    ev = HipEvents()
    stream = d.stream()
    e0, e1 = ev.new(), ev.new()
    d.synchronize()
    ev.record(e0, stream)
    d.run(args.steps)
    ev.record(e1, stream)
    ms = ev.elapsed_ms(e0, e1)
    zcs = args.steps * cells / (ms * 1e-3)
    print("timed %d cycles: %.3f ms/cycle, %.3e zone-cycles/s" % (args.steps, ms / args.steps,
                                                                   zcs), file=sys.stderr,
          flush=True)
    d.kernel_timing(True)
    d.run(args.steps)
    stats = d.kernel_stats()
    d.kernel_timing(False)
    tot = sum(v[0] for v in stats.values())
    for k, (t, c) in sorted(stats.items(), key=lambda kv: -kv[1][0]):
        print("  %-20s %8.3f ms total %5d launches %8.4f ms avg %5.1f%%"
              % (k, t, c, t / c, 100 * t / tot), file=sys.stderr)
    smr_ms = sum(v[0] for k, v in stats.items() if k.startswith("smr_"))
    dom = max(stats, key=lambda k: stats[k][0])
    dom_ms, dom_n = stats[dom]
    abytes = kernel_algorithmic_bytes(dom, m)
    roof = {"kernel": dom, "bound": "hbm",
            "achieved": abytes / (dom_ms / dom_n * 1e-3) / 1e9 if abytes else None,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": abytes / (dom_ms / dom_n * 1e-3) / HBM_PEAK if abytes else None,
            "traffic": None, "avg_launch_ms": dom_ms / dom_n,
            "algorithmic_bytes_per_launch": abytes}
    out = {"metric": SMR_METRIC, "value": zcs, "unit": "zone-cycles/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "MHD fast linear wave on athinput.linear_wave3d's refined "
                                  "mesh, root %dx%dx%d, MeshBlocks %d^3, five level-1 "
                                  "regions" % (args.n, args.n // 2, args.n // 2, args.bn),
                      "meshblocks": m.nblocks, "meshblocks_per_level": levels,
                      "cells": cells, "hip_graph_cycle": not args.no_graph},
           "kernels_ms_per_cycle": {k: v[0] / args.steps for k, v in stats.items()},
           "smr_program_share": smr_ms / tot,
           "roofline": roof, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def cpu_baseline_sr(n=128, ncycles=2, min_seconds=10.0):
    """The oracle on a bounded sample of the SR shear layer: 128^3 as 64 MeshBlocks of
    32^3, PPM + HLLE, reference primitive recovery."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import OracleMesh, lib as oralib
    from athena_amd.pgen.sr import ShearLayer, shear_mesh
    m = shear_mesh(n, 32)
    ic = ShearLayer(m)
    o = OracleMesh(m)
    o.load([ic.block(g) for g in range(m.nblocks)])
    o.initialize()
    t0 = time.perf_counter()
    done = 0
    while True:
        o.run(1)
        done += 1
        dt = time.perf_counter() - t0
        if done >= ncycles and dt >= min_seconds:
            break
    return {"value": done * n ** 3 / dt, "unit": "zone-cycles/s",
            "cores": int(oralib().ora_threads()), "kind": "port",
            "sample": "%d^3 cells as %d MeshBlocks of 32^3, %d VL2 cycles (%.2f s)"
                      % (n, m.nblocks, done, dt),
            "calibration": {"oracle_over_reference": ORACLE_OVER_REFERENCE,
                            "basis": CALIBRATION_BASIS,
                            "reference_equivalent_value":
                                done * n ** 3 / dt / ORACLE_OVER_REFERENCE}}


def copy_calibration():
    """Three plain device copies of 512 MiB: under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    they give the factors that turn those counters into bytes moved."""
    import torch
    a = torch.ones(64 << 20, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied %d bytes x 3" % (a.numel() * 8), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="turbulence",
                    choices=["turbulence", "sr_c2p", "ensemble", "sr_shear", "smr", "copy"],
                    help="turbulence: the zone-cycles/s throughput case (default); "
                         "sr_c2p: batched SR primitive recovery with the learned guess; "
                         "ensemble: 48 blast waves of uneven sizes; sr_shear: the SR shear "
                         "layer with the learned guess (use --steps 50); smr: the "
                         "reference's statically refined linear-wave mesh; copy: three "
                         "512 MiB device copies, the calibration of the HBM counters")
    ap.add_argument("--members", type=int, default=48, help="ensemble members")
    ap.add_argument("--n", type=int, default=256, help="mesh cells per direction")
    ap.add_argument("--bn", type=int, default=32, help="MeshBlock cells per direction")
    ap.add_argument("--c2p-n", type=int, default=128, help="states per direction (sr_c2p)")
    ap.add_argument("--c2p-warm-s", type=float, default=1.0,
                    help="seconds of untimed calls over every mode before timing (sr_c2p)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default thread-to-cell mapping (turbulence)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every kernel of a cycle eagerly instead of replaying the "
                         "cycle as one captured HIP graph (turbulence)")
    ap.add_argument("--flux-tile", type=int, default=-1,
                    help="flux kernel tile index after tuning (turbulence; 6 and 7: the "
                         "z-march over segments of 8 and 16 planes; -1: keep)")
    ap.add_argument("--fused-rows", type=int, default=0,
                    help="rows per band of the optional fused flux + divergence kernel "
                         "(turbulence, sr_shear; 0: separate flux and update kernels, the "
                         "default; -1: the tallest band a work-group holds)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported "
              "(run with --gpus 1)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "sr_c2p":
        return bench_sr_c2p(args)
    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "sr_shear":
        return bench_sr_shear(args)
    if args.workload == "copy":
        return copy_calibration()
    if args.workload == "smr":
        return bench_smr(args)

    from athena_amd.gpu import GpuMesh
    from athena_amd.pgen.turbulence import Turbulence, turbulence_mesh

    m = turbulence_mesh(args.n, args.bn)
    t0 = time.perf_counter()
    ic = Turbulence(m)
    d = GpuMesh(m)
    for g in range(m.nblocks):
        b = ic.block(g)
        for k in ("u", "b1", "b2", "b3"):
            d.set(g, k, b[k])
    del ic
    d.initialize()
    d.set_fused(args.fused_rows)
    print("setup %.1f s: %d MeshBlocks of %s, fused flux + divergence rows %d"
          % (time.perf_counter() - t0, m.nblocks, m.bnx, d.fused()), file=sys.stderr,
          flush=True)
    if not args.no_autotune:
        d.autotune(reps=3)
    if args.flux_tile >= 0:
        d.set_tiles([args.flux_tile] + list(d.tiles())[1:])
    tiles = [TILE_NAMES[t] for t in d.tiles()]
    print("%s tiles (flux, corner, update, cons2prim): %s"
          % ("default" if args.no_autotune else "autotuned", tiles), file=sys.stderr,
          flush=True)
    stream = d.stream()
    ev = HipEvents()
    d.set_graph(not args.no_graph)

    d.run(args.warmup)
    d.synchronize()

    # timed region: K cycles bracketed by events on the library stream
    e0, e1 = ev.new(), ev.new()
    d.synchronize()
    ev.record(e0, stream)
    d.run(args.steps)
    ev.record(e1, stream)
    ms = ev.elapsed_ms(e0, e1)
    d.synchronize()
    cells = m.nx[0] * m.nx[1] * m.nx[2]
    zcs = args.steps * cells / (ms * 1e-3)
    print("timed %d cycles: %.3f ms/cycle, %.3e zone-cycles/s" % (args.steps, ms / args.steps,
                                                                   zcs), file=sys.stderr,
          flush=True)

    # per-kernel durations over a second timed pass (events around every launch)
    d.kernel_timing(True)
    d.run(args.steps)
    stats = d.kernel_stats()
    d.kernel_timing(False)
    tot = sum(v[0] for v in stats.values())
    dom = max(stats, key=lambda k: stats[k][0])
    dom_ms, dom_n = stats[dom]
    dom_avg_s = dom_ms / dom_n * 1e-3
    abytes = kernel_algorithmic_bytes(dom, m)
    for k, (t, c) in sorted(stats.items(), key=lambda kv: -kv[1][0]):
        print("  %-14s %8.3f ms total %5d launches %8.4f ms avg %5.1f%%"
              % (k, t, c, t / c, 100 * t / tot), file=sys.stderr)
    measured = (args.n, args.bn) == (256, 32) and tiles == TRAFFIC_TILES
    roof = {"kernel": dom, "bound": "hbm",
            "achieved": abytes / dom_avg_s / 1e9 if abytes else None,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": abytes / dom_avg_s / HBM_PEAK if abytes else None,
            "traffic": MEASURED_TRAFFIC.get(dom) if measured else None,
            "avg_launch_ms": dom_ms / dom_n,
            "algorithmic_bytes_per_launch": abytes}
    if roof["traffic"]:
        roof["traffic_GBps"] = roof["traffic"] / dom_avg_s / 1e9
        roof["traffic_source"] = ("%s: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this "
                                  "workload (256^3 as 512 MeshBlocks of 32^3, tiles %s), "
                                  "not counters of this run" % (TRAFFIC_SOURCE, TRAFFIC_TILES))
    if measured and dom in MEASURED_FP64_FLOP:
        fl = MEASURED_FP64_FLOP[dom]
        roof["fp64_valu"] = {"flop_per_launch": fl, "achieved": fl / dom_avg_s / 1e12,
                             "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                             "frac": fl / dom_avg_s / FP64_PEAK,
                             "flop_source": "%s: SQ_INSTS_VALU_*_F64 counters of this "
                                            "workload, not of this run" % FP64_SOURCE}
    step_roof = {"bytes_per_zone_cycle": BYTES_PER_ZONE_CYCLE,
                 "achieved": zcs * BYTES_PER_ZONE_CYCLE / 1e9, "unit": "GB/s",
                 "frac": zcs * BYTES_PER_ZONE_CYCLE / HBM_PEAK}
    if measured and all(k in MEASURED_TRAFFIC for k in stats if k != "new_dt"):
        # the HBM bytes the cycle's launches actually move (the profile's counters times
        # this run's launches per cycle) over this run's cycle time: how busy HBM is,
        # beside the compulsory 536-byte figure above
        per_cycle = sum(MEASURED_TRAFFIC[k] * c / args.steps for k, (_, c) in stats.items()
                        if k != "new_dt")
        step_roof["measured_traffic"] = {
            "bytes_per_cycle": per_cycle,
            "achieved": per_cycle / (ms / args.steps * 1e-3) / 1e9, "unit": "GB/s",
            "frac": per_cycle / (ms / args.steps * 1e-3) / HBM_PEAK,
            "source": "per-launch FETCH_SIZE/WRITE_SIZE of %s" % TRAFFIC_SOURCE}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()

    out = {"metric": METRIC, "value": zcs, "unit": "zone-cycles/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "3D ideal MHD %d^3 periodic as %d MeshBlocks of %d^3, "
                                  "decaying turbulence (seeded k^-2 velocity, rms Mach 1), "
                                  "VL2+PLM+HLLD+CT" % (args.n, m.nblocks, args.bn),
                      "cells": cells, "meshblocks": m.nblocks, "nghost": m.ng,
                      "tiles": tiles, "fused_rows": d.fused(),
                      "hip_graph_cycle": not args.no_graph},
           "roofline": roof, "step_roofline": step_roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
