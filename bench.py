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

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value


def kernel_algorithmic_bytes(name, m):
    """Compulsory HBM bytes of one launch of a kernel over all blocks (doubles touched
    once: inputs read + outputs written), per the same accounting as the 536 B/zone-
    cycle figure.  Only the flux kernels are needed for the roofline line."""
    n1, n2, n3 = m.bnx
    nb = m.nblocks
    if name.startswith("flux_x"):
        d = int(name[6]) - 1
        order = name.split("_")[2]
        ext = {"dc": (1, 0), "plm": (2, 1), "ppm": (3, 2)}[order]   # stencil below/above
        faces = [n1 + 1, n2 + 2, n3 + 2] if d == 0 else ([n1 + 2, n2 + 1, n3 + 2] if d == 1
                                                        else [n1 + 2, n2 + 2, n3 + 1])
        cells = list(faces)
        cells[d] = [n1, n2, n3][d] + ext[0] + ext[1]
        nface = faces[0] * faces[1] * faces[2]
        ncell = cells[0] * cells[1] * cells[2]
        # reads: w(5) + bcc(3) per cell, bx per face; writes: flux(5) + 2 EMF + weight
        return 8 * nb * (8 * ncell + 1 * nface + 8 * nface)
    return None


def cpu_baseline(sample_blocks=(4, 4, 4), ncycles=2):
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
    o.run(ncycles)
    dt = time.perf_counter() - t0
    zc = ncycles * m.nx[0] * m.nx[1] * m.nx[2]
    return {"value": zc / dt, "unit": "zone-cycles/s", "cores": int(oralib().ora_threads()),
            "kind": "port",
            "sample": "%d^3 cells as %d MeshBlocks of 32^3, %d VL2 cycles (%.2f s)"
                      % (n, m.nblocks, ncycles, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="mesh cells per direction")
    ap.add_argument("--bn", type=int, default=32, help="MeshBlock cells per direction")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported "
              "(run with --gpus 1)" % args.gpus, file=sys.stderr)
        sys.exit(2)

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
    print("setup %.1f s: %d MeshBlocks of %s" % (time.perf_counter() - t0, m.nblocks,
                                                 m.bnx), file=sys.stderr, flush=True)
    stream = d.stream()
    ev = HipEvents()

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
    roof = {"kernel": dom, "bound": "hbm",
            "achieved": abytes / dom_avg_s / 1e9 if abytes else None,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": abytes / dom_avg_s / HBM_PEAK if abytes else None,
            "traffic": None,
            "avg_launch_ms": dom_ms / dom_n,
            "algorithmic_bytes_per_launch": abytes}
    step_roof = {"bytes_per_zone_cycle": BYTES_PER_ZONE_CYCLE,
                 "achieved": zcs * BYTES_PER_ZONE_CYCLE / 1e9, "unit": "GB/s",
                 "frac": zcs * BYTES_PER_ZONE_CYCLE / HBM_PEAK}

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
                      "cells": cells, "meshblocks": m.nblocks, "nghost": m.ng},
           "roofline": roof, "step_roofline": step_roof, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
