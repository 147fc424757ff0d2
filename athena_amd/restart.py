"""Restart files (.rst) in the reference's binary layout, for checkpoint/resume of a
device mesh and for exchanging states with the reference.

Layout written by RestartOutput::WriteOutputFile (outputs/restart.cpp:36-240) and read
by the restart Mesh constructor (mesh/mesh.cpp:619-1010), for MHD without GR, passive
scalars, radiation or user data:

* the input parameters as ParameterInput::ParameterDump prints them
  (parameter_input.cpp:835-860), ending with ``<par_end>``;
* int nbtotal, int root_level, RegionSize mesh_size (9 doubles + 3 ints, padded to 8
  bytes, athena.hpp:116-122), double time, double dt, int ncycle, uint64 datasize;
* the ID list: per MeshBlock a LogicalLocation (3 int64 + int level, padded,
  athena.hpp:74-79) and a double cost;
* per MeshBlock, in the same order: u, b.x1f, b.x2f, b.x3f as AthenaArray data.

MeshBlocks appear in the reference's global-id order, the Z-order traversal of the
block tree (meshblock_tree.cpp:87-110, 336-352); the device library numbers the blocks
of a uniform mesh row-major, so writing and reading map between the two orders.  A
statically refined mesh (``MeshDesc.locs``) already numbers its blocks in the tree's
order, with each block's level in its LogicalLocation, and its parameters say
``refinement = static``, as the reference writes them."""
from __future__ import annotations

import struct

import numpy as np

from .mesh import MeshDesc

REGION = struct.Struct("<9d3i4x")          # RegionSize
LOCCOST = struct.Struct("<3qi4xd")        # LogicalLocation + cost
assert REGION.size == 88 and LOCCOST.size == 40


def root_level(nbx):
    """Smallest L with 2^L >= the largest block count (mesh.cpp:306)."""
    nb = max(nbx)
    lev = 0
    while (1 << lev) < nb:
        lev += 1
    return lev


def ztree_order(nbx):
    """(l1, l2, l3) of every MeshBlock in the reference's global-id order: children of a
    tree node are visited i fastest, then j, then k, at every level."""
    lev = root_level(nbx)

    def key(loc):
        k = 0
        for b in range(lev):
            for d in range(3):
                k |= ((loc[d] >> b) & 1) << (3 * b + d)
        return k

    locs = [(l1, l2, l3) for l3 in range(nbx[2]) for l2 in range(nbx[1])
            for l1 in range(nbx[0])]
    return sorted(locs, key=key)


def block_order(m: MeshDesc):
    """((level, l1, l2, l3) of every MeshBlock in the reference's global-id order, the
    device gid of each, the root level)."""
    if m.smr:
        locs = [tuple(int(v) for v in r) for r in m.locs]
        return locs, list(range(m.nblocks)), int(m.root_level)
    lev = root_level(m.nbx)
    order = ztree_order(m.nbx)
    return [(lev,) + tuple(loc) for loc in order], [m.gid(*loc) for loc in order], lev


def default_params(m: MeshDesc, tlim=1.0e300, problem_id="gpu"):
    """Input blocks describing m, in the reference's athinput vocabulary."""
    mesh = {}
    for d in (1, 2, 3):
        mesh["nx%d" % d] = str(m.nx[d - 1])
        mesh["x%dmin" % d] = repr(float(m.xmin[d - 1]))
        mesh["x%dmax" % d] = repr(float(m.xmax[d - 1]))
        mesh["ix%d_bc" % d] = m.bc[2 * d - 2]
        mesh["ox%d_bc" % d] = m.bc[2 * d - 1]
    if m.smr:
        mesh["refinement"] = "static"
    mb = {"nx%d" % d: str(m.bnx[d - 1]) for d in (1, 2, 3)}
    hydro = {"gamma": repr(float(m.gamma))}
    dfl, pfl = m.floors
    hydro["dfloor"] = repr(float(dfl))
    hydro["pfloor"] = repr(float(pfl))
    time = {"cfl_number": repr(float(m.cfl)), "nlim": "-1", "tlim": repr(float(tlim)),
            "integrator": "vl2", "xorder": str(m.xorder)}
    return {"job": {"problem_id": problem_id}, "mesh": mesh, "meshblock": mb,
            "hydro": hydro, "time": time}


def param_dump(blocks):
    """ParameterInput::ParameterDump text: names and values padded to align."""
    out = ["#------------------------- PAR_DUMP -------------------------"]
    for name, lines in blocks.items():
        out.append("<%s>" % name)
        wn = max((len(k) for k in lines), default=0)
        wv = max((len(v) for v in lines.values()), default=0)
        for k, v in lines.items():
            out.append(k + " " * (wn - len(k) + 1) + "= " + v + " " * (wv - len(v) + 1))
    out.append("#------------------------- PAR_DUMP -------------------------")
    out.append("<par_end>")
    return "\n".join(out) + "\n"


def write(path, m: MeshDesc, get, time, dt, ncycle, params=None):
    """Writes the registers of mesh m (get(gid, what) -> array, row-major gids)."""
    text = param_dump(params if params is not None else default_params(m))
    order, gids, lev = block_order(m)
    shapes = [m.shape(k) for k in ("u", "b1", "b2", "b3")]
    datasize = 8 * sum(int(np.prod(s)) for s in shapes)
    with open(path, "wb") as f:
        f.write(text.encode())
        f.write(struct.pack("<ii", m.nblocks, lev))
        f.write(REGION.pack(*m.xmin, *m.xmax, 1.0, 1.0, 1.0, *m.nx))
        # written field by field as restart.cpp does: no padding between them
        f.write(struct.pack("<dd", time, dt) + struct.pack("<i", ncycle)
                + struct.pack("<Q", datasize))
        for level, l1, l2, l3 in order:
            f.write(LOCCOST.pack(l1, l2, l3, level, 1.0))
        for g in gids:
            for k in ("u", "b1", "b2", "b3"):
                f.write(np.ascontiguousarray(get(g, k), dtype="<f8").tobytes())


def read(path, m: MeshDesc):
    """Reads a restart file of mesh m: (header dict, {gid: {u, b1, b2, b3}}, params)."""
    raw = open(path, "rb").read()
    end = raw.index(b"<par_end>") + len(b"<par_end>\n")
    params = raw[:end].decode()
    p = end
    nbtotal, lev = struct.unpack_from("<ii", raw, p)
    p += 8
    region = REGION.unpack_from(raw, p)
    p += REGION.size
    time, dt = struct.unpack_from("<dd", raw, p)
    p += 16
    (ncycle,) = struct.unpack_from("<i", raw, p)
    p += 4
    (datasize,) = struct.unpack_from("<Q", raw, p)
    p += 8
    if nbtotal != m.nblocks or tuple(region[9:12]) != tuple(m.nx):
        raise ValueError("restart file does not describe this mesh")
    order, gids, _ = block_order(m)
    gid_of = dict(zip(order, gids))
    locs, keys = [], []
    for _ in range(nbtotal):
        l1, l2, l3, level, cost = LOCCOST.unpack_from(raw, p)
        if (level, l1, l2, l3) not in gid_of:
            raise ValueError("restart file lists a MeshBlock this mesh does not have")
        keys.append((level, l1, l2, l3))
        locs.append((level, l1, l2, l3) if m.smr else (l1, l2, l3))
        p += LOCCOST.size
    shapes = {k: m.shape(k) for k in ("u", "b1", "b2", "b3")}
    if datasize != 8 * sum(int(np.prod(s)) for s in shapes.values()):
        raise ValueError("restart block size does not match the mesh")
    blocks = {}
    for key in keys:
        b = {}
        for k in ("u", "b1", "b2", "b3"):
            n = int(np.prod(shapes[k]))
            b[k] = np.frombuffer(raw, dtype="<f8", count=n, offset=p).reshape(shapes[k]).copy()
            p += 8 * n
        blocks[gid_of[key]] = b
    head = {"nbtotal": nbtotal, "root_level": lev, "region": region, "time": time,
            "dt": dt, "ncycle": ncycle, "datasize": datasize, "locations": locs}
    return head, blocks, params
