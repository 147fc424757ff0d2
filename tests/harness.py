"""Shared helpers: run one problem on the oracle and/or the device library and compare
registers over the active region."""
from __future__ import annotations

import numpy as np

ACTIVE_KEYS = ("u", "w", "bcc", "b1", "b2", "b3")


def active(m, what, a):
    ng = m.ng
    n1, n2, n3 = m.bnx
    if what in ("u", "w", "bcc"):
        return a[:, ng:ng + n3, ng:ng + n2, ng:ng + n1]
    if what == "b1":
        return a[ng:ng + n3, ng:ng + n2, ng:ng + n1 + 1]
    if what == "b2":
        return a[ng:ng + n3, ng:ng + n2 + 1, ng:ng + n1]
    if what == "b3":
        return a[ng:ng + n3 + 1, ng:ng + n2, ng:ng + n1]
    raise KeyError(what)


def rel_l1(a, b):
    den = np.sum(np.abs(b))
    num = np.sum(np.abs(a - b))
    return num / den if den > 0 else num


def compare(m, get_a, get_b, keys=ACTIVE_KEYS, floor=0.0):
    """Per-variable relative L1 and bitwise-equal flags over all blocks.  With floor > 0,
    a component whose L1 norm is below `floor` times the largest norm of its register
    group (the components of u, w or bcc; the three face fields) holds only round-off of
    an analytically vanishing quantity: its difference is then measured relative to the
    group's largest norm instead of its own."""
    comps = {}
    groups = {}
    for k in keys:
        A = [active(m, k, get_a(g, k)) for g in range(m.nblocks)]
        B = [active(m, k, get_b(g, k)) for g in range(m.nblocks)]
        if k in ("u", "w", "bcc"):
            for n in range(A[0].shape[0]):
                name = "%s%d" % (k, n)
                comps[name] = (np.concatenate([x[n].ravel() for x in A]),
                               np.concatenate([x[n].ravel() for x in B]))
                groups[name] = k
        else:
            comps[k] = (np.concatenate([x.ravel() for x in A]),
                        np.concatenate([x.ravel() for x in B]))
            groups[k] = "b"
    norm = {name: float(np.sum(np.abs(b))) for name, (_, b) in comps.items()}
    big = {}
    for name, grp in groups.items():
        big[grp] = max(big.get(grp, 0.0), norm[name])
    out = {}
    for name, (a, b) in comps.items():
        den = norm[name]
        if floor > 0.0 and den < floor * big[groups[name]]:
            den = big[groups[name]]
        num = float(np.sum(np.abs(a - b)))
        out[name] = (num / den if den > 0 else num, bool(np.array_equal(a, b)))
    return out


def div_b(m, get):
    """max |div B| * dx over all blocks (CT keeps it at round-off)."""
    from athena_amd.mesh import block_coords
    worst = 0.0
    ng = m.ng
    n1, n2, n3 = m.bnx
    for g in range(m.nblocks):
        c = block_coords(m, g)
        b1, b2, b3 = get(g, "b1"), get(g, "b2"), get(g, "b3")
        ks, js, is_ = ng, ng, ng
        d1 = (b1[ks:ks + n3, js:js + n2, is_ + 1:is_ + n1 + 1] - b1[ks:ks + n3, js:js + n2, is_:is_ + n1]) / c.dx1f[is_]
        d2 = (b2[ks:ks + n3, js + 1:js + n2 + 1, is_:is_ + n1] - b2[ks:ks + n3, js:js + n2, is_:is_ + n1]) / c.dx2f[js]
        d3 = (b3[ks + 1:ks + n3 + 1, js:js + n2, is_:is_ + n1] - b3[ks:ks + n3, js:js + n2, is_:is_ + n1]) / c.dx3f[ks]
        worst = max(worst, float(np.max(np.abs(d1 + d2 + d3))) * c.dx1f[is_])
    return worst


def totals(m, get):
    """Volume sums of the conserved variables (history-style, outputs/history.cpp)."""
    from athena_amd.mesh import block_coords
    tot = np.zeros(5)
    for g in range(m.nblocks):
        c = block_coords(m, g)
        vol = c.dx1f[m.ng] * c.dx2f[m.ng] * c.dx3f[m.ng]
        u = active(m, "u", get(g, "u"))
        tot += u.reshape(5, -1).sum(axis=1) * vol
    return tot
