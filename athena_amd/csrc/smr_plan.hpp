//========================================================================================
// smr_plan.hpp -- host side of static mesh refinement: neighbour lists of a multilevel
// MeshBlock list and the per-MeshBlock step programs the device executes for
//   * the hydro flux correction (restricted fine fluxes on a coarse/fine face,
//     cc/flux_correction_cc.cpp:77-286),
//   * the EMF correction (same-level sums, clearing, restricted fine EMFs, averaging;
//     fc/flux_correction_fc.cpp:51-1548, receive order of ReceiveFluxCorrection :1610-1700),
//   * the boundary exchange with prolongation (cc/bvals_cc.cpp:287-520,
//     fc/bvals_fc.cpp:344-975, bvals_refine.cpp:98-673).
// Neighbour lists follow BoundaryBase::SearchAndSetNeighbors (bvals_base.cpp:298-700) and
// CountFineEdges (bvals_fc.cpp:1129-1197).  Each program is a list of box operations in
// the order the reference performs the corresponding loops, separated by barriers where
// a later box may touch what an earlier one wrote; the buffer ranges of the reference's
// pack and unpack routines become the source and destination boxes of one operation (the
// packed buffer is never materialised: the receiving MeshBlock reads the sender's array).
// Host only; every box is checked against the array it addresses before any launch.
//========================================================================================
#ifndef ATHENA_AMD_SMR_PLAN_HPP_
#define ATHENA_AMD_SMR_PLAN_HPP_

#include <algorithm>
#include <array>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace smrplan {

enum Op : int {
  BARRIER = 0,
  // boundary program (registers u, b of the updated state; w receives prolongated
  // primitives; cu/cb the coarse buffers)
  CC_COPY = 1,          // u(dst box)  <- u of block src at s0 + (dst - d0)
  CC_RESTRICT = 2,      // u(dst box)  <- restriction of src's u, coarse index s0 + (dst - d0)
  CC_TO_COARSE = 3,     // cu(dst box) <- u of block src at s0 + (dst - d0)
  CC_OWN_RESTRICT = 4,  // cu(dst box) <- restriction of own u at coarse index dst
  FC_COPY = 5, FC_RESTRICT = 6, FC_TO_COARSE = 7, FC_OWN_RESTRICT = 8,   // comp = face dir
  C2P_COARSE = 9,       // coarse box: bcc from cb, ConservedToPrimitive cu -> cw
  PROLONG_CC = 10,      // coarse box: cw -> w (8 fine cells each)
  PROLONG_FC_SHARED = 11,   // coarse face box of comp: cb -> b
  PROLONG_FC_INTERNAL = 12, // coarse cell box
  BCC_P2C = 13,         // fine box: bcc from b, PrimitiveToConserved w -> u
  // hydro flux correction
  FLUX_RESTRICT = 14,   // flux[comp](dst box, 5 vars) <- area average of src's fine fluxes
  // EMF correction (reads a snapshot of every block's EMFs, writes its own)
  EMF_ADD_SAME = 15,    // e[comp](dst) += snapshot e[comp] of src at s0 + (dst - d0)
  EMF_CLEAR = 16,       // e[comp](dst) = 0
  EMF_ADD_FINE = 17,    // e[comp](dst) += length average of src's two fine edges at
                        //                 s0 + 2 (dst - d0)
  EMF_SCALE = 18        // e[comp](dst) *= f
};

// One box operation.  Boxes are in (i, j, k) order; a source index is s0 + mul*(dst - d0).
struct Step {
  int op, comp, src;
  int d0[3], n[3];
  int s0[3];
  int mul;
  int blk;       // the MeshBlock whose registers the step writes
  double f;
};

enum { FACE = 1, EDGE = 2, CORNER = 3 };

struct Nbr {
  int gid, level;
  int ox1, ox2, ox3;
  int type;
  int fi1, fi2;
  int fid, eid;
};

struct BlockInfo {
  int level;
  std::int64_t lx[3];
  std::vector<Nbr> nb;
  int nblevel[3][3][3];
  bool edge_flag[12];
  int nedge_fine[12];
};

struct Shape {
  int nx[3], ng, cng;
  int is[3], ie[3], cis[3], cie[3];
  int nc[3], ncc[3];
};

inline Shape make_shape(const int *nx, int ng) {
  Shape S;
  S.ng = ng;
  S.cng = (ng + 1) / 2 + 1;   // cnghost (meshblock.cpp:92)
  for (int d = 0; d < 3; ++d) {
    S.nx[d] = nx[d];
    S.is[d] = ng; S.ie[d] = ng + nx[d] - 1;
    S.cis[d] = ng; S.cie[d] = ng + nx[d] / 2 - 1;
    S.nc[d] = nx[d] + 2 * ng;
    S.ncc[d] = nx[d] / 2 + 2 * ng;
  }
  return S;
}

struct Box {
  int lo[3], hi[3];   // inclusive, (i, j, k)
  int size(int d) const { return hi[d] - lo[d] + 1; }
  bool empty() const { return size(0) <= 0 || size(1) <= 0 || size(2) <= 0; }
};

inline Box box(int si, int ei, int sj, int ej, int sk, int ek) {
  Box b;
  b.lo[0] = si; b.hi[0] = ei; b.lo[1] = sj; b.hi[1] = ej; b.lo[2] = sk; b.hi[2] = ek;
  return b;
}

//----------------------------------------------------------------------------------------
// Neighbour search over the MeshBlock list (periodic mesh): FindNeighbor returns the leaf
// at the same level, the coarser leaf covering the location, or the refined node there.
class Tree {
 public:
  Tree(const int *nrb, int root_level, const std::vector<BlockInfo> &b) : root_(root_level) {
    for (int d = 0; d < 3; ++d) nrb_[d] = nrb[d];
    for (std::size_t g = 0; g < b.size(); ++g) {
      auto key = std::array<std::int64_t, 4>{b[g].level, b[g].lx[0], b[g].lx[1], b[g].lx[2]};
      if (!leaf_.emplace(key, static_cast<int>(g)).second)
        throw std::runtime_error("a MeshBlock location is listed twice");
    }
  }
  int leaf(int level, std::int64_t a, std::int64_t b, std::int64_t c) const {
    auto it = leaf_.find({level, a, b, c});
    return (it == leaf_.end()) ? -1 : it->second;
  }
  struct Found { bool refined; int gid, level; std::int64_t nl[3]; };
  Found find(const BlockInfo &s, int o1, int o2, int o3) const {
    const int o[3] = {o1, o2, o3};
    Found f{false, -1, s.level, {0, 0, 0}};
    for (int d = 0; d < 3; ++d) {
      const std::int64_t nrbl = static_cast<std::int64_t>(nrb_[d]) << (s.level - root_);
      std::int64_t l = s.lx[d] + o[d];
      if (l < 0) l += nrbl;
      if (l >= nrbl) l -= nrbl;
      f.nl[d] = l;
    }
    int g = leaf(s.level, f.nl[0], f.nl[1], f.nl[2]);
    if (g >= 0) { f.gid = g; return f; }
    if (s.level > root_) {
      g = leaf(s.level - 1, f.nl[0] >> 1, f.nl[1] >> 1, f.nl[2] >> 1);
      if (g >= 0) { f.gid = g; f.level = s.level - 1; return f; }
    }
    f.refined = true;
    f.level = s.level + 1;
    return f;
  }
  int child(const Found &f, int c1, int c2, int c3) const {
    const int g = leaf(f.level, 2 * f.nl[0] + c1, 2 * f.nl[1] + c2, 2 * f.nl[2] + c3);
    if (g < 0)
      throw std::runtime_error("MeshBlock list is not a tree with adjacent levels differing "
                               "by at most one");
    return g;
  }
 private:
  int nrb_[3];
  int root_;
  std::map<std::array<std::int64_t, 4>, int> leaf_;
};

inline void add_nbr(BlockInfo &s, int gid, int level, int o1, int o2, int o3, int type, int fi1,
                    int fi2) {
  Nbr n{gid, level, o1, o2, o3, type, fi1, fi2, -1, -1};
  if (type == FACE) n.fid = (o1 != 0) ? (o1 < 0 ? 0 : 1) : ((o2 != 0) ? (o2 < 0 ? 2 : 3)
                                                                     : (o3 < 0 ? 4 : 5));
  if (type == EDGE) {   // NeighborBlock::SetNeighbor (bvals_base.cpp:62-66)
    if (o3 == 0) n.eid = ((o1 + 1) >> 1) | ((o2 + 1) & 2);
    else if (o2 == 0) n.eid = 4 + (((o1 + 1) >> 1) | ((o3 + 1) & 2));
    else n.eid = 8 + (((o2 + 1) >> 1) | ((o3 + 1) & 2));
  }
  s.nb.push_back(n);
}

// SearchAndSetNeighbors + CountFineEdges for every block.
inline void search_neighbors(const int *nrb, int root_level, std::vector<BlockInfo> &bl) {
  const Tree T(nrb, root_level, bl);
  const int ed[3][2] = {{0, 1}, {0, 2}, {1, 2}};   // (x1,x2), (x1,x3), (x2,x3) edges
  for (BlockInfo &s : bl) {
    s.nb.clear();
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) s.nblevel[k][j][i] = -1;
    s.nblevel[1][1][1] = s.level;
    const int my[3] = {(s.lx[0] & 1) ? 1 : -1, (s.lx[1] & 1) ? 1 : -1, (s.lx[2] & 1) ? 1 : -1};
    for (int d = 0; d < 3; ++d)
      for (int n = -1; n <= 1; n += 2) {
        int o[3] = {0, 0, 0};
        o[d] = n;
        const Tree::Found f = T.find(s, o[0], o[1], o[2]);
        s.nblevel[o[2] + 1][o[1] + 1][o[0] + 1] = f.level;
        if (f.refined) {
          const int fface = 1 - (n + 1) / 2;
          for (int f2 = 0; f2 < 2; ++f2)
            for (int f1 = 0; f1 < 2; ++f1) {
              int c[3];
              if (d == 0) { c[0] = fface; c[1] = f1; c[2] = f2; }
              else if (d == 1) { c[0] = f1; c[1] = fface; c[2] = f2; }
              else { c[0] = f1; c[1] = f2; c[2] = fface; }
              add_nbr(s, T.child(f, c[0], c[1], c[2]), f.level, o[0], o[1], o[2], FACE, f1, f2);
            }
        } else {
          add_nbr(s, f.gid, f.level, o[0], o[1], o[2], FACE, 0, 0);
        }
      }
    for (int e = 0; e < 3; ++e) {
      const int da = ed[e][0], db = ed[e][1], dc = 3 - da - db;
      for (int mm = -1; mm <= 1; mm += 2)
        for (int n = -1; n <= 1; n += 2) {
          int o[3] = {0, 0, 0};
          o[da] = n;
          o[db] = mm;
          const Tree::Found f = T.find(s, o[0], o[1], o[2]);
          s.nblevel[o[2] + 1][o[1] + 1][o[0] + 1] = f.level;
          if (f.refined) {
            for (int f1 = 0; f1 < 2; ++f1) {
              int c[3];
              c[da] = 1 - (n + 1) / 2;
              c[db] = 1 - (mm + 1) / 2;
              c[dc] = f1;
              add_nbr(s, T.child(f, c[0], c[1], c[2]), f.level, o[0], o[1], o[2], EDGE, f1, 0);
            }
          } else if (f.level >= s.level || (my[da] == n && my[db] == mm)) {
            add_nbr(s, f.gid, f.level, o[0], o[1], o[2], EDGE, 0, 0);
          }
        }
    }
    for (int l = -1; l <= 1; l += 2)
      for (int mm = -1; mm <= 1; mm += 2)
        for (int n = -1; n <= 1; n += 2) {
          const Tree::Found f = T.find(s, n, mm, l);
          const int gid = f.refined ? T.child(f, 1 - (n + 1) / 2, 1 - (mm + 1) / 2, 1 - (l + 1) / 2)
                                    : f.gid;
          s.nblevel[l + 1][mm + 1][n + 1] = f.level;
          if (f.level >= s.level || (my[0] == n && my[1] == mm && my[2] == l))
            add_nbr(s, gid, f.level, n, mm, l, CORNER, 0, 0);
        }
    int eid = 0;
    for (int e = 0; e < 3; ++e) {
      const int da = ed[e][0], db = ed[e][1];
      for (int ob = -1; ob <= 1; ob += 2)
        for (int oa = -1; oa <= 1; oa += 2) {
          int nf = 0, fl = s.level;
          for (int qb = std::max(ob - 1, -1); qb <= std::min(ob + 1, 1); ++qb)
            for (int qa = std::max(oa - 1, -1); qa <= std::min(oa + 1, 1); ++qa) {
              int o[3] = {0, 0, 0};
              o[da] = qa;
              o[db] = qb;
              const int lv = s.nblevel[o[2] + 1][o[1] + 1][o[0] + 1];
              if (lv > fl) { fl++; nf = 0; }
              if (lv == fl) nf++;
            }
          s.edge_flag[eid] = (fl == s.level);
          s.nedge_fine[eid++] = nf;
        }
    }
    for (const Nbr &q : s.nb)
      if (std::abs(q.level - s.level) > 1)
        throw std::runtime_error("adjacent MeshBlocks differ by more than one level");
  }
}

// The entry of block `of` pointing back at block `to` across offset -o.
inline const Nbr &back(const std::vector<BlockInfo> &bl, int of, int to, int o1, int o2, int o3) {
  for (const Nbr &q : bl[of].nb)
    if (q.gid == to && q.ox1 == -o1 && q.ox2 == -o2 && q.ox3 == -o3) return q;
  throw std::runtime_error("neighbour lists are not symmetric");
}

//----------------------------------------------------------------------------------------
// Buffer ranges (the reference's, in each block's own index space).
inline int fi_x2(const Nbr &nb) { return (nb.ox1 != 0) ? nb.fi1 : nb.fi2; }
inline int fi_x3(const Nbr &nb) { return (nb.ox1 != 0 && nb.ox2 != 0) ? nb.fi1 : nb.fi2; }

// cell-centred (bvals_cc.cpp:287-540)
inline Box cc_same_send(const Shape &S, const Nbr &nb) {
  const int o[3] = {nb.ox1, nb.ox2, nb.ox3};
  Box b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = (o[d] > 0) ? S.ie[d] - S.ng + 1 : S.is[d];
    b.hi[d] = (o[d] < 0) ? S.is[d] + S.ng - 1 : S.ie[d];
  }
  return b;
}
inline Box cc_coarser_send(const Shape &S, const Nbr &nb) {
  const int o[3] = {nb.ox1, nb.ox2, nb.ox3}, cn = S.ng - 1;
  Box b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = (o[d] > 0) ? S.cie[d] - cn : S.cis[d];
    b.hi[d] = (o[d] < 0) ? S.cis[d] + cn : S.cie[d];
  }
  return b;
}
inline Box cc_finer_send(const Shape &S, const Nbr &nb) {
  const int o[3] = {nb.ox1, nb.ox2, nb.ox3}, cn = S.cng - 1;
  const int fi[3] = {nb.fi1, fi_x2(nb), fi_x3(nb)};
  Box b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = (o[d] > 0) ? S.ie[d] - cn : S.is[d];
    b.hi[d] = (o[d] < 0) ? S.is[d] + cn : S.ie[d];
    if (o[d] == 0) {
      if (fi[d] == 1) b.lo[d] += S.nx[d] / 2 - S.cng;
      else b.hi[d] -= S.nx[d] / 2 - S.cng;
    }
  }
  return b;
}
inline Box cc_same_recv(const Shape &S, const Nbr &nb) {
  const int o[3] = {nb.ox1, nb.ox2, nb.ox3};
  Box b;
  for (int d = 0; d < 3; ++d) {
    if (o[d] == 0) { b.lo[d] = S.is[d]; b.hi[d] = S.ie[d]; }
    else if (o[d] > 0) { b.lo[d] = S.ie[d] + 1; b.hi[d] = S.ie[d] + S.ng; }
    else { b.lo[d] = S.is[d] - S.ng; b.hi[d] = S.is[d] - 1; }
  }
  return b;
}
inline Box cc_coarser_recv(const Shape &S, const BlockInfo &s, const Nbr &nb) {
  const int o[3] = {nb.ox1, nb.ox2, nb.ox3}, cng = S.cng;
  Box b;
  for (int d = 0; d < 3; ++d) {
    if (o[d] == 0) {
      b.lo[d] = S.cis[d]; b.hi[d] = S.cie[d];
      if ((s.lx[d] & 1) == 0) b.hi[d] += cng; else b.lo[d] -= cng;
    } else if (o[d] > 0) { b.lo[d] = S.cie[d] + 1; b.hi[d] = S.cie[d] + cng; }
    else { b.lo[d] = S.cis[d] - cng; b.hi[d] = S.cis[d] - 1; }
  }
  return b;
}
inline Box cc_finer_recv(const Shape &S, const Nbr &nb) {
  const int o[3] = {nb.ox1, nb.ox2, nb.ox3};
  const int fi[3] = {nb.fi1, fi_x2(nb), fi_x3(nb)};
  Box b;
  for (int d = 0; d < 3; ++d) {
    if (o[d] == 0) {
      b.lo[d] = S.is[d]; b.hi[d] = S.ie[d];
      if (fi[d] == 1) b.lo[d] += S.nx[d] / 2; else b.hi[d] -= S.nx[d] / 2;
    } else if (o[d] > 0) { b.lo[d] = S.ie[d] + 1; b.hi[d] = S.ie[d] + S.ng; }
    else { b.lo[d] = S.is[d] - S.ng; b.hi[d] = S.is[d] - 1; }
  }
  return b;
}

// face-centred (bvals_fc.cpp:344-975), multilevel.  Component c: its own direction has
// one more face; `lo_face`/`hi_face` pick the face-direction variants of each routine.
// Same level, send (:344-399)
inline void fc_same_send(const Shape &S, const Nbr &nb, Box r[3]) {
  const int o[3] = {nb.ox1, nb.ox2, nb.ox3};
  const bool nf = nb.type != FACE;
  for (int c = 0; c < 3; ++c) {
    for (int d = 0; d < 3; ++d) {
      if (d == c) {
        if (o[d] == 0) { r[c].lo[d] = S.is[d]; r[c].hi[d] = S.ie[d] + 1; }
        else if (o[d] > 0) { r[c].lo[d] = S.ie[d] - S.ng + 1; r[c].hi[d] = S.ie[d]; }
        else { r[c].lo[d] = S.is[d] + 1; r[c].hi[d] = S.is[d] + S.ng; }
        if (nf) { if (o[d] > 0) r[c].hi[d]++; else if (o[d] < 0) r[c].lo[d]--; }
      } else {
        if (o[d] == 0) { r[c].lo[d] = S.is[d]; r[c].hi[d] = S.ie[d]; }
        else if (o[d] > 0) { r[c].lo[d] = S.ie[d] - S.ng + 1; r[c].hi[d] = S.ie[d]; }
        else { r[c].lo[d] = S.is[d]; r[c].hi[d] = S.is[d] + S.ng - 1; }
      }
    }
  }
}
// To coarser, send (:403-472), coarse indices, cng = NGHOST
inline void fc_coarser_send(const Shape &S, const Nbr &nb, Box r[3]) {
  const int o[3] = {nb.ox1, nb.ox2, nb.ox3}, cg = S.ng;
  const bool nf = nb.type != FACE;
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) {
      if (d == c) {
        if (o[d] == 0) { r[c].lo[d] = S.cis[d]; r[c].hi[d] = S.cie[d] + 1; }
        else if (o[d] > 0) { r[c].lo[d] = S.cie[d] - cg + 1; r[c].hi[d] = S.cie[d]; }
        else { r[c].lo[d] = S.cis[d] + 1; r[c].hi[d] = S.cis[d] + cg; }
        if (nf) { if (o[d] > 0) r[c].hi[d]++; else if (o[d] < 0) r[c].lo[d]--; }
      } else {
        if (o[d] == 0) { r[c].lo[d] = S.cis[d]; r[c].hi[d] = S.cie[d]; }
        else if (o[d] > 0) { r[c].lo[d] = S.cie[d] - cg + 1; r[c].hi[d] = S.cie[d]; }
        else { r[c].lo[d] = S.cis[d]; r[c].hi[d] = S.cis[d] + cg - 1; }
      }
    }
}
// To finer, send (:476-575)
inline void fc_finer_send(const Shape &S, const Nbr &nb, Box r[3]) {
  const int o[3] = {nb.ox1, nb.ox2, nb.ox3}, cn = S.cng - 1, cg = S.cng;
  const int fi[3] = {nb.fi1, fi_x2(nb), fi_x3(nb)};
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) {
      const int h = S.nx[d] / 2 - cg;
      if (d == c) {
        if (o[d] == 0) {
          r[c].lo[d] = S.is[d]; r[c].hi[d] = S.ie[d] + 1;
          if (fi[d] == 1) r[c].lo[d] += h; else r[c].hi[d] -= h;
        } else if (o[d] > 0) { r[c].lo[d] = S.ie[d] + 1 - cg; r[c].hi[d] = S.ie[d] + 1; }
        else { r[c].lo[d] = S.is[d]; r[c].hi[d] = S.is[d] + cg; }
      } else {
        if (o[d] == 0) {
          r[c].lo[d] = S.is[d]; r[c].hi[d] = S.ie[d];
          if (fi[d] == 1) r[c].lo[d] += h; else r[c].hi[d] -= h;
        } else if (o[d] > 0) { r[c].lo[d] = S.ie[d] - cn; r[c].hi[d] = S.ie[d]; }
        else { r[c].lo[d] = S.is[d]; r[c].hi[d] = S.is[d] + cn; }
      }
    }
}
// Same level, receive (:583-686), multilevel
inline void fc_same_recv(const Shape &S, const Nbr &nb, Box r[3]) {
  const int o[3] = {nb.ox1, nb.ox2, nb.ox3};
  const bool nf = nb.type != FACE;
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) {
      if (d == c) {
        if (o[d] == 0) { r[c].lo[d] = S.is[d]; r[c].hi[d] = S.ie[d] + 1; }
        else if (o[d] > 0) { r[c].lo[d] = S.ie[d] + 2; r[c].hi[d] = S.ie[d] + S.ng + 1; }
        else { r[c].lo[d] = S.is[d] - S.ng; r[c].hi[d] = S.is[d] - 1; }
        if (nf) { if (o[d] > 0) r[c].lo[d]--; else if (o[d] < 0) r[c].hi[d]++; }
      } else {
        if (o[d] == 0) { r[c].lo[d] = S.is[d]; r[c].hi[d] = S.ie[d]; }
        else if (o[d] > 0) { r[c].lo[d] = S.ie[d] + 1; r[c].hi[d] = S.ie[d] + S.ng; }
        else { r[c].lo[d] = S.is[d] - S.ng; r[c].hi[d] = S.is[d] - 1; }
      }
    }
}
// From coarser, receive into the coarse buffer (:690-807)
inline void fc_coarser_recv(const Shape &S, const BlockInfo &s, const Nbr &nb, Box r[3]) {
  const int o[3] = {nb.ox1, nb.ox2, nb.ox3}, cg = S.cng;
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) {
      const bool even = (s.lx[d] & 1) == 0;
      if (d == c) {
        if (o[d] == 0) {
          r[c].lo[d] = S.cis[d]; r[c].hi[d] = S.cie[d] + 1;
          if (even) r[c].hi[d] += cg; else r[c].lo[d] -= cg;
        } else if (o[d] > 0) { r[c].lo[d] = S.cie[d] + 1; r[c].hi[d] = S.cie[d] + 1 + cg; }
        else { r[c].lo[d] = S.cis[d] - cg; r[c].hi[d] = S.cis[d]; }
      } else {
        if (o[d] == 0) {
          r[c].lo[d] = S.cis[d]; r[c].hi[d] = S.cie[d];
          if (even) r[c].hi[d] += cg; else r[c].lo[d] -= cg;
        } else if (o[d] > 0) { r[c].lo[d] = S.cie[d] + 1; r[c].hi[d] = S.cie[d] + cg; }
        else { r[c].lo[d] = S.cis[d] - cg; r[c].hi[d] = S.cis[d] - 1; }
      }
    }
}
// From finer, receive (:813-975)
inline void fc_finer_recv(const Shape &S, const Nbr &nb, Box r[3]) {
  const int o[3] = {nb.ox1, nb.ox2, nb.ox3};
  const int fi[3] = {nb.fi1, fi_x2(nb), fi_x3(nb)};
  const bool nf = nb.type != FACE;
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) {
      const int h = S.nx[d] / 2;
      if (d == c) {
        if (o[d] == 0) {
          r[c].lo[d] = S.is[d]; r[c].hi[d] = S.ie[d] + 1;
          if (fi[d] == 1) r[c].lo[d] += h; else r[c].hi[d] -= h;
        } else if (o[d] > 0) { r[c].lo[d] = S.ie[d] + 2; r[c].hi[d] = S.ie[d] + S.ng + 1; }
        else { r[c].lo[d] = S.is[d] - S.ng; r[c].hi[d] = S.is[d] - 1; }
        if (nf) { if (o[d] > 0) r[c].lo[d]--; else if (o[d] < 0) r[c].hi[d]++; }
      } else {
        if (o[d] == 0) {
          r[c].lo[d] = S.is[d]; r[c].hi[d] = S.ie[d];
          if (fi[d] == 1) r[c].lo[d] += h; else r[c].hi[d] -= h;
        } else if (o[d] > 0) { r[c].lo[d] = S.ie[d] + 1; r[c].hi[d] = S.ie[d] + S.ng; }
        else { r[c].lo[d] = S.is[d] - S.ng; r[c].hi[d] = S.is[d] - 1; }
      }
    }
}

//----------------------------------------------------------------------------------------
// Program assembly.  A block's steps run in order on one work-group; a barrier goes in
// only where a step touches what an earlier step since the last barrier wrote, or writes
// what one read (the own-block arrays below; other blocks' arrays are only read, and only
// their active data, which no step of the launch writes).
enum ArrId { A_U, A_B0, A_CU = 4, A_CW, A_CBCC, A_CB0, A_W = 10, A_BCC, A_F0, A_E0 = 15 };

struct Acc {
  int arr;
  Box box;
  bool write;
  int op;        // the step's operation: restrictions of the same array write identical
                 // values wherever their boxes meet, so they do not conflict
};

struct Program {
  std::vector<Step> steps;
  std::vector<int> offs;     // steps of block g: [offs[g], offs[g+1])
  Shape S{};
  std::vector<Acc> pending;  // own-array accesses since the last barrier
  void begin_block() { pending.clear(); }
};

inline void check_box(const Box &b, const int *ext, const char *what) {
  for (int d = 0; d < 3; ++d)
    if (b.lo[d] < 0 || b.hi[d] >= ext[d])
      throw std::runtime_error(std::string("SMR plan: box outside the array (") + what + ")");
}

inline Box grow(Box b, int lo, int hi) {
  for (int d = 0; d < 3; ++d) { b.lo[d] += lo; b.hi[d] += hi; }
  return b;
}
// the fine cells under a box of coarse cells; face = d: the fine faces of component d
// under a box of coarse faces of that component (-1: cells)
inline Box fine_box(const Shape &S, const Box &c, int face = -1) {
  Box f;
  for (int d = 0; d < 3; ++d) {
    f.lo[d] = (c.lo[d] - S.cis[d]) * 2 + S.is[d];
    f.hi[d] = (c.hi[d] - S.cis[d]) * 2 + S.is[d] + (d == face ? 0 : 1);
  }
  return f;
}
inline bool overlap(const Box &a, const Box &b) {
  for (int d = 0; d < 3; ++d)
    if (a.hi[d] < b.lo[d] || b.hi[d] < a.lo[d]) return false;
  return true;
}

// The own-block accesses of one step (boxes that cover what it reads and writes).
inline void own_accesses(const Shape &S, const Step &st, const Box &dst, std::vector<Acc> &a) {
  const int c = st.comp, op = st.op;
  a.clear();
  auto add = [&](int arr, const Box &b, bool w) { a.push_back({arr, b, w, op}); };
  switch (op) {
    case CC_COPY: case CC_RESTRICT: add(A_U, dst, true); break;
    case CC_TO_COARSE: add(A_CU, dst, true); break;
    case CC_OWN_RESTRICT:
      add(A_CU, dst, true);
      add(A_U, fine_box(S, dst), false);
      break;
    case FC_COPY: case FC_RESTRICT: add(A_B0 + c, dst, true); break;
    case FC_TO_COARSE: add(A_CB0 + c, dst, true); break;
    case FC_OWN_RESTRICT:
      add(A_CB0 + c, dst, true);
      add(A_B0 + c, fine_box(S, dst, c), false);
      break;
    case C2P_COARSE:
      add(A_CU, dst, true);
      add(A_CW, dst, true);
      add(A_CBCC, dst, true);
      for (int q = 0; q < 3; ++q) {
        Box f = dst;
        f.hi[q] += 1;
        add(A_CB0 + q, f, false);
      }
      break;
    case PROLONG_CC:
      add(A_CW, grow(dst, -1, 1), false);
      add(A_W, fine_box(S, dst), true);
      break;
    case PROLONG_FC_SHARED: {
      Box cr = grow(dst, -1, 1);
      cr.lo[c] += 1;
      cr.hi[c] -= 1;
      add(A_CB0 + c, cr, false);
      add(A_B0 + c, fine_box(S, dst, c), true);
      break;
    }
    case PROLONG_FC_INTERNAL:
      for (int q = 0; q < 3; ++q) {
        Box f = fine_box(S, dst);
        f.hi[q] += 1;   // faces of the coarse cells, both ends
        add(A_B0 + q, f, true);
      }
      break;
    case BCC_P2C:
      add(A_W, dst, false);
      for (int q = 0; q < 3; ++q) {
        Box f = dst;
        f.hi[q] += 1;
        add(A_B0 + q, f, false);
      }
      add(A_U, dst, true);
      add(A_BCC, dst, true);
      break;
    case FLUX_RESTRICT: add(A_F0 + c, dst, true); break;
    default: add(A_E0 + c, dst, true); break;   // EMF read-modify-writes
  }
}

inline bool conflicts(const Acc &x, const Acc &y) {
  if (x.arr != y.arr || !(x.write || y.write) || !overlap(x.box, y.box)) return false;
  const bool same_restrict = x.write && y.write && x.op == y.op &&
                             (x.op == CC_OWN_RESTRICT || x.op == FC_OWN_RESTRICT);
  return !same_restrict;
}

inline void emit(Program &p, int op, int comp, int src, const Box &dst, const int *s0, int mul,
                 double f = 0.0) {
  if (dst.empty()) return;
  Step s{};
  s.op = op; s.comp = comp; s.src = src; s.mul = mul; s.f = f;
  for (int d = 0; d < 3; ++d) {
    s.d0[d] = dst.lo[d];
    s.n[d] = dst.size(d);
    s.s0[d] = s0 ? s0[d] : dst.lo[d];
  }
  std::vector<Acc> acc;
  own_accesses(p.S, s, dst, acc);
  bool conflict = false;
  for (const Acc &x : acc)
    for (const Acc &y : p.pending)
      if (conflicts(x, y)) conflict = true;
  if (conflict) {
    Step b{};
    b.op = BARRIER;
    p.steps.push_back(b);
    p.pending.clear();
  }
  p.steps.push_back(s);
  p.pending.insert(p.pending.end(), acc.begin(), acc.end());
}

// Array extents (i, j, k) of the registers a program addresses.
struct Ext {
  int cell[3], face[3][3], ccell[3], cface[3][3], edge[3][3];
};
inline Ext extents(const Shape &S) {
  Ext e;
  for (int d = 0; d < 3; ++d) { e.cell[d] = S.nc[d]; e.ccell[d] = S.ncc[d]; }
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) {
      e.face[c][d] = S.nc[d] + (d == c);
      e.cface[c][d] = S.ncc[d] + (d == c);
      e.edge[c][d] = S.nc[d] + (d != c);
    }
  return e;
}

// Source box of a copy or restriction, checked like the destination.
inline void check_src(const Box &dst, const int *s0, int mul, const int *ext, int reach,
                      const char *what) {
  Box s;
  for (int d = 0; d < 3; ++d) {
    s.lo[d] = s0[d];
    s.hi[d] = s0[d] + mul * (dst.size(d) - 1) + reach;
  }
  check_box(s, ext, what);
}

// fine index of a coarse index
inline int fine_of(const Shape &S, int d, int c) { return (c - S.cis[d]) * 2 + S.is[d]; }

// The boundary program of block g: sends' own coarse-buffer restrictions, receives in
// neighbour order, then ProlongateBoundaries for each coarser neighbour.
inline void boundary_program(const Shape &S, const std::vector<BlockInfo> &bl, int g,
                             Program &p) {
  p.S = S;
  p.begin_block();
  const Ext E = extents(S);
  const BlockInfo &s = bl[g];
  // LoadBoundaryBufferToCoarser restricts into the sender's own coarse buffer
  for (const Nbr &nb : s.nb) {
    if (nb.level >= s.level) continue;
    const Box rc = cc_coarser_send(S, nb);
    check_box(rc, E.ccell, "own cell restriction");
    emit(p, CC_OWN_RESTRICT, 0, -1, rc, nullptr, 1);
    Box rf[3];
    fc_coarser_send(S, nb, rf);
    for (int c = 0; c < 3; ++c) {
      check_box(rf[c], E.cface[c], "own face restriction");
      emit(p, FC_OWN_RESTRICT, c, -1, rf[c], nullptr, 1);
    }
  }
  for (const Nbr &nb : s.nb) {
    const Nbr &bk = back(bl, nb.gid, g, nb.ox1, nb.ox2, nb.ox3);
    Box rf[3], sf[3];
    if (nb.level == s.level) {
      const Box d = cc_same_recv(S, nb), sb = cc_same_send(S, bk);
      check_box(d, E.cell, "cell copy");
      check_src(d, sb.lo, 1, E.cell, 0, "cell copy source");
      emit(p, CC_COPY, 0, nb.gid, d, sb.lo, 1);
      fc_same_recv(S, nb, rf);
      fc_same_send(S, bk, sf);
      for (int c = 0; c < 3; ++c) {
        check_box(rf[c], E.face[c], "face copy");
        check_src(rf[c], sf[c].lo, 1, E.face[c], 0, "face copy source");
        emit(p, FC_COPY, c, nb.gid, rf[c], sf[c].lo, 1);
      }
    } else if (nb.level > s.level) {
      const Box d = cc_finer_recv(S, nb), sb = cc_coarser_send(S, bk);
      check_box(d, E.cell, "cell restriction");
      int fs[3];
      for (int q = 0; q < 3; ++q) fs[q] = fine_of(S, q, sb.lo[q]);
      check_src(d, fs, 2, E.cell, 1, "cell restriction source");
      emit(p, CC_RESTRICT, 0, nb.gid, d, sb.lo, 1);
      fc_finer_recv(S, nb, rf);
      fc_coarser_send(S, bk, sf);
      for (int c = 0; c < 3; ++c) {
        check_box(rf[c], E.face[c], "face restriction");
        for (int q = 0; q < 3; ++q) fs[q] = fine_of(S, q, sf[c].lo[q]);
        int reach[3];
        for (int q = 0; q < 3; ++q) reach[q] = (q == c) ? 0 : 1;
        Box src;
        for (int q = 0; q < 3; ++q) {
          src.lo[q] = fs[q];
          src.hi[q] = fs[q] + 2 * (rf[c].size(q) - 1) + reach[q];
        }
        check_box(src, E.face[c], "face restriction source");
        emit(p, FC_RESTRICT, c, nb.gid, rf[c], sf[c].lo, 1);
      }
    } else {
      const Box d = cc_coarser_recv(S, s, nb), sb = cc_finer_send(S, bk);
      check_box(d, E.ccell, "coarse receive");
      check_src(d, sb.lo, 1, E.cell, 0, "coarse receive source");
      emit(p, CC_TO_COARSE, 0, nb.gid, d, sb.lo, 1);
      fc_coarser_recv(S, s, nb, rf);
      fc_finer_send(S, bk, sf);
      for (int c = 0; c < 3; ++c) {
        check_box(rf[c], E.cface[c], "coarse face receive");
        check_src(rf[c], sf[c].lo, 1, E.face[c], 0, "coarse face receive source");
        emit(p, FC_TO_COARSE, c, nb.gid, rf[c], sf[c].lo, 1);
      }
    }
  }
  // ProlongateBoundaries (bvals_refine.cpp:98-246, 540-673), periodic mesh.  The reference
  // runs restriction, coarse recovery and prolongation neighbour by neighbour; here each
  // stage runs for all coarser neighbours before the next.  The two orders give the same
  // fine ghosts: the same-level restrictions write the same values whenever they repeat,
  // and every coarse cell a prolongation reads has been restricted or received before
  // either order's recovery of it (they differ only in coarse-buffer cells no
  // prolongation reads).
  const int lv = s.level;
  auto NL = [&](int a, int b, int c) { return s.nblevel[a][b][c]; };
  struct Pro { Box pr, fb[3]; };
  std::vector<Pro> pros;
  for (const Nbr &nb : s.nb) {
    if (nb.level >= s.level) continue;
    const int o[3] = {nb.ox1, nb.ox2, nb.ox3};
    int ns[3], ne[3];
    for (int d = 0; d < 3; ++d) { ns[d] = std::max(o[d] - 1, -1); ne[d] = std::min(o[d] + 1, 1); }
    for (int nk = ns[2]; nk <= ne[2]; ++nk)
      for (int nj = ns[1]; nj <= ne[1]; ++nj)
        for (int ni = ns[0]; ni <= ne[0]; ++ni) {
          const int nn[3] = {ni, nj, nk};
          if (ni == 0 && nj == 0 && nk == 0) continue;
          if (NL(nk + 1, nj + 1, ni + 1) != lv) continue;
          // RestrictGhostCellsOnSameLevel (bvals_refine.cpp:255-352)
          Box r;
          for (int d = 0; d < 3; ++d) {
            if (nn[d] == 0) {
              r.lo[d] = S.cis[d]; r.hi[d] = S.cie[d];
              if (o[d] == 1) r.lo[d] = S.cie[d]; else if (o[d] == -1) r.hi[d] = S.cis[d];
            } else if (nn[d] == 1) { r.lo[d] = r.hi[d] = S.cie[d] + 1; }
            else { r.lo[d] = r.hi[d] = S.cis[d] - 1; }
          }
          check_box(r, E.ccell, "same-level restriction");
          emit(p, CC_OWN_RESTRICT, 0, -1, r, nullptr, 1);
          for (int c = 0; c < 3; ++c) {
            Box rb = r;
            int rs = r.lo[c], re = r.hi[c] + 1;
            int lo_n[3] = {nk + 1, nj + 1, ni + 1}, hi_n[3] = {nk + 1, nj + 1, ni + 1};
            // index of direction c in [k][j][i] order: c=0 -> 2, c=1 -> 1, c=2 -> 0
            lo_n[2 - c] -= 1;
            hi_n[2 - c] += 1;
            if (rs == S.cis[c] && NL(lo_n[0], lo_n[1], lo_n[2]) < lv) rs++;
            if (re == S.cie[c] + 1 && NL(hi_n[0], hi_n[1], hi_n[2]) < lv) re--;
            rb.lo[c] = rs;
            rb.hi[c] = re;
            if (!rb.empty()) check_box(rb, E.cface[c], "same-level face restriction");
            emit(p, FC_OWN_RESTRICT, c, -1, rb, nullptr, 1);
          }
        }
    const int cn = S.cng - 1;
    Pro q;
    Box &pr = q.pr;
    for (int d = 0; d < 3; ++d) {
      if (o[d] == 0) {
        pr.lo[d] = S.cis[d]; pr.hi[d] = S.cie[d];
        if ((s.lx[d] & 1) == 0) pr.hi[d] += cn; else pr.lo[d] -= cn;
      } else if (o[d] > 0) { pr.lo[d] = S.cie[d] + 1; pr.hi[d] = S.cie[d] + cn; }
      else { pr.lo[d] = S.cis[d] - cn; pr.hi[d] = S.cis[d] - 1; }
    }
    check_box(pr, E.ccell, "cell prolongation");
    for (int d = 0; d < 3; ++d)
      if (pr.lo[d] < 1 || pr.hi[d] + 1 >= S.ncc[d] || fine_of(S, d, pr.hi[d]) + 1 >= S.nc[d])
        throw std::runtime_error("SMR plan: prolongation stencil outside the arrays");
    for (int c = 0; c < 3; ++c) {
      Box fb = pr;
      int lo = pr.lo[c], hi = pr.hi[c] + 1;
      int ln[3] = {o[2] + 1, o[1] + 1, o[0] + 1}, hn[3] = {o[2] + 1, o[1] + 1, o[0] + 1};
      ln[2 - c] = o[c];          // nblevel[...][ox - 1 + 1]
      hn[2 - c] = o[c] + 2;
      if (o[c] >= 0 && NL(ln[0], ln[1], ln[2]) >= lv) lo++;
      if (o[c] <= 0 && NL(hn[0], hn[1], hn[2]) >= lv) hi--;
      fb.lo[c] = lo;
      fb.hi[c] = hi;
      if (!fb.empty()) check_box(fb, E.cface[c], "face prolongation");
      q.fb[c] = fb;
    }
    pros.push_back(q);
  }
  // ApplyPhysicalBoundariesOnCoarseLevel: every side has a neighbour on a periodic mesh,
  // so the recovery covers each prolongation box grown by one coarse cell
  for (const Pro &q : pros) {
    const Box cr = grow(q.pr, -1, 1);
    check_box(cr, E.ccell, "coarse recovery");
    emit(p, C2P_COARSE, 0, -1, cr, nullptr, 1);
  }
  for (const Pro &q : pros) {
    emit(p, PROLONG_CC, 0, -1, q.pr, nullptr, 1);
    for (int c = 0; c < 3; ++c) emit(p, PROLONG_FC_SHARED, c, -1, q.fb[c], nullptr, 1);
  }
  for (const Pro &q : pros) emit(p, PROLONG_FC_INTERNAL, 0, -1, q.pr, nullptr, 1);
  for (const Pro &q : pros) {
    Box fr;
    for (int d = 0; d < 3; ++d) {
      fr.lo[d] = fine_of(S, d, q.pr.lo[d]);
      fr.hi[d] = fine_of(S, d, q.pr.hi[d]) + 1;
    }
    check_box(fr, E.cell, "fine recovery box");
    emit(p, BCC_P2C, 0, -1, fr, nullptr, 1);
  }
}

// The hydro flux correction of block g: restricted fluxes of each finer face neighbour
// replace its own on that quarter of the face (SetFluxBoundaryFromFiner).
inline void flux_program(const Shape &S, const std::vector<BlockInfo> &bl, int g, Program &p) {
  p.S = S;
  p.begin_block();
  const Ext E = extents(S);
  const BlockInfo &s = bl[g];
  for (const Nbr &nb : s.nb) {
    if (nb.type != FACE) break;
    if (nb.level <= s.level) continue;
    const int dir = nb.fid / 2, side = nb.fid & 1;
    const int t1 = (dir == 0) ? 1 : 0, t2 = (dir == 2) ? 1 : 2;   // transverse directions
    Box d;
    d.lo[dir] = d.hi[dir] = side ? S.ie[dir] + 1 : S.is[dir];
    d.lo[t1] = S.is[t1]; d.hi[t1] = S.ie[t1];
    d.lo[t2] = S.is[t2]; d.hi[t2] = S.ie[t2];
    if (nb.fi1 == 0) d.hi[t1] -= S.nx[t1] / 2; else d.lo[t1] += S.nx[t1] / 2;
    if (nb.fi2 == 0) d.hi[t2] -= S.nx[t2] / 2; else d.lo[t2] += S.nx[t2] / 2;
    check_box(d, E.face[dir], "flux correction");
    const Nbr &bk = back(bl, nb.gid, g, nb.ox1, nb.ox2, nb.ox3);
    const int bside = bk.fid & 1;
    int s0[3];
    s0[dir] = bside ? S.ie[dir] + 1 : S.is[dir];
    s0[t1] = S.is[t1];
    s0[t2] = S.is[t2];
    check_src(d, s0, 2, E.face[dir], 1, "flux correction source");
    emit(p, FLUX_RESTRICT, dir, nb.gid, d, s0, 2);
  }
}

// EMF boxes of a face (direction dir, side) or an edge of a block, per edge component.
inline void emf_face_boxes(const Shape &S, int dir, int side, Box out[3], bool has[3]) {
  for (int c = 0; c < 3; ++c) {
    has[c] = (c != dir);
    if (!has[c]) continue;
    for (int d = 0; d < 3; ++d) {
      out[c].lo[d] = S.is[d];
      out[c].hi[d] = S.ie[d] + (d != c ? 1 : 0);
    }
    out[c].lo[dir] = out[c].hi[dir] = side ? S.ie[dir] + 1 : S.is[dir];
  }
}
inline Box emf_edge_box(const Shape &S, int eid, int &comp) {
  const int lo = eid & 1, hi = (eid & 2) ? 1 : 0;
  int da, db;
  if (eid < 4) { da = 0; db = 1; comp = 2; }
  else if (eid < 8) { da = 0; db = 2; comp = 1; }
  else { da = 1; db = 2; comp = 0; }
  Box b;
  b.lo[comp] = S.is[comp]; b.hi[comp] = S.ie[comp];
  b.lo[da] = b.hi[da] = lo ? S.ie[da] + 1 : S.is[da];
  b.lo[db] = b.hi[db] = hi ? S.ie[db] + 1 : S.is[db];
  return b;
}

// The EMF correction of block g (ReceiveFluxCorrection): same-level sums in neighbour
// order, ClearCoarseFluxBoundary, the restricted EMFs of finer neighbours, then
// AverageFluxBoundary.
inline void emf_program(const Shape &S, const std::vector<BlockInfo> &bl, int g, Program &p) {
  p.S = S;
  p.begin_block();
  const Ext E = extents(S);
  const BlockInfo &s = bl[g];
  for (const Nbr &nb : s.nb) {
    if (nb.type != FACE && nb.type != EDGE) break;
    if (nb.level != s.level) continue;
    if (nb.type == EDGE && !s.edge_flag[nb.eid]) continue;
    const Nbr &bk = back(bl, nb.gid, g, nb.ox1, nb.ox2, nb.ox3);
    if (nb.type == FACE) {
      Box mine[3], theirs[3];
      bool has[3], h2[3];
      emf_face_boxes(S, nb.fid / 2, nb.fid & 1, mine, has);
      emf_face_boxes(S, bk.fid / 2, bk.fid & 1, theirs, h2);
      for (int c = 0; c < 3; ++c)
        if (has[c]) {
          check_box(mine[c], E.edge[c], "EMF sum");
          check_src(mine[c], theirs[c].lo, 1, E.edge[c], 0, "EMF sum source");
          emit(p, EMF_ADD_SAME, c, nb.gid, mine[c], theirs[c].lo, 1);
        }
    } else {
      int c, c2;
      const Box mine = emf_edge_box(S, nb.eid, c), theirs = emf_edge_box(S, bk.eid, c2);
      check_box(mine, E.edge[c], "EMF edge sum");
      check_src(mine, theirs.lo, 1, E.edge[c], 0, "EMF edge sum source");
      emit(p, EMF_ADD_SAME, c, nb.gid, mine, theirs.lo, 1);
    }
  }
  // ClearCoarseFluxBoundary (flux_correction_fc.cpp:1214-1352)
  for (int f = 0; f < 6; ++f) {
    const int dir = f / 2, side = f & 1;
    int o[3] = {0, 0, 0};
    o[dir] = side ? 1 : -1;
    if (s.nblevel[o[2] + 1][o[1] + 1][o[0] + 1] <= s.level) continue;
    Box b[3];
    bool has[3];
    emf_face_boxes(S, dir, side, b, has);
    for (int c = 0; c < 3; ++c)
      if (has[c]) {
        // interior of the face only: the transverse direction other than c loses its ends
        const int t = 3 - dir - c;
        b[c].lo[t] += 1;
        b[c].hi[t] -= 1;
        emit(p, EMF_CLEAR, c, -1, b[c], nullptr, 1);
      }
  }
  for (int e = 0; e < 12; ++e) {
    if (s.edge_flag[e]) continue;
    int c;
    const Box b = emf_edge_box(S, e, c);
    emit(p, EMF_CLEAR, c, -1, b, nullptr, 1);
  }
  for (const Nbr &nb : s.nb) {
    if (nb.type != FACE && nb.type != EDGE) break;
    if (nb.level != s.level + 1) continue;
    const Nbr &bk = back(bl, nb.gid, g, nb.ox1, nb.ox2, nb.ox3);
    if (nb.type == FACE) {
      const int dir = nb.fid / 2;
      const int t1 = (dir == 0) ? 1 : 0, t2 = (dir == 2) ? 1 : 2;
      Box b[3];
      bool has[3];
      emf_face_boxes(S, dir, nb.fid & 1, b, has);
      for (int c = 0; c < 3; ++c)
        if (has[c]) {
          // the quarter of the face this fine block covers (fi1 along t1, fi2 along t2)
          const int tt[2] = {t1, t2}, fi[2] = {nb.fi1, nb.fi2};
          for (int q = 0; q < 2; ++q) {
            const int t = tt[q], h = S.nx[t] / 2;
            if (fi[q] == 0) b[c].hi[t] = S.is[t] + h - 1 + (t != c ? 1 : 0);
            else b[c].lo[t] = S.is[t] + h;
          }
          int s0[3];
          s0[dir] = (bk.fid & 1) ? S.ie[dir] + 1 : S.is[dir];
          s0[t1] = S.is[t1];
          s0[t2] = S.is[t2];
          check_box(b[c], E.edge[c], "EMF fine sum");
          int reach = 1;
          Box src;
          for (int q = 0; q < 3; ++q) {
            src.lo[q] = s0[q];
            src.hi[q] = s0[q] + 2 * (b[c].size(q) - 1) + ((q == c) ? reach : 0);
          }
          check_box(src, E.edge[c], "EMF fine sum source");
          emit(p, EMF_ADD_FINE, c, nb.gid, b[c], s0, 2);
        }
    } else {
      int c, c2;
      Box b = emf_edge_box(S, nb.eid, c);
      const Box theirs = emf_edge_box(S, bk.eid, c2);
      const int h = S.nx[c] / 2;
      if (nb.fi1 == 0) b.hi[c] = S.is[c] + h - 1; else b.lo[c] = S.is[c] + h;
      int s0[3] = {theirs.lo[0], theirs.lo[1], theirs.lo[2]};
      s0[c] = S.is[c];
      check_box(b, E.edge[c], "EMF fine edge");
      Box src;
      for (int q = 0; q < 3; ++q) {
        src.lo[q] = s0[q];
        src.hi[q] = s0[q] + 2 * (b.size(q) - 1) + ((q == c) ? 1 : 0);
      }
      check_box(src, E.edge[c], "EMF fine edge source");
      emit(p, EMF_ADD_FINE, c, nb.gid, b, s0, 2);
    }
  }
  // AverageFluxBoundary (flux_correction_fc.cpp:1355-1548), periodic faces
  for (int f = 0; f < 6; ++f) {
    const int dir = f / 2, side = f & 1;
    int o[3] = {0, 0, 0};
    o[dir] = side ? 1 : -1;
    const int nl = s.nblevel[o[2] + 1][o[1] + 1][o[0] + 1];
    Box b[3];
    bool has[3];
    emf_face_boxes(S, dir, side, b, has);
    for (int c = 0; c < 3; ++c) {
      if (!has[c]) continue;
      const int t = 3 - dir - c;   // the transverse direction across component c's lines
      if (nl == s.level) {
        b[c].lo[t] += 1;
        b[c].hi[t] -= 1;
        emit(p, EMF_SCALE, c, -1, b[c], nullptr, 1, 0.5);
      } else if (nl > s.level) {
        b[c].lo[t] = b[c].hi[t] = S.is[t] + S.nx[t] / 2;
        emit(p, EMF_SCALE, c, -1, b[c], nullptr, 1, 0.5);
      }
    }
  }
  for (int e = 0; e < 12; ++e) {
    if (s.nedge_fine[e] == 1) continue;
    int c;
    const Box b = emf_edge_box(S, e, c);
    emit(p, EMF_SCALE, c, -1, b, nullptr, 1, 1.0 / static_cast<double>(s.nedge_fine[e]));
  }
}

// A program of all blocks as phases: phase k holds the k-th barrier-free group of every
// block's steps, so one launch runs a phase over all blocks at once.  A work-group of
// CHUNK_THREADS threads takes a chunk of up to CHUNK elements of one step (step index,
// first element); one element per thread measured fastest (2048 per group: 18 % slower).
constexpr int CHUNK_THREADS = 256;
constexpr int CHUNK = 256;

inline int step_elements(const Step &s) {
  const bool five = s.op == CC_COPY || s.op == CC_RESTRICT || s.op == CC_TO_COARSE ||
                    s.op == CC_OWN_RESTRICT || s.op == PROLONG_CC || s.op == FLUX_RESTRICT;
  return (five ? 5 : 1) * s.n[0] * s.n[1] * s.n[2];
}

struct Phased {
  std::vector<Step> steps;
  std::vector<int> chunk_step, chunk_first;   // per chunk
  std::vector<int> phase_chunks;              // chunk offset of each phase, plus the end
};

inline Phased to_phases(const Program &p, int nblocks) {
  std::vector<std::vector<std::vector<Step>>> groups(nblocks);
  std::size_t nph = 0;
  for (int g = 0; g < nblocks; ++g) {
    groups[g].emplace_back();
    for (int q = p.offs[g]; q < p.offs[g + 1]; ++q) {
      const Step &st = p.steps[q];
      if (st.op == BARRIER) {
        if (!groups[g].back().empty()) groups[g].emplace_back();
        continue;
      }
      Step c = st;
      c.blk = g;
      groups[g].back().push_back(c);
    }
    if (groups[g].back().empty()) groups[g].pop_back();
    nph = std::max(nph, groups[g].size());
  }
  Phased out;
  out.phase_chunks.push_back(0);
  for (std::size_t k = 0; k < nph; ++k) {
    for (int g = 0; g < nblocks; ++g) {
      if (k >= groups[g].size()) continue;
      for (const Step &st : groups[g][k]) {
        const int idx = static_cast<int>(out.steps.size());
        out.steps.push_back(st);
        const int ne = step_elements(st);
        for (int first = 0; first < ne; first += CHUNK) {
          out.chunk_step.push_back(idx);
          out.chunk_first.push_back(first);
        }
      }
    }
    out.phase_chunks.push_back(static_cast<int>(out.chunk_step.size()));
  }
  return out;
}

}  // namespace smrplan

#endif  // ATHENA_AMD_SMR_PLAN_HPP_
