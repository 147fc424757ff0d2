//========================================================================================
// refine.hip -- MeshRefinement operators for static mesh refinement on MI355X (gfx950):
// restriction of cell-centred values and face fields to a MeshBlock's coarse buffer,
// prolongation from the coarse buffer (minmod-limited linear for cells and for the fine
// faces on coarse faces), the divergence-preserving prolongation of the fine faces
// inside coarse cells (Toth & Roe 2002), and the restriction of edge EMFs used by the
// EMF correction between levels.  3-D, uniform-Cartesian branches of
// mesh_refinement.cpp:116-150, 272-300, 328-372, 402-430, 464-525, 835-872 (and the
// matching Cartesian blocks of ProlongateSharedFieldX2/X3), 1223-1338, with each
// expression in the reference's association order (built with -ffp-contract=off).
//
// Geometry (the reference's, meshblock.cpp:66,93-94): a block has nx cells per
// direction and ng ghosts (nc = nx + 2 ng); its coarse buffer has nx/2 cells and ng
// ghosts (ncc = nx/2 + 2 ng); fine index = (coarse index - ng) * 2 + ng.  A batch is nb
// blocks of one shape stored back to back: cell arrays (nvar, nc3, nc2, nc1) per block,
// face component d with one more entry along d, coordinate tables per block packed as
// x1f (n1+1), x1v (n1), dx1f (n1), then x2, x3.  Cartesian metrics: volume dx1 dx2 dx3,
// face areas dx2 dx3 / dx1 dx3 / dx1 dx2, edge lengths dx2, dx3; face-centre positions
// along transverse directions are the cell centres (cartesian.cpp:105-118).
//
// One thread per (block, component, coarse position), linear and i fastest.  Every
// index a launch touches is checked on the host against the arrays it addresses.
//========================================================================================
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/athena_gpu.h"
#include "physics.hpp"
#include "refine_ops.hpp"

namespace {

constexpr int RWG = 256;

struct RShape {
  int nx1, nx2, nx3, ng;
  int nc1, nc2, nc3, ncc1, ncc2, ncc3;
  long fcl, ccl;          // coordinate table lengths (fine, coarse) per block
};

struct RRange {
  int i0, i1, j0, j1, k0, k1;
  __host__ __device__ int ni() const { return i1 - i0 + 1; }
  __host__ __device__ int nj() const { return j1 - j0 + 1; }
  __host__ __device__ int nk() const { return k1 - k0 + 1; }
};

using refops::at;
using refops::RGrid;
using refops::View;

// thread -> (block, component, k, j, i) over nb * ncomp * |range|; false past the end
__device__ __forceinline__ bool rmap(const RRange &r, int nb, int ncomp, int &b, int &n,
                                     int &k, int &j, int &i) {
  const long t = static_cast<long>(blockIdx.x) * RWG + threadIdx.x;
  const long per = static_cast<long>(r.ni()) * r.nj() * r.nk();
  if (t >= per * ncomp * nb) return false;
  const long bn = t / per;
  long q = t - bn * per;
  b = static_cast<int>(bn / ncomp);
  n = static_cast<int>(bn - static_cast<long>(b) * ncomp);
  i = r.i0 + static_cast<int>(q % r.ni());
  q /= r.ni();
  j = r.j0 + static_cast<int>(q % r.nj());
  k = r.k0 + static_cast<int>(q / r.nj());
  return true;
}

__global__ void __launch_bounds__(RWG) k_restrict_cc(RShape S, RRange r, int nb, int nvar,
                                                      const double *fcoord,
                                                      const double *fine, double *coarse) {
  int b, n, ck, cj, ci;
  if (!rmap(r, nb, nvar, b, n, ck, cj, ci)) return;
  const RGrid F(fcoord + b * S.fcl, S.nc1, S.nc2, S.nc3);
  const long fn = static_cast<long>(S.nc1) * S.nc2 * S.nc3;
  const long cn = static_cast<long>(S.ncc1) * S.ncc2 * S.ncc3;
  const View f{const_cast<double *>(fine) + (static_cast<long>(b) * nvar + n) * fn, S.nc2, S.nc1};
  const int k = (ck - S.ng) * 2 + S.ng, j = (cj - S.ng) * 2 + S.ng, i = (ci - S.ng) * 2 + S.ng;
  coarse[(static_cast<long>(b) * nvar + n) * cn + at(ck, cj, ci, S.ncc2, S.ncc1)] =
      refops::restrict_cc_at(f, F, k, j, i);
}

template <int comp>
__global__ void __launch_bounds__(RWG) k_restrict_fc(RShape S, RRange r, int nb,
                                                      const double *fcoord,
                                                      const double *fine, double *coarse) {
  int b, n, ck, cj, ci;
  if (!rmap(r, nb, 1, b, n, ck, cj, ci)) return;
  const RGrid F(fcoord + b * S.fcl, S.nc1, S.nc2, S.nc3);
  const int fn1 = S.nc1 + (comp == 0), fn2 = S.nc2 + (comp == 1), fn3 = S.nc3 + (comp == 2);
  const int cn1 = S.ncc1 + (comp == 0), cn2 = S.ncc2 + (comp == 1), cn3 = S.ncc3 + (comp == 2);
  const View f{const_cast<double *>(fine) + static_cast<long>(b) * fn1 * fn2 * fn3, fn2, fn1};
  const int k = (ck - S.ng) * 2 + S.ng, j = (cj - S.ng) * 2 + S.ng, i = (ci - S.ng) * 2 + S.ng;
  coarse[static_cast<long>(b) * cn1 * cn2 * cn3 + at(ck, cj, ci, cn2, cn1)] =
      refops::restrict_fc_at(comp, f, F, k, j, i);
}

__global__ void __launch_bounds__(RWG) k_prolongate_cc(RShape S, RRange r, int nb, int nvar,
                                                        const double *fcoord,
                                                        const double *ccoord,
                                                        const double *coarse, double *fine) {
  int b, n, k, j, i;
  if (!rmap(r, nb, nvar, b, n, k, j, i)) return;
  const RGrid F(fcoord + b * S.fcl, S.nc1, S.nc2, S.nc3);
  const RGrid C(ccoord + b * S.ccl, S.ncc1, S.ncc2, S.ncc3);
  const long fn = static_cast<long>(S.nc1) * S.nc2 * S.nc3;
  const long cn = static_cast<long>(S.ncc1) * S.ncc2 * S.ncc3;
  const View c{const_cast<double *>(coarse) + (static_cast<long>(b) * nvar + n) * cn, S.ncc2,
               S.ncc1};
  const View f{fine + (static_cast<long>(b) * nvar + n) * fn, S.nc2, S.nc1};
  refops::prolong_cc_at(c, f, F, C, S.ng, k, j, i);
}

template <int comp>
__global__ void __launch_bounds__(RWG) k_prolongate_fc_shared(RShape S, RRange r, int nb,
                                                               const double *fcoord,
                                                               const double *ccoord,
                                                               const double *coarse,
                                                               double *fine) {
  int b, n, k, j, i;
  if (!rmap(r, nb, 1, b, n, k, j, i)) return;
  const RGrid F(fcoord + b * S.fcl, S.nc1, S.nc2, S.nc3);
  const RGrid C(ccoord + b * S.ccl, S.ncc1, S.ncc2, S.ncc3);
  const int fn1 = S.nc1 + (comp == 0), fn2 = S.nc2 + (comp == 1), fn3 = S.nc3 + (comp == 2);
  const int cn1 = S.ncc1 + (comp == 0), cn2 = S.ncc2 + (comp == 1), cn3 = S.ncc3 + (comp == 2);
  const View c{const_cast<double *>(coarse) + static_cast<long>(b) * cn1 * cn2 * cn3, cn2, cn1};
  const View f{fine + static_cast<long>(b) * fn1 * fn2 * fn3, fn2, fn1};
  refops::prolong_fc_shared_at(comp, c, f, F, C, S.ng, k, j, i);
}

__global__ void __launch_bounds__(RWG) k_prolongate_fc_internal(RShape S, RRange r, int nb,
                                                                 const double *fcoord,
                                                                 double *b1, double *b2,
                                                                 double *b3) {
  int b, n, k, j, i;
  if (!rmap(r, nb, 1, b, n, k, j, i)) return;
  const RGrid F(fcoord + b * S.fcl, S.nc1, S.nc2, S.nc3);
  const long n1f = static_cast<long>(S.nc3) * S.nc2 * (S.nc1 + 1);
  const long n2f = static_cast<long>(S.nc3) * (S.nc2 + 1) * S.nc1;
  const long n3f = static_cast<long>(S.nc3 + 1) * S.nc2 * S.nc1;
  const View X1{b1 + b * n1f, S.nc2, S.nc1 + 1};
  const View X2{b2 + b * n2f, S.nc2 + 1, S.nc1};
  const View X3{b3 + b * n3f, S.nc2, S.nc1};
  refops::prolong_fc_internal_at(X1, X2, X3, F, S.ng, k, j, i);
}

// The EMF restriction of FaceCenteredBoundaryVariable::LoadFluxBoundaryBufferToCoarser
// (flux_correction_fc.cpp:315-420): a coarse edge along comp is the length-weighted
// average of its two fine edges.  Edge arrays: along x1 (n3+1, n2+1, n1), along x2
// (n3+1, n2, n1+1), along x3 (n3, n2+1, n1+1).
template <int comp>
__global__ void __launch_bounds__(RWG) k_restrict_emf(RShape S, RRange r, int nb,
                                                       const double *fcoord,
                                                       const double *fine, double *coarse) {
  int b, n, ck, cj, ci;
  if (!rmap(r, nb, 1, b, n, ck, cj, ci)) return;
  const RGrid F(fcoord + b * S.fcl, S.nc1, S.nc2, S.nc3);
  const int fn1 = S.nc1 + (comp != 0), fn2 = S.nc2 + (comp != 1), fn3 = S.nc3 + (comp != 2);
  const int cn1 = S.ncc1 + (comp != 0), cn2 = S.ncc2 + (comp != 1), cn3 = S.ncc3 + (comp != 2);
  const View f{const_cast<double *>(fine) + static_cast<long>(b) * fn1 * fn2 * fn3, fn2, fn1};
  const int k = (ck - S.ng) * 2 + S.ng, j = (cj - S.ng) * 2 + S.ng, i = (ci - S.ng) * 2 + S.ng;
  coarse[static_cast<long>(b) * cn1 * cn2 * cn3 + at(ck, cj, ci, cn2, cn1)] =
      refops::restrict_emf_at(comp, f, F, k, j, i);
}

//----------------------------------------------------------------------------------------
// Host side: shapes, range checks, launches.

int rfail(const std::string &m) { return athgpu::set_error(m); }

bool make_shape(const int *shape, RShape &S) {
  if (!shape) return false;
  S.nx1 = shape[0]; S.nx2 = shape[1]; S.nx3 = shape[2]; S.ng = shape[3];
  if (S.nx1 < 2 || S.nx2 < 2 || S.nx3 < 2 || S.ng < 1) return false;
  if (S.nx1 % 2 || S.nx2 % 2 || S.nx3 % 2) return false;
  S.nc1 = S.nx1 + 2 * S.ng; S.nc2 = S.nx2 + 2 * S.ng; S.nc3 = S.nx3 + 2 * S.ng;
  S.ncc1 = S.nx1 / 2 + 2 * S.ng; S.ncc2 = S.nx2 / 2 + 2 * S.ng; S.ncc3 = S.nx3 / 2 + 2 * S.ng;
  S.fcl = 3L * S.nc1 + 1 + 3L * S.nc2 + 1 + 3L * S.nc3 + 1;
  S.ccl = 3L * S.ncc1 + 1 + 3L * S.ncc2 + 1 + 3L * S.ncc3 + 1;
  return true;
}

// Per direction: the coarse indices [lo, hi] must lie in [clo, chi] and the fine indices
// they address, f(lo) + flo .. f(hi) + fhi with f(c) = (c - ng) * 2 + ng, in [0, fmax].
bool dir_ok(int lo, int hi, int clo, int chi, int ng, int flo, int fhi, int fmax) {
  if (lo > hi || lo < clo || hi > chi) return false;
  const int a = (lo - ng) * 2 + ng + flo, b = (hi - ng) * 2 + ng + fhi;
  return a >= 0 && b <= fmax;
}

bool range_of(const int *r, RRange &R) {
  if (!r) return false;
  R = RRange{r[0], r[1], r[2], r[3], r[4], r[5]};
  return true;
}

int launch_check(hipStream_t s) {
  (void)s;
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return rfail(std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

dim3 rgrid(const RRange &R, int nb, int ncomp) {
  const long n = static_cast<long>(R.ni()) * R.nj() * R.nk() * ncomp * nb;
  return dim3(static_cast<unsigned>((n + RWG - 1) / RWG));
}

//----------------------------------------------------------------------------------------
// MeshBlock tree of a 3-D mesh with static refinement (meshblock_tree.cpp).  Nodes live in
// one vector and refer to their children by index; a node without children is a leaf,
// that is a MeshBlock.  Logical levels count from the node that covers the whole
// power-of-two box (level 0); the root grid sits at root_level.
struct TNode {
  int level;
  long long lx[3];
  int child[8];   // index n = i + 2 j + 4 k of the child at (2 lx + (i, j, k)), -1 if none
  bool leaf() const {
    for (int n = 0; n < 8; ++n) if (child[n] >= 0) return false;
    return true;
  }
};

// About 64 bytes per node: the cap bounds the host vector near 256 MiB.
constexpr long kTreeMaxNodes = 1L << 22;
constexpr int kTreeMaxLevel = 40;

struct Tree {
  std::vector<TNode> nd;
  long long nrb[3];
  bool periodic[6];   // inner/outer x1, x2, x3
  int root_level = 0;
  bool overflow = false;

  int make(int level, long long l1, long long l2, long long l3) {
    if (static_cast<long>(nd.size()) >= kTreeMaxNodes) { overflow = true; return -1; }
    TNode t;
    t.level = level; t.lx[0] = l1; t.lx[1] = l2; t.lx[2] = l3;
    for (int n = 0; n < 8; ++n) t.child[n] = -1;
    nd.push_back(t);
    return static_cast<int>(nd.size()) - 1;
  }

  // CreateRootGrid (meshblock_tree.cpp:87-112): children exist where they overlap the
  // nrb[0] x nrb[1] x nrb[2] root blocks.
  void create_root(int id) {
    const int level = nd[id].level;
    if (level == root_level || overflow) return;
    const long long fac = 1LL << (root_level - level - 1);
    for (int n = 0; n < 8; ++n) {
      const int i = n & 1, j = (n >> 1) & 1, k = (n >> 2) & 1;
      const long long c1 = nd[id].lx[0] * 2 + i, c2 = nd[id].lx[1] * 2 + j,
                      c3 = nd[id].lx[2] * 2 + k;
      if (c3 * fac < nrb[2] && c2 * fac < nrb[1] && c1 * fac < nrb[0]) {
        const int c = make(level + 1, c1, c2, c3);
        if (c < 0) return;
        nd[id].child[n] = c;
        create_root(c);
      }
    }
  }

  // Refine (meshblock_tree.cpp:166-258): the eight children, then every neighbour of the
  // refined node at its own level is made to exist (periodic faces wrap, others end the
  // search), which keeps neighbouring leaves within one level of each other.
  void refine(int id) {
    if (!nd[id].leaf() || overflow) return;
    const int level = nd[id].level;
    const long long l0[3] = {nd[id].lx[0], nd[id].lx[1], nd[id].lx[2]};
    for (int n = 0; n < 8; ++n) {
      const int c = make(level + 1, l0[0] * 2 + (n & 1), l0[1] * 2 + ((n >> 1) & 1),
                         l0[2] * 2 + ((n >> 2) & 1));
      if (c < 0) return;
      nd[id].child[n] = c;
    }
    long long nmax[3];
    for (int q = 0; q < 3; ++q) nmax[q] = nrb[q] << (level - root_level);
    for (int oz = -1; oz <= 1; ++oz)
      for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) {
          if (ox == 0 && oy == 0 && oz == 0) continue;
          const int o[3] = {ox, oy, oz};
          long long l[3];
          bool skip = false;
          for (int q = 0; q < 3 && !skip; ++q) {
            l[q] = l0[q] + o[q];
            if (l[q] < 0) { if (periodic[2 * q]) l[q] = nmax[q] - 1; else skip = true; }
            else if (l[q] >= nmax[q]) { if (periodic[2 * q + 1]) l[q] = 0; else skip = true; }
          }
          if (!skip) add(level, l);
          if (overflow) return;
        }
  }

  // AddMeshBlock (meshblock_tree.cpp:116-135): walk from the top, refining leaves on the
  // way, until a node of the requested level exists at l.
  void add(int level, const long long *l) {
    int id = 0;
    while (nd[id].level < level && !overflow) {
      if (nd[id].leaf()) refine(id);
      if (overflow) return;
      const int sh = level - nd[id].level - 1;
      const int n = static_cast<int>(((l[0] >> sh) & 1) + 2 * ((l[1] >> sh) & 1) +
                                     4 * ((l[2] >> sh) & 1));
      id = nd[id].child[n];
      if (id < 0) { overflow = true; return; }   // outside the root grid: not reachable
    }
  }

  // GetMeshBlockList (meshblock_tree.cpp:336-356): leaves depth first, children in
  // index order, which is the Z order of the MeshBlock ids.
  void leaves(int id, std::vector<int> &out) const {
    if (nd[id].leaf()) { out.push_back(id); return; }
    for (int n = 0; n < 8; ++n) if (nd[id].child[n] >= 0) leaves(nd[id].child[n], out);
  }

  // FindNeighbor (meshblock_tree.cpp:362-468): the node at the neighbour's position and
  // level, or the coarser leaf that contains it; -1 across a non-periodic face, -3 when
  // the tree has no node there (never for a tree built here).
  int neighbor(int id, int o1, int o2, int o3) const {
    const int ll = nd[id].level;
    const int o[3] = {o1, o2, o3};
    long long l[3];
    for (int q = 0; q < 3; ++q) {
      const long long nmax = nrb[q] << (ll - root_level);
      l[q] = nd[id].lx[q] + o[q];
      if (l[q] < 0) { if (!periodic[2 * q]) return -1; l[q] = nmax - 1; }
      else if (l[q] >= nmax) { if (!periodic[2 * q + 1]) return -1; l[q] = 0; }
    }
    if (ll < 1) return 0;
    int bt = 0;
    for (int level = 0; level < ll; ++level) {
      if (nd[bt].leaf()) return (level == ll - 1) ? bt : -3;
      const int sh = ll - level - 1;
      const int n = static_cast<int>(((l[0] >> sh) & 1) + 2 * ((l[1] >> sh) & 1) +
                                     4 * ((l[2] >> sh) & 1));
      bt = nd[bt].child[n];
      if (bt < 0) return -3;
    }
    return bt;
  }
};

// Uniform mesh generator at logical face index `index` of `nrange` (mesh.hpp:457-472 with
// sym_interval, :535-538).
double tree_face_x(long long index, long long nrange, double xmin, double xmax) {
  const long long no = index - nrange / 2, noc = index - (nrange + 1) / 2;
  const double x = static_cast<double>(no + noc) / (2.0 * static_cast<double>(nrange));
  return 0.5 * (xmin + xmax) + (x * xmax - x * xmin);
}

// First v in [lo, hi) with pred(v) true, or hi; pred must be false then true along v.
template <class Pred>
long long first_true(long long lo, long long hi, Pred pred) {
  while (lo < hi) {
    const long long mid = lo + (hi - lo) / 2;
    if (pred(mid)) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// The logical index range [lo, hi) at ref_lev that a refinement region covers along one
// direction, widened to whole sibling pairs (mesh.cpp:400-447).  The reference scans the
// indices in order for the first right face beyond rmin and the first at or beyond rmax;
// face positions grow with the index, so a bisection finds the same indices in
// O(ref_lev) steps for deep levels.
void region_range(long long nrb, int ref_lev, double rmin, double rmax, double xmin,
                  double xmax, long long *lo, long long *hi) {
  const long long lxmax = nrb * (1LL << ref_lev);
  long long a = first_true(0, lxmax, [&](long long v) {
    return tree_face_x(v + 1, lxmax, xmin, xmax) > rmin;
  });
  long long b = first_true(a, lxmax, [&](long long v) {
    return tree_face_x(v + 1, lxmax, xmin, xmax) >= rmax;
  });
  if (a % 2 == 1) --a;
  if (b % 2 == 0) ++b;
  *lo = a; *hi = b;
}

}  // namespace

extern "C" {

int ath_gpu_restrict_cc(const int *shape, int nb, int nvar, const double *fcoord,
                        const double *fine, double *coarse, const int *range, void *stream) {
  RShape S;
  RRange R;
  if (!make_shape(shape, S) || !range_of(range, R) || nb < 1 || nvar < 1 || !fcoord ||
      !fine || !coarse)
    return rfail("ath_gpu_restrict_cc: bad argument");
  // coarse cells in the buffer; fine cells i, i+1 of each in the block
  if (!dir_ok(R.i0, R.i1, 0, S.ncc1 - 1, S.ng, 0, 1, S.nc1 - 1) ||
      !dir_ok(R.j0, R.j1, 0, S.ncc2 - 1, S.ng, 0, 1, S.nc2 - 1) ||
      !dir_ok(R.k0, R.k1, 0, S.ncc3 - 1, S.ng, 0, 1, S.nc3 - 1))
    return rfail("ath_gpu_restrict_cc: range outside the arrays");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_restrict_cc, rgrid(R, nb, nvar), dim3(RWG), 0, st, S, R, nb, nvar,
                     fcoord, fine, coarse);
  return launch_check(st);
}

int ath_gpu_restrict_fc(const int *shape, int nb, int comp, const double *fcoord,
                        const double *fine, double *coarse, const int *range, void *stream) {
  RShape S;
  RRange R;
  if (!make_shape(shape, S) || !range_of(range, R) || nb < 1 || comp < 0 || comp > 2 ||
      !fcoord || !fine || !coarse)
    return rfail("ath_gpu_restrict_fc: bad argument");
  // along comp: coarse faces 0..ncc and fine faces 0..nc; across: like cells
  const int lo[3] = {R.i0, R.j0, R.k0}, hi[3] = {R.i1, R.j1, R.k1};
  const int nc[3] = {S.nc1, S.nc2, S.nc3}, ncc[3] = {S.ncc1, S.ncc2, S.ncc3};
  for (int d = 0; d < 3; ++d) {
    const bool ok = (d == comp) ? dir_ok(lo[d], hi[d], 0, ncc[d], S.ng, 0, 0, nc[d])
                                : dir_ok(lo[d], hi[d], 0, ncc[d] - 1, S.ng, 0, 1, nc[d] - 1);
    if (!ok) return rfail("ath_gpu_restrict_fc: range outside the arrays");
  }
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g = rgrid(R, nb, 1);
  if (comp == 0)
    hipLaunchKernelGGL(k_restrict_fc<0>, g, dim3(RWG), 0, st, S, R, nb, fcoord, fine, coarse);
  else if (comp == 1)
    hipLaunchKernelGGL(k_restrict_fc<1>, g, dim3(RWG), 0, st, S, R, nb, fcoord, fine, coarse);
  else
    hipLaunchKernelGGL(k_restrict_fc<2>, g, dim3(RWG), 0, st, S, R, nb, fcoord, fine, coarse);
  return launch_check(st);
}

int ath_gpu_prolongate_cc(const int *shape, int nb, int nvar, const double *fcoord,
                          const double *ccoord, const double *coarse, double *fine,
                          const int *range, void *stream) {
  RShape S;
  RRange R;
  if (!make_shape(shape, S) || !range_of(range, R) || nb < 1 || nvar < 1 || !fcoord ||
      !ccoord || !coarse || !fine)
    return rfail("ath_gpu_prolongate_cc: bad argument");
  // coarse cells with both neighbours in the buffer; fine cells fi, fi+1 in the block
  if (!dir_ok(R.i0, R.i1, 1, S.ncc1 - 2, S.ng, 0, 1, S.nc1 - 1) ||
      !dir_ok(R.j0, R.j1, 1, S.ncc2 - 2, S.ng, 0, 1, S.nc2 - 1) ||
      !dir_ok(R.k0, R.k1, 1, S.ncc3 - 2, S.ng, 0, 1, S.nc3 - 1))
    return rfail("ath_gpu_prolongate_cc: range outside the arrays");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_prolongate_cc, rgrid(R, nb, nvar), dim3(RWG), 0, st, S, R, nb, nvar,
                     fcoord, ccoord, coarse, fine);
  return launch_check(st);
}

int ath_gpu_prolongate_fc_shared(const int *shape, int nb, int comp, const double *fcoord,
                                 const double *ccoord, const double *coarse, double *fine,
                                 const int *range, void *stream) {
  RShape S;
  RRange R;
  if (!make_shape(shape, S) || !range_of(range, R) || nb < 1 || comp < 0 || comp > 2 ||
      !fcoord || !ccoord || !coarse || !fine)
    return rfail("ath_gpu_prolongate_fc_shared: bad argument");
  // along comp: coarse faces 0..ncc onto fine faces 0..nc (no neighbours read); across:
  // coarse positions with both neighbours, fine fi, fi+1 in the block
  const int lo[3] = {R.i0, R.j0, R.k0}, hi[3] = {R.i1, R.j1, R.k1};
  const int nc[3] = {S.nc1, S.nc2, S.nc3}, ncc[3] = {S.ncc1, S.ncc2, S.ncc3};
  for (int d = 0; d < 3; ++d) {
    const bool ok = (d == comp) ? dir_ok(lo[d], hi[d], 0, ncc[d], S.ng, 0, 0, nc[d])
                                : dir_ok(lo[d], hi[d], 1, ncc[d] - 2, S.ng, 0, 1, nc[d] - 1);
    if (!ok) return rfail("ath_gpu_prolongate_fc_shared: range outside the arrays");
  }
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g = rgrid(R, nb, 1);
  if (comp == 0)
    hipLaunchKernelGGL(k_prolongate_fc_shared<0>, g, dim3(RWG), 0, st, S, R, nb, fcoord,
                       ccoord, coarse, fine);
  else if (comp == 1)
    hipLaunchKernelGGL(k_prolongate_fc_shared<1>, g, dim3(RWG), 0, st, S, R, nb, fcoord,
                       ccoord, coarse, fine);
  else
    hipLaunchKernelGGL(k_prolongate_fc_shared<2>, g, dim3(RWG), 0, st, S, R, nb, fcoord,
                       ccoord, coarse, fine);
  return launch_check(st);
}

int ath_gpu_prolongate_fc_internal(const int *shape, int nb, const double *fcoord,
                                   double *b1, double *b2, double *b3, const int *range,
                                   void *stream) {
  RShape S;
  RRange R;
  if (!make_shape(shape, S) || !range_of(range, R) || nb < 1 || !fcoord || !b1 || !b2 ||
      !b3)
    return rfail("ath_gpu_prolongate_fc_internal: bad argument");
  // each coarse cell reads fine faces f .. f+2 along each direction and cells f, f+1
  if (!dir_ok(R.i0, R.i1, 0, S.ncc1 - 1, S.ng, 0, 1, S.nc1 - 1) ||
      !dir_ok(R.j0, R.j1, 0, S.ncc2 - 1, S.ng, 0, 1, S.nc2 - 1) ||
      !dir_ok(R.k0, R.k1, 0, S.ncc3 - 1, S.ng, 0, 1, S.nc3 - 1))
    return rfail("ath_gpu_prolongate_fc_internal: range outside the arrays");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_prolongate_fc_internal, rgrid(R, nb, 1), dim3(RWG), 0, st, S, R, nb,
                     fcoord, b1, b2, b3);
  return launch_check(st);
}

int ath_gpu_restrict_emf(const int *shape, int nb, int comp, const double *fcoord,
                         const double *fine, double *coarse, const int *range, void *stream) {
  RShape S;
  RRange R;
  if (!make_shape(shape, S) || !range_of(range, R) || nb < 1 || comp < 0 || comp > 2 ||
      !fcoord || !fine || !coarse)
    return rfail("ath_gpu_restrict_emf: bad argument");
  // along comp: coarse edges 0..ncc-1 over fine edges f, f+1; across: edge positions
  // at faces, coarse 0..ncc and fine 0..nc
  const int lo[3] = {R.i0, R.j0, R.k0}, hi[3] = {R.i1, R.j1, R.k1};
  const int nc[3] = {S.nc1, S.nc2, S.nc3}, ncc[3] = {S.ncc1, S.ncc2, S.ncc3};
  for (int d = 0; d < 3; ++d) {
    const bool ok = (d == comp) ? dir_ok(lo[d], hi[d], 0, ncc[d] - 1, S.ng, 0, 1, nc[d] - 1)
                                : dir_ok(lo[d], hi[d], 0, ncc[d], S.ng, 0, 0, nc[d]);
    if (!ok) return rfail("ath_gpu_restrict_emf: range outside the arrays");
  }
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g = rgrid(R, nb, 1);
  if (comp == 0)
    hipLaunchKernelGGL(k_restrict_emf<0>, g, dim3(RWG), 0, st, S, R, nb, fcoord, fine, coarse);
  else if (comp == 1)
    hipLaunchKernelGGL(k_restrict_emf<1>, g, dim3(RWG), 0, st, S, R, nb, fcoord, fine, coarse);
  else
    hipLaunchKernelGGL(k_restrict_emf<2>, g, dim3(RWG), 0, st, S, R, nb, fcoord, fine, coarse);
  return launch_check(st);
}

int ath_gpu_tree_build(const int *nrb, const int *bc, const double *xmin, const double *xmax,
                   const ath_refine_region *reg, int nreg, long long *locs, int *nbr,
                   int cap, int *nblocks, int *root_level) {
  if (!nrb || !bc || !xmin || !xmax || (nreg > 0 && !reg) || nreg < 0 || !nblocks)
    return rfail("ath_gpu_tree_build: bad argument");
  Tree T;
  long long nbmax = 1;
  for (int q = 0; q < 3; ++q) {
    if (nrb[q] < 1) return rfail("ath_gpu_tree_build: root blocks per direction must be >= 1");
    if (!(xmin[q] < xmax[q])) return rfail("ath_gpu_tree_build: xmin must be below xmax");
    T.nrb[q] = nrb[q];
    nbmax = (nrb[q] > nbmax) ? nrb[q] : nbmax;
    if ((bc[2 * q] != 0 && bc[2 * q] != 1) || (bc[2 * q + 1] != 0 && bc[2 * q + 1] != 1) ||
        ((bc[2 * q] == 0) != (bc[2 * q + 1] == 0)))
      return rfail("ath_gpu_tree_build: bc flags are 0 (periodic, on both faces) or 1 (outflow)");
    T.periodic[2 * q] = bc[2 * q] == 0;
    T.periodic[2 * q + 1] = bc[2 * q + 1] == 0;
  }
  for (T.root_level = 0; (1LL << T.root_level) < nbmax; ++T.root_level) {}
  for (int r = 0; r < nreg; ++r) {
    const ath_refine_region &R = reg[r];
    if (R.level < 1 || R.level + T.root_level > kTreeMaxLevel)
      return rfail("ath_gpu_tree_build: region level must be in 1 .. " +
                   std::to_string(kTreeMaxLevel - T.root_level));
    for (int q = 0; q < 3; ++q) {
      if (R.xmin[q] > R.xmax[q]) return rfail("ath_gpu_tree_build: invalid refinement region");
      if (R.xmin[q] < xmin[q] || R.xmax[q] > xmax[q])
        return rfail("ath_gpu_tree_build: refinement region must lie inside the mesh");
    }
  }
  {   // the MeshBlocks the regions alone ask for, before any node is made
    double want = static_cast<double>(nrb[0]) * nrb[1] * nrb[2];
    for (int r = 0; r < nreg; ++r) {
      long long lo[3], hi[3];
      double v = 1.0;
      for (int q = 0; q < 3; ++q) {
        region_range(nrb[q], reg[r].level, reg[r].xmin[q], reg[r].xmax[q], xmin[q], xmax[q],
                     &lo[q], &hi[q]);
        v *= static_cast<double>(hi[q] - lo[q]);
      }
      want += v;
    }
    if (want > static_cast<double>(kTreeMaxNodes))
      return rfail("ath_gpu_tree_build: the regions need about " +
                   std::to_string(static_cast<long long>(want)) + " MeshBlocks, more than " +
                   std::to_string(kTreeMaxNodes) + " tree nodes");
  }
  T.make(0, 0, 0, 0);
  T.create_root(0);
  for (int r = 0; r < nreg && !T.overflow; ++r) {
    const ath_refine_region &R = reg[r];
    long long lo[3], hi[3];
    for (int q = 0; q < 3; ++q)
      region_range(nrb[q], R.level, R.xmin[q], R.xmax[q], xmin[q], xmax[q], &lo[q], &hi[q]);
    const int lrlev = R.level + T.root_level;
    for (long long k = lo[2]; k < hi[2] && !T.overflow; k += 2)
      for (long long j = lo[1]; j < hi[1] && !T.overflow; j += 2)
        for (long long i = lo[0]; i < hi[0] && !T.overflow; i += 2) {
          const long long l[3] = {i, j, k};
          T.add(lrlev, l);
        }
  }
  if (T.overflow)
    return rfail("ath_gpu_tree_build: more than " + std::to_string(kTreeMaxNodes) + " tree nodes");
  std::vector<int> lv;
  T.leaves(0, lv);
  *nblocks = static_cast<int>(lv.size());
  if (root_level) *root_level = T.root_level;
  if (static_cast<long>(lv.size()) > cap)
    return rfail("ath_gpu_tree_build: " + std::to_string(lv.size()) + " MeshBlocks, cap " +
                 std::to_string(cap));
  std::vector<int> gid(T.nd.size(), -1);
  for (size_t g = 0; g < lv.size(); ++g) gid[lv[g]] = static_cast<int>(g);
  for (size_t g = 0; g < lv.size(); ++g) {
    const TNode &t = T.nd[lv[g]];
    if (locs) {
      locs[4 * g] = t.level;
      for (int q = 0; q < 3; ++q) locs[4 * g + 1 + q] = t.lx[q];
    }
    if (!nbr) continue;
    for (int o3 = -1; o3 <= 1; ++o3)
      for (int o2 = -1; o2 <= 1; ++o2)
        for (int o1 = -1; o1 <= 1; ++o1) {
          const int b = T.neighbor(lv[g], o1, o2, o3);
          int v;
          if (b == -3) return rfail("ath_gpu_tree_build: neighbour search failed");
          if (b < 0) v = -1;
          else if (!T.nd[b].leaf()) v = -2;
          else v = gid[b];
          nbr[27 * g + (o3 + 1) * 9 + (o2 + 1) * 3 + (o1 + 1)] = v;
        }
  }
  return 0;
}

}  // extern "C"
