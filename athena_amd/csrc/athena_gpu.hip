// athena_gpu.hip -- MI355X (gfx950) implementation of Athena++'s per-MeshBlock MHD
// update behind the C ABI declared in include/athena_gpu.h.
//
// Data layout in HBM: one pool per register, blocks back to back, each block an
// AthenaArray ((n,k,j,i), i fastest).  Kernels take the whole mesh in one launch
// (work-group id -> (MeshBlock, chunk)), so a VL2 stage is five launches regardless of
// block count: fluxes x1/x2/x3 (reconstruct + Riemann) -> corner EMFs -> update (u and
// b, with the same-level EMF averaging folded in) -> ghost exchange (a gather from the
// neighbours' active cells) -> primitive recovery (+ CFL reduction on the last stage).
// The mesh clock (time, dt) lives in device memory, so cycles run without host
// synchronisation.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/athena_gpu.h"
#include "mlp.hpp"
#include "physics.hpp"
#include "refine_ops.hpp"
#include "smr_plan.hpp"

using namespace athgpu;

namespace {

thread_local std::string g_err;

int fail(const std::string &msg) {
  g_err = msg;
  return 1;
}

}  // namespace

// The library's error text, shared with the other translation units (refine.hip).
int athgpu::set_error(const std::string &msg) { return fail(msg); }

namespace {

#define HIPCHK(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(std::string(#expr) + ": " + hipGetErrorString(e_));              \
  } while (0)

constexpr int WG = 256;

// Geometry and constants shared by all kernels of one mesh (passed by value).
struct MeshK {
  int nx1, nx2, nx3, ng;
  int nc1, nc2, nc3;
  int is, ie, js, je, ks, ke;
  int nblocks;
  long ncell;                 // nc1*nc2*nc3
  long nf1, nf2, nf3;         // x1f, x2f, x3f face-array sizes per block
  long ne1, ne2, ne3;         // x1e, x2e, x3e edge-array sizes per block
  int cstride;                // doubles of coordinate table per block
  int coff[3];                // offset of direction d inside the table
  const int *nbr;             // [nblocks][27] same-level neighbour gid, -1 across an
                              // outflow face of the mesh
  int outflow;                // 1: some mesh face is an outflow boundary
  int smr;                    // 1: statically refined mesh (EMFs corrected beforehand)
  const double *coord;        // [nblocks][cstride]: per d: dxf[nc], pf[nc], mf[nc]
  EosK eos;
  double cfl;
  int xcd_remap;              // per-kernel XCD-contiguous numbering, bits XCD_FLUX ..
  int c2p_mode;               // SR primitive recovery: C2PMode
  MlpK mlp;                   // network of C2P_LEARNED
  unsigned long long *c2p_stats;   // null, or C2P_STAT_SLOTS x [evals, fallbacks, cells]
};

// The SR recovery counters are spread over this many slots of three accumulators, chosen
// by wavefront, so that their atomics do not queue on one address; a read sums the slots.
constexpr int C2P_STAT_SLOTS = 4096;

struct Clock {
  double time, dt, tlim;
  double ncycle;
};

// Element indices within one block's array, written as (uniform plane term) + (row
// term): with k uniform across a work-group the plane term stays in scalar registers.
__device__ __forceinline__ long cidx(const MeshK &M, int k, int j, int i) {
  return static_cast<long>(k) * (M.nc2 * M.nc1) + (j * M.nc1 + i);
}
__device__ __forceinline__ long f1idx(const MeshK &M, int k, int j, int i) {
  return static_cast<long>(k) * (M.nc2 * (M.nc1 + 1)) + (j * (M.nc1 + 1) + i);
}
__device__ __forceinline__ long f2idx(const MeshK &M, int k, int j, int i) {
  return static_cast<long>(k) * ((M.nc2 + 1) * M.nc1) + (j * M.nc1 + i);
}
__device__ __forceinline__ long f3idx(const MeshK &M, int k, int j, int i) {
  return static_cast<long>(k) * (M.nc2 * M.nc1) + (j * M.nc1 + i);
}
// edge arrays: x1e (nc3+1, nc2+1, nc1), x2e (nc3+1, nc2, nc1+1), x3e (nc3, nc2+1, nc1+1)
__device__ __forceinline__ long e1idx(const MeshK &M, int k, int j, int i) {
  return static_cast<long>(k) * ((M.nc2 + 1) * M.nc1) + (j * M.nc1 + i);
}
__device__ __forceinline__ long e2idx(const MeshK &M, int k, int j, int i) {
  return static_cast<long>(k) * (M.nc2 * (M.nc1 + 1)) + (j * (M.nc1 + 1) + i);
}
__device__ __forceinline__ long e3idx(const MeshK &M, int k, int j, int i) {
  return static_cast<long>(k) * ((M.nc2 + 1) * (M.nc1 + 1)) + (j * (M.nc1 + 1) + i);
}
__device__ __forceinline__ int nbr(const MeshK &M, int g, int o1, int o2, int o3) {
  return M.nbr[g * 27 + (o3 + 1) * 9 + (o2 + 1) * 3 + (o1 + 1)];
}
__device__ __forceinline__ const double *ctab(const MeshK &M, int g, int d) {
  return M.coord + static_cast<long>(g) * M.cstride + M.coff[d];
}

// XCD-aware work-group numbering.  The hardware hands work-groups to the 8 XCDs
// round-robin by id, and each XCD has its own L2; renumbering so that the ids one XCD
// receives form one contiguous logical range keeps a MeshBlock (and its stencil
// neighbours) inside one L2.  Bijective for any grid size.
constexpr int NXCD = 8;
// Bits of MeshK::xcd_remap (environment variable ATH_XCD_REMAP, a decimal mask; 31 remaps
// every kernel): which kernels number their work-groups XCD-contiguously.
enum XcdBits { XCD_FLUX = 1, XCD_CORNER = 2, XCD_UPDATE = 4, XCD_GHOST = 8, XCD_C2P = 16 };

__device__ __forceinline__ int xcd_logical_wg(int remap) {
  const int w = blockIdx.x, n = gridDim.x;
  if (!remap) return w;
  const int q = n / NXCD, r = n % NXCD;
  const int x = w % NXCD, l = w / NXCD;
  return x * q + (x < r ? x : r) + l;
}

// Work decomposition of one launch over a box of ni x nj x nk positions per MeshBlock:
// nch work-groups per block (nwg = nblocks * nch), each either a linear chunk of WG
// positions, i fastest (tx == 0), or a tx x ty x tz tile (tx*ty*tz == WG).  Only the
// thread-to-position mapping changes; every position is computed by the same arithmetic.
struct TileK {
  int tx, ty, tz;
};

__device__ __forceinline__ bool map_work(int lw, int nwg, int nblocks, int ni, int nj, int nk,
                                         TileK T, int &g, int &i, int &j, int &k) {
  const int nch = nwg / nblocks;
  g = lw / nch;
  const int c = lw - g * nch;
  if (T.tx <= 0) {
    const int t = c * WG + threadIdx.x;
    i = t % ni;
    const int r = t / ni;
    j = r % nj;
    k = r / nj;
    return t < ni * nj * nk;
  }
  const int nti = (ni + T.tx - 1) / T.tx, ntj = (nj + T.ty - 1) / T.ty;
  const int ti = c % nti, r = c / nti;
  i = ti * T.tx + static_cast<int>(threadIdx.x) % T.tx;
  j = (r % ntj) * T.ty + (static_cast<int>(threadIdx.x) / T.tx) % T.ty;
  k = (r / ntj) * T.tz + static_cast<int>(threadIdx.x) / (T.tx * T.ty);
  return i < ni && j < nj && k < nk;
}

// Position boxes of the stage kernels per MeshBlock (first index lo, extent n):
//   fluxes:   cells [is-1, ie+1]^3, each computing the faces on its low sides that lie in
//             the reference's ranges (calculate_fluxes.cpp:68-76, 176-182, 299-302);
//   corner:   edges [is, ie+1]^3 (calculate_corner_e.cpp:187-229);
//   update:   [is, ie+1]^3: the active cell and its three low faces (CT ranges);
//   cons2prim: every cell including ghosts.
enum WorkKind { KK_FLUX = 0, KK_CORNER = 1, KK_UPDATE = 2, KK_C2P = 3, KK_N = 4 };

__host__ __device__ __forceinline__ void kind_box(const MeshK &M, int kind, int *lo, int *n) {
  if (kind == KK_FLUX) {
    lo[0] = M.is - 1; lo[1] = M.js - 1; lo[2] = M.ks - 1;
    n[0] = M.nx1 + 2; n[1] = M.nx2 + 2; n[2] = M.nx3 + 2;
  } else if (kind == KK_C2P) {
    lo[0] = lo[1] = lo[2] = 0;
    n[0] = M.nc1; n[1] = M.nc2; n[2] = M.nc3;
  } else {
    lo[0] = M.is; lo[1] = M.js; lo[2] = M.ks;
    n[0] = M.nx1 + 1; n[1] = M.nx2 + 1; n[2] = M.nx3 + 1;
  }
}

// map_work over kind's box; (i, j, k) are block cell indices.
__device__ __forceinline__ bool map_kind(const MeshK &M, int kind, int lw, int nwg, TileK T,
                                         int &g, int &i, int &j, int &k) {
  int lo[3], n[3];
  kind_box(M, kind, lo, n);
  const bool ok = map_work(lw, nwg, M.nblocks, n[0], n[1], n[2], T, g, i, j, k);
  i += lo[0];
  j += lo[1];
  k += lo[2];
  return ok;
}

//========================================================================================
// Reconstruction of one variable at one interface along DIR.
//   q[] holds the variable at cells c-3 .. c+2 (q[3] is the cell on the right of the
//   face, q[2] the cell on the left).  Produces the left state (right edge of cell c-1)
//   and right state (left edge of cell c).
//   ORDER 1: donor cell (dc.cpp:24-103).
//   ORDER 2: PLM, uniform VL limiter (plm.cpp:105-112, faces :149-155 / :300-308).
//   ORDER 3: PPM + Colella-Sekora limiter (ppm.cpp, uniform Cartesian branch).

__device__ __forceinline__ double vl_slope(double qm, double q0, double qp) {
  double dwl = (q0 - qm);
  double dwr = (qp - q0);
  double dw2 = dwl * dwr;
  double dwm = 2.0 * dw2 / (dwl + dwr);
  if (dw2 <= 0.0) dwm = 0.0;
  return dwm;
}

// PPM edge values of the cell q[2] given q[0..4] = cells (c-2..c+2):
// returns (qminus, qplus) = (left-edge, right-edge) values after limiting.
__device__ __forceinline__ void ppm_cell(const double *q, double &qminus, double &qplus) {
  const double C2 = 1.25;
  const double c1 = 0.5, c2 = 0.5, c3 = 0.5, c4 = 0.5, c5 = 1.0 / 6.0, c6 = -1.0 / 6.0;
  const double q_im2 = q[0], q_im1 = q[1], q0 = q[2], q_ip1 = q[3], q_ip2 = q[4];
  double qa = (q0 - q_im1);
  double qb = (q_ip1 - q0);
  double dd_im1 = c1 * qa + c2 * (q_im1 - q_im2);
  double dd = c1 * qb + c2 * qa;
  double dd_ip1 = c1 * (q_ip2 - q_ip1) + c2 * qb;
  double dph = (c3 * q_im1 + c4 * q0) + (c5 * dd_im1 + c6 * dd);
  double dph_ip1 = (c3 * q0 + c4 * q_ip1) + (c5 * dd + c6 * dd_ip1);
  double d2qc_im1 = q_im2 + q0 - 2.0 * q_im1;
  double d2qc = q_im1 + q_ip1 - 2.0 * q0;
  double d2qc_ip1 = q0 + q_ip2 - 2.0 * q_ip1;
  {
    double qa_tmp = dph - q_im1, qb_tmp = q0 - dph;
    double a = 3.0 * (q_im1 + q0 - 2.0 * dph), b = d2qc_im1, c = d2qc, qd = 0.0;
    if (sgn(a) == sgn(b) && sgn(a) == sgn(c))
      qd = sgn(a) * dmin(C2 * fabs(b), dmin(C2 * fabs(c), fabs(a)));
    double t = 0.5 * (q_im1 + q0) - qd / 6.0;
    if (qa_tmp * qb_tmp < 0.0) dph = t;
  }
  {
    double qa_tmp = dph_ip1 - q0, qb_tmp = q_ip1 - dph_ip1;
    double a = 3.0 * (q0 + q_ip1 - 2.0 * dph_ip1), b = d2qc, c = d2qc_ip1, qd = 0.0;
    if (sgn(a) == sgn(b) && sgn(a) == sgn(c))
      qd = sgn(a) * dmin(C2 * fabs(b), dmin(C2 * fabs(c), fabs(a)));
    double t = 0.5 * (q0 + q_ip1) - qd / 6.0;
    if (qa_tmp * qb_tmp < 0.0) dph_ip1 = t;
  }
  double d2qf = 6.0 * (dph + dph_ip1 - 2.0 * q0);
  qminus = dph;
  qplus = dph_ip1;
  double dqf_minus = q0 - qminus;
  double dqf_plus = qplus - q0;
  double qa_tmp = dqf_minus * dqf_plus;
  double qb_tmp = (q_ip1 - q0) * (q0 - q_im1);
  double a = d2qc_im1, b = d2qc, c = d2qc_ip1, qd = d2qf, qe = 0.0;
  if (sgn(a) == sgn(b) && sgn(a) == sgn(c) && sgn(a) == sgn(qd))
    qe = sgn(qd) * dmin(dmin(C2 * fabs(a), C2 * fabs(b)), dmin(C2 * fabs(c), fabs(qd)));
  a = dmax(fabs(q_im1), fabs(q_im2));
  b = dmax(dmax(fabs(q0), fabs(q_ip1)), fabs(q_ip2));
  double rho = 0.0;
  if (fabs(qd) > (1.0e-12) * dmax(a, b)) rho = qe / qd;
  double tmp_m = q0 - rho * dqf_minus;
  double tmp_p = q0 + rho * dqf_plus;
  double tmp2_m = q0 - 2.0 * dqf_plus;
  double tmp2_p = q0 + 2.0 * dqf_minus;
  if ((qa_tmp <= 0.0 || qb_tmp <= 0.0)) {
    if (rho <= (1.0 - (1.0e-12))) {
      qminus = tmp_m;
      qplus = tmp_p;
    }
  } else {
    if (fabs(dqf_minus) >= 2.0 * fabs(dqf_plus)) qminus = tmp2_m;
    if (fabs(dqf_plus) >= 2.0 * fabs(dqf_minus)) qplus = tmp2_p;
  }
}

//========================================================================================
// Kernel 1: interface fluxes (Hydro::CalculateFluxes + the Riemann solver +
// GetWeightForCT).  flux_face computes one interface along DIR; k_flux3 runs one
// thread per cell of the union range and computes the three faces on the cell's low
// sides, so w and bcc are fetched from HBM once per stage instead of once per
// direction.  k_flux keeps the one-direction form for the per-function entry.

// Per-direction output registers.
struct FaceOut {
  const double *bn;   // normal face field b.xNf
  double *flx, *ey, *ez, *wct;
};

template <int DIR, int ORDER, int RS>
__device__ __forceinline__ void face_solve(const MeshK &M, int g, int k, int j, int i,
                                           double *wl, double *wr, const FaceOut &o,
                                           double dt, bool emit, double *fx);

// One interface along DIR at face (k, j, i) of block g: reconstruction, the Riemann
// solver, and the face EMFs and CT weight of that face written to o (when emit); the five
// fluxes are returned in fx[] in conserved order (IDN, IM1, IM2, IM3, IEN).
template <int DIR, int ORDER, int RS>
__device__ __forceinline__ void face_flux(const MeshK &M, int g, int k, int j, int i,
                                          const double *__restrict__ wb,
                                          const double *__restrict__ bb, const FaceOut &o,
                                          double dt, bool emit, double *fx) {
  const long s = (DIR == 0) ? 1 : ((DIR == 1) ? M.nc1 : static_cast<long>(M.nc1) * M.nc2);
  const long c = cidx(M, k, j, i);
  constexpr int ivx = IVX + DIR;
  constexpr int ivy = IVX + ((ivx - IVX) + 1) % 3;
  constexpr int ivz = IVX + ((ivx - IVX) + 2) % 3;
  constexpr int cby = (DIR == 0) ? IB2 : ((DIR == 1) ? IB3 : IB1);
  constexpr int cbz = (DIR == 0) ? IB3 : ((DIR == 1) ? IB1 : IB2);
  const int fc = (DIR == 0) ? i : ((DIR == 1) ? j : k);    // face/cell index along DIR
  const double *ct = ctab(M, g, DIR);
  const int ncd = (DIR == 0) ? M.nc1 : ((DIR == 1) ? M.nc2 : M.nc3);

  double wl[NWAVE], wr[NWAVE];
#pragma unroll
  for (int v = 0; v < NWAVE; ++v) {
    const double *src;
    if (v == IDN) src = wb;
    else if (v == IVX) src = wb + ivx * M.ncell;
    else if (v == IVY) src = wb + ivy * M.ncell;
    else if (v == IVZ) src = wb + ivz * M.ncell;
    else if (v == IPR) src = wb + IPR * M.ncell;
    else if (v == IBY) src = bb + cby * M.ncell;
    else src = bb + cbz * M.ncell;
    if (ORDER == 1) {
      wl[v] = src[c - s];
      wr[v] = src[c];
    } else if (ORDER == 2) {
      double qm2 = src[c - 2 * s], qm1 = src[c - s], q0 = src[c], qp1 = src[c + s];
      double dl = vl_slope(qm2, qm1, q0);
      double dr = vl_slope(qm1, q0, qp1);
      wl[v] = qm1 + ct[ncd + (fc - 1)] * dl;        // pf of cell fc-1
      wr[v] = q0 - ct[2 * ncd + fc] * dr;           // mf of cell fc
    } else {
      double q[6];
#pragma unroll
      for (int m = 0; m < 6; ++m) q[m] = src[c + (m - 3) * s];
      double qmL, qpL, qmR, qpR;
      ppm_cell(q, qmL, qpL);        // cell c-1: q[0..4]
      ppm_cell(q + 1, qmR, qpR);    // cell c:   q[1..5]
      wl[v] = qpL;
      wr[v] = qmR;
    }
  }
  face_solve<DIR, ORDER, RS>(M, g, k, j, i, wl, wr, o, dt, emit, fx);
}

// The rest of one interface once its states are reconstructed: the primitive floors
// (PPM), the Riemann solver, the fluxes returned in fx[] in conserved order, and the face
// EMFs and CT weight written to o (when emit).  wl and wr are in the rotated order of DIR.
template <int DIR, int ORDER, int RS>
__device__ __forceinline__ void face_solve(const MeshK &M, int g, int k, int j, int i,
                                           double *wl, double *wr, const FaceOut &o,
                                           double dt, bool emit, double *fx) {
  constexpr int ivx = IVX + DIR;
  constexpr int ivy = IVX + ((ivx - IVX) + 1) % 3;
  constexpr int ivz = IVX + ((ivx - IVX) + 2) % 3;
  const int fc = (DIR == 0) ? i : ((DIR == 1) ? j : k);
  const double *ct = ctab(M, g, DIR);
  if (ORDER == 3) {   // ApplyPrimitiveFloors on both reconstructed states
    wl[IDN] = (wl[IDN] > M.eos.dfloor) ? wl[IDN] : M.eos.dfloor;
    wl[IPR] = (wl[IPR] > M.eos.pfloor) ? wl[IPR] : M.eos.pfloor;
    wr[IDN] = (wr[IDN] > M.eos.dfloor) ? wr[IDN] : M.eos.dfloor;
    wr[IPR] = (wr[IPR] > M.eos.pfloor) ? wr[IPR] : M.eos.pfloor;
  }

  long fi;
  if (DIR == 0) fi = f1idx(M, k, j, i);
  else if (DIR == 1) fi = f2idx(M, k, j, i);
  else fi = f3idx(M, k, j, i);
  const long nf = (DIR == 0) ? M.nf1 : ((DIR == 1) ? M.nf2 : M.nf3);
  const double bx = o.bn[static_cast<long>(g) * nf + fi];

  double f[NWAVE];
  if (RS == 0) hlld(wl, wr, bx, M.eos, f);
  else if (RS == 1) hlle(wl, wr, bx, M.eos, f);
  else if (RS == 2) sr_hlle(wl, wr, bx, M.eos, f);
  else sr_hlld(wl, wr, bx, M.eos, f);

  fx[IDN] = f[IDN];
  fx[ivx] = f[IVX];
  fx[ivy] = f[IVY];
  fx[ivz] = f[IVZ];
  fx[IEN] = f[IEN];
  if (emit) {
    o.ey[static_cast<long>(g) * nf + fi] = -f[IBY];
    o.ez[static_cast<long>(g) * nf + fi] = f[IBZ];
    const double dxw = ct[fc];
    o.wct[static_cast<long>(g) * nf + fi] = ct_weight(f[IDN], wl[IDN], wr[IDN], dxw, dt);
  }
}

// face_flux with the fluxes written to Hydro::flux[DIR] (o.flx) as well.
template <int DIR, int ORDER, int RS>
__device__ __forceinline__ void flux_face(const MeshK &M, int g, int k, int j, int i,
                                          const double *__restrict__ wb,
                                          const double *__restrict__ bb, const FaceOut &o,
                                          double dt) {
  double fx[NHYDRO];
  face_flux<DIR, ORDER, RS>(M, g, k, j, i, wb, bb, o, dt, true, fx);
  long fi;
  if (DIR == 0) fi = f1idx(M, k, j, i);
  else if (DIR == 1) fi = f2idx(M, k, j, i);
  else fi = f3idx(M, k, j, i);
  const long nf = (DIR == 0) ? M.nf1 : ((DIR == 1) ? M.nf2 : M.nf3);
  double *fb = o.flx + static_cast<long>(g) * NHYDRO * nf;
#pragma unroll
  for (int n = 0; n < NHYDRO; ++n) fb[n * nf + fi] = fx[n];
}

template <int DIR, int ORDER, int RS>
__global__ void __launch_bounds__(WG) k_flux(MeshK M, const double *__restrict__ w,
                                              const double *__restrict__ bcc, FaceOut o,
                                              const Clock *__restrict__ clk) {
  const int lw = xcd_logical_wg(M.xcd_remap & XCD_FLUX);
  const int nch = gridDim.x / M.nblocks;
  const int g = lw / nch;
  int ni, nj, nk, i0, j0, k0;
  if constexpr (DIR == 0) { ni = M.nx1 + 1; nj = M.nx2 + 2; nk = M.nx3 + 2; i0 = M.is; j0 = M.js - 1; k0 = M.ks - 1; }
  else if constexpr (DIR == 1) { ni = M.nx1 + 2; nj = M.nx2 + 1; nk = M.nx3 + 2; i0 = M.is - 1; j0 = M.js; k0 = M.ks - 1; }
  else { ni = M.nx1 + 2; nj = M.nx2 + 2; nk = M.nx3 + 1; i0 = M.is - 1; j0 = M.js - 1; k0 = M.ks; }
  const int t = (lw % nch) * WG + threadIdx.x;
  if (t >= ni * nj * nk) return;
  const int i = i0 + t % ni;
  const int r = t / ni;
  const int j = j0 + r % nj;
  const int k = k0 + r / nj;
  flux_face<DIR, ORDER, RS>(M, g, k, j, i, w + static_cast<long>(g) * NHYDRO * M.ncell,
                            bcc + static_cast<long>(g) * 3 * M.ncell, o, clk->dt);
}

// One thread per cell of [is-1,ie+1]x[js-1,je+1]x[ks-1,ke+1]; each computes the faces on
// its low sides that lie in the reference's loop ranges (calculate_fluxes.cpp:68-76,
// 176-182, 299-302).
template <int ORDER, int RS>
__device__ __forceinline__ void flux3_body(const MeshK &M, const double *__restrict__ w,
                                           const double *__restrict__ bcc, const FaceOut &o1,
                                           const FaceOut &o2, const FaceOut &o3, double dt,
                                           int lw, int nwg, TileK T) {
  int g, i, j, k;
  if (!map_kind(M, KK_FLUX, lw, nwg, T, g, i, j, k)) return;
  const double *wb = w + static_cast<long>(g) * NHYDRO * M.ncell;
  const double *bb = bcc + static_cast<long>(g) * 3 * M.ncell;
  if (i >= M.is) flux_face<0, ORDER, RS>(M, g, k, j, i, wb, bb, o1, dt);
  if (j >= M.js) flux_face<1, ORDER, RS>(M, g, k, j, i, wb, bb, o2, dt);
  if (k >= M.ks) flux_face<2, ORDER, RS>(M, g, k, j, i, wb, bb, o3, dt);
}

template <int ORDER, int RS>
__global__ void __launch_bounds__(WG) k_flux3(MeshK M, const double *__restrict__ w,
                                               const double *__restrict__ bcc, FaceOut o1,
                                               FaceOut o2, FaceOut o3,
                                               const Clock *__restrict__ clk, TileK T) {
  flux3_body<ORDER, RS>(M, w, bcc, o1, o2, o3, clk->dt, xcd_logical_wg(M.xcd_remap & XCD_FLUX),
                        gridDim.x, T);
}

// Reconstruction of cell (k, j, i) along DIR: its low-face value qm[v] and high-face value
// qp[v] for the seven face-state variables in the rotated order of DIR.  The interface at
// index fc along DIR takes wl = qp of cell fc-1 and wr = qm of cell fc, the same
// expressions face_flux evaluates (PLM: pf and mf of plm.cpp:149-155; PPM: ppm_cell).
template <int DIR, int ORDER>
__device__ __forceinline__ void recon_cell(const MeshK &M, int g, int k, int j, int i,
                                           const double *__restrict__ wb,
                                           const double *__restrict__ bb, double *qm,
                                           double *qp) {
  const long s = (DIR == 0) ? 1 : ((DIR == 1) ? M.nc1 : static_cast<long>(M.nc1) * M.nc2);
  const long c = cidx(M, k, j, i);
  constexpr int ivx = IVX + DIR;
  constexpr int ivy = IVX + ((ivx - IVX) + 1) % 3;
  constexpr int ivz = IVX + ((ivx - IVX) + 2) % 3;
  constexpr int cby = (DIR == 0) ? IB2 : ((DIR == 1) ? IB3 : IB1);
  constexpr int cbz = (DIR == 0) ? IB3 : ((DIR == 1) ? IB1 : IB2);
  const int fc = (DIR == 0) ? i : ((DIR == 1) ? j : k);
  const double *ct = ctab(M, g, DIR);
  const int ncd = (DIR == 0) ? M.nc1 : ((DIR == 1) ? M.nc2 : M.nc3);
#pragma unroll
  for (int v = 0; v < NWAVE; ++v) {
    const double *src;
    if (v == IDN) src = wb;
    else if (v == IVX) src = wb + ivx * M.ncell;
    else if (v == IVY) src = wb + ivy * M.ncell;
    else if (v == IVZ) src = wb + ivz * M.ncell;
    else if (v == IPR) src = wb + IPR * M.ncell;
    else if (v == IBY) src = bb + cby * M.ncell;
    else src = bb + cbz * M.ncell;
    if (ORDER == 1) {
      qm[v] = qp[v] = src[c];
    } else if (ORDER == 2) {
      const double q0 = src[c];
      const double d = vl_slope(src[c - s], q0, src[c + s]);
      qp[v] = q0 + ct[ncd + fc] * d;
      qm[v] = q0 - ct[2 * ncd + fc] * d;
    } else {
      double q[5];
#pragma unroll
      for (int m = 0; m < 5; ++m) q[m] = src[c + (m - 2) * s];
      ppm_cell(q, qm[v], qp[v]);
    }
  }
}

// Work split of the z-march flux kernel: each MeshBlock's flux box (ni x nj columns of
// nk planes) in nseg segments of nearly equal height; positions run over (block, segment,
// column), column = (j - js + 1) * ni + (i - is + 1).  A wavefront owns FZM_OWN of them
// and its lane 0 repeats the previous wavefront's last position.
constexpr int FZM_OWN = 63;

struct FzmGeom {
  int ni, nj, nk, nseg, hmax;   // box, segments per block, tallest segment
  long ncol, npos, nwave;
};

__host__ __device__ __forceinline__ FzmGeom fzm_geom(const MeshK &M, int seg) {
  FzmGeom G;
  G.ni = M.nx1 + 2;
  G.nj = M.nx2 + 2;
  G.nk = M.nx3 + 2;
  G.nseg = (G.nk + seg - 1) / seg;
  G.hmax = (G.nk + G.nseg - 1) / G.nseg;
  G.ncol = static_cast<long>(G.ni) * G.nj;
  G.npos = G.ncol * G.nseg * M.nblocks;
  G.nwave = (G.npos + FZM_OWN - 1) / FZM_OWN;
  return G;
}

// Position of lane `lane` of wavefront `wv`: block g, column (j, i) and planes k0..k1;
// false for lanes past the last position (they compute a clamped copy and store nothing).
__host__ __device__ __forceinline__ bool fzm_pos(const MeshK &M, const FzmGeom &G, long wv,
                                                 int lane, int &g, int &j, int &i, int &k0,
                                                 int &k1) {
  long p = wv * FZM_OWN + lane - 1;
  const bool in = p < G.npos;
  p = (p < 0) ? 0 : (in ? p : G.npos - 1);
  const long bs = p / G.ncol;
  const int col = static_cast<int>(p - bs * G.ncol);
  g = static_cast<int>(bs / G.nseg);
  const int sg = static_cast<int>(bs - static_cast<long>(g) * G.nseg);
  j = M.js - 1 + col / G.ni;
  i = M.is - 1 + col % G.ni;
  k0 = M.ks - 1 + (sg * G.nk) / G.nseg;
  k1 = M.ks - 1 + ((sg + 1) * G.nk) / G.nseg - 1;
  return in;
}

// Kernel 1, z-march form (tile option of the flux kind).  Each lane owns a column (j, i)
// of the flux box of one MeshBlock and walks a segment of its planes, computing at each
// plane the same faces as k_flux3 for cell (k, j, i).  Lanes run i fastest over
// consecutive columns, so the lane before holds cell (k, j, i-1) whenever the x1 face
// (k, j, i) exists (i >= is): every lane reconstructs its own cell along x1 and passes its
// high-face states one lane up, and the x1 face takes its left states from there.  Along
// x3 the lane keeps the high-face states of cell (k-1, j, i) from the previous plane.  So
// each cell is reconstructed once along x1 and x3 instead of twice; along x2 both cells of
// the face are reconstructed by the lane, as in k_flux3.  Every state is the value
// face_flux computes for that face, so the fluxes, EMFs and weights are the same bits.
// Not bounded to k_flux3's 3 wavefronts per SIMD: at 168 VGPRs the spills cost more
// (PLM launch 3.40 against 2.84 ms at 256^3) than the third wavefront gains.
template <int ORDER, int RS>
__global__ void __launch_bounds__(WG) k_flux_zm(MeshK M, const double *__restrict__ w,
                                                 const double *__restrict__ bcc, FaceOut o1,
                                                 FaceOut o2, FaceOut o3,
                                                 const Clock *__restrict__ clk, int seg) {
  const FzmGeom G = fzm_geom(M, seg);
  const int lane = static_cast<int>(threadIdx.x) & 63;
  const long wv = static_cast<long>(blockIdx.x) * (WG / 64) + (threadIdx.x >> 6);
  if (wv >= G.nwave) return;   // whole wavefronts only
  int g0, j0, i0, k0, k1;
  const bool in = fzm_pos(M, G, wv, lane, g0, j0, i0, k0, k1);
  const bool own = in && lane > 0;
  const double dt = clk->dt;

  double p3[NWAVE] = {};   // high-face x3 states of cell (k-1, j, i)
  if (k0 >= M.ks) {
    double m3[NWAVE];
    recon_cell<2, ORDER>(M, g0, k0 - 1, j0, i0, w + static_cast<long>(g0) * NHYDRO * M.ncell,
                         bcc + static_cast<long>(g0) * 3 * M.ncell, m3, p3);
  }
#pragma unroll 1
  for (int h = 0; h < G.hmax; ++h) {
    const bool live = own && k0 + h <= k1;
    int k = (k0 + h <= k1) ? k0 + h : k1;
    int g = g0, j = j0, i = i0;
    // opaque indices: the stencil addresses are formed afresh at each plane instead of
    // being held across the loop as per-lane pointers, which would take dozens of
    // registers and cost occupancy
    asm volatile("" : "+v"(k), "+v"(g), "+v"(j), "+v"(i));
    const double *wb = w + static_cast<long>(g) * NHYDRO * M.ncell;
    const double *bb = bcc + static_cast<long>(g) * 3 * M.ncell;
    double wl[NWAVE], wr[NWAVE], fx[NHYDRO];
    // x1: this lane's cell, left states from the lane before
    recon_cell<0, ORDER>(M, g, k, j, i, wb, bb, wr, wl);
#pragma unroll
    for (int v = 0; v < NWAVE; ++v) wl[v] = __shfl_up(wl[v], 1, 64);
    if (live && i >= M.is) {
      face_solve<0, ORDER, RS>(M, g, k, j, i, wl, wr, o1, dt, true, fx);
      const long fi = f1idx(M, k, j, i);
      double *fb = o1.flx + static_cast<long>(g) * NHYDRO * M.nf1;
#pragma unroll
      for (int n = 0; n < NHYDRO; ++n) fb[n * M.nf1 + fi] = fx[n];
    }
    // x2: both cells (indices made opaque again, so that the x1 addresses are not kept
    // alive through the x1 solve for reuse here)
    asm volatile("" : "+v"(k), "+v"(g), "+v"(j), "+v"(i));
    wb = w + static_cast<long>(g) * NHYDRO * M.ncell;
    bb = bcc + static_cast<long>(g) * 3 * M.ncell;
    if (live && j >= M.js) {
      double t[NWAVE];
      recon_cell<1, ORDER>(M, g, k, j - 1, i, wb, bb, t, wl);
      recon_cell<1, ORDER>(M, g, k, j, i, wb, bb, wr, t);
      face_solve<1, ORDER, RS>(M, g, k, j, i, wl, wr, o2, dt, true, fx);
      const long fi = f2idx(M, k, j, i);
      double *fb = o2.flx + static_cast<long>(g) * NHYDRO * M.nf2;
#pragma unroll
      for (int n = 0; n < NHYDRO; ++n) fb[n * M.nf2 + fi] = fx[n];
    }
    // x3: low states of this cell, high states of the cell below carried in p3
    asm volatile("" : "+v"(k), "+v"(g), "+v"(j), "+v"(i));
    wb = w + static_cast<long>(g) * NHYDRO * M.ncell;
    bb = bcc + static_cast<long>(g) * 3 * M.ncell;
    recon_cell<2, ORDER>(M, g, k, j, i, wb, bb, wr, wl);
    if (live && k >= M.ks) {
      face_solve<2, ORDER, RS>(M, g, k, j, i, p3, wr, o3, dt, true, fx);
      const long fi = f3idx(M, k, j, i);
      double *fb = o3.flx + static_cast<long>(g) * NHYDRO * M.nf3;
#pragma unroll
      for (int n = 0; n < NHYDRO; ++n) fb[n * M.nf3 + fi] = fx[n];
    }
#pragma unroll
    for (int v = 0; v < NWAVE; ++v) p3[v] = wl[v];
  }
}

//========================================================================================
// Kernel 1 (production form): the fluxes fused with Hydro::AddFluxDivergence
// (add_flux_divergence.cpp:39-96).  A work-group owns a band of rows of one MeshBlock and
// marches up its planes k = ks-1 .. ke+1; thread t owns the column (j, i) of row
// rho = t / (nx1+2), column c = t % (nx1+2) of the band (i = is-1+c).  At plane k the
// thread computes the faces on the low sides of cell (k, j, i) that lie in the
// reference's flux ranges (calculate_fluxes.cpp:68-76, 176-182, 299-302), as k_flux3
// does per cell, and writes their EMFs and CT weights as k_flux3 does; one more row of
// threads computes the x2 faces of the next band's first row, which the update of the
// band's last row needs (the next band writes their EMFs).  The fluxes themselves never
// leave the chip: the x3 flux stays with its column, so the thread that owns an active
// cell first finishes cell (k-1, j, i) with it, then the x1 and x2 fluxes go through LDS
// to start cell (k, j, i), whose x1 + x2 divergence terms and low x3 flux are carried in
// LDS to the next plane.  The divergence is summed in the reference's order (x1, x2, x3
// terms), so U comes out bit for bit as from k_update; Hydro::flux is not written.
struct FdivBand {
  int ja, jb;     // active rows whose cells the band updates
  int r0, r1;     // rows of its x1 and x3 faces (the flux box rows of the band)
  int q1;         // row of the extra x2 faces (jb + 1); computed by the band iff q1 > r1
};

// Threads per row of a band and the band count of a block with bands of up to `rows`
// active rows: as even as possible, nband = ceil(nx2 / rows).
__host__ __device__ __forceinline__ int fdiv_nband(const MeshK &M, int rows) {
  return (M.nx2 + rows - 1) / rows;
}

__host__ __device__ __forceinline__ FdivBand fdiv_band(const MeshK &M, int rows, int b) {
  const int nband = fdiv_nband(M, rows);
  FdivBand B;
  B.ja = M.js + (b * M.nx2) / nband;
  B.jb = M.js + ((b + 1) * M.nx2) / nband - 1;
  B.r0 = (b == 0) ? M.js - 1 : B.ja;
  B.r1 = (b == nband - 1) ? M.je + 1 : B.jb;
  B.q1 = B.jb + 1;
  return B;
}

// Thread rows a band of `rows` active rows needs at most: its ghost row (edge bands)
// and the extra x2 row.
__host__ __device__ __forceinline__ int fdiv_thread_rows(int rows) { return rows + 2; }

// LDS doubles of one work-group: per thread the x1 and x2 fluxes of its column at the
// current plane and the carry of its owned cell (x1 + x2 terms, low x3 flux).
constexpr int FDIV_LDS_DOUBLES = 4 * NHYDRO * WG;

// The duties of thread t of band B.
struct FdivThread {
  bool live;          // a column of the band (or of its extra x2 row)
  int rho, c, j, i;   // band row, column, block row and column
  bool x1, x2, x3;    // faces it computes (x3 only for k >= ks)
  bool emit2;         // writes the EMFs and weight of its x2 face
  bool own;           // updates cell (., j, i)
};

__host__ __device__ __forceinline__ FdivThread fdiv_thread(const MeshK &M, const FdivBand &B,
                                                           int t) {
  FdivThread T;
  const int ni = M.nx1 + 2;
  const int R = B.r1 - B.r0 + 1;
  const int nrow = R + ((B.q1 > B.r1) ? 1 : 0);
  T.rho = t / ni;
  T.c = t - T.rho * ni;
  T.live = T.rho < nrow;
  T.i = M.is - 1 + T.c;
  T.j = (T.rho < R) ? B.r0 + T.rho : B.q1;
  T.x1 = T.live && T.rho < R && T.c >= 1;
  T.x2 = T.live && T.j >= M.js;
  T.x3 = T.live && T.rho < R;
  T.emit2 = T.j <= B.r1;
  T.own = T.live && T.rho < R && T.j >= B.ja && T.j <= B.jb && T.i >= M.is && T.i <= M.ie;
  return T;
}

template <int ORDER, int RS>
__global__ void __launch_bounds__(WG) k_flux_div(MeshK M, const double *__restrict__ w,
                                                     const double *__restrict__ bcc,
                                                     FaceOut o1, FaceOut o2, FaceOut o3,
                                                     const double *ubase, double *uout,
                                                     double beta, int first_stage,
                                                     const Clock *__restrict__ clk,
                                                     int rows) {
  extern __shared__ double lds[];
  const int nband = fdiv_nband(M, rows);
  const int g = static_cast<int>(blockIdx.x) / nband;
  const FdivBand B = fdiv_band(M, rows, static_cast<int>(blockIdx.x) - g * nband);
  const int t = static_cast<int>(threadIdx.x);
  const FdivThread T = fdiv_thread(M, B, t);
  const double dt = clk->dt;
  const double wght = beta * dt;
  const int ni = M.nx1 + 2;
  const int j0 = T.j, i0 = T.i;
  const double *wb = w + static_cast<long>(g) * NHYDRO * M.ncell;
  const double *bb = bcc + static_cast<long>(g) * 3 * M.ncell;
  const long ub = static_cast<long>(g) * NHYDRO * M.ncell;
  const double *dx1 = ctab(M, g, 0), *dx2 = ctab(M, g, 1), *dx3 = ctab(M, g, 2);
  double *F1 = lds;                        // [n][t]: x1 flux of the thread's face
  double *F2 = lds + NHYDRO * WG;          // [n][t]: x2 flux
  double *CD = lds + 2 * NHYDRO * WG;      // [n][t]: x1 + x2 terms of the owned cell
  double *CP = lds + 3 * NHYDRO * WG;      // [n][t]: its low x3 flux

  for (int k = M.ks - 1; k <= M.ke + 1; ++k) {
    // the column indices are laundered each plane so that nothing derived from them is
    // hoisted out of the march and held in registers across the Riemann solves
    int j = j0, i = i0;
    asm volatile("" : "+v"(j), "+v"(i));
    double fx[NHYDRO];
    if (T.x3 && k >= M.ks) {
      face_flux<2, ORDER, RS>(M, g, k, j, i, wb, bb, o3, dt, true, fx);
      if (T.own) {
        if (k - 1 >= M.ks) {               // cell (k-1, j, i): its x3 term, then U
          const int kk = k - 1;
          const double a3 = dx1[i] * dx2[j];
          const double vol = dx1[i] * dx2[j] * dx3[kk];
          const long c = cidx(M, kk, j, i);
#pragma unroll
          for (int n = 0; n < NHYDRO; ++n) {
            double dflx = CD[n * WG + t];
            dflx += (a3 * fx[n] - a3 * CP[n * WG + t]);
            double u0 = ubase[ub + n * M.ncell + c];
            if (first_stage) u0 = 0.0 + 1.0 * u0;
            uout[ub + n * M.ncell + c] = u0 - wght * dflx / vol;
          }
        }
#pragma unroll
        for (int n = 0; n < NHYDRO; ++n) CP[n * WG + t] = fx[n];
      }
    }
    if (T.x1) {
      face_flux<0, ORDER, RS>(M, g, k, j, i, wb, bb, o1, dt, true, fx);
#pragma unroll
      for (int n = 0; n < NHYDRO; ++n) F1[n * WG + t] = fx[n];
    }
    if (T.x2) {
      face_flux<1, ORDER, RS>(M, g, k, j, i, wb, bb, o2, dt, T.emit2, fx);
#pragma unroll
      for (int n = 0; n < NHYDRO; ++n) F2[n * WG + t] = fx[n];
    }
    __syncthreads();
    if (T.own && k >= M.ks && k <= M.ke) {  // cell (k, j, i): x1 and x2 terms
      const double a1 = dx2[j] * dx3[k];
      const double a2 = dx1[i] * dx3[k];
#pragma unroll
      for (int n = 0; n < NHYDRO; ++n) {
        double d = (a1 * F1[n * WG + t + 1] - a1 * F1[n * WG + t]);
        d += (a2 * F2[n * WG + t + ni] - a2 * F2[n * WG + t]);
        CD[n * WG + t] = d;
      }
    }
    __syncthreads();
  }
}

//========================================================================================
// Kernel 2: corner EMFs, Field::ComputeCornerE 3-D branch (calculate_corner_e.cpp:129-230)
// One thread per edge position (k,j,i) in [ks..ke+1]x[js..je+1]x[is..ie+1]; the
// cell-centred E (=-v x B, SR: divided by u0) is recomputed on the fly.
template <bool SR>
__device__ __forceinline__ double cc_e(const MeshK &M, const double *wb, const double *bb,
                                       int comp, long c) {
  const long n = M.ncell;
  double vx = wb[IVX * n + c], vy = wb[IVY * n + c], vz = wb[IVZ * n + c];
  double b1 = bb[IB1 * n + c], b2 = bb[IB2 * n + c], b3 = bb[IB3 * n + c];
  if (!SR) {
    if (comp == 0) return vz * b2 - vy * b3;
    if (comp == 1) return vx * b3 - vz * b1;
    return vy * b1 - vx * b2;
  } else {
    double u0 = sqrt(1.0 + sq(vx) + sq(vy) + sq(vz));
    if (comp == 0) return (vz * b2 - vy * b3) / u0;
    if (comp == 1) return (vx * b3 - vz * b1) / u0;
    return (vy * b1 - vx * b2) / u0;
  }
}

// Inputs and outputs of the corner-EMF kernel.
struct CornerArgs {
  const double *w, *bcc;
  const double *e2x1, *e3x1, *e1x2, *e3x2, *e1x3, *e2x3;
  const double *wx1, *wx2, *wx3;
  double *e1, *e2, *e3;
};

// The three corner EMFs at edge position (k,j,i) from accessors of the face EMFs and
// upwind weights (W1F.. E3X1, each called as (k, j, i)) and of the cell-centred E
// (CC(comp, k, j, i)), handed to out(comp, value) one by one; the same expressions for
// every thread-to-position mapping.
template <class FW1, class FW2, class FW3, class FE2X1, class FE3X1, class FE1X2, class FE3X2,
          class FE1X3, class FE2X3, class FCC, class FOut>
__device__ __forceinline__ void corner_edges(int k, int j, int i, FW1 W1F, FW2 W2F, FW3 W3F,
                                             FE2X1 E2X1, FE3X1 E3X1, FE1X2 E1X2,
                                             FE3X2 E3X2, FE1X3 E1X3, FE2X3 E2X3, FCC CC,
                                             FOut out) {
  {
    double de1_l3 = (1.0 - W2F(k - 1, j, i)) * (E1X3(k, j, i) - CC(0, k - 1, j, i)) +
                    (W2F(k - 1, j, i)) * (E1X3(k, j - 1, i) - CC(0, k - 1, j - 1, i));
    double de1_r3 = (1.0 - W2F(k, j, i)) * (E1X3(k, j, i) - CC(0, k, j, i)) +
                    (W2F(k, j, i)) * (E1X3(k, j - 1, i) - CC(0, k, j - 1, i));
    double de1_l2 = (1.0 - W3F(k, j - 1, i)) * (E1X2(k, j, i) - CC(0, k, j - 1, i)) +
                    (W3F(k, j - 1, i)) * (E1X2(k - 1, j, i) - CC(0, k - 1, j - 1, i));
    double de1_r2 = (1.0 - W3F(k, j, i)) * (E1X2(k, j, i) - CC(0, k, j, i)) +
                    (W3F(k, j, i)) * (E1X2(k - 1, j, i) - CC(0, k - 1, j, i));
    out(0, 0.25 * (de1_l3 + de1_r3 + de1_l2 + de1_r2 + E1X2(k - 1, j, i) + E1X2(k, j, i) +
                E1X3(k, j - 1, i) + E1X3(k, j, i)));
  }
  {
    double de2_l3 = (1.0 - W1F(k - 1, j, i)) * (E2X3(k, j, i) - CC(1, k - 1, j, i)) +
                    (W1F(k - 1, j, i)) * (E2X3(k, j, i - 1) - CC(1, k - 1, j, i - 1));
    double de2_r3 = (1.0 - W1F(k, j, i)) * (E2X3(k, j, i) - CC(1, k, j, i)) +
                    (W1F(k, j, i)) * (E2X3(k, j, i - 1) - CC(1, k, j, i - 1));
    double de2_l1 = (1.0 - W3F(k, j, i - 1)) * (E2X1(k, j, i) - CC(1, k, j, i - 1)) +
                    (W3F(k, j, i - 1)) * (E2X1(k - 1, j, i) - CC(1, k - 1, j, i - 1));
    double de2_r1 = (1.0 - W3F(k, j, i)) * (E2X1(k, j, i) - CC(1, k, j, i)) +
                    (W3F(k, j, i)) * (E2X1(k - 1, j, i) - CC(1, k - 1, j, i));
    out(1, 0.25 * (de2_l3 + de2_r3 + de2_l1 + de2_r1 + E2X3(k, j, i - 1) + E2X3(k, j, i) +
                E2X1(k - 1, j, i) + E2X1(k, j, i)));
  }
  {
    double de3_l2 = (1.0 - W1F(k, j - 1, i)) * (E3X2(k, j, i) - CC(2, k, j - 1, i)) +
                    (W1F(k, j - 1, i)) * (E3X2(k, j, i - 1) - CC(2, k, j - 1, i - 1));
    double de3_r2 = (1.0 - W1F(k, j, i)) * (E3X2(k, j, i) - CC(2, k, j, i)) +
                    (W1F(k, j, i)) * (E3X2(k, j, i - 1) - CC(2, k, j, i - 1));
    double de3_l1 = (1.0 - W2F(k, j, i - 1)) * (E3X1(k, j, i) - CC(2, k, j, i - 1)) +
                    (W2F(k, j, i - 1)) * (E3X1(k, j - 1, i) - CC(2, k, j - 1, i - 1));
    double de3_r1 = (1.0 - W2F(k, j, i)) * (E3X1(k, j, i) - CC(2, k, j, i)) +
                    (W2F(k, j, i)) * (E3X1(k, j - 1, i) - CC(2, k, j - 1, i));
    out(2, 0.25 * (de3_l1 + de3_r1 + de3_l2 + de3_r2 + E3X2(k, j, i - 1) + E3X2(k, j, i) +
                E3X1(k, j - 1, i) + E3X1(k, j, i)));
  }
}

template <bool SR>
__device__ __forceinline__ void corner_body(const MeshK &M, const CornerArgs &ca, int lw,
                                            int nwg, TileK T) {
  int g, i, j, k;
  if (!map_kind(M, KK_CORNER, lw, nwg, T, g, i, j, k)) return;
  const double *__restrict__ w = ca.w;
  const double *__restrict__ bcc = ca.bcc;
  const double *wb = w + static_cast<long>(g) * NHYDRO * M.ncell;
  const double *bb = bcc + static_cast<long>(g) * 3 * M.ncell;
  const long o1 = static_cast<long>(g) * M.nf1, o2 = static_cast<long>(g) * M.nf2;
  const long o3 = static_cast<long>(g) * M.nf3;
  const double *__restrict__ wx1 = ca.wx1 + o1, *__restrict__ wx2 = ca.wx2 + o2;
  const double *__restrict__ wx3 = ca.wx3 + o3;
  const double *__restrict__ e2x1 = ca.e2x1 + o1, *__restrict__ e3x1 = ca.e3x1 + o1;
  const double *__restrict__ e1x2 = ca.e1x2 + o2, *__restrict__ e3x2 = ca.e3x2 + o2;
  const double *__restrict__ e1x3 = ca.e1x3 + o3, *__restrict__ e2x3 = ca.e2x3 + o3;
  double *__restrict__ e1 = ca.e1, *__restrict__ e2 = ca.e2, *__restrict__ e3 = ca.e3;
  auto W1F = [&](int kk, int jj, int ii) { return wx1[f1idx(M, kk, jj, ii)]; };
  auto W2F = [&](int kk, int jj, int ii) { return wx2[f2idx(M, kk, jj, ii)]; };
  auto W3F = [&](int kk, int jj, int ii) { return wx3[f3idx(M, kk, jj, ii)]; };
  auto E2X1 = [&](int kk, int jj, int ii) { return e2x1[f1idx(M, kk, jj, ii)]; };
  auto E3X1 = [&](int kk, int jj, int ii) { return e3x1[f1idx(M, kk, jj, ii)]; };
  auto E1X2 = [&](int kk, int jj, int ii) { return e1x2[f2idx(M, kk, jj, ii)]; };
  auto E3X2 = [&](int kk, int jj, int ii) { return e3x2[f2idx(M, kk, jj, ii)]; };
  auto E1X3 = [&](int kk, int jj, int ii) { return e1x3[f3idx(M, kk, jj, ii)]; };
  auto E2X3 = [&](int kk, int jj, int ii) { return e2x3[f3idx(M, kk, jj, ii)]; };
  auto cc = [&](int comp, int kk, int jj, int ii) {
    return cc_e<SR>(M, wb, bb, comp, cidx(M, kk, jj, ii));
  };
  auto out = [&](int comp, double v) {
    if (comp == 0) e1[static_cast<long>(g) * M.ne1 + e1idx(M, k, j, i)] = v;
    else if (comp == 1) e2[static_cast<long>(g) * M.ne2 + e2idx(M, k, j, i)] = v;
    else e3[static_cast<long>(g) * M.ne3 + e3idx(M, k, j, i)] = v;
  };
  corner_edges(k, j, i, W1F, W2F, W3F, E2X1, E3X1, E1X2, E3X2, E1X3, E2X3, cc, out);
}

// Z-marching corner EMFs (tile option with tx < 0, tz = planes per segment): a
// work-group owns a TI x TJ tile of (i, j) edge positions (TI = the box width up to 256,
// TJ = 256 / (TI+1) - 1 rows but at least one, so that for TI < 255 its 256 threads
// stage the (TI+1) x (TJ+1) cells in one pass) and walks up a segment of planes.  The cell-centred E of the
// (TI+1) x (TJ+1) cells of a plane that the tile's edges touch is evaluated once into
// LDS (planes k-1 and k, double-buffered), and every thread carries the four face values
// of plane k-1 that its edges need from the previous step, so each array element is
// read from memory about once instead of once per neighbouring work-group.
struct MarchGeom {
  int ti, tj, nti, ntj, ks, nks;
};

__host__ __device__ __forceinline__ MarchGeom march_geom(const MeshK &M, int ks) {
  int lo[3], n[3];
  kind_box(M, KK_CORNER, lo, n);
  MarchGeom G;
  G.ti = (n[0] < WG) ? n[0] : WG;
  G.tj = WG / (G.ti + 1) - 1;   // stage (ti+1) x (tj+1) cells in one pass
  if (G.tj < 1) G.tj = 1;
  if (G.tj > n[1]) G.tj = n[1];
  G.nti = (n[0] + G.ti - 1) / G.ti;
  G.ntj = (n[1] + G.tj - 1) / G.tj;
  G.ks = ks;
  G.nks = (n[2] + ks - 1) / ks;
  return G;
}

__host__ __device__ __forceinline__ int march_lds_doubles(const MarchGeom &G) {
  return 2 * 3 * (G.ti + 1) * (G.tj + 1);
}

template <bool SR>
__device__ __forceinline__ void corner_march(const MeshK &M, const CornerArgs &ca, int lw,
                                             int nwg, int ks) {
  extern __shared__ double lds_cc[];
  int lo[3], n[3];
  kind_box(M, KK_CORNER, lo, n);
  const MarchGeom G = march_geom(M, ks);
  const int per = nwg / M.nblocks;               // = G.nti * G.ntj * G.nks
  const int g = lw / per;
  int r = lw - g * per;
  const int seg = r % G.nks;
  r /= G.nks;
  const int tjx = r % G.ntj, tix = r / G.ntj;
  const int i0 = lo[0] + tix * G.ti, j0 = lo[1] + tjx * G.tj;
  const int k0 = lo[2] + seg * G.ks;
  const int k1 = (k0 + G.ks < lo[2] + n[2]) ? k0 + G.ks : lo[2] + n[2];
  const int iend = lo[0] + n[0] - 1, jend = lo[1] + n[1] - 1;   // last edge positions
  const int t = threadIdx.x;
  const int i = i0 + t % G.ti, j = j0 + t / G.ti;
  const bool act = (t / G.ti) < G.tj && i <= iend && j <= jend;
  const int sw = G.ti + 1, ncs = sw * (G.tj + 1);   // staged cells i0-1.., j0-1.. per plane
  const double *wb = ca.w + static_cast<long>(g) * NHYDRO * M.ncell;
  const double *bb = ca.bcc + static_cast<long>(g) * 3 * M.ncell;
  const long o1 = static_cast<long>(g) * M.nf1, o2 = static_cast<long>(g) * M.nf2;
  const long o3 = static_cast<long>(g) * M.nf3;
  auto stage = [&](double *dst, int kk) {
    for (int c = t; c < ncs; c += WG) {
      const int ci = i0 - 1 + c % sw, cj = j0 - 1 + c / sw;
      if (ci > iend || cj > jend) continue;      // beyond the box: never read
      const long q = cidx(M, kk, cj, ci);
      dst[c] = cc_e<SR>(M, wb, bb, 0, q);
      dst[ncs + c] = cc_e<SR>(M, wb, bb, 1, q);
      dst[2 * ncs + c] = cc_e<SR>(M, wb, bb, 2, q);
    }
  };
  // face values of plane k-1 at (j, i): W1F, W2F, E1X2, E2X1
  double w1p = 0.0, w2p = 0.0, e12p = 0.0, e21p = 0.0;
  if (act) {
    w1p = ca.wx1[o1 + f1idx(M, k0 - 1, j, i)];
    w2p = ca.wx2[o2 + f2idx(M, k0 - 1, j, i)];
    e12p = ca.e1x2[o2 + f2idx(M, k0 - 1, j, i)];
    e21p = ca.e2x1[o1 + f1idx(M, k0 - 1, j, i)];
  }
  stage(lds_cc, k0 - 1);
  int cur = 1;
  for (int kq = k0; kq < k1; ++kq) {
    const int k = __builtin_amdgcn_readfirstlane(kq);   // the same for the whole work-group
    double *C = lds_cc + cur * 3 * ncs;
    const double *P = lds_cc + (cur ^ 1) * 3 * ncs;
    stage(C, k);
    __syncthreads();
    if (act) {
      const double w1c = ca.wx1[o1 + f1idx(M, k, j, i)];
      const double w2c = ca.wx2[o2 + f2idx(M, k, j, i)];
      const double e12c = ca.e1x2[o2 + f2idx(M, k, j, i)];
      const double e21c = ca.e2x1[o1 + f1idx(M, k, j, i)];
      auto carried = [&](const double *a, long o, int comp, double prev, double curv) {
        return [=, &M](int kk, int jj, int ii) {
          if (kk < k) return prev;
          if (jj == j && ii == i) return curv;
          const long x = (comp == 0) ? f1idx(M, kk, jj, ii)
                                     : ((comp == 1) ? f2idx(M, kk, jj, ii) : f3idx(M, kk, jj, ii));
          return a[o + x];
        };
      };
      auto plain = [&](const double *a, long o, int comp) {
        return [=, &M](int kk, int jj, int ii) {
          const long x = (comp == 0) ? f1idx(M, kk, jj, ii)
                                     : ((comp == 1) ? f2idx(M, kk, jj, ii) : f3idx(M, kk, jj, ii));
          return a[o + x];
        };
      };
      auto cc = [&](int comp, int kk, int jj, int ii) {
        const double *B = (kk < k) ? P : C;
        return B[comp * ncs + (jj - j0 + 1) * sw + (ii - i0 + 1)];
      };
      auto out = [&](int comp, double v) {
        if (comp == 0) ca.e1[static_cast<long>(g) * M.ne1 + e1idx(M, k, j, i)] = v;
        else if (comp == 1) ca.e2[static_cast<long>(g) * M.ne2 + e2idx(M, k, j, i)] = v;
        else ca.e3[static_cast<long>(g) * M.ne3 + e3idx(M, k, j, i)] = v;
      };
      corner_edges(k, j, i, carried(ca.wx1, o1, 0, w1p, w1c), carried(ca.wx2, o2, 1, w2p, w2c),
                   plain(ca.wx3, o3, 2), carried(ca.e2x1, o1, 0, e21p, e21c),
                   plain(ca.e3x1, o1, 0), carried(ca.e1x2, o2, 1, e12p, e12c),
                   plain(ca.e3x2, o2, 1), plain(ca.e1x3, o3, 2), plain(ca.e2x3, o3, 2), cc, out);
      w1p = w1c;
      w2p = w2c;
      e12p = e12c;
      e21p = e21c;
    }
    __syncthreads();
    cur ^= 1;
  }
}

template <bool SR>
__global__ void __launch_bounds__(WG) k_corner(MeshK M, CornerArgs ca, TileK T) {
  corner_body<SR>(M, ca, xcd_logical_wg(M.xcd_remap & XCD_CORNER), gridDim.x, T);
}

template <bool SR>
__global__ void __launch_bounds__(WG) k_corner_march(MeshK M, CornerArgs ca, TileK T) {
  corner_march<SR>(M, ca, xcd_logical_wg(M.xcd_remap & XCD_CORNER), gridDim.x, T.tz);
}

//========================================================================================
// Same-level EMF flux correction on a uniform mesh, evaluated where CT reads the EMF.
// A surface EMF is own + face neighbour(s) + edge neighbour, summed in the reference's
// neighbour order (x1 faces, x2 faces, x3 faces, then the edge) and scaled by 1/2
// (face) or 1/nedge_fine (edge: 1/4 with every neighbour present; outflow faces below)
// (flux_correction_fc.cpp:51-…, 689-…,
// 1355-1548; bvals_base.cpp:327-700).  Raw EMFs are never modified, so every thread
// reads the un-averaged neighbour values exactly like the packed send buffers.
// Scale of an edge EMF shared by cnt blocks (1/nedge_fine, bvals_fc.cpp:1129-1193 and
// flux_correction_fc.cpp:1478-1545; a block alone on its edge keeps its value).
__device__ __forceinline__ double edge_scale(double v, int cnt) {
  if (cnt == 1) return v;
  return (cnt == 2) ? v * 0.5 : ((cnt == 4) ? v * 0.25 : v * (1.0 / 3.0));
}

// With outflow faces a neighbour may be missing (nbr = -1): it adds nothing, a face with
// no neighbour is not averaged (flux_correction_fc.cpp:1363-1366) and an edge is scaled
// by the number of blocks around it.
__device__ __forceinline__ double emf3(const MeshK &M, const double *e, int g, int k,
                                       int j, int i) {
  double v = e[static_cast<long>(g) * M.ne3 + e3idx(M, k, j, i)];
  if (M.smr) return v;
  const int o1 = (i == M.is) ? -1 : ((i == M.ie + 1) ? 1 : 0);
  const int o2 = (j == M.js) ? -1 : ((j == M.je + 1) ? 1 : 0);
  if (!o1 && !o2) return v;
  if (!M.outflow) {
    if (o1) v += e[static_cast<long>(nbr(M, g, o1, 0, 0)) * M.ne3 + e3idx(M, k, j, i - o1 * M.nx1)];
    if (o2) v += e[static_cast<long>(nbr(M, g, 0, o2, 0)) * M.ne3 + e3idx(M, k, j - o2 * M.nx2, i)];
    if (o1 && o2) {
      v += e[static_cast<long>(nbr(M, g, o1, o2, 0)) * M.ne3 +
             e3idx(M, k, j - o2 * M.nx2, i - o1 * M.nx1)];
      return v * 0.25;
    }
    return v * 0.5;
  }
  const int na = o1 ? nbr(M, g, o1, 0, 0) : -1, nb = o2 ? nbr(M, g, 0, o2, 0) : -1;
  if (na >= 0) v += e[static_cast<long>(na) * M.ne3 + e3idx(M, k, j, i - o1 * M.nx1)];
  if (nb >= 0) v += e[static_cast<long>(nb) * M.ne3 + e3idx(M, k, j - o2 * M.nx2, i)];
  if (o1 && o2) {
    const int ne = nbr(M, g, o1, o2, 0);
    if (ne >= 0)
      v += e[static_cast<long>(ne) * M.ne3 + e3idx(M, k, j - o2 * M.nx2, i - o1 * M.nx1)];
    return edge_scale(v, 1 + (na >= 0) + (nb >= 0) + (ne >= 0));
  }
  return (na >= 0 || nb >= 0) ? v * 0.5 : v;
}
__device__ __forceinline__ double emf2(const MeshK &M, const double *e, int g, int k,
                                       int j, int i) {
  double v = e[static_cast<long>(g) * M.ne2 + e2idx(M, k, j, i)];
  if (M.smr) return v;
  const int o1 = (i == M.is) ? -1 : ((i == M.ie + 1) ? 1 : 0);
  const int o3 = (k == M.ks) ? -1 : ((k == M.ke + 1) ? 1 : 0);
  if (!o1 && !o3) return v;
  if (!M.outflow) {
    if (o1) v += e[static_cast<long>(nbr(M, g, o1, 0, 0)) * M.ne2 + e2idx(M, k, j, i - o1 * M.nx1)];
    if (o3) v += e[static_cast<long>(nbr(M, g, 0, 0, o3)) * M.ne2 + e2idx(M, k - o3 * M.nx3, j, i)];
    if (o1 && o3) {
      v += e[static_cast<long>(nbr(M, g, o1, 0, o3)) * M.ne2 +
             e2idx(M, k - o3 * M.nx3, j, i - o1 * M.nx1)];
      return v * 0.25;
    }
    return v * 0.5;
  }
  const int na = o1 ? nbr(M, g, o1, 0, 0) : -1, nb = o3 ? nbr(M, g, 0, 0, o3) : -1;
  if (na >= 0) v += e[static_cast<long>(na) * M.ne2 + e2idx(M, k, j, i - o1 * M.nx1)];
  if (nb >= 0) v += e[static_cast<long>(nb) * M.ne2 + e2idx(M, k - o3 * M.nx3, j, i)];
  if (o1 && o3) {
    const int ne = nbr(M, g, o1, 0, o3);
    if (ne >= 0)
      v += e[static_cast<long>(ne) * M.ne2 + e2idx(M, k - o3 * M.nx3, j, i - o1 * M.nx1)];
    return edge_scale(v, 1 + (na >= 0) + (nb >= 0) + (ne >= 0));
  }
  return (na >= 0 || nb >= 0) ? v * 0.5 : v;
}
__device__ __forceinline__ double emf1(const MeshK &M, const double *e, int g, int k,
                                       int j, int i) {
  double v = e[static_cast<long>(g) * M.ne1 + e1idx(M, k, j, i)];
  if (M.smr) return v;
  const int o2 = (j == M.js) ? -1 : ((j == M.je + 1) ? 1 : 0);
  const int o3 = (k == M.ks) ? -1 : ((k == M.ke + 1) ? 1 : 0);
  if (!o2 && !o3) return v;
  if (!M.outflow) {
    if (o2) v += e[static_cast<long>(nbr(M, g, 0, o2, 0)) * M.ne1 + e1idx(M, k, j - o2 * M.nx2, i)];
    if (o3) v += e[static_cast<long>(nbr(M, g, 0, 0, o3)) * M.ne1 + e1idx(M, k - o3 * M.nx3, j, i)];
    if (o2 && o3) {
      v += e[static_cast<long>(nbr(M, g, 0, o2, o3)) * M.ne1 +
             e1idx(M, k - o3 * M.nx3, j - o2 * M.nx2, i)];
      return v * 0.25;
    }
    return v * 0.5;
  }
  const int na = o2 ? nbr(M, g, 0, o2, 0) : -1, nb = o3 ? nbr(M, g, 0, 0, o3) : -1;
  if (na >= 0) v += e[static_cast<long>(na) * M.ne1 + e1idx(M, k, j - o2 * M.nx2, i)];
  if (nb >= 0) v += e[static_cast<long>(nb) * M.ne1 + e1idx(M, k - o3 * M.nx3, j, i)];
  if (o2 && o3) {
    const int ne = nbr(M, g, 0, o2, o3);
    if (ne >= 0)
      v += e[static_cast<long>(ne) * M.ne1 + e1idx(M, k - o3 * M.nx3, j - o2 * M.nx2, i)];
    return edge_scale(v, 1 + (na >= 0) + (nb >= 0) + (ne >= 0));
  }
  return (na >= 0 || nb >= 0) ? v * 0.5 : v;
}

//========================================================================================
// Kernel 3a: conserved update, MeshBlock::WeightedAve + Hydro::AddFluxDivergence
// (weighted_ave.cpp:35-…, add_flux_divergence.cpp:39-96) for active cells; the ghost
// copies are written afterwards by the exchange kernel (k_ghost).
// first_stage: out = (0 + 1*base) - wght*div F   (u1 zeroed, delta = 1, then swapped)
// otherwise:   out = base - wght*div F           (in place allowed: out == base)
__device__ __forceinline__ void update_cell_u(const MeshK &M, int g, int k, int j, int i,
                                              const double *base, double *out,
                                              const double *__restrict__ f1,
                                              const double *__restrict__ f2,
                                              const double *__restrict__ f3, double wght,
                                              int first_stage) {
  const double *dx1 = ctab(M, g, 0), *dx2 = ctab(M, g, 1), *dx3 = ctab(M, g, 2);
  const double a1 = dx2[j] * dx3[k];
  const double a2 = dx1[i] * dx3[k];
  const double a3 = dx1[i] * dx2[j];
  const double vol = dx1[i] * dx2[j] * dx3[k];
  const long c = cidx(M, k, j, i);
  const long b0 = static_cast<long>(g) * NHYDRO * M.ncell;
  const double *F1 = f1 + static_cast<long>(g) * NHYDRO * M.nf1;
  const double *F2 = f2 + static_cast<long>(g) * NHYDRO * M.nf2;
  const double *F3 = f3 + static_cast<long>(g) * NHYDRO * M.nf3;
  const long x1 = f1idx(M, k, j, i), x2 = f2idx(M, k, j, i), x3 = f3idx(M, k, j, i);
  const long sx2 = M.nc1, sx3 = static_cast<long>(M.nc2) * M.nc1;
#pragma unroll
  for (int n = 0; n < NHYDRO; ++n) {
    double dflx = (a1 * F1[n * M.nf1 + x1 + 1] - a1 * F1[n * M.nf1 + x1]);
    dflx += (a2 * F2[n * M.nf2 + x2 + sx2] - a2 * F2[n * M.nf2 + x2]);
    dflx += (a3 * F3[n * M.nf3 + x3 + sx3] - a3 * F3[n * M.nf3 + x3]);
    double u0 = base[b0 + n * M.ncell + c];
    if (first_stage) u0 = 0.0 + 1.0 * u0;
    out[b0 + n * M.ncell + c] = u0 - wght * dflx / vol;
  }
}

// Face-field update, WeightedAve(FaceField) + Field::CT (ct.cpp:31-116) with averaged
// surface EMFs.
template <int comp>
__device__ __forceinline__ void update_face_b(const MeshK &M, int g, int k, int j, int i,
                                              const double *bin, double *outp,
                                              const double *__restrict__ e1,
                                              const double *__restrict__ e2,
                                              const double *__restrict__ e3, double wght,
                                              int first_stage) {
  const double *dx1 = ctab(M, g, 0), *dx2 = ctab(M, g, 1), *dx3 = ctab(M, g, 2);
  double v;
  long nf;
  if (comp == 0) {
    nf = M.nf1;
    v = bin[static_cast<long>(g) * nf + f1idx(M, k, j, i)];
    if (first_stage) v = 0.0 + 1.0 * v;
    const double area = dx2[j] * dx3[k];
    const double l3 = dx3[k], l2 = dx2[j];
    v -= (wght / area) * (l3 * emf3(M, e3, g, k, j + 1, i) - l3 * emf3(M, e3, g, k, j, i));
    v += (wght / area) * (l2 * emf2(M, e2, g, k + 1, j, i) - l2 * emf2(M, e2, g, k, j, i));
  } else if (comp == 1) {
    nf = M.nf2;
    v = bin[static_cast<long>(g) * nf + f2idx(M, k, j, i)];
    if (first_stage) v = 0.0 + 1.0 * v;
    const double area = dx1[i] * dx3[k];
    const double l3 = dx3[k], l1 = dx1[i];
    v += (wght / area) * (l3 * emf3(M, e3, g, k, j, i + 1) - l3 * emf3(M, e3, g, k, j, i));
    v -= (wght / area) * (l1 * emf1(M, e1, g, k + 1, j, i) - l1 * emf1(M, e1, g, k, j, i));
  } else {
    nf = M.nf3;
    v = bin[static_cast<long>(g) * nf + f3idx(M, k, j, i)];
    if (first_stage) v = 0.0 + 1.0 * v;
    const double area = dx1[i] * dx2[j];
    const double l2 = dx2[j], l1 = dx1[i];
    v -= (wght / area) * (l2 * emf2(M, e2, g, k, j, i + 1) - l2 * emf2(M, e2, g, k, j, i));
    v += (wght / area) * (l1 * emf1(M, e1, g, k, j + 1, i) - l1 * emf1(M, e1, g, k, j, i));
  }
  long fi;
  if (comp == 0) fi = f1idx(M, k, j, i);
  else if (comp == 1) fi = f2idx(M, k, j, i);
  else fi = f3idx(M, k, j, i);
  outp[static_cast<long>(g) * nf + fi] = v;
}

// One thread per (k,j,i) in [ks,ke+1]x[js,je+1]x[is,ie+1]: the active cell's conserved
// update and the three faces on its low sides that are active faces.
struct UpdateArgs {
  const double *ubase;
  double *uout;
  const double *f1, *f2, *f3;
  const double *bin[3];
  double *bout[3];
  const double *e1, *e2, *e3;
};

__device__ __forceinline__ void update_body(const MeshK &M, const UpdateArgs &a, double beta,
                                            int first_stage, double dt, int lw, int nwg,
                                            TileK T, bool do_u = true) {
  int g, i, j, k;
  if (!map_kind(M, KK_UPDATE, lw, nwg, T, g, i, j, k)) return;
  const double wght = beta * dt;
  const bool ci = i <= M.ie, cj = j <= M.je, ck = k <= M.ke;
  if (do_u && ci && cj && ck)
    update_cell_u(M, g, k, j, i, a.ubase, a.uout, a.f1, a.f2, a.f3, wght, first_stage);
  if (cj && ck)
    update_face_b<0>(M, g, k, j, i, a.bin[0], a.bout[0], a.e1, a.e2, a.e3, wght, first_stage);
  if (ci && ck)
    update_face_b<1>(M, g, k, j, i, a.bin[1], a.bout[1], a.e1, a.e2, a.e3, wght, first_stage);
  if (ci && cj)
    update_face_b<2>(M, g, k, j, i, a.bin[2], a.bout[2], a.e1, a.e2, a.e3, wght, first_stage);
}

__global__ void __launch_bounds__(WG) k_update(MeshK M, UpdateArgs a, double beta,
                                                int first_stage,
                                                const Clock *__restrict__ clk, TileK T) {
  update_body(M, a, beta, first_stage, clk->dt, xcd_logical_wg(M.xcd_remap & XCD_UPDATE),
              gridDim.x, T);
}

// Field::CT alone (after k_flux_div has updated U).
__global__ void __launch_bounds__(WG) k_update_b(MeshK M, UpdateArgs a, double beta,
                                                  int first_stage,
                                                  const Clock *__restrict__ clk, TileK T) {
  update_body(M, a, beta, first_stage, clk->dt, xcd_logical_wg(M.xcd_remap & XCD_UPDATE),
              gridDim.x, T,
              false);
}

//========================================================================================
// Kernel 3b: same-level ghost exchange on a periodic uniform mesh as a gather
// (cc/bvals_cc.cpp:287-427, fc/bvals_fc.cpp:344-688): every ghost position of the
// cell-centred register (5 components) and of the three face-field components takes the
// value of its source, the active cell or face of a neighbour MeshBlock that the
// reference's send buffers carry there.  Every array of a block has NGHOST ghost layers
// on each side of each direction, active positions [ng, n-ng-1] (for a face component
// along its normal that is [s, e+1], the shared faces staying with their owner), and a
// ghost position idx < ng copies idx + nx of the low neighbour, idx > n-ng-1 copies
// idx - nx of the high neighbour.  One thread per ghost position (all 5 components of a
// cell-centred one); the ghost positions are numbered in memory order, so consecutive
// threads write consecutive addresses: the ng low ghost planes, per active plane its ng
// low ghost rows, the 2*ng ends of each active row and its ng high ghost rows, then the
// ng high planes.
struct GhostArgs {
  double *u;
  double *b[3];
};

__host__ __device__ __forceinline__ long ghost_count(int n1, int n2, int n3, int ng) {
  return static_cast<long>(n1) * n2 * n3 -
         static_cast<long>(n1 - 2 * ng) * (n2 - 2 * ng) * (n3 - 2 * ng);
}

// ghost number t of an (n3, n2, n1) array -> (k, j, i)
__host__ __device__ __forceinline__ void ghost_pos(int t, int n1, int n2, int n3, int ng,
                                                   int &i, int &j, int &k) {
  const int plane = n1 * n2;
  const int m2 = n2 - 2 * ng, m3 = n3 - 2 * ng;
  const int rows = ng * n1, ends = m2 * 2 * ng;
  const int per = 2 * rows + ends;               // ghost positions of one active plane
  if (t < ng * plane || t >= ng * plane + m3 * per) {   // a whole ghost plane
    if (t >= ng * plane) t -= m3 * per - m3 * plane;    // high planes follow the active
    const int q = t / plane, r = t - q * plane;
    k = q;
    j = r / n1;
    i = r - j * n1;
    return;
  }
  t -= ng * plane;
  const int q = t / per;
  int r = t - q * per;
  k = ng + q;
  if (r < rows) {
    j = r / n1;
    i = r - j * n1;
  } else if (r < rows + ends) {
    r -= rows;
    const int row = r / (2 * ng), x = r - row * 2 * ng;
    j = ng + row;
    i = (x < ng) ? x : n1 - 2 * ng + x;
  } else {
    r -= rows + ends;
    const int row = r / n1;
    j = n2 - ng + row;
    i = r - row * n1;
  }
}

__host__ __device__ __forceinline__ int ghost_off(int idx, int n, int ng) {
  return (idx < ng) ? -1 : ((idx > n - ng - 1) ? 1 : 0);
}

// Work items of one block: the cell ghost positions (each copies the 5 components), then
// the x1, x2, x3 face ghost positions.
__host__ __device__ __forceinline__ long ghost_items(const MeshK &M) {
  return ghost_count(M.nc1, M.nc2, M.nc3, M.ng) + ghost_count(M.nc1 + 1, M.nc2, M.nc3, M.ng) +
         ghost_count(M.nc1, M.nc2 + 1, M.nc3, M.ng) + ghost_count(M.nc1, M.nc2, M.nc3 + 1, M.ng);
}

__device__ __forceinline__ void ghost_body(const MeshK &M, const GhostArgs &a, int lw,
                                           int nwg) {
  const int nch = nwg / M.nblocks;
  const int g = lw / nch;
  int t = (lw - g * nch) * WG + threadIdx.x;
  const int ng = M.ng;
  const int cc = static_cast<int>(ghost_count(M.nc1, M.nc2, M.nc3, ng));
  int n1 = M.nc1, n2 = M.nc2, n3 = M.nc3, ncomp;
  double *arr;
  long stride, cstride;                          // block stride, component stride
  if (t < cc) {
    arr = a.u; ncomp = NHYDRO; stride = NHYDRO * M.ncell; cstride = M.ncell;
  } else {
    t -= cc;
    ncomp = 1; cstride = 0;
    const int c1 = static_cast<int>(ghost_count(M.nc1 + 1, M.nc2, M.nc3, ng));
    const int c2 = static_cast<int>(ghost_count(M.nc1, M.nc2 + 1, M.nc3, ng));
    const int c3 = static_cast<int>(ghost_count(M.nc1, M.nc2, M.nc3 + 1, ng));
    if (t < c1) {
      n1 += 1; arr = a.b[0]; stride = M.nf1;
    } else if (t < c1 + c2) {
      t -= c1; n2 += 1; arr = a.b[1]; stride = M.nf2;
    } else if (t < c1 + c2 + c3) {
      t -= c1 + c2; n3 += 1; arr = a.b[2]; stride = M.nf3;
    } else {
      return;
    }
  }
  int i, j, k;
  ghost_pos(t, n1, n2, n3, ng, i, j, k);
  int p1 = ghost_off(i, n1, ng), p2 = ghost_off(j, n2, ng), p3 = ghost_off(k, n3, ng);
  int si = i - p1 * M.nx1, sj = j - p2 * M.nx2, sk = k - p3 * M.nx3;
  if (M.outflow) {
    // No neighbour across an outflow face: the zero-gradient copy of the block's own
    // boundary layer (cc/outflow_cc.cpp, fc/outflow_fc.cpp), the last active cell or, for
    // a face component along its normal, the boundary face (n - ng - 1 = ie + 1).
    // Directions that keep a neighbour still take its values, so an edge or corner ghost
    // gets what the reference's exchange followed by the boundary copy leaves there.
    if (p1 && nbr(M, g, p1, 0, 0) < 0) { si = (p1 < 0) ? ng : n1 - ng - 1; p1 = 0; }
    if (p2 && nbr(M, g, 0, p2, 0) < 0) { sj = (p2 < 0) ? ng : n2 - ng - 1; p2 = 0; }
    if (p3 && nbr(M, g, 0, 0, p3) < 0) { sk = (p3 < 0) ? ng : n3 - ng - 1; p3 = 0; }
  }
  const int src = nbr(M, g, p1, p2, p3);
  const long own = (static_cast<long>(k) * n2 + j) * n1 + i;
  const long from = (static_cast<long>(sk) * n2 + sj) * n1 + si;
  double *dst = arr + static_cast<long>(g) * stride + own;
  const double *s0 = arr + static_cast<long>(src) * stride + from;
  if (ncomp == NHYDRO) {
    double v[NHYDRO];
#pragma unroll
    for (int q = 0; q < NHYDRO; ++q) v[q] = s0[q * cstride];
#pragma unroll
    for (int q = 0; q < NHYDRO; ++q) dst[q * cstride] = v[q];
  } else {
    dst[0] = s0[0];
  }
}

__global__ void __launch_bounds__(WG) k_ghost(MeshK M, GhostArgs a) {
  ghost_body(M, a, xcd_logical_wg(M.xcd_remap & XCD_GHOST), gridDim.x);
}

//========================================================================================
// Kernel 3c: BoundaryValues::ApplyPhysicalBoundaries (bvals.cpp:453-…) for outflow
// faces, after ConservedToPrimitive.  Every cell ghost position across an outflow face
// takes the primitives of the cell the boundary functions copy from (the boundary layer
// in each outflow direction, cc/outflow_cc.cpp, applied to the primitives as
// time_integrator.cpp:2291 and mesh.cpp:1868 select) and rebuilds its conserved
// variables with EquationOfState::PrimitiveToConserved (adiabatic_mhd.cpp:99-136) from
// them and its own cell-centred field, which the ConservedToPrimitive kernel has already
// computed from the copied face ghosts (bvals.cpp:517-523); SR meshes use the SR
// PrimitiveToConserved (adiabatic_mhd_sr.cpp:276-332).  The source cell lies on the
// active side of every outflow face, so no thread reads what another one writes.
struct PhysArgs {
  double *u, *prim;
  const double *bcc;
};

template <bool SR>
__device__ __forceinline__ void phys_bnd_body(const MeshK &M, const PhysArgs &a, int lw,
                                              int nwg) {
  const int nch = nwg / M.nblocks;
  const int g = lw / nch;
  const int t = (lw - g * nch) * WG + threadIdx.x;
  const int n1 = M.nc1, n2 = M.nc2, n3 = M.nc3, ng = M.ng;
  if (t >= ghost_count(n1, n2, n3, ng)) return;
  int i, j, k;
  ghost_pos(t, n1, n2, n3, ng, i, j, k);
  const int p1 = ghost_off(i, n1, ng), p2 = ghost_off(j, n2, ng), p3 = ghost_off(k, n3, ng);
  int si = i, sj = j, sk = k;
  bool phys = false;
  if (p1 && nbr(M, g, p1, 0, 0) < 0) { si = (p1 < 0) ? M.is : M.ie; phys = true; }
  if (p2 && nbr(M, g, 0, p2, 0) < 0) { sj = (p2 < 0) ? M.js : M.je; phys = true; }
  if (p3 && nbr(M, g, 0, 0, p3) < 0) { sk = (p3 < 0) ? M.ks : M.ke; phys = true; }
  if (!phys) return;
  const long nc = M.ncell;
  const long b0 = static_cast<long>(g) * NHYDRO * nc;
  const long c = cidx(M, k, j, i), src = cidx(M, sk, sj, si);
  double w[NHYDRO];
#pragma unroll
  for (int q = 0; q < NHYDRO; ++q) w[q] = a.prim[b0 + q * nc + src];
#pragma unroll
  for (int q = 0; q < NHYDRO; ++q) a.prim[b0 + q * nc + c] = w[q];
  const double *bc = a.bcc + static_cast<long>(g) * 3 * nc + c;
  const double bcc1 = bc[0], bcc2 = bc[nc], bcc3 = bc[2 * nc];
  double *u = a.u + b0 + c;
  if (!SR) {
    u[IDN * nc] = w[IDN];
    u[IM1 * nc] = w[IVX] * w[IDN];
    u[IM2 * nc] = w[IVY] * w[IDN];
    u[IM3 * nc] = w[IVZ] * w[IDN];
    u[IEN * nc] = w[IPR] * M.eos.igm1 +
                  0.5 * (w[IDN] * (w[IVX] * w[IVX] + w[IVY] * w[IVY] + w[IVZ] * w[IVZ]) +
                         (bcc1 * bcc1 + bcc2 * bcc2 + bcc3 * bcc3));
  } else {
    const double red = M.eos.gamma / (M.eos.gamma - 1.0);
    const double u1 = w[IVX], u2 = w[IVY], u3 = w[IVZ];
    const double u0 = sqrt(1.0 + sq(u1) + sq(u2) + sq(u3));
    const double b0 = bcc1 * u1 + bcc2 * u2 + bcc3 * u3;
    const double b1 = (bcc1 + b0 * u1) / u0;
    const double b2 = (bcc2 + b0 * u2) / u0;
    const double b3 = (bcc3 + b0 * u3) / u0;
    const double b_sq = -sq(b0) + sq(b1) + sq(b2) + sq(b3);
    const double wtot_u02 = (w[IDN] + red * w[IPR] + b_sq) * u0 * u0;
    u[IDN * nc] = w[IDN] * u0;
    u[IEN * nc] = wtot_u02 - b0 * b0 - (w[IPR] + 0.5 * b_sq);
    u[IM1 * nc] = wtot_u02 * u1 / u0 - b0 * b1;
    u[IM2 * nc] = wtot_u02 * u2 / u0 - b0 * b2;
    u[IM3 * nc] = wtot_u02 * u3 / u0 - b0 * b3;
  }
}

template <bool SR>
__global__ void __launch_bounds__(WG) k_phys_bnd(MeshK M, PhysArgs a) {
  phys_bnd_body<SR>(M, a, blockIdx.x, gridDim.x);
}

//========================================================================================
// Kernel 4: ConservedToPrimitive over every cell (ghosts included,
// time_integrator.cpp:2206-2212) with Field::CalculateCellCenteredField
// (field.cpp:112-180, uniform branch), writing prim (the output register) and bcc.
// With dt_partial != nullptr the active cells also evaluate Hydro::NewBlockTimeStep
// (new_blockdt.cpp:45-230) and each work-group writes its minimum.
// Inputs and outputs of the ConservedToPrimitive kernel.
struct C2PArgs {
  double *u, *prim, *bcc;
  const double *b1, *b2, *b3;
  double *dt_partial;         // null: no NewBlockTimeStep
  const double *lseed;        // learned pressure seeds (C2P_LEARNED)
};

// ConservedToPrimitive of one cell (k, j, i) of block g: bcc from the faces, the
// primitives into prim (the conserved variables rewritten where a floor applies) and,
// with with_dt, the cell's Hydro::NewBlockTimeStep minimum into *my_dt.
template <bool SR, bool LEARNED>
__device__ __forceinline__ void c2p_cell(const MeshK &M, const C2PArgs &a, int g, int k, int j,
                                         int i, bool with_dt, double *my_dt, C2PStats *st) {
  double *__restrict__ u = a.u;
  double *__restrict__ prim = a.prim;
  double *__restrict__ bcc = a.bcc;
  const double *__restrict__ b1 = a.b1;
  const double *__restrict__ b2 = a.b2;
  const double *__restrict__ b3 = a.b3;
  const double *__restrict__ lseed = a.lseed;
  const long n = M.ncell;
  const long c = cidx(M, k, j, i);
  const double *B1 = b1 + static_cast<long>(g) * M.nf1;
  const double *B2 = b2 + static_cast<long>(g) * M.nf2;
  const double *B3 = b3 + static_cast<long>(g) * M.nf3;
  const double b1l = B1[f1idx(M, k, j, i)], b1r = B1[f1idx(M, k, j, i + 1)];
  const double b2l = B2[f2idx(M, k, j, i)], b2r = B2[f2idx(M, k, j + 1, i)];
  const double b3l = B3[f3idx(M, k, j, i)], b3r = B3[f3idx(M, k + 1, j, i)];
  double bc[3];
  bc[IB1] = 0.5 * b1l + 0.5 * b1r;
  bc[IB2] = 0.5 * b2l + 0.5 * b2r;
  bc[IB3] = 0.5 * b3l + 0.5 * b3r;
  double *bb = bcc + static_cast<long>(g) * 3 * n;
  bb[IB1 * n + c] = bc[IB1];
  bb[IB2 * n + c] = bc[IB2];
  bb[IB3 * n + c] = bc[IB3];
  double *ub = u + static_cast<long>(g) * NHYDRO * n;
  double *pb = prim + static_cast<long>(g) * NHYDRO * n;
  double w[NHYDRO];
  if (!SR) {
    // adiabatic_mhd.cpp:41-90
    const double gm1 = M.eos.gamma - 1.0;
    double u_d = ub[IDN * n + c];
    const double u_m1 = ub[IM1 * n + c], u_m2 = ub[IM2 * n + c], u_m3 = ub[IM3 * n + c];
    double u_e = ub[IEN * n + c];
    const double ud_in = u_d, ue_in = u_e;
    u_d = (u_d > M.eos.dfloor) ? u_d : M.eos.dfloor;
    w[IDN] = u_d;
    double di = 1.0 / u_d;
    w[IVX] = u_m1 * di;
    w[IVY] = u_m2 * di;
    w[IVZ] = u_m3 * di;
    double pb_ = 0.5 * (sq(bc[IB1]) + sq(bc[IB2]) + sq(bc[IB3]));
    double e_k = 0.5 * di * (sq(u_m1) + sq(u_m2) + sq(u_m3));
    double w_p = gm1 * (u_e - e_k - pb_);
    u_e = (w_p > M.eos.pfloor) ? u_e : ((M.eos.pfloor / gm1) + e_k + pb_);
    w_p = (w_p > M.eos.pfloor) ? w_p : M.eos.pfloor;
    w[IPR] = w_p;
    if (u_d != ud_in) ub[IDN * n + c] = u_d;
    if (u_e != ue_in) ub[IEN * n + c] = u_e;
  } else {
    double cons[NHYDRO];
#pragma unroll
    for (int q = 0; q < NHYDRO; ++q) { cons[q] = ub[q * n + c]; w[q] = pb[q * n + c]; }
    double cin[NHYDRO];
#pragma unroll
    for (int q = 0; q < NHYDRO; ++q) cin[q] = cons[q];
    sr_cons2prim_cell(cons, w, bc, M.eos, LEARNED ? C2P_LEARNED : M.c2p_mode,
                      LEARNED ? lseed[static_cast<long>(g) * n + c] : 0.0, st);
#pragma unroll
    for (int q = 0; q < NHYDRO; ++q)
      if (cons[q] != cin[q]) ub[q * n + c] = cons[q];
  }
#pragma unroll
  for (int q = 0; q < NHYDRO; ++q) pb[q * n + c] = w[q];

  if (with_dt && i >= M.is && i <= M.ie && j >= M.js && j <= M.je && k >= M.ks &&
      k <= M.ke) {
    const double *dx1 = ctab(M, g, 0), *dx2 = ctab(M, g, 1), *dx3 = ctab(M, g, 2);
    double dt1 = dx1[i], dt2 = dx2[j], dt3 = dx3[k];
    if (!SR) {
      double wi[NWAVE];
      wi[IDN] = w[IDN]; wi[IVX] = w[IVX]; wi[IVY] = w[IVY]; wi[IVZ] = w[IVZ];
      wi[IPR] = w[IPR];
      double bx = bc[IB1] + fabs(b1l - bc[IB1]);
      wi[IBY] = bc[IB2]; wi[IBZ] = bc[IB3];
      double cf = fast_speed(wi, bx, M.eos.gamma);
      dt1 /= dmax(0.0, (fabs(wi[IVX]) + cf));
      wi[IBY] = bc[IB3]; wi[IBZ] = bc[IB1];
      bx = bc[IB2] + fabs(b2l - bc[IB2]);
      cf = fast_speed(wi, bx, M.eos.gamma);
      dt2 /= dmax(0.0, (fabs(wi[IVY]) + cf));
      wi[IBY] = bc[IB1]; wi[IBZ] = bc[IB2];
      bx = bc[IB3] + fabs(b3l - bc[IB3]);
      cf = fast_speed(wi, bx, M.eos.gamma);
      dt3 /= dmax(0.0, (fabs(wi[IVZ]) + cf));
    }
    *my_dt = dmin(dmin(dt1, dt2), dt3);
  }
}

// Every thread of the work-group runs the whole body (no early return): the dt
// reduction at its end synchronises the work-group.  `slot` is the work-group's entry
// in dt_partial.
template <bool SR, bool LEARNED>
__device__ __forceinline__ void c2p_body(const MeshK &M, const C2PArgs &a, int lw, int nwg,
                                         int slot, TileK T) {
  double *__restrict__ dt_partial = a.dt_partial;
  int g, i, j, k;
  const bool live = map_kind(M, KK_C2P, lw, nwg, T, g, i, j, k);
  double my_dt = DBL_MAX;
  C2PStats st{0, 0};
  if (live) c2p_cell<SR, LEARNED>(M, a, g, k, j, i, dt_partial != nullptr, &my_dt, &st);
  if (SR && M.c2p_stats) {   // wavefront sums, one atomic per wavefront into its slot
    unsigned long long ev = st.evals, fb = st.fallback, nc = live ? 1 : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      ev += __shfl_down(ev, off, 64);
      fb += __shfl_down(fb, off, 64);
      nc += __shfl_down(nc, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      const int sl = (static_cast<int>(blockIdx.x) * (WG / 64) + (threadIdx.x >> 6)) %
                     C2P_STAT_SLOTS;
      unsigned long long *c = M.c2p_stats + 3 * sl;
      atomicAdd(&c[0], ev);
      atomicAdd(&c[1], fb);
      atomicAdd(&c[2], nc);
    }
  }
  if (dt_partial) {
    // work-group minimum: wavefront (64 lanes) shuffle, then LDS across 4 wavefronts
    __shared__ double sm[WG / 64];
    double v = my_dt;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_down(v, off, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = sm[0];
      for (int q = 1; q < WG / 64; ++q) m = dmin(m, sm[q]);
      dt_partial[slot] = m;
    }
    __syncthreads();   // sm is reused by the next work item of a persistent launch
  }
}

// At least 3 wavefronts per SIMD: the SR recovery needs ~190 VGPRs unbounded (2 waves);
// capped at 168 it spills a few registers to scratch and runs faster (the Newtonian
// instantiation needs 62 and is unaffected).
constexpr int C2P_MIN_WAVES = 3;

template <bool SR, bool LEARNED>
__global__ void __launch_bounds__(WG, C2P_MIN_WAVES) k_cons2prim(MeshK M, C2PArgs a, TileK T) {
  c2p_body<SR, LEARNED>(M, a, xcd_logical_wg(M.xcd_remap & XCD_C2P), gridDim.x, blockIdx.x, T);
}


// Learned gas-pressure seeds for C2P_LEARNED, cell for cell as k_cons2prim, from the
// conserved state and the cell-centred field it is about to see.  Wavefront-collective
// BF16 MFMA inference; a kernel of its own so that the FP64 recovery keeps its register
// budget.  Dead lanes carry a dummy state.
__device__ __forceinline__ void mesh_seed_body(const MeshK &M, const double *__restrict__ u,
                                               const double *__restrict__ b1,
                                               const double *__restrict__ b2,
                                               const double *__restrict__ b3,
                                               double *__restrict__ lseed, int lw, int nwg,
                                               TileK T) {
  int g, i, j, k;
  const bool live = map_kind(M, KK_C2P, lw, nwg, T, g, i, j, k);
  double cs[NHYDRO] = {1.0, 0.0, 0.0, 0.0, 2.0}, bs[3] = {0.0, 0.0, 0.0};
  const long c = live ? cidx(M, k, j, i) : 0;
  if (live) {
    const double *B1 = b1 + static_cast<long>(g) * M.nf1;
    const double *B2 = b2 + static_cast<long>(g) * M.nf2;
    const double *B3 = b3 + static_cast<long>(g) * M.nf3;
    bs[IB1] = 0.5 * B1[f1idx(M, k, j, i)] + 0.5 * B1[f1idx(M, k, j, i + 1)];
    bs[IB2] = 0.5 * B2[f2idx(M, k, j, i)] + 0.5 * B2[f2idx(M, k, j + 1, i)];
    bs[IB3] = 0.5 * B3[f3idx(M, k, j, i)] + 0.5 * B3[f3idx(M, k + 1, j, i)];
    const double *ub = u + static_cast<long>(g) * NHYDRO * M.ncell;
#pragma unroll
    for (int q = 0; q < NHYDRO; ++q) cs[q] = ub[q * static_cast<long>(M.ncell) + c];
  }
  float x[MLP_IN];
  mlp_features(cs, bs, M.mlp.P, x);
  const double sd = mlp_seed(cs, mlp_wave_infer_bf16(M.mlp, x));
  if (live) lseed[static_cast<long>(g) * M.ncell + c] = sd;
}

__global__ void __launch_bounds__(WG) k_mesh_seed(MeshK M, const double *__restrict__ u,
                                                  const double *__restrict__ b1,
                                                  const double *__restrict__ b2,
                                                  const double *__restrict__ b3,
                                                  double *__restrict__ lseed, TileK T) {
  mesh_seed_body(M, u, b1, b2, b3, lseed, xcd_logical_wg(M.xcd_remap & XCD_C2P), gridDim.x, T);
}

//========================================================================================
// History sums (outputs/history.cpp:84-152, MHD without gravity/scalars): volume-
// weighted mass, momenta, kinetic energy by direction, total energy and magnetic energy
// by direction over the active cells.  Per-work-group partial sums, then one work-group
// adds the partials in a fixed order, so the result does not depend on scheduling.
constexpr int NHIST = 11;

__global__ void __launch_bounds__(WG) k_history(MeshK M, const double *__restrict__ u,
                                                 const double *__restrict__ bcc,
                                                 double *__restrict__ part) {
  const int nact = M.nx1 * M.nx2 * M.nx3;
  const int nch = gridDim.x / M.nblocks;
  const int g = blockIdx.x / nch;
  const int t = (blockIdx.x % nch) * WG + threadIdx.x;
  double h[NHIST];
#pragma unroll
  for (int q = 0; q < NHIST; ++q) h[q] = 0.0;
  if (t < nact) {
    const int i = M.is + t % M.nx1;
    const int j = M.js + (t / M.nx1) % M.nx2;
    const int k = M.ks + t / (M.nx1 * M.nx2);
    const double *dx1 = ctab(M, g, 0), *dx2 = ctab(M, g, 1), *dx3 = ctab(M, g, 2);
    const double vol = dx1[i] * dx2[j] * dx3[k];
    const long n = M.ncell, c = cidx(M, k, j, i);
    const double *ub = u + static_cast<long>(g) * NHYDRO * n;
    const double *bb = bcc + static_cast<long>(g) * 3 * n;
    const double u_d = ub[IDN * n + c], u_mx = ub[IM1 * n + c];
    const double u_my = ub[IM2 * n + c], u_mz = ub[IM3 * n + c], u_e = ub[IEN * n + c];
    const double b1 = bb[IB1 * n + c], b2 = bb[IB2 * n + c], b3 = bb[IB3 * n + c];
    h[0] = vol * u_d;
    h[1] = vol * u_mx;
    h[2] = vol * u_my;
    h[3] = vol * u_mz;
    h[4] = vol * 0.5 * sq(u_mx) / u_d;
    h[5] = vol * 0.5 * sq(u_my) / u_d;
    h[6] = vol * 0.5 * sq(u_mz) / u_d;
    h[7] = vol * u_e;
    h[8] = vol * 0.5 * b1 * b1;
    h[9] = vol * 0.5 * b2 * b2;
    h[10] = vol * 0.5 * b3 * b3;
  }
  __shared__ double sm[NHIST][WG / 64];
#pragma unroll
  for (int q = 0; q < NHIST; ++q) {
    double v = h[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sm[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < NHIST) {
    double v = sm[threadIdx.x][0];
    for (int w = 1; w < WG / 64; ++w) v += sm[threadIdx.x][w];
    part[static_cast<long>(blockIdx.x) * NHIST + threadIdx.x] = v;
  }
}

__global__ void __launch_bounds__(1024) k_history_final(const double *__restrict__ part,
                                                         long npart, double *out) {
  __shared__ double sm[NHIST][1024 / 64];
  for (int q = 0; q < NHIST; ++q) {
    double v = 0.0;
    for (long p = threadIdx.x; p < npart; p += 1024) v += part[p * NHIST + q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sm[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < NHIST) {
    double v = sm[threadIdx.x][0];
    for (int w = 1; w < 1024 / 64; ++w) v += sm[threadIdx.x][w];
    out[threadIdx.x] = v;
  }
}

// Final reduction + Mesh::NewTimeStep (mesh.cpp:1236-1282) and, when advance != 0,
// the end-of-cycle clock update of the main loop (main.cpp:535-541).
__device__ __forceinline__ void new_dt_body(const double *__restrict__ part, long npart,
                                            Clock *clk, double cfl, int advance) {
  __shared__ double sm[1024 / 64];
  double v = DBL_MAX;
  for (long q = threadIdx.x; q < npart; q += 1024) v = dmin(v, part[q]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sm[0];
    for (int q = 1; q < 1024 / 64; ++q) m = dmin(m, sm[q]);
    m *= cfl;   // per-block min_dt_hyperbolic *= cfl_number; min commutes with it
    Clock c = *clk;
    if (advance) {
      c.ncycle += 1.0;
      c.time += c.dt;
    }
    double dt = 2.0 * c.dt;
    dt = dmin(dt, m);
    if (c.time < c.tlim && (c.tlim - c.time) < dt) dt = c.tlim - c.time;
    c.dt = dt;
    *clk = c;
  }
}

__global__ void __launch_bounds__(1024) k_new_dt(const double *__restrict__ part, long npart,
                                                  Clock *clk, double cfl, int advance) {
  new_dt_body(part, npart, clk, cfl, advance);
}

//========================================================================================
// Ensembles: independent meshes (uneven sizes) advanced together.  Each stage kernel is
// launched once for the whole ensemble; work item v of a launch is (member, work-group
// of that member) from a precomputed map, and a launch smaller than the map loops over
// it (persistent work-groups).  The member table holds every kernel argument of one
// stage, so a whole cycle is a fixed sequence of launches that a HIP graph can replay.
struct EnsMember {
  MeshK M;
  const double *w, *bcc;
  FaceOut fo[3];
  CornerArgs ca;
  UpdateArgs ua;
  GhostArgs ga;
  C2PArgs pa;
  Clock *clk;
  int nwg[4];       // work-groups of each work kind
  TileK tile[4];
  int nwg_ghost;    // work-groups of the ghost exchange
};

__device__ __forceinline__ int2 ens_item(const int2 *__restrict__ map, int v) {
  const int2 mw = map[v];   // the same for every lane: keep it in scalar registers
  return make_int2(__builtin_amdgcn_readfirstlane(mw.x), __builtin_amdgcn_readfirstlane(mw.y));
}

template <int ORDER, int RS>
__global__ void __launch_bounds__(WG) k_flux3_ens(const EnsMember *__restrict__ E,
                                                   const int2 *__restrict__ map, int nwork) {
  for (int v = blockIdx.x; v < nwork; v += gridDim.x) {
    const int2 mw = ens_item(map, v);
    const EnsMember &e = E[mw.x];
    flux3_body<ORDER, RS>(e.M, e.w, e.bcc, e.fo[0], e.fo[1], e.fo[2], e.clk->dt, mw.y,
                          e.nwg[0], e.tile[0]);
  }
}

template <bool SR>
__global__ void __launch_bounds__(WG) k_corner_ens(const EnsMember *__restrict__ E,
                                                    const int2 *__restrict__ map, int nwork) {
  for (int v = blockIdx.x; v < nwork; v += gridDim.x) {
    const int2 mw = ens_item(map, v);
    const EnsMember &e = E[mw.x];
    corner_body<SR>(e.M, e.ca, mw.y, e.nwg[1], e.tile[1]);
  }
}

template <bool SR>
__global__ void __launch_bounds__(WG) k_corner_march_ens(const EnsMember *__restrict__ E,
                                                         const int2 *__restrict__ map,
                                                         int nwork) {
  for (int v = blockIdx.x; v < nwork; v += gridDim.x) {
    const int2 mw = ens_item(map, v);
    const EnsMember &e = E[mw.x];
    corner_march<SR>(e.M, e.ca, mw.y, e.nwg[1], e.tile[1].tz);
  }
}

__global__ void __launch_bounds__(WG) k_ghost_ens(const EnsMember *__restrict__ E,
                                                   const int2 *__restrict__ map, int nwork) {
  for (int v = blockIdx.x; v < nwork; v += gridDim.x) {
    const int2 mw = ens_item(map, v);
    const EnsMember &e = E[mw.x];
    ghost_body(e.M, e.ga, mw.y, e.nwg_ghost);
  }
}

__global__ void __launch_bounds__(WG) k_update_ens(const EnsMember *__restrict__ E,
                                                    const int2 *__restrict__ map, int nwork,
                                                    double beta, int first_stage) {
  for (int v = blockIdx.x; v < nwork; v += gridDim.x) {
    const int2 mw = ens_item(map, v);
    const EnsMember &e = E[mw.x];
    update_body(e.M, e.ua, beta, first_stage, e.clk->dt, mw.y, e.nwg[2], e.tile[2]);
  }
}

template <bool SR, bool LEARNED>
__global__ void __launch_bounds__(WG, C2P_MIN_WAVES) k_c2p_ens(const EnsMember *__restrict__ E,
                                                 const int2 *__restrict__ map, int nwork) {
  for (int v = blockIdx.x; v < nwork; v += gridDim.x) {
    const int2 mw = ens_item(map, v);
    const EnsMember &e = E[mw.x];
    c2p_body<SR, LEARNED>(e.M, e.pa, mw.y, e.nwg[3], mw.y, e.tile[3]);
  }
}

// The learned pressure seeds of every member (k_mesh_seed over the recovery's work map).
__global__ void __launch_bounds__(WG) k_seed_ens(const EnsMember *__restrict__ E,
                                                 const int2 *__restrict__ map, int nwork) {
  for (int v = blockIdx.x; v < nwork; v += gridDim.x) {
    const int2 mw = ens_item(map, v);
    const EnsMember &e = E[mw.x];
    // the recovery only reads the seed register, which this kernel fills
    mesh_seed_body(e.M, e.pa.u, e.pa.b1, e.pa.b2, e.pa.b3, const_cast<double *>(e.pa.lseed),
                   mw.y, e.nwg[3], e.tile[3]);
  }
}

// One work-group per member: its dt reduction and clock update.
__global__ void __launch_bounds__(1024) k_new_dt_ens(const EnsMember *__restrict__ E,
                                                      int advance) {
  const EnsMember &e = E[blockIdx.x];
  new_dt_body(e.pa.dt_partial, e.nwg[3], e.clk, e.M.cfl, advance);
}

//========================================================================================
// Per-MeshBlock kernels over caller-owned arrays (ath_gpu_block_*): one MeshBlock whose
// registers the caller keeps in device memory.  The MeshK describes that one block
// (g = 0 everywhere) with no neighbours, so the EMF accessors of the update return the
// caller's EMFs unchanged: the reference's CT and AddFluxDivergence see only their own
// MeshBlock's arrays, the neighbour correction having been applied before by the caller.

// Reconstruction::{DonorCell,PiecewiseLinear,PiecewiseParabolic}X{DIR+1} on pencil
// (k, j) for cells il .. iu (dc.cpp:24-103, plm.cpp:27-463, ppm.cpp:44-923): each cell's
// right-edge value goes to the left state of the face above it and its left-edge value
// to the right state of its own face.  X1 writes wl(n, i+1) and wr(n, i); X2 and X3 write
// wl(n, i) (the face above cell j or k) and wr(n, i) (its own face).  Pencils (NWAVE,
// nc1).  The transverse field components are (IB2, IB3) along x1, (IB3, IB1) along x2
// and (IB1, IB2) along x3.
template <int DIR, int ORDER>
__global__ void __launch_bounds__(WG) k_blk_recon(MeshK M, int k, int j, int il, int iu,
                                                  const double *__restrict__ w,
                                                  const double *__restrict__ bcc,
                                                  double *__restrict__ wl,
                                                  double *__restrict__ wr) {
  const int i = il + static_cast<int>(blockIdx.x * WG + threadIdx.x);
  if (i > iu) return;
  const long s = (DIR == 0) ? 1 : ((DIR == 1) ? M.nc1 : static_cast<long>(M.nc1) * M.nc2);
  const long c = cidx(M, k, j, i);
  constexpr int cby = (DIR == 0) ? IB2 : ((DIR == 1) ? IB3 : IB1);
  constexpr int cbz = (DIR == 0) ? IB3 : ((DIR == 1) ? IB1 : IB2);
  const int fc = (DIR == 0) ? i : ((DIR == 1) ? j : k);
  const double *ct = ctab(M, 0, DIR);
  const int ncd = (DIR == 0) ? M.nc1 : ((DIR == 1) ? M.nc2 : M.nc3);
  const int il_out = (DIR == 0) ? i + 1 : i;     // pencil index of the left state
  for (int v = 0; v < NWAVE; ++v) {
    const double *src;
    if (v < NHYDRO) src = w + v * M.ncell;
    else if (v == IBY) src = bcc + cby * M.ncell;
    else src = bcc + cbz * M.ncell;
    double lo, hi;   // left-edge and right-edge values of the cell
    if (ORDER == 1) {
      lo = hi = src[c];
    } else if (ORDER == 2) {
      const double q0 = src[c];
      const double dwm = vl_slope(src[c - s], q0, src[c + s]);
      hi = q0 + ct[ncd + fc] * dwm;
      lo = q0 - ct[2 * ncd + fc] * dwm;
    } else {
      double q[5];
#pragma unroll
      for (int m = 0; m < 5; ++m) q[m] = src[c + (m - 2) * s];
      ppm_cell(q, lo, hi);
      if (v == IDN) {
        lo = (lo > M.eos.dfloor) ? lo : M.eos.dfloor;
        hi = (hi > M.eos.dfloor) ? hi : M.eos.dfloor;
      } else if (v == IPR) {
        lo = (lo > M.eos.pfloor) ? lo : M.eos.pfloor;
        hi = (hi > M.eos.pfloor) ? hi : M.eos.pfloor;
      }
    }
    wl[v * M.nc1 + il_out] = hi;
    wr[v * M.nc1 + i] = lo;
  }
}

// Hydro::RiemannSolver(k, j, il, iu, ivx, bx, wl, wr, flx, ey, ez, wct, dxw) of hlld.cpp,
// hlle_mhd.cpp, hlle_mhd_rel.cpp or hlld_rel.cpp (the solver loop and its epilogue):
// states rotated by ivx, outputs at face (k, j, i) of the face arrays along ivx, the CT
// weight from dxw(i) and dt (hydro.cpp:154-158).
template <int DIR, int RS>
__global__ void __launch_bounds__(WG) k_blk_riemann(MeshK M, int k, int j, int il, int iu,
                                                    const double *__restrict__ bx,
                                                    const double *__restrict__ wl,
                                                    const double *__restrict__ wr,
                                                    double *__restrict__ flx,
                                                    double *__restrict__ ey,
                                                    double *__restrict__ ez,
                                                    double *__restrict__ wct,
                                                    const double *__restrict__ dxw, double dt) {
  const int i = il + static_cast<int>(blockIdx.x * WG + threadIdx.x);
  if (i > iu) return;
  constexpr int ivx = IVX + DIR;
  constexpr int ivy = IVX + ((ivx - IVX) + 1) % 3;
  constexpr int ivz = IVX + ((ivx - IVX) + 2) % 3;
  const int p = M.nc1;
  double a[NWAVE], b[NWAVE], f[NWAVE];
  a[IDN] = wl[IDN * p + i]; a[IVX] = wl[ivx * p + i]; a[IVY] = wl[ivy * p + i];
  a[IVZ] = wl[ivz * p + i]; a[IPR] = wl[IPR * p + i]; a[IBY] = wl[IBY * p + i];
  a[IBZ] = wl[IBZ * p + i];
  b[IDN] = wr[IDN * p + i]; b[IVX] = wr[ivx * p + i]; b[IVY] = wr[ivy * p + i];
  b[IVZ] = wr[ivz * p + i]; b[IPR] = wr[IPR * p + i]; b[IBY] = wr[IBY * p + i];
  b[IBZ] = wr[IBZ * p + i];
  long fi, nf;
  if (DIR == 0) { fi = f1idx(M, k, j, i); nf = M.nf1; }
  else if (DIR == 1) { fi = f2idx(M, k, j, i); nf = M.nf2; }
  else { fi = f3idx(M, k, j, i); nf = M.nf3; }
  const double bxi = bx[fi];
  if (RS == 0) hlld(a, b, bxi, M.eos, f);
  else if (RS == 1) hlle(a, b, bxi, M.eos, f);
  else if (RS == 2) sr_hlle(a, b, bxi, M.eos, f);
  else sr_hlld(a, b, bxi, M.eos, f);
  flx[IDN * nf + fi] = f[IDN];
  flx[ivx * nf + fi] = f[IVX];
  flx[ivy * nf + fi] = f[IVY];
  flx[ivz * nf + fi] = f[IVZ];
  flx[IEN * nf + fi] = f[IEN];
  ey[fi] = -f[IBY];
  ez[fi] = f[IBZ];
  wct[fi] = ct_weight(f[IDN], a[IDN], b[IDN], dxw[i], dt);
}

// Hydro::CalculateFluxes on one block with dt by value (the fused three-direction body).
template <int ORDER, int RS>
__global__ void __launch_bounds__(WG) k_blk_flux3(MeshK M, const double *__restrict__ w,
                                                  const double *__restrict__ bcc, FaceOut o1,
                                                  FaceOut o2, FaceOut o3, double dt) {
  flux3_body<ORDER, RS>(M, w, bcc, o1, o2, o3, dt, blockIdx.x, gridDim.x, TileK{0, 0, 0});
}

// Field::CT (what & 2) and Hydro::AddFluxDivergence (what & 1) of one block, in place on
// the caller's b_out / u_out.
__global__ void __launch_bounds__(WG) k_blk_update(MeshK M, UpdateArgs a, double wght,
                                                   int what) {
  int g, i, j, k;
  if (!map_kind(M, KK_UPDATE, blockIdx.x, gridDim.x, TileK{0, 0, 0}, g, i, j, k)) return;
  const bool ci = i <= M.ie, cj = j <= M.je, ck = k <= M.ke;
  if ((what & 1) && ci && cj && ck)
    update_cell_u(M, g, k, j, i, a.uout, a.uout, a.f1, a.f2, a.f3, wght, 0);
  if (what & 2) {
    if (cj && ck)
      update_face_b<0>(M, g, k, j, i, a.bout[0], a.bout[0], a.e1, a.e2, a.e3, wght, 0);
    if (ci && ck)
      update_face_b<1>(M, g, k, j, i, a.bout[1], a.bout[1], a.e1, a.e2, a.e3, wght, 0);
    if (ci && cj)
      update_face_b<2>(M, g, k, j, i, a.bout[2], a.bout[2], a.e1, a.e2, a.e3, wght, 0);
  }
}

// EquationOfState::ConservedToPrimitive over the box lo[] .. lo[] + n[] - 1 of one block.
template <bool SR>
__global__ void __launch_bounds__(WG, C2P_MIN_WAVES) k_blk_c2p(MeshK M, C2PArgs a, int il,
                                                               int jl, int kl, int ni, int nj,
                                                               int nk) {
  const int t = static_cast<int>(blockIdx.x * WG + threadIdx.x);
  if (t >= ni * nj * nk) return;
  const int i = il + t % ni, j = jl + (t / ni) % nj, k = kl + t / (ni * nj);
  double dt = DBL_MAX;
  c2p_cell<SR, false>(M, a, 0, k, j, i, false, &dt, nullptr);
}

// Hydro::NewBlockTimeStep (new_blockdt.cpp:45-230) of one block from its primitives,
// cell-centred and face fields: per work-group minima of the active cells' time steps
// into part[], then k_blk_dt_final takes their minimum and applies cfl_number.
template <bool SR>
__global__ void __launch_bounds__(WG) k_blk_dt(MeshK M, const double *__restrict__ w,
                                               const double *__restrict__ bcc,
                                               const double *__restrict__ b1,
                                               const double *__restrict__ b2,
                                               const double *__restrict__ b3,
                                               double *__restrict__ part) {
  const int t = static_cast<int>(blockIdx.x * WG + threadIdx.x);
  double my = DBL_MAX;
  if (t < M.nx1 * M.nx2 * M.nx3) {
    const int i = M.is + t % M.nx1, j = M.js + (t / M.nx1) % M.nx2;
    const int k = M.ks + t / (M.nx1 * M.nx2);
    const double *dx1 = ctab(M, 0, 0), *dx2 = ctab(M, 0, 1), *dx3 = ctab(M, 0, 2);
    double dt1 = dx1[i], dt2 = dx2[j], dt3 = dx3[k];
    if (!SR) {
      const long n = M.ncell, c = cidx(M, k, j, i);
      const double bc1 = bcc[IB1 * n + c], bc2 = bcc[IB2 * n + c], bc3 = bcc[IB3 * n + c];
      double wi[NWAVE];
      wi[IDN] = w[IDN * n + c]; wi[IVX] = w[IVX * n + c]; wi[IVY] = w[IVY * n + c];
      wi[IVZ] = w[IVZ * n + c]; wi[IPR] = w[IPR * n + c];
      double bx = bc1 + fabs(b1[f1idx(M, k, j, i)] - bc1);
      wi[IBY] = bc2; wi[IBZ] = bc3;
      double cf = fast_speed(wi, bx, M.eos.gamma);
      dt1 /= dmax(0.0, (fabs(wi[IVX]) + cf));
      wi[IBY] = bc3; wi[IBZ] = bc1;
      bx = bc2 + fabs(b2[f2idx(M, k, j, i)] - bc2);
      cf = fast_speed(wi, bx, M.eos.gamma);
      dt2 /= dmax(0.0, (fabs(wi[IVY]) + cf));
      wi[IBY] = bc1; wi[IBZ] = bc2;
      bx = bc3 + fabs(b3[f3idx(M, k, j, i)] - bc3);
      cf = fast_speed(wi, bx, M.eos.gamma);
      dt3 /= dmax(0.0, (fabs(wi[IVZ]) + cf));
    }
    my = dmin(dmin(dt1, dt2), dt3);
  }
  __shared__ double sm[WG / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) my = dmin(my, __shfl_down(my, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = my;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sm[0];
    for (int q = 1; q < WG / 64; ++q) m = dmin(m, sm[q]);
    part[blockIdx.x] = m;
  }
}

__global__ void __launch_bounds__(1024) k_blk_dt_final(const double *__restrict__ part,
                                                       int npart, double cfl,
                                                       double *__restrict__ dt_out) {
  __shared__ double sm[1024 / 64];
  double v = DBL_MAX;
  for (int q = threadIdx.x; q < npart; q += 1024) v = dmin(v, part[q]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_down(v, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sm[0];
    for (int q = 1; q < 1024 / 64; ++q) m = dmin(m, sm[q]);
    *dt_out = m * cfl;   // min_dt_hyperbolic *= cfl_number (new_blockdt.cpp:154)
  }
}

// Batched single-interface Riemann solves (known-answer vectors)
__global__ void k_riemann_batch(int solver, EosK eos, int n, const double *wl,
                                const double *wr, const double *bx, double *out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double a[NWAVE], b[NWAVE], f[NWAVE];
  for (int v = 0; v < NWAVE; ++v) { a[v] = wl[q * NWAVE + v]; b[v] = wr[q * NWAVE + v]; }
  if (solver == 0) hlld(a, b, bx[q], eos, f);
  else if (solver == 1) hlle(a, b, bx[q], eos, f);
  else if (solver == 2) sr_hlle(a, b, bx[q], eos, f);
  else sr_hlld(a, b, bx[q], eos, f);
  for (int v = 0; v < NWAVE; ++v) out[q * NWAVE + v] = f[v];
}

__global__ void k_sr_c2p_batch(EosK eos, int n, double *cons, double *prim,
                               const double *bcc, int *iters) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double c[5], p[5], b[3];
  for (int v = 0; v < 5; ++v) { c[v] = cons[q * 5 + v]; p[v] = prim[q * 5 + v]; }
  for (int v = 0; v < 3; ++v) b[v] = bcc[q * 3 + v];
  iters[q] = sr_cons2prim_cell(c, p, b, eos, C2P_REFERENCE, 0.0, nullptr);
  for (int v = 0; v < 5; ++v) { cons[q * 5 + v] = c[v]; prim[q * 5 + v] = p[v]; }
}

//========================================================================================
// Static mesh refinement: the step programs built on the host (smr_plan.hpp) run phase
// by phase, one launch per phase over the steps of every MeshBlock in it, a work-group
// per chunk of 256 elements of one step.  A step writes only its own block's registers;
// it reads other blocks' active data (boundary exchange, flux correction) or a snapshot of
// their EMFs (EMF correction), which nothing in the same launch writes.

struct SmrArgs {
  const smrplan::Step *steps;
  const int *chunk_step, *chunk_first;
  double *u, *w, *bcc;              // the updated state; w takes prolongated primitives
  double *b[3];
  double *cu, *cw, *cbcc, *cb[3];   // coarse buffers: conserved, primitive, bcc, faces
  const double *fco, *cco;          // fine / coarse coordinate tables per block
  long fcl, ccl;
  double *flux[3];
  double *e[3];
  const double *es[3];              // EMF snapshot
};

struct SmrDims {
  int c1, c2, c3;                   // coarse cells per direction (nx/2 + 2 ng)
  long ccell;
  long cf[3];                       // coarse face-array sizes
};

__device__ __forceinline__ SmrDims smr_dims(const MeshK &M) {
  SmrDims D;
  D.c1 = M.nx1 / 2 + 2 * M.ng; D.c2 = M.nx2 / 2 + 2 * M.ng; D.c3 = M.nx3 / 2 + 2 * M.ng;
  D.ccell = static_cast<long>(D.c1) * D.c2 * D.c3;
  D.cf[0] = static_cast<long>(D.c3) * D.c2 * (D.c1 + 1);
  D.cf[1] = static_cast<long>(D.c3) * (D.c2 + 1) * D.c1;
  D.cf[2] = static_cast<long>(D.c3 + 1) * D.c2 * D.c1;
  return D;
}

// views of one block's arrays
__device__ __forceinline__ refops::View cell_view(const MeshK &M, double *p, int g, int n) {
  return refops::View{p + (static_cast<long>(g) * NHYDRO + n) * M.ncell, M.nc2, M.nc1};
}
__device__ __forceinline__ refops::View face_view(const MeshK &M, double *p, int g, int c) {
  const long nf = (c == 0) ? M.nf1 : ((c == 1) ? M.nf2 : M.nf3);
  return refops::View{p + static_cast<long>(g) * nf, M.nc2 + (c == 1), M.nc1 + (c == 0)};
}
__device__ __forceinline__ refops::View edge_view(const MeshK &M, const double *p, int g,
                                                  int c) {
  const long ne = (c == 0) ? M.ne1 : ((c == 1) ? M.ne2 : M.ne3);
  return refops::View{const_cast<double *>(p) + static_cast<long>(g) * ne, M.nc2 + (c != 1),
                      M.nc1 + (c != 0)};
}
__device__ __forceinline__ refops::View ccell_view(const SmrDims &D, double *p, int g, int n,
                                                   int nvar) {
  return refops::View{p + (static_cast<long>(g) * nvar + n) * D.ccell, D.c2, D.c1};
}
__device__ __forceinline__ refops::View cface_view(const SmrDims &D, double *p, int g, int c) {
  return refops::View{p + static_cast<long>(g) * D.cf[c], D.c2 + (c == 1), D.c1 + (c == 0)};
}

// ConservedToPrimitive on a coarse-buffer cell (adiabatic_mhd.cpp:41-90 with the coarse
// face fields, as ApplyPhysicalBoundariesOnCoarseLevel calls it).
__device__ __forceinline__ void smr_coarse_c2p(const MeshK &M, const SmrDims &D,
                                               const SmrArgs &a, int g, int k, int j, int i) {
  const refops::View B1 = cface_view(D, a.cb[0], g, 0), B2 = cface_view(D, a.cb[1], g, 1);
  const refops::View B3 = cface_view(D, a.cb[2], g, 2);
  const double bc1 = 0.5 * B1(k, j, i) + 0.5 * B1(k, j, i + 1);
  const double bc2 = 0.5 * B2(k, j, i) + 0.5 * B2(k, j + 1, i);
  const double bc3 = 0.5 * B3(k, j, i) + 0.5 * B3(k + 1, j, i);
  ccell_view(D, a.cbcc, g, 0, 3)(k, j, i) = bc1;
  ccell_view(D, a.cbcc, g, 1, 3)(k, j, i) = bc2;
  ccell_view(D, a.cbcc, g, 2, 3)(k, j, i) = bc3;
  double &u_d = ccell_view(D, a.cu, g, IDN, NHYDRO)(k, j, i);
  double &u_e = ccell_view(D, a.cu, g, IEN, NHYDRO)(k, j, i);
  const double u_m1 = ccell_view(D, a.cu, g, IM1, NHYDRO)(k, j, i);
  const double u_m2 = ccell_view(D, a.cu, g, IM2, NHYDRO)(k, j, i);
  const double u_m3 = ccell_view(D, a.cu, g, IM3, NHYDRO)(k, j, i);
  const double gm1 = M.eos.gamma - 1.0;
  u_d = (u_d > M.eos.dfloor) ? u_d : M.eos.dfloor;
  ccell_view(D, a.cw, g, IDN, NHYDRO)(k, j, i) = u_d;
  const double di = 1.0 / u_d;
  ccell_view(D, a.cw, g, IVX, NHYDRO)(k, j, i) = u_m1 * di;
  ccell_view(D, a.cw, g, IVY, NHYDRO)(k, j, i) = u_m2 * di;
  ccell_view(D, a.cw, g, IVZ, NHYDRO)(k, j, i) = u_m3 * di;
  const double pb = 0.5 * (bc1 * bc1 + bc2 * bc2 + bc3 * bc3);
  const double e_k = 0.5 * di * (u_m1 * u_m1 + u_m2 * u_m2 + u_m3 * u_m3);
  const double w_p = gm1 * (u_e - e_k - pb);
  u_e = (w_p > M.eos.pfloor) ? u_e : ((M.eos.pfloor / gm1) + e_k + pb);
  ccell_view(D, a.cw, g, IPR, NHYDRO)(k, j, i) = (w_p > M.eos.pfloor) ? w_p : M.eos.pfloor;
}

// bcc from the faces and PrimitiveToConserved on a fine cell (field.cpp:112-180 uniform
// branch, adiabatic_mhd.cpp:99-136).
__device__ __forceinline__ void smr_bcc_p2c(const MeshK &M, const SmrArgs &a, int g, int k,
                                            int j, int i) {
  const refops::View B1 = face_view(M, a.b[0], g, 0), B2 = face_view(M, a.b[1], g, 1);
  const refops::View B3 = face_view(M, a.b[2], g, 2);
  const double bc1 = 0.5 * B1(k, j, i) + 0.5 * B1(k, j, i + 1);
  const double bc2 = 0.5 * B2(k, j, i) + 0.5 * B2(k, j + 1, i);
  const double bc3 = 0.5 * B3(k, j, i) + 0.5 * B3(k + 1, j, i);
  const long c = cidx(M, k, j, i), bb = static_cast<long>(g) * 3 * M.ncell;
  a.bcc[bb + c] = bc1;
  a.bcc[bb + M.ncell + c] = bc2;
  a.bcc[bb + 2 * M.ncell + c] = bc3;
  const long ub = static_cast<long>(g) * NHYDRO * M.ncell;
  const double w_d = a.w[ub + c], w_vx = a.w[ub + M.ncell + c];
  const double w_vy = a.w[ub + 2 * M.ncell + c], w_vz = a.w[ub + 3 * M.ncell + c];
  const double w_p = a.w[ub + 4 * M.ncell + c];
  a.u[ub + c] = w_d;
  a.u[ub + M.ncell + c] = w_vx * w_d;
  a.u[ub + 2 * M.ncell + c] = w_vy * w_d;
  a.u[ub + 3 * M.ncell + c] = w_vz * w_d;
  a.u[ub + 4 * M.ncell + c] =
      w_p * M.eos.igm1 + 0.5 * (w_d * (w_vx * w_vx + w_vy * w_vy + w_vz * w_vz) +
                                (bc1 * bc1 + bc2 * bc2 + bc3 * bc3));
}

// Area-weighted restriction of the fine fluxes of block src whose lowest face is
// (k, j, i) (LoadFluxBoundaryBufferToCoarser, flux_correction_cc.cpp:77-166).
__device__ __forceinline__ double smr_flux_restrict(const MeshK &M, const SmrArgs &a, int dir,
                                                    int src, int n, int k, int j, int i) {
  const long nf = (dir == 0) ? M.nf1 : ((dir == 1) ? M.nf2 : M.nf3);
  const refops::View f{a.flux[dir] + (static_cast<long>(src) * NHYDRO + n) * nf,
                       M.nc2 + (dir == 1), M.nc1 + (dir == 0)};
  const double *d1 = ctab(M, src, 0), *d2 = ctab(M, src, 1), *d3 = ctab(M, src, 2);
  if (dir == 0) {
    const double amm = d2[j] * d3[k], amp = d2[j + 1] * d3[k];
    const double apm = d2[j] * d3[k + 1], app = d2[j + 1] * d3[k + 1];
    const double tarea = amm + amp + apm + app;
    return (f(k, j, i) * amm + f(k, j + 1, i) * amp + f(k + 1, j, i) * apm +
            f(k + 1, j + 1, i) * app) / tarea;
  } else if (dir == 1) {
    const double s00 = d1[i] * d3[k], s01 = d1[i + 1] * d3[k];
    const double s10 = d1[i] * d3[k + 1], s11 = d1[i + 1] * d3[k + 1];
    const double tarea = s00 + s01 + s10 + s11;
    return (f(k, j, i) * s00 + f(k, j, i + 1) * s01 + f(k + 1, j, i) * s10 +
            f(k + 1, j, i + 1) * s11) / tarea;
  }
  const double s00 = d1[i] * d2[j], s01 = d1[i + 1] * d2[j];
  const double s10 = d1[i] * d2[j + 1], s11 = d1[i + 1] * d2[j + 1];
  const double tarea = s00 + s01 + s10 + s11;
  return (f(k, j, i) * s00 + f(k, j, i + 1) * s01 + f(k, j + 1, i) * s10 +
          f(k, j + 1, i + 1) * s11) / tarea;
}

__global__ void __launch_bounds__(smrplan::CHUNK_THREADS) k_smr(MeshK M, SmrArgs a, int chunk0) {
  const int ch = chunk0 + blockIdx.x;
  const smrplan::Step st = a.steps[a.chunk_step[ch]];
  const int g = st.blk;
  const SmrDims D = smr_dims(M);
  const refops::RGrid Fg(a.fco + g * a.fcl, M.nc1, M.nc2, M.nc3);
  const refops::RGrid Cg(a.cco + g * a.ccl, D.c1, D.c2, D.c3);
  {
    const bool nv5 = st.op == smrplan::CC_COPY || st.op == smrplan::CC_RESTRICT ||
                     st.op == smrplan::CC_TO_COARSE || st.op == smrplan::CC_OWN_RESTRICT ||
                     st.op == smrplan::PROLONG_CC || st.op == smrplan::FLUX_RESTRICT;
    const int nb = st.n[0] * st.n[1] * st.n[2];
    const int total = nv5 ? NHYDRO * nb : nb;
    const int t_end = min(total, a.chunk_first[ch] + smrplan::CHUNK);
    for (int t = a.chunk_first[ch] + static_cast<int>(threadIdx.x); t < t_end;
         t += smrplan::CHUNK_THREADS) {
      const int n = t / nb;
      int q = t - n * nb;
      const int di = q % st.n[0];
      q /= st.n[0];
      const int dj = q % st.n[1], dk = q / st.n[1];
      const int i = st.d0[0] + di, j = st.d0[1] + dj, k = st.d0[2] + dk;
      const int si = st.s0[0] + st.mul * di, sj = st.s0[1] + st.mul * dj;
      const int sk = st.s0[2] + st.mul * dk;
      const int c = st.comp;
      switch (st.op) {
        case smrplan::CC_COPY:
          cell_view(M, a.u, g, n)(k, j, i) = cell_view(M, a.u, st.src, n)(sk, sj, si);
          break;
        case smrplan::CC_RESTRICT: {
          const refops::RGrid Fs(a.fco + st.src * a.fcl, M.nc1, M.nc2, M.nc3);
          cell_view(M, a.u, g, n)(k, j, i) = refops::restrict_cc_at(
              cell_view(M, a.u, st.src, n), Fs, (sk - M.ng) * 2 + M.ng, (sj - M.ng) * 2 + M.ng,
              (si - M.ng) * 2 + M.ng);
          break;
        }
        case smrplan::CC_TO_COARSE:
          ccell_view(D, a.cu, g, n, NHYDRO)(k, j, i) = cell_view(M, a.u, st.src, n)(sk, sj, si);
          break;
        case smrplan::CC_OWN_RESTRICT:
          ccell_view(D, a.cu, g, n, NHYDRO)(k, j, i) = refops::restrict_cc_at(
              cell_view(M, a.u, g, n), Fg, (k - M.ng) * 2 + M.ng, (j - M.ng) * 2 + M.ng,
              (i - M.ng) * 2 + M.ng);
          break;
        case smrplan::FC_COPY:
          face_view(M, a.b[c], g, c)(k, j, i) = face_view(M, a.b[c], st.src, c)(sk, sj, si);
          break;
        case smrplan::FC_RESTRICT: {
          const refops::RGrid Fs(a.fco + st.src * a.fcl, M.nc1, M.nc2, M.nc3);
          face_view(M, a.b[c], g, c)(k, j, i) = refops::restrict_fc_at(
              c, face_view(M, a.b[c], st.src, c), Fs, (sk - M.ng) * 2 + M.ng,
              (sj - M.ng) * 2 + M.ng, (si - M.ng) * 2 + M.ng);
          break;
        }
        case smrplan::FC_TO_COARSE:
          cface_view(D, a.cb[c], g, c)(k, j, i) = face_view(M, a.b[c], st.src, c)(sk, sj, si);
          break;
        case smrplan::FC_OWN_RESTRICT:
          cface_view(D, a.cb[c], g, c)(k, j, i) = refops::restrict_fc_at(
              c, face_view(M, a.b[c], g, c), Fg, (k - M.ng) * 2 + M.ng, (j - M.ng) * 2 + M.ng,
              (i - M.ng) * 2 + M.ng);
          break;
        case smrplan::C2P_COARSE:
          smr_coarse_c2p(M, D, a, g, k, j, i);
          break;
        case smrplan::PROLONG_CC:
          refops::prolong_cc_at(ccell_view(D, a.cw, g, n, NHYDRO), cell_view(M, a.w, g, n), Fg,
                                Cg, M.ng, k, j, i);
          break;
        case smrplan::PROLONG_FC_SHARED:
          refops::prolong_fc_shared_at(c, cface_view(D, a.cb[c], g, c),
                                       face_view(M, a.b[c], g, c), Fg, Cg, M.ng, k, j, i);
          break;
        case smrplan::PROLONG_FC_INTERNAL:
          refops::prolong_fc_internal_at(face_view(M, a.b[0], g, 0), face_view(M, a.b[1], g, 1),
                                         face_view(M, a.b[2], g, 2), Fg, M.ng, k, j, i);
          break;
        case smrplan::BCC_P2C:
          smr_bcc_p2c(M, a, g, k, j, i);
          break;
        case smrplan::FLUX_RESTRICT: {
          const long nf = (c == 0) ? M.nf1 : ((c == 1) ? M.nf2 : M.nf3);
          const refops::View f{a.flux[c] + (static_cast<long>(g) * NHYDRO + n) * nf,
                               M.nc2 + (c == 1), M.nc1 + (c == 0)};
          f(k, j, i) = smr_flux_restrict(M, a, c, st.src, n, sk, sj, si);
          break;
        }
        case smrplan::EMF_ADD_SAME:
          edge_view(M, a.e[c], g, c)(k, j, i) += edge_view(M, a.es[c], st.src, c)(sk, sj, si);
          break;
        case smrplan::EMF_CLEAR:
          edge_view(M, a.e[c], g, c)(k, j, i) = 0.0;
          break;
        case smrplan::EMF_ADD_FINE: {
          const refops::RGrid Fs(a.fco + st.src * a.fcl, M.nc1, M.nc2, M.nc3);
          edge_view(M, a.e[c], g, c)(k, j, i) +=
              refops::restrict_emf_at(c, edge_view(M, a.es[c], st.src, c), Fs, sk, sj, si);
          break;
        }
        case smrplan::EMF_SCALE:
          edge_view(M, a.e[c], g, c)(k, j, i) *= st.f;
          break;
        default:
          break;
      }
    }
  }
}

//========================================================================================
// Host side

double mesh_gen_x(long index, long nrange) {   // mesh.hpp:457-472, sym_interval = true
  long noffset = index - nrange / 2;
  long noffset_ceil = index - (nrange + 1) / 2;
  return static_cast<double>(noffset + noffset_ceil) / (2.0 * nrange);
}
double uniform_gen(double x, double xmin, double xmax) {   // mesh.hpp:535-547
  return static_cast<double>(0.5) * (xmin + xmax) + (x * xmax - x * xmin);
}

}  // namespace

// Device view of a network created by ath_ml_create (mlp.hip).
MlpK ath_ml_device_view(ath_mlp_t m);

struct ath_mesh {
  ath_mesh_desc d;
  MeshK M;
  int nbx[3];
  hipStream_t stream = nullptr;
  // registers
  double *u = nullptr, *u1 = nullptr, *w = nullptr, *w1 = nullptr, *bcc = nullptr;
  double *b[3] = {nullptr, nullptr, nullptr}, *bb[3] = {nullptr, nullptr, nullptr};
  double *flux[3] = {nullptr, nullptr, nullptr};
  double *e2x1 = nullptr, *e3x1 = nullptr, *e1x2 = nullptr, *e3x2 = nullptr;
  double *e1x3 = nullptr, *e2x3 = nullptr, *wx[3] = {nullptr, nullptr, nullptr};
  double *e[3] = {nullptr, nullptr, nullptr};
  double *dtpart = nullptr;
  long ndtpart = 0;           // capacity of dtpart
  long ndt_used = 0;          // entries written by the last cons2prim with dt
  // per work kind: linear chunks, the corner kernel's z-march over 8 planes
  TileK tile[4] = {{0, 0, 0}, {-1, 0, 8}, {0, 0, 0}, {0, 0, 0}};
  int fdiv_rows = 0;          // > 0: stages use k_flux_div with bands of this many rows
  // ath_gpu_cycle as one captured HIP graph per cycle (ath_gpu_mesh_set_graph)
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<const void *> graph_key;   // what the captured cycle depends on
  long nstage = 0;            // stages run (the register configuration's parity)
  // static mesh refinement (ath_gpu_smr_mesh_create)
  bool smr = false;
  std::vector<smrplan::BlockInfo> smr_blocks;
  // the flux, EMF and boundary programs: steps, per-chunk (step, first element), and the
  // chunk offset of each phase (one launch per phase)
  smrplan::Step *smr_steps[3] = {nullptr, nullptr, nullptr};
  int *smr_chunk_step[3] = {nullptr, nullptr, nullptr};
  int *smr_chunk_first[3] = {nullptr, nullptr, nullptr};
  std::vector<int> smr_phases[3];
  double *cu = nullptr, *cw = nullptr, *cbcc = nullptr, *cb[3] = {nullptr, nullptr, nullptr};
  double *fco = nullptr, *cco = nullptr, *esnap[3] = {nullptr, nullptr, nullptr};
  long fcl = 0, ccl = 0;
  int *nbr_d = nullptr;
  double *coord_d = nullptr;
  Clock *clk = nullptr;
  unsigned long long *c2p_stats_d = nullptr;   // SR recovery counters, allocated on demand
  double *lseed = nullptr;   // learned pressure seeds (one scalar register), on demand
  double *hist_d = nullptr;  // history partial sums, on demand
  long nhist = 0;
  std::vector<void *> allocs;
  // optional per-kernel event timing (ath_gpu_kernel_timing)
  bool prof_on = false;
  struct Rec { int id; hipEvent_t a, b; };
  std::vector<Rec> prof_rec;
  double prof_ms[64] = {0};
  long prof_cnt[64] = {0};

  int alloc(double **p, long n) {
    HIPCHK(hipMalloc(reinterpret_cast<void **>(p), sizeof(double) * n));
    HIPCHK(hipMemsetAsync(*p, 0, sizeof(double) * n, stream));
    allocs.push_back(*p);
    return 0;
  }
};

namespace {

// kernel ids for the per-kernel timer
enum KernelId {
  KID_FLUX = 0,          // + 3*dir + (order-1): 9 flux variants
  KID_CORNER = 9, KID_UPDATE_U = 10, KID_SEED = 11, KID_CONS2PRIM = 12,
  KID_NEW_DT = 13, KID_GHOST = 14, KID_FLUX3 = 15,   // + (order-1): 3 fused variants
  KID_PHYS_BND = 18, KID_FDIV = 19,                  // + (order-1): fused flux + divergence
  KID_UPDATE_B = 22, KID_SMR_FLUX = 23, KID_SMR_EMF = 24, KID_SMR_BND = 25,
  KID_COUNT = 26
};
const char *kernel_name(int id) {
  static const char *flux[9] = {
      "flux_x1_dc", "flux_x1_plm", "flux_x1_ppm", "flux_x2_dc", "flux_x2_plm", "flux_x2_ppm",
      "flux_x3_dc", "flux_x3_plm", "flux_x3_ppm"};
  if (id >= 0 && id < 9) return flux[id];
  switch (id) {
    case KID_CORNER: return "corner_emf";
    case KID_UPDATE_U: return "update_u_b_ct";
    case KID_SEED: return "c2p_learned_seed";
    case KID_CONS2PRIM: return "cons2prim";
    case KID_NEW_DT: return "new_dt";
    case KID_GHOST: return "ghost_exchange";
    case KID_FLUX3: return "flux_xyz_dc";
    case KID_FLUX3 + 1: return "flux_xyz_plm";
    case KID_FLUX3 + 2: return "flux_xyz_ppm";
    case KID_PHYS_BND: return "physical_boundary";
    case KID_FDIV: return "flux_div_dc";
    case KID_FDIV + 1: return "flux_div_plm";
    case KID_FDIV + 2: return "flux_div_ppm";
    case KID_UPDATE_B: return "update_b_ct";
    case KID_SMR_FLUX: return "smr_flux_correction";
    case KID_SMR_EMF: return "smr_emf_correction";
    case KID_SMR_BND: return "smr_boundary";
    default: return "unknown";
  }
}

struct ProfScope {
  ath_mesh *m;
  int id;
  hipEvent_t b = nullptr;
  ProfScope(ath_mesh *mm, int kid) : m(mm), id(kid) {
    if (!m->prof_on) return;
    hipEvent_t a;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      m->prof_on = false;
      return;
    }
    (void)hipEventRecord(a, m->stream);
    m->prof_rec.push_back({id, a, b});
  }
  ~ProfScope() {
    if (m->prof_on && b) (void)hipEventRecord(b, m->stream);
  }
};

FaceOut face_out(ath_mesh *m, int dir) {
  FaceOut o;
  o.bn = m->b[dir];
  o.flx = m->flux[dir];
  o.wct = m->wx[dir];
  if (dir == 0) { o.ey = m->e3x1; o.ez = m->e2x1; }
  else if (dir == 1) { o.ey = m->e1x2; o.ez = m->e3x2; }
  else { o.ey = m->e2x3; o.ez = m->e1x3; }
  return o;
}

// Riemann solver template index: 0 HLLD, 1 HLLE (Newtonian), 2 SR HLLE, 3 SR HLLD.
int solver_code(const ath_mesh_desc &d) {
  if (d.relativistic) return (d.rsolver == 0) ? 3 : 2;
  return d.rsolver;
}


// Work-groups per MeshBlock of kind `kind` under tile T (map_work's nch).
long kind_chunks(const MeshK &M, int kind, TileK T) {
  int lo[3], n[3];
  kind_box(M, kind, lo, n);
  if (T.tx < 0 && kind == KK_CORNER) {
    const MarchGeom G = march_geom(M, T.tz);
    return static_cast<long>(G.nti) * G.ntj * G.nks;
  }
  if (T.tx <= 0) return (static_cast<long>(n[0]) * n[1] * n[2] + WG - 1) / WG;
  return static_cast<long>((n[0] + T.tx - 1) / T.tx) * ((n[1] + T.ty - 1) / T.ty) *
         ((n[2] + T.tz - 1) / T.tz);
}

// Tile shapes the autotuner chooses from (index 0: linear chunks); indices 6 and 7 are
// the z-march over segments of 8 and 16 planes: the corner kernel's (k_corner_march) and,
// on single meshes, the flux kernel's (k_flux_zm).
constexpr int NTILES = 8;
constexpr TileK TILES[NTILES] = {{0, 0, 0}, {64, 4, 1}, {32, 8, 1}, {16, 16, 1}, {16, 4, 4},
                                 {8, 8, 4}, {-1, 0, 8}, {-1, 0, 16}};

bool tile_ok(int kind, int t) { return TILES[t].tx >= 0 || kind == KK_CORNER; }
// A single mesh also runs the flux kernel's z-march (k_flux_zm); ensembles do not.
bool mesh_tile_ok(int kind, int t) { return tile_ok(kind, t) || kind == KK_FLUX; }

// Dynamic LDS bytes of a corner launch with tile T.
size_t corner_lds(const MeshK &M, TileK T) {
  return (T.tx < 0) ? sizeof(double) * march_lds_doubles(march_geom(M, T.tz)) : 0;
}

// Register configuration of a mesh (the pointers the VL2 stages swap).
struct Regs {
  double *u, *u1, *w, *w1, *b[3], *bb[3];
};
Regs regs_of(const ath_mesh *m) {
  Regs r;
  r.u = m->u; r.u1 = m->u1; r.w = m->w; r.w1 = m->w1;
  for (int c = 0; c < 3; ++c) { r.b[c] = m->b[c]; r.bb[c] = m->bb[c]; }
  return r;
}

// Every kernel argument of one VL2 stage given the register configuration at its start:
// flux/corner read w, b; the update writes u1, b1 and then the registers swap
// (IntegrateHydro/IntegrateField: WeightedAve into zeroed u1/b1 and SwapAthenaArray;
// stage 2 is in place on the swapped-in U^n, B^n); ConservedToPrimitive(u, w, b, w1,
// bcc) then writes the output register w1 and w swaps with w1.
struct StageArgs {
  int order, first;
  double beta;
  const double *w, *bcc;
  FaceOut fo[3];
  CornerArgs ca;
  UpdateArgs ua;
  GhostArgs ga;
  C2PArgs pa;
};

StageArgs stage_args(const ath_mesh *m, const Regs &r, int stg) {
  StageArgs a;
  a.order = (stg == 1) ? 1 : m->d.xorder;
  a.first = (stg == 1) ? 1 : 0;
  a.beta = (stg == 1) ? 0.5 : 1.0;
  a.w = r.w;
  a.bcc = m->bcc;
  for (int d = 0; d < 3; ++d) {
    FaceOut &o = a.fo[d];
    o.bn = r.b[d];
    o.flx = m->flux[d];
    o.wct = m->wx[d];
    if (d == 0) { o.ey = m->e3x1; o.ez = m->e2x1; }
    else if (d == 1) { o.ey = m->e1x2; o.ez = m->e3x2; }
    else { o.ey = m->e2x3; o.ez = m->e1x3; }
  }
  a.ca = CornerArgs{r.w, m->bcc, m->e2x1, m->e3x1, m->e1x2, m->e3x2, m->e1x3, m->e2x3,
                    m->wx[0], m->wx[1], m->wx[2], m->e[0], m->e[1], m->e[2]};
  UpdateArgs &ua = a.ua;
  ua.ubase = (stg == 1) ? r.u : r.u1;
  ua.uout = r.u1;
  ua.f1 = m->flux[0]; ua.f2 = m->flux[1]; ua.f3 = m->flux[2];
  for (int c = 0; c < 3; ++c) {
    ua.bin[c] = (stg == 1) ? r.b[c] : r.bb[c];
    ua.bout[c] = r.bb[c];
  }
  ua.e1 = m->e[0]; ua.e2 = m->e[1]; ua.e3 = m->e[2];
  a.ga = GhostArgs{ua.uout, {ua.bout[0], ua.bout[1], ua.bout[2]}};
  // after the update's swap, u is the updated state and b the updated faces
  a.pa = C2PArgs{r.u1, r.w1, m->bcc, r.bb[0], r.bb[1], r.bb[2],
                 (stg == 2) ? m->dtpart : nullptr, m->lseed};
  return a;
}

dim3 kind_grid(const ath_mesh *m, int kind) {
  return dim3(static_cast<unsigned>(kind_chunks(m->M, kind, m->tile[kind]) * m->M.nblocks));
}

int launch_flux_args(ath_mesh *m, int order, const double *w, const double *bcc,
                     const FaceOut *fo) {
  const MeshK &M = m->M;
  const int rs = solver_code(m->d);
  const TileK T = m->tile[KK_FLUX];
  // z-march over segments of T.tz planes for PLM and PPM; a donor-cell launch has nothing
  // to share and keeps linear chunks (the z-march measured 9-11 % slower there)
  const bool zm = T.tx < 0 && order > 1;
  const dim3 grid = zm ? dim3(static_cast<unsigned>((fzm_geom(M, T.tz).nwave + WG / 64 - 1) /
                                                    (WG / 64)))
                       : kind_grid(m, KK_FLUX);
  ProfScope ps(m, KID_FLUX3 + (order - 1));
#define LF(O, R)                                                                          \
  hipLaunchKernelGGL((k_flux3<O, R>), grid, dim3(WG), 0, m->stream, M, w, bcc, fo[0], fo[1], \
                     fo[2], m->clk, T)
#define LZ(O, R)                                                                          \
  hipLaunchKernelGGL((k_flux_zm<O, R>), grid, dim3(WG), 0, m->stream, M, w, bcc, fo[0],   \
                     fo[1], fo[2], m->clk, T.tz)
#define LF_O(R)                                    \
  do {                                             \
    if (order == 1) LF(1, R);                      \
    else if (order == 2 && zm) LZ(2, R);           \
    else if (order == 2) LF(2, R);                 \
    else if (zm) LZ(3, R);                         \
    else LF(3, R);                                 \
  } while (0)
  if (rs == 0) LF_O(0);
  else if (rs == 1) LF_O(1);
  else if (rs == 2) LF_O(2);
  else LF_O(3);
#undef LF_O
#undef LZ
#undef LF
  HIPCHK(hipGetLastError());
  return 0;
}

// All three directions in one launch (the production path), on the current registers.
int launch_flux(ath_mesh *m, int order) {
  const StageArgs a = stage_args(m, regs_of(m), 2);
  return launch_flux_args(m, order, a.w, a.bcc, a.fo);
}

// One direction per launch (Hydro::CalculateFluxes direction by direction).
int launch_flux_dirs(ath_mesh *m, int order) {
  const MeshK &M = m->M;
  const int rs = solver_code(m->d);
  for (int dir = 0; dir < 3; ++dir) {
    long nface;
    if (dir == 0) nface = static_cast<long>(M.nx1 + 1) * (M.nx2 + 2) * (M.nx3 + 2);
    else if (dir == 1) nface = static_cast<long>(M.nx1 + 2) * (M.nx2 + 1) * (M.nx3 + 2);
    else nface = static_cast<long>(M.nx1 + 2) * (M.nx2 + 2) * (M.nx3 + 1);
    dim3 grid(((nface + WG - 1) / WG) * M.nblocks);
    FaceOut o = face_out(m, dir);
#define LF(D, O, R)                                                                       \
  hipLaunchKernelGGL((k_flux<D, O, R>), grid, dim3(WG), 0, m->stream, M, m->w, m->bcc, o, \
                     m->clk)
#define LF_O(D, R)                      \
  do {                                  \
    if (order == 1) LF(D, 1, R);        \
    else if (order == 2) LF(D, 2, R);   \
    else LF(D, 3, R);                   \
  } while (0)
#define LF_R(D)                         \
  do {                                  \
    if (rs == 0) LF_O(D, 0);            \
    else if (rs == 1) LF_O(D, 1);       \
    else if (rs == 2) LF_O(D, 2);       \
    else LF_O(D, 3);                    \
  } while (0)
    {
      ProfScope ps(m, KID_FLUX + 3 * dir + (order - 1));
      if (dir == 0) LF_R(0);
      else if (dir == 1) LF_R(1);
      else LF_R(2);
    }
#undef LF_R
#undef LF_O
#undef LF
    HIPCHK(hipGetLastError());
  }
  return 0;
}

int launch_corner_args(ath_mesh *m, const CornerArgs &ca) {
  const MeshK &M = m->M;
  const dim3 grid = kind_grid(m, KK_CORNER);
  ProfScope ps(m, KID_CORNER);
  const TileK T = m->tile[KK_CORNER];
  const size_t lds = corner_lds(M, T);
  if (T.tx < 0 && m->d.relativistic)
    hipLaunchKernelGGL((k_corner_march<true>), grid, dim3(WG), lds, m->stream, M, ca, T);
  else if (T.tx < 0)
    hipLaunchKernelGGL((k_corner_march<false>), grid, dim3(WG), lds, m->stream, M, ca, T);
  else if (m->d.relativistic)
    hipLaunchKernelGGL((k_corner<true>), grid, dim3(WG), 0, m->stream, M, ca, T);
  else
    hipLaunchKernelGGL((k_corner<false>), grid, dim3(WG), 0, m->stream, M, ca, T);
  HIPCHK(hipGetLastError());
  return 0;
}

int launch_corner(ath_mesh *m) {
  return launch_corner_args(m, stage_args(m, regs_of(m), 2).ca);
}

int launch_cons2prim_args(ath_mesh *m, const C2PArgs &pa) {
  const MeshK &M = m->M;
  const dim3 grid = kind_grid(m, KK_C2P);
  const TileK T = m->tile[KK_C2P];
  if (pa.dt_partial) m->ndt_used = static_cast<long>(grid.x);
  ProfScope ps(m, KID_CONS2PRIM);
  if (m->d.relativistic && M.c2p_mode == C2P_LEARNED) {
    {
      ProfScope pss(m, KID_SEED);
      hipLaunchKernelGGL(k_mesh_seed, grid, dim3(WG), 0, m->stream, M, pa.u, pa.b1, pa.b2,
                         pa.b3, m->lseed, T);
    }
    hipLaunchKernelGGL((k_cons2prim<true, true>), grid, dim3(WG), 0, m->stream, M, pa, T);
  } else if (m->d.relativistic) {
    hipLaunchKernelGGL((k_cons2prim<true, false>), grid, dim3(WG), 0, m->stream, M, pa, T);
  } else {
    hipLaunchKernelGGL((k_cons2prim<false, false>), grid, dim3(WG), 0, m->stream, M, pa, T);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// ConservedToPrimitive(u, ., b, prim_out, bcc) on the current registers.
int launch_cons2prim(ath_mesh *m, double *prim_out, bool with_dt) {
  C2PArgs pa{m->u, prim_out, m->bcc, m->b[0], m->b[1], m->b[2],
             with_dt ? m->dtpart : nullptr, m->lseed};
  return launch_cons2prim_args(m, pa);
}

int launch_new_dt(ath_mesh *m, int advance) {
  ProfScope ps(m, KID_NEW_DT);
  hipLaunchKernelGGL(k_new_dt, dim3(1), dim3(1024), 0, m->stream, m->dtpart, m->ndt_used,
                     m->clk, m->d.cfl, advance);
  HIPCHK(hipGetLastError());
  return 0;
}

int launch_update_args(ath_mesh *m, const UpdateArgs &ua, double beta, int first,
                       bool do_u = true) {
  const dim3 grid = kind_grid(m, KK_UPDATE);
  const TileK T = m->tile[KK_UPDATE];
  ProfScope ps(m, do_u ? KID_UPDATE_U : KID_UPDATE_B);
  if (do_u)
    hipLaunchKernelGGL(k_update, grid, dim3(WG), 0, m->stream, m->M, ua, beta, first, m->clk,
                       T);
  else
    hipLaunchKernelGGL(k_update_b, grid, dim3(WG), 0, m->stream, m->M, ua, beta, first,
                       m->clk, T);
  HIPCHK(hipGetLastError());
  return 0;
}

// Rows per band of k_flux_div for a mesh: about 128 owned cells per work-group, at most
// one per thread; 0 where a row of the block holds more active cells than a work-group
// has threads.
// Whether bands of up to `rows` active rows fit one work-group of k_flux_div: every
// thread row of the band (its ghost row at a mesh edge and the extra x2 row included)
// takes nx1 + 2 threads.
bool fdiv_rows_ok(const MeshK &M, int rows) {
  return rows >= 1 && rows <= M.nx2 &&
         static_cast<long>(fdiv_thread_rows(rows)) * (M.nx1 + 2) <= WG;
}

// Band height that ath_gpu_mesh_set_fused(-1) picks: as many rows as a work-group holds;
// 0 (separate flux and update kernels) where that is fewer than 2 rows, whose extra x2 row
// would add a third more Riemann solves.
int fdiv_default_rows(const MeshK &M) {
  int rows = WG / (M.nx1 + 2) - 2;
  if (rows > M.nx2) rows = M.nx2;
  return (rows >= 2 && fdiv_rows_ok(M, rows)) ? rows : 0;
}

int launch_flux_div_args(ath_mesh *m, const StageArgs &a) {
  const MeshK &M = m->M;
  const int rows = m->fdiv_rows;
  const int rs = solver_code(m->d);
  const dim3 grid(static_cast<unsigned>(fdiv_nband(M, rows) * M.nblocks));
  const size_t lds = sizeof(double) * FDIV_LDS_DOUBLES;
  const UpdateArgs &ua = a.ua;
  ProfScope ps(m, KID_FDIV + (a.order - 1));
#define LF(O, R)                                                                          \
  hipLaunchKernelGGL((k_flux_div<O, R>), grid, dim3(WG), lds, m->stream, M, a.w, a.bcc,     \
                     a.fo[0], a.fo[1], a.fo[2], ua.ubase, ua.uout, a.beta, a.first, m->clk, \
                     rows)
#define LF_O(R)                         \
  do {                                  \
    if (a.order == 1) LF(1, R);         \
    else if (a.order == 2) LF(2, R);    \
    else LF(3, R);                      \
  } while (0)
  if (rs == 0) LF_O(0);
  else if (rs == 1) LF_O(1);
  else if (rs == 2) LF_O(2);
  else LF_O(3);
#undef LF_O
#undef LF
  HIPCHK(hipGetLastError());
  return 0;
}

int ghost_grid(const MeshK &M) {
  return static_cast<int>((ghost_items(M) + WG - 1) / WG) * M.nblocks;
}

int launch_ghost_args(ath_mesh *m, const GhostArgs &ga) {
  ProfScope ps(m, KID_GHOST);
  hipLaunchKernelGGL(k_ghost, dim3(ghost_grid(m->M)), dim3(WG), 0, m->stream, m->M, ga);
  HIPCHK(hipGetLastError());
  return 0;
}

// Outflow faces after ConservedToPrimitive (nothing to do on a periodic mesh).
int launch_phys_bnd(ath_mesh *m, const C2PArgs &pa) {
  const MeshK &M = m->M;
  if (!M.outflow) return 0;
  ProfScope ps(m, KID_PHYS_BND);
  const int grid =
      static_cast<int>((ghost_count(M.nc1, M.nc2, M.nc3, M.ng) + WG - 1) / WG) * M.nblocks;
  if (m->d.relativistic)
    hipLaunchKernelGGL(k_phys_bnd<true>, dim3(grid), dim3(WG), 0, m->stream, M,
                       PhysArgs{pa.u, pa.prim, pa.bcc});
  else
    hipLaunchKernelGGL(k_phys_bnd<false>, dim3(grid), dim3(WG), 0, m->stream, M,
                       PhysArgs{pa.u, pa.prim, pa.bcc});
  HIPCHK(hipGetLastError());
  return 0;
}

void swap_stage_regs(Regs &r) {   // the two swaps of one stage
  std::swap(r.u, r.u1);
  for (int c = 0; c < 3; ++c) std::swap(r.b[c], r.bb[c]);
  std::swap(r.w, r.w1);
}

void set_regs(ath_mesh *m, const Regs &r) {
  m->u = r.u; m->u1 = r.u1; m->w = r.w; m->w1 = r.w1;
  for (int c = 0; c < 3; ++c) { m->b[c] = r.b[c]; m->bb[c] = r.bb[c]; }
}

SmrArgs smr_args(const ath_mesh *m, int prog, double *u, double *const *b, double *w) {
  SmrArgs a{};
  a.steps = m->smr_steps[prog];
  a.chunk_step = m->smr_chunk_step[prog];
  a.chunk_first = m->smr_chunk_first[prog];
  a.u = u;
  a.w = w;
  a.bcc = m->bcc;
  for (int c = 0; c < 3; ++c) {
    a.b[c] = b ? b[c] : nullptr;
    a.cb[c] = m->cb[c];
    a.flux[c] = m->flux[c];
    a.e[c] = m->e[c];
    a.es[c] = m->esnap[c];
  }
  a.cu = m->cu; a.cw = m->cw; a.cbcc = m->cbcc;
  a.fco = m->fco; a.cco = m->cco;
  a.fcl = m->fcl; a.ccl = m->ccl;
  return a;
}

int launch_smr(ath_mesh *m, int prog, const SmrArgs &a) {
  static const int kid[3] = {KID_SMR_FLUX, KID_SMR_EMF, KID_SMR_BND};
  ProfScope ps(m, kid[prog]);
  const std::vector<int> &ph = m->smr_phases[prog];
  for (std::size_t k = 0; k + 1 < ph.size(); ++k) {
    const int nch = ph[k + 1] - ph[k];
    if (nch <= 0) continue;
    hipLaunchKernelGGL(k_smr, dim3(nch), dim3(smrplan::CHUNK_THREADS), 0, m->stream, m->M, a,
                       ph[k]);
    HIPCHK(hipGetLastError());
  }
  return 0;
}

// One VL2 stage on a statically refined mesh: fluxes, their correction on coarse/fine
// faces, corner EMFs and their correction, the update, the boundary program (exchange,
// restriction, prolongation), and the primitive recovery over every cell.
int stage_smr(ath_mesh *m, int stg) {
  Regs r = regs_of(m);
  const StageArgs a = stage_args(m, r, stg);
  if (launch_flux_args(m, a.order, a.w, a.bcc, a.fo)) return 1;
  if (launch_smr(m, 0, smr_args(m, 0, nullptr, nullptr, nullptr))) return 1;
  if (launch_corner_args(m, a.ca)) return 1;
  const long ne[3] = {m->M.ne1, m->M.ne2, m->M.ne3};
  for (int c = 0; c < 3; ++c)
    HIPCHK(hipMemcpyAsync(m->esnap[c], m->e[c], sizeof(double) * ne[c] * m->M.nblocks,
                          hipMemcpyDeviceToDevice, m->stream));
  if (launch_smr(m, 1, smr_args(m, 1, nullptr, nullptr, nullptr))) return 1;
  if (launch_update_args(m, a.ua, a.beta, a.first, true)) return 1;
  // boundary program on the updated state; prolongated primitives go to the register the
  // recovery treats as the previous primitives
  if (launch_smr(m, 2, smr_args(m, 2, a.ga.u, a.ga.b, r.w))) return 1;
  if (launch_cons2prim_args(m, a.pa)) return 1;
  swap_stage_regs(r);
  set_regs(m, r);
  m->nstage += 1;
  return 0;
}

int stage(ath_mesh *m, int stg) {
  if (m->smr) return stage_smr(m, stg);
  Regs r = regs_of(m);
  const StageArgs a = stage_args(m, r, stg);
  const bool fused = m->fdiv_rows > 0;
  if (fused) {
    if (launch_flux_div_args(m, a)) return 1;
  } else if (launch_flux_args(m, a.order, a.w, a.bcc, a.fo)) {
    return 1;
  }
  if (launch_corner_args(m, a.ca)) return 1;
  if (launch_update_args(m, a.ua, a.beta, a.first, !fused)) return 1;
  if (launch_ghost_args(m, a.ga)) return 1;
  if (launch_cons2prim_args(m, a.pa)) return 1;
  if (launch_phys_bnd(m, a.pa)) return 1;
  swap_stage_regs(r);
  set_regs(m, r);
  m->nstage += 1;
  return 0;
}

// Everything a captured cycle bakes in: the register pointers at the start of a cycle
// (a cycle's two stages swap every pair twice, so they are the same at the start of the
// next), the fused band height, the tile shapes, the recovery mode and its network.
std::vector<const void *> cycle_key(const ath_mesh *m) {
  std::vector<const void *> k = {m->u, m->u1, m->w, m->w1, m->b[0], m->b[1], m->b[2],
                                 m->bb[0], m->bb[1], m->bb[2], m->lseed, m->M.mlp.P,
                                 m->M.c2p_stats};
  k.push_back(reinterpret_cast<const void *>(static_cast<intptr_t>(m->fdiv_rows)));
  k.push_back(reinterpret_cast<const void *>(static_cast<intptr_t>(m->M.c2p_mode)));
  for (int q = 0; q < KK_N; ++q) {
    const TileK &t = m->tile[q];
    k.push_back(reinterpret_cast<const void *>(
        static_cast<intptr_t>(((t.tx + 2) << 20) | (t.ty << 10) | t.tz)));
  }
  return k;
}

void drop_cycle_graph(ath_mesh *m) {
  if (m->gexec) (void)hipGraphExecDestroy(m->gexec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  m->gexec = nullptr;
  m->graph = nullptr;
  m->graph_key.clear();
}

// One VL2 cycle (two stages and the new dt), launched eagerly.
int cycle_eager(ath_mesh *m) {
  if (stage(m, 1)) return 1;
  if (stage(m, 2)) return 1;
  return launch_new_dt(m, 1);
}

// One VL2 cycle replayed from the captured graph, capturing it first when the mesh's
// configuration differs from the captured one.  The host register bookkeeping advances
// as the eager cycle would: two stages, registers back where they were.
int cycle_graph(ath_mesh *m) {
  const std::vector<const void *> key = cycle_key(m);
  if (!m->gexec || key != m->graph_key) {
    drop_cycle_graph(m);
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
    const long nst = m->nstage;
    const int rc = cycle_eager(m);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(m->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return 1;
    }
    if (ec != hipSuccess) return fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    m->graph = g;
    HIPCHK(hipGraphInstantiate(&m->gexec, m->graph, nullptr, nullptr, 0));
    m->graph_key = key;
    // capturing recorded the cycle without running it: undo the bookkeeping it advanced
    if (cycle_key(m) != key) return fail("a captured cycle must leave the registers in place");
    m->nstage = nst;
  }
  HIPCHK(hipGraphLaunch(m->gexec, m->stream));
  m->nstage += 2;
  return 0;
}

int block_count(ath_mesh *m, int what, long *n, double ***pool) {
  const MeshK &M = m->M;
  switch (what) {
    case ATH_U: *n = NHYDRO * M.ncell; *pool = &m->u; return 0;
    case ATH_W: *n = NHYDRO * M.ncell; *pool = &m->w; return 0;
    case ATH_W1: *n = NHYDRO * M.ncell; *pool = &m->w1; return 0;
    case ATH_BCC: *n = 3 * M.ncell; *pool = &m->bcc; return 0;
    case ATH_B1: *n = M.nf1; *pool = &m->b[0]; return 0;
    case ATH_B2: *n = M.nf2; *pool = &m->b[1]; return 0;
    case ATH_B3: *n = M.nf3; *pool = &m->b[2]; return 0;
    case ATH_FLUX1: *n = NHYDRO * M.nf1; *pool = &m->flux[0]; return 0;
    case ATH_FLUX2: *n = NHYDRO * M.nf2; *pool = &m->flux[1]; return 0;
    case ATH_FLUX3: *n = NHYDRO * M.nf3; *pool = &m->flux[2]; return 0;
    case ATH_E1: *n = M.ne1; *pool = &m->e[0]; return 0;
    case ATH_E2: *n = M.ne2; *pool = &m->e[1]; return 0;
    case ATH_E3: *n = M.ne3; *pool = &m->e[2]; return 0;
    case ATH_WGHT1: *n = M.nf1; *pool = &m->wx[0]; return 0;
    case ATH_WGHT2: *n = M.nf2; *pool = &m->wx[1]; return 0;
    case ATH_WGHT3: *n = M.nf3; *pool = &m->wx[2]; return 0;
    default: return fail("unknown register code " + std::to_string(what));
  }
}

}  // namespace

//========================================================================================
// Ensemble host side

struct ath_ensemble {
  std::vector<ath_mesh *> mem;
  int xorder = 2, rs = 0, sr = 0;
  int c2p_mode = 0;           // primitive-recovery mode shared by the members
  bool learned = false;       // c2p_mode == C2P_LEARNED: seed kernel before the recovery
  hipStream_t stream = nullptr;
  EnsMember *tab_d[2] = {nullptr, nullptr};   // stage-1 and stage-2 member tables
  int2 *map_d[KK_N] = {nullptr, nullptr, nullptr, nullptr};
  int nwork[KK_N] = {0, 0, 0, 0};
  int2 *map_ghost_d = nullptr;                // work items of the ghost exchange
  int nwork_ghost = 0;
  int tile_idx[KK_N] = {0, 6, 0, 0};          // index into TILES
  int persist[KK_N] = {0, 0, 0, 0};           // 0: one work item per work-group; else grid
  std::vector<long> parity;                   // members' stage count parity at creation
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~ath_ensemble() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto *p : tab_d) if (p) (void)hipFree(p);
    for (auto *p : map_d) if (p) (void)hipFree(p);
    if (map_ghost_d) (void)hipFree(map_ghost_d);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

void ens_drop_graph(ath_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

// (Re)builds the member tables and work maps for the current tile choices.
int ens_build(ath_ensemble *e) {
  HIPCHK(hipStreamSynchronize(e->stream));
  ens_drop_graph(e);
  const int n = static_cast<int>(e->mem.size());
  std::vector<EnsMember> tab[2];
  std::vector<int2> map[KK_N], mapg;
  for (int q = 0; q < n; ++q) {
    ath_mesh *m = e->mem[q];
    Regs r = regs_of(m);
    for (int stg = 1; stg <= 2; ++stg) {
      const StageArgs a = stage_args(m, r, stg);
      EnsMember em;
      em.M = m->M;
      em.w = a.w;
      em.bcc = a.bcc;
      for (int d = 0; d < 3; ++d) em.fo[d] = a.fo[d];
      em.ca = a.ca;
      em.ua = a.ua;
      em.ga = a.ga;
      em.nwg_ghost = ghost_grid(m->M);
      em.pa = a.pa;
      em.clk = m->clk;
      for (int k = 0; k < KK_N; ++k) {
        em.tile[k] = TILES[e->tile_idx[k]];
        em.nwg[k] = static_cast<int>(kind_chunks(m->M, k, em.tile[k]) * m->M.nblocks);
      }
      if (em.nwg[KK_C2P] > m->ndtpart) return fail("dt partial buffer too small");
      tab[stg - 1].push_back(em);
      swap_stage_regs(r);
    }
    for (int k = 0; k < KK_N; ++k) {
      const int nw = tab[0].back().nwg[k];
      for (int w = 0; w < nw; ++w) map[k].push_back(make_int2(q, w));
    }
    for (int w = 0; w < tab[0].back().nwg_ghost; ++w) mapg.push_back(make_int2(q, w));
  }
  for (int t = 0; t < 2; ++t) {
    if (e->tab_d[t]) HIPCHK(hipFree(e->tab_d[t]));
    e->tab_d[t] = nullptr;
    HIPCHK(hipMalloc(reinterpret_cast<void **>(&e->tab_d[t]), sizeof(EnsMember) * n));
    HIPCHK(hipMemcpy(e->tab_d[t], tab[t].data(), sizeof(EnsMember) * n,
                     hipMemcpyHostToDevice));
  }
  for (int k = 0; k < KK_N; ++k) {
    if (e->map_d[k]) HIPCHK(hipFree(e->map_d[k]));
    e->map_d[k] = nullptr;
    e->nwork[k] = static_cast<int>(map[k].size());
    HIPCHK(hipMalloc(reinterpret_cast<void **>(&e->map_d[k]), sizeof(int2) * map[k].size()));
    HIPCHK(hipMemcpy(e->map_d[k], map[k].data(), sizeof(int2) * map[k].size(),
                     hipMemcpyHostToDevice));
  }
  if (e->map_ghost_d) HIPCHK(hipFree(e->map_ghost_d));
  e->map_ghost_d = nullptr;
  e->nwork_ghost = static_cast<int>(mapg.size());
  HIPCHK(hipMalloc(reinterpret_cast<void **>(&e->map_ghost_d), sizeof(int2) * mapg.size()));
  HIPCHK(hipMemcpy(e->map_ghost_d, mapg.data(), sizeof(int2) * mapg.size(),
                   hipMemcpyHostToDevice));
  return 0;
}

dim3 ens_grid(const ath_ensemble *e, int kind) {
  const int p = e->persist[kind];
  return dim3(static_cast<unsigned>((p > 0 && p < e->nwork[kind]) ? p : e->nwork[kind]));
}

int ens_launch_kind(ath_ensemble *e, int kind, int stg) {
  const EnsMember *T = e->tab_d[stg - 1];
  const dim3 g = ens_grid(e, kind);
  const int2 *mp = e->map_d[kind];
  const int nw = e->nwork[kind];
  if (kind == KK_FLUX) {
    const int order = (stg == 1) ? 1 : e->xorder;
#define LE(O, R) hipLaunchKernelGGL((k_flux3_ens<O, R>), g, dim3(WG), 0, e->stream, T, mp, nw)
#define LE_O(R)                         \
  do {                                  \
    if (order == 1) LE(1, R);           \
    else if (order == 2) LE(2, R);      \
    else LE(3, R);                      \
  } while (0)
    if (e->rs == 0) LE_O(0);
    else if (e->rs == 1) LE_O(1);
    else if (e->rs == 2) LE_O(2);
    else LE_O(3);
#undef LE_O
#undef LE
  } else if (kind == KK_CORNER) {
    const TileK tk = TILES[e->tile_idx[kind]];
    size_t lds = 0;
    for (auto *m : e->mem) lds = std::max(lds, corner_lds(m->M, tk));
    if (tk.tx < 0 && e->sr)
      hipLaunchKernelGGL(k_corner_march_ens<true>, g, dim3(WG), lds, e->stream, T, mp, nw);
    else if (tk.tx < 0)
      hipLaunchKernelGGL(k_corner_march_ens<false>, g, dim3(WG), lds, e->stream, T, mp, nw);
    else if (e->sr)
      hipLaunchKernelGGL(k_corner_ens<true>, g, dim3(WG), 0, e->stream, T, mp, nw);
    else
      hipLaunchKernelGGL(k_corner_ens<false>, g, dim3(WG), 0, e->stream, T, mp, nw);
  } else if (kind == KK_UPDATE) {
    hipLaunchKernelGGL(k_update_ens, g, dim3(WG), 0, e->stream, T, mp, nw,
                       (stg == 1) ? 0.5 : 1.0, (stg == 1) ? 1 : 0);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_ghost_ens, dim3(static_cast<unsigned>(e->nwork_ghost)), dim3(WG), 0,
                       e->stream, T, e->map_ghost_d, e->nwork_ghost);
  } else if (e->learned) {
    hipLaunchKernelGGL(k_seed_ens, g, dim3(WG), 0, e->stream, T, mp, nw);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL((k_c2p_ens<true, true>), g, dim3(WG), 0, e->stream, T, mp, nw);
  } else {
    if (e->sr) hipLaunchKernelGGL((k_c2p_ens<true, false>), g, dim3(WG), 0, e->stream, T, mp, nw);
    else hipLaunchKernelGGL((k_c2p_ens<false, false>), g, dim3(WG), 0, e->stream, T, mp, nw);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// One VL2 cycle of every member: 5 launches per stage (the update followed by the ghost
// exchange) and one dt launch.
int ens_record_cycle(ath_ensemble *e) {
  for (int stg = 1; stg <= 2; ++stg)
    for (int k = 0; k < KK_N; ++k)
      if (ens_launch_kind(e, k, stg)) return 1;
  hipLaunchKernelGGL(k_new_dt_ens, dim3(static_cast<unsigned>(e->mem.size())), dim3(1024), 0,
                     e->stream, e->tab_d[1], 1);
  HIPCHK(hipGetLastError());
  return 0;
}

int ens_check(ath_ensemble *e) {
  for (size_t q = 0; q < e->mem.size(); ++q) {
    if ((e->mem[q]->nstage & 1) != e->parity[q])
      return fail("a member ran a single VL2 stage after the ensemble was created; its "
                  "registers no longer match the ensemble's tables");
    HIPCHK(hipStreamSynchronize(e->mem[q]->stream));
  }
  return 0;
}

}  // namespace

//========================================================================================
// One MeshBlock over caller-owned arrays (ath_gpu_block_*): its geometry, coordinate
// table and equation of state on the device, as a MeshBlock owns its Coordinates and
// EquationOfState.

struct ath_block {
  MeshK M;
  int rs = 0;              // solver_code
  int relativistic = 0;
  int *nbr_d = nullptr;
  double *coord_d = nullptr;
  double *scratch_d = nullptr;   // this device's copy of `scratch` (ath_gpu_block_new_dt_noalloc)
};

// Per-work-group minima of ath_gpu_block_new_dt_noalloc, one double per 256 active cells:
// at most kDtScratchMax partials (16,777,216 active cells per MeshBlock).  Every device
// holds its own copy until the process ends; host code never names the array directly
// but uses the address hipGetSymbolAddress returns for the current device.
constexpr int kDtScratchMax = 65536;
static __device__ __align__(256) double scratch[kDtScratchMax];

namespace {

int blk_check_stream(hipStream_t st) {
  HIPCHK(hipGetLastError());
  (void)st;
  return 0;
}

bool in_box(int lo, int hi, int n) { return lo >= 0 && hi < n && lo <= hi; }

// Device address of `scratch` on the current device, looked up once per device.
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_addr;

int scratch_address(double **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (static_cast<int>(g_scratch_addr.size()) <= dev) g_scratch_addr.resize(dev + 1, nullptr);
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_addr[dev];
  return 0;
}

// Threads of a pencil launch over il .. iu.
dim3 pencil_grid(int il, int iu) { return dim3(static_cast<unsigned>((iu - il + WG) / WG)); }

}  // namespace

//========================================================================================
// C ABI

extern "C" {

const char *ath_gpu_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace {

// Coordinate table of one grid in the refinement operators' layout (x1f, x1v, dx1f, then
// x2, x3) for a block at logical location lx, `shift` levels above the root grid; coarse:
// the coarse buffer's grid (nx/2 cells, faces at every second fine face; the coarse flag
// of the Coordinates constructor, coordinates.cpp:30-146).
void level_table(const ath_mesh_desc &d, const long *lx, int shift, bool coarse,
                 std::vector<double> &out) {
  out.clear();
  for (int q = 0; q < 3; ++q) {
    const int nxb = d.bnx[q], ng = d.ng, n = coarse ? nxb / 2 : nxb, nc = n + 2 * ng;
    const long nbq = static_cast<long>(d.nx[q] / d.bnx[q]) << shift;
    const long nroot = static_cast<long>(d.nx[q]) << shift;
    const long l = lx[q];
    const double bmin = (l == 0) ? d.xmin[q] : uniform_gen(mesh_gen_x(l, nbq), d.xmin[q], d.xmax[q]);
    const double bmax = (l == nbq - 1) ? d.xmax[q]
                                       : uniform_gen(mesh_gen_x(l + 1, nbq), d.xmin[q], d.xmax[q]);
    const int il = ng, iu = ng + n - 1;
    const double dx = (bmax - bmin) / (iu - il + 1);
    std::vector<double> xf(nc + 1), xv(nc), dxf(nc, dx);
    for (int i = il - ng; i <= iu + ng + 1; ++i) {
      const long noff = coarse ? (i - il) * 2 + l * nxb : i - il + l * nxb;
      xf[i] = uniform_gen(mesh_gen_x(noff, nroot), d.xmin[q], d.xmax[q]);
    }
    xf[il] = bmin;
    xf[iu + 1] = bmax;
    for (int i = 0; i < nc; ++i) xv[i] = 0.5 * (xf[i + 1] + xf[i]);
    out.insert(out.end(), xf.begin(), xf.end());
    out.insert(out.end(), xv.begin(), xv.end());
    out.insert(out.end(), dxf.begin(), dxf.end());
  }
}

// Step programs, coarse buffers, coordinate tables and the EMF snapshot of a refined mesh.
int smr_setup(ath_mesh *m, const std::vector<smrplan::BlockInfo> &sb, int root_level) {
  const MeshK &M = m->M;
  const int nb = M.nblocks;
  const smrplan::Shape S = smrplan::make_shape(m->d.bnx, m->d.ng);
  m->smr_blocks = sb;
  smrplan::Program prog[3];
  try {
    for (int q = 0; q < 3; ++q) prog[q].offs.push_back(0);
    for (int g = 0; g < nb; ++g) {
      smrplan::flux_program(S, sb, g, prog[0]);
      prog[0].offs.push_back(static_cast<int>(prog[0].steps.size()));
      smrplan::emf_program(S, sb, g, prog[1]);
      prog[1].offs.push_back(static_cast<int>(prog[1].steps.size()));
      smrplan::boundary_program(S, sb, g, prog[2]);
      prog[2].offs.push_back(static_cast<int>(prog[2].steps.size()));
    }
  } catch (const std::exception &e) {
    return fail(e.what());
  }
  std::vector<double> fco, cco, t;
  for (int g = 0; g < nb; ++g) {
    const long lx[3] = {static_cast<long>(sb[g].lx[0]), static_cast<long>(sb[g].lx[1]),
                        static_cast<long>(sb[g].lx[2])};
    level_table(m->d, lx, sb[g].level - root_level, false, t);
    if (g == 0) m->fcl = static_cast<long>(t.size());
    fco.insert(fco.end(), t.begin(), t.end());
    level_table(m->d, lx, sb[g].level - root_level, true, t);
    if (g == 0) m->ccl = static_cast<long>(t.size());
    cco.insert(cco.end(), t.begin(), t.end());
  }
  const long cc = static_cast<long>(S.ncc[0]) * S.ncc[1] * S.ncc[2];
  const long cf[3] = {static_cast<long>(S.ncc[2]) * S.ncc[1] * (S.ncc[0] + 1),
                      static_cast<long>(S.ncc[2]) * (S.ncc[1] + 1) * S.ncc[0],
                      static_cast<long>(S.ncc[2] + 1) * S.ncc[1] * S.ncc[0]};
  const long ne[3] = {M.ne1, M.ne2, M.ne3};
  int rc = 0;
  rc |= m->alloc(&m->cu, NHYDRO * cc * nb);
  rc |= m->alloc(&m->cw, NHYDRO * cc * nb);
  rc |= m->alloc(&m->cbcc, 3 * cc * nb);
  for (int c = 0; c < 3 && !rc; ++c) {
    rc |= m->alloc(&m->cb[c], cf[c] * nb);
    rc |= m->alloc(&m->esnap[c], ne[c] * nb);
  }
  if (!rc) rc |= m->alloc(&m->fco, static_cast<long>(fco.size()));
  if (!rc) rc |= m->alloc(&m->cco, static_cast<long>(cco.size()));
  for (int q = 0; q < 3 && !rc; ++q) {
    const smrplan::Phased ph = smrplan::to_phases(prog[q], nb);
    const size_t nst = std::max<size_t>(ph.steps.size(), 1);
    const size_t nch = std::max<size_t>(ph.chunk_step.size(), 1);
    void *ps = nullptr, *pc = nullptr, *pf = nullptr;
    if (hipMalloc(&ps, sizeof(smrplan::Step) * nst) != hipSuccess ||
        hipMalloc(&pc, sizeof(int) * nch) != hipSuccess ||
        hipMalloc(&pf, sizeof(int) * nch) != hipSuccess) {
      if (ps) (void)hipFree(ps);
      if (pc) (void)hipFree(pc);
      if (pf) (void)hipFree(pf);
      return fail("hipMalloc failed for SMR programs");
    }
    m->allocs.push_back(ps);
    m->allocs.push_back(pc);
    m->allocs.push_back(pf);
    m->smr_steps[q] = static_cast<smrplan::Step *>(ps);
    m->smr_chunk_step[q] = static_cast<int *>(pc);
    m->smr_chunk_first[q] = static_cast<int *>(pf);
    m->smr_phases[q] = ph.phase_chunks;
    if (!ph.steps.empty()) {
      HIPCHK(hipMemcpyAsync(ps, ph.steps.data(), sizeof(smrplan::Step) * ph.steps.size(),
                            hipMemcpyHostToDevice, m->stream));
      HIPCHK(hipMemcpyAsync(pc, ph.chunk_step.data(), sizeof(int) * ph.chunk_step.size(),
                            hipMemcpyHostToDevice, m->stream));
      HIPCHK(hipMemcpyAsync(pf, ph.chunk_first.data(), sizeof(int) * ph.chunk_first.size(),
                            hipMemcpyHostToDevice, m->stream));
    }
    // the uploads read the host vectors: finish them before ph goes out of scope
    HIPCHK(hipStreamSynchronize(m->stream));
  }
  if (rc) return 1;
  HIPCHK(hipMemcpyAsync(m->fco, fco.data(), sizeof(double) * fco.size(), hipMemcpyHostToDevice,
                        m->stream));
  HIPCHK(hipMemcpyAsync(m->cco, cco.data(), sizeof(double) * cco.size(), hipMemcpyHostToDevice,
                        m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

// The MeshBlock list of a refined mesh: every location inside the root grid at its level,
// none listed twice or inside another, and together they cover the root grid exactly.
int smr_blocks_of(const ath_mesh_desc &d, const long long *locs, int nblocks, int root_level,
                  std::vector<smrplan::BlockInfo> &sb) {
  if (nblocks < 1 || !locs) return fail("a refined mesh needs at least one MeshBlock");
  int nrb[3], maxlev = root_level;
  for (int q = 0; q < 3; ++q) nrb[q] = d.nx[q] / d.bnx[q];
  sb.resize(nblocks);
  std::map<std::array<long long, 4>, int> seen;
  for (int g = 0; g < nblocks; ++g) {
    smrplan::BlockInfo &b = sb[g];
    b.level = static_cast<int>(locs[4 * g]);
    if (b.level < root_level || b.level > root_level + 20)
      return fail("MeshBlock level outside [root_level, root_level + 20]");
    maxlev = std::max(maxlev, b.level);
    for (int q = 0; q < 3; ++q) {
      b.lx[q] = locs[4 * g + 1 + q];
      if (b.lx[q] < 0 || b.lx[q] >= (static_cast<long long>(nrb[q]) << (b.level - root_level)))
        return fail("MeshBlock location outside the mesh");
    }
    if (!seen.emplace(std::array<long long, 4>{b.level, b.lx[0], b.lx[1], b.lx[2]}, g).second)
      return fail("a MeshBlock location is listed twice");
  }
  long double vol = 0.0L;
  for (const auto &b : sb) {
    vol += std::ldexp(1.0L, 3 * (maxlev - b.level));
    for (int up = 1; up <= b.level - root_level; ++up)
      if (seen.count({b.level - up, b.lx[0] >> up, b.lx[1] >> up, b.lx[2] >> up}))
        return fail("a MeshBlock lies inside another listed MeshBlock");
  }
  const long double want = std::ldexp(static_cast<long double>(nrb[0]) * nrb[1] * nrb[2],
                                      3 * (maxlev - root_level));
  if (vol != want) return fail("the MeshBlocks do not cover the mesh");
  try {
    smrplan::search_neighbors(nrb, root_level, sb);
  } catch (const std::exception &e) {
    return fail(e.what());
  }
  return 0;
}

// The device mirror of a uniform mesh (sb null) or of a statically refined one (sb: the
// MeshBlock list with its neighbour lists, root_level the logical level of the root grid).
int mesh_build(const ath_mesh_desc *desc, const std::vector<smrplan::BlockInfo> *sb,
               int root_level, ath_mesh_t *out) {
  if (!desc || !out) return fail("null argument");
  *out = nullptr;
  const ath_mesh_desc &d = *desc;
  for (int q = 0; q < 3; ++q) {
    if (d.bnx[q] <= 0 || d.nx[q] <= 0 || d.nx[q] % d.bnx[q])
      return fail("mesh nx must be a positive multiple of meshblock nx in every direction");
    if (d.bnx[q] < 2) return fail("meshblocks must have at least 2 cells per direction (3-D only)");
    if (!(d.xmax[q] > d.xmin[q])) return fail("xmax must exceed xmin");
  }
  if (d.ng < 2) return fail("nghost must be >= 2");
  if (d.xorder < 1 || d.xorder > 3) return fail("xorder must be 1, 2 or 3");
  if (d.xorder == 3 && d.ng < 3)
    return fail("xorder=3 (PPM) reconstruction selected, but nghost < 3");
  if (d.rsolver < 0 || d.rsolver > 1) return fail("rsolver must be 0 (HLLD) or 1 (HLLE)");
  if (!(d.gamma > 1.0)) return fail("gamma must exceed 1");
  if (!(d.cfl > 0.0) || d.cfl > 0.5) return fail("cfl_number must be in (0, 0.5] for 3-D VL2");
  for (int q = 0; q < 3; ++q)
    if (d.bnx[q] < d.ng) return fail("meshblock nx must be >= nghost");
  for (int q = 0; q < 3; ++q) {
    if (d.bc[2 * q] < 0 || d.bc[2 * q] > 1 || d.bc[2 * q + 1] < 0 || d.bc[2 * q + 1] > 1)
      return fail("boundary flags must be 0 (periodic) or 1 (outflow)");
    if ((d.bc[2 * q] == 0) != (d.bc[2 * q + 1] == 0))
      return fail("a periodic boundary needs periodic on the opposite face too");
  }

  ath_mesh *m = new ath_mesh;
  m->d = d;
  MeshK &M = m->M;
  M.nx1 = d.bnx[0]; M.nx2 = d.bnx[1]; M.nx3 = d.bnx[2]; M.ng = d.ng;
  M.nc1 = M.nx1 + 2 * M.ng; M.nc2 = M.nx2 + 2 * M.ng; M.nc3 = M.nx3 + 2 * M.ng;
  M.is = M.ng; M.ie = M.ng + M.nx1 - 1;
  M.js = M.ng; M.je = M.ng + M.nx2 - 1;
  M.ks = M.ng; M.ke = M.ng + M.nx3 - 1;
  for (int q = 0; q < 3; ++q) m->nbx[q] = d.nx[q] / d.bnx[q];
  M.nblocks = sb ? static_cast<int>(sb->size()) : m->nbx[0] * m->nbx[1] * m->nbx[2];
  M.ncell = static_cast<long>(M.nc1) * M.nc2 * M.nc3;
  M.nf1 = static_cast<long>(M.nc3) * M.nc2 * (M.nc1 + 1);
  M.nf2 = static_cast<long>(M.nc3) * (M.nc2 + 1) * M.nc1;
  M.nf3 = static_cast<long>(M.nc3 + 1) * M.nc2 * M.nc1;
  M.ne1 = static_cast<long>(M.nc3 + 1) * (M.nc2 + 1) * M.nc1;
  M.ne2 = static_cast<long>(M.nc3 + 1) * M.nc2 * (M.nc1 + 1);
  M.ne3 = static_cast<long>(M.nc3) * (M.nc2 + 1) * (M.nc1 + 1);
  M.coff[0] = 0;
  M.coff[1] = 3 * M.nc1;
  M.coff[2] = 3 * (M.nc1 + M.nc2);
  M.cstride = 3 * (M.nc1 + M.nc2 + M.nc3);
  const double fmin = static_cast<double>(std::numeric_limits<float>::min());
  M.eos.gamma = d.gamma;
  M.eos.igm1 = 1.0 / (d.gamma - 1.0);
  M.eos.dfloor = (d.dfloor < 0.0) ? std::sqrt(1024 * fmin) : d.dfloor;
  M.eos.pfloor = (d.pfloor < 0.0) ? std::sqrt(1024 * fmin) : d.pfloor;
  M.eos.sigma_max = (d.sigma_max < 0.0) ? 0.0 : d.sigma_max;
  M.eos.beta_min = (d.beta_min < 0.0) ? 0.0 : d.beta_min;
  M.eos.gamma_max = (d.gamma_max < 0.0) ? 1000.0 : d.gamma_max;
  M.cfl = d.cfl;
  {
    const char *env = std::getenv("ATH_XCD_REMAP");
    M.xcd_remap = env ? (std::atoi(env) & 31) : 0;
  }
  M.c2p_mode = C2P_REFERENCE;
  M.mlp = MlpK{};
  M.c2p_stats = nullptr;
  M.outflow = 0;
  for (int f = 0; f < 6; ++f) M.outflow |= d.bc[f];
  M.smr = sb ? 1 : 0;
  m->smr = sb != nullptr;
  m->fdiv_rows = 0;   // separate flux and update kernels: measured faster (see the header)

  // neighbour table and coordinates (coordinates.cpp:129-146, cartesian.cpp:23-72)
  const int nb = M.nblocks;
  std::vector<int> nbrh(static_cast<size_t>(nb) * 27);
  std::vector<double> coordh(static_cast<size_t>(nb) * M.cstride);
  for (int g = 0; g < nb; ++g) {
    long lx[3] = {g % m->nbx[0], (g / m->nbx[0]) % m->nbx[1], g / (m->nbx[0] * m->nbx[1])};
    int shift = 0;   // levels above the root grid
    if (sb) {
      for (int q = 0; q < 3; ++q) lx[q] = static_cast<long>((*sb)[g].lx[q]);
      shift = (*sb)[g].level - root_level;
    }
    for (int o3 = -1; o3 <= 1; ++o3)
      for (int o2 = -1; o2 <= 1; ++o2)
        for (int o1 = -1; o1 <= 1; ++o1) {
          const int o[3] = {o1, o2, o3};
          bool phys = false;   // across an outflow face of the mesh: no neighbour
          for (int q = 0; q < 3; ++q) {
            const int l = lx[q] + o[q];
            if ((l < 0 && d.bc[2 * q]) || (l >= m->nbx[q] && d.bc[2 * q + 1])) phys = true;
          }
          const long l1 = (lx[0] + o1 + m->nbx[0]) % m->nbx[0];
          const long l2 = (lx[1] + o2 + m->nbx[1]) % m->nbx[1];
          const long l3 = (lx[2] + o3 + m->nbx[2]) % m->nbx[2];
          // refined meshes exchange through their step programs: every side has a
          // neighbour (periodic), recorded as the block itself
          nbrh[g * 27 + (o3 + 1) * 9 + (o2 + 1) * 3 + (o1 + 1)] =
              sb ? g : (phys ? -1 : static_cast<int>(l1 + m->nbx[0] * (l2 + m->nbx[1] * l3)));
        }
    for (int q = 0; q < 3; ++q) {
      const int nx = d.bnx[q], ng = d.ng, nc = nx + 2 * ng;
      const long nbq = static_cast<long>(m->nbx[q]) << shift;   // blocks along q, this level
      const long nroot = static_cast<long>(d.nx[q]) << shift;    // cells along q, this level
      const long l = lx[q];
      double bmin = (l == 0) ? d.xmin[q] : uniform_gen(mesh_gen_x(l, nbq), d.xmin[q], d.xmax[q]);
      double bmax = (l == nbq - 1) ? d.xmax[q]
                                   : uniform_gen(mesh_gen_x(l + 1, nbq), d.xmin[q], d.xmax[q]);
      const int il = ng, iu = ng + nx - 1;
      const double dx = (bmax - bmin) / (iu - il + 1);
      std::vector<double> xf(nc + 1), xv(nc);
      for (int i = il - ng; i <= iu + ng + 1; ++i)
        xf[i] = uniform_gen(mesh_gen_x(i - il + l * nx, nroot), d.xmin[q], d.xmax[q]);
      xf[il] = bmin;
      xf[iu + 1] = bmax;
      for (int i = 0; i < nc; ++i) xv[i] = 0.5 * (xf[i + 1] + xf[i]);
      double *tab = coordh.data() + static_cast<size_t>(g) * M.cstride + M.coff[q];
      for (int i = 0; i < nc; ++i) {
        tab[i] = dx;                                   // dxNf
        tab[nc + i] = (xf[i + 1] - xv[i]) / dx;        // PLM right-face factor
        tab[2 * nc + i] = (xv[i] - xf[i]) / dx;        // PLM left-face factor
      }
    }
  }

  int rc = 0;
  auto cleanup = [&]() {
    for (void *p : m->allocs) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    delete m;
    return fail("hipStreamCreate failed");
  }
  const long nc5 = NHYDRO * M.ncell * nb;
  rc |= m->alloc(&m->u, nc5);
  rc |= m->alloc(&m->u1, nc5);
  rc |= m->alloc(&m->w, nc5);
  rc |= m->alloc(&m->w1, nc5);
  rc |= m->alloc(&m->bcc, 3 * M.ncell * nb);
  const long nf[3] = {M.nf1, M.nf2, M.nf3};
  for (int q = 0; q < 3 && !rc; ++q) {
    rc |= m->alloc(&m->b[q], nf[q] * nb);
    rc |= m->alloc(&m->bb[q], nf[q] * nb);
    rc |= m->alloc(&m->flux[q], NHYDRO * nf[q] * nb);
    rc |= m->alloc(&m->wx[q], nf[q] * nb);
  }
  if (!rc) {
    rc |= m->alloc(&m->e2x1, nf[0] * nb);
    rc |= m->alloc(&m->e3x1, nf[0] * nb);
    rc |= m->alloc(&m->e1x2, nf[1] * nb);
    rc |= m->alloc(&m->e3x2, nf[1] * nb);
    rc |= m->alloc(&m->e1x3, nf[2] * nb);
    rc |= m->alloc(&m->e2x3, nf[2] * nb);
    rc |= m->alloc(&m->e[0], M.ne1 * nb);
    rc |= m->alloc(&m->e[1], M.ne2 * nb);
    rc |= m->alloc(&m->e[2], M.ne3 * nb);
    m->ndtpart = 0;   // room for any tile shape of the autotuner
    for (int q = 0; q < NTILES; ++q)
      m->ndtpart = std::max(m->ndtpart, kind_chunks(M, KK_C2P, TILES[q]) * nb);
    rc |= m->alloc(&m->dtpart, m->ndtpart);
    rc |= m->alloc(&m->coord_d, static_cast<long>(coordh.size()));
  }
  if (!rc) {
    if (hipMalloc(reinterpret_cast<void **>(&m->nbr_d), sizeof(int) * nbrh.size()) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&m->clk), sizeof(Clock)) != hipSuccess) {
      rc = fail("hipMalloc failed for tables");
    } else {
      m->allocs.push_back(m->nbr_d);
      m->allocs.push_back(m->clk);
    }
  }
  if (!rc) {
    Clock c0{0.0, std::numeric_limits<double>::max(), std::numeric_limits<double>::max(), 0.0};
    if (hipMemcpyAsync(m->nbr_d, nbrh.data(), sizeof(int) * nbrh.size(), hipMemcpyHostToDevice,
                       m->stream) != hipSuccess ||
        hipMemcpyAsync(m->coord_d, coordh.data(), sizeof(double) * coordh.size(),
                       hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipMemcpyAsync(m->clk, &c0, sizeof(Clock), hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
      rc = fail("upload of mesh tables failed");
  }
  if (rc) {
    std::string keep = g_err;
    cleanup();
    g_err = keep;
    return 1;
  }
  M.nbr = m->nbr_d;
  M.coord = m->coord_d;
  if (sb && smr_setup(m, *sb, root_level)) {
    std::string keep = g_err;
    cleanup();
    g_err = keep;
    return 1;
  }
  *out = m;
  return 0;
}

}  // namespace

extern "C" {

int ath_gpu_mesh_create(const ath_mesh_desc *desc, ath_mesh_t *out) {
  return mesh_build(desc, nullptr, 0, out);
}

int ath_gpu_smr_mesh_create(const ath_mesh_desc *desc, const long long *locs, int nblocks,
                            int root_level, ath_mesh_t *out) {
  if (!desc || !out) return fail("null argument");
  *out = nullptr;
  const ath_mesh_desc &d = *desc;
  for (int f = 0; f < 6; ++f)
    if (d.bc[f] != 0) return fail("refined meshes are periodic on every face");
  if (d.relativistic) return fail("refined meshes are Newtonian");
  for (int q = 0; q < 3; ++q) {
    if (d.bnx[q] <= 0 || d.nx[q] <= 0 || d.nx[q] % d.bnx[q])
      return fail("mesh nx must be a positive multiple of meshblock nx in every direction");
    if (d.bnx[q] % 2 || d.bnx[q] < 2 * d.ng)
      return fail("refined meshes need even MeshBlock sizes of at least 2 nghost");
  }
  if (d.ng % 2)   // MeshRefinement::MeshRefinement (mesh_refinement.cpp:42-48)
    return fail("nghost must be even with mesh refinement");
  std::vector<smrplan::BlockInfo> sb;
  if (smr_blocks_of(d, locs, nblocks, root_level, sb)) return 1;
  return mesh_build(desc, &sb, root_level, out);
}

int ath_gpu_smr_plan(const ath_mesh_desc *desc, const long long *locs, int nblocks,
                     int root_level, int gid, int *nbr_out, int cap, int *nnbr, long *nsteps) {
  if (!desc || !nnbr || !nsteps) return fail("null argument");
  const ath_mesh_desc &d = *desc;
  for (int q = 0; q < 3; ++q)
    if (d.bnx[q] <= 0 || d.nx[q] % d.bnx[q] || d.bnx[q] % 2 || d.bnx[q] < 2 * d.ng)
      return fail("refined meshes need even MeshBlock sizes of at least 2 nghost dividing nx");
  if (d.ng % 2) return fail("nghost must be even with mesh refinement");
  std::vector<smrplan::BlockInfo> sb;
  if (smr_blocks_of(d, locs, nblocks, root_level, sb)) return 1;
  if (gid < 0 || gid >= nblocks) return fail("block id out of range");
  const smrplan::Shape S = smrplan::make_shape(d.bnx, d.ng);
  smrplan::Program prog[3];
  try {
    for (int q = 0; q < 3; ++q) prog[q].offs.push_back(0);
    for (int g = 0; g < nblocks; ++g) {
      smrplan::flux_program(S, sb, g, prog[0]);
      smrplan::emf_program(S, sb, g, prog[1]);
      smrplan::boundary_program(S, sb, g, prog[2]);
      for (int q = 0; q < 3; ++q) prog[q].offs.push_back(static_cast<int>(prog[q].steps.size()));
    }
  } catch (const std::exception &e) {
    return fail(e.what());
  }
  for (int q = 0; q < 3; ++q) {
    const smrplan::Phased ph = smrplan::to_phases(prog[q], nblocks);
    nsteps[q] = static_cast<long>(ph.steps.size());
    nsteps[3 + q] = static_cast<long>(ph.phase_chunks.size()) - 1;
  }
  const std::vector<smrplan::Nbr> &nb = sb[gid].nb;
  *nnbr = static_cast<int>(nb.size());
  for (int q = 0; q < *nnbr && q < cap && nbr_out; ++q) {
    const smrplan::Nbr &e = nb[q];
    const int v[9] = {e.gid, e.level, e.ox1, e.ox2, e.ox3, e.type, e.fi1, e.fi2,
                      (e.type == smrplan::EDGE) ? e.eid : e.fid};
    for (int t = 0; t < 9; ++t) nbr_out[9 * q + t] = v[t];
  }
  return 0;
}

int ath_gpu_smr_block_level(ath_mesh_t m, int gid, int *level) {
  if (!m || !level) return fail("null argument");
  if (gid < 0 || gid >= m->M.nblocks) return fail("block id out of range");
  *level = m->smr ? m->smr_blocks[gid].level : 0;
  return 0;
}

int ath_gpu_mesh_destroy(ath_mesh_t m) {
  if (!m) return 0;
  (void)hipStreamSynchronize(m->stream);
  for (auto &r : m->prof_rec) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  drop_cycle_graph(m);
  for (void *p : m->allocs) (void)hipFree(p);
  if (m->hist_d) (void)hipFree(m->hist_d);
  (void)hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

int ath_gpu_nblocks(ath_mesh_t m, int *nb) {
  if (!m || !nb) return fail("null argument");
  *nb = m->M.nblocks;
  return 0;
}

int ath_gpu_block_size(ath_mesh_t m, int what, long *n) {
  if (!m || !n) return fail("null argument");
  double **pool;
  return block_count(m, what, n, &pool);
}

int ath_gpu_set_block(ath_mesh_t m, int gid, int what, const double *host) {
  if (!m || !host) return fail("null argument");
  if (gid < 0 || gid >= m->M.nblocks) return fail("block id out of range");
  long n;
  double **pool;
  if (block_count(m, what, &n, &pool)) return 1;
  HIPCHK(hipMemcpyAsync(*pool + static_cast<long>(gid) * n, host, sizeof(double) * n,
                        hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int ath_gpu_get_block(ath_mesh_t m, int gid, int what, double *host) {
  if (!m || !host) return fail("null argument");
  if (gid < 0 || gid >= m->M.nblocks) return fail("block id out of range");
  long n;
  double **pool;
  if (block_count(m, what, &n, &pool)) return 1;
  HIPCHK(hipMemcpyAsync(host, *pool + static_cast<long>(gid) * n, sizeof(double) * n,
                        hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int ath_gpu_initialize(ath_mesh_t m, double tlim) {
  if (!m) return fail("null argument");
  Clock c0{0.0, std::numeric_limits<double>::max(), tlim, 0.0};
  HIPCHK(hipMemcpyAsync(m->clk, &c0, sizeof(Clock), hipMemcpyHostToDevice, m->stream));
  if (m->smr) {   // boundary program; prolongated primitives land in w, then recovered
    if (launch_smr(m, 2, smr_args(m, 2, m->u, m->b, m->w))) return 1;
  } else if (launch_ghost_args(m, GhostArgs{m->u, {m->b[0], m->b[1], m->b[2]}})) {
    return 1;
  }
  // ConservedToPrimitive(u, w1, b, w, bcc): output register is w
  if (launch_cons2prim(m, m->w, true)) return 1;
  if (launch_phys_bnd(m, C2PArgs{m->u, m->w, m->bcc, m->b[0], m->b[1], m->b[2], nullptr,
                                 m->lseed}))
    return 1;
  if (launch_new_dt(m, 0)) return 1;
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int ath_gpu_vl2_stage(ath_mesh_t m, int stg) {
  if (!m) return fail("null argument");
  if (stg != 1 && stg != 2) return fail("stage must be 1 or 2");
  if (stage(m, stg)) return 1;
  if (stg == 2 && launch_new_dt(m, 1)) return 1;
  return 0;
}

int ath_gpu_cycle(ath_mesh_t m, int ncycles, double tlim, int *done) {
  if (!m) return fail("null argument");
  int n = 0;
  const bool bounded = std::isfinite(tlim);
  if (bounded) {   // keep the device tlim in step with the caller's
    HIPCHK(hipMemcpyAsync(&m->clk->tlim, &tlim, sizeof(double), hipMemcpyHostToDevice,
                          m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
  }
  for (; n < ncycles; ++n) {
    if (bounded) {
      Clock c;
      HIPCHK(hipMemcpyAsync(&c, m->clk, sizeof(Clock), hipMemcpyDeviceToHost, m->stream));
      HIPCHK(hipStreamSynchronize(m->stream));
      if (!(c.time < tlim)) break;
    }
    if (m->use_graph && !m->prof_on) {
      if (cycle_graph(m)) return 1;
    } else if (cycle_eager(m)) {
      return 1;
    }
  }
  if (done) *done = n;
  return 0;
}

int ath_gpu_mesh_set_graph(ath_mesh_t m, int on) {
  if (!m) return fail("null argument");
  HIPCHK(hipStreamSynchronize(m->stream));
  m->use_graph = on != 0;
  if (!m->use_graph) drop_cycle_graph(m);
  return 0;
}

int ath_gpu_get_clock(ath_mesh_t m, double *time, double *dt, int *ncycle) {
  if (!m) return fail("null argument");
  Clock c;
  HIPCHK(hipMemcpyAsync(&c, m->clk, sizeof(Clock), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  if (time) *time = c.time;
  if (dt) *dt = c.dt;
  if (ncycle) *ncycle = static_cast<int>(c.ncycle);
  return 0;
}

int ath_gpu_set_clock(ath_mesh_t m, double time, double dt, int ncycle) {
  if (!m) return fail("null argument");
  Clock c;
  HIPCHK(hipMemcpyAsync(&c, m->clk, sizeof(Clock), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  c.time = time;
  c.dt = dt;
  c.ncycle = ncycle;
  HIPCHK(hipMemcpyAsync(m->clk, &c, sizeof(Clock), hipMemcpyHostToDevice, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int ath_gpu_calculate_fluxes(ath_mesh_t m, int order) {
  if (!m) return fail("null argument");
  if (order < 1 || order > 3) return fail("order must be 1, 2 or 3");
  if (order == 3 && m->M.ng < 3) return fail("PPM needs nghost >= 3");
  return launch_flux_dirs(m, order);
}

int ath_gpu_compute_corner_e(ath_mesh_t m) {
  if (!m) return fail("null argument");
  return launch_corner(m);
}

int ath_gpu_stream(ath_mesh_t m, void **stream) {
  if (!m || !stream) return fail("null argument");
  *stream = m->stream;
  return 0;
}

int ath_gpu_synchronize(ath_mesh_t m) {
  if (!m) return fail("null argument");
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int ath_gpu_ensemble_create(const ath_mesh_t *members, int n, ath_ensemble_t *out) {
  if (!members || !out || n <= 0) return fail("bad argument");
  *out = nullptr;
  ath_ensemble *e = new ath_ensemble;
  for (int q = 0; q < n; ++q) {
    ath_mesh *m = members[q];
    if (!m) { delete e; return fail("null member"); }
    for (auto *o : e->mem)
      if (o == m) { delete e; return fail("a mesh appears twice in the ensemble"); }
    const int rs = solver_code(m->d);
    if (q == 0) {
      e->xorder = m->d.xorder;
      e->rs = rs;
      e->sr = m->d.relativistic;
      e->c2p_mode = m->M.c2p_mode;
      e->learned = m->d.relativistic && m->M.c2p_mode == C2P_LEARNED;
    } else if (m->d.xorder != e->xorder || rs != e->rs || m->d.relativistic != e->sr) {
      delete e;
      return fail("ensemble members must share xorder, Riemann solver and dynamics");
    } else if (m->M.c2p_mode != e->c2p_mode) {
      delete e;
      return fail("ensemble members must share the primitive-recovery mode");
    }
    if (m->M.outflow) {
      delete e;
      return fail("ensemble members must be periodic meshes");
    }
    e->mem.push_back(m);
    e->parity.push_back(m->nstage & 1);
  }
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return fail("hipStreamCreate failed");
  }
  if (ens_check(e) || ens_build(e)) {
    delete e;
    return 1;
  }
  *out = e;
  return 0;
}

int ath_gpu_ensemble_destroy(ath_ensemble_t e) {
  if (!e) return 0;
  (void)hipStreamSynchronize(e->stream);
  delete e;
  return 0;
}

int ath_gpu_ensemble_size(ath_ensemble_t e, int *nmembers, long *active_cells) {
  if (!e) return fail("null argument");
  long cells = 0;
  for (auto *m : e->mem) cells += static_cast<long>(m->M.nx1) * m->M.nx2 * m->M.nx3 * m->M.nblocks;
  if (nmembers) *nmembers = static_cast<int>(e->mem.size());
  if (active_cells) *active_cells = cells;
  return 0;
}

int ath_gpu_ensemble_cycle(ath_ensemble_t e, int ncycles, int use_graph) {
  if (!e) return fail("null argument");
  if (ens_check(e)) return 1;
  if (use_graph && !e->exec) {
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const int rc = ens_record_cycle(e);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return 1;
    }
    HIPCHK(ce);
    e->graph = g;
    HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
  }
  for (int c = 0; c < ncycles; ++c) {
    if (use_graph) HIPCHK(hipGraphLaunch(e->exec, e->stream));
    else if (ens_record_cycle(e)) return 1;
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  return 0;
}

int ath_gpu_ensemble_get_tiles(ath_ensemble_t e, int *tile_idx, int *persist) {
  if (!e) return fail("null argument");
  for (int k = 0; k < KK_N; ++k) {
    if (tile_idx) tile_idx[k] = e->tile_idx[k];
    if (persist) persist[k] = e->persist[k];
  }
  return 0;
}

int ath_gpu_ensemble_set_tiles(ath_ensemble_t e, const int *tile_idx, const int *persist) {
  if (!e) return fail("null argument");
  for (int k = 0; k < KK_N; ++k) {
    if (tile_idx && (tile_idx[k] < 0 || tile_idx[k] >= NTILES))
      return fail("tile index out of range");
    if (tile_idx && !tile_ok(k, tile_idx[k]))
      return fail("the z-march tile applies to the corner kernel only");
    if (persist && persist[k] < 0) return fail("persistent grid must be >= 0");
  }
  for (int k = 0; k < KK_N; ++k) {
    if (tile_idx) e->tile_idx[k] = tile_idx[k];
    if (persist) e->persist[k] = persist[k];
  }
  return ens_build(e);
}

// Times every (tile shape, persistent grid) pair of each work kind on the ensemble's own
// data and keeps the fastest.  The trial launches write only scratch registers (fluxes,
// EMFs, u1, b1, w1, bcc); the members' W and bcc are then recomputed from U and B by
// the initialisation path and their clocks restored, so the state after tuning is the
// state before it.
int ath_gpu_ensemble_autotune(ath_ensemble_t e, int reps, double *best_ms) {
  if (!e) return fail("null argument");
  if (reps <= 0) reps = 3;
  if (ens_check(e)) return 1;
  std::vector<Clock> clocks(e->mem.size());
  for (size_t q = 0; q < e->mem.size(); ++q)
    HIPCHK(hipMemcpy(&clocks[q], e->mem[q]->clk, sizeof(Clock), hipMemcpyDeviceToHost));
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  int ncu = 0;
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int grids[3] = {0, 4 * ncu, 8 * ncu};
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  int rc = 0;
  for (int k = 0; k < KK_N && !rc; ++k) {
    float best = 1e30f;
    int bt = e->tile_idx[k], bp = e->persist[k];
    for (int t = 0; t < NTILES && !rc; ++t)
      for (int pg = 0; pg < 3 && !rc; ++pg) {
        if (!tile_ok(k, t)) continue;
        e->tile_idx[k] = t;
        e->persist[k] = grids[pg];
        if (ens_build(e)) { rc = 1; break; }
        // stage 1 of the flux/corner/update kinds reads only valid registers; the
        // recovery reads the u1 that the stage-1 update wrote
        if (k == KK_C2P && ens_launch_kind(e, KK_UPDATE, 1)) { rc = 1; break; }
        if (ens_launch_kind(e, k, 1)) { rc = 1; break; }   // warm-up
        if (hipEventRecord(a, e->stream) != hipSuccess) { rc = fail("hipEventRecord"); break; }
        for (int r = 0; r < reps && !rc; ++r) rc = ens_launch_kind(e, k, 1);
        if (rc || hipEventRecord(b, e->stream) != hipSuccess ||
            hipEventSynchronize(b) != hipSuccess) { rc = rc ? rc : fail("event"); break; }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (ms < best) { best = ms; bt = t; bp = grids[pg]; }
      }
    e->tile_idx[k] = bt;
    e->persist[k] = bp;
    if (best_ms) best_ms[k] = best / reps;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return 1;
  HIPCHK(hipStreamSynchronize(e->stream));
  for (size_t q = 0; q < e->mem.size(); ++q) {   // W, bcc from U, B; clock back
    ath_mesh *m = e->mem[q];
    if (ath_gpu_initialize(m, clocks[q].tlim)) return 1;
    HIPCHK(hipMemcpy(m->clk, &clocks[q], sizeof(Clock), hipMemcpyHostToDevice));
  }
  return ens_build(e);
}

int ath_gpu_history(ath_mesh_t m, double *sums) {
  if (!m || !sums) return fail("null argument");
  const MeshK &M = m->M;
  const long nact = static_cast<long>(M.nx1) * M.nx2 * M.nx3;
  const long nch = (nact + WG - 1) / WG;
  const long nwg = nch * M.nblocks;
  if (!m->hist_d || m->nhist < nwg) {
    if (m->hist_d) HIPCHK(hipFree(m->hist_d));
    m->hist_d = nullptr;
    HIPCHK(hipMalloc(reinterpret_cast<void **>(&m->hist_d),
                     sizeof(double) * (nwg + 1) * NHIST));
    m->nhist = nwg;
  }
  double *out = m->hist_d + nwg * NHIST;
  hipLaunchKernelGGL(k_history, dim3(static_cast<unsigned>(nwg)), dim3(WG), 0, m->stream, M,
                     m->u, m->bcc, m->hist_d);
  hipLaunchKernelGGL(k_history_final, dim3(1), dim3(1024), 0, m->stream, m->hist_d, nwg, out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(sums, out, sizeof(double) * NHIST, hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return 0;
}

int ath_gpu_kernel_box(int kind, int nx, int ng, int *lo, int *n) {
  if (kind < 0 || kind >= KK_N || nx <= 0 || ng < 0 || !lo || !n) return fail("bad argument");
  MeshK M{};
  M.nx1 = M.nx2 = M.nx3 = nx;
  M.ng = ng;
  M.nc1 = M.nc2 = M.nc3 = nx + 2 * ng;
  M.is = M.js = M.ks = ng;
  M.ie = M.je = M.ke = ng + nx - 1;
  kind_box(M, kind, lo, n);
  return 0;
}

int ath_gpu_flux_march_position(const int *shape, int seg, long wave, int lane, int *pos,
                                long *nwave) {
  if (!shape || !pos || !nwave) return fail("null argument");
  if (shape[0] < 1 || shape[1] < 1 || shape[2] < 1 || shape[3] < 1 || shape[4] < 1 ||
      seg < 1)
    return fail("bad argument");
  MeshK M{};
  M.nx1 = shape[0]; M.nx2 = shape[1]; M.nx3 = shape[2];
  M.ng = shape[3];
  M.nblocks = shape[4];
  M.is = M.js = M.ks = M.ng;
  M.ie = M.ng + M.nx1 - 1; M.je = M.ng + M.nx2 - 1; M.ke = M.ng + M.nx3 - 1;
  const FzmGeom G = fzm_geom(M, seg);
  *nwave = G.nwave;
  if (wave < 0 || wave >= G.nwave || lane < 0 || lane > 63) return fail("position out of range");
  int g, j, i, k0, k1;
  pos[0] = fzm_pos(M, G, wave, lane, g, j, i, k0, k1) ? 1 : 0;
  pos[1] = g; pos[2] = j; pos[3] = i; pos[4] = k0; pos[5] = k1; pos[6] = G.hmax;
  return 0;
}

int ath_gpu_ghost_position(int nx, int ng, int array, int t, int *pos) {
  if (!pos) return fail("null argument");
  if (nx < 1 || ng < 1 || array < 0 || array > 3) return fail("bad argument");
  const int nc = nx + 2 * ng;
  const int n1 = nc + (array == 1), n2 = nc + (array == 2), n3 = nc + (array == 3);
  if (t < 0 || t >= ghost_count(n1, n2, n3, ng)) return fail("ghost number out of range");
  int i, j, k;
  ghost_pos(t, n1, n2, n3, ng, i, j, k);
  const int p1 = ghost_off(i, n1, ng), p2 = ghost_off(j, n2, ng), p3 = ghost_off(k, n3, ng);
  const int v[9] = {i, j, k, i - p1 * nx, j - p2 * nx, k - p3 * nx, p1, p2, p3};
  for (int q = 0; q < 9; ++q) pos[q] = v[q];
  return 0;
}

int ath_gpu_mesh_set_tiles(ath_mesh_t m, const int *tile_idx) {
  if (!m || !tile_idx) return fail("null argument");
  for (int k = 0; k < KK_N; ++k)
    if (tile_idx[k] < 0 || tile_idx[k] >= NTILES) return fail("tile index out of range");
  for (int k = 0; k < KK_N; ++k)
    if (!mesh_tile_ok(k, tile_idx[k]))
      return fail("the z-march tiles apply to the flux and corner kernels only");
  HIPCHK(hipStreamSynchronize(m->stream));
  for (int k = 0; k < KK_N; ++k) m->tile[k] = TILES[tile_idx[k]];
  return 0;
}

int ath_gpu_mesh_get_tiles(ath_mesh_t m, int *tile_idx) {
  if (!m || !tile_idx) return fail("null argument");
  for (int k = 0; k < KK_N; ++k) {
    tile_idx[k] = 0;
    for (int t = 0; t < NTILES; ++t)
      if (TILES[t].tx == m->tile[k].tx && TILES[t].ty == m->tile[k].ty &&
          TILES[t].tz == m->tile[k].tz)
        tile_idx[k] = t;
  }
  return 0;
}

// Single-mesh autotune: a one-member ensemble chooses the tile shape of every kernel
// kind (persistent grids do not apply to the single-mesh launches); the mesh's state is
// left as it was.
int ath_gpu_mesh_autotune(ath_mesh_t m, int reps, double *best_ms) {
  if (!m) return fail("null argument");
  // the tuner times the ensemble kernels, which have no physical-boundary step
  if (m->M.outflow) return fail("autotune: meshes with outflow faces keep their tiles");
  if (m->smr) return fail("autotune: refined meshes keep their tiles");
  ath_mesh_t one[1] = {m};
  ath_ensemble_t e = nullptr;
  if (ath_gpu_ensemble_create(one, 1, &e)) return 1;
  const int nop[KK_N] = {0, 0, 0, 0};
  int rc = ath_gpu_ensemble_set_tiles(e, nullptr, nop);
  // only the tile shape is searched: persistent grids stay off
  if (!rc) {
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    std::vector<Clock> clk(1);
    HIPCHK(hipMemcpy(&clk[0], m->clk, sizeof(Clock), hipMemcpyDeviceToHost));
    for (int k = 0; k < KK_N && !rc; ++k) {
      float best = 1e30f;
      int bt = 0;
      for (int t = 0; t < NTILES && !rc; ++t) {
        if (!tile_ok(k, t)) continue;
        e->tile_idx[k] = t;
        if (ens_build(e)) { rc = 1; break; }
        if (k == KK_C2P && ens_launch_kind(e, KK_UPDATE, 1)) { rc = 1; break; }
        if (ens_launch_kind(e, k, 1)) { rc = 1; break; }
        if (hipEventRecord(a, e->stream) != hipSuccess) { rc = fail("hipEventRecord"); break; }
        for (int r = 0; r < (reps > 0 ? reps : 3) && !rc; ++r) rc = ens_launch_kind(e, k, 1);
        if (rc || hipEventRecord(b, e->stream) != hipSuccess ||
            hipEventSynchronize(b) != hipSuccess) { rc = rc ? rc : fail("event"); break; }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (ms < best) { best = ms; bt = t; }
      }
      e->tile_idx[k] = bt;
      m->tile[k] = TILES[bt];
      if (best_ms) best_ms[k] = best / (reps > 0 ? reps : 3);
    }
    // the flux z-march exists for single meshes only: time the chosen flux tile and the
    // z-march tiles with the mesh's own flux launch at its reconstruction order (the
    // launches write only the flux, EMF and weight registers)
    if (!rc) {
      HIPCHK(hipStreamSynchronize(e->stream));
      const int nrep = reps > 0 ? reps : 3;
      auto time_flux = [&](TileK T, float *ms) {
        m->tile[KK_FLUX] = T;
        if (launch_flux(m, m->d.xorder)) return 1;   // warm-up
        if (hipEventRecord(a, m->stream) != hipSuccess) return fail("hipEventRecord");
        for (int r = 0; r < nrep; ++r)
          if (launch_flux(m, m->d.xorder)) return 1;
        if (hipEventRecord(b, m->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess)
          return fail("event");
        (void)hipEventElapsedTime(ms, a, b);
        return 0;
      };
      const TileK chosen = m->tile[KK_FLUX];
      TileK best = chosen;
      float tbest = 0.0f;
      rc = time_flux(chosen, &tbest);
      for (int t = 0; t < NTILES && !rc; ++t) {
        if (tile_ok(KK_FLUX, t) || !mesh_tile_ok(KK_FLUX, t)) continue;   // z-march tiles
        float tz = 0.0f;
        rc = time_flux(TILES[t], &tz);
        if (!rc && tz < tbest) {
          tbest = tz;
          best = TILES[t];
        }
      }
      m->tile[KK_FLUX] = best;
      if (!rc && best_ms && (best.tx < 0)) best_ms[KK_FLUX] = tbest / nrep;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!rc) {
      HIPCHK(hipStreamSynchronize(e->stream));
      if (ath_gpu_initialize(m, clk[0].tlim)) rc = 1;
      else HIPCHK(hipMemcpy(m->clk, &clk[0], sizeof(Clock), hipMemcpyHostToDevice));
    }
  }
  ath_gpu_ensemble_destroy(e);
  return rc;
}

int ath_gpu_ensemble_stream(ath_ensemble_t e, void **stream) {
  if (!e || !stream) return fail("null argument");
  *stream = e->stream;
  return 0;
}

int ath_gpu_mesh_set_c2p(ath_mesh_t m, ath_mlp_t mlp, int mode) {
  if (!m) return fail("null argument");
  if (mode < C2P_REFERENCE || mode > C2P_LEARNED) return fail("mode must be 0, 1 or 2");
  if (mode != C2P_REFERENCE && !m->d.relativistic)
    return fail("solver modes other than 0 apply to relativistic meshes only");
  if (mode == C2P_LEARNED && !mlp) return fail("the learned mode needs an MLP");
  HIPCHK(hipStreamSynchronize(m->stream));
  if (mode == C2P_LEARNED && !m->lseed) {
    HIPCHK(hipMalloc(reinterpret_cast<void **>(&m->lseed),
                     sizeof(double) * m->M.ncell * m->M.nblocks));
    m->allocs.push_back(m->lseed);
  }
  m->M.c2p_mode = mode;
  m->M.mlp = ath_ml_device_view(mlp);
  return 0;
}

int ath_gpu_mesh_c2p_stats(ath_mesh_t m, int enable, unsigned long long *totals) {
  if (!m) return fail("null argument");
  HIPCHK(hipStreamSynchronize(m->stream));
  const size_t bytes = 3 * C2P_STAT_SLOTS * sizeof(unsigned long long);
  if (!m->c2p_stats_d) {
    HIPCHK(hipMalloc(reinterpret_cast<void **>(&m->c2p_stats_d), bytes));
    m->allocs.push_back(m->c2p_stats_d);
    HIPCHK(hipMemset(m->c2p_stats_d, 0, bytes));
  }
  if (totals) {
    std::vector<unsigned long long> h(3 * C2P_STAT_SLOTS);
    HIPCHK(hipMemcpy(h.data(), m->c2p_stats_d, bytes, hipMemcpyDeviceToHost));
    totals[0] = totals[1] = totals[2] = 0;
    for (int sl = 0; sl < C2P_STAT_SLOTS; ++sl)
      for (int q = 0; q < 3; ++q) totals[q] += h[3 * sl + q];
  }
  if (enable > 0) HIPCHK(hipMemset(m->c2p_stats_d, 0, bytes));
  if (enable >= 0) m->M.c2p_stats = enable ? m->c2p_stats_d : nullptr;
  return 0;
}

int ath_gpu_kernel_timing(ath_mesh_t m, int enable) {
  if (!m) return fail("null argument");
  HIPCHK(hipStreamSynchronize(m->stream));
  for (auto &r : m->prof_rec) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  m->prof_rec.clear();
  for (int q = 0; q < 64; ++q) { m->prof_ms[q] = 0.0; m->prof_cnt[q] = 0; }
  m->prof_on = enable != 0;
  return 0;
}

int ath_gpu_kernel_stats(ath_mesh_t m, int id, char *name, int namelen, double *total_ms,
                         long *count) {
  if (!m) return fail("null argument");
  if (id < 0 || id >= KID_COUNT) return fail("kernel id out of range");
  HIPCHK(hipStreamSynchronize(m->stream));
  for (auto &r : m->prof_rec) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
    m->prof_ms[r.id] += ms;
    m->prof_cnt[r.id] += 1;
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  m->prof_rec.clear();
  if (name && namelen > 0) {
    std::snprintf(name, namelen, "%s", kernel_name(id));
  }
  if (total_ms) *total_ms = m->prof_ms[id];
  if (count) *count = m->prof_cnt[id];
  return 0;
}

int ath_gpu_riemann_batch(int solver, double gamma, int n, const double *wl,
                          const double *wr, const double *bx, double *flux) {
  if (solver < 0 || solver > 3) return fail("solver must be 0, 1, 2 or 3");
  if (n <= 0) return 0;
  if (!wl || !wr || !bx || !flux) return fail("null argument");
  EosK eos{};
  eos.gamma = gamma;
  eos.igm1 = 1.0 / (gamma - 1.0);
  double *d = nullptr;
  const size_t n7 = static_cast<size_t>(n) * NWAVE;
  HIPCHK(hipMalloc(reinterpret_cast<void **>(&d), sizeof(double) * (3 * n7 + n)));
  double *dwl = d, *dwr = d + n7, *dout = d + 2 * n7, *dbx = d + 3 * n7;
  int rc = 0;
  if (hipMemcpy(dwl, wl, sizeof(double) * n7, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dwr, wr, sizeof(double) * n7, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dbx, bx, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail("upload failed");
  } else {
    hipLaunchKernelGGL(k_riemann_batch, dim3((n + 255) / 256), dim3(256), 0, 0, solver, eos,
                       n, dwl, dwr, dbx, dout);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(flux, dout, sizeof(double) * n7, hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail("riemann batch failed");
  }
  (void)hipFree(d);
  return rc;
}

int ath_gpu_sr_con2prim_batch(int n, double gamma, double dfloor, double pfloor,
                              double sigma_max, double beta_min, double gamma_max,
                              double *cons, double *prim, const double *bcc, int *iters) {
  if (n <= 0) return 0;
  if (!cons || !prim || !bcc || !iters) return fail("null argument");
  EosK eos{};
  eos.gamma = gamma;
  eos.igm1 = 1.0 / (gamma - 1.0);
  eos.dfloor = dfloor;
  eos.pfloor = pfloor;
  eos.sigma_max = sigma_max;
  eos.beta_min = beta_min;
  eos.gamma_max = gamma_max;
  double *d = nullptr;
  int *di = nullptr;
  const size_t n5 = static_cast<size_t>(n) * 5;
  HIPCHK(hipMalloc(reinterpret_cast<void **>(&d), sizeof(double) * (2 * n5 + 3 * n)));
  if (hipMalloc(reinterpret_cast<void **>(&di), sizeof(int) * n) != hipSuccess) {
    (void)hipFree(d);
    return fail("hipMalloc failed");
  }
  int rc = 0;
  double *dc = d, *dp = d + n5, *db = d + 2 * n5;
  if (hipMemcpy(dc, cons, sizeof(double) * n5, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dp, prim, sizeof(double) * n5, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db, bcc, sizeof(double) * 3 * n, hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail("upload failed");
  } else {
    hipLaunchKernelGGL(k_sr_c2p_batch, dim3((n + 255) / 256), dim3(256), 0, 0, eos, n, dc, dp,
                       db, di);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(cons, dc, sizeof(double) * n5, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(prim, dp, sizeof(double) * n5, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(iters, di, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail("sr con2prim batch failed");
  }
  (void)hipFree(d);
  (void)hipFree(di);
  return rc;
}


int ath_gpu_mesh_set_fused(ath_mesh_t m, int rows) {
  if (!m) return fail("null argument");
  const MeshK &M = m->M;
  if (rows < 0) rows = fdiv_default_rows(M);
  if (rows > 0 && m->smr) return fail("fused flux kernel: refined meshes use separate kernels");
  if (rows > 0 && !fdiv_rows_ok(M, rows))
    return fail("fused flux kernel: rows must be in 1 .. nx2 with (rows + 2) * (nx1 + 2) <= " +
                std::to_string(WG));
  HIPCHK(hipStreamSynchronize(m->stream));
  m->fdiv_rows = rows;
  return 0;
}

int ath_gpu_mesh_get_fused(ath_mesh_t m, int *rows) {
  if (!m || !rows) return fail("null argument");
  *rows = m->fdiv_rows;
  return 0;
}

int ath_gpu_flux_div_plan(const int *nx, int ng, int rows, int *faces, long max_faces,
                          long *nfaces, int *cells, long max_cells, long *ncells,
                          long *lds_doubles) {
  if (!nx || !nfaces || !ncells || !lds_doubles) return fail("null argument");
  if (ng < 1 || nx[0] < 1 || nx[1] < 1 || nx[2] < 1) return fail("bad block shape");
  MeshK M{};
  M.nx1 = nx[0]; M.nx2 = nx[1]; M.nx3 = nx[2]; M.ng = ng;
  M.nc1 = M.nx1 + 2 * ng; M.nc2 = M.nx2 + 2 * ng; M.nc3 = M.nx3 + 2 * ng;
  M.is = ng; M.ie = ng + M.nx1 - 1; M.js = ng; M.je = ng + M.nx2 - 1;
  M.ks = ng; M.ke = ng + M.nx3 - 1;
  if (!fdiv_rows_ok(M, rows))
    return fail("rows must be in 1 .. nx2 with (rows + 2) * (nx1 + 2) <= " +
                std::to_string(WG));
  *lds_doubles = FDIV_LDS_DOUBLES;
  *nfaces = 0;
  *ncells = 0;
  auto face = [&](int dir, int k, int j, int i) {
    if (faces && *nfaces < max_faces) {
      int *f = faces + 4 * *nfaces;
      f[0] = dir; f[1] = k; f[2] = j; f[3] = i;
    }
    ++*nfaces;
  };
  const int ni = M.nx1 + 2;
  std::vector<FdivThread> th(WG);
  for (int b = 0; b < fdiv_nband(M, rows); ++b) {
    const FdivBand B = fdiv_band(M, rows, b);
    for (int t = 0; t < WG; ++t) th[t] = fdiv_thread(M, B, t);
    // every live thread of the band's rows lies in its work-group
    const int R = B.r1 - B.r0 + 1, nrow = R + ((B.q1 > B.r1) ? 1 : 0);
    if (nrow * ni > WG) return fail("a band needs more threads than a work-group has");
    for (int k = M.ks - 1; k <= M.ke + 1; ++k) {
      for (int t = 0; t < WG; ++t) {
        const FdivThread &T = th[t];
        if (T.x3 && k >= M.ks) face(2, k, T.j, T.i);
        if (T.x1) face(0, k, T.j, T.i);
        if (T.x2 && T.emit2) face(1, k, T.j, T.i);
        if (!T.own) continue;
        // the owner's reads: F1[t], F1[t+1], F2[t], F2[t+ni] must hold the x1 faces i and
        // i+1 and the x2 faces j and j+1 of its cell, computed by those threads
        if (k >= M.ks && k <= M.ke) {
          const int src[4] = {t, t + 1, t, t + ni};
          const int dir[4] = {0, 0, 1, 1};
          const int dj[4] = {0, 0, 0, 1}, di[4] = {0, 1, 0, 0};
          for (int q = 0; q < 4; ++q) {
            if (src[q] < 0 || src[q] >= WG) return fail("LDS read out of range");
            const FdivThread &S = th[src[q]];
            const bool has = dir[q] == 0 ? S.x1 : S.x2;
            if (!has || S.j != T.j + dj[q] || S.i != T.i + di[q])
              return fail("an owned cell reads a flux its neighbour does not compute");
          }
        }
        if (!T.x3) return fail("an owned cell's column computes no x3 faces");
        if (k - 1 >= M.ks && k - 1 <= M.ke) {
          if (cells && *ncells < max_cells) {
            int *c = cells + 3 * *ncells;
            c[0] = k - 1; c[1] = T.j; c[2] = T.i;
          }
          ++*ncells;
        }
      }
    }
  }
  return 0;
}

//----------------------------------------------------------------------------------------
// Per-MeshBlock entries over caller-owned device arrays

int ath_gpu_block_create(const ath_block_desc *d, ath_block_t *out) {
  if (!d || !out) return fail("null argument");
  *out = nullptr;
  for (int q = 0; q < 3; ++q)
    if (d->nx[q] < 2) return fail("block nx must be >= 2 in every direction (3-D only)");
  if (d->ng < 2) return fail("nghost must be >= 2");
  for (int q = 0; q < 3; ++q)
    if (d->nx[q] < d->ng) return fail("block nx must be >= nghost");
  if (d->rsolver < 0 || d->rsolver > 1) return fail("rsolver must be 0 (HLLD) or 1 (HLLE)");
  if (!(d->gamma > 1.0)) return fail("gamma must exceed 1");
  const double *xf[3] = {d->x1f, d->x2f, d->x3f}, *xv[3] = {d->x1v, d->x2v, d->x3v};
  const double *dxf[3] = {d->dx1f, d->dx2f, d->dx3f};
  for (int q = 0; q < 3; ++q)
    if (!xf[q] || !xv[q] || !dxf[q]) return fail("null coordinate array");
  ath_block *b = new ath_block;
  MeshK &M = b->M;
  M = MeshK{};
  M.nx1 = d->nx[0]; M.nx2 = d->nx[1]; M.nx3 = d->nx[2]; M.ng = d->ng;
  M.nc1 = M.nx1 + 2 * M.ng; M.nc2 = M.nx2 + 2 * M.ng; M.nc3 = M.nx3 + 2 * M.ng;
  M.is = M.ng; M.ie = M.ng + M.nx1 - 1;
  M.js = M.ng; M.je = M.ng + M.nx2 - 1;
  M.ks = M.ng; M.ke = M.ng + M.nx3 - 1;
  M.nblocks = 1;
  M.ncell = static_cast<long>(M.nc1) * M.nc2 * M.nc3;
  M.nf1 = static_cast<long>(M.nc3) * M.nc2 * (M.nc1 + 1);
  M.nf2 = static_cast<long>(M.nc3) * (M.nc2 + 1) * M.nc1;
  M.nf3 = static_cast<long>(M.nc3 + 1) * M.nc2 * M.nc1;
  M.ne1 = static_cast<long>(M.nc3 + 1) * (M.nc2 + 1) * M.nc1;
  M.ne2 = static_cast<long>(M.nc3 + 1) * M.nc2 * (M.nc1 + 1);
  M.ne3 = static_cast<long>(M.nc3) * (M.nc2 + 1) * (M.nc1 + 1);
  M.coff[0] = 0;
  M.coff[1] = 3 * M.nc1;
  M.coff[2] = 3 * (M.nc1 + M.nc2);
  M.cstride = 3 * (M.nc1 + M.nc2 + M.nc3);
  const double fmin = static_cast<double>(std::numeric_limits<float>::min());
  M.eos.gamma = d->gamma;
  M.eos.igm1 = 1.0 / (d->gamma - 1.0);
  M.eos.dfloor = (d->dfloor < 0.0) ? std::sqrt(1024 * fmin) : d->dfloor;
  M.eos.pfloor = (d->pfloor < 0.0) ? std::sqrt(1024 * fmin) : d->pfloor;
  M.eos.sigma_max = (d->sigma_max < 0.0) ? 0.0 : d->sigma_max;
  M.eos.beta_min = (d->beta_min < 0.0) ? 0.0 : d->beta_min;
  M.eos.gamma_max = (d->gamma_max < 0.0) ? 1000.0 : d->gamma_max;
  M.cfl = d->cfl;
  M.outflow = 1;          // no neighbour on any side: the EMF accessors return raw values
  M.c2p_mode = C2P_REFERENCE;
  b->relativistic = d->relativistic ? 1 : 0;
  b->rs = d->relativistic ? ((d->rsolver == 0) ? 3 : 2) : d->rsolver;
  // coordinate table: per direction dxNf, the PLM face factors (plm.cpp:149-155)
  std::vector<double> tab(M.cstride);
  const int nc[3] = {M.nc1, M.nc2, M.nc3};
  for (int q = 0; q < 3; ++q) {
    double *t = tab.data() + M.coff[q];
    for (int i = 0; i < nc[q]; ++i) {
      t[i] = dxf[q][i];
      t[nc[q] + i] = (xf[q][i + 1] - xv[q][i]) / dxf[q][i];
      t[2 * nc[q] + i] = (xv[q][i] - xf[q][i]) / dxf[q][i];
    }
  }
  std::vector<int> nb(27, -1);
  if (hipMalloc(reinterpret_cast<void **>(&b->coord_d), sizeof(double) * tab.size()) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&b->nbr_d), sizeof(int) * 27) != hipSuccess ||
      hipMemcpy(b->coord_d, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->nbr_d, nb.data(), sizeof(int) * 27, hipMemcpyHostToDevice) != hipSuccess) {
    if (b->coord_d) (void)hipFree(b->coord_d);
    if (b->nbr_d) (void)hipFree(b->nbr_d);
    delete b;
    return fail("ath_gpu_block_create: device allocation or upload failed");
  }
  M.coord = b->coord_d;
  M.nbr = b->nbr_d;
  if (scratch_address(&b->scratch_d) != 0) {
    (void)hipFree(b->coord_d);
    (void)hipFree(b->nbr_d);
    delete b;
    return 1;
  }
  *out = b;
  return 0;
}

int ath_gpu_block_destroy(ath_block_t b) {
  if (!b) return 0;
  (void)hipFree(b->coord_d);
  (void)hipFree(b->nbr_d);
  delete b;
  return 0;
}

int ath_gpu_block_reconstruct(ath_block_t b, int dir, int order, int k, int j, int il, int iu,
                              const double *w, const double *bcc, double *wl, double *wr,
                              void *stream) {
  if (!b || !w || !bcc || !wl || !wr) return fail("null argument");
  if (dir < 0 || dir > 2) return fail("dir must be 0, 1 or 2");
  if (order < 1 || order > 3) return fail("order must be 1, 2 or 3");
  const MeshK &M = b->M;
  if (order == 3 && M.ng < 3) return fail("PPM needs nghost >= 3");
  const int e = order - 1;   // stencil half-width
  const int pos[3] = {0, j, k}, n[3] = {M.nc1, M.nc2, M.nc3};
  if (!in_box(il, iu, M.nc1) || !in_box(j, j, M.nc2) || !in_box(k, k, M.nc3))
    return fail("pencil outside the block");
  if (dir == 0 && (il - e < 0 || iu + std::max(e, 1) >= M.nc1))
    return fail("reconstruction stencil leaves the block along x1");
  if (dir > 0 && (pos[dir] - e < 0 || pos[dir] + e >= n[dir]))
    return fail("reconstruction stencil leaves the block");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g = pencil_grid(il, iu);
#define LR(D, O) hipLaunchKernelGGL((k_blk_recon<D, O>), g, dim3(WG), 0, st, M, k, j, il, iu, w, bcc, wl, wr)
#define LR_O(D)                 \
  do {                          \
    if (order == 1) LR(D, 1);   \
    else if (order == 2) LR(D, 2); \
    else LR(D, 3);              \
  } while (0)
  if (dir == 0) LR_O(0);
  else if (dir == 1) LR_O(1);
  else LR_O(2);
#undef LR_O
#undef LR
  return blk_check_stream(st);
}

int ath_gpu_block_riemann(ath_block_t b, int k, int j, int il, int iu, int ivx,
                          const double *bx, const double *wl, const double *wr, double *flx,
                          double *ey, double *ez, double *wct, const double *dxw, double dt,
                          void *stream) {
  if (!b || !bx || !wl || !wr || !flx || !ey || !ez || !wct || !dxw) return fail("null argument");
  if (ivx < 1 || ivx > 3) return fail("ivx must be IVX, IVY or IVZ (1, 2, 3)");
  const MeshK &M = b->M;
  const int d = ivx - 1;
  // face-array extents along (i, j, k) for the direction of ivx
  const int n1 = M.nc1 + (d == 0), n2 = M.nc2 + (d == 1), n3 = M.nc3 + (d == 2);
  if (!in_box(il, iu, n1) || iu >= M.nc1 || !in_box(j, j, n2) || !in_box(k, k, n3))
    return fail("pencil outside the face arrays");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g = pencil_grid(il, iu);
  const int rs = b->rs;
#define LRS(D, R) hipLaunchKernelGGL((k_blk_riemann<D, R>), g, dim3(WG), 0, st, M, k, j, il, iu, bx, wl, wr, flx, ey, ez, wct, dxw, dt)
#define LRS_R(D)                   \
  do {                             \
    if (rs == 0) LRS(D, 0);        \
    else if (rs == 1) LRS(D, 1);   \
    else if (rs == 2) LRS(D, 2);   \
    else LRS(D, 3);                \
  } while (0)
  if (d == 0) LRS_R(0);
  else if (d == 1) LRS_R(1);
  else LRS_R(2);
#undef LRS_R
#undef LRS
  return blk_check_stream(st);
}

int ath_gpu_block_calculate_fluxes(ath_block_t b, const double *w, const double *b1,
                                   const double *b2, const double *b3, const double *bcc,
                                   int order, double dt, const ath_block_regs *r,
                                   void *stream) {
  if (!b || !w || !b1 || !b2 || !b3 || !bcc || !r) return fail("null argument");
  if (order < 1 || order > 3) return fail("order must be 1, 2 or 3");
  const MeshK &M = b->M;
  if (order == 3 && M.ng < 3) return fail("PPM needs nghost >= 3");
  for (int q = 0; q < 3; ++q)
    if (!r->flux[q] || !r->wght[q]) return fail("null flux or weight array");
  if (!r->e3_x1f || !r->e2_x1f || !r->e1_x2f || !r->e3_x2f || !r->e2_x3f || !r->e1_x3f)
    return fail("null face EMF array");
  FaceOut o[3];
  o[0] = FaceOut{b1, r->flux[0], r->e3_x1f, r->e2_x1f, r->wght[0]};
  o[1] = FaceOut{b2, r->flux[1], r->e1_x2f, r->e3_x2f, r->wght[1]};
  o[2] = FaceOut{b3, r->flux[2], r->e2_x3f, r->e1_x3f, r->wght[2]};
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g(static_cast<unsigned>(kind_chunks(M, KK_FLUX, TileK{0, 0, 0})));
  const int rs = b->rs;
#define LF(O, R) hipLaunchKernelGGL((k_blk_flux3<O, R>), g, dim3(WG), 0, st, M, w, bcc, o[0], o[1], o[2], dt)
#define LF_O(R)                         \
  do {                                  \
    if (order == 1) LF(1, R);           \
    else if (order == 2) LF(2, R);      \
    else LF(3, R);                      \
  } while (0)
  if (rs == 0) LF_O(0);
  else if (rs == 1) LF_O(1);
  else if (rs == 2) LF_O(2);
  else LF_O(3);
#undef LF_O
#undef LF
  return blk_check_stream(st);
}

int ath_gpu_block_compute_corner_e(ath_block_t b, const double *w, const double *bcc,
                                   const ath_block_regs *r, void *stream) {
  if (!b || !w || !bcc || !r) return fail("null argument");
  for (int q = 0; q < 3; ++q)
    if (!r->e[q] || !r->wght[q]) return fail("null edge EMF or weight array");
  if (!r->e3_x1f || !r->e2_x1f || !r->e1_x2f || !r->e3_x2f || !r->e2_x3f || !r->e1_x3f)
    return fail("null face EMF array");
  const MeshK &M = b->M;
  CornerArgs ca{w, bcc, r->e2_x1f, r->e3_x1f, r->e1_x2f, r->e3_x2f, r->e1_x3f, r->e2_x3f,
                r->wght[0], r->wght[1], r->wght[2], r->e[0], r->e[1], r->e[2]};
  hipStream_t st = static_cast<hipStream_t>(stream);
  const TileK T{0, 0, 0};
  const dim3 g(static_cast<unsigned>(kind_chunks(M, KK_CORNER, T)));
  if (b->relativistic) hipLaunchKernelGGL((k_corner<true>), g, dim3(WG), 0, st, M, ca, T);
  else hipLaunchKernelGGL((k_corner<false>), g, dim3(WG), 0, st, M, ca, T);
  return blk_check_stream(st);
}

static int blk_update(ath_block_t b, const ath_block_regs *r, double wght, double *u_out,
                      double *const *b_out, int what, void *stream) {
  const MeshK &M = b->M;
  UpdateArgs a{};
  if (what & 1) {
    for (int q = 0; q < 3; ++q)
      if (!r->flux[q]) return fail("null flux array");
    a.uout = u_out;
    a.ubase = u_out;
    a.f1 = r->flux[0]; a.f2 = r->flux[1]; a.f3 = r->flux[2];
  }
  if (what & 2) {
    for (int q = 0; q < 3; ++q)
      if (!r->e[q] || !b_out[q]) return fail("null edge EMF or field array");
    for (int q = 0; q < 3; ++q) { a.bin[q] = b_out[q]; a.bout[q] = b_out[q]; }
    a.e1 = r->e[0]; a.e2 = r->e[1]; a.e3 = r->e[2];
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g(static_cast<unsigned>(kind_chunks(M, KK_UPDATE, TileK{0, 0, 0})));
  hipLaunchKernelGGL(k_blk_update, g, dim3(WG), 0, st, M, a, wght, what);
  return blk_check_stream(st);
}

int ath_gpu_block_ct(ath_block_t b, double wght, const ath_block_regs *r, double *b1_out,
                     double *b2_out, double *b3_out, void *stream) {
  if (!b || !r || !b1_out || !b2_out || !b3_out) return fail("null argument");
  double *bo[3] = {b1_out, b2_out, b3_out};
  return blk_update(b, r, wght, nullptr, bo, 2, stream);
}

int ath_gpu_block_add_flux_divergence(ath_block_t b, double wght, const ath_block_regs *r,
                                      double *u_out, void *stream) {
  if (!b || !r || !u_out) return fail("null argument");
  double *bo[3] = {nullptr, nullptr, nullptr};
  return blk_update(b, r, wght, u_out, bo, 1, stream);
}

int ath_gpu_block_cons2prim(ath_block_t b, double *cons, const double *prim_old,
                            const double *b1, const double *b2, const double *b3,
                            double *prim, double *bcc, int il, int iu, int jl, int ju, int kl,
                            int ku, void *stream) {
  (void)prim_old;   // both recoveries read their seed (SR) from prim itself
  if (!b || !cons || !b1 || !b2 || !b3 || !prim || !bcc) return fail("null argument");
  const MeshK &M = b->M;
  if (!in_box(il, iu, M.nc1) || !in_box(jl, ju, M.nc2) || !in_box(kl, ku, M.nc3))
    return fail("range outside the block");
  C2PArgs a{cons, prim, bcc, b1, b2, b3, nullptr, nullptr};
  const int ni = iu - il + 1, nj = ju - jl + 1, nk = ku - kl + 1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g(static_cast<unsigned>((static_cast<long>(ni) * nj * nk + WG - 1) / WG));
  if (b->relativistic)
    hipLaunchKernelGGL((k_blk_c2p<true>), g, dim3(WG), 0, st, M, a, il, jl, kl, ni, nj, nk);
  else
    hipLaunchKernelGGL((k_blk_c2p<false>), g, dim3(WG), 0, st, M, a, il, jl, kl, ni, nj, nk);
  return blk_check_stream(st);
}

int ath_gpu_block_new_dt(ath_block_t b, const double *w, const double *bcc, const double *b1,
                         const double *b2, const double *b3, double *dt_out, void *stream) {
  if (!b || !w || !bcc || !b1 || !b2 || !b3 || !dt_out) return fail("null argument");
  const MeshK &M = b->M;
  const int npart = (M.nx1 * M.nx2 * M.nx3 + WG - 1) / WG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double *part = nullptr;
  HIPCHK(hipMalloc(reinterpret_cast<void **>(&part), sizeof(double) * npart));
  if (b->relativistic)
    hipLaunchKernelGGL((k_blk_dt<true>), dim3(npart), dim3(WG), 0, st, M, w, bcc, b1, b2, b3, part);
  else
    hipLaunchKernelGGL((k_blk_dt<false>), dim3(npart), dim3(WG), 0, st, M, w, bcc, b1, b2, b3, part);
  hipLaunchKernelGGL(k_blk_dt_final, dim3(1), dim3(1024), 0, st, part, npart, M.cfl, dt_out);
  const hipError_t le = hipGetLastError();
  const hipError_t se = hipStreamSynchronize(st);   // part is freed below
  (void)hipFree(part);
  if (le != hipSuccess) return fail(std::string("ath_gpu_block_new_dt: ") + hipGetErrorString(le));
  if (se != hipSuccess) return fail(std::string("ath_gpu_block_new_dt: ") + hipGetErrorString(se));
  return 0;
}

int ath_gpu_block_new_dt_noalloc(ath_block_t b, const double *w, const double *bcc,
                                 const double *b1, const double *b2, const double *b3,
                                 double *dt_out, void *stream) {
  if (!b || !w || !bcc || !b1 || !b2 || !b3 || !dt_out) return fail("null argument");
  const MeshK &M = b->M;
  const long ncell = static_cast<long>(M.nx1) * M.nx2 * M.nx3;
  const long npart = (ncell + WG - 1) / WG;
  if (npart > kDtScratchMax)
    return fail("ath_gpu_block_new_dt_noalloc: " + std::to_string(ncell) +
                " active cells exceed the scratch limit of " +
                std::to_string(static_cast<long>(kDtScratchMax) * WG));
  double *part = b->scratch_d;   // every slot 0 .. npart-1 is written before it is read
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->relativistic)
    hipLaunchKernelGGL((k_blk_dt<true>), dim3(npart), dim3(WG), 0, st, M, w, bcc, b1, b2, b3, part);
  else
    hipLaunchKernelGGL((k_blk_dt<false>), dim3(npart), dim3(WG), 0, st, M, w, bcc, b1, b2, b3, part);
  hipLaunchKernelGGL(k_blk_dt_final, dim3(1), dim3(1024), 0, st, part, static_cast<int>(npart),
                     M.cfl, dt_out);
  return blk_check_stream(st);
}

}  // extern "C"
