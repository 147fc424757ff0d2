//========================================================================================
// refine_ops.hpp -- the MeshRefinement operators at one position, shared by the batched
// refinement kernels (refine.hip) and the static-mesh-refinement boundary program
// (athena_gpu.hip).  3-D uniform-Cartesian branches of mesh_refinement.cpp:116-150
// (RestrictCellCenteredValues), 272-300, 328-372, 402-430 (RestrictFieldX1..3), 464-525
// (ProlongateCellCenteredValues), 835-872 and the matching Cartesian blocks of
// ProlongateSharedFieldX2/X3, 1223-1338 (ProlongateInternalField), each expression in the
// reference's association order (the library is built with -ffp-contract=off).
//
// A fine grid has nc = nx + 2 ng cells per direction, its coarse buffer ncc = nx/2 + 2 ng;
// fine index = (coarse index - ng) * 2 + ng.  Coordinate tables are packed x1f (n1+1),
// x1v (n1), dx1f (n1), then x2, x3.  Cartesian metrics: volume dx1 dx2 dx3, face areas
// dx2 dx3 / dx1 dx3 / dx1 dx2; face-centre positions along transverse directions are the
// cell centres (cartesian.cpp:105-118).
//========================================================================================
#ifndef ATHENA_AMD_REFINE_OPS_HPP_
#define ATHENA_AMD_REFINE_OPS_HPP_

#include <hip/hip_runtime.h>

namespace refops {

// Coordinate tables of one grid (n1 x n2 x n3 cells).
struct RGrid {
  const double *x1f, *x1v, *dx1f, *x2f, *x2v, *dx2f, *x3f, *x3v, *dx3f;
  __device__ RGrid(const double *c, int n1, int n2, int n3) {
    x1f = c; x1v = x1f + n1 + 1; dx1f = x1v + n1;
    x2f = dx1f + n1; x2v = x2f + n2 + 1; dx2f = x2v + n2;
    x3f = dx2f + n2; x3v = x3f + n3 + 1; dx3f = x3v + n3;
  }
};

__device__ __forceinline__ double rsign(double x) { return (x < 0.0) ? -1.0 : 1.0; }
__device__ __forceinline__ double rmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double minmod_c(double gm, double gp) {
  return 0.5 * (rsign(gm) + rsign(gp)) * rmin(fabs(gm), fabs(gp));
}

__device__ __forceinline__ long at(int k, int j, int i, int n2, int n1) {
  return (static_cast<long>(k) * n2 + j) * n1 + i;
}

// A 3-D array view (n3 planes of n2 rows of n1) over one variable of one block.
struct View {
  double *p;
  int n2, n1;
  __device__ double &operator()(int k, int j, int i) const { return p[at(k, j, i, n2, n1)]; }
};

// Volume-weighted average of the eight fine cells whose lowest corner is (k, j, i).
__device__ __forceinline__ double restrict_cc_at(const View &f, const RGrid &F, int k, int j,
                                                 int i) {
  auto vol = [&](int kk, int jj, int ii) { return F.dx1f[ii] * F.dx2f[jj] * F.dx3f[kk]; };
  const double v000 = vol(k, j, i), v010 = vol(k, j + 1, i);
  const double v001 = vol(k, j, i + 1), v011 = vol(k, j + 1, i + 1);
  const double v100 = vol(k + 1, j, i), v110 = vol(k + 1, j + 1, i);
  const double v101 = vol(k + 1, j, i + 1), v111 = vol(k + 1, j + 1, i + 1);
  const double tvol = ((v000 + v010) + (v001 + v011)) + ((v100 + v110) + (v101 + v111));
  return (((f(k, j, i) * v000 + f(k, j + 1, i) * v010) +
           (f(k, j, i + 1) * v001 + f(k, j + 1, i + 1) * v011)) +
          ((f(k + 1, j, i) * v100 + f(k + 1, j + 1, i) * v110) +
           (f(k + 1, j, i + 1) * v101 + f(k + 1, j + 1, i + 1) * v111))) /
         tvol;
}

// Area-weighted average of the four fine faces of component comp whose lowest corner is
// the fine face (k, j, i).
__device__ __forceinline__ double restrict_fc_at(int comp, const View &f, const RGrid &F, int k,
                                                 int j, int i) {
  if (comp == 0) {
    const double a00 = F.dx2f[j] * F.dx3f[k], a01 = F.dx2f[j + 1] * F.dx3f[k];
    const double a10 = F.dx2f[j] * F.dx3f[k + 1], a11 = F.dx2f[j + 1] * F.dx3f[k + 1];
    const double tarea = a00 + a01 + a10 + a11;
    return (f(k, j, i) * a00 + f(k, j + 1, i) * a01 + f(k + 1, j, i) * a10 +
            f(k + 1, j + 1, i) * a11) / tarea;
  } else if (comp == 1) {
    const double a00 = F.dx1f[i] * F.dx3f[k], a01 = F.dx1f[i + 1] * F.dx3f[k];
    const double a10 = F.dx1f[i] * F.dx3f[k + 1], a11 = F.dx1f[i + 1] * F.dx3f[k + 1];
    const double tarea = a00 + a01 + a10 + a11;
    return (f(k, j, i) * a00 + f(k, j, i + 1) * a01 + f(k + 1, j, i) * a10 +
            f(k + 1, j, i + 1) * a11) / tarea;
  }
  const double a00 = F.dx1f[i] * F.dx2f[j], a01 = F.dx1f[i + 1] * F.dx2f[j];
  const double a10 = F.dx1f[i] * F.dx2f[j + 1], a11 = F.dx1f[i + 1] * F.dx2f[j + 1];
  const double tarea = a00 + a01 + a10 + a11;
  return (f(k, j, i) * a00 + f(k, j, i + 1) * a01 + f(k, j + 1, i) * a10 +
          f(k, j + 1, i + 1) * a11) / tarea;
}

// The eight fine cells of coarse cell (k, j, i): minmod-limited gradients on the coarse
// grid evaluated at the fine centres.
__device__ __forceinline__ void prolong_cc_at(const View &c, const View &f, const RGrid &F,
                                              const RGrid &C, int ng, int k, int j, int i) {
  const int fk = (k - ng) * 2 + ng, fj = (j - ng) * 2 + ng, fi = (i - ng) * 2 + ng;
  const double x3m = C.x3v[k - 1], x3c = C.x3v[k], x3p = C.x3v[k + 1];
  const double dx3m = x3c - x3m, dx3p = x3p - x3c;
  const double dx3fm = x3c - F.x3v[fk], dx3fp = F.x3v[fk + 1] - x3c;
  const double x2m = C.x2v[j - 1], x2c = C.x2v[j], x2p = C.x2v[j + 1];
  const double dx2m = x2c - x2m, dx2p = x2p - x2c;
  const double dx2fm = x2c - F.x2v[fj], dx2fp = F.x2v[fj + 1] - x2c;
  const double x1m = C.x1v[i - 1], x1c = C.x1v[i], x1p = C.x1v[i + 1];
  const double dx1m = x1c - x1m, dx1p = x1p - x1c;
  const double dx1fm = x1c - F.x1v[fi], dx1fp = F.x1v[fi + 1] - x1c;
  const double cc = c(k, j, i);
  const double gx1c = minmod_c((cc - c(k, j, i - 1)) / dx1m, (c(k, j, i + 1) - cc) / dx1p);
  const double gx2c = minmod_c((cc - c(k, j - 1, i)) / dx2m, (c(k, j + 1, i) - cc) / dx2p);
  const double gx3c = minmod_c((cc - c(k - 1, j, i)) / dx3m, (c(k + 1, j, i) - cc) / dx3p);
  f(fk, fj, fi) = cc - (gx1c * dx1fm + gx2c * dx2fm + gx3c * dx3fm);
  f(fk, fj, fi + 1) = cc + (gx1c * dx1fp - gx2c * dx2fm - gx3c * dx3fm);
  f(fk, fj + 1, fi) = cc - (gx1c * dx1fm - gx2c * dx2fp + gx3c * dx3fm);
  f(fk, fj + 1, fi + 1) = cc + (gx1c * dx1fp + gx2c * dx2fp - gx3c * dx3fm);
  f(fk + 1, fj, fi) = cc - (gx1c * dx1fm + gx2c * dx2fm - gx3c * dx3fp);
  f(fk + 1, fj, fi + 1) = cc + (gx1c * dx1fp - gx2c * dx2fm + gx3c * dx3fp);
  f(fk + 1, fj + 1, fi) = cc - (gx1c * dx1fm - gx2c * dx2fp - gx3c * dx3fp);
  f(fk + 1, fj + 1, fi + 1) = cc + (gx1c * dx1fp + gx2c * dx2fp + gx3c * dx3fp);
}

// The four fine faces of component comp on coarse face (k, j, i): the coarse value with
// minmod-limited transverse gradients.
__device__ __forceinline__ void prolong_fc_shared_at(int comp, const View &c, const View &f,
                                                     const RGrid &F, const RGrid &C, int ng,
                                                     int k, int j, int i) {
  const int fk = (k - ng) * 2 + ng, fj = (j - ng) * 2 + ng, fi = (i - ng) * 2 + ng;
  const double cc = c(k, j, i);
  if (comp == 0) {
    const double x3m = C.x3v[k - 1], x3c = C.x3v[k], x3p = C.x3v[k + 1];
    const double dx3m = x3c - x3m, dx3p = x3p - x3c;
    const double fx3m = F.x3v[fk], fx3p = F.x3v[fk + 1];
    const double x2m = C.x2v[j - 1], x2c = C.x2v[j], x2p = C.x2v[j + 1];
    const double dx2m = x2c - x2m, dx2p = x2p - x2c;
    const double fx2m = F.x2v[fj], fx2p = F.x2v[fj + 1];
    const double gx2c = minmod_c((cc - c(k, j - 1, i)) / dx2m, (c(k, j + 1, i) - cc) / dx2p);
    const double gx3c = minmod_c((cc - c(k - 1, j, i)) / dx3m, (c(k + 1, j, i) - cc) / dx3p);
    f(fk, fj, fi) = cc - gx2c * (x2c - fx2m) - gx3c * (x3c - fx3m);
    f(fk, fj + 1, fi) = cc + gx2c * (fx2p - x2c) - gx3c * (x3c - fx3m);
    f(fk + 1, fj, fi) = cc - gx2c * (x2c - fx2m) + gx3c * (fx3p - x3c);
    f(fk + 1, fj + 1, fi) = cc + gx2c * (fx2p - x2c) + gx3c * (fx3p - x3c);
  } else if (comp == 1) {
    const double x3m = C.x3v[k - 1], x3c = C.x3v[k], x3p = C.x3v[k + 1];
    const double dx3m = x3c - x3m, dx3p = x3p - x3c;
    const double fx3m = F.x3v[fk], fx3p = F.x3v[fk + 1];
    const double x1m = C.x1v[i - 1], x1c = C.x1v[i], x1p = C.x1v[i + 1];
    const double dx1m = x1c - x1m, dx1p = x1p - x1c;
    const double fx1m = F.x1v[fi], fx1p = F.x1v[fi + 1];
    const double gx1c = minmod_c((cc - c(k, j, i - 1)) / dx1m, (c(k, j, i + 1) - cc) / dx1p);
    const double gx3c = minmod_c((cc - c(k - 1, j, i)) / dx3m, (c(k + 1, j, i) - cc) / dx3p);
    f(fk, fj, fi) = cc - gx1c * (x1c - fx1m) - gx3c * (x3c - fx3m);
    f(fk, fj, fi + 1) = cc + gx1c * (fx1p - x1c) - gx3c * (x3c - fx3m);
    f(fk + 1, fj, fi) = cc - gx1c * (x1c - fx1m) + gx3c * (fx3p - x3c);
    f(fk + 1, fj, fi + 1) = cc + gx1c * (fx1p - x1c) + gx3c * (fx3p - x3c);
  } else {
    const double x2m = C.x2v[j - 1], x2c = C.x2v[j], x2p = C.x2v[j + 1];
    const double dx2m = x2c - x2m, dx2p = x2p - x2c;
    const double fx2m = F.x2v[fj], fx2p = F.x2v[fj + 1];
    const double x1m = C.x1v[i - 1], x1c = C.x1v[i], x1p = C.x1v[i + 1];
    const double dx1m = x1c - x1m, dx1p = x1p - x1c;
    const double fx1m = F.x1v[fi], fx1p = F.x1v[fi + 1];
    const double gx1c = minmod_c((cc - c(k, j, i - 1)) / dx1m, (c(k, j, i + 1) - cc) / dx1p);
    const double gx2c = minmod_c((cc - c(k, j - 1, i)) / dx2m, (c(k, j + 1, i) - cc) / dx2p);
    f(fk, fj, fi) = cc - gx1c * (x1c - fx1m) - gx2c * (x2c - fx2m);
    f(fk, fj, fi + 1) = cc + gx1c * (fx1p - x1c) - gx2c * (x2c - fx2m);
    f(fk, fj + 1, fi) = cc - gx1c * (x1c - fx1m) + gx2c * (fx2p - x2c);
    f(fk, fj + 1, fi + 1) = cc + gx1c * (fx1p - x1c) + gx2c * (fx2p - x2c);
  }
}

// The twelve fine faces inside coarse cell (k, j, i) from the 24 fine faces on its surface,
// preserving the divergence (Toth & Roe 2002).
__device__ __forceinline__ void prolong_fc_internal_at(const View &X1, const View &X2,
                                                       const View &X3, const RGrid &F, int ng,
                                                       int k, int j, int i) {
  auto A1 = [&](int kk, int jj) { return F.dx2f[jj] * F.dx3f[kk]; };
  auto A2 = [&](int kk, int ii) { return F.dx1f[ii] * F.dx3f[kk]; };
  auto A3 = [&](int jj, int ii) { return F.dx1f[ii] * F.dx2f[jj]; };
  const int fk = (k - ng) * 2 + ng, fj = (j - ng) * 2 + ng, fi = (i - ng) * 2 + ng;
  double Uxx = 0.0, Vyy = 0.0, Wzz = 0.0;
  double Uxyz = 0.0, Vxyz = 0.0, Wxyz = 0.0;
#pragma unroll
  for (int jj = 0; jj < 2; jj++) {
    const int js = 2 * jj - 1, fjj = fj + jj, fjp = fj + 2 * jj;
#pragma unroll
    for (int ii = 0; ii < 2; ii++) {
      const int is = 2 * ii - 1, fii = fi + ii, fip = fi + 2 * ii;
      Uxx += is * (js * (X2(fk, fjp, fii) * A2(fk, fii) + X2(fk + 1, fjp, fii) * A2(fk + 1, fii)) +
                   (X3(fk + 2, fjj, fii) * A3(fjj, fii) - X3(fk, fjj, fii) * A3(fjj, fii)));
      Vyy += js * ((X3(fk + 2, fjj, fii) * A3(fjj, fii) - X3(fk, fjj, fii) * A3(fjj, fii)) +
                   is * (X1(fk, fjj, fip) * A1(fk, fjj) + X1(fk + 1, fjj, fip) * A1(fk + 1, fjj)));
      Wzz += is * (X1(fk + 1, fjj, fip) * A1(fk + 1, fjj) - X1(fk, fjj, fip) * A1(fk, fjj)) +
             js * (X2(fk + 1, fjp, fii) * A2(fk + 1, fii) - X2(fk, fjp, fii) * A2(fk, fii));
      Uxyz += is * js * (X1(fk + 1, fjj, fip) * A1(fk + 1, fjj) - X1(fk, fjj, fip) * A1(fk, fjj));
      Vxyz += is * js * (X2(fk + 1, fjp, fii) * A2(fk + 1, fii) - X2(fk, fjp, fii) * A2(fk, fii));
      Wxyz += is * js * (X3(fk + 2, fjj, fii) * A3(fjj, fii) - X3(fk, fjj, fii) * A3(fjj, fii));
    }
  }
  const double s1 = F.dx1f[fi] + F.dx1f[fi + 1], s2 = F.dx2f[fj] + F.dx2f[fj + 1];
  const double s3 = F.dx3f[fk] + F.dx3f[fk + 1];
  const double Sdx1 = s1 * s1, Sdx2 = s2 * s2, Sdx3 = s3 * s3;
  Uxx *= 0.125; Vyy *= 0.125; Wzz *= 0.125;
  Uxyz *= 0.125 / (Sdx2 + Sdx3);
  Vxyz *= 0.125 / (Sdx1 + Sdx3);
  Wxyz *= 0.125 / (Sdx1 + Sdx2);
#pragma unroll
  for (int kk = 0; kk < 2; ++kk)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const double t3 = kk ? 1.0 : -1.0, t2 = jj ? 1.0 : -1.0;
      X1(fk + kk, fj + jj, fi + 1) =
          (0.5 * (X1(fk + kk, fj + jj, fi) * A1(fk + kk, fj + jj) +
                  X1(fk + kk, fj + jj, fi + 2) * A1(fk + kk, fj + jj)) +
           Uxx + t3 * Sdx3 * Vxyz + t2 * Sdx2 * Wxyz) / A1(fk + kk, fj + jj);
    }
#pragma unroll
  for (int kk = 0; kk < 2; ++kk)
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) {
      const double t3 = kk ? 1.0 : -1.0, t1 = ii ? 1.0 : -1.0;
      X2(fk + kk, fj + 1, fi + ii) =
          (0.5 * (X2(fk + kk, fj, fi + ii) * A2(fk + kk, fi + ii) +
                  X2(fk + kk, fj + 2, fi + ii) * A2(fk + kk, fi + ii)) +
           Vyy + t3 * Sdx3 * Uxyz + t1 * Sdx1 * Wxyz) / A2(fk + kk, fi + ii);
    }
#pragma unroll
  for (int jj = 0; jj < 2; ++jj)
#pragma unroll
    for (int ii = 0; ii < 2; ++ii) {
      const double t2 = jj ? 1.0 : -1.0, t1 = ii ? 1.0 : -1.0;
      X3(fk + 1, fj + jj, fi + ii) =
          (0.5 * (X3(fk + 2, fj + jj, fi + ii) * A3(fj + jj, fi + ii) +
                  X3(fk, fj + jj, fi + ii) * A3(fj + jj, fi + ii)) +
           Wzz + t2 * Sdx2 * Uxyz + t1 * Sdx1 * Vxyz) / A3(fj + jj, fi + ii);
    }
}

// Length-weighted average of the two fine edges (k, j, i) and its successor along comp
// (the EMF restriction of LoadFluxBoundaryBufferToCoarser, flux_correction_fc.cpp:315-573).
__device__ __forceinline__ double restrict_emf_at(int comp, const View &f, const RGrid &F,
                                                  int k, int j, int i) {
  if (comp == 0) {
    const double l1 = F.dx1f[i], l2 = F.dx1f[i + 1];
    return (f(k, j, i) * l1 + f(k, j, i + 1) * l2) / (l1 + l2);
  } else if (comp == 1) {
    const double l1 = F.dx2f[j], l2 = F.dx2f[j + 1];
    return (f(k, j, i) * l1 + f(k, j + 1, i) * l2) / (l1 + l2);
  }
  const double l1 = F.dx3f[k], l2 = F.dx3f[k + 1];
  return (f(k, j, i) * l1 + f(k + 1, j, i) * l2) / (l1 + l2);
}

}  // namespace refops

#endif  // ATHENA_AMD_REFINE_OPS_HPP_
