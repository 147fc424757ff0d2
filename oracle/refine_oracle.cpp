//========================================================================================
// refine_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the 3-D, uniform-Cartesian branches of Athena++'s MeshRefinement
// operators (mesh/mesh_refinement.cpp), the checker for the device versions in
// athena_amd/csrc/refine.hip.  Only tests/ may load it.
//
// Layout (the reference's): a fine block has nx cells per direction and ng ghosts
// (nc = nx + 2 ng, active cells is = ng ..); its coarse buffer has nx/2 cells and ng
// ghosts (ncc = nx/2 + 2 ng, cis = ng); fine index = (coarse index - cis) * 2 + is
// (meshblock.cpp:66,93-94).  Cell arrays are (nvar, n3, n2, n1), face component d has
// one more entry along d.  Coordinates per grid: x1f (n1+1), x1v (n1), dx1f (n1), then
// the same for x2 and x3, packed in that order.  Cartesian metrics: volume
// dx1f dx2f dx3f, face areas dx2f dx3f / dx1f dx3f / dx1f dx2f, edge lengths dx2f, dx3f
// (coordinates.cpp:377-545); face-centre positions x?s? = the cell centres
// (cartesian.cpp:105-118), so the shared-face prolongation takes the Cartesian
// (not flux-interpolating) branch (mesh_refinement.cpp:50-60).
//
// Parity pinning (tests/test_refine_oracle.py): restriction of cell averages of a
// linear field and of constants is exact; prolongation of linear data reproduces the
// field at the fine centres; restriction after prolongation is the identity for cells
// and shared faces; the divergence-preserving internal prolongation leaves every fine
// cell with the divergence of its coarse parent and reproduces linear fields.
//========================================================================================
#include <algorithm>
#include <cmath>

namespace {

using Real = double;

inline Real SIGN(Real x) { return (x < 0.0) ? -1.0 : 1.0; }
inline Real SQR(Real x) { return x * x; }

struct Grid {                    // one grid's coordinate tables
  int n1, n2, n3;
  const Real *x1f, *x1v, *dx1f, *x2f, *x2v, *dx2f, *x3f, *x3v, *dx3f;
  Grid(const Real *c, int a, int b, int d) : n1(a), n2(b), n3(d) {
    x1f = c; x1v = x1f + n1 + 1; dx1f = x1v + n1;
    x2f = dx1f + n1; x2v = x2f + n2 + 1; dx2f = x2v + n2;
    x3f = dx2f + n2; x3v = x3f + n3 + 1; dx3f = x3v + n3;
  }
};

inline long coord_len(int n1, int n2, int n3) {
  return 3L * n1 + 1 + 3L * n2 + 1 + 3L * n3 + 1;
}

struct Arr3 {                    // (k, j, i) view
  Real *p;
  int n2, n1;
  Real &operator()(int k, int j, int i) const {
    return p[(static_cast<long>(k) * n2 + j) * n1 + i];
  }
};

struct Shape {
  int nx1, nx2, nx3, ng;
  int nc1, nc2, nc3, ncc1, ncc2, ncc3;
  int is, js, ks, cis, cjs, cks;
  explicit Shape(const int *s) {
    nx1 = s[0]; nx2 = s[1]; nx3 = s[2]; ng = s[3];
    nc1 = nx1 + 2 * ng; nc2 = nx2 + 2 * ng; nc3 = nx3 + 2 * ng;
    ncc1 = nx1 / 2 + 2 * ng; ncc2 = nx2 / 2 + 2 * ng; ncc3 = nx3 / 2 + 2 * ng;
    is = js = ks = ng;
    cis = cjs = cks = ng;
  }
};

}  // namespace

extern "C" {

// RestrictCellCenteredValues, 3-D branch (mesh_refinement.cpp:116-150).
void ora_restrict_cc(const int *shape, int nvar, const double *fcoord, const double *fine,
                     double *coarse, const int *r) {
  const Shape S(shape);
  const Grid F(fcoord, S.nc1, S.nc2, S.nc3);
  const long fn = static_cast<long>(S.nc1) * S.nc2 * S.nc3;
  const long cn = static_cast<long>(S.ncc1) * S.ncc2 * S.ncc3;
  auto vol = [&](int k, int j, int i) { return F.dx1f[i] * F.dx2f[j] * F.dx3f[k]; };
  for (int n = 0; n < nvar; ++n) {
    Arr3 f{const_cast<Real *>(fine) + n * fn, S.nc2, S.nc1};
    Arr3 c{coarse + n * cn, S.ncc2, S.ncc1};
    for (int ck = r[4]; ck <= r[5]; ++ck) {
      const int k = (ck - S.cks) * 2 + S.ks;
      for (int cj = r[2]; cj <= r[3]; ++cj) {
        const int j = (cj - S.cjs) * 2 + S.js;
        for (int ci = r[0]; ci <= r[1]; ++ci) {
          const int i = (ci - S.cis) * 2 + S.is;
          const Real v000 = vol(k, j, i), v010 = vol(k, j + 1, i);
          const Real v001 = vol(k, j, i + 1), v011 = vol(k, j + 1, i + 1);
          const Real v100 = vol(k + 1, j, i), v110 = vol(k + 1, j + 1, i);
          const Real v101 = vol(k + 1, j, i + 1), v111 = vol(k + 1, j + 1, i + 1);
          const Real tvol = ((v000 + v010) + (v001 + v011)) + ((v100 + v110) + (v101 + v111));
          c(ck, cj, ci) = (((f(k, j, i) * v000 + f(k, j + 1, i) * v010) +
                            (f(k, j, i + 1) * v001 + f(k, j + 1, i + 1) * v011)) +
                           ((f(k + 1, j, i) * v100 + f(k + 1, j + 1, i) * v110) +
                            (f(k + 1, j, i + 1) * v101 + f(k + 1, j + 1, i + 1) * v111))) /
                          tvol;
        }
      }
    }
  }
}

// RestrictFieldX1/X2/X3, 3-D branches (mesh_refinement.cpp:272-300, 328-372, 402-430).
// comp 0, 1, 2: the x1, x2, x3 face component; r is the coarse range of that component.
void ora_restrict_fc(const int *shape, int comp, const double *fcoord, const double *fine,
                     double *coarse, const int *r) {
  const Shape S(shape);
  const Grid F(fcoord, S.nc1, S.nc2, S.nc3);
  const int fn1 = S.nc1 + (comp == 0), fn2 = S.nc2 + (comp == 1);
  const int cn1 = S.ncc1 + (comp == 0), cn2 = S.ncc2 + (comp == 1);
  Arr3 f{const_cast<Real *>(fine), fn2, fn1};
  Arr3 c{coarse, cn2, cn1};
  for (int ck = r[4]; ck <= r[5]; ++ck) {
    const int k = (ck - S.cks) * 2 + S.ks;
    for (int cj = r[2]; cj <= r[3]; ++cj) {
      const int j = (cj - S.cjs) * 2 + S.js;
      for (int ci = r[0]; ci <= r[1]; ++ci) {
        const int i = (ci - S.cis) * 2 + S.is;
        if (comp == 0) {
          const Real a00 = F.dx2f[j] * F.dx3f[k], a01 = F.dx2f[j + 1] * F.dx3f[k];
          const Real a10 = F.dx2f[j] * F.dx3f[k + 1], a11 = F.dx2f[j + 1] * F.dx3f[k + 1];
          const Real tarea = a00 + a01 + a10 + a11;
          c(ck, cj, ci) = (f(k, j, i) * a00 + f(k, j + 1, i) * a01 + f(k + 1, j, i) * a10 +
                           f(k + 1, j + 1, i) * a11) / tarea;
        } else if (comp == 1) {
          const Real a00 = F.dx1f[i] * F.dx3f[k], a01 = F.dx1f[i + 1] * F.dx3f[k];
          const Real a10 = F.dx1f[i] * F.dx3f[k + 1], a11 = F.dx1f[i + 1] * F.dx3f[k + 1];
          const Real tarea = a00 + a01 + a10 + a11;
          c(ck, cj, ci) = (f(k, j, i) * a00 + f(k, j, i + 1) * a01 + f(k + 1, j, i) * a10 +
                           f(k + 1, j, i + 1) * a11) / tarea;
        } else {
          const Real a00 = F.dx1f[i] * F.dx2f[j], a01 = F.dx1f[i + 1] * F.dx2f[j];
          const Real a10 = F.dx1f[i] * F.dx2f[j + 1], a11 = F.dx1f[i + 1] * F.dx2f[j + 1];
          const Real tarea = a00 + a01 + a10 + a11;
          c(ck, cj, ci) = (f(k, j, i) * a00 + f(k, j, i + 1) * a01 + f(k, j + 1, i) * a10 +
                           f(k, j + 1, i + 1) * a11) / tarea;
        }
      }
    }
  }
}

// ProlongateCellCenteredValues, 3-D branch (mesh_refinement.cpp:464-525): minmod-limited
// gradients on the coarse grid, evaluated at the eight fine centres.
void ora_prolongate_cc(const int *shape, int nvar, const double *fcoord,
                       const double *ccoord, const double *coarse, double *fine,
                       const int *r) {
  const Shape S(shape);
  const Grid F(fcoord, S.nc1, S.nc2, S.nc3);
  const Grid C(ccoord, S.ncc1, S.ncc2, S.ncc3);
  const long fn = static_cast<long>(S.nc1) * S.nc2 * S.nc3;
  const long cn = static_cast<long>(S.ncc1) * S.ncc2 * S.ncc3;
  for (int n = 0; n < nvar; ++n) {
    Arr3 f{fine + n * fn, S.nc2, S.nc1};
    Arr3 c{const_cast<Real *>(coarse) + n * cn, S.ncc2, S.ncc1};
    for (int k = r[4]; k <= r[5]; ++k) {
      const int fk = (k - S.cks) * 2 + S.ks;
      const Real x3m = C.x3v[k - 1], x3c = C.x3v[k], x3p = C.x3v[k + 1];
      const Real dx3m = x3c - x3m, dx3p = x3p - x3c;
      const Real dx3fm = x3c - F.x3v[fk], dx3fp = F.x3v[fk + 1] - x3c;
      for (int j = r[2]; j <= r[3]; ++j) {
        const int fj = (j - S.cjs) * 2 + S.js;
        const Real x2m = C.x2v[j - 1], x2c = C.x2v[j], x2p = C.x2v[j + 1];
        const Real dx2m = x2c - x2m, dx2p = x2p - x2c;
        const Real dx2fm = x2c - F.x2v[fj], dx2fp = F.x2v[fj + 1] - x2c;
        for (int i = r[0]; i <= r[1]; ++i) {
          const int fi = (i - S.cis) * 2 + S.is;
          const Real x1m = C.x1v[i - 1], x1c = C.x1v[i], x1p = C.x1v[i + 1];
          const Real dx1m = x1c - x1m, dx1p = x1p - x1c;
          const Real dx1fm = x1c - F.x1v[fi], dx1fp = F.x1v[fi + 1] - x1c;
          const Real cc = c(k, j, i);
          const Real gx1m = (cc - c(k, j, i - 1)) / dx1m, gx1p = (c(k, j, i + 1) - cc) / dx1p;
          const Real gx1c = 0.5 * (SIGN(gx1m) + SIGN(gx1p)) * std::min(std::abs(gx1m), std::abs(gx1p));
          const Real gx2m = (cc - c(k, j - 1, i)) / dx2m, gx2p = (c(k, j + 1, i) - cc) / dx2p;
          const Real gx2c = 0.5 * (SIGN(gx2m) + SIGN(gx2p)) * std::min(std::abs(gx2m), std::abs(gx2p));
          const Real gx3m = (cc - c(k - 1, j, i)) / dx3m, gx3p = (c(k + 1, j, i) - cc) / dx3p;
          const Real gx3c = 0.5 * (SIGN(gx3m) + SIGN(gx3p)) * std::min(std::abs(gx3m), std::abs(gx3p));
          f(fk, fj, fi) = cc - (gx1c * dx1fm + gx2c * dx2fm + gx3c * dx3fm);
          f(fk, fj, fi + 1) = cc + (gx1c * dx1fp - gx2c * dx2fm - gx3c * dx3fm);
          f(fk, fj + 1, fi) = cc - (gx1c * dx1fm - gx2c * dx2fp + gx3c * dx3fm);
          f(fk, fj + 1, fi + 1) = cc + (gx1c * dx1fp + gx2c * dx2fp - gx3c * dx3fm);
          f(fk + 1, fj, fi) = cc - (gx1c * dx1fm + gx2c * dx2fm - gx3c * dx3fp);
          f(fk + 1, fj, fi + 1) = cc + (gx1c * dx1fp - gx2c * dx2fm + gx3c * dx3fp);
          f(fk + 1, fj + 1, fi) = cc - (gx1c * dx1fm - gx2c * dx2fp - gx3c * dx3fp);
          f(fk + 1, fj + 1, fi + 1) = cc + (gx1c * dx1fp + gx2c * dx2fp + gx3c * dx3fp);
        }
      }
    }
  }
}

// ProlongateSharedFieldX1/X2/X3, 3-D Cartesian branches (mesh_refinement.cpp:835-872 and
// the corresponding blocks of :914-1056, :1058-1221): the coarse face value with
// minmod-limited transverse gradients at the four fine faces on the coarse face.
void ora_prolongate_fc_shared(const int *shape, int comp, const double *fcoord,
                              const double *ccoord, const double *coarse, double *fine,
                              const int *r) {
  const Shape S(shape);
  const Grid F(fcoord, S.nc1, S.nc2, S.nc3);
  const Grid C(ccoord, S.ncc1, S.ncc2, S.ncc3);
  const int fn1 = S.nc1 + (comp == 0), fn2 = S.nc2 + (comp == 1);
  const int cn1 = S.ncc1 + (comp == 0), cn2 = S.ncc2 + (comp == 1);
  Arr3 f{fine, fn2, fn1};
  Arr3 c{const_cast<Real *>(coarse), cn2, cn1};
  auto mm = [](Real gm, Real gp) {
    return 0.5 * (SIGN(gm) + SIGN(gp)) * std::min(std::abs(gm), std::abs(gp));
  };
  for (int k = r[4]; k <= r[5]; ++k) {
    const int fk = (k - S.cks) * 2 + S.ks;
    for (int j = r[2]; j <= r[3]; ++j) {
      const int fj = (j - S.cjs) * 2 + S.js;
      for (int i = r[0]; i <= r[1]; ++i) {
        const int fi = (i - S.cis) * 2 + S.is;
        const Real cc = c(k, j, i);
        if (comp == 0) {
          const Real x3m = C.x3v[k - 1], x3c = C.x3v[k], x3p = C.x3v[k + 1];
          const Real dx3m = x3c - x3m, dx3p = x3p - x3c;
          const Real fx3m = F.x3v[fk], fx3p = F.x3v[fk + 1];
          const Real x2m = C.x2v[j - 1], x2c = C.x2v[j], x2p = C.x2v[j + 1];
          const Real dx2m = x2c - x2m, dx2p = x2p - x2c;
          const Real fx2m = F.x2v[fj], fx2p = F.x2v[fj + 1];
          const Real gx2c = mm((cc - c(k, j - 1, i)) / dx2m, (c(k, j + 1, i) - cc) / dx2p);
          const Real gx3c = mm((cc - c(k - 1, j, i)) / dx3m, (c(k + 1, j, i) - cc) / dx3p);
          f(fk, fj, fi) = cc - gx2c * (x2c - fx2m) - gx3c * (x3c - fx3m);
          f(fk, fj + 1, fi) = cc + gx2c * (fx2p - x2c) - gx3c * (x3c - fx3m);
          f(fk + 1, fj, fi) = cc - gx2c * (x2c - fx2m) + gx3c * (fx3p - x3c);
          f(fk + 1, fj + 1, fi) = cc + gx2c * (fx2p - x2c) + gx3c * (fx3p - x3c);
        } else if (comp == 1) {
          const Real x3m = C.x3v[k - 1], x3c = C.x3v[k], x3p = C.x3v[k + 1];
          const Real dx3m = x3c - x3m, dx3p = x3p - x3c;
          const Real fx3m = F.x3v[fk], fx3p = F.x3v[fk + 1];
          const Real x1m = C.x1v[i - 1], x1c = C.x1v[i], x1p = C.x1v[i + 1];
          const Real dx1m = x1c - x1m, dx1p = x1p - x1c;
          const Real fx1m = F.x1v[fi], fx1p = F.x1v[fi + 1];
          const Real gx1c = mm((cc - c(k, j, i - 1)) / dx1m, (c(k, j, i + 1) - cc) / dx1p);
          const Real gx3c = mm((cc - c(k - 1, j, i)) / dx3m, (c(k + 1, j, i) - cc) / dx3p);
          f(fk, fj, fi) = cc - gx1c * (x1c - fx1m) - gx3c * (x3c - fx3m);
          f(fk, fj, fi + 1) = cc + gx1c * (fx1p - x1c) - gx3c * (x3c - fx3m);
          f(fk + 1, fj, fi) = cc - gx1c * (x1c - fx1m) + gx3c * (fx3p - x3c);
          f(fk + 1, fj, fi + 1) = cc + gx1c * (fx1p - x1c) + gx3c * (fx3p - x3c);
        } else {
          const Real x2m = C.x2v[j - 1], x2c = C.x2v[j], x2p = C.x2v[j + 1];
          const Real dx2m = x2c - x2m, dx2p = x2p - x2c;
          const Real fx2m = F.x2v[fj], fx2p = F.x2v[fj + 1];
          const Real x1m = C.x1v[i - 1], x1c = C.x1v[i], x1p = C.x1v[i + 1];
          const Real dx1m = x1c - x1m, dx1p = x1p - x1c;
          const Real fx1m = F.x1v[fi], fx1p = F.x1v[fi + 1];
          const Real gx1c = mm((cc - c(k, j, i - 1)) / dx1m, (c(k, j, i + 1) - cc) / dx1p);
          const Real gx2c = mm((cc - c(k, j - 1, i)) / dx2m, (c(k, j + 1, i) - cc) / dx2p);
          f(fk, fj, fi) = cc - gx1c * (x1c - fx1m) - gx2c * (x2c - fx2m);
          f(fk, fj, fi + 1) = cc + gx1c * (fx1p - x1c) - gx2c * (x2c - fx2m);
          f(fk, fj + 1, fi) = cc - gx1c * (x1c - fx1m) + gx2c * (fx2p - x2c);
          f(fk, fj + 1, fi + 1) = cc + gx1c * (fx1p - x1c) + gx2c * (fx2p - x2c);
        }
      }
    }
  }
}

// ProlongateInternalField, 3-D branch (mesh_refinement.cpp:1223-1338): the twelve fine
// faces inside each coarse cell from the 24 fine faces on its surface, preserving the
// divergence (Toth & Roe 2002).  b1, b2, b3: the fine block's x1f, x2f, x3f arrays; r is
// the coarse cell range.
void ora_prolongate_fc_internal(const int *shape, const double *fcoord, double *b1,
                                double *b2, double *b3, const int *r) {
  const Shape S(shape);
  const Grid F(fcoord, S.nc1, S.nc2, S.nc3);
  Arr3 x1f{b1, S.nc2, S.nc1 + 1}, x2f{b2, S.nc2 + 1, S.nc1}, x3f{b3, S.nc2, S.nc1};
  auto A1 = [&](int k, int j, int i) { (void)i; return F.dx2f[j] * F.dx3f[k]; };
  auto A2 = [&](int k, int j, int i) { (void)j; return F.dx1f[i] * F.dx3f[k]; };
  auto A3 = [&](int k, int j, int i) { (void)k; return F.dx1f[i] * F.dx2f[j]; };
  for (int k = r[4]; k <= r[5]; ++k) {
    const int fk = (k - S.cks) * 2 + S.ks;
    for (int j = r[2]; j <= r[3]; ++j) {
      const int fj = (j - S.cjs) * 2 + S.js;
      for (int i = r[0]; i <= r[1]; ++i) {
        const int fi = (i - S.cis) * 2 + S.is;
        Real Uxx = 0.0, Vyy = 0.0, Wzz = 0.0;
        Real Uxyz = 0.0, Vxyz = 0.0, Wxyz = 0.0;
        for (int jj = 0; jj < 2; jj++) {
          const int js = 2 * jj - 1, fjj = fj + jj, fjp = fj + 2 * jj;
          for (int ii = 0; ii < 2; ii++) {
            const int is = 2 * ii - 1, fii = fi + ii, fip = fi + 2 * ii;
            Uxx += is * (js * (x2f(fk, fjp, fii) * A2(fk, fjp, fii) +
                               x2f(fk + 1, fjp, fii) * A2(fk + 1, fjp, fii)) +
                         (x3f(fk + 2, fjj, fii) * A3(fk + 2, fjj, fii) -
                          x3f(fk, fjj, fii) * A3(fk, fjj, fii)));
            Vyy += js * ((x3f(fk + 2, fjj, fii) * A3(fk + 2, fjj, fii) -
                          x3f(fk, fjj, fii) * A3(fk, fjj, fii)) +
                         is * (x1f(fk, fjj, fip) * A1(fk, fjj, fip) +
                               x1f(fk + 1, fjj, fip) * A1(fk + 1, fjj, fip)));
            Wzz += is * (x1f(fk + 1, fjj, fip) * A1(fk + 1, fjj, fip) -
                         x1f(fk, fjj, fip) * A1(fk, fjj, fip)) +
                   js * (x2f(fk + 1, fjp, fii) * A2(fk + 1, fjp, fii) -
                         x2f(fk, fjp, fii) * A2(fk, fjp, fii));
            Uxyz += is * js * (x1f(fk + 1, fjj, fip) * A1(fk + 1, fjj, fip) -
                               x1f(fk, fjj, fip) * A1(fk, fjj, fip));
            Vxyz += is * js * (x2f(fk + 1, fjp, fii) * A2(fk + 1, fjp, fii) -
                               x2f(fk, fjp, fii) * A2(fk, fjp, fii));
            Wxyz += is * js * (x3f(fk + 2, fjj, fii) * A3(fk + 2, fjj, fii) -
                               x3f(fk, fjj, fii) * A3(fk, fjj, fii));
          }
        }
        const Real Sdx1 = SQR(F.dx1f[fi] + F.dx1f[fi + 1]);
        const Real Sdx2 = SQR(F.dx2f[fj] + F.dx2f[fj + 1]);
        const Real Sdx3 = SQR(F.dx3f[fk] + F.dx3f[fk + 1]);
        Uxx *= 0.125; Vyy *= 0.125; Wzz *= 0.125;
        Uxyz *= 0.125 / (Sdx2 + Sdx3);
        Vxyz *= 0.125 / (Sdx1 + Sdx3);
        Wxyz *= 0.125 / (Sdx1 + Sdx2);
        for (int kk = 0; kk < 2; ++kk)
          for (int jj = 0; jj < 2; ++jj) {
            const Real s3 = kk ? 1.0 : -1.0, s2 = jj ? 1.0 : -1.0;
            x1f(fk + kk, fj + jj, fi + 1) =
                (0.5 * (x1f(fk + kk, fj + jj, fi) * A1(fk + kk, fj + jj, fi) +
                        x1f(fk + kk, fj + jj, fi + 2) * A1(fk + kk, fj + jj, fi + 2)) +
                 Uxx + s3 * Sdx3 * Vxyz + s2 * Sdx2 * Wxyz) /
                A1(fk + kk, fj + jj, fi + 1);
          }
        for (int kk = 0; kk < 2; ++kk)
          for (int ii = 0; ii < 2; ++ii) {
            const Real s3 = kk ? 1.0 : -1.0, s1 = ii ? 1.0 : -1.0;
            x2f(fk + kk, fj + 1, fi + ii) =
                (0.5 * (x2f(fk + kk, fj, fi + ii) * A2(fk + kk, fj, fi + ii) +
                        x2f(fk + kk, fj + 2, fi + ii) * A2(fk + kk, fj + 2, fi + ii)) +
                 Vyy + s3 * Sdx3 * Uxyz + s1 * Sdx1 * Wxyz) /
                A2(fk + kk, fj + 1, fi + ii);
          }
        for (int jj = 0; jj < 2; ++jj)
          for (int ii = 0; ii < 2; ++ii) {
            const Real s2 = jj ? 1.0 : -1.0, s1 = ii ? 1.0 : -1.0;
            x3f(fk + 1, fj + jj, fi + ii) =
                (0.5 * (x3f(fk + 2, fj + jj, fi + ii) * A3(fk + 2, fj + jj, fi + ii) +
                        x3f(fk, fj + jj, fi + ii) * A3(fk, fj + jj, fi + ii)) +
                 Wzz + s2 * Sdx2 * Uxyz + s1 * Sdx1 * Vxyz) /
                A3(fk + 1, fj + jj, fi + ii);
          }
      }
    }
  }
}

// The EMF restriction of FaceCenteredBoundaryVariable::LoadFluxBoundaryBufferToCoarser
// (flux_correction_fc.cpp:315-420, 3-D Cartesian): a coarse edge along direction comp is
// the length-weighted average of the two fine edges it contains.  Edge arrays of a grid
// with n cells: along x1 (n3+1, n2+1, n1), along x2 (n3+1, n2, n1+1), along x3
// (n3, n2+1, n1+1).
void ora_restrict_emf(const int *shape, int comp, const double *fcoord, const double *fine,
                      double *coarse, const int *r) {
  const Shape S(shape);
  const Grid F(fcoord, S.nc1, S.nc2, S.nc3);
  const int fn1 = S.nc1 + (comp != 0), fn2 = S.nc2 + (comp != 1);
  const int cn1 = S.ncc1 + (comp != 0), cn2 = S.ncc2 + (comp != 1);
  Arr3 f{const_cast<Real *>(fine), fn2, fn1};
  Arr3 c{coarse, cn2, cn1};
  for (int ck = r[4]; ck <= r[5]; ++ck) {
    const int k = (ck - S.cks) * 2 + S.ks;
    for (int cj = r[2]; cj <= r[3]; ++cj) {
      const int j = (cj - S.cjs) * 2 + S.js;
      for (int ci = r[0]; ci <= r[1]; ++ci) {
        const int i = (ci - S.cis) * 2 + S.is;
        if (comp == 0) {
          const Real l1 = F.dx1f[i], l2 = F.dx1f[i + 1];
          c(ck, cj, ci) = (f(k, j, i) * l1 + f(k, j, i + 1) * l2) / (l1 + l2);
        } else if (comp == 1) {
          const Real l1 = F.dx2f[j], l2 = F.dx2f[j + 1];
          c(ck, cj, ci) = (f(k, j, i) * l1 + f(k, j + 1, i) * l2) / (l1 + l2);
        } else {
          const Real l1 = F.dx3f[k], l2 = F.dx3f[k + 1];
          c(ck, cj, ci) = (f(k, j, i) * l1 + f(k + 1, j, i) * l2) / (l1 + l2);
        }
      }
    }
  }
}

}  // extern "C"
