//========================================================================================
// oracle_core.hpp -- TEST INFRASTRUCTURE ONLY: shared types of the CPU restatement
// (athena_oracle.cpp, athena_oracle_ext.cpp).  AthenaArray-style storage and the
// per-MeshBlock register set of Hydro/Field (hydro.cpp:31-60, field.cpp:27-105).
//========================================================================================
#ifndef ORACLE_CORE_HPP_
#define ORACLE_CORE_HPP_
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace ora {

using Real = double;

constexpr int IDN = 0, IM1 = 1, IM2 = 2, IM3 = 3, IEN = 4;
constexpr int IVX = 1, IVY = 2, IVZ = 3, IPR = 4, IBY = 5, IBZ = 6;
constexpr int IB1 = 0, IB2 = 1, IB3 = 2;
constexpr int NHYDRO = 5, NWAVE = 7;
constexpr Real TINY_NUMBER = 1.0e-20;
constexpr Real ONE_3RD = 0.33333333333333333;   // defs.hpp.in:144

inline Real SQR(Real x) { return x * x; }
// std::min / std::max semantics: min(a,b) = (b<a)?b:a ; max(a,b) = (a<b)?b:a
inline Real rmin(Real a, Real b) { return (b < a) ? b : a; }
inline Real rmax(Real a, Real b) { return (a < b) ? b : a; }

// AthenaArray-style storage, row-major with i fastest (athena_arrays.hpp:140-143)
struct Arr {
  int n4 = 0, n3 = 0, n2 = 0, n1 = 0;
  std::vector<Real> v;
  void alloc(int a, int b, int c, int d) {
    n4 = a; n3 = b; n2 = c; n1 = d;
    v.assign(static_cast<size_t>(a) * b * c * d, 0.0);
  }
  Real &operator()(int n, int k, int j, int i) {
    return v[i + static_cast<size_t>(n1) * (j + static_cast<size_t>(n2) *
                                                   (k + static_cast<size_t>(n3) * n))];
  }
  Real &operator()(int k, int j, int i) { return (*this)(0, k, j, i); }
  void zero() { std::fill(v.begin(), v.end(), 0.0); }
  void swap(Arr &o) { v.swap(o.v); }
};

}  // namespace ora

extern "C" {
// Keep in sync with tests/oracle_bind.py (ctypes mirror).
typedef struct {
  int nx[3];          // mesh cells
  int bnx[3];         // meshblock cells
  int ng;             // NGHOST
  int xorder;         // 1 = DC, 2 = PLM, 3 = PPM
  int rsolver;        // 0 = HLLD, 1 = HLLE
  int relativistic;   // 0 = Newtonian, 1 = special relativistic
  double xmin[3], xmax[3];
  double gamma;
  double dfloor, pfloor;   // <0 selects the reference default sqrt(1024*FLT_MIN)
  double cfl;
  double sigma_max, beta_min, gamma_max;   // SR floors (<0 -> reference defaults)
  int bc[6];          // ix1, ox1, ix2, ox2, ix3, ox3: 0 periodic, 1 outflow
} ora_cfg;
}

namespace ora {

struct Block {
  int gid, lx[3];
  int nx1, nx2, nx3, ng;
  int is, ie, js, je, ks, ke;
  int nc1, nc2, nc3;
  Real xbmin[3], xbmax[3];
  std::vector<Real> x1f, x2f, x3f, x1v, x2v, x3v, dx1f, dx2f, dx3f, dx1v, dx2v, dx3v;
  Arr u, u1, w, w1, bcc;
  Arr b1f, b2f, b3f, bb1f, bb2f, bb3f;  // b and register b1
  Arr flux[3];
  Arr e2x1, e3x1, e1x2, e3x2, e1x3, e2x3;  // face EMFs from the Riemann solver
  Arr wx1, wx2, wx3;                       // CT upwind weights
  Arr e1, e2, e3;                          // corner (edge) EMFs
  Arr cce;                                 // cell-centred E (3, nc3, nc2, nc1)
  int nb[3][3][3];                         // neighbour gid at offset (ox3+1, ox2+1, ox1+1),
                                           // -1 across an outflow mesh boundary
  Real new_dt;
};

struct Mesh {
  ora_cfg c;
  int nbx[3];
  std::vector<Block> blocks;
  Real time = 0.0, dt = std::numeric_limits<Real>::max();
  int ncycle = 0;
  Real dfloor, pfloor, gm1;
  Real sigma_max, beta_min, gamma_max;
};


// Declarations shared across the two translation units.
Real FastMagnetosonicSpeed(const Real prim[NWAVE], Real bx, Real gamma);
void HLLE1(const Real wli[NWAVE], const Real wri[NWAVE], Real bxi, Real gamma,
           Real flxi[NWAVE]);
void RiemannSRHLLD(Mesh &m, const Real wl[NWAVE], const Real wr[NWAVE], Real bbx,
                   Real flxi[NWAVE]);
void RiemannSR(Mesh &m, const Real wli[NWAVE], const Real wri[NWAVE], Real bxi,
               Real flxi[NWAVE]);
void PiecewiseParabolic(Mesh &m, Block &b, int d, int k, int j, int il, int iu, Arr &w,
                        Arr &bcc, Arr &wl, Arr &wr);
void Cons2PrimSR(Mesh &m, Block &b, Arr &cons, Arr &prim_old, Arr &prim, int il, int iu,
                 int jl, int ju, int kl, int ku);
void CellCenteredField(Block &b, int il, int iu, int jl, int ju, int kl, int ku);
void Prim2ConsSRCell(const Real prim[5], const Real bc[3], Real gamma_adi, Real cons[5]);
Real Wval(Block &b, Arr &w, Arr &bcc, int d, int n, int k, int j, int i);
inline Real SIGN(Real x) { return ((x) < 0.0) ? -1.0 : 1.0; }
constexpr Real PI = 3.1415926535897932;

}  // namespace ora
#endif  // ORACLE_CORE_HPP_
