// physics.hpp -- device-side MHD physics for the MI355X per-MeshBlock update.
//
// Single-interface Riemann solvers, fast speeds, the CT upwind weight and SR
// primitive recovery.  Every expression keeps the association order of the Athena++
// routine it replaces (cited per function) and the library is compiled with
// -ffp-contract=off, so results agree bit for bit with the reference's g++ -O3 build
// wherever only +,-,*,/ and sqrt are involved.
//
// State vectors are "rotated": index IVX is the interface-normal velocity, IBY/IBZ are
// the transverse field components, exactly as the reference's wli[]/wri[] arrays.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace athgpu {

// Sets the text ath_gpu_last_error returns and returns 1 (defined in athena_gpu.hip).
int set_error(const std::string &msg);

constexpr int IDN = 0, IM1 = 1, IM2 = 2, IM3 = 3, IEN = 4;
constexpr int IVX = 1, IVY = 2, IVZ = 3, IPR = 4, IBY = 5, IBZ = 6;
constexpr int IB1 = 0, IB2 = 1, IB3 = 2;
constexpr int NHYDRO = 5, NWAVE = 7;

// std::min/std::max semantics of the reference (NaN-propagation identical)
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double sq(double x) { return x * x; }
__device__ __forceinline__ double sgn(double x) { return (x < 0.0) ? -1.0 : 1.0; }

struct EosK {
  double gamma;   // adiabatic index
  double igm1;    // 1/(gamma-1)
  double dfloor, pfloor;
  double sigma_max, beta_min, gamma_max;   // SR only
};

// EquationOfState::FastMagnetosonicSpeed (adiabatic_mhd.cpp:149-156)
__device__ __forceinline__ double fast_speed(const double *w, double bx, double gamma) {
  double asq = gamma * w[IPR];
  double vaxsq = bx * bx;
  double ct2 = (w[IBY] * w[IBY] + w[IBZ] * w[IBZ]);
  double qsq = vaxsq + ct2 + asq;
  double tmp = vaxsq + ct2 - asq;
  return sqrt(0.5 * (qsq + sqrt(tmp * tmp + 4.0 * asq * ct2)) / w[IDN]);
}

// Hydro::GetWeightForCT (hydro.cpp:154-158)
__device__ __forceinline__ double ct_weight(double dflx, double rhol, double rhor,
                                            double dx, double dt) {
  double v_over_c = (1024.0) * dt * dflx / (dx * (rhol + rhor));
  double tmp_min = dmin(0.5, v_over_c);
  return 0.5 + dmax(-0.5, tmp_min);
}

//----------------------------------------------------------------------------------------
// HLLD (Miyoshi & Kusano 2005), hlld.cpp:36-376.  f[7] receives the rotated flux.
struct Cons7 { double d, mx, my, mz, e, by, bz; };

__device__ __forceinline__ void hlld(const double *wl, const double *wr, double bx,
                                     const EosK &eos, double *f) {
  const double small = 1.0e-4;
  const double igm1 = eos.igm1;
  double bxsq = bx * bx;
  double pbl = 0.5 * (bxsq + (sq(wl[IBY]) + sq(wl[IBZ])));
  double pbr = 0.5 * (bxsq + (sq(wr[IBY]) + sq(wr[IBZ])));
  double kel = 0.5 * wl[IDN] * (sq(wl[IVX]) + (sq(wl[IVY]) + sq(wl[IVZ])));
  double ker = 0.5 * wr[IDN] * (sq(wr[IVX]) + (sq(wr[IVY]) + sq(wr[IVZ])));

  Cons7 ul, ur;
  ul.d = wl[IDN]; ul.mx = wl[IVX] * ul.d; ul.my = wl[IVY] * ul.d; ul.mz = wl[IVZ] * ul.d;
  ul.e = wl[IPR] * igm1 + kel + pbl; ul.by = wl[IBY]; ul.bz = wl[IBZ];
  ur.d = wr[IDN]; ur.mx = wr[IVX] * ur.d; ur.my = wr[IVY] * ur.d; ur.mz = wr[IVZ] * ur.d;
  ur.e = wr[IPR] * igm1 + ker + pbr; ur.by = wr[IBY]; ur.bz = wr[IBZ];

  double cfl = fast_speed(wl, bx, eos.gamma);
  double cfr = fast_speed(wr, bx, eos.gamma);
  double s0 = dmin(wl[IVX] - cfl, wr[IVX] - cfr);
  double s4 = dmax(wl[IVX] + cfl, wr[IVX] + cfr);

  double ptl = wl[IPR] + pbl;
  double ptr = wr[IPR] + pbr;

  // side fluxes F_L, F_R
  Cons7 fl, fr;
  fl.d = ul.mx;
  fl.mx = ul.mx * wl[IVX] + ptl - bxsq;
  fl.my = ul.my * wl[IVX] - bx * ul.by;
  fl.mz = ul.mz * wl[IVX] - bx * ul.bz;
  fl.e = wl[IVX] * (ul.e + ptl - bxsq) - bx * (wl[IVY] * ul.by + wl[IVZ] * ul.bz);
  fl.by = ul.by * wl[IVX] - bx * wl[IVY];
  fl.bz = ul.bz * wl[IVX] - bx * wl[IVZ];
  fr.d = ur.mx;
  fr.mx = ur.mx * wr[IVX] + ptr - bxsq;
  fr.my = ur.my * wr[IVX] - bx * ur.by;
  fr.mz = ur.mz * wr[IVX] - bx * ur.bz;
  fr.e = wr[IVX] * (ur.e + ptr - bxsq) - bx * (wr[IVY] * ur.by + wr[IVZ] * ur.bz);
  fr.by = ur.by * wr[IVX] - bx * wr[IVY];
  fr.bz = ur.bz * wr[IVX] - bx * wr[IVZ];

  // supersonic cases need nothing else
  if (s0 >= 0.0) {
    f[IDN] = fl.d; f[IVX] = fl.mx; f[IVY] = fl.my; f[IVZ] = fl.mz;
    f[IEN] = fl.e; f[IBY] = fl.by; f[IBZ] = fl.bz;
    return;
  }
  if (s4 <= 0.0) {
    f[IDN] = fr.d; f[IVX] = fr.mx; f[IVY] = fr.my; f[IVZ] = fr.mz;
    f[IEN] = fr.e; f[IBY] = fr.by; f[IBZ] = fr.bz;
    return;
  }

  double sdl = s0 - wl[IVX];
  double sdr = s4 - wr[IVX];
  double s2 = (sdr * ur.mx - sdl * ul.mx + (ptl - ptr)) / (sdr * ur.d - sdl * ul.d);
  double sdml = s0 - s2;
  double sdmr = s4 - s2;
  double sdml_inv = 1.0 / sdml;
  double sdmr_inv = 1.0 / sdmr;

  Cons7 lst, rst;   // U*_L, U*_R
  lst.d = ul.d * sdl * sdml_inv;
  rst.d = ur.d * sdr * sdmr_inv;
  double lst_dinv = 1.0 / lst.d;
  double rst_dinv = 1.0 / rst.d;
  double sqrtdl = sqrt(lst.d);
  double sqrtdr = sqrt(rst.d);
  double s1 = s2 - fabs(bx) / sqrtdl;
  double s3 = s2 + fabs(bx) / sqrtdr;

  double ptstl = ptl + ul.d * sdl * (s2 - wl[IVX]);
  double ptstr = ptr + ur.d * sdr * (s2 - wr[IVX]);
  double ptst = 0.5 * (ptstr + ptstl);

  lst.mx = lst.d * s2;
  if (fabs(ul.d * sdl * sdml - bxsq) < (small) * ptst) {
    lst.my = lst.d * wl[IVY]; lst.mz = lst.d * wl[IVZ];
    lst.by = ul.by; lst.bz = ul.bz;
  } else {
    double t = bx * (sdl - sdml) / (ul.d * sdl * sdml - bxsq);
    lst.my = lst.d * (wl[IVY] - ul.by * t);
    lst.mz = lst.d * (wl[IVZ] - ul.bz * t);
    t = (ul.d * sq(sdl) - bxsq) / (ul.d * sdl * sdml - bxsq);
    lst.by = ul.by * t; lst.bz = ul.bz * t;
  }
  double vbstl = (lst.mx * bx + (lst.my * lst.by + lst.mz * lst.bz)) * lst_dinv;
  lst.e = (sdl * ul.e - ptl * wl[IVX] + ptst * s2 +
           bx * (wl[IVX] * bx + (wl[IVY] * ul.by + wl[IVZ] * ul.bz) - vbstl)) * sdml_inv;

  rst.mx = rst.d * s2;
  if (fabs(ur.d * sdr * sdmr - bxsq) < (small) * ptst) {
    rst.my = rst.d * wr[IVY]; rst.mz = rst.d * wr[IVZ];
    rst.by = ur.by; rst.bz = ur.bz;
  } else {
    double t = bx * (sdr - sdmr) / (ur.d * sdr * sdmr - bxsq);
    rst.my = rst.d * (wr[IVY] - ur.by * t);
    rst.mz = rst.d * (wr[IVZ] - ur.bz * t);
    t = (ur.d * sq(sdr) - bxsq) / (ur.d * sdr * sdmr - bxsq);
    rst.by = ur.by * t; rst.bz = ur.bz * t;
  }
  double vbstr = (rst.mx * bx + (rst.my * rst.by + rst.mz * rst.bz)) * rst_dinv;
  rst.e = (sdr * ur.e - ptr * wr[IVX] + ptst * s2 +
           bx * (wr[IVX] * bx + (wr[IVY] * ur.by + wr[IVZ] * ur.bz) - vbstr)) * sdmr_inv;

  // single-star fluxes F*_L = F_L + S_L (U*_L - U_L) and F*_R likewise
  if (s1 >= 0.0) {
    f[IDN] = fl.d + s0 * (lst.d - ul.d);
    f[IVX] = fl.mx + s0 * (lst.mx - ul.mx);
    f[IVY] = fl.my + s0 * (lst.my - ul.my);
    f[IVZ] = fl.mz + s0 * (lst.mz - ul.mz);
    f[IEN] = fl.e + s0 * (lst.e - ul.e);
    f[IBY] = fl.by + s0 * (lst.by - ul.by);
    f[IBZ] = fl.bz + s0 * (lst.bz - ul.bz);
    return;
  }
  if (s3 <= 0.0) {
    f[IDN] = fr.d + s4 * (rst.d - ur.d);
    f[IVX] = fr.mx + s4 * (rst.mx - ur.mx);
    f[IVY] = fr.my + s4 * (rst.my - ur.my);
    f[IVZ] = fr.mz + s4 * (rst.mz - ur.mz);
    f[IEN] = fr.e + s4 * (rst.e - ur.e);
    f[IBY] = fr.by + s4 * (rst.by - ur.by);
    f[IBZ] = fr.bz + s4 * (rst.bz - ur.bz);
    return;
  }

  // double-star states (M&K eqs. 59-63)
  double invsumd = 1.0 / (sqrtdl + sqrtdr);
  double bxsig = (bx > 0.0 ? 1.0 : -1.0);
  Cons7 ldst, rdst;
  ldst.d = lst.d; rdst.d = rst.d;
  ldst.mx = lst.mx; rdst.mx = rst.mx;
  double t = invsumd * (sqrtdl * (lst.my * lst_dinv) + sqrtdr * (rst.my * rst_dinv) +
                        bxsig * (rst.by - lst.by));
  ldst.my = ldst.d * t; rdst.my = rdst.d * t;
  t = invsumd * (sqrtdl * (lst.mz * lst_dinv) + sqrtdr * (rst.mz * rst_dinv) +
                 bxsig * (rst.bz - lst.bz));
  ldst.mz = ldst.d * t; rdst.mz = rdst.d * t;
  t = invsumd * (sqrtdl * rst.by + sqrtdr * lst.by +
                 bxsig * sqrtdl * sqrtdr * ((rst.my * rst_dinv) - (lst.my * lst_dinv)));
  ldst.by = rdst.by = t;
  t = invsumd * (sqrtdl * rst.bz + sqrtdr * lst.bz +
                 bxsig * sqrtdl * sqrtdr * ((rst.mz * rst_dinv) - (lst.mz * lst_dinv)));
  ldst.bz = rdst.bz = t;
  t = s2 * bx + (ldst.my * ldst.by + ldst.mz * ldst.bz) / ldst.d;
  ldst.e = lst.e - sqrtdl * bxsig * (vbstl - t);
  rdst.e = rst.e + sqrtdr * bxsig * (vbstr - t);

  if (s2 >= 0.0) {   // F**_L = F_L + S_L(U*_L-U_L) + S*_L(U**_L-U*_L)
    f[IDN] = fl.d + s0 * (lst.d - ul.d) + s1 * (ldst.d - lst.d);
    f[IVX] = fl.mx + s0 * (lst.mx - ul.mx) + s1 * (ldst.mx - lst.mx);
    f[IVY] = fl.my + s0 * (lst.my - ul.my) + s1 * (ldst.my - lst.my);
    f[IVZ] = fl.mz + s0 * (lst.mz - ul.mz) + s1 * (ldst.mz - lst.mz);
    f[IEN] = fl.e + s0 * (lst.e - ul.e) + s1 * (ldst.e - lst.e);
    f[IBY] = fl.by + s0 * (lst.by - ul.by) + s1 * (ldst.by - lst.by);
    f[IBZ] = fl.bz + s0 * (lst.bz - ul.bz) + s1 * (ldst.bz - lst.bz);
  } else {
    f[IDN] = fr.d + s4 * (rst.d - ur.d) + s3 * (rdst.d - rst.d);
    f[IVX] = fr.mx + s4 * (rst.mx - ur.mx) + s3 * (rdst.mx - rst.mx);
    f[IVY] = fr.my + s4 * (rst.my - ur.my) + s3 * (rdst.my - rst.my);
    f[IVZ] = fr.mz + s4 * (rst.mz - ur.mz) + s3 * (rdst.mz - rst.mz);
    f[IEN] = fr.e + s4 * (rst.e - ur.e) + s3 * (rdst.e - rst.e);
    f[IBY] = fr.by + s4 * (rst.by - ur.by) + s3 * (rdst.by - rst.by);
    f[IBZ] = fr.bz + s4 * (rst.bz - ur.bz) + s3 * (rdst.bz - rst.bz);
  }
}

//----------------------------------------------------------------------------------------
// HLLE for adiabatic MHD, hlle_mhd.cpp:25-182
__device__ __forceinline__ void hlle(const double *wl, const double *wr, double bx,
                                     const EosK &eos, double *f) {
  const double gm1 = eos.gamma - 1.0;
  double sqrtdl = sqrt(wl[IDN]);
  double sqrtdr = sqrt(wr[IDN]);
  double isdlpdr = 1.0 / (sqrtdl + sqrtdr);
  double rd = sqrtdl * sqrtdr;
  double rvx = (sqrtdl * wl[IVX] + sqrtdr * wr[IVX]) * isdlpdr;
  double rvy = (sqrtdl * wl[IVY] + sqrtdr * wr[IVY]) * isdlpdr;
  double rvz = (sqrtdl * wl[IVZ] + sqrtdr * wr[IVZ]) * isdlpdr;
  double rby = (sqrtdr * wl[IBY] + sqrtdl * wr[IBY]) * isdlpdr;
  double rbz = (sqrtdr * wl[IBZ] + sqrtdl * wr[IBZ]) * isdlpdr;
  double x = 0.5 * (sq(wl[IBY] - wr[IBY]) + sq(wl[IBZ] - wr[IBZ])) / (sq(sqrtdl + sqrtdr));
  double y = 0.5 * (wl[IDN] + wr[IDN]) / rd;
  double pbl = 0.5 * (bx * bx + sq(wl[IBY]) + sq(wl[IBZ]));
  double pbr = 0.5 * (bx * bx + sq(wr[IBY]) + sq(wr[IBZ]));
  double el = wl[IPR] / gm1 + 0.5 * wl[IDN] * (sq(wl[IVX]) + sq(wl[IVY]) + sq(wl[IVZ])) + pbl;
  double er = wr[IPR] / gm1 + 0.5 * wr[IDN] * (sq(wr[IVX]) + sq(wr[IVY]) + sq(wr[IVZ])) + pbr;
  double hroe = ((el + wl[IPR] + pbl) / sqrtdl + (er + wr[IPR] + pbr) / sqrtdr) * isdlpdr;
  double cl = fast_speed(wl, bx, eos.gamma);
  double cr = fast_speed(wr, bx, eos.gamma);
  double btsq = sq(rby) + sq(rbz);
  double vaxsq = bx * bx / rd;
  double bt_starsq = (gm1 - (gm1 - 1.0) * y) * btsq;
  double hp = hroe - (vaxsq + btsq / rd);
  double vsq = sq(rvx) + sq(rvy) + sq(rvz);
  double twid_asq = dmax((gm1 * (hp - 0.5 * vsq) - (gm1 - 1.0) * x), 0.0);
  double ct2 = bt_starsq / rd;
  double tsum = vaxsq + ct2 + twid_asq;
  double tdif = vaxsq + ct2 - twid_asq;
  double cf2_cs2 = sqrt(tdif * tdif + 4.0 * twid_asq * ct2);
  double cfsq = 0.5 * (tsum + cf2_cs2);
  double a = sqrt(cfsq);
  double al = dmin((rvx - a), (wl[IVX] - cl));
  double ar = dmax((rvx + a), (wr[IVX] + cr));
  double bp = ar > 0.0 ? ar : 0.0;
  double bm = al < 0.0 ? al : 0.0;
  double vxl = wl[IVX] - bm;
  double vxr = wr[IVX] - bp;
  double fl[NWAVE], fr[NWAVE];
  fl[IDN] = wl[IDN] * vxl;
  fr[IDN] = wr[IDN] * vxr;
  fl[IVX] = wl[IDN] * wl[IVX] * vxl + pbl - sq(bx);
  fr[IVX] = wr[IDN] * wr[IVX] * vxr + pbr - sq(bx);
  fl[IVY] = wl[IDN] * wl[IVY] * vxl - bx * wl[IBY];
  fr[IVY] = wr[IDN] * wr[IVY] * vxr - bx * wr[IBY];
  fl[IVZ] = wl[IDN] * wl[IVZ] * vxl - bx * wl[IBZ];
  fr[IVZ] = wr[IDN] * wr[IVZ] * vxr - bx * wr[IBZ];
  fl[IVX] += wl[IPR];
  fr[IVX] += wr[IPR];
  fl[IEN] = el * vxl + wl[IVX] * (wl[IPR] + pbl - bx * bx);
  fr[IEN] = er * vxr + wr[IVX] * (wr[IPR] + pbr - bx * bx);
  fl[IEN] -= bx * (wl[IBY] * wl[IVY] + wl[IBZ] * wl[IVZ]);
  fr[IEN] -= bx * (wr[IBY] * wr[IVY] + wr[IBZ] * wr[IVZ]);
  fl[IBY] = wl[IBY] * vxl - bx * wl[IVY];
  fr[IBY] = wr[IBY] * vxr - bx * wr[IVY];
  fl[IBZ] = wl[IBZ] * vxl - bx * wl[IVZ];
  fr[IBZ] = wr[IBZ] * vxr - bx * wr[IVZ];
  double t = 0.0;
  if (bp != bm) t = 0.5 * (bp + bm) / (bp - bm);
#pragma unroll
  for (int n = 0; n < NWAVE; ++n) f[n] = 0.5 * (fl[n] + fr[n]) + (fl[n] - fr[n]) * t;
}

//----------------------------------------------------------------------------------------
// Special relativity.  Flat-space fast magnetosonic speeds,
// EquationOfState::FastMagnetosonicSpeedsGR with g00=-1, g01=0, g11=1
// (adiabatic_mhd_sr.cpp:350-370).
__device__ __forceinline__ void sr_fast_speeds(double gamma_adi, double wgas, double pgas,
                                               double u0, double u1, double b_sq,
                                               double *lp, double *lm) {
  const double g00 = -1.0, g01 = 0.0, g11 = 1.0;
  double cs_sq = gamma_adi * pgas / wgas;
  double va_sq = b_sq / (b_sq + wgas);
  double cms_sq = cs_sq + va_sq - cs_sq * va_sq;
  double a = sq(u0) - (g00 + sq(u0)) * cms_sq;
  double b = -2.0 * (u0 * u1 - (g01 + u0 * u1) * cms_sq);
  double c = sq(u1) - (g11 + sq(u1)) * cms_sq;
  double a1 = b / a;
  double a0 = c / a;
  double s = sqrt(dmax(sq(a1) - 4.0 * a0, 0.0));
  *lp = (a1 >= 0.0) ? -2.0 * a0 / (a1 + s) : (-a1 + s) / 2.0;
  *lm = (a1 >= 0.0) ? (-a1 - s) / 2.0 : -2.0 * a0 / (a1 - s);
}

// One side of the SR HLLE: conserved state and flux in the rotated frame
// (hlle_mhd_rel.cpp:170-270)
__device__ __forceinline__ void sr_side(const double *w, double bb1, double gamma_adi,
                                        double *cons, double *flux, double *lp,
                                        double *lm) {
  double rho = w[IDN], pgas = w[IPR];
  double u0 = sqrt(1.0 + sq(w[IVX]) + sq(w[IVY]) + sq(w[IVZ]));
  double u1 = w[IVX], u2 = w[IVY], u3 = w[IVZ];
  double bb2 = w[IBY], bb3 = w[IBZ];
  double b0 = bb1 * u1 + bb2 * u2 + bb3 * u3;
  double b1 = (bb1 + b0 * u1) / u0;
  double b2 = (bb2 + b0 * u2) / u0;
  double b3 = (bb3 + b0 * u3) / u0;
  double b_sq = -sq(b0) + sq(b1) + sq(b2) + sq(b3);
  double wgas = rho + gamma_adi / (gamma_adi - 1.0) * pgas;
  sr_fast_speeds(gamma_adi, wgas, pgas, u0, u1, b_sq, lp, lm);
  double wtot = wgas + b_sq;
  double ptot = pgas + 0.5 * b_sq;
  cons[IDN] = rho * u0;
  cons[IEN] = wtot * u0 * u0 - b0 * b0 - ptot;
  cons[IVX] = wtot * u1 * u0 - b1 * b0;
  cons[IVY] = wtot * u2 * u0 - b2 * b0;
  cons[IVZ] = wtot * u3 * u0 - b3 * b0;
  cons[IBY] = b2 * u0 - b0 * u2;
  cons[IBZ] = b3 * u0 - b0 * u3;
  flux[IDN] = rho * u1;
  flux[IEN] = wtot * u0 * u1 - b0 * b1;
  flux[IVX] = wtot * u1 * u1 - b1 * b1 + ptot;
  flux[IVY] = wtot * u2 * u1 - b2 * b1;
  flux[IVZ] = wtot * u3 * u1 - b3 * b1;
  flux[IBY] = b2 * u1 - b1 * u2;
  flux[IBZ] = b3 * u1 - b1 * u3;
}

// SR HLLE, HLLETransforming (hlle_mhd_rel.cpp:103-310) with v_interface = 0
__device__ __forceinline__ void sr_hlle(const double *wl, const double *wr, double bx,
                                        const EosK &eos, double *f) {
  double cl[NWAVE], fl[NWAVE], cr[NWAVE], fr[NWAVE];
  double lpl, lml, lpr, lmr;
  sr_side(wl, bx, eos.gamma, cl, fl, &lpl, &lml);
  sr_side(wr, bx, eos.gamma, cr, fr, &lpr, &lmr);
  double lambda_l = dmin(lml, lmr);
  double lambda_r = dmax(lpl, lpr);
  if (lambda_l >= 0.0) {
#pragma unroll
    for (int n = 0; n < NWAVE; ++n) f[n] = fl[n];
  } else if (lambda_r <= 0.0) {
#pragma unroll
    for (int n = 0; n < NWAVE; ++n) f[n] = fr[n];
  } else {
    double inv = 1.0 / (lambda_r - lambda_l);
#pragma unroll
    for (int n = 0; n < NWAVE; ++n)
      f[n] = (lambda_r * fl[n] - lambda_l * fr[n] + lambda_l * lambda_r * (cr[n] - cl[n])) * inv;
  }
}

//----------------------------------------------------------------------------------------
// SR HLLD (Mignone, Ugliano & Bodo 2009; hlld_rel.cpp:116-927, special relativity,
// v_interface = 0), one interface in the rotated frame.  The total pressure of the
// contact is found by a secant iteration on the jump residual, started from an HLL
// estimate; any non-finite or non-positive intermediate falls back to the HLL flux.

// Alfven-region state for a trial total pressure (MUB 21-35, 43); side -1 left, +1 right.
struct SrAlf {
  double vx, vy, vz, by, bz, eta, kx, ky, kz;
};

__device__ __forceinline__ SrAlf sr_alfven(const double *r, double lambda, double bbx,
                                           double ptot, int side) {
  SrAlf a;
  const double aa = r[IVX] - lambda * r[IEN] + ptot * (1.0 - sq(lambda));
  const double gg = sq(r[IBY]) + sq(r[IBZ]);
  const double cc = r[IVY] * r[IBY] + r[IVZ] * r[IBZ];
  const double qq = -aa - gg + sq(bbx) * (1.0 - sq(lambda));
  const double x_inv = 1.0 / (bbx * (aa * lambda * bbx + cc) - (aa + gg) * (lambda * ptot + r[IEN]));
  a.vx = (bbx * (aa * bbx + lambda * cc) - (aa + gg) * (ptot + r[IVX])) * x_inv;
  a.vy = (qq * r[IVY] + r[IBY] * (cc + bbx * (lambda * r[IVX] - r[IEN]))) * x_inv;
  a.vz = (qq * r[IVZ] + r[IBZ] * (cc + bbx * (lambda * r[IVX] - r[IEN]))) * x_inv;
  a.by = (r[IBY] - bbx * a.vy) / (lambda - a.vx);
  a.bz = (r[IBZ] - bbx * a.vz) / (lambda - a.vx);
  const double v_rm = a.vx * r[IVX] + a.vy * r[IVY] + a.vz * r[IVZ];
  const double wtot = ptot + (r[IEN] - v_rm) / (lambda - a.vx);
  a.eta = (side < 0) ? -copysign(sqrt(wtot), bbx) : copysign(sqrt(wtot), bbx);
  const double den_inv = 1.0 / (lambda * ptot + r[IEN] + bbx * a.eta);
  a.kx = (r[IVX] + ptot + lambda * bbx * a.eta) * den_inv;
  a.ky = (r[IVY] + r[IBY] * a.eta) * den_inv;
  a.kz = (r[IVZ] + r[IBZ] * a.eta) * den_inv;
  return a;
}

// Jump residual across the contact for a trial pressure (MUB 45, 48, 49).
__device__ __forceinline__ double sr_contact_res(const SrAlf &l, const SrAlf &rr, double bbx) {
  const double delta_kx = rr.kx - l.kx + 1.0e-12;
  const double bbx_c = bbx * delta_kx;
  const double by_c = rr.by * (rr.kx - rr.vx) - l.by * (l.kx - l.vx) + bbx * (rr.vy - l.vy);
  const double bz_c = rr.bz * (rr.kx - rr.vx) - l.bz * (l.kx - l.vx) + bbx * (rr.vz - l.vz);
  const double k_sq_l = sq(l.kx) + sq(l.ky) + sq(l.kz);
  double kdb = l.kx * bbx_c + l.ky * by_c + l.kz * bz_c;
  const double y_l = (1.0 - k_sq_l) / (l.eta * delta_kx - kdb);
  const double k_sq_r = sq(rr.kx) + sq(rr.ky) + sq(rr.kz);
  kdb = rr.kx * bbx_c + rr.ky * by_c + rr.kz * bz_c;
  const double y_r = (1.0 - k_sq_r) / (rr.eta * delta_kx - kdb);
  return delta_kx * (1.0 - bbx * (y_r - y_l));
}

// Energy residual of a trial total enthalpy and its derivative (Mignone & McKinney 2007
// A1-A18; hlld_rel.cpp:945-999).
__device__ __forceinline__ double sr_e_res(double w, double dd, double ee, double m_sq,
                                           double bb_sq, double ss_sq, double gp) {
  const double v_sq = (m_sq + ss_sq / sq(w) * (2.0 * w + bb_sq)) / sq(w + bb_sq);
  const double gamma_sq = 1.0 / (1.0 - v_sq);
  const double gl = sqrt(gamma_sq);
  const double chi = (1.0 - v_sq) * (w - gl * dd);
  const double pgas = chi / gp;
  const double ee_calc = w - pgas + 0.5 * bb_sq * (1.0 + v_sq) - ss_sq / (2.0 * sq(w));
  return ee_calc - ee;
}
__device__ __forceinline__ double sr_e_res_prime(double w, double dd, double m_sq,
                                                 double bb_sq, double ss_sq, double gp) {
  const double v_sq = (m_sq + ss_sq / sq(w) * (2.0 * w + bb_sq)) / sq(w + bb_sq);
  const double gamma_sq = 1.0 / (1.0 - v_sq);
  const double gl = sqrt(gamma_sq);
  const double chi = (1.0 - v_sq) * (w - gl * dd);
  const double w_cu = sq(w) * w;
  const double w_b_cu = sq(w + bb_sq) * (w + bb_sq);
  const double dv_sq_dw = -2.0 / (w_cu * w_b_cu) *
                          (ss_sq * (3.0 * w * (w + bb_sq) + sq(bb_sq)) + m_sq * w_cu);
  const double dchi_dw = 1.0 - v_sq - gl / 2.0 * (dd + 2.0 * gl * chi) * dv_sq_dw;
  const double drho_dw = -gl * dd / 2.0 * dv_sq_dw;
  const double dpgas_dw = 1.0 / gp * dchi_dw + 0.0 * drho_dw;
  return 1.0 - dpgas_dw + 0.5 * bb_sq * dv_sq_dw + ss_sq / w_cu;
}

__device__ inline void sr_hlld(const double *wl, const double *wr, double bbx,
                               const EosK &eos, double *f) {
  const double gamma_adi = eos.gamma;
  const double gp = gamma_adi / (gamma_adi - 1.0);
  double cons_l[NWAVE], flux_l[NWAVE], cons_r[NWAVE], flux_r[NWAVE];
  double lpl, lml, lpr, lmr;
  sr_side(wl, bbx, gamma_adi, cons_l, flux_l, &lpl, &lml);
  sr_side(wr, bbx, gamma_adi, cons_r, flux_r, &lpr, &lmr);
  const double lambda_l = dmin(lml, lmr);
  const double lambda_r = dmax(lpl, lpr);
  if (lambda_l >= 0.0) {   // the L and R regions need nothing of the fan
#pragma unroll
    for (int n = 0; n < NWAVE; ++n) f[n] = flux_l[n];
    return;
  }
  if (lambda_r <= 0.0) {
#pragma unroll
    for (int n = 0; n < NWAVE; ++n) f[n] = flux_r[n];
    return;
  }
  double r_l[NWAVE], r_r[NWAVE], cons_hll[NWAVE], flux_hll[NWAVE];
#pragma unroll
  for (int n = 0; n < NWAVE; ++n) {
    r_l[n] = lambda_l * cons_l[n] - flux_l[n];
    r_r[n] = lambda_r * cons_r[n] - flux_r[n];
  }
  const double ldi = 1.0 / (lambda_r - lambda_l);
#pragma unroll
  for (int n = 0; n < NWAVE; ++n) {
    cons_hll[n] = (r_r[n] - r_l[n]) * ldi;
    flux_hll[n] = (lambda_l * r_r[n] - lambda_r * r_l[n]) * ldi;
  }
  double ptot_hll;
  {
    const double m_sq = sq(cons_hll[IVX]) + sq(cons_hll[IVY]) + sq(cons_hll[IVZ]);
    const double bb_sq = sq(bbx) + sq(cons_hll[IBY]) + sq(cons_hll[IBZ]);
    const double m_dot_bb = cons_hll[IVX] * bbx + cons_hll[IVY] * cons_hll[IBY] +
                            cons_hll[IVZ] * cons_hll[IBZ];
    const double ss_sq = sq(m_dot_bb);
    const double a1 = 4.0 / 3.0 * (bb_sq - cons_hll[IEN]);
    const double a0 = 0.33333333333333333 * (m_sq + bb_sq * (bb_sq - 2.0 * cons_hll[IEN]));
    const double s2 = sq(a1) - 4.0 * a0;
    const double s = (s2 < 0.0) ? 0.0 : sqrt(s2);
    double w = (s2 >= 0.0 && a1 >= 0.0) ? -2.0 * a0 / (a1 + s) : 0.5 * (-a1 + s);
    double res = sr_e_res(w, cons_hll[IDN], cons_hll[IEN], m_sq, bb_sq, ss_sq, gp);
    for (int n = 0; n < 2; ++n) {
      const double der = sr_e_res_prime(w, cons_hll[IDN], m_sq, bb_sq, ss_sq, gp);
      const double delta = -res / der;
      w = w + delta;
      res = sr_e_res(w, cons_hll[IDN], cons_hll[IEN], m_sq, bb_sq, ss_sq, gp);
    }
    const double v_sq = (m_sq + ss_sq / sq(w) * (2.0 * w + bb_sq)) / sq(w + bb_sq);
    const double gamma_lor_sq = 1.0 / (1.0 - v_sq);
    const double gamma_lor = sqrt(gamma_lor_sq);
    const double chi = (1.0 - v_sq) * (w - gamma_lor * cons_hll[IDN]);
    const double pgas = 1.0 / gp * chi;
    const double vx = (cons_hll[IVX] + m_dot_bb / w * bbx) / (w + bb_sq);
    const double vy = (cons_hll[IVY] + m_dot_bb / w * cons_hll[IBY]) / (w + bb_sq);
    const double vz = (cons_hll[IVZ] + m_dot_bb / w * cons_hll[IBZ]) / (w + bb_sq);
    const double v_bb = vx * bbx + vy * cons_hll[IBY] + vz * cons_hll[IBZ];
    const double b_sq = bb_sq / gamma_lor_sq + sq(v_bb);
    ptot_hll = pgas + 0.5 * b_sq;
  }
  double ptot_init;
  if (sq(bbx) / ptot_hll < 0.01) {   // weak field: quadratic estimate (MUB 55)
    const double a1 = cons_hll[IEN] - flux_hll[IVX];
    const double a0 = cons_hll[IVX] * flux_hll[IEN] - flux_hll[IVX] * cons_hll[IEN];
    const double s2 = sq(a1) - 4.0 * a0;
    const double s = (s2 < 0.0) ? 0.0 : sqrt(s2);
    ptot_init = (s2 >= 0.0 && a1 >= 0.0) ? -2.0 * a0 / (a1 + s) : (-a1 + s) / 2.0;
  } else {
    ptot_init = ptot_hll;
  }
  bool to_hlle = !isfinite(ptot_init) || ptot_init <= 0.0;
  SrAlf al = sr_alfven(r_l, lambda_l, bbx, ptot_init, -1);
  SrAlf ar = sr_alfven(r_r, lambda_r, bbx, ptot_init, +1);
  const double res_0 = sr_contact_res(al, ar, bbx);
  const double ptot_1 = ptot_init * (1.0 + 1.0e-5);
  al = sr_alfven(r_l, lambda_l, bbx, ptot_1, -1);
  ar = sr_alfven(r_r, lambda_r, bbx, ptot_1, +1);
  double ptot_last = ptot_init, ptot_n = ptot_1, res_last = res_0;
  double res_n = sr_contact_res(al, ar, bbx);
  const double tol_ptot = 1.0e-8 * ptot_init;
  for (int n = 0; n < 4; ++n) {
    const double ptot_old = ptot_last;
    ptot_last = ptot_n;
    const double res_old = res_last;
    res_last = res_n;
    const bool iterate = fabs(res_last) > 1.0e-12 && fabs(ptot_last - ptot_old) > tol_ptot;
    ptot_n = iterate ? (res_last * ptot_old - res_old * ptot_last) / (res_last - res_old)
                     : ptot_last;
    al = sr_alfven(r_l, lambda_l, bbx, ptot_n, -1);
    ar = sr_alfven(r_r, lambda_r, bbx, ptot_n, +1);
    res_n = sr_contact_res(al, ar, bbx);
  }
  const double ptot_c = ptot_n;
  const double lambda_al = al.kx, lambda_ar = ar.kx;
  if (!isfinite(lambda_al) || !isfinite(lambda_ar) || !isfinite(ptot_c) || ptot_c <= 0.0)
    to_hlle = true;
  const double *fi = flux_hll;
  double fsel[NWAVE], csel[NWAVE];
  if (!to_hlle) {
    // aL / aR states (MUB 32-34) and fluxes (MUB 11-12); contact (MUB 45-52)
    double cons_al[NWAVE], cons_ar[NWAVE];
    cons_al[IBY] = al.by; cons_al[IBZ] = al.bz;
    cons_ar[IBY] = ar.by; cons_ar[IBZ] = ar.bz;
    const double v_bb_al = al.vx * bbx + al.vy * cons_al[IBY] + al.vz * cons_al[IBZ];
    cons_al[IDN] = r_l[IDN] / (lambda_l - al.vx);
    cons_al[IEN] = (r_l[IEN] + ptot_c * al.vx - v_bb_al * bbx) / (lambda_l - al.vx);
    cons_al[IVX] = (cons_al[IEN] + ptot_c) * al.vx - v_bb_al * bbx;
    cons_al[IVY] = (cons_al[IEN] + ptot_c) * al.vy - v_bb_al * cons_al[IBY];
    cons_al[IVZ] = (cons_al[IEN] + ptot_c) * al.vz - v_bb_al * cons_al[IBZ];
    const double v_bb_ar = ar.vx * bbx + ar.vy * cons_ar[IBY] + ar.vz * cons_ar[IBZ];
    cons_ar[IDN] = r_r[IDN] / (lambda_r - ar.vx);
    cons_ar[IEN] = (r_r[IEN] + ptot_c * ar.vx - v_bb_ar * bbx) / (lambda_r - ar.vx);
    cons_ar[IVX] = (cons_ar[IEN] + ptot_c) * ar.vx - v_bb_ar * bbx;
    cons_ar[IVY] = (cons_ar[IEN] + ptot_c) * ar.vy - v_bb_ar * cons_ar[IBY];
    cons_ar[IVZ] = (cons_ar[IEN] + ptot_c) * ar.vz - v_bb_ar * cons_ar[IBZ];
    if (lambda_al >= 0.0 - 1.0e-6) {           // aL region
#pragma unroll
      for (int n = 0; n < NWAVE; ++n) {
        csel[n] = cons_al[n];
        fsel[n] = lambda_l * cons_al[n] - r_l[n];
      }
    } else if (lambda_ar <= 0.0 + 1.0e-6) {    // aR region
#pragma unroll
      for (int n = 0; n < NWAVE; ++n) {
        csel[n] = cons_ar[n];
        fsel[n] = lambda_r * cons_ar[n] - r_r[n];
      }
    } else {                                   // contact region
      double cons_c[NWAVE];
      const double denom_c_inv = 1.0 / (lambda_ar - lambda_al + 1.0e-12);
      double numer_al = cons_al[IBY] * (lambda_al - al.vx) + bbx * al.vy;
      double numer_ar = cons_ar[IBY] * (lambda_ar - ar.vx) + bbx * ar.vy;
      cons_c[IBY] = (numer_ar - numer_al) * denom_c_inv;
      numer_al = cons_al[IBZ] * (lambda_al - al.vx) + bbx * al.vz;
      numer_ar = cons_ar[IBZ] * (lambda_ar - ar.vx) + bbx * ar.vz;
      cons_c[IBZ] = (numer_ar - numer_al) * denom_c_inv;
      const double k_sq_l = sq(al.kx) + sq(al.ky) + sq(al.kz);
      const double k_bc_l = al.kx * bbx + al.ky * cons_c[IBY] + al.kz * cons_c[IBZ];
      const double vx_cl = al.kx - bbx * (1.0 - k_sq_l) / (al.eta - k_bc_l);
      const double vy_cl = al.ky - cons_c[IBY] * (1.0 - k_sq_l) / (al.eta - k_bc_l);
      const double vz_cl = al.kz - cons_c[IBZ] * (1.0 - k_sq_l) / (al.eta - k_bc_l);
      const double k_sq_r = sq(ar.kx) + sq(ar.ky) + sq(ar.kz);
      const double k_bc_r = ar.kx * bbx + ar.ky * cons_c[IBY] + ar.kz * cons_c[IBZ];
      const double vx_cr = ar.kx - bbx * (1.0 - k_sq_r) / (ar.eta - k_bc_r);
      const double vy_cr = ar.ky - cons_c[IBY] * (1.0 - k_sq_r) / (ar.eta - k_bc_r);
      const double vz_cr = ar.kz - cons_c[IBZ] * (1.0 - k_sq_r) / (ar.eta - k_bc_r);
      const double vx_c = 0.5 * (vx_cl + vx_cr);
      const double vy_c = 0.5 * (vy_cl + vy_cr);
      const double vz_c = 0.5 * (vz_cl + vz_cr);
      const double v_bb_c = vx_c * bbx + vy_c * cons_c[IBY] + vz_c * cons_c[IBZ];
      const bool cl = vx_c >= 0.0;
      const double *ca = cl ? cons_al : cons_ar;
      const double lam_a = cl ? lambda_al : lambda_ar, vxa = cl ? al.vx : ar.vx;
      cons_c[IDN] = ca[IDN] * (lam_a - vxa) / (lam_a - vx_c);
      cons_c[IEN] = (lam_a * ca[IEN] - ca[IVX] + ptot_c * vx_c - v_bb_c * bbx) / (lam_a - vx_c);
      cons_c[IVX] = (cons_c[IEN] + ptot_c) * vx_c - v_bb_c * bbx;
      cons_c[IVY] = (cons_c[IEN] + ptot_c) * vy_c - v_bb_c * cons_c[IBY];
      cons_c[IVZ] = (cons_c[IEN] + ptot_c) * vz_c - v_bb_c * cons_c[IBZ];
      const double lam_f = cl ? lambda_l : lambda_r;
      const double *rs = cl ? r_l : r_r;
#pragma unroll
      for (int n = 0; n < NWAVE; ++n) {
        const double fa = lam_f * ca[n] - rs[n];
        csel[n] = cons_c[n];
        fsel[n] = fa + lam_a * (cons_c[n] - ca[n]);
      }
    }
    bool bad = false;
#pragma unroll
    for (int n = 0; n < NWAVE; ++n)
      if (!isfinite(csel[n]) || !isfinite(fsel[n])) bad = true;
    if (!bad) fi = fsel;
  }
#pragma unroll
  for (int n = 0; n < NWAVE; ++n) f[n] = fi[n];
}

//----------------------------------------------------------------------------------------
// SR primitive recovery (Newman & Hamlin 2014), adiabatic_mhd_sr.cpp:404-522.
// NHC holds the per-cell invariants the iteration works on.
struct NHC {
  double dd, ee, mm_sq, mm1, mm2, mm3, bb_sq, bb1, bb2, bb3, tt;
  double d, pgas_min;
};

constexpr double NH_TOL = 1.0e-12;
constexpr double NH_A_MIN = 1.0e-12;
constexpr double NH_V_SQ_MAX = 1.0 - 1.0e-12;
constexpr double NH_PI = 3.1415926535897932;

__device__ __forceinline__ void nh_prepare(NHC &c) {
  const double pgas_uniform_min = 1.0e-12;
  double d = 0.5 * (c.mm_sq * c.bb_sq - sq(c.tt));   // (NH 5.7)
  d = dmax(d, 0.0);
  double pgas_min = cbrt(27.0 / 4.0 * d) - c.ee - 0.5 * c.bb_sq;
  c.d = d;
  c.pgas_min = dmax(pgas_min, pgas_uniform_min);
}

// Primitives from a converged gas pressure (adiabatic_mhd_sr.cpp:486-521).
__device__ inline bool nh_finish(const NHC &c, double pgas, double gamma_adi, double *prim,
                                 double *p_gamma_lor, double *p_pmag) {
  prim[IPR] = pgas;
  if (!isfinite(pgas)) return false;
  double a = c.ee + pgas + 0.5 * c.bb_sq;
  a = dmax(a, NH_A_MIN);
  double phi = acos(1.0 / a * sqrt(27.0 * c.d / (4.0 * a)));
  double eee = a / 3.0 - 2.0 / 3.0 * a * cos(2.0 / 3.0 * (phi + NH_PI));
  double ll = eee - c.bb_sq;
  double v_sq = (c.mm_sq * sq(ll) + sq(c.tt) * (c.bb_sq + 2.0 * ll)) / sq(ll * (c.bb_sq + ll));
  v_sq = dmin(dmax(v_sq, 0.0), NH_V_SQ_MAX);
  double gamma_sq = 1.0 / (1.0 - v_sq);
  double gamma = sqrt(gamma_sq);
  prim[IDN] = c.dd / gamma;
  if (!isfinite(prim[IDN])) return false;
  double ss = c.tt / ll;
  double v1 = (c.mm1 + ss * c.bb1) / (ll + c.bb_sq);
  double v2 = (c.mm2 + ss * c.bb2) / (ll + c.bb_sq);
  double v3 = (c.mm3 + ss * c.bb3) / (ll + c.bb_sq);
  prim[IVX] = gamma * v1;
  prim[IVY] = gamma * v2;
  prim[IVZ] = gamma * v3;
  if (!isfinite(prim[IVX]) || !isfinite(prim[IVY]) || !isfinite(prim[IVZ])) return false;
  *p_gamma_lor = gamma;
  *p_pmag = 0.5 * (c.bb_sq / gamma_sq + sq(ss));
  return true;
}

// Intermediates of one evaluation of the fixed-point map, enough to finish the
// primitives at the evaluated pressure without solving the cubic again.
struct NHI {
  double ll, v_sq, gamma_sq, gamma;
};

// Primitives at pgas from the intermediates of nh_g(pgas): the tail of nh_finish.
__device__ inline bool nh_finish_i(const NHC &c, double pgas, const NHI &im, double *prim,
                                   double *p_gamma_lor, double *p_pmag) {
  prim[IPR] = pgas;
  if (!isfinite(pgas)) return false;
  const double ll = im.ll, gamma = im.gamma;
  prim[IDN] = c.dd / gamma;
  if (!isfinite(prim[IDN])) return false;
  double ss = c.tt / ll;
  double v1 = (c.mm1 + ss * c.bb1) / (ll + c.bb_sq);
  double v2 = (c.mm2 + ss * c.bb2) / (ll + c.bb_sq);
  double v3 = (c.mm3 + ss * c.bb3) / (ll + c.bb_sq);
  prim[IVX] = gamma * v1;
  prim[IVY] = gamma * v2;
  prim[IVZ] = gamma * v3;
  if (!isfinite(prim[IVX]) || !isfinite(prim[IVY]) || !isfinite(prim[IVZ])) return false;
  *p_gamma_lor = gamma;
  *p_pmag = 0.5 * (c.bb_sq / im.gamma_sq + sq(ss));
  return true;
}

// The fixed-point map of the NH iteration, pgas -> pgas' (NH 5.7, 5.10, 5.11, 5.5, 5.2,
// 3.1, 5.1, 4.5, 4.1), with its derivative when dg != nullptr and its intermediates when
// im != nullptr.
__device__ inline double nh_g(const NHC &c, double p, double gamma_adi, double *dg,
                              NHI *im = nullptr) {
  double a = c.ee + p + 0.5 * c.bb_sq;
  const bool aclamp = !(a >= NH_A_MIN);
  a = dmax(a, NH_A_MIN);
  const double s = 1.0 / a * sqrt(27.0 * c.d / (4.0 * a));
  const double phi = acos(s);
  const double th = 2.0 / 3.0 * (phi + NH_PI);
  const double cth = cos(th);
  const double eee = a / 3.0 - 2.0 / 3.0 * a * cth;
  const double ll = eee - c.bb_sq;
  const double N = c.mm_sq * sq(ll) + sq(c.tt) * (c.bb_sq + 2.0 * ll);
  const double Q = sq(ll * (c.bb_sq + ll));
  double v_sq = N / Q;
  const bool vclamp = !(v_sq >= 0.0 && v_sq <= NH_V_SQ_MAX);
  v_sq = dmin(dmax(v_sq, 0.0), NH_V_SQ_MAX);
  const double gamma_sq = 1.0 / (1.0 - v_sq);
  const double gamma = sqrt(gamma_sq);
  if (im) *im = NHI{ll, v_sq, gamma_sq, gamma};
  const double wgas = ll / gamma_sq;
  const double rho = c.dd / gamma;
  const double k = (gamma_adi - 1.0) / gamma_adi;
  const double g = k * (wgas - rho);
  if (dg) {
    double deee = 0.0;
    if (!aclamp) {
      const double dsda = -1.5 * s / a;
      const double dphida = -dsda / sqrt(dmax(1.0 - s * s, 1.0e-300));
      // th = 2/3 (phi + pi) lies in [2pi/3, pi] for phi = acos(s) in [0, pi/2]: sin >= 0
      const double sth = sqrt(dmax(1.0 - cth * cth, 0.0));
      deee = 1.0 / 3.0 - 2.0 / 3.0 * cth + 2.0 / 3.0 * a * sth * (2.0 / 3.0) * dphida;
    }
    double dvsq = 0.0;
    if (!vclamp) {
      const double dN = 2.0 * c.mm_sq * ll + 2.0 * sq(c.tt);
      const double dQ = 2.0 * ll * (c.bb_sq + ll) * (c.bb_sq + 2.0 * ll);
      dvsq = (dN * Q - N * dQ) / (Q * Q);
    }
    const double dW = (1.0 - v_sq) - ll * dvsq;
    const double drho = -c.dd / (2.0 * sqrt(1.0 - v_sq)) * dvsq;
    *dg = k * (dW - drho) * deee;
  }
  return g;
}

// The reference iteration: fixed point with Aitken extrapolation every third iterate.
// Same operation order as adiabatic_mhd_sr.cpp:437-483.
__device__ inline bool nh_solve_c(const NHC &c, double gamma_adi, double pgas_old,
                                  double *prim, double *p_gamma_lor, double *p_pmag,
                                  int *iters, int *evals) {
  const int max_iterations = 15;
  const double rr_max = 1.0 - 1.0e-12;
  const double pgas_min = c.pgas_min;
  double p0 = dmax(pgas_old, pgas_min), p1 = 0.0, p2 = 0.0;   // pgas[0..2]
  int n;
  for (n = 0; n < max_iterations; ++n) {
    const int r = n % 3;
    if (r != 2) {
      double pc = (r == 0) ? p0 : p1;
      double pn = dmax(nh_g(c, pc, gamma_adi, nullptr), pgas_min);
      *evals += 1;
      if (r == 0) p1 = pn; else p2 = pn;
      if (pn > pgas_min && fabs(pn - pc) < NH_TOL) break;
    } else {
      double rr = (p2 - p1) / (p1 - p0);
      if (!isfinite(rr) || fabs(rr) > rr_max) continue;
      p0 = p1 + (p2 - p1) / (1.0 - rr);
      p0 = dmax(p0, pgas_min);
      if (p0 > pgas_min && fabs(p0 - p2) < NH_TOL) break;
    }
  }
  *iters = n;
  if (n == max_iterations) return false;
  const int rn = (n + 1) % 3;
  double pgas = (rn == 0) ? p0 : ((rn == 1) ? p1 : p2);
  return nh_finish(c, pgas, gamma_adi, prim, p_gamma_lor, p_pmag);
}

// Newton iteration on the same fixed-point residual r(p) = p - g(p) with the analytic
// derivative of g; no counterpart in the reference.  Converges at the first iterate p
// with |g(p) - p| < 1e-12, the reference's tolerance on successive fixed-point iterates,
// and returns the primitives at that p from the intermediates of its evaluation (one
// cubic solve fewer than finishing at g(p)).  Takes a fixed-point step where 1 - g' is
// not positive.
__device__ inline bool newton_solve_c(const NHC &c, double gamma_adi, double seed,
                                      double *prim, double *p_gamma_lor, double *p_pmag,
                                      int *iters, int *evals) {
  const int max_iterations = 15;
  const double pgas_min = c.pgas_min;
  double p = dmax(seed, pgas_min);
  NHI im{0.0, 0.0, 1.0, 1.0};
  int n;
  bool conv = false;
  for (n = 0; n < max_iterations; ++n) {
    double dg;
    const double g = dmax(nh_g(c, p, gamma_adi, &dg, &im), pgas_min);
    *evals += 1;
    if (g > pgas_min && fabs(g - p) < NH_TOL) { conv = true; break; }
    const double den = 1.0 - dg;
    double pn = (den > 0.0 && isfinite(den)) ? p - (p - g) / den : g;
    if (!isfinite(pn)) pn = g;
    p = dmax(pn, pgas_min);
  }
  *iters = n + (conv ? 1 : 0);
  if (!conv) return false;
  return nh_finish_i(c, p, im, prim, p_gamma_lor, p_pmag);
}

__device__ inline bool nh_solve(double dd, double ee, double mm_sq, double mm1, double mm2,
                                double mm3, double bb_sq, double bb1, double bb2,
                                double bb3, double tt, double gamma_adi, double pgas_old,
                                double *prim, double *p_gamma_lor, double *p_pmag,
                                int *iters, int *evals) {
  NHC c{dd, ee, mm_sq, mm1, mm2, mm3, bb_sq, bb1, bb2, bb3, tt, 0.0, 0.0};
  nh_prepare(c);
  return nh_solve_c(c, gamma_adi, pgas_old, prim, p_gamma_lor, p_pmag, iters, evals);
}

// Solver selection for the first solve of sr_cons2prim_cell.
enum C2PMode {
  C2P_REFERENCE = 0,   // NH seeded from the output register (the reference)
  C2P_NEWTON = 1,      // Newton seeded from the output register
  C2P_LEARNED = 2      // Newton seeded by the learned guess; reference NH on failure
};

// Cost counters of one cell's recovery: evaluations of the fixed-point map over all
// solves of the cell, and whether the learned-seed Newton solve fell back to NH.
struct C2PStats {
  int evals;
  int fallback;
};

// One cell of the SR ConservedToPrimitive wrapper (adiabatic_mhd_sr.cpp:72-272):
// floors, NH solve seeded from the old content of the output array, fallbacks,
// gamma_max ceiling and the conserved rewrite when anything was fixed.
// cons[5] in/out, prim[5] seed in / result out.  Returns first-solve NH count, or -1
// when the first solve failed.  mode selects the first solver (C2PMode); learned_seed is
// the gas-pressure seed of C2P_LEARNED; st (may be null) receives the cost counters.
__device__ inline int sr_cons2prim_cell(double *cons, double *prim, const double *bcc,
                                        const EosK &eos, int mode, double learned_seed,
                                        C2PStats *st) {
  const double mm_sq_ee_sq_max = 1.0 - 1.0e-12;
  const double gamma_adi = eos.gamma;
  double dd = cons[IDN];
  double ee = cons[IEN];
  double mm0 = sq(cons[IM1]) + sq(cons[IM2]) + sq(cons[IM3]);
  double mm1 = cons[IM1], mm2 = cons[IM2], mm3 = cons[IM3];
  double bb0 = sq(bcc[IB1]) + sq(bcc[IB2]) + sq(bcc[IB3]);
  double bb1 = bcc[IB1], bb2 = bcc[IB2], bb3 = bcc[IB3];
  double tt = cons[IM1] * bcc[IB1] + cons[IM2] * bcc[IB2] + cons[IM3] * bcc[IB3];

  bool fixed = false;
  double dfl = eos.dfloor, pfl = eos.pfloor;
  if (dd < dfl) { dd = dfl; fixed = true; }
  double ee_min = dfl + pfl / (gamma_adi - 1.0) + 0.5 * bb0;
  if (ee < ee_min) { ee = ee_min; fixed = true; }
  double mm_sq_max = mm_sq_ee_sq_max * sq(ee);
  if (mm0 > mm_sq_max) {
    double factor = sqrt(mm_sq_max / mm0);
    mm0 = mm_sq_max;
    mm1 *= factor; mm2 *= factor; mm3 *= factor; tt *= factor;
    fixed = true;
  }
  double rho_old = dmax(prim[IDN], dfl);
  double pgas_old = dmax(prim[IPR], pfl);
  double u1_old = prim[IVX], u2_old = prim[IVY], u3_old = prim[IVZ];
  NHC nc{dd, ee, mm0, mm1, mm2, mm3, bb0, bb1, bb2, bb3, tt, 0.0, 0.0};
  nh_prepare(nc);
  double gamma = 0.0, pmag = 0.0;
  int iters = 0, evals = 0, fb = 0;
  bool success;
  if (mode == C2P_REFERENCE) {
    success = nh_solve_c(nc, gamma_adi, pgas_old, prim, &gamma, &pmag, &iters, &evals);
  } else {
    const double seed = (mode == C2P_LEARNED) ? learned_seed : pgas_old;
    success = newton_solve_c(nc, gamma_adi, seed, prim, &gamma, &pmag, &iters, &evals);
    if (!success && mode == C2P_LEARNED) {   // the plain solver is the fallback
      fb = 1;
      int it2 = 0;
      success = nh_solve_c(nc, gamma_adi, pgas_old, prim, &gamma, &pmag, &it2, &evals);
      iters += it2;
    }
  }
  int first = success ? iters : -1;
  if (!success) {
    prim[IDN] = rho_old; prim[IPR] = pgas_old;
    prim[IVX] = u1_old; prim[IVY] = u2_old; prim[IVZ] = u3_old;
    fixed = true;
  }
  if (eos.sigma_max > 0.0) dfl = dmax(dfl, 2.0 * pmag / eos.sigma_max);
  if (eos.beta_min > 0.0) pfl = dmax(pfl, eos.beta_min * pmag);
  double rho_add = dmax(dfl - prim[IDN], 0.0);
  double pgas_add = dmax(pfl - prim[IPR], 0.0);
  if (success && (rho_add > 0.0 || pgas_add > 0.0)) {
    double wgas_add = rho_add + gamma_adi / (gamma_adi - 1.0) * pgas_add;
    dd += rho_add * gamma;
    ee += wgas_add * sq(gamma) + pgas_add;
    success = nh_solve(dd, ee, mm0, mm1, mm2, mm3, bb0, bb1, bb2, bb3, tt, gamma_adi,
                       pgas_old, prim, &gamma, &pmag, &iters, &evals);
    if (!success) {
      prim[IDN] = rho_old; prim[IPR] = pgas_old;
      prim[IVX] = u1_old; prim[IVY] = u2_old; prim[IVZ] = u3_old;
    }
    fixed = true;
  }
  if (!success) gamma = sqrt(1.0 + sq(prim[IVX]) + sq(prim[IVY]) + sq(prim[IVZ]));
  bool velocity_ceiling = false;
  if (gamma > eos.gamma_max) {
    double factor = sqrt((sq(eos.gamma_max) - 1.0) / (sq(gamma) - 1.0));
    prim[IVX] *= factor; prim[IVY] *= factor; prim[IVZ] *= factor;
    fixed = true;
    velocity_ceiling = true;
  }
  if (velocity_ceiling) {
    double b0 = prim[IVX] * bcc[IB1] + prim[IVY] * bcc[IB2] + prim[IVZ] * bcc[IB3];
    pmag = 0.5 * (bb0 / sq(gamma) + sq(b0 / gamma));
  }
  dfl = eos.dfloor;
  if (eos.sigma_max > 0.0) dfl = dmax(dfl, 2.0 * pmag / eos.sigma_max);
  pfl = eos.pfloor;
  if (eos.beta_min > 0.0) pfl = dmax(pfl, eos.beta_min * pmag);
  if (prim[IDN] < dfl) { prim[IDN] = dfl; fixed = true; }
  if (prim[IPR] < pfl) { prim[IPR] = pfl; fixed = true; }
  if (!success) {
    prim[IDN] = dfl; prim[IPR] = pfl;
    prim[IVX] = prim[IVY] = prim[IVZ] = 0.0;
  }
  if (fixed) {
    double u1 = prim[IVX], u2 = prim[IVY], u3 = prim[IVZ];
    double rho = prim[IDN], pgas = prim[IPR];
    gamma = sqrt(1.0 + sq(u1) + sq(u2) + sq(u3));
    double b0 = bcc[IB1] * u1 + bcc[IB2] * u2 + bcc[IB3] * u3;
    double b1 = (bcc[IB1] + b0 * u1) / gamma;
    double b2 = (bcc[IB2] + b0 * u2) / gamma;
    double b3 = (bcc[IB3] + b0 * u3) / gamma;
    double b_sq = -sq(b0) + sq(b1) + sq(b2) + sq(b3);
    double wgas = rho + gamma_adi / (gamma_adi - 1.0) * pgas;
    double wtot = wgas + b_sq;
    pmag = 0.5 * b_sq;
    cons[IDN] = rho * gamma;
    cons[IEN] = wtot * sq(gamma) - sq(b0) - (pgas + pmag);
    cons[IM1] = wtot * gamma * u1 - b0 * b1;
    cons[IM2] = wtot * gamma * u2 - b0 * b2;
    cons[IM3] = wtot * gamma * u3 - b0 * b3;
  }
  if (st) { st->evals = evals; st->fallback = fb; }
  return first;
}

}  // namespace athgpu
