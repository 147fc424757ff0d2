//========================================================================================
// athena_oracle_ext.cpp -- TEST INFRASTRUCTURE ONLY (see athena_oracle.cpp header).
// HLLE (Newtonian MHD), PPM reconstruction and the special-relativistic pieces of the
// CPU restatement.
//========================================================================================
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

#include "oracle_core.hpp"

namespace ora {

//----------------------------------------------------------------------------------------
// HLLE for adiabatic MHD, one interface (hlle_mhd.cpp:25-182)
void HLLE1(const Real wli[NWAVE], const Real wri[NWAVE], Real bxi, Real gamma,
           Real flxi[NWAVE]) {
  Real wroe[NWAVE], fl[NWAVE], fr[NWAVE];
  Real gm1 = gamma - 1.0;
  Real sqrtdl = std::sqrt(wli[IDN]);
  Real sqrtdr = std::sqrt(wri[IDN]);
  Real isdlpdr = 1.0 / (sqrtdl + sqrtdr);

  wroe[IDN] = sqrtdl * sqrtdr;
  wroe[IVX] = (sqrtdl * wli[IVX] + sqrtdr * wri[IVX]) * isdlpdr;
  wroe[IVY] = (sqrtdl * wli[IVY] + sqrtdr * wri[IVY]) * isdlpdr;
  wroe[IVZ] = (sqrtdl * wli[IVZ] + sqrtdr * wri[IVZ]) * isdlpdr;
  wroe[IBY] = (sqrtdr * wli[IBY] + sqrtdl * wri[IBY]) * isdlpdr;
  wroe[IBZ] = (sqrtdr * wli[IBZ] + sqrtdl * wri[IBZ]) * isdlpdr;
  Real x = 0.5 * (SQR(wli[IBY] - wri[IBY]) + SQR(wli[IBZ] - wri[IBZ])) /
           (SQR(sqrtdl + sqrtdr));
  Real y = 0.5 * (wli[IDN] + wri[IDN]) / wroe[IDN];

  Real pbl = 0.5 * (bxi * bxi + SQR(wli[IBY]) + SQR(wli[IBZ]));
  Real pbr = 0.5 * (bxi * bxi + SQR(wri[IBY]) + SQR(wri[IBZ]));
  Real el = wli[IPR] / gm1 +
            0.5 * wli[IDN] * (SQR(wli[IVX]) + SQR(wli[IVY]) + SQR(wli[IVZ])) + pbl;
  Real er = wri[IPR] / gm1 +
            0.5 * wri[IDN] * (SQR(wri[IVX]) + SQR(wri[IVY]) + SQR(wri[IVZ])) + pbr;
  Real hroe = ((el + wli[IPR] + pbl) / sqrtdl + (er + wri[IPR] + pbr) / sqrtdr) * isdlpdr;

  Real cl = FastMagnetosonicSpeed(wli, bxi, gamma);
  Real cr = FastMagnetosonicSpeed(wri, bxi, gamma);

  Real btsq = SQR(wroe[IBY]) + SQR(wroe[IBZ]);
  Real vaxsq = bxi * bxi / wroe[IDN];
  Real bt_starsq = (gm1 - (gm1 - 1.0) * y) * btsq;
  Real hp = hroe - (vaxsq + btsq / wroe[IDN]);
  Real vsq = SQR(wroe[IVX]) + SQR(wroe[IVY]) + SQR(wroe[IVZ]);
  Real twid_asq = rmax((gm1 * (hp - 0.5 * vsq) - (gm1 - 1.0) * x), 0.0);
  Real ct2 = bt_starsq / wroe[IDN];
  Real tsum = vaxsq + ct2 + twid_asq;
  Real tdif = vaxsq + ct2 - twid_asq;
  Real cf2_cs2 = std::sqrt(tdif * tdif + 4.0 * twid_asq * ct2);
  Real cfsq = 0.5 * (tsum + cf2_cs2);
  Real a = std::sqrt(cfsq);

  Real al = rmin((wroe[IVX] - a), (wli[IVX] - cl));
  Real ar = rmax((wroe[IVX] + a), (wri[IVX] + cr));
  Real bp = ar > 0.0 ? ar : 0.0;
  Real bm = al < 0.0 ? al : 0.0;

  Real vxl = wli[IVX] - bm;
  Real vxr = wri[IVX] - bp;

  fl[IDN] = wli[IDN] * vxl;
  fr[IDN] = wri[IDN] * vxr;
  fl[IVX] = wli[IDN] * wli[IVX] * vxl + pbl - SQR(bxi);
  fr[IVX] = wri[IDN] * wri[IVX] * vxr + pbr - SQR(bxi);
  fl[IVY] = wli[IDN] * wli[IVY] * vxl - bxi * wli[IBY];
  fr[IVY] = wri[IDN] * wri[IVY] * vxr - bxi * wri[IBY];
  fl[IVZ] = wli[IDN] * wli[IVZ] * vxl - bxi * wli[IBZ];
  fr[IVZ] = wri[IDN] * wri[IVZ] * vxr - bxi * wri[IBZ];
  fl[IVX] += wli[IPR];
  fr[IVX] += wri[IPR];
  fl[IEN] = el * vxl + wli[IVX] * (wli[IPR] + pbl - bxi * bxi);
  fr[IEN] = er * vxr + wri[IVX] * (wri[IPR] + pbr - bxi * bxi);
  fl[IEN] -= bxi * (wli[IBY] * wli[IVY] + wli[IBZ] * wli[IVZ]);
  fr[IEN] -= bxi * (wri[IBY] * wri[IVY] + wri[IBZ] * wri[IVZ]);
  fl[IBY] = wli[IBY] * vxl - bxi * wli[IVY];
  fr[IBY] = wri[IBY] * vxr - bxi * wri[IVY];
  fl[IBZ] = wli[IBZ] * vxl - bxi * wli[IVZ];
  fr[IBZ] = wri[IBZ] * vxr - bxi * wri[IVZ];

  Real tmp = 0.0;
  if (bp != bm) tmp = 0.5 * (bp + bm) / (bp - bm);
  for (int n = 0; n < NWAVE; ++n)
    flxi[n] = 0.5 * (fl[n] + fr[n]) + (fl[n] - fr[n]) * tmp;
}

}  // namespace ora

namespace ora {

//----------------------------------------------------------------------------------------
// PPM with the Colella-Sekora extremum-preserving limiter, uniform Cartesian branch
// (ppm.cpp:44-334 X1, :344-633 X2, :642-923 X3; coefficients reconstruction.cpp:483-505)
void PiecewiseParabolic(Mesh &m, Block &b, int d, int k, int j, int il, int iu, Arr &w,
                        Arr &bcc, Arr &wl, Arr &wr) {
  const Real C2 = 1.25;
  const Real c1 = 0.5, c2 = 0.5, c3 = 0.5, c4 = 0.5, c5 = 1.0 / 6.0, c6 = -1.0 / 6.0;
  const int sk = (d == 2), sj = (d == 1), si = (d == 0);
  const Real dfl = m.dfloor, pfl = m.pfloor;
  for (int n = 0; n < NWAVE; ++n) {
    for (int i = il; i <= iu; ++i) {
      Real q_im2 = Wval(b, w, bcc, d, n, k - 2 * sk, j - 2 * sj, i - 2 * si);
      Real q_im1 = Wval(b, w, bcc, d, n, k - sk, j - sj, i - si);
      Real q = Wval(b, w, bcc, d, n, k, j, i);
      Real q_ip1 = Wval(b, w, bcc, d, n, k + sk, j + sj, i + si);
      Real q_ip2 = Wval(b, w, bcc, d, n, k + 2 * sk, j + 2 * sj, i + 2 * si);

      Real qa = (q - q_im1);
      Real qb = (q_ip1 - q);
      Real dd_im1 = c1 * qa + c2 * (q_im1 - q_im2);
      Real dd = c1 * qb + c2 * qa;
      Real dd_ip1 = c1 * (q_ip2 - q_ip1) + c2 * qb;
      Real dph = (c3 * q_im1 + c4 * q) + (c5 * dd_im1 + c6 * dd);
      Real dph_ip1 = (c3 * q + c4 * q_ip1) + (c5 * dd + c6 * dd_ip1);

      Real d2qc_im1 = q_im2 + q - 2.0 * q_im1;
      Real d2qc = q_im1 + q_ip1 - 2.0 * q;
      Real d2qc_ip1 = q + q_ip2 - 2.0 * q_ip1;
      {
        Real qa_tmp = dph - q_im1;
        Real qb_tmp = q - dph;
        Real qa2 = 3.0 * (q_im1 + q - 2.0 * dph);
        Real qb2 = d2qc_im1;
        Real qc2 = d2qc;
        Real qd = 0.0;
        if (SIGN(qa2) == SIGN(qb2) && SIGN(qa2) == SIGN(qc2)) {
          qd = SIGN(qa2) * rmin(C2 * std::abs(qb2), rmin(C2 * std::abs(qc2), std::abs(qa2)));
        }
        Real dph_tmp = 0.5 * (q_im1 + q) - qd / 6.0;
        if (qa_tmp * qb_tmp < 0.0) dph = dph_tmp;
      }
      {
        Real qa_tmp = dph_ip1 - q;
        Real qb_tmp = q_ip1 - dph_ip1;
        Real qa2 = 3.0 * (q + q_ip1 - 2.0 * dph_ip1);
        Real qb2 = d2qc;
        Real qc2 = d2qc_ip1;
        Real qd = 0.0;
        if (SIGN(qa2) == SIGN(qb2) && SIGN(qa2) == SIGN(qc2)) {
          qd = SIGN(qa2) * rmin(C2 * std::abs(qb2), rmin(C2 * std::abs(qc2), std::abs(qa2)));
        }
        Real dphip1_tmp = 0.5 * (q + q_ip1) - qd / 6.0;
        if (qa_tmp * qb_tmp < 0.0) dph_ip1 = dphip1_tmp;
      }
      Real d2qf = 6.0 * (dph + dph_ip1 - 2.0 * q);
      Real qminus = dph, qplus = dph_ip1;
      Real dqf_minus = q - qminus;
      Real dqf_plus = qplus - q;
      {
        Real qa_tmp = dqf_minus * dqf_plus;
        Real qb_tmp = (q_ip1 - q) * (q - q_im1);
        Real qa3 = d2qc_im1, qb3 = d2qc, qc3 = d2qc_ip1, qd = d2qf, qe = 0.0;
        if (SIGN(qa3) == SIGN(qb3) && SIGN(qa3) == SIGN(qc3) && SIGN(qa3) == SIGN(qd)) {
          qe = SIGN(qd) * rmin(rmin(C2 * std::abs(qa3), C2 * std::abs(qb3)),
                               rmin(C2 * std::abs(qc3), std::abs(qd)));
        }
        qa3 = rmax(std::abs(q_im1), std::abs(q_im2));
        qb3 = rmax(rmax(std::abs(q), std::abs(q_ip1)), std::abs(q_ip2));
        Real rho = 0.0;
        if (std::abs(qd) > (1.0e-12) * rmax(qa3, qb3)) rho = qe / qd;
        Real tmp_m = q - rho * dqf_minus;
        Real tmp_p = q + rho * dqf_plus;
        Real tmp2_m = q - 2.0 * dqf_plus;
        Real tmp2_p = q + 2.0 * dqf_minus;
        if ((qa_tmp <= 0.0 || qb_tmp <= 0.0)) {
          if (rho <= (1.0 - (1.0e-12))) {
            qminus = tmp_m;
            qplus = tmp_p;
          }
        } else {
          if (std::abs(dqf_minus) >= 2.0 * std::abs(dqf_plus)) qminus = tmp2_m;
          if (std::abs(dqf_plus) >= 2.0 * std::abs(dqf_minus)) qplus = tmp2_p;
        }
      }
      if (d == 0) { wl(n, 0, 0, i + 1) = qplus; wr(n, 0, 0, i) = qminus; }
      else { wl(n, 0, 0, i) = qplus; wr(n, 0, 0, i) = qminus; }
    }
  }
  // ApplyPrimitiveFloors (adiabatic_mhd.cpp:162-172; SR :531-552 identical form)
  for (int i = il; i <= iu; ++i) {
    int il_ = (d == 0) ? i + 1 : i;
    Real &wd = wl(IDN, 0, 0, il_), &wp = wl(IPR, 0, 0, il_);
    wd = (wd > dfl) ? wd : dfl;
    wp = (wp > pfl) ? wp : pfl;
    Real &rd = wr(IDN, 0, 0, i), &rp = wr(IPR, 0, 0, i);
    rd = (rd > dfl) ? rd : dfl;
    rp = (rp > pfl) ? rp : pfl;
  }
}

//----------------------------------------------------------------------------------------
// SR: EquationOfState::FastMagnetosonicSpeedsGR in flat space (adiabatic_mhd_sr.cpp:350-370)
void FastSpeedsSR(Real gamma_adi, Real wgas, Real pgas, Real u0, Real u1, Real b_sq,
                  Real g00, Real g01, Real g11, Real *lp, Real *lm) {
  Real cs_sq = gamma_adi * pgas / wgas;
  Real va_sq = b_sq / (b_sq + wgas);
  Real cms_sq = cs_sq + va_sq - cs_sq * va_sq;
  Real a = SQR(u0) - (g00 + SQR(u0)) * cms_sq;
  Real b = -2.0 * (u0 * u1 - (g01 + u0 * u1) * cms_sq);
  Real c = SQR(u1) - (g11 + SQR(u1)) * cms_sq;
  Real a1 = b / a;
  Real a0 = c / a;
  Real s = std::sqrt(rmax(SQR(a1) - 4.0 * a0, static_cast<Real>(0.0)));
  *lp = (a1 >= 0.0) ? -2.0 * a0 / (a1 + s) : (-a1 + s) / 2.0;
  *lm = (a1 >= 0.0) ? (-a1 - s) / 2.0 : -2.0 * a0 / (a1 - s);
}

// SR HLLE, one interface in the rotated frame (hlle_mhd_rel.cpp:103-~310,
// HLLETransforming with ivx = IVX and v_interface = 0)
void RiemannSR(Mesh &m, const Real wl[NWAVE], const Real wr[NWAVE], Real bb1,
               Real flxi[NWAVE]) {
  const Real gamma_adi = m.c.gamma;
  Real rho_l = wl[IDN], pgas_l = wl[IPR];
  Real u_l[4];
  u_l[0] = std::sqrt(1.0 + SQR(wl[IVX]) + SQR(wl[IVY]) + SQR(wl[IVZ]));
  u_l[1] = wl[IVX]; u_l[2] = wl[IVY]; u_l[3] = wl[IVZ];
  Real bb2_l = wl[IBY], bb3_l = wl[IBZ];
  Real rho_r = wr[IDN], pgas_r = wr[IPR];
  Real u_r[4];
  u_r[0] = std::sqrt(1.0 + SQR(wr[IVX]) + SQR(wr[IVY]) + SQR(wr[IVZ]));
  u_r[1] = wr[IVX]; u_r[2] = wr[IVY]; u_r[3] = wr[IVZ];
  Real bb2_r = wr[IBY], bb3_r = wr[IBZ];

  Real b_l[4];
  b_l[0] = bb1 * u_l[1] + bb2_l * u_l[2] + bb3_l * u_l[3];
  b_l[1] = (bb1 + b_l[0] * u_l[1]) / u_l[0];
  b_l[2] = (bb2_l + b_l[0] * u_l[2]) / u_l[0];
  b_l[3] = (bb3_l + b_l[0] * u_l[3]) / u_l[0];
  Real b_sq_l = -SQR(b_l[0]) + SQR(b_l[1]) + SQR(b_l[2]) + SQR(b_l[3]);
  Real b_r[4];
  b_r[0] = bb1 * u_r[1] + bb2_r * u_r[2] + bb3_r * u_r[3];
  b_r[1] = (bb1 + b_r[0] * u_r[1]) / u_r[0];
  b_r[2] = (bb2_r + b_r[0] * u_r[2]) / u_r[0];
  b_r[3] = (bb3_r + b_r[0] * u_r[3]) / u_r[0];
  Real b_sq_r = -SQR(b_r[0]) + SQR(b_r[1]) + SQR(b_r[2]) + SQR(b_r[3]);

  Real wgas_l = rho_l + gamma_adi / (gamma_adi - 1.0) * pgas_l;
  Real lambda_m_l, lambda_p_l;
  FastSpeedsSR(gamma_adi, wgas_l, pgas_l, u_l[0], u_l[1], b_sq_l, -1.0, 0.0, 1.0,
               &lambda_p_l, &lambda_m_l);
  Real wgas_r = rho_r + gamma_adi / (gamma_adi - 1.0) * pgas_r;
  Real lambda_m_r, lambda_p_r;
  FastSpeedsSR(gamma_adi, wgas_r, pgas_r, u_r[0], u_r[1], b_sq_r, -1.0, 0.0, 1.0,
               &lambda_p_r, &lambda_m_r);
  Real lambda_l = rmin(lambda_m_l, lambda_m_r);
  Real lambda_r = rmax(lambda_p_l, lambda_p_r);

  Real cons_l[NWAVE], flux_l[NWAVE], cons_r[NWAVE], flux_r[NWAVE];
  Real wtot_l = wgas_l + b_sq_l;
  Real ptot_l = pgas_l + 0.5 * b_sq_l;
  cons_l[IDN] = rho_l * u_l[0];
  cons_l[IEN] = wtot_l * u_l[0] * u_l[0] - b_l[0] * b_l[0] - ptot_l;
  cons_l[IVX] = wtot_l * u_l[1] * u_l[0] - b_l[1] * b_l[0];
  cons_l[IVY] = wtot_l * u_l[2] * u_l[0] - b_l[2] * b_l[0];
  cons_l[IVZ] = wtot_l * u_l[3] * u_l[0] - b_l[3] * b_l[0];
  cons_l[IBY] = b_l[2] * u_l[0] - b_l[0] * u_l[2];
  cons_l[IBZ] = b_l[3] * u_l[0] - b_l[0] * u_l[3];
  flux_l[IDN] = rho_l * u_l[1];
  flux_l[IEN] = wtot_l * u_l[0] * u_l[1] - b_l[0] * b_l[1];
  flux_l[IVX] = wtot_l * u_l[1] * u_l[1] - b_l[1] * b_l[1] + ptot_l;
  flux_l[IVY] = wtot_l * u_l[2] * u_l[1] - b_l[2] * b_l[1];
  flux_l[IVZ] = wtot_l * u_l[3] * u_l[1] - b_l[3] * b_l[1];
  flux_l[IBY] = b_l[2] * u_l[1] - b_l[1] * u_l[2];
  flux_l[IBZ] = b_l[3] * u_l[1] - b_l[1] * u_l[3];

  Real wtot_r = wgas_r + b_sq_r;
  Real ptot_r = pgas_r + 0.5 * b_sq_r;
  cons_r[IDN] = rho_r * u_r[0];
  cons_r[IEN] = wtot_r * u_r[0] * u_r[0] - b_r[0] * b_r[0] - ptot_r;
  cons_r[IVX] = wtot_r * u_r[1] * u_r[0] - b_r[1] * b_r[0];
  cons_r[IVY] = wtot_r * u_r[2] * u_r[0] - b_r[2] * b_r[0];
  cons_r[IVZ] = wtot_r * u_r[3] * u_r[0] - b_r[3] * b_r[0];
  cons_r[IBY] = b_r[2] * u_r[0] - b_r[0] * u_r[2];
  cons_r[IBZ] = b_r[3] * u_r[0] - b_r[0] * u_r[3];
  flux_r[IDN] = rho_r * u_r[1];
  flux_r[IEN] = wtot_r * u_r[0] * u_r[1] - b_r[0] * b_r[1];
  flux_r[IVX] = wtot_r * u_r[1] * u_r[1] - b_r[1] * b_r[1] + ptot_r;
  flux_r[IVY] = wtot_r * u_r[2] * u_r[1] - b_r[2] * b_r[1];
  flux_r[IVZ] = wtot_r * u_r[3] * u_r[1] - b_r[3] * b_r[1];
  flux_r[IBY] = b_r[2] * u_r[1] - b_r[1] * u_r[2];
  flux_r[IBZ] = b_r[3] * u_r[1] - b_r[1] * u_r[3];

  Real lambda_diff_inv = 1.0 / (lambda_r - lambda_l);
  Real flux_hll[NWAVE];
  for (int n = 0; n < NWAVE; ++n)
    flux_hll[n] = (lambda_r * flux_l[n] - lambda_l * flux_r[n] +
                   lambda_l * lambda_r * (cons_r[n] - cons_l[n])) * lambda_diff_inv;
  const Real v_interface = 0.0;
  const Real *fi;
  if (lambda_l >= v_interface) fi = flux_l;
  else if (lambda_r <= v_interface) fi = flux_r;
  else fi = flux_hll;
  for (int n = 0; n < NWAVE; ++n) flxi[n] = fi[n];
}

//----------------------------------------------------------------------------------------
// SR primitive recovery: Newman & Hamlin (2014) iteration
// (adiabatic_mhd_sr.cpp:404-522 ConservedToPrimitiveNormal)
// Returns iteration count through *iters (n at exit) for diagnostics.
bool NHSolve(Real dd, Real ee, Real mm_sq, Real mm1, Real mm2, Real mm3, Real bb_sq,
             Real bb1, Real bb2, Real bb3, Real tt, Real gamma_adi, Real pgas_old,
             Real prim[5], Real *p_gamma_lor, Real *p_pmag, int *iters) {
  const int max_iterations = 15;
  const Real tol = 1.0e-12;
  const Real pgas_uniform_min = 1.0e-12;
  const Real a_min = 1.0e-12;
  const Real v_sq_max = 1.0 - 1.0e-12;
  const Real rr_max = 1.0 - 1.0e-12;

  Real d = 0.5 * (mm_sq * bb_sq - SQR(tt));
  d = rmax(d, 0.0);
  Real pgas_min = std::cbrt(27.0 / 4.0 * d) - ee - 0.5 * bb_sq;
  pgas_min = rmax(pgas_min, pgas_uniform_min);

  Real pgas[3];
  pgas[0] = rmax(pgas_old, pgas_min);
  int n;
  for (n = 0; n < max_iterations; ++n) {
    Real a = 0.0;
    if (n % 3 != 2) {
      a = ee + pgas[n % 3] + 0.5 * bb_sq;
      a = rmax(a, a_min);
    }
    if (n % 3 != 2) {
      Real phi = std::acos(1.0 / a * std::sqrt(27.0 * d / (4.0 * a)));
      Real eee = a / 3.0 - 2.0 / 3.0 * a * std::cos(2.0 / 3.0 * (phi + PI));
      Real ll = eee - bb_sq;
      Real v_sq = (mm_sq * SQR(ll) + SQR(tt) * (bb_sq + 2.0 * ll)) / SQR(ll * (bb_sq + ll));
      v_sq = rmin(rmax(v_sq, static_cast<Real>(0.0)), v_sq_max);
      Real gamma_sq = 1.0 / (1.0 - v_sq);
      Real gamma = std::sqrt(gamma_sq);
      Real wgas = ll / gamma_sq;
      Real rho = dd / gamma;
      pgas[(n + 1) % 3] = (gamma_adi - 1.0) / gamma_adi * (wgas - rho);
      pgas[(n + 1) % 3] = rmax(pgas[(n + 1) % 3], pgas_min);
    }
    if (n % 3 != 2) {
      if (pgas[(n + 1) % 3] > pgas_min && std::abs(pgas[(n + 1) % 3] - pgas[n % 3]) < tol)
        break;
    }
    if (n % 3 == 2) {
      Real rr = (pgas[2] - pgas[1]) / (pgas[1] - pgas[0]);
      if (!std::isfinite(rr) || std::abs(rr) > rr_max) continue;
      pgas[0] = pgas[1] + (pgas[2] - pgas[1]) / (1.0 - rr);
      pgas[0] = rmax(pgas[0], pgas_min);
      if (pgas[0] > pgas_min && std::abs(pgas[0] - pgas[2]) < tol) break;
    }
  }
  *iters = n;
  if (n == max_iterations) return false;
  prim[IPR] = pgas[(n + 1) % 3];
  if (!std::isfinite(prim[IPR])) return false;
  Real a = ee + prim[IPR] + 0.5 * bb_sq;
  a = rmax(a, a_min);
  Real phi = std::acos(1.0 / a * std::sqrt(27.0 * d / (4.0 * a)));
  Real eee = a / 3.0 - 2.0 / 3.0 * a * std::cos(2.0 / 3.0 * (phi + PI));
  Real ll = eee - bb_sq;
  Real v_sq = (mm_sq * SQR(ll) + SQR(tt) * (bb_sq + 2.0 * ll)) / SQR(ll * (bb_sq + ll));
  v_sq = rmin(rmax(v_sq, static_cast<Real>(0.0)), v_sq_max);
  Real gamma_sq = 1.0 / (1.0 - v_sq);
  Real gamma = std::sqrt(gamma_sq);
  prim[IDN] = dd / gamma;
  if (!std::isfinite(prim[IDN])) return false;
  Real ss = tt / ll;
  Real v1 = (mm1 + ss * bb1) / (ll + bb_sq);
  Real v2 = (mm2 + ss * bb2) / (ll + bb_sq);
  Real v3 = (mm3 + ss * bb3) / (ll + bb_sq);
  prim[IVX] = gamma * v1;
  prim[IVY] = gamma * v2;
  prim[IVZ] = gamma * v3;
  if (!std::isfinite(prim[IVX]) || !std::isfinite(prim[IVY]) || !std::isfinite(prim[IVZ]))
    return false;
  *p_gamma_lor = gamma;
  *p_pmag = 0.5 * (bb_sq / gamma_sq + SQR(ss));
  return true;
}

// One cell of EquationOfState::ConservedToPrimitive, SR (adiabatic_mhd_sr.cpp:72-272).
// cons[5] in/out, prim[5] in (seed, previous content of the output array)/out,
// bcc[3] in.  Returns NH iteration count of the first solve (or -1 on failure).
int Cons2PrimSRCell(Real cons[5], Real prim[5], const Real bcc[3], Real gamma_adi,
                    Real density_floor, Real pressure_floor, Real sigma_max,
                    Real beta_min, Real gamma_max) {
  const Real mm_sq_ee_sq_max = 1.0 - 1.0e-12;
  Real dd = cons[IDN];
  Real ee = cons[IEN];
  Real mm0 = SQR(cons[IM1]) + SQR(cons[IM2]) + SQR(cons[IM3]);
  Real mm1 = cons[IM1], mm2 = cons[IM2], mm3 = cons[IM3];
  Real bb0 = SQR(bcc[IB1]) + SQR(bcc[IB2]) + SQR(bcc[IB3]);
  Real bb1 = bcc[IB1], bb2 = bcc[IB2], bb3 = bcc[IB3];
  Real tt = cons[IM1] * bcc[IB1] + cons[IM2] * bcc[IB2] + cons[IM3] * bcc[IB3];

  bool fixed = false;
  Real density_floor_local = density_floor;
  Real pressure_floor_local = pressure_floor;
  Real dd_min = density_floor_local;
  if (dd < dd_min) { dd = dd_min; fixed = true; }
  Real ee_min = density_floor_local + pressure_floor_local / (gamma_adi - 1.0) + 0.5 * bb0;
  if (ee < ee_min) { ee = ee_min; fixed = true; }
  Real mm_sq_max = mm_sq_ee_sq_max * SQR(ee);
  if (mm0 > mm_sq_max) {
    Real factor = std::sqrt(mm_sq_max / mm0);
    mm0 = mm_sq_max;
    mm1 *= factor; mm2 *= factor; mm3 *= factor; tt *= factor;
    fixed = true;
  }
  Real rho_old = rmax(prim[IDN], density_floor_local);
  Real pgas_old = rmax(prim[IPR], pressure_floor_local);
  Real u1_old = prim[IVX], u2_old = prim[IVY], u3_old = prim[IVZ];

  Real gamma = 0.0, pmag = 0.0;
  int iters = 0;
  bool success = NHSolve(dd, ee, mm0, mm1, mm2, mm3, bb0, bb1, bb2, bb3, tt, gamma_adi,
                         pgas_old, prim, &gamma, &pmag, &iters);
  int first_iters = success ? iters : -1;
  if (!success) {
    prim[IDN] = rho_old; prim[IPR] = pgas_old;
    prim[IVX] = u1_old; prim[IVY] = u2_old; prim[IVZ] = u3_old;
    fixed = true;
  }
  if (sigma_max > 0.0) density_floor_local = rmax(density_floor_local, 2.0 * pmag / sigma_max);
  if (beta_min > 0.0) pressure_floor_local = rmax(pressure_floor_local, beta_min * pmag);
  Real rho_add = rmax(density_floor_local - prim[IDN], static_cast<Real>(0.0));
  Real pgas_add = rmax(pressure_floor_local - prim[IPR], static_cast<Real>(0.0));
  if (success && (rho_add > 0.0 || pgas_add > 0.0)) {
    Real wgas_add = rho_add + gamma_adi / (gamma_adi - 1.0) * pgas_add;
    dd += rho_add * gamma;
    ee += wgas_add * SQR(gamma) + pgas_add;
    success = NHSolve(dd, ee, mm0, mm1, mm2, mm3, bb0, bb1, bb2, bb3, tt, gamma_adi,
                      pgas_old, prim, &gamma, &pmag, &iters);
    if (!success) {
      prim[IDN] = rho_old; prim[IPR] = pgas_old;
      prim[IVX] = u1_old; prim[IVY] = u2_old; prim[IVZ] = u3_old;
    }
    fixed = true;
  }
  Real &u1 = prim[IVX], &u2 = prim[IVY], &u3 = prim[IVZ];
  if (!success) gamma = std::sqrt(1.0 + SQR(u1) + SQR(u2) + SQR(u3));
  bool velocity_ceiling = false;
  if (gamma > gamma_max) {
    Real factor = std::sqrt((SQR(gamma_max) - 1.0) / (SQR(gamma) - 1.0));
    u1 *= factor; u2 *= factor; u3 *= factor;
    fixed = true;
    velocity_ceiling = true;
  }
  if (velocity_ceiling) {
    Real b0 = u1 * bcc[IB1] + u2 * bcc[IB2] + u3 * bcc[IB3];
    pmag = 0.5 * (bb0 / SQR(gamma) + SQR(b0 / gamma));
  }
  density_floor_local = density_floor;
  if (sigma_max > 0.0) density_floor_local = rmax(density_floor_local, 2.0 * pmag / sigma_max);
  pressure_floor_local = pressure_floor;
  if (beta_min > 0.0) pressure_floor_local = rmax(pressure_floor_local, beta_min * pmag);
  Real &rho = prim[IDN], &pgas = prim[IPR];
  if (rho < density_floor_local) { rho = density_floor_local; fixed = true; }
  if (pgas < pressure_floor_local) { pgas = pressure_floor_local; fixed = true; }
  if (!success) {
    rho = density_floor_local;
    pgas = pressure_floor_local;
    u1 = u2 = u3 = 0.0;
  }
  if (fixed) {
    gamma = std::sqrt(1.0 + SQR(u1) + SQR(u2) + SQR(u3));
    Real b0 = bcc[IB1] * u1 + bcc[IB2] * u2 + bcc[IB3] * u3;
    Real b1 = (bcc[IB1] + b0 * u1) / gamma;
    Real b2 = (bcc[IB2] + b0 * u2) / gamma;
    Real b3 = (bcc[IB3] + b0 * u3) / gamma;
    Real b_sq = -SQR(b0) + SQR(b1) + SQR(b2) + SQR(b3);
    Real wgas = rho + gamma_adi / (gamma_adi - 1.0) * pgas;
    Real wtot = wgas + b_sq;
    pmag = 0.5 * b_sq;
    cons[IDN] = rho * gamma;
    cons[IEN] = wtot * SQR(gamma) - SQR(b0) - (pgas + pmag);
    cons[IM1] = wtot * gamma * u1 - b0 * b1;
    cons[IM2] = wtot * gamma * u2 - b0 * b2;
    cons[IM3] = wtot * gamma * u3 - b0 * b3;
  }
  return first_iters;
}

void Cons2PrimSR(Mesh &m, Block &b, Arr &cons, Arr &prim_old, Arr &prim, int il, int iu,
                 int jl, int ju, int kl, int ku) {
  (void)prim_old;   // the reference seeds from the output array (adiabatic_mhd_sr.cpp:143)
  CellCenteredField(b, il, iu, jl, ju, kl, ku);
  const Real g = m.c.gamma, dfl = m.dfloor, pfl = m.pfloor;
  const Real smax = m.sigma_max, bmin = m.beta_min, gmax = m.gamma_max;
  for (int k = kl; k <= ku; ++k)
    for (int j = jl; j <= ju; ++j)
      for (int i = il; i <= iu; ++i) {
        Real c[5], p[5], bc[3];
        for (int n = 0; n < 5; ++n) { c[n] = cons(n, k, j, i); p[n] = prim(n, k, j, i); }
        for (int n = 0; n < 3; ++n) bc[n] = b.bcc(n, k, j, i);
        Cons2PrimSRCell(c, p, bc, g, dfl, pfl, smax, bmin, gmax);
        for (int n = 0; n < 5; ++n) { cons(n, k, j, i) = c[n]; prim(n, k, j, i) = p[n]; }
      }
}

// PrimitiveToConserved, SR (adiabatic_mhd_sr.cpp:284-332), one cell
void Prim2ConsSRCell(const Real prim[5], const Real bc[3], Real gamma_adi, Real cons[5]) {
  Real gamma_adi_red = gamma_adi / (gamma_adi - 1.0);
  Real rho = prim[IDN], pgas = prim[IPR];
  Real u1 = prim[IVX], u2 = prim[IVY], u3 = prim[IVZ];
  Real bb1 = bc[IB1], bb2 = bc[IB2], bb3 = bc[IB3];
  Real u0 = std::sqrt(1.0 + SQR(u1) + SQR(u2) + SQR(u3));
  Real b0 = bb1 * u1 + bb2 * u2 + bb3 * u3;
  Real b1 = (bb1 + b0 * u1) / u0;
  Real b2 = (bb2 + b0 * u2) / u0;
  Real b3 = (bb3 + b0 * u3) / u0;
  Real b_sq = -SQR(b0) + SQR(b1) + SQR(b2) + SQR(b3);
  Real wtot_u02 = (rho + gamma_adi_red * pgas + b_sq) * u0 * u0;
  cons[IDN] = rho * u0;
  cons[IEN] = wtot_u02 - b0 * b0 - (pgas + 0.5 * b_sq);
  cons[IM1] = wtot_u02 * u1 / u0 - b0 * b1;
  cons[IM2] = wtot_u02 * u2 / u0 - b0 * b2;
  cons[IM3] = wtot_u02 * u3 / u0 - b0 * b3;
}

}  // namespace ora

extern "C" {
// Batch SR primitive recovery for known-answer vectors: cons (n,5) in/out,
// prim (n,5) seed in / result out, bcc (n,3); iters (n) first-solve NH count or -1.
void ora_sr_con2prim(int n, double gamma_adi, double dfloor, double pfloor,
                     double sigma_max, double beta_min, double gamma_max, double *cons,
                     double *prim, const double *bcc, int *iters) {
#pragma omp parallel for schedule(static)
  for (int q = 0; q < n; ++q)
    iters[q] = ora::Cons2PrimSRCell(cons + 5 * q, prim + 5 * q, bcc + 3 * q, gamma_adi,
                                    dfloor, pfloor, sigma_max, beta_min, gamma_max);
}
void ora_sr_prim2cons(int n, double gamma_adi, const double *prim, const double *bcc,
                      double *cons) {
  for (int q = 0; q < n; ++q) ora::Prim2ConsSRCell(prim + 5 * q, bcc + 3 * q, gamma_adi, cons + 5 * q);
}
}
