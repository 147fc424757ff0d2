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
// SR HLLD (Mignone, Ugliano & Bodo 2009), one interface in the rotated frame:
// hlld_rel.cpp:116-927 (HLLDTransforming, special relativity, v_interface = 0).
namespace {

// State between a fast wave and an Alfven wave for a trial total pressure ptot
// (MUB 21-35, 43): the Alfven-region velocity, transverse field, eta and K.
struct SRAlfven {
  Real vx, vy, vz, by, bz, eta, kx, ky, kz;
};

// side = -1: left fast wave (eta has the sign of -Bx), +1: right.
SRAlfven SRAlfvenState(const Real r[NWAVE], Real lambda, Real bbx, Real ptot, int side) {
  SRAlfven a;
  Real aa = r[IVX] - lambda * r[IEN] + ptot * (1.0 - SQR(lambda));                 // 26
  Real gg = SQR(r[IBY]) + SQR(r[IBZ]);                                             // 27
  Real cc = r[IVY] * r[IBY] + r[IVZ] * r[IBZ];                                     // 28
  Real qq = -aa - gg + SQR(bbx) * (1.0 - SQR(lambda));                             // 29
  Real x_inv = 1.0 / (bbx * (aa * lambda * bbx + cc)
                      - (aa + gg) * (lambda * ptot + r[IEN]));                     // 30
  a.vx = (bbx * (aa * bbx + lambda * cc) - (aa + gg) * (ptot + r[IVX])) * x_inv;   // 23
  a.vy = (qq * r[IVY] + r[IBY] * (cc + bbx * (lambda * r[IVX] - r[IEN]))) * x_inv;  // 24
  a.vz = (qq * r[IVZ] + r[IBZ] * (cc + bbx * (lambda * r[IVX] - r[IEN]))) * x_inv;  // 24
  a.by = (r[IBY] - bbx * a.vy) / (lambda - a.vx);                                   // 21
  a.bz = (r[IBZ] - bbx * a.vz) / (lambda - a.vx);
  Real v_rm = a.vx * r[IVX] + a.vy * r[IVY] + a.vz * r[IVZ];
  Real wtot = ptot + (r[IEN] - v_rm) / (lambda - a.vx);                            // 31
  a.eta = (side < 0) ? -std::copysign(std::sqrt(wtot), bbx)
                     : std::copysign(std::sqrt(wtot), bbx);                         // 35
  Real den_inv = 1.0 / (lambda * ptot + r[IEN] + bbx * a.eta);                      // 43
  a.kx = (r[IVX] + ptot + lambda * bbx * a.eta) * den_inv;
  a.ky = (r[IVY] + r[IBY] * a.eta) * den_inv;
  a.kz = (r[IVZ] + r[IBZ] * a.eta) * den_inv;
  return a;
}

// Contact-jump residual for a trial total pressure (MUB 45, 48, 49).
Real SRContactResidual(const SRAlfven &l, const SRAlfven &rr, Real bbx) {
  const Real delta_kx_aug = 1.0e-12;
  Real delta_kx = rr.kx - l.kx + delta_kx_aug;
  Real bbx_c_dk = bbx * delta_kx;
  Real by_c_dk = rr.by * (rr.kx - rr.vx) - l.by * (l.kx - l.vx) + bbx * (rr.vy - l.vy);
  Real bz_c_dk = rr.bz * (rr.kx - rr.vx) - l.bz * (l.kx - l.vx) + bbx * (rr.vz - l.vz);
  Real k_sq_l = SQR(l.kx) + SQR(l.ky) + SQR(l.kz);
  Real kdb = l.kx * bbx_c_dk + l.ky * by_c_dk + l.kz * bz_c_dk;
  Real y_l = (1.0 - k_sq_l) / (l.eta * delta_kx - kdb);
  Real k_sq_r = SQR(rr.kx) + SQR(rr.ky) + SQR(rr.kz);
  kdb = rr.kx * bbx_c_dk + rr.ky * by_c_dk + rr.kz * bz_c_dk;
  Real y_r = (1.0 - k_sq_r) / (rr.eta * delta_kx - kdb);
  return delta_kx * (1.0 - bbx * (y_r - y_l));
}

// hlld_rel.cpp:945-959 and 975-999 (Mignone & McKinney 2007, A1-A18)
Real SREResidual(Real w, Real dd, Real ee, Real m_sq, Real bb_sq, Real ss_sq, Real gp) {
  Real v_sq = (m_sq + ss_sq / SQR(w) * (2.0 * w + bb_sq)) / SQR(w + bb_sq);
  Real gamma_sq = 1.0 / (1.0 - v_sq);
  Real gamma_lorentz = std::sqrt(gamma_sq);
  Real chi = (1.0 - v_sq) * (w - gamma_lorentz * dd);
  Real pgas = chi / gp;
  Real ee_calc = w - pgas + 0.5 * bb_sq * (1.0 + v_sq) - ss_sq / (2.0 * SQR(w));
  return ee_calc - ee;
}
Real SREResidualPrime(Real w, Real dd, Real m_sq, Real bb_sq, Real ss_sq, Real gp) {
  Real v_sq = (m_sq + ss_sq / SQR(w) * (2.0 * w + bb_sq)) / SQR(w + bb_sq);
  Real gamma_sq = 1.0 / (1.0 - v_sq);
  Real gamma_lorentz = std::sqrt(gamma_sq);
  Real chi = (1.0 - v_sq) * (w - gamma_lorentz * dd);
  Real w_cu = SQR(w) * w;
  Real w_b_cu = SQR(w + bb_sq) * (w + bb_sq);
  Real dv_sq_dw = -2.0 / (w_cu * w_b_cu) *
                  (ss_sq * (3.0 * w * (w + bb_sq) + SQR(bb_sq)) + m_sq * w_cu);
  Real dchi_dw = 1.0 - v_sq - gamma_lorentz / 2.0 * (dd + 2.0 * gamma_lorentz * chi) * dv_sq_dw;
  Real drho_dw = -gamma_lorentz * dd / 2.0 * dv_sq_dw;
  Real dpgas_dchi = 1.0 / gp;
  Real dpgas_drho = 0.0;
  Real dpgas_dw = dpgas_dchi * dchi_dw + dpgas_drho * drho_dw;
  return 1.0 - dpgas_dw + 0.5 * bb_sq * dv_sq_dw + ss_sq / w_cu;
}

// The L/R conserved state and flux of one side (MUB 8, 15).
void SRSideState(const Real w[NWAVE], Real bbx, Real gamma_adi, Real cons[NWAVE],
                 Real flux[NWAVE], Real *lp, Real *lm) {
  const Real gamma_prime = gamma_adi / (gamma_adi - 1.0);
  Real rho = w[IDN], pgas = w[IPR];
  Real u[4];
  u[0] = std::sqrt(1.0 + SQR(w[IVX]) + SQR(w[IVY]) + SQR(w[IVZ]));
  u[1] = w[IVX]; u[2] = w[IVY]; u[3] = w[IVZ];
  Real b[4];
  b[0] = bbx * u[1] + w[IBY] * u[2] + w[IBZ] * u[3];
  b[1] = (bbx + b[0] * u[1]) / u[0];
  b[2] = (w[IBY] + b[0] * u[2]) / u[0];
  b[3] = (w[IBZ] + b[0] * u[3]) / u[0];
  Real b_sq = -SQR(b[0]) + SQR(b[1]) + SQR(b[2]) + SQR(b[3]);
  Real wgas = rho + gamma_adi / (gamma_adi - 1.0) * pgas;
  FastSpeedsSR(gamma_adi, wgas, pgas, u[0], u[1], b_sq, -1.0, 0.0, 1.0, lp, lm);
  Real wtot = rho + gamma_prime * pgas + b_sq;
  Real ptot = pgas + 0.5 * b_sq;
  cons[IDN] = rho * u[0];
  cons[IEN] = wtot * u[0] * u[0] - b[0] * b[0] - ptot;
  cons[IVX] = wtot * u[1] * u[0] - b[1] * b[0];
  cons[IVY] = wtot * u[2] * u[0] - b[2] * b[0];
  cons[IVZ] = wtot * u[3] * u[0] - b[3] * b[0];
  cons[IBY] = b[2] * u[0] - b[0] * u[2];
  cons[IBZ] = b[3] * u[0] - b[0] * u[3];
  flux[IDN] = rho * u[1];
  flux[IEN] = wtot * u[0] * u[1] - b[0] * b[1];
  flux[IVX] = wtot * u[1] * u[1] - b[1] * b[1] + ptot;
  flux[IVY] = wtot * u[2] * u[1] - b[2] * b[1];
  flux[IVZ] = wtot * u[3] * u[1] - b[3] * b[1];
  flux[IBY] = b[2] * u[1] - b[1] * u[2];
  flux[IBZ] = b[3] * u[1] - b[1] * u[3];
}

}  // namespace

void RiemannSRHLLD(Mesh &m, const Real wl[NWAVE], const Real wr[NWAVE], Real bbx,
                   Real flxi[NWAVE]) {
  const Real p_transition = 0.01;
  const Real vc_extension = 1.0e-6;
  const Real delta_kx_aug = 1.0e-12;
  const int num_secant = 4;
  const Real initial_offset = 1.0e-5;
  const Real tol_res = 1.0e-12;
  const Real gamma_adi = m.c.gamma;
  const Real gamma_prime = gamma_adi / (gamma_adi - 1.0);

  Real cons_l[NWAVE], flux_l[NWAVE], cons_r[NWAVE], flux_r[NWAVE];
  Real lp_l, lm_l, lp_r, lm_r;
  SRSideState(wl, bbx, gamma_adi, cons_l, flux_l, &lp_l, &lm_l);
  SRSideState(wr, bbx, gamma_adi, cons_r, flux_r, &lp_r, &lm_r);
  Real lambda_l = std::min(lm_l, lm_r);                                    // MB 55
  Real lambda_r = std::max(lp_l, lp_r);
  Real r_l[NWAVE], r_r[NWAVE], cons_hll[NWAVE], flux_hll[NWAVE];
  for (int n = 0; n < NWAVE; ++n) r_l[n] = lambda_l * cons_l[n] - flux_l[n];   // MUB 12
  for (int n = 0; n < NWAVE; ++n) r_r[n] = lambda_r * cons_r[n] - flux_r[n];
  Real lambda_diff_inv = 1.0 / (lambda_r - lambda_l);
  for (int n = 0; n < NWAVE; ++n) cons_hll[n] = (r_r[n] - r_l[n]) * lambda_diff_inv;
  for (int n = 0; n < NWAVE; ++n)
    flux_hll[n] = (lambda_l * r_r[n] - lambda_r * r_l[n]) * lambda_diff_inv;

  // total pressure of the HLL state: two Newton steps on the enthalpy (MM A26-A27)
  Real ptot_hll;
  {
    Real m_sq = SQR(cons_hll[IVX]) + SQR(cons_hll[IVY]) + SQR(cons_hll[IVZ]);
    Real bb_sq = SQR(bbx) + SQR(cons_hll[IBY]) + SQR(cons_hll[IBZ]);
    Real m_dot_bb = cons_hll[IVX] * bbx + cons_hll[IVY] * cons_hll[IBY] +
                    cons_hll[IVZ] * cons_hll[IBZ];
    Real ss_sq = SQR(m_dot_bb);
    Real a1 = 4.0 / 3.0 * (bb_sq - cons_hll[IEN]);
    Real a0 = ONE_3RD * (m_sq + bb_sq * (bb_sq - 2.0 * cons_hll[IEN]));
    Real s2 = SQR(a1) - 4.0 * a0;
    Real s = (s2 < 0.0) ? 0.0 : std::sqrt(s2);
    Real w_new = (s2 >= 0.0 && a1 >= 0.0) ? -2.0 * a0 / (a1 + s) : 0.5 * (-a1 + s);
    Real res_new = SREResidual(w_new, cons_hll[IDN], cons_hll[IEN], m_sq, bb_sq, ss_sq,
                               gamma_prime);
    for (int n = 0; n < 2; ++n) {
      Real w_old = w_new, res_old = res_new;
      Real der = SREResidualPrime(w_old, cons_hll[IDN], m_sq, bb_sq, ss_sq, gamma_prime);
      Real delta = -res_old / der;
      w_new = w_old + delta;
      res_new = SREResidual(w_new, cons_hll[IDN], cons_hll[IEN], m_sq, bb_sq, ss_sq,
                            gamma_prime);
    }
    Real w = w_new;
    Real v_sq = (m_sq + ss_sq / SQR(w) * (2.0 * w + bb_sq)) / SQR(w + bb_sq);
    Real gamma_lor_sq = 1.0 / (1.0 - v_sq);
    Real gamma_lor = std::sqrt(gamma_lor_sq);
    Real chi = (1.0 - v_sq) * (w - gamma_lor * cons_hll[IDN]);
    Real pgas = 1.0 / gamma_prime * chi;
    Real vx = (cons_hll[IVX] + m_dot_bb / w * bbx) / (w + bb_sq);
    Real vy = (cons_hll[IVY] + m_dot_bb / w * cons_hll[IBY]) / (w + bb_sq);
    Real vz = (cons_hll[IVZ] + m_dot_bb / w * cons_hll[IBZ]) / (w + bb_sq);
    Real v_bb = vx * bbx + vy * cons_hll[IBY] + vz * cons_hll[IBZ];
    Real b_sq = bb_sq / gamma_lor_sq + SQR(v_bb);
    ptot_hll = pgas + 0.5 * b_sq;
  }
  Real ptot_init;                                                          // MUB 53, 55
  if (SQR(bbx) / ptot_hll < p_transition) {
    Real a1 = cons_hll[IEN] - flux_hll[IVX];
    Real a0 = cons_hll[IVX] * flux_hll[IEN] - flux_hll[IVX] * cons_hll[IEN];
    Real s2 = SQR(a1) - 4.0 * a0;
    Real s = (s2 < 0.0) ? 0.0 : std::sqrt(s2);
    ptot_init = (s2 >= 0.0 && a1 >= 0.0) ? -2.0 * a0 / (a1 + s) : (-a1 + s) / 2.0;
  } else {
    ptot_init = ptot_hll;
  }
  bool to_hlle = !std::isfinite(ptot_init) || ptot_init <= 0.0;

  // secant iteration on the contact residual
  Real ptot_0 = ptot_init;
  SRAlfven al = SRAlfvenState(r_l, lambda_l, bbx, ptot_0, -1);
  SRAlfven ar = SRAlfvenState(r_r, lambda_r, bbx, ptot_0, +1);
  Real res_0 = SRContactResidual(al, ar, bbx);
  Real ptot_1 = ptot_init * (1.0 + initial_offset);
  al = SRAlfvenState(r_l, lambda_l, bbx, ptot_1, -1);
  ar = SRAlfvenState(r_r, lambda_r, bbx, ptot_1, +1);
  Real res_1 = SRContactResidual(al, ar, bbx);
  Real ptot_last = ptot_0, ptot_n = ptot_1, res_last = res_0, res_n = res_1;
  Real tol_ptot = 1.0e-8 * ptot_init;
  for (int n = 0; n < num_secant; ++n) {
    Real ptot_old = ptot_last;
    ptot_last = ptot_n;
    Real res_old = res_last;
    res_last = res_n;
    bool iterate = std::abs(res_last) > tol_res && std::abs(ptot_last - ptot_old) > tol_ptot;
    ptot_n = iterate ? (res_last * ptot_old - res_old * ptot_last) / (res_last - res_old)
                     : ptot_last;
    al = SRAlfvenState(r_l, lambda_l, bbx, ptot_n, -1);
    ar = SRAlfvenState(r_r, lambda_r, bbx, ptot_n, +1);
    res_n = SRContactResidual(al, ar, bbx);
  }
  Real ptot_c = ptot_n;
  Real lambda_al = al.kx, lambda_ar = ar.kx;
  if (!std::isfinite(lambda_al) || !std::isfinite(lambda_ar) || !std::isfinite(ptot_c) ||
      ptot_c <= 0.0)
    to_hlle = true;

  // aL and aR regions (MUB 32-34) and their fluxes (MUB 11-12)
  Real cons_al[NWAVE], cons_ar[NWAVE], flux_al[NWAVE], flux_ar[NWAVE];
  cons_al[IBY] = al.by; cons_al[IBZ] = al.bz;
  cons_ar[IBY] = ar.by; cons_ar[IBZ] = ar.bz;
  Real v_bb_al = al.vx * bbx + al.vy * cons_al[IBY] + al.vz * cons_al[IBZ];
  cons_al[IDN] = r_l[IDN] / (lambda_l - al.vx);
  cons_al[IEN] = (r_l[IEN] + ptot_c * al.vx - v_bb_al * bbx) / (lambda_l - al.vx);
  cons_al[IVX] = (cons_al[IEN] + ptot_c) * al.vx - v_bb_al * bbx;
  cons_al[IVY] = (cons_al[IEN] + ptot_c) * al.vy - v_bb_al * cons_al[IBY];
  cons_al[IVZ] = (cons_al[IEN] + ptot_c) * al.vz - v_bb_al * cons_al[IBZ];
  Real v_bb_ar = ar.vx * bbx + ar.vy * cons_ar[IBY] + ar.vz * cons_ar[IBZ];
  cons_ar[IDN] = r_r[IDN] / (lambda_r - ar.vx);
  cons_ar[IEN] = (r_r[IEN] + ptot_c * ar.vx - v_bb_ar * bbx) / (lambda_r - ar.vx);
  cons_ar[IVX] = (cons_ar[IEN] + ptot_c) * ar.vx - v_bb_ar * bbx;
  cons_ar[IVY] = (cons_ar[IEN] + ptot_c) * ar.vy - v_bb_ar * cons_ar[IBY];
  cons_ar[IVZ] = (cons_ar[IEN] + ptot_c) * ar.vz - v_bb_ar * cons_ar[IBZ];
  for (int n = 0; n < NWAVE; ++n) flux_al[n] = lambda_l * cons_al[n] - r_l[n];
  for (int n = 0; n < NWAVE; ++n) flux_ar[n] = lambda_r * cons_ar[n] - r_r[n];

  // contact region (MUB 45, 47, 50-52), v_c averaged over the two sides
  Real cons_c[NWAVE], flux_c[NWAVE];
  Real denom_c_inv = 1.0 / (lambda_ar - lambda_al + delta_kx_aug);
  Real numer_al = cons_al[IBY] * (lambda_al - al.vx) + bbx * al.vy;
  Real numer_ar = cons_ar[IBY] * (lambda_ar - ar.vx) + bbx * ar.vy;
  cons_c[IBY] = (numer_ar - numer_al) * denom_c_inv;
  numer_al = cons_al[IBZ] * (lambda_al - al.vx) + bbx * al.vz;
  numer_ar = cons_ar[IBZ] * (lambda_ar - ar.vx) + bbx * ar.vz;
  cons_c[IBZ] = (numer_ar - numer_al) * denom_c_inv;
  Real k_sq_l = SQR(al.kx) + SQR(al.ky) + SQR(al.kz);
  Real k_bc_l = al.kx * bbx + al.ky * cons_c[IBY] + al.kz * cons_c[IBZ];
  Real vx_cl = al.kx - bbx * (1.0 - k_sq_l) / (al.eta - k_bc_l);
  Real vy_cl = al.ky - cons_c[IBY] * (1.0 - k_sq_l) / (al.eta - k_bc_l);
  Real vz_cl = al.kz - cons_c[IBZ] * (1.0 - k_sq_l) / (al.eta - k_bc_l);
  Real k_sq_r = SQR(ar.kx) + SQR(ar.ky) + SQR(ar.kz);
  Real k_bc_r = ar.kx * bbx + ar.ky * cons_c[IBY] + ar.kz * cons_c[IBZ];
  Real vx_cr = ar.kx - bbx * (1.0 - k_sq_r) / (ar.eta - k_bc_r);
  Real vy_cr = ar.ky - cons_c[IBY] * (1.0 - k_sq_r) / (ar.eta - k_bc_r);
  Real vz_cr = ar.kz - cons_c[IBZ] * (1.0 - k_sq_r) / (ar.eta - k_bc_r);
  Real vx_c = 0.5 * (vx_cl + vx_cr);
  Real vy_c = 0.5 * (vy_cl + vy_cr);
  Real vz_c = 0.5 * (vz_cl + vz_cr);
  Real v_bb_c = vx_c * bbx + vy_c * cons_c[IBY] + vz_c * cons_c[IBZ];
  const bool cl = vx_c >= 0.0;
  if (cl) {
    cons_c[IDN] = cons_al[IDN] * (lambda_al - al.vx) / (lambda_al - vx_c);
    cons_c[IEN] = (lambda_al * cons_al[IEN] - cons_al[IVX] + ptot_c * vx_c - v_bb_c * bbx) /
                  (lambda_al - vx_c);
  } else {
    cons_c[IDN] = cons_ar[IDN] * (lambda_ar - ar.vx) / (lambda_ar - vx_c);
    cons_c[IEN] = (lambda_ar * cons_ar[IEN] - cons_ar[IVX] + ptot_c * vx_c - v_bb_c * bbx) /
                  (lambda_ar - vx_c);
  }
  cons_c[IVX] = (cons_c[IEN] + ptot_c) * vx_c - v_bb_c * bbx;
  cons_c[IVY] = (cons_c[IEN] + ptot_c) * vy_c - v_bb_c * cons_c[IBY];
  cons_c[IVZ] = (cons_c[IEN] + ptot_c) * vz_c - v_bb_c * cons_c[IBZ];
  for (int n = 0; n < NWAVE; ++n)
    flux_c[n] = cl ? flux_al[n] + lambda_al * (cons_c[n] - cons_al[n])
                   : flux_ar[n] + lambda_ar * (cons_c[n] - cons_ar[n]);

  // region of the wave fan containing the interface (hlld_rel.cpp:829-880)
  const Real v_interface = 0.0;
  const Real *ci, *fi;
  if (lambda_l >= v_interface) { ci = cons_l; fi = flux_l; }
  else if (lambda_r <= v_interface) { ci = cons_r; fi = flux_r; }
  else if (to_hlle) { ci = cons_hll; fi = flux_hll; }
  else if (lambda_al >= v_interface - vc_extension) { ci = cons_al; fi = flux_al; }
  else if (lambda_ar <= v_interface + vc_extension) { ci = cons_ar; fi = flux_ar; }
  else { ci = cons_c; fi = flux_c; }
  if (!to_hlle) {
    bool bad = false;
    for (int n = 0; n < NWAVE; ++n)
      if (!std::isfinite(ci[n]) || !std::isfinite(fi[n])) bad = true;
    if (bad) fi = flux_hll;
  }
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
