//========================================================================================
// mlp_oracle.cpp -- TEST INFRASTRUCTURE ONLY.  CPU checker of the learned-guess network
// used by the SR primitive recovery (athena_amd/csrc/mlp.hpp, mlp.hip).  The network has
// no counterpart in the reference, so nothing here restates reference code; this file
// is pinned by central finite differences of its own loss (gradient), by a closed form
// of the forward pass on one-hot parameters, and by brute force of the feature formulas
// in the tests.
//
// Parameter layout (float vector, used here in double):
//   W1[64][6] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3 | x_mean[6] x_scale[6] y_mean y_scale
//========================================================================================
#include <cmath>
#include <vector>

namespace {
constexpr int NI = 6, NH = 64;
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH;
constexpr int OW3 = OB2 + NH, OB3 = OW3 + NH, NW = OB3 + 1;
constexpr int OXM = NW, OXS = OXM + NI, OYM = OXS + NI, OYS = OYM + 1;

// Forward pass of one normalised sample: returns the normalised output and keeps the
// activations.
double forward(const double *P, const double *x, double *h1, double *h2) {
  for (int i = 0; i < NH; ++i) {
    double z = P[OB1 + i];
    for (int k = 0; k < NI; ++k) z += P[OW1 + i * NI + k] * x[k];
    h1[i] = std::tanh(z);
  }
  double y = P[OB3];
  for (int i = 0; i < NH; ++i) {
    double z = P[OB2 + i];
    for (int j = 0; j < NH; ++j) z += P[OW2 + i * NH + j] * h1[j];
    h2[i] = std::tanh(z);
    y += P[OW3 + i] * h2[i];
  }
  return y;
}
}  // namespace

extern "C" {

int ora_mlp_nw(void) { return NW; }

// De-normalised outputs y (n) for normalised inputs x (n,6).
void ora_mlp_forward(const double *P, int n, const double *x, double *y) {
  double h1[NH], h2[NH];
  for (int b = 0; b < n; ++b) y[b] = forward(P, x + b * NI, h1, h2) * P[OYS] + P[OYM];
}

// loss = mean over b of (net(x_b) - t_b)^2 in normalised output units, and its gradient
// (NW doubles) by back-propagation.
void ora_mlp_grad(const double *P, int n, const double *x, const double *t, double *grad,
                  double *loss) {
  std::vector<double> g(NW, 0.0);
  double L = 0.0, h1[NH], h2[NH], dz2[NH];
  for (int b = 0; b < n; ++b) {
    const double *xb = x + b * NI;
    const double e = forward(P, xb, h1, h2) - t[b];
    L += e * e / n;
    const double dy = 2.0 * e / n;
    g[OB3] += dy;
    for (int i = 0; i < NH; ++i) {
      g[OW3 + i] += dy * h2[i];
      dz2[i] = dy * P[OW3 + i] * (1.0 - h2[i] * h2[i]);
      g[OB2 + i] += dz2[i];
      for (int j = 0; j < NH; ++j) g[OW2 + i * NH + j] += dz2[i] * h1[j];
    }
    for (int j = 0; j < NH; ++j) {
      double s = 0.0;
      for (int i = 0; i < NH; ++i) s += P[OW2 + i * NH + j] * dz2[i];
      const double dz1 = s * (1.0 - h1[j] * h1[j]);
      g[OB1 + j] += dz1;
      for (int k = 0; k < NI; ++k) g[OW1 + j * NI + k] += dz1 * xb[k];
    }
  }
  for (int q = 0; q < NW; ++q) grad[q] = g[q];
  *loss = L;
}

// The six raw inputs of one cell (cons: D, M1, M2, M3, E; bcc: B1, B2, B3).
void ora_mlp_features(int n, const double *cons, const double *bcc, double *x) {
  for (int q = 0; q < n; ++q) {
    const double *c = cons + 5 * q, *bb = bcc + 3 * q;
    const double dd = c[0], ee = c[4];
    const double mm = c[1] * c[1] + c[2] * c[2] + c[3] * c[3];
    const double b2 = bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2];
    const double tt = c[1] * bb[0] + c[2] * bb[1] + c[3] * bb[2];
    const double m_e = std::sqrt(mm) / ee;
    double *o = x + NI * q;
    o[0] = std::log10(std::fmax(ee / dd - 1.0, 1.0e-12));
    o[1] = m_e;
    o[2] = std::log10(std::fmax(1.0 - m_e * m_e, 1.0e-12));
    o[3] = std::log10(b2 / ee + 1.0e-12);
    o[4] = tt * tt / (mm * b2 + 1.0e-300);
    o[5] = std::log10(std::fmax(std::sqrt(std::fmax(ee * ee - mm, 0.0)) / dd - 1.0, 1.0e-12));
  }
}

}  // extern "C"
