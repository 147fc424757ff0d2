// mlp.hip -- learned initial guess for SR primitive recovery.
//
// A 6 -> 64 -> 64 -> 1 tanh MLP maps rotation-invariant combinations of the conserved
// state (D, E, |M|^2, |B|^2, M.B) to log10(p_gas / (E - D)).  It is trained in FP32 by
// the kernels below (per-sample forward/backward, per-parameter gradient reduction over
// the batch, Adam), and inferred in BF16 on MFMA (v_mfma_f32_32x32x8_bf16) inside the
// primitive-recovery kernel, where it seeds an FP64 Newton iteration on the reference's
// fixed-point residual; the reference Newman-Hamlin iteration stays the fallback and
// its residual test decides convergence.  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/athena_gpu.h"
#include "mlp.hpp"
#include "physics.hpp"

using namespace athgpu;

namespace {

thread_local std::string m_err;
int mfail(const std::string &s) {
  m_err = s;
  return 1;
}
#define MCHK(expr)                                                                 \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) return mfail(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

//----------------------------------------------------------------------------------------
// Training kernels (FP32)

// Per-sample forward and backward pass.  Stores the activations and the
// back-propagated pre-activation gradients needed by the weight-gradient reduction.
// loss = mean over the batch of (y - t)^2.
__global__ void __launch_bounds__(256) k_mlp_fb(const float *__restrict__ P, int n,
                                                const float *__restrict__ X,
                                                const float *__restrict__ T,
                                                const int *__restrict__ idx, float *H1,
                                                float *H2, float *DZ1, float *DZ2, float *DY,
                                                float *L, float inv_n) {
  __shared__ float sW2[MLP_H * MLP_H];
  __shared__ float sW1[MLP_H * MLP_IN], sb1[MLP_H], sb2[MLP_H], sW3[MLP_H];
  for (int q = threadIdx.x; q < MLP_H * MLP_H; q += blockDim.x) sW2[q] = P[MLP_OFF_W2 + q];
  for (int q = threadIdx.x; q < MLP_H * MLP_IN; q += blockDim.x) sW1[q] = P[MLP_OFF_W1 + q];
  for (int q = threadIdx.x; q < MLP_H; q += blockDim.x) {
    sb1[q] = P[MLP_OFF_B1 + q];
    sb2[q] = P[MLP_OFF_B2 + q];
    sW3[q] = P[MLP_OFF_W3 + q];
  }
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  const int src = idx ? idx[b] : b;
  float x[MLP_IN];
  for (int k = 0; k < MLP_IN; ++k) x[k] = X[static_cast<long>(src) * MLP_IN + k];
  float h1[MLP_H], h2[MLP_H];
  for (int i = 0; i < MLP_H; ++i) {
    float z = sb1[i];
    for (int k = 0; k < MLP_IN; ++k) z += sW1[i * MLP_IN + k] * x[k];
    h1[i] = tanhf(z);
    H1[static_cast<long>(b) * MLP_H + i] = h1[i];
  }
  float y = P[MLP_OFF_B3];
  for (int i = 0; i < MLP_H; ++i) {
    float z = sb2[i];
    for (int j = 0; j < MLP_H; ++j) z += sW2[i * MLP_H + j] * h1[j];
    h2[i] = tanhf(z);
    H2[static_cast<long>(b) * MLP_H + i] = h2[i];
    y += sW3[i] * h2[i];
  }
  const float e = y - T[src];
  L[b] = e * e * inv_n;
  const float dy = 2.0f * e * inv_n;
  DY[b] = dy;
  float dz2[MLP_H];
  for (int i = 0; i < MLP_H; ++i) {
    dz2[i] = dy * sW3[i] * (1.0f - h2[i] * h2[i]);
    DZ2[static_cast<long>(b) * MLP_H + i] = dz2[i];
  }
  for (int j = 0; j < MLP_H; ++j) {
    float s = 0.0f;
    for (int i = 0; i < MLP_H; ++i) s += sW2[i * MLP_H + j] * dz2[i];
    DZ1[static_cast<long>(b) * MLP_H + j] = s * (1.0f - h1[j] * h1[j]);
  }
}

// Per-parameter gradient over one chunk of the batch: block (x, y) sums the outer
// products of samples [y*chunk, (y+1)*chunk) into G[y][q]; the chunks are added in a fixed
// order by k_adam / k_gsum, so the result does not depend on scheduling.
__global__ void __launch_bounds__(256) k_mlp_grad(int n, int chunk, const float *__restrict__ X,
                                                  const int *__restrict__ idx,
                                                  const float *__restrict__ H1,
                                                  const float *__restrict__ H2,
                                                  const float *__restrict__ DZ1,
                                                  const float *__restrict__ DZ2,
                                                  const float *__restrict__ DY, float *G) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= MLP_NW) return;
  const int b0 = blockIdx.y * chunk;
  const int b1 = min(n, b0 + chunk);
  float s = 0.0f;
  if (q < MLP_OFF_B1) {                         // W1[i][k]
    const int i = q / MLP_IN, k = q % MLP_IN;
    for (int b = b0; b < b1; ++b) {
      const int src = idx ? idx[b] : b;
      s += DZ1[static_cast<long>(b) * MLP_H + i] * X[static_cast<long>(src) * MLP_IN + k];
    }
  } else if (q < MLP_OFF_W2) {                  // b1[i]
    const int i = q - MLP_OFF_B1;
    for (int b = b0; b < b1; ++b) s += DZ1[static_cast<long>(b) * MLP_H + i];
  } else if (q < MLP_OFF_B2) {                  // W2[i][j]
    const int r = q - MLP_OFF_W2, i = r / MLP_H, j = r % MLP_H;
    for (int b = b0; b < b1; ++b)
      s += DZ2[static_cast<long>(b) * MLP_H + i] * H1[static_cast<long>(b) * MLP_H + j];
  } else if (q < MLP_OFF_W3) {                  // b2[i]
    const int i = q - MLP_OFF_B2;
    for (int b = b0; b < b1; ++b) s += DZ2[static_cast<long>(b) * MLP_H + i];
  } else if (q < MLP_OFF_B3) {                  // W3[i]
    const int i = q - MLP_OFF_W3;
    for (int b = b0; b < b1; ++b) s += DY[b] * H2[static_cast<long>(b) * MLP_H + i];
  } else {                                      // b3
    for (int b = b0; b < b1; ++b) s += DY[b];
  }
  G[static_cast<long>(blockIdx.y) * MLP_NW + q] = s;
}

__global__ void k_gsum(const float *__restrict__ Gp, int nchunk, float *G) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= MLP_NW) return;
  float s = 0.0f;
  for (int c = 0; c < nchunk; ++c) s += Gp[static_cast<long>(c) * MLP_NW + q];
  G[q] = s;
}

__global__ void k_sum(const float *__restrict__ v, int n, float *out) {
  __shared__ float sm[256];
  float s = 0.0f;
  for (int q = threadIdx.x; q < n; q += 256) s += v[q];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

// Adam (Kingma & Ba 2015) on the weights only.
__global__ void k_adam(float *P, const float *__restrict__ Gp, int nchunk, float *Mo, float *Vo,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= MLP_NW) return;
  float g = 0.0f;
  for (int c = 0; c < nchunk; ++c) g += Gp[static_cast<long>(c) * MLP_NW + q];
  const float m = b1 * Mo[q] + (1.0f - b1) * g;
  const float v = b2 * Vo[q] + (1.0f - b2) * g * g;
  Mo[q] = m;
  Vo[q] = v;
  P[q] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

//----------------------------------------------------------------------------------------
// Inference

// FP32 reference inference (one thread per sample), used to check the MFMA path.
__global__ void k_mlp_infer_f32(const float *__restrict__ P, int n,
                                const float *__restrict__ X, float *Y) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  float x[MLP_IN], h1[MLP_H];
  for (int k = 0; k < MLP_IN; ++k) x[k] = X[static_cast<long>(b) * MLP_IN + k];
  for (int i = 0; i < MLP_H; ++i) {
    float z = P[MLP_OFF_B1 + i];
    for (int k = 0; k < MLP_IN; ++k) z += P[MLP_OFF_W1 + i * MLP_IN + k] * x[k];
    h1[i] = tanhf(z);
  }
  float y = P[MLP_OFF_B3];
  for (int i = 0; i < MLP_H; ++i) {
    float z = P[MLP_OFF_B2 + i];
    for (int j = 0; j < MLP_H; ++j) z += P[MLP_OFF_W2 + i * MLP_H + j] * h1[j];
    y += P[MLP_OFF_W3 + i] * tanhf(z);
  }
  Y[b] = y * P[MLP_OFF_YS] + P[MLP_OFF_YM];
}

// Packs W1 (hi and lo parts) and W2 into bf16 quads in the A-operand order of the
// inference MFMAs.
__global__ void k_mlp_pack(const float *__restrict__ P, uint2 *pk) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= MLP_PK1 + MLP_PK2) return;
  float w[4];
  if (t < MLP_PK1) {   // W1 part `part`, tile mb, lane l -> row 32mb + l%32, k 4*(l/32)+j
    const int part = t / 128, mb = (t / 64) % 2, l = t % 64;
    const int row = 32 * mb + (l & 31);
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * (l >> 5) + j;
      const float v = (k < MLP_IN) ? P[MLP_OFF_W1 + row * MLP_IN + k] : 0.0f;
      const float vh = bf_lo(pack2(v, 0.0f));
      w[j] = part ? v - vh : v;
    }
  } else {             // W2: tile mb, K-step s -> k 8s + 4*(l/32) + j
    const int u = t - MLP_PK1;
    const int l = u % 64, s = (u / 64) % 8, mb = u / 512;
    const int row = 32 * mb + (l & 31);
    for (int j = 0; j < 4; ++j) w[j] = P[MLP_OFF_W2 + row * MLP_H + 8 * s + 4 * (l >> 5) + j];
  }
  pk[t] = make_uint2(pack2(w[0], w[1]), pack2(w[2], w[3]));
}

__global__ void __launch_bounds__(256) k_mlp_infer_bf16(MlpK K, int n,
                                                        const float *__restrict__ X,
                                                        float *Y) {
  const long b = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  float x[MLP_IN];
  for (int k = 0; k < MLP_IN; ++k) x[k] = (b < n) ? X[b * MLP_IN + k] : 0.0f;
  const float y = mlp_wave_infer_bf16(K, x);
  if (b < n) Y[b] = y;
}

// Raw (un-normalised) features of n cells.
__global__ void k_mlp_features(int n, const double *__restrict__ cons,
                               const double *__restrict__ bcc, float *X) {
  const long q = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double c[5], b[3];
  for (int v = 0; v < 5; ++v) c[v] = cons[q * 5 + v];
  for (int v = 0; v < 3; ++v) b[v] = bcc[q * 3 + v];
  float x[MLP_IN];
  mlp_raw_features(c, b, x);
  for (int k = 0; k < MLP_IN; ++k) X[q * MLP_IN + k] = x[k];
}

//----------------------------------------------------------------------------------------
// Learned gas-pressure seeds of n cells: wavefront-collective BF16 MFMA inference.  A
// kernel of its own so that the FP64 solver kernel keeps its own register budget.
__global__ void __launch_bounds__(256) k_mlp_seed(MlpK K, int n, const double *__restrict__ cons,
                                                  const double *__restrict__ bcc,
                                                  double *__restrict__ seed) {
  const long q = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool live = q < n;
  double c[5] = {1.0, 0.0, 0.0, 0.0, 2.0}, b[3] = {0.0, 0.0, 0.0};
  if (live) {
    for (int v = 0; v < 5; ++v) c[v] = cons[q * 5 + v];
    for (int v = 0; v < 3; ++v) b[v] = bcc[q * 3 + v];
  }
  float x[MLP_IN];
  mlp_features(c, b, K.P, x);
  const double s = mlp_seed(c, mlp_wave_infer_bf16(K, x));
  if (live) seed[q] = s;
}

// Batched SR primitive recovery with the solver mode of physics.hpp (reference NH, Newton
// from the register seed, or Newton from the learned seed in `seed`).
// bounded to 3 wavefronts per SIMD: ~190 VGPRs unbounded; at 168 a few registers spill
// and the solve runs 4-8 % faster
template <int MODE>
__global__ void __launch_bounds__(256, 3) k_sr_c2p_modes(EosK eos, int n, double *cons,
                                                      double *prim, const double *bcc,
                                                      const double *__restrict__ seed,
                                                      int *iters, int *evals, int *fallback) {
  const long q = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double c[5], cin[5], p[5], b[3];
  for (int v = 0; v < 5; ++v) { c[v] = cin[v] = cons[q * 5 + v]; p[v] = prim[q * 5 + v]; }
  for (int v = 0; v < 3; ++v) b[v] = bcc[q * 3 + v];
  C2PStats st;
  const int it = sr_cons2prim_cell(c, p, b, eos, MODE, MODE == C2P_LEARNED ? seed[q] : 0.0, &st);
  if (iters) iters[q] = it;
  if (evals) evals[q] = st.evals;
  if (fallback) fallback[q] = st.fallback;
  for (int v = 0; v < 5; ++v) {
    if (c[v] != cin[v]) cons[q * 5 + v] = c[v];   // only where floors changed it
    prim[q * 5 + v] = p[v];
  }
}

}  // namespace

struct ath_mlp {
  float *P = nullptr;    // parameters + normalisation (MLP_NP floats)
  uint2 *pk = nullptr;   // packed bf16 weights for the MFMA inference
  float *Mo = nullptr, *Vo = nullptr;
  double *seed = nullptr;   // learned seeds of the batched recovery (grow-only)
  long nseed = 0;
  long step = 0;
  hipStream_t stream = nullptr;
  MlpK k() const {
    MlpK K;
    K.P = P;
    K.pk = pk;
    return K;
  }
};

namespace {

int pack(ath_mlp *m) {
  hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_PK1 + MLP_PK2 + 255) / 256), dim3(256), 0, m->stream,
                     m->P, m->pk);
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(m->stream));
  return 0;
}

// Device scratch of one forward/backward pass over `batch` samples with `nchunk`
// gradient chunks.
struct FbBuf {
  float *base = nullptr;
  float *H1, *H2, *DZ1, *DZ2, *DY, *L, *Gp;
  int alloc(int batch, int nchunk) {
    const size_t b = static_cast<size_t>(batch);
    const size_t tot = 4 * b * MLP_H + 2 * b + static_cast<size_t>(nchunk) * MLP_NW;
    if (hipMalloc(reinterpret_cast<void **>(&base), sizeof(float) * tot) != hipSuccess)
      return mfail("hipMalloc failed");
    H1 = base;
    H2 = H1 + b * MLP_H;
    DZ1 = H2 + b * MLP_H;
    DZ2 = DZ1 + b * MLP_H;
    DY = DZ2 + b * MLP_H;
    L = DY + b;
    Gp = L + b;
    return 0;
  }
  ~FbBuf() {
    if (base) (void)hipFree(base);
  }
};

constexpr int GRAD_CHUNK = 256;   // samples per gradient partial sum

void launch_fb(ath_mlp *m, const FbBuf &f, int batch, const float *X, const float *T,
               const int *idx) {
  const int nchunk = (batch + GRAD_CHUNK - 1) / GRAD_CHUNK;
  hipLaunchKernelGGL(k_mlp_fb, dim3((batch + 255) / 256), dim3(256), 0, m->stream, m->P, batch,
                     X, T, idx, f.H1, f.H2, f.DZ1, f.DZ2, f.DY, f.L, 1.0f / batch);
  hipLaunchKernelGGL(k_mlp_grad, dim3((MLP_NW + 255) / 256, nchunk), dim3(256), 0, m->stream,
                     batch, GRAD_CHUNK, X, idx, f.H1, f.H2, f.DZ1, f.DZ2, f.DY, f.Gp);
}

template <class T>
struct DevBuf {
  T *p = nullptr;
  int alloc(size_t n) {
    if (hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * (n ? n : 1)) != hipSuccess)
      return mfail("hipMalloc failed");
    return 0;
  }
  int up(const T *h, size_t n) {
    if (hipMemcpy(p, h, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess)
      return mfail("upload failed");
    return 0;
  }
  int down(T *h, size_t n) const {
    if (hipMemcpy(h, p, sizeof(T) * n, hipMemcpyDeviceToHost) != hipSuccess)
      return mfail("download failed");
    return 0;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

EosK make_eos(double gamma, double dfloor, double pfloor, double sigma_max, double beta_min,
              double gamma_max) {
  EosK eos{};
  eos.gamma = gamma;
  eos.igm1 = 1.0 / (gamma - 1.0);
  eos.dfloor = dfloor;
  eos.pfloor = pfloor;
  eos.sigma_max = sigma_max;
  eos.beta_min = beta_min;
  eos.gamma_max = gamma_max;
  return eos;
}

}  // namespace

extern "C" {

const char *ath_ml_last_error(void) { return m_err.c_str(); }

int ath_ml_param_count(void) { return MLP_NP; }

int ath_ml_create(ath_mlp_t *out) {
  if (!out) return mfail("null argument");
  ath_mlp *m = new ath_mlp;
  if (hipMalloc(reinterpret_cast<void **>(&m->P), sizeof(float) * MLP_NP) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&m->pk), sizeof(uint2) * (MLP_PK1 + MLP_PK2)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&m->Mo), sizeof(float) * MLP_NW) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&m->Vo), sizeof(float) * MLP_NW) != hipSuccess ||
      hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    ath_ml_destroy(m);
    return mfail("allocation of MLP buffers failed");
  }
  if (hipMemset(m->P, 0, sizeof(float) * MLP_NP) != hipSuccess ||
      hipMemset(m->Mo, 0, sizeof(float) * MLP_NW) != hipSuccess ||
      hipMemset(m->Vo, 0, sizeof(float) * MLP_NW) != hipSuccess || pack(m)) {
    ath_ml_destroy(m);
    return mfail("initialisation of MLP buffers failed");
  }
  *out = m;
  return 0;
}

int ath_ml_destroy(ath_mlp_t m) {
  if (!m) return 0;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  if (m->P) (void)hipFree(m->P);
  if (m->pk) (void)hipFree(m->pk);
  if (m->Mo) (void)hipFree(m->Mo);
  if (m->Vo) (void)hipFree(m->Vo);
  if (m->seed) (void)hipFree(m->seed);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

int ath_ml_set_params(ath_mlp_t m, const float *host) {
  if (!m || !host) return mfail("null argument");
  MCHK(hipMemcpy(m->P, host, sizeof(float) * MLP_NP, hipMemcpyHostToDevice));
  MCHK(hipMemset(m->Mo, 0, sizeof(float) * MLP_NW));
  MCHK(hipMemset(m->Vo, 0, sizeof(float) * MLP_NW));
  m->step = 0;
  return pack(m);
}

int ath_ml_get_params(ath_mlp_t m, float *host) {
  if (!m || !host) return mfail("null argument");
  MCHK(hipStreamSynchronize(m->stream));
  MCHK(hipMemcpy(host, m->P, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost));
  return 0;
}

int ath_ml_grad(ath_mlp_t m, int n, const float *x, const float *t, float *grad,
                float *loss) {
  if (!m || !x || !t || n <= 0) return mfail("bad argument");
  const int nchunk = (n + GRAD_CHUNK - 1) / GRAD_CHUNK;
  FbBuf f;
  DevBuf<float> X, T, G, S;
  if (f.alloc(n, nchunk) || X.alloc(static_cast<size_t>(n) * MLP_IN) || T.alloc(n) ||
      G.alloc(MLP_NW) || S.alloc(1) || X.up(x, static_cast<size_t>(n) * MLP_IN) || T.up(t, n))
    return 1;
  launch_fb(m, f, n, X.p, T.p, nullptr);
  hipLaunchKernelGGL(k_gsum, dim3((MLP_NW + 255) / 256), dim3(256), 0, m->stream, f.Gp, nchunk,
                     G.p);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, m->stream, f.L, n, S.p);
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(m->stream));
  if ((grad && G.down(grad, MLP_NW)) || (loss && S.down(loss, 1))) return 1;
  return 0;
}

int ath_ml_train(ath_mlp_t m, int n, const float *x, const float *t, int steps, int batch,
                 const int *order, float lr, float *loss_hist) {
  if (!m || !x || !t || !order || n <= 0 || steps <= 0 || batch <= 0)
    return mfail("bad argument");
  const size_t nord = static_cast<size_t>(steps) * batch;
  for (size_t q = 0; q < nord; ++q)
    if (order[q] < 0 || order[q] >= n) return mfail("sample index out of range");
  const int nchunk = (batch + GRAD_CHUNK - 1) / GRAD_CHUNK;
  FbBuf f;
  DevBuf<float> X, T, LH;
  DevBuf<int> I;
  if (f.alloc(batch, nchunk) || X.alloc(static_cast<size_t>(n) * MLP_IN) || T.alloc(n) ||
      LH.alloc(steps) || I.alloc(nord) || X.up(x, static_cast<size_t>(n) * MLP_IN) ||
      T.up(t, n) || I.up(order, nord))
    return 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int s = 0; s < steps; ++s) {
    const int *idx = I.p + static_cast<size_t>(s) * batch;
    launch_fb(m, f, batch, X.p, T.p, idx);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, m->stream, f.L, batch, LH.p + s);
    m->step += 1;
    const float c1 = 1.0f - std::pow(b1, static_cast<float>(m->step));
    const float c2 = 1.0f - std::pow(b2, static_cast<float>(m->step));
    hipLaunchKernelGGL(k_adam, dim3((MLP_NW + 255) / 256), dim3(256), 0, m->stream, m->P, f.Gp,
                       nchunk, m->Mo, m->Vo, lr, b1, b2, eps, c1, c2);
    MCHK(hipGetLastError());
  }
  MCHK(hipStreamSynchronize(m->stream));
  if (loss_hist && LH.down(loss_hist, steps)) return 1;
  return pack(m);
}

int ath_ml_infer(ath_mlp_t m, int n, const float *x, float *y, int bf16) {
  if (!m || !x || !y || n <= 0) return mfail("bad argument");
  DevBuf<float> X, Y;
  if (X.alloc(static_cast<size_t>(n) * MLP_IN) || Y.alloc(n) ||
      X.up(x, static_cast<size_t>(n) * MLP_IN))
    return 1;
  if (bf16)
    hipLaunchKernelGGL(k_mlp_infer_bf16, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->k(),
                       n, X.p, Y.p);
  else
    hipLaunchKernelGGL(k_mlp_infer_f32, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->P, n,
                       X.p, Y.p);
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(m->stream));
  return Y.down(y, n);
}

int ath_ml_features(int n, const double *cons, const double *bcc, float *x) {
  if (n <= 0) return 0;
  if (!cons || !bcc || !x) return mfail("null argument");
  DevBuf<double> C, B;
  DevBuf<float> X;
  if (C.alloc(static_cast<size_t>(n) * 5) || B.alloc(static_cast<size_t>(n) * 3) ||
      X.alloc(static_cast<size_t>(n) * MLP_IN) || C.up(cons, static_cast<size_t>(n) * 5) ||
      B.up(bcc, static_cast<size_t>(n) * 3))
    return 1;
  hipLaunchKernelGGL(k_mlp_features, dim3((n + 255) / 256), dim3(256), 0, 0, n, C.p, B.p, X.p);
  MCHK(hipGetLastError());
  MCHK(hipDeviceSynchronize());
  return X.down(x, static_cast<size_t>(n) * MLP_IN);
}

int ath_gpu_sr_con2prim_device(ath_mlp_t mlp, int mode, int n, double gamma, double dfloor,
                               double pfloor, double sigma_max, double beta_min,
                               double gamma_max, double *cons, double *prim,
                               const double *bcc, int *iters, int *evals, int *fallback,
                               void *stream) {
  if (mode < 0 || mode > 2) return mfail("mode must be 0, 1 or 2");
  if (mode == 2 && !mlp) return mfail("the learned mode needs an MLP");
  if (n <= 0) return 0;
  if (!cons || !prim || !bcc) return mfail("null argument");
  const EosK eos = make_eos(gamma, dfloor, pfloor, sigma_max, beta_min, gamma_max);
  const dim3 grid((n + 255) / 256), wg(256);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (mode == C2P_REFERENCE) {
    hipLaunchKernelGGL(k_sr_c2p_modes<C2P_REFERENCE>, grid, wg, 0, s, eos, n, cons, prim, bcc,
                       nullptr, iters, evals, fallback);
  } else if (mode == C2P_NEWTON) {
    hipLaunchKernelGGL(k_sr_c2p_modes<C2P_NEWTON>, grid, wg, 0, s, eos, n, cons, prim, bcc,
                       nullptr, iters, evals, fallback);
  } else {
    if (mlp->nseed < n) {   // grows once; later calls of this size allocate nothing
      MCHK(hipStreamSynchronize(s));
      if (mlp->seed) MCHK(hipFree(mlp->seed));
      mlp->seed = nullptr;
      mlp->nseed = 0;
      MCHK(hipMalloc(reinterpret_cast<void **>(&mlp->seed), sizeof(double) * n));
      mlp->nseed = n;
    }
    hipLaunchKernelGGL(k_mlp_seed, grid, wg, 0, s, mlp->k(), n, cons, bcc, mlp->seed);
    hipLaunchKernelGGL(k_sr_c2p_modes<C2P_LEARNED>, grid, wg, 0, s, eos, n, cons, prim, bcc,
                       mlp->seed, iters, evals, fallback);
  }
  MCHK(hipGetLastError());
  return 0;
}

int ath_gpu_sr_con2prim_modes(ath_mlp_t mlp, int mode, int n, double gamma, double dfloor,
                              double pfloor, double sigma_max, double beta_min,
                              double gamma_max, double *cons, double *prim,
                              const double *bcc, int *iters, int *evals, int *fallback) {
  if (n <= 0) return 0;
  if (!cons || !prim || !bcc) return mfail("null argument");
  const size_t n5 = static_cast<size_t>(n) * 5;
  DevBuf<double> C, Pr, B;
  DevBuf<int> I;
  if (C.alloc(n5) || Pr.alloc(n5) || B.alloc(static_cast<size_t>(n) * 3) ||
      I.alloc(3 * static_cast<size_t>(n)) || C.up(cons, n5) || Pr.up(prim, n5) ||
      B.up(bcc, static_cast<size_t>(n) * 3))
    return 1;
  if (ath_gpu_sr_con2prim_device(mlp, mode, n, gamma, dfloor, pfloor, sigma_max, beta_min,
                                 gamma_max, C.p, Pr.p, B.p, I.p, I.p + n, I.p + 2 * n, nullptr))
    return 1;
  MCHK(hipDeviceSynchronize());
  if (C.down(cons, n5) || Pr.down(prim, n5)) return 1;
  std::vector<int> h(3 * static_cast<size_t>(n));
  if (I.down(h.data(), h.size())) return 1;
  if (iters) std::memcpy(iters, h.data(), sizeof(int) * n);
  if (evals) std::memcpy(evals, h.data() + n, sizeof(int) * n);
  if (fallback) std::memcpy(fallback, h.data() + 2 * n, sizeof(int) * n);
  return 0;
}

}  // extern "C"

// Device view of the network for the mesh kernels.
MlpK ath_ml_device_view(ath_mlp_t m) {
  MlpK K;
  if (m) K = m->k();
  return K;
}
