// mlp.hpp -- device side of the learned initial guess for SR primitive recovery.
//
// Network: x (6 features) -> 64 tanh -> 64 tanh -> y, y = log10(p_gas / (E - D)) after
// the output de-normalisation.  Parameters are one float array:
//   W1[64][6] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3 |
//   x_mean[6] | x_scale[6] | y_mean | y_scale
// Features are normalised as (x - x_mean) * x_scale, the output as y * y_scale + y_mean.
//
// Inference runs one wavefront per 64 cells on v_mfma_f32_32x32x8_bf16: hidden units are
// the M dimension, cells the N dimension.  The accumulator layout of that instruction
// (lane l, register r holds row 8*(r/4) + 4*(l/32) + r%4 of column l%32) is exactly the
// B-operand layout of the next layer's K-steps (lane l holds k = 4*(l/32) + j of column
// l%32), so activations stay in registers between the layers; only the weights are read,
// pre-packed in A-operand order as bf16 (four per lane per instruction).  The first
// layer is computed in split bf16 (x = x_hi + x_lo, W1 = W1_hi + W1_lo, three MFMAs per
// tile): the inputs span several decades and their single-bf16 rounding alone would
// cost ~0.04 dex of accuracy in the predicted pressure.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace athgpu {

constexpr int MLP_IN = 6;
constexpr int MLP_H = 64;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;    // 384
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;             // 448
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;     // 4544
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;             // 4608
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;             // 4672
constexpr int MLP_NW = MLP_OFF_B3 + 1;                      // 4673 trained parameters
constexpr int MLP_OFF_XM = MLP_NW;
constexpr int MLP_OFF_XS = MLP_OFF_XM + MLP_IN;
constexpr int MLP_OFF_YM = MLP_OFF_XS + MLP_IN;
constexpr int MLP_OFF_YS = MLP_OFF_YM + 1;
constexpr int MLP_NP = MLP_OFF_YS + 1;                      // 4687 floats in all
// Packed bf16 weights, in units of 4 bf16 (8 bytes) per lane per MFMA:
//   layer 1: {hi, lo} x 2 row tiles x 64 lanes; layer 2: 2 row tiles x 8 K-steps x 64.
constexpr int MLP_PK1 = 2 * 2 * 64;
constexpr int MLP_PK2 = 2 * 8 * 64;

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

struct MlpK {
  const float *P = nullptr;      // MLP_NP floats
  const uint2 *pk = nullptr;     // MLP_PK1 + MLP_PK2 packed bf16 quads
};

typedef __bf16 bf16x2v_t __attribute__((ext_vector_type(2)));
typedef float f32x2v_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2v_t __attribute__((ext_vector_type(2)));

// Two floats to two bf16 (round to nearest even, v_cvt_pk_bf16_f32), a in the low half.
__device__ __forceinline__ unsigned int pack2(float a, float b) {
  return __builtin_bit_cast(unsigned int, __builtin_convertvector(f32x2v_t{a, b}, bf16x2v_t));
}
__device__ __forceinline__ float bf_lo(unsigned int w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(unsigned int w) { return __uint_as_float(w & 0xffff0000u); }
__device__ __forceinline__ bf16x4_t quad(unsigned int lo, unsigned int hi) {
  return __builtin_bit_cast(bf16x4_t, u32x2v_t{lo, hi});
}
__device__ __forceinline__ bf16x4_t quad(uint2 w) { return quad(w.x, w.y); }

// tanh(x) = 1 - 2 / (2^(2x log2 e) + 1) with the hardware exp2 and reciprocal; absolute
// error ~1e-7, far below the bf16 rounding of the activations it feeds; saturates
// correctly (exp2 -> inf gives 1, exp2 -> 0 gives -1).
__device__ __forceinline__ float fast_tanh(float x) {
  const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
  return __builtin_fmaf(-2.0f, __builtin_amdgcn_rcpf(e + 1.0f), 1.0f);
}

// The six rotation- and scale-invariant inputs, from the conserved state (D, E, M) and
// the cell-centred field B:
//   log10(E/D - 1), |M|/E, log10(1 - M^2/E^2), log10(B^2/E), (M.B)^2/(M^2 B^2),
//   log10(sqrt(E^2 - M^2)/D - 1).
// The last one is formed in FP64 because for cold, fast flows it is a small difference
// of nearly equal numbers that the network could not form from the others.
__device__ __forceinline__ float flog10(float v) {   // v_log_f32 (log2) times log10(2)
  return __builtin_amdgcn_logf(v) * 0.30102999566398120f;
}

__device__ __forceinline__ void mlp_raw_features(const double *c, const double *b,
                                                 float *x) {
  const double dd = c[0], ee = c[4];
  const double mm = c[1] * c[1] + c[2] * c[2] + c[3] * c[3];
  const double bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
  const double tt = c[1] * b[0] + c[2] * b[1] + c[3] * b[2];
  const double idd = 1.0 / dd, iee = 1.0 / ee;
  const double m2 = mm * iee * iee;
  x[0] = flog10(static_cast<float>(fmax(ee * idd - 1.0, 1.0e-12)));
  x[1] = sqrtf(static_cast<float>(m2));
  x[2] = flog10(static_cast<float>(fmax(1.0 - m2, 1.0e-12)));
  x[3] = flog10(static_cast<float>(bb * iee + 1.0e-12));
  x[4] = static_cast<float>(tt * tt / (mm * bb + 1.0e-300));
  x[5] = flog10(static_cast<float>(fmax(sqrt(fmax(ee * ee - mm, 0.0)) * idd - 1.0, 1.0e-12)));
}

__device__ __forceinline__ void mlp_features(const double *c, const double *b,
                                             const float *P, float *x) {
  mlp_raw_features(c, b, x);
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k) {
    float v = (x[k] - P[MLP_OFF_XM + k]) * P[MLP_OFF_XS + k];
    x[k] = isfinite(v) ? v : 0.0f;
  }
}

// Gas-pressure seed from the network output y = log10(p / (E - D)).
__device__ __forceinline__ double mlp_seed(const double *c, float y) {
  const double emd = c[4] - c[0];
  const double s = emd * static_cast<double>(__builtin_amdgcn_exp2f(y * 3.3219280948873623f));
  return (emd > 0.0 && isfinite(s)) ? s : 0.0;
}

// Wavefront-collective BF16 MFMA inference: every lane of the wavefront must call it
// (inactive lanes pass any finite features); lane l receives the output for the features
// it passed, de-normalised.
__device__ __forceinline__ float mlp_wave_infer_bf16(const MlpK &K, const float *x) {
  const int lane = threadIdx.x & 63;
  const int hi = lane >> 5;            // which half-wave
  const float *P = K.P;
  // B operands of layer 1: column n = lane%32 (+32 nb), k = 4*hi + j.  Cell n's features
  // live in lane n (nb = 0) or lane n + 32 (nb = 1); features 6, 7 are zero padding.
  bf16x4_t xh[2], xl[2];
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    const int src = (lane & 31) + 32 * nb;
    float f[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float v0 = __shfl(x[j], src, 64);
      const float v1 = (j < 2) ? __shfl(x[4 + j], src, 64) : 0.0f;
      f[j] = hi ? v1 : v0;
    }
    const unsigned int h01 = pack2(f[0], f[1]), h23 = pack2(f[2], f[3]);
    const unsigned int l01 = pack2(f[0] - bf_lo(h01), f[1] - bf_hi(h01));
    const unsigned int l23 = pack2(f[2] - bf_lo(h23), f[3] - bf_hi(h23));
    xh[nb] = quad(h01, h23);
    xl[nb] = quad(l01, l23);
  }
  // layer 1: H1[64 x 64 cells] = tanh(W1 X + b1), 2 row tiles x 2 column tiles
  bf16x4_t h1b[2][8];   // [nb][K-step of layer 2]
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) {
    const bf16x4_t wh = quad(K.pk[mb * 64 + lane]);
    const bf16x4_t wl = quad(K.pk[128 + mb * 64 + lane]);
    f32x16_t bias;
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
      const float4 b4 = *reinterpret_cast<const float4 *>(P + MLP_OFF_B1 + 32 * mb + 8 * s4 + 4 * hi);
      bias[4 * s4 + 0] = b4.x; bias[4 * s4 + 1] = b4.y;
      bias[4 * s4 + 2] = b4.z; bias[4 * s4 + 3] = b4.w;
    }
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
      f32x16_t acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(wh, xh[nb], bias, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(wl, xh[nb], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(wh, xl[nb], acc, 0, 0, 0);
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4)
        h1b[nb][4 * mb + s4] = quad(pack2(fast_tanh(acc[4 * s4 + 0]), fast_tanh(acc[4 * s4 + 1])),
                                    pack2(fast_tanh(acc[4 * s4 + 2]), fast_tanh(acc[4 * s4 + 3])));
    }
  }
  // layer 2 and the output layer: y = W3 . tanh(W2 H1 + b2) + b3
  float ysum[2] = {0.0f, 0.0f};
  const uint2 *pk2 = K.pk + MLP_PK1;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) {
    f32x16_t bias;
    float w3[16];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
      const int row = 32 * mb + 8 * s4 + 4 * hi;
      const float4 b4 = *reinterpret_cast<const float4 *>(P + MLP_OFF_B2 + row);
      const float4 w4 = *reinterpret_cast<const float4 *>(P + MLP_OFF_W3 + row);
      bias[4 * s4 + 0] = b4.x; bias[4 * s4 + 1] = b4.y;
      bias[4 * s4 + 2] = b4.z; bias[4 * s4 + 3] = b4.w;
      w3[4 * s4 + 0] = w4.x; w3[4 * s4 + 1] = w4.y; w3[4 * s4 + 2] = w4.z; w3[4 * s4 + 3] = w4.w;
    }
    f32x16_t acc[2] = {bias, bias};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const bf16x4_t wa = quad(pk2[(mb * 8 + s) * 64 + lane]);
      acc[0] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(wa, h1b[0][s], acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(wa, h1b[1][s], acc[1], 0, 0, 0);
    }
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
      for (int r = 0; r < 16; ++r) ysum[nb] = __builtin_fmaf(w3[r], fast_tanh(acc[nb][r]), ysum[nb]);
  }
  // each column's 64 rows are split between lanes l and l ^ 32
  ysum[0] += __shfl_xor(ysum[0], 32, 64);
  ysum[1] += __shfl_xor(ysum[1], 32, 64);
  const float yn = (hi ? ysum[1] : ysum[0]) + P[MLP_OFF_B3];
  return yn * P[MLP_OFF_YS] + P[MLP_OFF_YM];
}

}  // namespace athgpu
