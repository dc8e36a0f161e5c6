// Learned pseudorange likelihood: a 8 -> 64 -> 64 -> 2 MLP per (particle,
// satellite) pair on the gfx950 matrix cores (bf16 operands, binary32
// accumulation), laid out so that one wavefront evaluates 16 pairs per pass
// with no data exchange between the layers.
//
// Network (parameters as a binary32 vector, offsets below):
//   x  = [clamp(0.1 e, -20, 20), f_1..f_7]      f: per-satellite features
//   z1 = W1 x + b1 = w_e x_0 + c_s,  c_s = b1 + sum_j W1[:,1+j] f_j  (per satellite)
//   z2 = W2 relu(z1) + b2,  z3 = W3 relu(z2) + b3
//   mu = 10 z3_0,  l = clamp(z3_1, -4, 12),  ll = -(e - mu)^2 exp(-l) / 2 - l / 2
//
// MFMA v_mfma_f32_16x16x32_bf16 operand layout (lane l, g = l / 16):
//   A (16 x 32): lane holds row l%16, k = 8g .. 8g+7
//   B (32 x 16): lane holds column l%16, k = 8g .. 8g+7
//   C (16 x 16): lane holds column l%16, rows 4g .. 4g+3
// Every layer is computed transposed (units x pairs), so a lane always holds
// one pair (column l%16) and 16 of its units: 16 mb + 4 g + j (mb = 0..3,
// j = 0..3).  The next layer takes those as its B operand directly by
// ordering its K dimension as unit(ks, g, s) = 16 (2 ks + s/4) + 4 g + s%4
// (K-step ks = 0..1, slot s = 0..7); its A operand (the weights) is stored in
// the same K order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define GG_MLP_H 64
#define GG_MLP_NF 8
#define GG_MLP_OFF_W1 0
#define GG_MLP_OFF_B1 (GG_MLP_OFF_W1 + GG_MLP_H * GG_MLP_NF)
#define GG_MLP_OFF_W2 (GG_MLP_OFF_B1 + GG_MLP_H)
#define GG_MLP_OFF_B2 (GG_MLP_OFF_W2 + GG_MLP_H * GG_MLP_H)
#define GG_MLP_OFF_W3 (GG_MLP_OFF_B2 + GG_MLP_H)
#define GG_MLP_OFF_B3 (GG_MLP_OFF_W3 + 2 * GG_MLP_H)
#define GG_MLP_NPARAMS (GG_MLP_OFF_B3 + 2)
#define GG_MLP_LV_MIN (-4.0f)
#define GG_MLP_LV_MAX (12.0f)
#define GG_MLP_MU_SCALE 10.0f
#define GG_MLP_X_CLAMP 20.0f
// bf16 codes of |x| <= 20 (0x0000..0x41A0 and their negatives)
#define GG_MLP_CODES_HALF 16801
#define GG_MLP_CODES (2 * GG_MLP_CODES_HALF)

typedef __bf16 gg_bf16x8 __attribute__((ext_vector_type(8)));
typedef float gg_f32x4 __attribute__((ext_vector_type(4)));

// FP8 E4M3 (OCP E4M3FN: 3 mantissa bits, exponents -6..8, subnormals down to
// 2^-9, largest finite 448) value nearest to x / sc, ties to even, saturated at
// +-448, times the power-of-two scale sc; the result is exact in bf16.
__host__ __device__ inline float gg_q_e4m3(float x, float sc) {
  const float y = x / sc;
  const float a = fabsf(y);
  if (!(a > 0.0f)) return x * 0.0f;  // zero keeps its sign, NaN stays NaN
  if (a >= 448.0f) return copysignf(448.0f, y) * sc;
  int e;
  frexpf(a, &e);                               // a in [2^(e-1), 2^e)
  const int lead = (e - 1 < -6) ? -6 : e - 1;  // exponent of the quantum's binade
  const float step = ldexpf(1.0f, lead - 3);
  const float q = fminf(rintf(a / step) * step, 448.0f);
  return copysignf(q, y) * sc;
}

// Packed weights (built by gg_k_mlp_pack from the binary32 parameters).
struct gg_mlp_w {
  gg_bf16x8 A2[4][2][64];   // W2 as A fragments: [mb][ks][lane]
  gg_bf16x8 A2t[4][2][64];  // W2^T as A fragments (backward)
  gg_bf16x8 A3[2][64];      // W3 (rows 0, 1; rows 2..15 zero)
  float b2[GG_MLP_H];
  float b3[4];
  float w_e[GG_MLP_H];      // bf16(W1[:, 0]) as binary32
  float W1[GG_MLP_H][GG_MLP_NF];
  float b1[GG_MLP_H];
  float W3b[2][GG_MLP_H];   // bf16(W3) as binary32 (backward)
};

static_assert(sizeof(gg_mlp_w) % 16 == 0, "gg_mlp_w is copied in 16-byte words");

__host__ __device__ __forceinline__ int gg_mlp_unit(int ks, int g, int s) {
  return 16 * (2 * ks + (s >> 2)) + 4 * g + (s & 3);
}

__device__ __forceinline__ gg_f32x4 gg_mfma(const gg_bf16x8& a, const gg_bf16x8& b,
                                            const gg_f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// First-layer bias of one satellite: c = b1 + sum_j W1[:,1+j] f_j, binary32,
// j left to right.
__device__ __forceinline__ float gg_mlp_c(const gg_mlp_w& w, int u, const float* f) {
#pragma clang fp contract(off)
  float c = w.b1[u];
#pragma unroll
  for (int j = 0; j < GG_MLP_NF - 1; ++j) c = c + w.W1[u][1 + j] * f[j];
  return c;
}

// Satellite features 1..7 from the unit vector to the satellite in the local
// ENU frame: el / (pi/2), sin az, cos az, (cn0 - 35) / 10, LOS flag,
// n_sat / 20, ESS / N.
__device__ __forceinline__ void gg_mlp_features(float ux, float uy, float uz, float cn0, float flag,
                                                int ns, float essf, float* f) {
  const float ch = sqrtf(ux * ux + uy * uy);
  f[0] = asinf(fminf(fmaxf(uz, -1.0f), 1.0f)) * 0.63661977236758134f;
  f[1] = (ch > 0.0f) ? ux / ch : 0.0f;
  f[2] = (ch > 0.0f) ? uy / ch : 1.0f;
  f[3] = (cn0 - 35.0f) * 0.1f;
  f[4] = flag;
  f[5] = (float)ns * 0.05f;
  f[6] = essf;
}

// Index of the bf16 code of a scaled residual (|x| <= 20) in a code table.
__device__ __forceinline__ int gg_mlp_code_index(float x) {
  const uint32_t b = (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)x);
  return (b & 0x8000u) ? GG_MLP_CODES_HALF + (int)(b & 0x7FFFu) : (int)b;
}

// The binary32 value of table entry idx (exactly representable in bf16).
__device__ __forceinline__ float gg_mlp_code_value(int idx) {
  const uint32_t b = (idx < GG_MLP_CODES_HALF) ? (uint32_t)idx
                                                : (0x8000u | (uint32_t)(idx - GG_MLP_CODES_HALF));
  return __builtin_bit_cast(float, b << 16);
}

__device__ __forceinline__ float gg_mlp_xe(float e) {
#pragma clang fp contract(off)
  return fminf(fmaxf(e * 0.1f, -GG_MLP_X_CLAMP), GG_MLP_X_CLAMP);
}

typedef short gg_i16x2 __attribute__((ext_vector_type(2)));
typedef float gg_f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 gg_bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned int gg_w32x4 __attribute__((ext_vector_type(4)));

// Two binary32 values to a packed bf16 pair (round to nearest even), low half
// from a: one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t gg_cvt_pk_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((gg_f32x2){a, b}, gg_bf16x2));
}

// ReLU of a packed bf16 pair: a signed 16-bit max with 0 on the bit patterns
// (negative values have the sign bit set).  Rounding and ReLU commute, so
// relu16(cvt(x)) = bf16(max(x, 0)).
__device__ __forceinline__ uint32_t gg_relu_pk_bf16(uint32_t v) {
  const gg_i16x2 zero = {0, 0};
  return __builtin_bit_cast(uint32_t,
                            __builtin_elementwise_max(__builtin_bit_cast(gg_i16x2, v), zero));
}

// ReLU of a lane's 16 units (C layout, c[mb][j]) packed as the two B
// fragments of the next layer.
__device__ __forceinline__ void gg_mlp_relu_pack(const gg_f32x4 (&c)[4], gg_bf16x8 (&b)[2]) {
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    const gg_f32x4 lo = c[2 * ks];
    const gg_f32x4 hi = c[2 * ks + 1];
    gg_w32x4 v;
    v[0] = gg_relu_pk_bf16(gg_cvt_pk_bf16(lo[0], lo[1]));
    v[1] = gg_relu_pk_bf16(gg_cvt_pk_bf16(lo[2], lo[3]));
    v[2] = gg_relu_pk_bf16(gg_cvt_pk_bf16(hi[0], hi[1]));
    v[3] = gg_relu_pk_bf16(gg_cvt_pk_bf16(hi[2], hi[3]));
    b[ks] = __builtin_bit_cast(gg_bf16x8, v);
  }
}

// One wavefront, 16 pairs (pair = lane % 16).  xe: the pair's scaled residual
// (used by lanes 0..15); a1: the satellite's first-layer A fragments.  lds_w:
// the packed weights (LDS or global).  Outputs the layer activations the
// backward pass needs and, in lanes 0..15, z3_0 / z3_1.
struct gg_mlp_act {
  gg_f32x4 z1[4];
  gg_f32x4 z2[4];
  gg_bf16x8 h1[2];
  gg_bf16x8 h2[2];
  float z30, z31;
};

// A lane's share of the layer-2/3 weights, held in registers across passes.
struct gg_mlp_regs {
  gg_bf16x8 A2[4][2];
  gg_bf16x8 A3[2];
  gg_f32x4 b2[4];  // C-layout bias of the lane's units 16 mb + 4 g + j
  float b30, b31;
};

__device__ __forceinline__ void gg_mlp_load_regs(const gg_mlp_w& w, gg_mlp_regs& r) {
  const int l = threadIdx.x & 63;
  const int g = l >> 4;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    r.A2[mb][0] = w.A2[mb][0][l];
    r.A2[mb][1] = w.A2[mb][1][l];
#pragma unroll
    for (int j = 0; j < 4; ++j) r.b2[mb][j] = w.b2[16 * mb + 4 * g + j];
  }
  r.A3[0] = w.A3[0][l];
  r.A3[1] = w.A3[1][l];
  r.b30 = (g == 0) ? w.b3[0] : 0.0f;
  r.b31 = (g == 0) ? w.b3[1] : 0.0f;
}

__device__ __forceinline__ void gg_mlp_fwd16(const gg_mlp_regs& w, const gg_bf16x8 (&a1)[4],
                                             float xe, gg_mlp_act& act) {
  const int l = threadIdx.x & 63;
  const int g = l >> 4;
  gg_bf16x8 b1;
#pragma unroll
  for (int s = 0; s < 8; ++s) b1[s] = (__bf16)0.0f;
  if (g == 0) {
    b1[0] = (__bf16)xe;
    b1[1] = (__bf16)1.0f;
  }
  const gg_f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) act.z1[mb] = gg_mfma(a1[mb], b1, zero);
  gg_mlp_relu_pack(act.z1, act.h1);
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    gg_f32x4 c = gg_mfma(w.A2[mb][0], act.h1[0], w.b2[mb]);
    act.z2[mb] = gg_mfma(w.A2[mb][1], act.h1[1], c);
  }
  gg_mlp_relu_pack(act.z2, act.h2);
  gg_f32x4 c3 = {w.b30, w.b31, 0.0f, 0.0f};
  c3 = gg_mfma(w.A3[0], act.h2[0], c3);
  c3 = gg_mfma(w.A3[1], act.h2[1], c3);
  act.z30 = c3[0];
  act.z31 = c3[1];
}

// First-layer A fragments of one satellite: lanes 0..15 hold k = 0 (w_e) and
// k = 1 (c_s) of unit 16 mb + lane; everything else is zero.
__device__ __forceinline__ void gg_mlp_a1(const float* w_e, const float* c_s,
                                          gg_bf16x8 (&a1)[4]) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
    for (int s = 0; s < 8; ++s) a1[mb][s] = (__bf16)0.0f;
    if (l < 16) {
      a1[mb][0] = (__bf16)w_e[16 * mb + l];
      a1[mb][1] = (__bf16)c_s[16 * mb + l];
    }
  }
}

// Pair log-likelihood from the network output (lanes 0..15).
__device__ __forceinline__ float gg_mlp_ll(float e, float z30, float z31) {
#pragma clang fp contract(off)
  const float mu = GG_MLP_MU_SCALE * z30;
  const float lv = fminf(fmaxf(z31, GG_MLP_LV_MIN), GG_MLP_LV_MAX);
  const float r = e - mu;
  return -0.5f * r * r * __expf(-lv) - 0.5f * lv;
}
