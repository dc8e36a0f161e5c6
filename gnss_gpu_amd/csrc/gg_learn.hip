// Learned likelihood: parameter handle, weight packing, evaluation and
// training kernels (the network and its layout are in gg_mlp.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "gg_common_host.h"
#include "gg_device.h"
#include "gg_mlp.h"
#include "gnss_gpu_amd.h"

struct gg_mlp {
  float* params = nullptr;  // [GG_MLP_NPARAMS] binary32 master copy
  gg_mlp_w* w = nullptr;    // packed bf16 fragments
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  float* grad = nullptr;    // [GG_MLP_NPARAMS]
  float* partial = nullptr; // [partial_cap][GG_MLP_NPARAMS]
  int partial_cap = 0;
  float* loss = nullptr;    // [partial_cap]
  // training set
  float* tf = nullptr;      // [E][S][7]
  float* tr0 = nullptr;     // [E][S]
  float* tu = nullptr;      // [E][S][3]
  int32_t* tns = nullptr;   // [E] satellites per epoch
  int n_train = 0, s_max = 0;
  float* cand = nullptr;    // [batch][K][4]
  int cand_cap = 0;
  int32_t* batch_idx = nullptr;
  int adam_t = 0;
  int fp8 = 0;              // FP8 E4M3 weight mode (gg_k_mlp_pack)
  hipStream_t stream = nullptr;
};

namespace {

// Per-tensor scale 2^k of the FP8 (OCP E4M3FN) weight mode: the smallest
// power of two with max|W| / 2^k <= 448 (1 for an all-zero tensor).
__device__ float gg_e4m3_scale(const float* W, int n, float* lds) {
  float m = 0.0f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmaxf(m, fabsf(W[i]));
  lds[threadIdx.x] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float mm = 0.0f;
    for (int i = 0; i < (int)blockDim.x; ++i) mm = fmaxf(mm, lds[i]);
    float sc = 1.0f;
    if (mm > 0.0f) {
      int e;
      frexpf(mm / 448.0f, &e);  // mm / 448 in [2^(e-1), 2^e)
      sc = ldexpf(1.0f, e);
      if (ldexpf(1.0f, e - 1) >= mm / 448.0f) sc = ldexpf(1.0f, e - 1);
    }
    lds[blockDim.x] = sc;
  }
  __syncthreads();
  const float sc = lds[blockDim.x];
  __syncthreads();
  return sc;
}

// Packs the binary32 master parameters into the matrix-core fragments.  With
// fp8 set, W1, W2 and W3 are first replaced by their FP8 E4M3 values times a
// per-tensor power-of-two scale (gg_q_e4m3); those values are exact in bf16,
// so the bf16 matrix-core path computes with the FP8 weights unchanged.
__global__ void gg_k_mlp_pack(const float* p, gg_mlp_w* w, int fp8) {
  __shared__ float lds[257];
  const int t = threadIdx.x;
  const float* W1 = p + GG_MLP_OFF_W1;
  const float* W2 = p + GG_MLP_OFF_W2;
  const float* W3 = p + GG_MLP_OFF_W3;
  float s1 = 1.0f, s2 = 1.0f, s3 = 1.0f;
  if (fp8) {
    s1 = gg_e4m3_scale(W1, GG_MLP_H * GG_MLP_NF, lds);
    s2 = gg_e4m3_scale(W2, GG_MLP_H * GG_MLP_H, lds);
    s3 = gg_e4m3_scale(W3, 2 * GG_MLP_H, lds);
  }
  auto q1 = [&](int i) { return fp8 ? gg_q_e4m3(W1[i], s1) : W1[i]; };
  auto q2 = [&](int i) { return fp8 ? gg_q_e4m3(W2[i], s2) : W2[i]; };
  auto q3 = [&](int i) { return fp8 ? gg_q_e4m3(W3[i], s3) : W3[i]; };
  for (int i = t; i < 4 * 2 * 64; i += blockDim.x) {
    const int mb = i / 128, ks = (i / 64) % 2, l = i % 64;
    gg_bf16x8 a, at;
    for (int s = 0; s < 8; ++s) {
      const int u = gg_mlp_unit(ks, l >> 4, s);
      a[s] = (__bf16)q2((16 * mb + (l & 15)) * GG_MLP_H + u);
      at[s] = (__bf16)q2(u * GG_MLP_H + 16 * mb + (l & 15));
    }
    w->A2[mb][ks][l] = a;
    w->A2t[mb][ks][l] = at;
  }
  for (int i = t; i < 2 * 64; i += blockDim.x) {
    const int ks = i / 64, l = i % 64;
    gg_bf16x8 a;
    for (int s = 0; s < 8; ++s)
      a[s] = ((l & 15) < 2) ? (__bf16)q3((l & 15) * GG_MLP_H + gg_mlp_unit(ks, l >> 4, s))
                            : (__bf16)0.0f;
    w->A3[ks][l] = a;
  }
  for (int u = t; u < GG_MLP_H; u += blockDim.x) {
    w->b2[u] = p[GG_MLP_OFF_B2 + u];
    w->w_e[u] = (float)(__bf16)q1(u * GG_MLP_NF);
    for (int j = 0; j < GG_MLP_NF; ++j) w->W1[u][j] = q1(u * GG_MLP_NF + j);
    w->b1[u] = p[GG_MLP_OFF_B1 + u];
    w->W3b[0][u] = (float)(__bf16)q3(u);
    w->W3b[1][u] = (float)(__bf16)q3(GG_MLP_H + u);
  }
  if (t < 4) w->b3[t] = (t < 2) ? p[GG_MLP_OFF_B3 + t] : 0.0f;
}

// Evaluation of groups of 16 pairs that share one satellite's features: one
// wavefront per group.
__global__ __launch_bounds__(256) void gg_k_mlp_eval(const gg_mlp_w* __restrict__ w, int nb,
                                                     const float* __restrict__ e,
                                                     const float* __restrict__ f,
                                                     float* __restrict__ ll, float* __restrict__ z3) {
  __shared__ float sc[4][GG_MLP_H];
  const int wave = threadIdx.x >> 6;
  const int l = threadIdx.x & 63;
  const int b = blockIdx.x * 4 + wave;
  const bool live = b < nb;
  if (live) {
    float fs[GG_MLP_NF - 1];
    for (int j = 0; j < GG_MLP_NF - 1; ++j) fs[j] = f[(int64_t)b * (GG_MLP_NF - 1) + j];
    sc[wave][l] = gg_mlp_c(*w, l, fs);
  }
  __syncthreads();
  if (!live) return;
  gg_bf16x8 a1[4];
  gg_mlp_a1(w->w_e, sc[wave], a1);
  gg_mlp_regs wr;
  gg_mlp_load_regs(*w, wr);
  const float ev = (l < 16) ? e[(int64_t)b * 16 + l] : 0.0f;
  gg_mlp_act act;
  gg_mlp_fwd16(wr, a1, gg_mlp_xe(ev), act);
  if (l < 16) {
    ll[(int64_t)b * 16 + l] = gg_mlp_ll(ev, act.z30, act.z31);
    z3[((int64_t)b * 16 + l) * 2 + 0] = act.z30;
    z3[((int64_t)b * 16 + l) * 2 + 1] = act.z31;
  }
}

#define GG_MLP_K 64       // candidates per training epoch
#define GG_MLP_SMAX 16    // satellites per training epoch (at most)
#define GG_MLP_TAU 2.0f   // target kernel width [m]
#define GG_PURPOSE_MLP_CAND 7

// Candidate receiver offsets of a training batch (candidate 0 = the truth) and
// the batch's epoch indices, from the handle's Philox stream.
__global__ void gg_k_mlp_cands(uint64_t seed, uint32_t step, int batch, int n_train, float sh,
                               float sv, float scb, float* cand, int32_t* idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch * GG_MLP_K) return;
  const int b = i / GG_MLP_K, k = i % GG_MLP_K;
  float z[4];
  gg_normal4(seed, (uint32_t)i, step, GG_PURPOSE_MLP_CAND, 0u, z);
  float* c = cand + (int64_t)i * 4;
  c[0] = (k == 0) ? 0.0f : sh * z[0];
  c[1] = (k == 0) ? 0.0f : sh * z[1];
  c[2] = (k == 0) ? 0.0f : sv * z[2];
  c[3] = (k == 0) ? 0.0f : scb * z[3];
  if (k == 0) {
    const gg_u32x4 r = gg_draw(seed, (uint32_t)b, step, GG_PURPOSE_MLP_CAND, 1u);
    idx[b] = (int32_t)(((uint64_t)r.v[0] * (uint64_t)(uint32_t)n_train) >> 32);
  }
}

// e = r0 + u . d - dcb in binary32, in the oracle's order.
__device__ __forceinline__ float gg_cand_residual(float r0, const float* u, const float* c) {
#pragma clang fp contract(off)
  float ud = u[0] * c[0] + u[1] * c[1];
  ud = ud + u[2] * c[2];
  return (r0 + ud) - c[3];
}

__device__ __forceinline__ float gg_sum16(float v) {
#pragma unroll
  for (int d = 1; d < 16; d <<= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// One workgroup per training epoch (4 wavefronts): forward over the S x 64
// (satellite, candidate) pairs, softmax over the candidates, backward.  Each
// wavefront writes its gradient (scaled by inv_batch) to partial[4 b + wave].
// Wavefront w handles satellites w, w+4, w+8, w+12, 16 candidates per pass.
__global__ __launch_bounds__(256) void gg_k_mlp_train(
    const gg_mlp_w* __restrict__ w, const float* __restrict__ tf, const float* __restrict__ tr0,
    const float* __restrict__ tu, const int32_t* __restrict__ tns, int s_max,
    const int32_t* __restrict__ idx, const float* __restrict__ cand, float inv_batch,
    float* __restrict__ partial, float* __restrict__ loss) {
#pragma clang fp contract(off)
  __shared__ float sf[GG_MLP_SMAX][GG_MLP_NF - 1];
  __shared__ float sr0[GG_MLP_SMAX];
  __shared__ float su[GG_MLP_SMAX][3];
  __shared__ float sc[GG_MLP_SMAX][GG_MLP_H];
  __shared__ float sll[GG_MLP_SMAX][GG_MLP_K];
  __shared__ float scand[GG_MLP_K][4];
  __shared__ float sdL[GG_MLP_K];
  __shared__ float sdc[GG_MLP_SMAX][GG_MLP_H];
  __shared__ __align__(16) __bf16 stz[4][GG_MLP_H][32];  // dz2^T staging per wave
  __shared__ __align__(16) __bf16 sth[4][GG_MLP_H][32];  // h1^T staging per wave
  const int b = blockIdx.x;
  const int t = threadIdx.x;
  const int wave = t >> 6;
  const int l = t & 63;
  const int g = l >> 4;
  const int ep = idx ? idx[b] : b;
  const int ns = min(tns[ep], GG_MLP_SMAX);
  for (int i = t; i < ns * (GG_MLP_NF - 1); i += 256)
    sf[i / (GG_MLP_NF - 1)][i % (GG_MLP_NF - 1)] =
        tf[((int64_t)ep * s_max + i / (GG_MLP_NF - 1)) * (GG_MLP_NF - 1) + i % (GG_MLP_NF - 1)];
  for (int i = t; i < ns; i += 256) sr0[i] = tr0[(int64_t)ep * s_max + i];
  for (int i = t; i < ns * 3; i += 256) su[i / 3][i % 3] = tu[((int64_t)ep * s_max + i / 3) * 3 + i % 3];
  for (int i = t; i < GG_MLP_K * 4; i += 256)
    scand[i / 4][i % 4] = cand[(int64_t)b * GG_MLP_K * 4 + i];
  for (int i = t; i < GG_MLP_SMAX * GG_MLP_H; i += 256) sdc[i / GG_MLP_H][i % GG_MLP_H] = 0.0f;
  __syncthreads();
  for (int i = t; i < ns * GG_MLP_H; i += 256)
    sc[i / GG_MLP_H][i % GG_MLP_H] = gg_mlp_c(*w, i % GG_MLP_H, sf[i / GG_MLP_H]);
  __syncthreads();
  gg_mlp_regs wr;
  gg_mlp_load_regs(*w, wr);
  // ---- forward
  for (int s = wave; s < ns; s += 4) {
    gg_bf16x8 a1[4];
    gg_mlp_a1(w->w_e, sc[s], a1);
    for (int cb = 0; cb < GG_MLP_K / 16; ++cb) {
      const int k = 16 * cb + (l & 15);
      const float e = gg_cand_residual(sr0[s], su[s], scand[k]);
      gg_mlp_act act;
      gg_mlp_fwd16(wr, a1, gg_mlp_xe(e), act);
      if (l < 16) sll[s][k] = gg_mlp_ll(e, act.z30, act.z31);
    }
  }
  __syncthreads();
  // ---- softmax over the candidates against the target kernel (wave 0)
  if (wave == 0) {
    const int k = l;
    float L = 0.0f;
    for (int s = 0; s < ns; ++s) L = L + sll[s][k];
    const float* c = scand[k];
    const float d2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2] + c[3] * c[3];
    float m = L, dmin = d2;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      m = fmaxf(m, __shfl_xor(m, d, 64));
      dmin = fminf(dmin, __shfl_xor(dmin, d, 64));
    }
    const float pe = __expf(L - m);
    const float te = __expf(-(d2 - dmin) / (2.0f * GG_MLP_TAU * GG_MLP_TAU));
    float sp = pe, st = te;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      sp += __shfl_xor(sp, d, 64);
      st += __shfl_xor(st, d, 64);
    }
    const float pk = pe / sp, tk = te / st;
    sdL[k] = (pk - tk) * inv_batch;
    float lk = -tk * ((L - m) - __logf(sp));
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) lk += __shfl_xor(lk, d, 64);
    if (k == 0) loss[b] = lk;
  }
  __syncthreads();
  // ---- backward
  gg_f32x4 aw2[4][4];
  float adb2[16], adw3[2][16], adwe[16], adb3[2] = {0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    adb2[i] = 0.0f;
    adw3[0][i] = adw3[1][i] = 0.0f;
    adwe[i] = 0.0f;
  }
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) aw2[mb][nb] = (gg_f32x4){0.0f, 0.0f, 0.0f, 0.0f};
  for (int s = wave; s < ns; s += 4) {
    gg_bf16x8 a1[4];
    gg_mlp_a1(w->w_e, sc[s], a1);
    float adc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) adc[i] = 0.0f;
    for (int cb = 0; cb < GG_MLP_K / 16; ++cb) {
      const int k = 16 * cb + (l & 15);
      const float e = gg_cand_residual(sr0[s], su[s], scand[k]);
      const float xe = gg_mlp_xe(e);
      gg_mlp_act act;
      gg_mlp_fwd16(wr, a1, xe, act);
      float dz30 = 0.0f, dz31 = 0.0f;
      if (l < 16) {
        const float gl = sdL[k];
        const float mu = GG_MLP_MU_SCALE * act.z30;
        const float lv = fminf(fmaxf(act.z31, GG_MLP_LV_MIN), GG_MLP_LV_MAX);
        const float r = e - mu;
        const float ex = __expf(-lv);
        dz30 = GG_MLP_MU_SCALE * r * ex * gl;
        const bool inside = act.z31 > GG_MLP_LV_MIN && act.z31 < GG_MLP_LV_MAX;
        dz31 = inside ? (0.5f * r * r * ex - 0.5f) * gl : 0.0f;
        adb3[0] += dz30;
        adb3[1] += dz31;
      }
      dz30 = __shfl(dz30, l & 15, 64);
      dz31 = __shfl(dz31, l & 15, 64);
      const float xeb = (float)(__bf16)__shfl(xe, l & 15, 64);
      gg_f32x4 dz2c[4];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int u = 16 * mb + 4 * g + j;
          const float h2 = (float)act.h2[mb >> 1][4 * (mb & 1) + j];
          adw3[0][4 * mb + j] += dz30 * h2;
          adw3[1][4 * mb + j] += dz31 * h2;
          const float dh2 = w->W3b[0][u] * dz30 + w->W3b[1][u] * dz31;
          const float dz2 = (act.z2[mb][j] > 0.0f) ? dh2 : 0.0f;
          dz2c[mb][j] = dz2;
          adb2[4 * mb + j] += dz2;
        }
      gg_bf16x8 bz2[2];
#pragma unroll
      for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int q = 0; q < 8; ++q) bz2[ks][q] = (__bf16)dz2c[2 * ks + (q >> 2)][q & 3];
      const gg_f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        gg_f32x4 dh1 = gg_mfma(w->A2t[mb][0][l], bz2[0], zero);
        dh1 = gg_mfma(w->A2t[mb][1][l], bz2[1], dh1);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float dz1 = (act.z1[mb][j] > 0.0f) ? dh1[j] : 0.0f;
          adwe[4 * mb + j] += dz1 * xeb;
          adc[4 * mb + j] += dz1;
        }
      }
      // dW2 += dz2^T h1 over pairs: stage 16 pairs per half, MFMA per 32 pairs
      const int col = 16 * (cb & 1) + (l & 15);
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int u = 16 * mb + 4 * g + j;
          stz[wave][u][col] = bz2[mb >> 1][4 * (mb & 1) + j];
          sth[wave][u][col] = act.h1[mb >> 1][4 * (mb & 1) + j];
        }
      if (cb & 1) {
        __threadfence_block();
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
          const gg_bf16x8 a = *reinterpret_cast<const gg_bf16x8*>(&stz[wave][16 * mb + (l & 15)][8 * g]);
#pragma unroll
          for (int nb = 0; nb < 4; ++nb) {
            const gg_bf16x8 bb =
                *reinterpret_cast<const gg_bf16x8*>(&sth[wave][16 * nb + (l & 15)][8 * g]);
            aw2[mb][nb] = gg_mfma(a, bb, aw2[mb][nb]);
          }
        }
        __threadfence_block();
        __builtin_amdgcn_wave_barrier();
      }
    }
    // per-satellite first-layer bias gradient
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const float v = gg_sum16(adc[i]);
      if ((l & 15) == 0) sdc[s][16 * (i >> 2) + 4 * g + (i & 3)] = v;
    }
  }
  // ---- write this wavefront's gradient
  float* out = partial + (int64_t)(4 * b + wave) * GG_MLP_NPARAMS;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        out[GG_MLP_OFF_W2 + (16 * mb + 4 * g + j) * GG_MLP_H + 16 * nb + (l & 15)] = aw2[mb][nb][j];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int u = 16 * (i >> 2) + 4 * g + (i & 3);
    const float db2 = gg_sum16(adb2[i]);
    const float dw30 = gg_sum16(adw3[0][i]);
    const float dw31 = gg_sum16(adw3[1][i]);
    const float dwe = gg_sum16(adwe[i]);
    if ((l & 15) == 0) {
      out[GG_MLP_OFF_B2 + u] = db2;
      out[GG_MLP_OFF_W3 + u] = dw30;
      out[GG_MLP_OFF_W3 + GG_MLP_H + u] = dw31;
      out[GG_MLP_OFF_W1 + u * GG_MLP_NF] = dwe;
    }
  }
  {
    const float d0 = gg_sum16(l < 16 ? adb3[0] : 0.0f);
    const float d1 = gg_sum16(l < 16 ? adb3[1] : 0.0f);
    if (l == 0) {
      out[GG_MLP_OFF_B3] = d0;
      out[GG_MLP_OFF_B3 + 1] = d1;
    }
  }
  __syncthreads();
  // first-layer weights of the satellite features and b1 (wave 0), zeros elsewhere
  if (wave == 0) {
    const int u = l;
    float db1 = 0.0f;
    float dw[GG_MLP_NF - 1];
#pragma unroll
    for (int j = 0; j < GG_MLP_NF - 1; ++j) dw[j] = 0.0f;
    for (int s = 0; s < ns; ++s) {
      const float dc = sdc[s][u];
      db1 += dc;
#pragma unroll
      for (int j = 0; j < GG_MLP_NF - 1; ++j) dw[j] += dc * sf[s][j];
    }
    out[GG_MLP_OFF_B1 + u] = db1;
#pragma unroll
    for (int j = 0; j < GG_MLP_NF - 1; ++j) out[GG_MLP_OFF_W1 + u * GG_MLP_NF + 1 + j] = dw[j];
  } else {
    out[GG_MLP_OFF_B1 + l] = 0.0f;
#pragma unroll
    for (int j = 0; j < GG_MLP_NF - 1; ++j) out[GG_MLP_OFF_W1 + l * GG_MLP_NF + 1 + j] = 0.0f;
  }
}

// grad = sum of the partial rows in order; optionally an Adam step.
__global__ void gg_k_mlp_reduce(const float* __restrict__ partial, int rows,
                                const float* __restrict__ loss, int n_loss, float* grad,
                                float* params, float* m, float* v, int adam_t, float lr,
                                float* loss_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0 && loss_out) {
    float s = 0.0f;
    for (int i = 0; i < n_loss; ++i) s += loss[i];
    *loss_out = s / (float)n_loss;
  }
  if (p >= GG_MLP_NPARAMS) return;
  float gsum = 0.0f;
  for (int r = 0; r < rows; ++r) gsum += partial[(int64_t)r * GG_MLP_NPARAMS + p];
  grad[p] = gsum;
  if (adam_t > 0) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float mi = b1 * m[p] + (1.0f - b1) * gsum;
    const float vi = b2 * v[p] + (1.0f - b2) * gsum * gsum;
    m[p] = mi;
    v[p] = vi;
    const float mh = mi / (1.0f - powf(b1, (float)adam_t));
    const float vh = vi / (1.0f - powf(b2, (float)adam_t));
    params[p] -= lr * mh / (sqrtf(vh) + 1e-8f);
  }
}

int pack(gg_mlp* h) {
  hipLaunchKernelGGL(gg_k_mlp_pack, dim3(1), dim3(256), 0, h->stream, h->params, h->w, h->fp8);
  GG_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int alloc(T** p, size_t n) {
  *p = nullptr;
  GG_HIP(hipMalloc((void**)p, (n ? n : 1) * sizeof(T)));
  // the handle's stream does not synchronise with the null stream
  GG_HIP(hipMemsetAsync(*p, 0, (n ? n : 1) * sizeof(T), nullptr));
  GG_HIP(hipStreamSynchronize(nullptr));
  return 0;
}

}  // namespace

const void* gg_mlp_packed_weights(const gg_mlp* m) { return m ? (const void*)m->w : nullptr; }

extern "C" {

int gg_mlp_create(gg_mlp** out) {
  if (!out) return gg_fail(-2, "null argument");
  *out = nullptr;
  gg_mlp* h = new gg_mlp();
  int rc = 0;
  if (!rc) rc = alloc(&h->params, GG_MLP_NPARAMS);
  if (!rc) rc = alloc(&h->w, 1);
  if (!rc) rc = alloc(&h->adam_m, GG_MLP_NPARAMS);
  if (!rc) rc = alloc(&h->adam_v, GG_MLP_NPARAMS);
  if (!rc) rc = alloc(&h->grad, GG_MLP_NPARAMS);
  if (!rc && hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    rc = gg_fail(-1, "hipStreamCreate failed");
  if (rc) {
    gg_mlp_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

int gg_mlp_destroy(gg_mlp* h) {
  if (!h) return 0;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  void* bufs[] = {h->params, h->w,  h->adam_m, h->adam_v, h->grad,  h->partial,
                  h->loss,   h->tf, h->tr0,    h->tu,     h->tns,   h->cand,
                  h->batch_idx};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int gg_mlp_set_params(gg_mlp* h, const float* p) {
  if (!h || !p) return gg_fail(-2, "null argument");
  GG_HIP(hipMemcpyAsync(h->params, p, sizeof(float) * GG_MLP_NPARAMS, hipMemcpyHostToDevice,
                        h->stream));
  GG_HIP(hipMemsetAsync(h->adam_m, 0, sizeof(float) * GG_MLP_NPARAMS, h->stream));
  GG_HIP(hipMemsetAsync(h->adam_v, 0, sizeof(float) * GG_MLP_NPARAMS, h->stream));
  h->adam_t = 0;
  if (pack(h)) return -1;
  GG_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int gg_mlp_set_fp8(gg_mlp* h, int32_t enable) {
  if (!h) return gg_fail(-2, "null handle");
  h->fp8 = enable ? 1 : 0;
  if (pack(h)) return -1;
  GG_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int gg_mlp_get_params(gg_mlp* h, float* p) {
  if (!h || !p) return gg_fail(-2, "null argument");
  GG_HIP(hipMemcpyAsync(p, h->params, sizeof(float) * GG_MLP_NPARAMS, hipMemcpyDeviceToHost,
                        h->stream));
  GG_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int gg_mlp_eval(gg_mlp* h, int32_t n_groups, const float* e, const float* f, float* ll,
                float* z3) {
  if (!h || !e || !f || !ll || !z3) return gg_fail(-2, "null argument");
  if (n_groups <= 0) return gg_fail(-2, "n_groups must be positive");
  const size_t ne = (size_t)n_groups * 16;
  float *de = nullptr, *df = nullptr, *dll = nullptr, *dz = nullptr;
  int rc = 0;
  if (!rc) rc = alloc(&de, ne);
  if (!rc) rc = alloc(&df, (size_t)n_groups * (GG_MLP_NF - 1));
  if (!rc) rc = alloc(&dll, ne);
  if (!rc) rc = alloc(&dz, ne * 2);
  if (!rc && (hipMemcpy(de, e, ne * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(df, f, (size_t)n_groups * (GG_MLP_NF - 1) * sizeof(float),
                        hipMemcpyHostToDevice) != hipSuccess))
    rc = gg_fail(-1, "hipMemcpy failed");
  if (!rc) {
    hipLaunchKernelGGL(gg_k_mlp_eval, dim3((n_groups + 3) / 4), dim3(256), 0, h->stream, h->w,
                       n_groups, de, df, dll, dz);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess)
      rc = gg_fail(-1, "mlp evaluation failed");
  }
  if (!rc && (hipMemcpy(ll, dll, ne * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(z3, dz, ne * 2 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess))
    rc = gg_fail(-1, "hipMemcpy failed");
  for (float* b : {de, df, dll, dz})
    if (b) (void)hipFree(b);
  return rc;
}

int gg_mlp_load_data(gg_mlp* h, int32_t n_epochs, int32_t s_max, const float* f,
                     const float* r0, const float* u, const int32_t* n_sat) {
  if (!h || !f || !r0 || !u || !n_sat) return gg_fail(-2, "null argument");
  if (n_epochs <= 0) return gg_fail(-2, "n_epochs must be positive");
  if (s_max <= 0 || s_max > GG_MLP_SMAX) return gg_fail(-2, "s_max must be in [1, 16]");
  for (int e = 0; e < n_epochs; ++e)
    if (n_sat[e] < 0 || n_sat[e] > s_max) return gg_fail(-2, "n_sat must be in [0, s_max]");
  GG_HIP(hipStreamSynchronize(h->stream));
  for (void* b : {(void*)h->tf, (void*)h->tr0, (void*)h->tu, (void*)h->tns})
    if (b) GG_HIP(hipFree(b));
  h->tf = h->tr0 = h->tu = nullptr;
  h->tns = nullptr;
  const size_t es = (size_t)n_epochs * s_max;
  int rc = 0;
  if (!rc) rc = alloc(&h->tf, es * (GG_MLP_NF - 1));
  if (!rc) rc = alloc(&h->tr0, es);
  if (!rc) rc = alloc(&h->tu, es * 3);
  if (!rc) rc = alloc(&h->tns, (size_t)n_epochs);
  if (rc) return rc;
  GG_HIP(hipMemcpy(h->tf, f, es * (GG_MLP_NF - 1) * sizeof(float), hipMemcpyHostToDevice));
  GG_HIP(hipMemcpy(h->tr0, r0, es * sizeof(float), hipMemcpyHostToDevice));
  GG_HIP(hipMemcpy(h->tu, u, es * 3 * sizeof(float), hipMemcpyHostToDevice));
  GG_HIP(hipMemcpy(h->tns, n_sat, (size_t)n_epochs * sizeof(int32_t), hipMemcpyHostToDevice));
  h->n_train = n_epochs;
  h->s_max = s_max;
  return 0;
}

static int ensure_batch(gg_mlp* h, int batch) {
  if (batch <= h->cand_cap) return 0;
  GG_HIP(hipStreamSynchronize(h->stream));
  for (void* b : {(void*)h->cand, (void*)h->batch_idx, (void*)h->partial, (void*)h->loss})
    if (b) GG_HIP(hipFree(b));
  h->cand = nullptr;
  h->batch_idx = nullptr;
  h->partial = nullptr;
  h->loss = nullptr;
  int rc = 0;
  if (!rc) rc = alloc(&h->cand, (size_t)batch * GG_MLP_K * 4);
  if (!rc) rc = alloc(&h->batch_idx, (size_t)batch);
  if (!rc) rc = alloc(&h->partial, (size_t)batch * 4 * GG_MLP_NPARAMS);
  if (!rc) rc = alloc(&h->loss, (size_t)batch);
  if (rc) return rc;
  h->cand_cap = batch;
  return 0;
}

int gg_mlp_train(gg_mlp* h, int32_t n_steps, int32_t batch, float lr, uint64_t seed,
                 float cand_sigma_h, float cand_sigma_v, float cand_sigma_cb, float* loss_out) {
  if (!h) return gg_fail(-2, "null handle");
  if (h->n_train <= 0) return gg_fail(-3, "no training data loaded");
  if (n_steps <= 0 || batch <= 0 || batch > 65536) return gg_fail(-2, "bad n_steps or batch");
  if (ensure_batch(h, batch)) return -1;
  float* dloss = nullptr;
  if (alloc(&dloss, (size_t)n_steps)) return -1;
  int rc = 0;
  for (int it = 0; it < n_steps && !rc; ++it) {
    const uint32_t step = (uint32_t)(h->adam_t + 1);
    hipLaunchKernelGGL(gg_k_mlp_cands, dim3((batch * GG_MLP_K + 255) / 256), dim3(256), 0,
                       h->stream, seed, step, batch, h->n_train, cand_sigma_h, cand_sigma_v,
                       cand_sigma_cb, h->cand, h->batch_idx);
    hipLaunchKernelGGL(gg_k_mlp_train, dim3(batch), dim3(256), 0, h->stream, h->w, h->tf,
                       h->tr0, h->tu, h->tns, h->s_max, h->batch_idx, h->cand,
                       1.0f / (float)batch, h->partial, h->loss);
    h->adam_t += 1;
    hipLaunchKernelGGL(gg_k_mlp_reduce, dim3((GG_MLP_NPARAMS + 255) / 256), dim3(256), 0,
                       h->stream, h->partial, 4 * batch, h->loss, batch, h->grad, h->params,
                       h->adam_m, h->adam_v, h->adam_t, lr, dloss + it);
    if (hipGetLastError() != hipSuccess) rc = gg_fail(-1, "training launch failed");
    if (!rc) rc = pack(h);
  }
  if (!rc && loss_out &&
      hipMemcpyAsync(loss_out, dloss, sizeof(float) * n_steps, hipMemcpyDeviceToHost,
                     h->stream) != hipSuccess)
    rc = gg_fail(-1, "hipMemcpy failed");
  if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = gg_fail(-1, "training failed");
  (void)hipFree(dloss);
  return rc;
}

int gg_mlp_loss_grad(gg_mlp* h, int32_t n_epochs, const float* cand, float* grad,
                     float* loss) {
  if (!h || !cand || !grad || !loss) return gg_fail(-2, "null argument");
  if (h->n_train <= 0) return gg_fail(-3, "no training data loaded");
  if (n_epochs <= 0 || n_epochs > h->n_train) return gg_fail(-2, "n_epochs out of range");
  if (ensure_batch(h, n_epochs)) return -1;
  GG_HIP(hipMemcpyAsync(h->cand, cand, sizeof(float) * (size_t)n_epochs * GG_MLP_K * 4,
                        hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(gg_k_mlp_train, dim3(n_epochs), dim3(256), 0, h->stream, h->w, h->tf,
                     h->tr0, h->tu, h->tns, h->s_max, (const int32_t*)nullptr, h->cand, 1.0f,
                     h->partial, h->loss);
  hipLaunchKernelGGL(gg_k_mlp_reduce, dim3((GG_MLP_NPARAMS + 255) / 256), dim3(256), 0,
                     h->stream, h->partial, 4 * n_epochs, h->loss, n_epochs, h->grad, h->params,
                     h->adam_m, h->adam_v, 0, 0.0f, (float*)nullptr);
  GG_HIP(hipGetLastError());
  GG_HIP(hipMemcpyAsync(grad, h->grad, sizeof(float) * GG_MLP_NPARAMS, hipMemcpyDeviceToHost,
                        h->stream));
  GG_HIP(hipMemcpyAsync(loss, h->loss, sizeof(float) * n_epochs, hipMemcpyDeviceToHost,
                        h->stream));
  GG_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

}  // extern "C"
