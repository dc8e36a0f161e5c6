// Shared device-side building blocks for the particle-filter and WLS kernels.
//
// Everything in this header that feeds the resampling decision (the f64 exp,
// the fixed-order block scans, Philox) is written so that a NumPy twin using
// only IEEE-754 +,-,*,/ reproduces it bit for bit: floating-point contraction
// is switched off for these functions, and every reduction/scan uses one fixed
// association order that does not depend on scheduling.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

#include "gnss_gpu_amd_layout.h"

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel Random Numbers: As Easy
// as 1, 2, 3", SC'11).  Counter-based: a draw is a pure function of
// (counter, key), so any launch geometry gives the same numbers.
// ---------------------------------------------------------------------------
struct gg_u32x4 {
  uint32_t v[4];
};

__host__ __device__ __forceinline__ gg_u32x4 gg_philox4x32_10(uint32_t c0, uint32_t c1,
                                                              uint32_t c2, uint32_t c3,
                                                              uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  gg_u32x4 out;
  out.v[0] = c0;
  out.v[1] = c1;
  out.v[2] = c2;
  out.v[3] = c3;
  return out;
}

// Draw `call` of stream (seed, purpose) for element `idx` at step `step`.
__host__ __device__ __forceinline__ gg_u32x4 gg_draw(uint64_t seed, uint32_t idx, uint32_t step,
                                                     uint32_t purpose, uint32_t call) {
  return gg_philox4x32_10(idx, step, purpose, call, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// uint32 -> float uniform in the open interval (0, 1): the top 23 bits plus
// one half, scaled by 2^-23, so every value is exact in binary32 and the
// largest is 1 - 2^-24.
__host__ __device__ __forceinline__ float gg_u01f(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  // the same value in two instructions: a funnel shift puts 0x3f800000 above
  // the top 23 bits (1 + m 2^-23 in [1, 2)), then one exact subtraction of
  // 1 - 2^-24 leaves (m + 0.5) 2^-23
  const float f = __uint_as_float(__builtin_amdgcn_alignbit(0x7fu, x, 9u));
  return f - 0.99999994039535522f;  // 1 - 2^-24
#else
  return ((float)(x >> 9) + 0.5f) * 1.1920928955078125e-07f;  // 2^-23
#endif
}

// two uint32 -> double uniform in (0, 1]; exact in binary64.
__host__ __device__ __forceinline__ double gg_u01d(uint32_t hi, uint32_t lo) {
  const uint64_t m = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
  return (double)(m + 1) * 1.1102230246251565e-16;  // 2^-53
}

// Box-Muller pair from two (0,1) uniforms, binary32, on the hardware
// transcendental units: v_log_f32 (log2), v_sqrt_f32, and v_sin/v_cos_f32,
// whose argument is in revolutions, so sin(2 pi u) is one instruction.
// Relative error of the normals is about 1e-6 (noise generation only).
__device__ __forceinline__ void gg_box_muller(float u1, float u2, float& z0, float& z1) {
  const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
  z0 = r * __builtin_amdgcn_cosf(u2);
  z1 = r * __builtin_amdgcn_sinf(u2);
}

// Four standard normals from one Philox block.
__device__ __forceinline__ void gg_normal4(uint64_t seed, uint32_t idx, uint32_t step,
                                           uint32_t purpose, uint32_t call, float z[4]) {
  const gg_u32x4 r = gg_draw(seed, idx, step, purpose, call);
  gg_box_muller(gg_u01f(r.v[0]), gg_u01f(r.v[1]), z[0], z[1]);
  gg_box_muller(gg_u01f(r.v[2]), gg_u01f(r.v[3]), z[2], z[3]);
}

// ---------------------------------------------------------------------------
// xoshiro128+ (Blackman and Vigna, "Scrambled linear pseudorandom number
// generators", ACM TOMS 2021), seeded per stream by one Philox4x32-10 block.
// Long runs of uniforms for one thread (the predict noise of its particles,
// a Megopolis chain's acceptance draws) cost one Philox block plus about
// seven full-rate 32-bit operations per output, where a Philox block per four
// outputs costs twenty 64-bit multiplies.  The top 23 bits feed gg_u01f; the
// generator's weak lowest bits are never used.  The stream stays a pure
// function of (seed, index, step, purpose, call): the seed block is counter-
// based, the state advances only inside one thread.
// ---------------------------------------------------------------------------
struct gg_xo {
  uint32_t s0, s1, s2, s3;
};

__host__ __device__ __forceinline__ gg_xo gg_xo_seed(uint64_t seed, uint32_t idx, uint32_t step,
                                                     uint32_t purpose, uint32_t call) {
  const gg_u32x4 r = gg_draw(seed, idx, step, purpose, call);
  gg_xo x;
  x.s0 = r.v[0];
  x.s1 = r.v[1];
  x.s2 = r.v[2];
  x.s3 = r.v[3];
  // the all-zero state is the generator's one fixed point
  x.s0 |= ((x.s0 | x.s1 | x.s2 | x.s3) == 0u) ? 1u : 0u;
  return x;
}

__host__ __device__ __forceinline__ uint32_t gg_xo_next(gg_xo& x) {
  const uint32_t out = x.s0 + x.s3;
  const uint32_t t = x.s1 << 9;
#if defined(__HIP_DEVICE_COMPILE__)
  // the same update with three-input exclusive ors (v_bitop3_b32, truth
  // table 0x96): s1 ^ s2 ^ s0 and s2 ^ s0 ^ t without the shared s2 ^ s0
  const uint32_t s1 = __builtin_amdgcn_bitop3_b32(x.s1, x.s2, x.s0, 0x96);
  const uint32_t s2 = __builtin_amdgcn_bitop3_b32(x.s2, x.s0, t, 0x96);
  const uint32_t s3 = x.s3 ^ x.s1;
  x.s0 = x.s0 ^ s3;
  x.s1 = s1;
  x.s2 = s2;
  x.s3 = (s3 << 11) | (s3 >> 21);
#else
  x.s2 ^= x.s0;
  x.s3 ^= x.s1;
  x.s1 ^= x.s2;
  x.s0 ^= x.s3;
  x.s2 ^= t;
  x.s3 = (x.s3 << 11) | (x.s3 >> 21);
#endif
  return out;
}

// Two standard normals (one Box-Muller pair) from the next two outputs.
__device__ __forceinline__ void gg_xo_normal2(gg_xo& x, float& z0, float& z1) {
  const float u1 = gg_u01f(gg_xo_next(x));
  const float u2 = gg_u01f(gg_xo_next(x));
  gg_box_muller(u1, u2, z0, z1);
}

// ---------------------------------------------------------------------------
// Deterministic binary64 exp.  Cody-Waite reduction x = k ln2 + r with ln2
// split into a 32-bit head and a tail, then the degree-13 Taylor polynomial of
// exp(r) on |r| <= ln2/2 evaluated by Horner, then exact scaling by 2^k.
// Only correctly rounded IEEE operations are used, in a fixed order, so the
// NumPy twin in the oracle reproduces every bit.  Max error about 1 ulp.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ double gg_exp64(double x) {
#pragma clang fp contract(off)
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_huge_val();
  if (x < -745.1332191019412) return 0.0;
  const double k = __builtin_rint(x * 0x1.71547652b82fep+0);
  const double r = (x - k * 0x1.62e42fee00000p-1) - k * 0x1.a39ef35793c76p-33;
  double p = 0x1.6124613a86d09p-33;
  p = p * r + 0x1.1eed8eff8d898p-29;
  p = p * r + 0x1.ae64567f544e4p-26;
  p = p * r + 0x1.27e4fb7789f5cp-22;
  p = p * r + 0x1.71de3a556c734p-19;
  p = p * r + 0x1.a01a01a01a01ap-16;
  p = p * r + 0x1.a01a01a01a01ap-13;
  p = p * r + 0x1.6c16c16c16c17p-10;
  p = p * r + 0x1.1111111111111p-7;
  p = p * r + 0x1.5555555555555p-5;
  p = p * r + 0x1.5555555555555p-3;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return __builtin_ldexp(p, (int)k);
}

// Deterministic binary32 exp with the same construction (16-bit ln2 head,
// degree-7 Taylor polynomial), every step a correctly rounded fused
// multiply-add: r = (x - k ln2_hi) - k ln2_lo and the Horner steps, 13
// instructions on the exact path.  The oracle's twin emulates each fused
// multiply-add exactly (binary64 sum, round to odd, then to binary32).
__host__ __device__ __forceinline__ float gg_exp32(float x) {
#pragma clang fp contract(off)
  if (x != x) return x;
  if (x > 88.72283f) return __builtin_huge_valf();
  if (x < -103.97208f) return 0.0f;
  const float k = __builtin_rintf(x * 0x1.715476p+0f);
  const float r = __builtin_fmaf(-k, 0x1.7f7d1cp-20f, __builtin_fmaf(-k, 0x1.62e4p-1f, x));
  float p = 0x1.a01a02p-13f;
  p = __builtin_fmaf(p, r, 0x1.6c16c2p-10f);
  p = __builtin_fmaf(p, r, 0x1.111112p-7f);
  p = __builtin_fmaf(p, r, 0x1.555556p-5f);
  p = __builtin_fmaf(p, r, 0x1.555556p-3f);
  p = __builtin_fmaf(p, r, 0.5f);
  p = __builtin_fmaf(p, r, 1.0f);
  p = __builtin_fmaf(p, r, 1.0f);
  return __builtin_ldexpf(p, (int)k);
}

// ---------------------------------------------------------------------------
// Fixed-order block scan of 256 doubles (4 wavefronts of 64 lanes).
//   1. Hillis-Steele inclusive scan inside each wavefront:
//        for d in 1,2,4,8,16,32: x[l] = x[l-d] + x[l]   (l >= d)
//   2. wavefront totals T0..T3 combined left to right: O1=T0, O2=T0+T1,
//      O3=(T0+T1)+T2; lane value in wave w>0 becomes O_w + x.
// Returns the inclusive value of this thread; `total` gets thread 255's value
// and `excl` the inclusive value of thread t-1 (0 for t=0).
// `lds` must hold at least 8 + 256 doubles and is reused on return.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double gg_block_scan256(double v, double* lds, double& excl,
                                                   double& total) {
#pragma clang fp contract(off)
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  double x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double y = __shfl_up(x, d, 64);
    if (lane >= d) x = y + x;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  if (t == 0) {
    const double o1 = lds[0];
    const double o2 = o1 + lds[1];
    const double o3 = o2 + lds[2];
    lds[4] = o1;
    lds[5] = o2;
    lds[6] = o3;
    lds[7] = o3 + lds[3];
  }
  __syncthreads();
  if (w > 0) x = lds[3 + w] + x;
  lds[8 + t] = x;
  __syncthreads();
  excl = (t == 0) ? 0.0 : lds[8 + t - 1];
  total = lds[7];
  __syncthreads();
  return x;
}

// NV independent scans of the same form as gg_block_scan256 in one pass (the
// same operations per value, so each result equals gg_block_scan256's bit for
// bit).  `lds` must hold at least NV * (8 + 256) doubles.
template <int NV>
__device__ __forceinline__ void gg_block_scan256_n(const double (&v)[NV], double* lds,
                                                   double (&excl)[NV], double (&total)[NV]) {
#pragma clang fp contract(off)
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  double x[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) x[k] = v[k];
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const double y = __shfl_up(x[k], d, 64);
      if (lane >= d) x[k] = y + x[k];
    }
  }
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * 264 + w] = x[k];
  __syncthreads();
  if (t < NV) {
    double* l = lds + t * 264;
    const double o1 = l[0];
    const double o2 = o1 + l[1];
    const double o3 = o2 + l[2];
    l[4] = o1;
    l[5] = o2;
    l[6] = o3;
    l[7] = o3 + l[3];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double* l = lds + k * 264;
    if (w > 0) x[k] = l[3 + w] + x[k];
    l[8 + t] = x[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double* l = lds + k * 264;
    excl[k] = (t == 0) ? 0.0 : l[8 + t - 1];
    total[k] = l[7];
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Cross-lane exchanges on the DPP network (register to register, no LDS
// traffic).  dpp_ctrl values of the gfx9 encoding:
//   0xB1  quad_perm [1,0,3,2]   lane ^ 1
//   0x4E  quad_perm [2,3,0,1]   lane ^ 2
//   0x141 row_half_mirror       lane i <-> 7 - i inside each group of 8
//   0x140 row_mirror            lane i <-> 15 - i inside each row of 16
//   0x111, 0x112, 0x114, 0x118  row_shr 1, 2, 4, 8 (inside rows)
//   0x142, 0x143                row_bcast 15, 31
// After the lane^1 and lane^2 steps every lane of a quad holds the same
// value, so the half-mirror partner stands in for lane ^ 4; likewise the
// mirror partner for lane ^ 8.  ds_swizzle in bit mode (and 0x1f, xor 0x10)
// gives lane ^ 16 inside each half of the wavefront.
// ---------------------------------------------------------------------------
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int gg_dpp_i32(int x, int old = 0) {
  return __builtin_amdgcn_update_dpp(old, x, CTRL, ROW_MASK, 0xF, false);
}

template <int CTRL>
__device__ __forceinline__ float gg_dpp_f32(float x) {
  return __builtin_bit_cast(float, gg_dpp_i32<CTRL>(__builtin_bit_cast(int, x)));
}

template <int CTRL>
__device__ __forceinline__ double gg_dpp_f64(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)gg_dpp_i32<CTRL>((int)(uint32_t)u);
  const uint32_t hi = (uint32_t)gg_dpp_i32<CTRL>((int)(uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint64_t)lo);
}

__device__ __forceinline__ int gg_swz16_i32(int x) { return __builtin_amdgcn_ds_swizzle(x, 0x401F); }

__device__ __forceinline__ float gg_swz16_f32(float x) {
  return __builtin_bit_cast(float, gg_swz16_i32(__builtin_bit_cast(int, x)));
}

__device__ __forceinline__ double gg_swz16_f64(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)gg_swz16_i32((int)(uint32_t)u);
  const uint32_t hi = (uint32_t)gg_swz16_i32((int)(uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint64_t)lo);
}

// Butterfly over 32 lanes: afterwards lanes 0..31 hold op over lanes 0..31
// and lanes 32..63 op over lanes 32..63, each in the balanced pairwise order.
#define GG_BUTTERFLY32(T, x, OP)          \
  do {                                    \
    x = OP(x, gg_dpp_##T<0xB1>(x));       \
    x = OP(x, gg_dpp_##T<0x4E>(x));       \
    x = OP(x, gg_dpp_##T<0x141>(x));      \
    x = OP(x, gg_dpp_##T<0x140>(x));      \
    x = OP(x, gg_swz16_##T(x));           \
  } while (0)

// Lane 0's value of a double, in every lane of the wavefront.
__device__ __forceinline__ double gg_readlane0_f64(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, 0);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 0);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint64_t)lo);
}

__device__ __forceinline__ double gg_add_f64(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ float gg_add_f32(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ double gg_max_f64(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float gg_max_f32(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ int gg_add_i32(int a, int b) { return a + b; }
__device__ __forceinline__ int gg_max_i32(int a, int b) { return max(a, b); }

// Block-wide max of a double (exact; any order gives the same result).
__device__ __forceinline__ double gg_block_max256(double v, double* lds) {
  const int t = threadIdx.x;
  GG_BUTTERFLY32(f64, v, gg_max_f64);
  if ((t & 31) == 0) lds[t >> 5] = v;
  __syncthreads();
  const double m = fmax(fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3])),
                        fmax(fmax(lds[4], lds[5]), fmax(lds[6], lds[7])));
  __syncthreads();
  return m;
}

// Block-wide maxima of two doubles in one pass (each as gg_block_max256).
__device__ __forceinline__ void gg_block_max256x2(double a, double b, double* lds, double& ma,
                                                  double& mb) {
  const int t = threadIdx.x;
  GG_BUTTERFLY32(f64, a, gg_max_f64);
  GG_BUTTERFLY32(f64, b, gg_max_f64);
  if ((t & 31) == 0) {
    lds[t >> 5] = a;
    lds[8 + (t >> 5)] = b;
  }
  __syncthreads();
  ma = fmax(fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3])),
            fmax(fmax(lds[4], lds[5]), fmax(lds[6], lds[7])));
  mb = fmax(fmax(fmax(lds[8], lds[9]), fmax(lds[10], lds[11])),
            fmax(fmax(lds[12], lds[13]), fmax(lds[14], lds[15])));
  __syncthreads();
}

// Block-wide max of a float (exact, NaN ignored as fmaxf does).
__device__ __forceinline__ float gg_block_max256f(float v, double* lds) {
  const int t = threadIdx.x;
  float* l = reinterpret_cast<float*>(lds);
  GG_BUTTERFLY32(f32, v, gg_max_f32);
  if ((t & 31) == 0) l[t >> 5] = v;
  __syncthreads();
  const float m = fmaxf(fmaxf(fmaxf(l[0], l[1]), fmaxf(l[2], l[3])),
                        fmaxf(fmaxf(l[4], l[5]), fmaxf(l[6], l[7])));
  __syncthreads();
  return m;
}

__device__ __forceinline__ int gg_block_max256_i(int v, int* lds) {
  const int t = threadIdx.x;
  GG_BUTTERFLY32(i32, v, gg_max_i32);
  if ((t & 31) == 0) lds[t >> 5] = v;
  __syncthreads();
  const int m = max(max(max(lds[0], lds[1]), max(lds[2], lds[3])),
                    max(max(lds[4], lds[5]), max(lds[6], lds[7])));
  __syncthreads();
  return m;
}

__device__ __forceinline__ int gg_block_sum256_i(int v, int* lds) {
  const int t = threadIdx.x;
  GG_BUTTERFLY32(i32, v, gg_add_i32);
  if ((t & 31) == 0) lds[t >> 5] = v;
  __syncthreads();
  const int m = ((lds[0] + lds[1]) + (lds[2] + lds[3])) + ((lds[4] + lds[5]) + (lds[6] + lds[7]));
  __syncthreads();
  return m;
}

// Inclusive prefix max of ints over the block (exact, order-free): row scan
// with row_shr 1, 2, 4, 8, then row_bcast 15 / 31 across the rows.
__device__ __forceinline__ int gg_block_prefix_max256_i(int v, int* lds) {
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  v = max(v, gg_dpp_i32<0x111>(v, INT_MIN));
  v = max(v, gg_dpp_i32<0x112>(v, INT_MIN));
  v = max(v, gg_dpp_i32<0x114>(v, INT_MIN));
  v = max(v, gg_dpp_i32<0x118>(v, INT_MIN));
  v = max(v, gg_dpp_i32<0x142, 0xA>(v, INT_MIN));
  v = max(v, gg_dpp_i32<0x143, 0xC>(v, INT_MIN));
  if (lane == 63) lds[w] = v;
  __syncthreads();
  int pre = INT_MIN;
  for (int k = 0; k < w; ++k) pre = max(pre, lds[k]);
  v = max(v, pre);
  __syncthreads();
  return v;
}
