// Particle-filter kernels for gfx950 (CDNA4, 64-lane wavefronts).
//
// One workgroup = 256 threads = 4 wavefronts; each thread owns GG_ITEMS (4)
// consecutive particles, so a workgroup covers one tile of GG_TILE (1024)
// particles and every SoA load/store is a 16-byte vector access that the
// wavefront coalesces into one contiguous 1 KiB transaction.
//
// The systematic step for a batch of filters is two launches, each over all
// filters:
//   gg_k_weight    predict (optional) + pseudorange likelihood + per-tile
//                  partials (max, sums of exp32(lw - m_b), ESS and mean sums).
//   gg_k_sysres    every work item combines its filter's partials (max, total,
//                  ESS, CDF tile offsets) itself and takes the resampling
//                  decision; the first item of each filter writes the decision
//                  and estimate; when resampling, each item finds its
//                  interval of outputs (one tile's, at most GG_QCHUNK) by a
//                  search over the merged tile ends and grid points, computes
//                  that tile's CDF ends, the ancestors and the state gather;
//                  the last item of each filter to
//                  finish adds the item sums and finalises (estimate, buffer
//                  swap).
// Besides the records and sums that pass through the device-coherent level
// (device-scope atomic stores and loads, two-level arrival counters; no L2
// writeback), kernel boundaries are the only cross-workgroup synchronisation:
// the L2s of the 8 XCDs are not coherent, an agent-scope release in every
// workgroup writes back its XCD's whole L2, and waiting on other workgroups
// would need every one of them resident.
// Megopolis: gg_k_weight, gg_k_decide (one workgroup per filter), gg_k_mego_w
// and gg_k_mego (chains, state copy and finalize), or B launches of
// gg_k_mego_ref and the ancestor-driven gather.
#pragma once

#include "gg_device.h"
#include "gg_mlp.h"

// Last-arrival detection without waiting: each of n_items work items calls
// gg_arrive_last once (one thread); exactly one call, the last to arrive,
// returns true.  Two levels of device-scope atomic counters (32 groups by
// item mod 32, then the groups) keep the arrivals on one address to about
// n_items / 32: a single counter serialises every arrival (about 11 ns each
// here).  Each counter sits in its own 128-byte line and is back at zero when
// the last arrival returns.  Callers make their data visible first: stores
// through the device-coherent level (device-scope atomic stores) completed
// before the call.
#define GG_ARRIVE_SUBS 32
#define GG_ARRIVE_STRIDE 32  // uint32 per counter (128 bytes)
#define GG_ARRIVE_WORDS ((GG_ARRIVE_SUBS + 1) * GG_ARRIVE_STRIDE)
// Per arrival sub-group: the state sums in fixed point, units of 2^-24
// (words 0 .. GG_MAX_D-1) and, in the last word, a flag for sums that did
// not fit.
#define GG_ACC_W 8
#define GG_ACC_SCALE 0x1p24
static_assert(GG_MAX_D + 1 <= GG_ACC_W, "fixed-point sum record too small");

__device__ __forceinline__ bool gg_arrive_last(uint32_t* c, int item, int n_items) {
  const int j = item % GG_ARRIVE_SUBS;
  const uint32_t group = (uint32_t)((n_items - 1 - j) / GG_ARRIVE_SUBS + 1);
  uint32_t* sub = c + GG_ARRIVE_STRIDE * (1 + j);
  if (__hip_atomic_fetch_add(sub, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u != group)
    return false;
  __hip_atomic_store(sub, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t groups = (uint32_t)(n_items < GG_ARRIVE_SUBS ? n_items : GG_ARRIVE_SUBS);
  if (__hip_atomic_fetch_add(c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u != groups)
    return false;
  __hip_atomic_store(c, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

struct gg_arena {
  float* state;           // [2][ncomp][P]
  float* lw;              // [P]
  int32_t* E;             // [P] tile-local prefix max of the output ends
  float* mw;              // [P] Megopolis weights exp32(lw - M) (same memory as E)
  float* aos;             // [P][8] Megopolis: the state of every particle, packed per particle
  int32_t* anc;           // [2][P]
  double* part;           // [n_tiles_total][GG_NPART]
  double* p4;             // [n_tiles_total][4]: m_b, S_b, S2_b, 0 (the combine's inputs, packed)
  double* g;              // [n_tiles_total]
  double* Pb;             // [n_tiles_total]
  double* csum;           // [GG_MAX_D][n_chunks]: state sums of each gather chunk
  uint32_t* fin;          // [n_filters][GG_ARRIVE_WORDS]: gather arrival counters
  int64_t* acc;           // [n_filters][GG_ARRIVE_SUBS][GG_ACC_W]: fixed-point gather sums
  double* scratch_f;      // [n_filters][16] misc per-filter results
  const gg_filter* filt;  // [n_filters]
  gg_dyn* dyn;            // [n_filters]
  const int32_t* tile_filter;   // [n_tiles_total]
  const int32_t* chunk_filter;  // [n_chunks_total]
  // epoch slot s of the satellite table:
  //   satd + s*5*sat_rows: xyz[sat_rows][3] | pr[sat_rows] | w[sat_rows]   (binary64)
  //   sati + s*(sat_rows+n_filters): sys[sat_rows] | n_sat[n_filters]      (int32)
  const double* satd;
  const int32_t* sati;
  const double* vel_guide;  // [n_slots][n_filters][3] (ECEF)
  // learned likelihood: packed network and per-row satellite inputs
  //   satf + s*2*sat_rows: cn0[sat_rows] | los_flag[sat_rows]   (binary32)
  const gg_mlp_w* mlp;
  const float* satf;
  float2* mlp_table;  // [sat_rows][GG_MLP_CODES] network outputs per bf16 code
  int64_t P;
  int32_t ncomp;
  int32_t n_isb;
  int32_t has_vel;
  int32_t sat_rows;
  int32_t n_filters;
  int32_t n_tiles;   // tiles of all filters
  int32_t n_chunks;  // resampling work items of all filters
  // 1 in a Megopolis step: its likelihood launch writes the predicted state to
  // the packed records (aos), which the chains read; after a step that
  // resampled it writes them only (gg_dyn.aos_only), and the Megopolis weight
  // launch copies them into the current state buffer if the step does not
  // resample
  int32_t aos_step;
};

__device__ __forceinline__ float* gg_comp(const gg_arena& a, int buf, int comp) {
  return a.state + ((int64_t)buf * a.ncomp + comp) * a.P;
}

// Satellite geometry staged in LDS as binary32 values relative to the frame
// anchor, grouped by inter-system-bias component: the reference system first,
// then ISB 0, 1, 2, each group in the epoch row's order.  Record k is GG_REC
// (three) float4:
//   rec[3k]     = (-2 ux, -2 uy, -2 uz, 1/rho)   u: unit vector anchor -> sat,
//                                                rho = |sat - anchor|
//   rec[3k + 1] = (c, wl, row index bits, isb bits)
//     c  = pr - rho - cb_anchor - isb_anchor of the satellite's system
//     wl = the likelihood's weight factor: -w / (2 sigma^2) (Gaussian),
//          w / (nu sigma^2) (Student-t), w (Huber, learned)
//     row index: position of the satellite in its epoch row; isb: ISB index,
//     -1 for the reference system.
//   rec[3k + 2] = (s c, s/2, -s/(8 rho), s/(16 rho^2)) with s = sqrt(|wl|):
//     the residual scaled by s, e' = s e, for the Gaussian and Student-t
//     terms (-e'^2 and log(1 + e'^2)), with s h(t) expanded in q (t = q/rho).
// gend[g]: end of group g (g = 0 the reference system, g = 1 + ISB index).
struct gg_sat_view {
  const float4* rec;
  float* cs;     // learned likelihood: first-layer bias per record [cap][64]
  int gend[GG_MAX_ISB + 1];
  float irmax;   // largest 1/rho of the staged satellites
};

// ---------------------------------------------------------------------------
// Host-style helpers that run per block (binary64, identical in all blocks).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void gg_rot_local(const double* R, const double* v, double* out) {
  out[0] = R[0] * v[0] + R[1] * v[1] + R[2] * v[2];
  out[1] = R[3] * v[0] + R[4] * v[1] + R[5] * v[2];
  out[2] = R[6] * v[0] + R[7] * v[1] + R[8] * v[2];
}

__device__ __forceinline__ void gg_rot_ecef(const double* R, const double* v, double* out) {
  out[0] = R[0] * v[0] + R[3] * v[1] + R[6] * v[2];
  out[1] = R[1] * v[0] + R[4] * v[1] + R[7] * v[2];
  out[2] = R[2] * v[0] + R[5] * v[1] + R[8] * v[2];
}

// Velocity bookkeeping of one predict call, computed identically by every
// block of a filter.  v_ref is the part of the motion that moves the anchor;
// particles only integrate (v_i - v_ref) * dt.
struct gg_pred_plan {
  double v_new_u[3];   // new uniform velocity (mode 0 staying uniform)
  double v_ref[3];     // local-frame velocity carried by the anchor
  float vg[3];         // guide, binary32 local frame
  float vref_f[3];
  float alpha;
  int mode_in;
  int mode_out;
  int per_particle_noise;  // sample velocity noise per particle
  // uniform-covariance Cholesky factor of sigma_pos^2 I + dt^2 Sigma_v (kf)
  float L[6];  // l00 l10 l11 l20 l21 l22
  double vcov_new[6];
};

__device__ __forceinline__ void gg_chol3(const double* a /*a00 a01 a02 a11 a12 a22*/, double* L) {
  const double l00 = sqrt(fmax(a[0], 0.0));
  double l10 = 0.0, l20 = 0.0;
  if (l00 > 1e-12) {
    l10 = a[1] / l00;
    l20 = a[2] / l00;
  }
  const double l11 = sqrt(fmax(a[3] - l10 * l10, 0.0));
  double l21 = 0.0;
  if (l11 > 1e-12) l21 = (a[4] - l20 * l10) / l11;
  const double l22 = sqrt(fmax(a[5] - l20 * l20 - l21 * l21, 0.0));
  L[0] = l00;
  L[1] = l10;
  L[2] = l11;
  L[3] = l20;
  L[4] = l21;
  L[5] = l22;
}

__device__ __forceinline__ gg_pred_plan gg_make_plan(const gg_dyn* dy, const gg_step_args& s,
                                                     const double* vg_ecef) {
  gg_pred_plan pl;
  double vg[3];
  gg_rot_local(dy->R, vg_ecef, vg);
  const double alpha = fmin(1.0, fmax(0.0, (double)s.alpha));
  pl.alpha = (float)alpha;
  pl.mode_in = dy->vel_mode;
  pl.per_particle_noise = (!s.velocity_kf && s.sigma_vel > 0.0f) ? 1 : 0;
  for (int k = 0; k < 3; ++k) {
    pl.vg[k] = (float)vg[k];
    pl.v_new_u[k] = (1.0 - alpha) * dy->vel_u[k] + alpha * vg[k];
  }
  if (pl.mode_in == 0 && !pl.per_particle_noise) {
    pl.mode_out = 0;
    for (int k = 0; k < 3; ++k) {
      pl.v_ref[k] = pl.v_new_u[k];
      pl.vref_f[k] = 0.0f;
    }
  } else if (pl.mode_in == 0) {
    pl.mode_out = 1;
    for (int k = 0; k < 3; ++k) {
      pl.vref_f[k] = (float)pl.v_new_u[k];
      pl.v_ref[k] = (double)pl.vref_f[k];
    }
  } else {
    pl.mode_out = pl.mode_in;
    for (int k = 0; k < 3; ++k) {
      pl.vref_f[k] = pl.vg[k];
      pl.v_ref[k] = (double)pl.vg[k];
    }
  }
  // uniform covariance path
  const double dt = s.dt;
  const double qdt = fmax((double)s.qv, 0.0) * fmax(dt, 0.0);
  for (int k = 0; k < 6; ++k) pl.vcov_new[k] = dy->vcov_u[k];
  for (int k = 0; k < 6; ++k) pl.L[k] = 0.0f;
  if (s.velocity_kf) {
    const double sp2 = (double)s.sigma_pos * (double)s.sigma_pos;
    const double dt2 = dt * dt;
    const double* c = dy->vcov_u;
    double a6[6] = {sp2 + dt2 * c[0], dt2 * c[1], dt2 * c[2], sp2 + dt2 * c[3], dt2 * c[4],
                    sp2 + dt2 * c[5]};
    double L[6];
    gg_chol3(a6, L);
    for (int k = 0; k < 6; ++k) pl.L[k] = (float)L[k];
    pl.vcov_new[0] += qdt;
    pl.vcov_new[3] += qdt;
    pl.vcov_new[5] += qdt;
  }
  return pl;
}

// Predict noise.  The thread that owns particles 4g .. 4g+3 of a filter draws
// them from one stream per predict call: xoshiro128+ seeded by the Philox
// block (filter seed, g, step, GG_PURPOSE_PREDICT, call).  Call 0 gives the
// 4 D normals z[q D + c] of component c (x, y, z in the local ENU frame, clock
// bias, then the inter-system biases) of particle 4g + q, as consecutive
// Box-Muller pairs; call 1 the 12 velocity-noise normals zv[3 q + k].  The
// stream hands out one normal at a time (a pair may straddle two particles),
// so no thread holds more than one pending normal.
struct gg_normal_stream {
  gg_xo x;
  float carry;
  bool has;
  __device__ __forceinline__ gg_normal_stream(uint64_t seed, uint32_t g, uint32_t step,
                                              uint32_t call)
      : x(gg_xo_seed(seed, g, step, GG_PURPOSE_PREDICT, call)), carry(0.0f), has(false) {}
  __device__ __forceinline__ float next() {
    if (has) {
      has = false;
      return carry;
    }
    float z0;
    gg_xo_normal2(x, z0, carry);
    has = true;
    return z0;
  }
};

// Noise of the fast predict path (uniform velocity carried by the anchor, no
// velocity filter): zn[q][c] = standard normal of component c of particle
// 4g + q.  It depends on nothing the step computes, so the weight kernel
// draws it before it waits for the step frame.
template <int NISB>
__device__ __forceinline__ void gg_noise4_fast(uint64_t seed, uint32_t g, uint32_t step,
                                               float (&zn)[GG_ITEMS][4 + NISB]) {
  gg_normal_stream zs(seed, g, step, 0u);
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q)
#pragma unroll
    for (int c = 0; c < 4 + NISB; ++c) zn[q][c] = zs.next();
}

// p += sigma_c z (one fused multiply-add per component): sigma_pos for x, y,
// z, sigma_cb for the clock bias, sigma_isb for the inter-system biases.
template <int D>
__device__ __forceinline__ void gg_apply_noise4_fast(const float (&zn)[GG_ITEMS][D],
                                                     const gg_step_args& s,
                                                     float (&p)[GG_ITEMS][D]) {
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q)
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const float sg = (c < 3) ? s.sigma_pos : (c == 3 ? s.sigma_cb : s.sigma_isb);
      p[q][c] = __builtin_fmaf(zn[q][c], sg, p[q][c]);
    }
}

// Predict one particle in place (binary32 offsets) with its normals z
// [4 + NISB] (position, clock bias, ISBs) and zv [3] (velocity noise).
template <int NISB>
__device__ __forceinline__ void gg_predict_one(const gg_arena& a, int cur, int64_t slot,
                                               const gg_step_args& s, const gg_pred_plan& pl,
                                               const float (&z0)[4 + NISB], const float (&zv)[3],
                                               float (&p)[4 + NISB]) {
  const float dt = (float)s.dt;
  float dv[3] = {0.0f, 0.0f, 0.0f};
  if (pl.mode_out >= 1) {
    float v[3];
    if (pl.mode_in == 0) {
      for (int k = 0; k < 3; ++k) v[k] = pl.vref_f[k];
    } else {
      for (int k = 0; k < 3; ++k) {
        const float vi = gg_comp(a, cur, 4 + NISB + k)[slot];
        v[k] = (1.0f - pl.alpha) * vi + pl.alpha * pl.vg[k];
      }
    }
    if (pl.per_particle_noise)
      for (int k = 0; k < 3; ++k) v[k] += zv[k] * s.sigma_vel;
    for (int k = 0; k < 3; ++k) {
      gg_comp(a, cur, 4 + NISB + k)[slot] = v[k];
      dv[k] = (v[k] - pl.vref_f[k]) * dt;
    }
  }
  float n0, n1, n2;
  if (s.velocity_kf) {
    float L[6];
    if (pl.mode_out == 2) {
      float* cv[6];
      for (int k = 0; k < 6; ++k) cv[k] = gg_comp(a, cur, 7 + NISB + k);
      const double sp2 = (double)s.sigma_pos * (double)s.sigma_pos;
      const double dt2 = s.dt * s.dt;
      double a6[6] = {sp2 + dt2 * cv[0][slot], dt2 * cv[1][slot], dt2 * cv[2][slot],
                      sp2 + dt2 * cv[3][slot], dt2 * cv[4][slot], sp2 + dt2 * cv[5][slot]};
      double Ld[6];
      gg_chol3(a6, Ld);
      for (int k = 0; k < 6; ++k) L[k] = (float)Ld[k];
      const float qdt = fmaxf(s.qv, 0.0f) * fmaxf(dt, 0.0f);
      cv[0][slot] += qdt;
      cv[3][slot] += qdt;
      cv[5][slot] += qdt;
    } else {
      for (int k = 0; k < 6; ++k) L[k] = pl.L[k];
    }
    n0 = L[0] * z0[0];
    n1 = L[1] * z0[0] + L[2] * z0[1];
    n2 = L[3] * z0[0] + L[4] * z0[1] + L[5] * z0[2];
  } else {
    n0 = z0[0] * s.sigma_pos;
    n1 = z0[1] * s.sigma_pos;
    n2 = z0[2] * s.sigma_pos;
  }
  p[0] += dv[0] + n0;
  p[1] += dv[1] + n1;
  p[2] += dv[2] + n2;
  p[3] += z0[3] * s.sigma_cb;
#pragma unroll
  for (int k = 0; k < NISB; ++k) p[4 + k] += z0[4 + k] * s.sigma_isb;
}

// Predict the thread's four particles (i0 = 4g).  With a uniform velocity
// carried by the anchor (plan mode 0) and no velocity filter every particle
// only adds its noise, as straight-line code; every other mode goes particle
// by particle.  Both branches take the same normals for the same particle.
template <int NISB>
__device__ __forceinline__ void gg_predict4(const gg_arena& a, int cur, int64_t base, int n,
                                            uint64_t seed, int i0, uint32_t step,
                                            const gg_step_args& s, const gg_pred_plan& pl,
                                            float (&p)[GG_ITEMS][4 + NISB]) {
  constexpr int D = 4 + NISB;
  const uint32_t g = (uint32_t)i0 >> 2;
  if (pl.mode_out == 0 && !s.velocity_kf) {
    float zn[GG_ITEMS][D];
    gg_noise4_fast<NISB>(seed, g, step, zn);
    gg_apply_noise4_fast<D>(zn, s, p);
    return;
  }
  gg_normal_stream zs(seed, g, step, 0u);
  gg_normal_stream vs(seed, g, step, 1u);
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
    float z[D], zv[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int c = 0; c < D; ++c) z[c] = zs.next();
    if (pl.per_particle_noise)
#pragma unroll
      for (int k = 0; k < 3; ++k) zv[k] = vs.next();
    if (i0 + q < n) gg_predict_one<NISB>(a, cur, base + q, s, pl, z, zv, p[q]);
  }
}

// Satellite row r relative to the frame anchor in the filter's ENU frame
// (binary64): d, |d| and 1/|d|.
__device__ __forceinline__ void gg_sat_local(const double* sd, const gg_dyn* dy,
                                             const double* anchor, int64_t r, double (&d)[3],
                                             double& rho, double& inv_rho) {
  double dx[3] = {sd[r * 3 + 0] - anchor[0], sd[r * 3 + 1] - anchor[1], sd[r * 3 + 2] - anchor[2]};
  gg_rot_local(dy->R, dx, d);
  rho = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  inv_rho = 1.0 / rho;
}

__device__ __forceinline__ int gg_sat_group(const gg_arena& a, int sys) {
  return (sys >= 1 && sys <= a.n_isb) ? sys : 0;
}

// Record of epoch-row satellite k (row r = sat_off + k) at position dest.
__device__ __forceinline__ float gg_stage_one(const gg_arena& a, const double* sd, const gg_dyn* dy,
                                              const double* anchor, const double* cb_anchor,
                                              const double* isb_anchor, int64_t r, int k, int gk,
                                              double wfac, float4* rec, int dest) {
  double d[3], rho, inv_rho;
  gg_sat_local(sd, dy, anchor, r, d, rho, inv_rho);
  const int ib = gk - 1;
  double c = sd[3 * a.sat_rows + r] - rho - *cb_anchor;
  if (ib >= 0) c -= isb_anchor[ib];
  const float ir = (float)inv_rho;
  const double wl = sd[4 * a.sat_rows + r] * wfac;
  rec[GG_REC * dest] = make_float4(-2.0f * (float)(d[0] * inv_rho),
                                   -2.0f * (float)(d[1] * inv_rho),
                                   -2.0f * (float)(d[2] * inv_rho), ir);
  rec[GG_REC * dest + 1] = make_float4((float)c, (float)wl, __int_as_float(k), __int_as_float(ib));
  // the scaled form: s = sqrt(|wl|), s c and the series of s h(q) in q
  const double sw = sqrt(fabs(wl));
  rec[GG_REC * dest + 2] = make_float4((float)(sw * c), (float)(0.5 * sw),
                                       (float)(-0.125 * sw * inv_rho),
                                       (float)(0.0625 * sw * inv_rho * inv_rho));
  return ir;
}

// Stage the filter's satellites for this epoch slot in LDS (records grouped by
// ISB component, see gg_sat_view).  anchor[3], cb_anchor, isb_anchor[3]: the
// step frame's binary64 origins (usually in LDS).  LDS: 36 bytes per
// satellite of the filter's capacity (records + group scratch), then the
// learned likelihood's first-layer biases.  Up to 64 satellites are placed by
// the first wavefront alone (ballots give each satellite its position and the
// group ends); larger rows go through LDS group indices.
__device__ __forceinline__ int gg_stage_sats(const gg_arena& a, int f, const gg_filter& F,
                                             const gg_dyn* dy, const double* anchor,
                                             const double* cb_anchor, const double* isb_anchor,
                                             int slot, const gg_step_args& s, gg_sat_view& sv,
                                             char* smem) {
  __shared__ int sgend[GG_MAX_ISB + 1];
  __shared__ float sirmax;
  const double* sd = a.satd + (int64_t)slot * 5 * a.sat_rows;
  const int32_t* si = a.sati + (int64_t)slot * (a.sat_rows + a.n_filters);
  const int ns = si[a.sat_rows + f];
  const int cap = F.sat_cap;
  const int t = threadIdx.x;
  float4* rec = reinterpret_cast<float4*>(smem);
  int* grp = reinterpret_cast<int*>(rec + GG_REC * cap);
  // weight factor of the likelihood (binary64, rounded once)
  const double sig2 = (double)s.sigma_pr * (double)s.sigma_pr;
  const double wfac = (s.lik == GG_LIK_GAUSS || s.lik == GG_LIK_GMM) ? -0.5 / sig2
                      : (s.lik == GG_LIK_STUDENT) ? 1.0 / ((double)s.nu * sig2) : 1.0;
  if (ns <= 64) {
    if (t < 64) {
      int gk = GG_MAX_ISB + 1;  // lanes without a satellite: after every group
      if (t < ns) gk = gg_sat_group(a, si[F.sat_off + t]);
      int base = 0, below = 0, run = 0;
#pragma unroll
      for (int g = 0; g <= GG_MAX_ISB; ++g) {
        const uint64_t m = __builtin_amdgcn_ballot_w64(gk == g);
        const int c = __builtin_popcountll(m);
        if (g == gk) {
          base = run;
          below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        }
        run += c;
        if (t == 0) sgend[g] = run;
      }
      float ir = 0.0f;
      if (t < ns)
        ir = gg_stage_one(a, sd, dy, anchor, cb_anchor, isb_anchor, F.sat_off + t, t, gk, wfac, rec,
                          base + below);
      GG_BUTTERFLY32(f32, ir, gg_max_f32);
      if (t == 0)
        sirmax = fmaxf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(ir), 0)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ir), 32)));
    }
  } else {
    for (int k = t; k < ns; k += blockDim.x) grp[k] = gg_sat_group(a, si[F.sat_off + k]);
    __syncthreads();
    if (t == 0) {
      int cnt[GG_MAX_ISB + 1] = {0, 0, 0, 0};
      for (int k = 0; k < ns; ++k) cnt[grp[k]] += 1;
      int run = 0;
      for (int g = 0; g <= GG_MAX_ISB; ++g) {
        run += cnt[g];
        sgend[g] = run;
      }
      sirmax = 0.0f;
    }
    float irm = 0.0f;
    for (int k = t; k < ns; k += blockDim.x) {
      const int gk = grp[k];
      int dest = 0;
      for (int j = 0; j < ns; ++j) {
        const int gj = grp[j];
        dest += (gj < gk || (gj == gk && j < k)) ? 1 : 0;
      }
      irm = fmaxf(irm, gg_stage_one(a, sd, dy, anchor, cb_anchor, isb_anchor, F.sat_off + k, k, gk,
                                    wfac, rec, dest));
    }
    __syncthreads();
    if (irm > 0.0f) atomicMax(reinterpret_cast<int*>(&sirmax), __float_as_int(irm));
  }
  __syncthreads();
  sv.rec = rec;
  sv.cs = reinterpret_cast<float*>(grp + cap);
#pragma unroll
  for (int g = 0; g <= GG_MAX_ISB; ++g) sv.gend[g] = __builtin_amdgcn_readfirstlane(sgend[g]);
  sv.irmax = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(sirmax)));
  return ns;
}

// Pseudorange residual pr - (|sat - x| + bias) of a binary32 offset particle x
// (pp = |x|^2; bias = clock bias + the ISB of the satellite's system).
// |sat - x| - rho = q h with q = pp/rho - 2 u.x, t = q/rho and
// h = 1 / (1 + sqrt(1 + t)); the 2e7 m range is never formed in binary32.
// FORM selects how h is evaluated (gg_range_form picks it per thread):
//   GG_FORM_EXACT  the closed form;
//   GG_FORM_QUAD   h = 1/2 - t/8 + t^2/16, truncation error (5/128) |t|^3,
//                  below 1e-10 relative for |t| <= 2^-10;
//   GG_FORM_LIN    h = 1/2 - t/8, truncation error t^2/16 (relative t^2/8),
//                  below 1e-10 relative for |t| <= 2^-15.3.
#define GG_FORM_EXACT 0
#define GG_FORM_QUAD 1
#define GG_FORM_LIN 2
template <int FORM>
__device__ __forceinline__ float gg_residual(const float4& g, const float4& m, const float* p,
                                             float pp, float bias) {
  const float ud2 = fmaf(g.z, p[2], fmaf(g.y, p[1], g.x * p[0]));  // -2 u.x
  const float q = fmaf(pp, g.w, ud2);
  const float t = q * g.w;
  float h;
  if (FORM == GG_FORM_LIN)
    h = fmaf(t, -0.125f, 0.5f);
  else if (FORM == GG_FORM_QUAD)
    h = fmaf(fmaf(t, 0.0625f, -0.125f), t, 0.5f);
  else
    h = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_sqrtf(1.0f + t));
  return fmaf(-q, h, m.x) - bias;
}

// The residual scaled by s = sqrt(|wl|) from the records rec[3k] (g) and
// rec[3k + 2] (r): e' = s c - q s h - s bias, the 2e7 m range never formed.
// Series forms: s h = s/2 - (s/(8 rho)) q [+ (s/(16 rho^2)) q^2] (gg_residual's
// series with t = q/rho); otherwise s h = 2 (s/2) / (1 + sqrt(1 + t)).
template <int FORM>
__device__ __forceinline__ float gg_residual_scaled(const float4& g, const float4& r, const float* p,
                                                    float pp, float bias) {
  const float ud2 = fmaf(g.z, p[2], fmaf(g.y, p[1], g.x * p[0]));  // -2 u.x
  const float q = fmaf(pp, g.w, ud2);
  float sh;
  if (FORM == GG_FORM_LIN)
    sh = fmaf(q, r.z, r.y);
  else if (FORM == GG_FORM_QUAD)
    sh = fmaf(fmaf(q, r.w, r.z), q, r.y);
  else
    sh = (r.y + r.y) * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_sqrtf(fmaf(q, g.w, 1.0f)));
  return fmaf(-(r.y + r.y), bias, fmaf(-q, sh, r.x));
}

// Range form of a thread whose particles have |x|^2 <= ppmax, against staged
// satellites with 1/rho <= irmax.  |t| <= (pp/rho + 2 |x|)/rho, so
// pp (1/rho)^2 <= 2^-25 gives |t| <= 2^-11.4 (the quadratic series) and
// pp (1/rho)^2 <= 2^-33 gives |t| <= 2^-15.4 (the linear one): |x| within
// about 3.5 km and 220 m of the anchor at GNSS ranges.  The margins cover the
// binary32 rounding of q and t.
__device__ __forceinline__ int gg_range_form(float ppmax, float irmax) {
  const float v = ppmax * irmax * irmax;
  return (v <= 1.16415321826934814e-10f) ? GG_FORM_LIN             // 2^-33
         : (v <= 2.98023223876953125e-08f) ? GG_FORM_QUAD          // 2^-25
                                           : GG_FORM_EXACT;
}

// |x|^2 of the thread's particles and the range form that holds for all of
// them (gg_range_form).
template <int D>
__device__ __forceinline__ int gg_pp4(const gg_sat_view& sv, const float (&p)[GG_ITEMS][D],
                                      float (&pp)[GG_ITEMS]) {
  float ppmax = 0.0f;
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
    pp[q] = fmaf(p[q][2], p[q][2], fmaf(p[q][1], p[q][1], p[q][0] * p[q][0]));
    ppmax = fmaxf(ppmax, pp[q]);
  }
  return gg_range_form(ppmax, sv.irmax);
}

// Record k's bias term for group G (-1: reference system) of particle p.
template <int NISB, int G>
__device__ __forceinline__ float gg_group_bias(const float* p) {
  if constexpr (G < 0 || G >= NISB)
    return p[3];
  else
    return p[3] + p[4 + G];
}

// Calls fn(group tag, k0, k1) for the record ranges of the reference system
// and of ISB components 0 .. NISB-1, in that order (the group index is a
// compile-time constant inside fn).
template <int G>
struct gg_grp {
  static constexpr int value = G;
};

template <int NISB, typename Fn>
__device__ __forceinline__ void gg_for_groups(const gg_sat_view& sv, Fn&& fn) {
  fn(gg_grp<-1>{}, 0, sv.gend[0]);
  if constexpr (NISB > 0) fn(gg_grp<0>{}, sv.gend[0], sv.gend[1]);
  if constexpr (NISB > 1) fn(gg_grp<1>{}, sv.gend[1], sv.gend[2]);
  if constexpr (NISB > 2) fn(gg_grp<2>{}, sv.gend[2], sv.gend[3]);
}

// natural log on the hardware log2 unit (relative error ~1e-7 for the
// products of (1 + r^2/(nu s^2)) factors used here)
__device__ __forceinline__ float gg_lnf(float x) {
  return __builtin_amdgcn_logf(x) * 0.6931471805599453f;
}

// Likelihood of one particle with the NLOS gate (residuals beyond gate_thr
// are dropped unless fewer than min(ns, 4) would remain).  Satellites in
// record order; the exact range form.
template <int NISB>
__device__ __forceinline__ float gg_loglik(const gg_sat_view& sv, int ns, const float* p,
                                           const gg_step_args& s) {
  const float pp = fmaf(p[2], p[2], fmaf(p[1], p[1], p[0] * p[0]));
  const float sig = s.sigma_pr;
  bool gate = s.gate_thr > 0.0f;
  if (gate) {
    int kept = 0;
    gg_for_groups<NISB>(sv, [&](auto grp, int k0, int k1) {
      const float bias = gg_group_bias<NISB, decltype(grp)::value>(p);
      for (int k = k0; k < k1; ++k) {
        const float e =
            gg_residual<GG_FORM_EXACT>(sv.rec[GG_REC * k], sv.rec[GG_REC * k + 1], p, pp, bias);
        kept += (fabsf(e) <= s.gate_thr) ? 1 : 0;
      }
    });
    const int min_kept = ns < 4 ? ns : 4;
    if (kept < min_kept) gate = false;
  }
  float lw = 0.0f;
  if (s.lik == GG_LIK_STUDENT) {
    // sum of logs = log of the product; the product is taken over groups of
    // three factors so that it stays far from binary32 overflow.
    float acc = 0.0f, prod = 1.0f;
    int cnt = 0;
    gg_for_groups<NISB>(sv, [&](auto grp, int k0, int k1) {
      const float bias = gg_group_bias<NISB, decltype(grp)::value>(p);
      for (int k = k0; k < k1; ++k) {
        const float4 m = sv.rec[GG_REC * k + 1];
        const float e = gg_residual<GG_FORM_EXACT>(sv.rec[GG_REC * k], m, p, pp, bias);
        if (gate && fabsf(e) > s.gate_thr) continue;
        prod *= fmaf(m.y * e, e, 1.0f);
        if (++cnt == 3) {
          acc += gg_lnf(prod);
          prod = 1.0f;
          cnt = 0;
        }
      }
    });
    acc += gg_lnf(prod);
    lw = -0.5f * (s.nu + 1.0f) * acc;
  } else {
    const float hk = s.huber_k;
    const float isig = 1.0f / sig;
    gg_for_groups<NISB>(sv, [&](auto grp, int k0, int k1) {
      const float bias = gg_group_bias<NISB, decltype(grp)::value>(p);
      for (int k = k0; k < k1; ++k) {
        const float4 m = sv.rec[GG_REC * k + 1];
        const float e = gg_residual<GG_FORM_EXACT>(sv.rec[GG_REC * k], m, p, pp, bias);
        if (gate && fabsf(e) > s.gate_thr) continue;
        if (s.lik == GG_LIK_HUBER) {
          const float z = fabsf(e) * isig;
          const float rho = (z < hk) ? 0.5f * z * z : hk * z - 0.5f * hk * hk;
          lw -= m.y * rho;
        } else {
          lw = fmaf(m.y * e, e, lw);
        }
      }
    });
  }
  return lw;
}

// Likelihood of the thread's four particles with the satellites in the outer
// loop, so that the four residual chains interleave; a group's bias is formed
// once per particle.  With the NLOS gate on, the kept set differs per
// particle and gg_loglik is used.
template <int NISB, int FORM>
__device__ __forceinline__ void gg_loglik4_form(const gg_sat_view& sv,
                                                const float (&p)[GG_ITEMS][4 + NISB],
                                                const float (&pp)[GG_ITEMS],
                                                const gg_step_args& s, float (&out)[GG_ITEMS]) {
  if (s.lik == GG_LIK_STUDENT) {
    float acc[GG_ITEMS], prod[GG_ITEMS];
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) {
      acc[q] = 0.0f;
      prod[q] = 1.0f;
    }
    int cnt = 0;
    gg_for_groups<NISB>(sv, [&](auto grp, int k0, int k1) {
      float bias[GG_ITEMS];
#pragma unroll
      for (int q = 0; q < GG_ITEMS; ++q) bias[q] = gg_group_bias<NISB, decltype(grp)::value>(p[q]);
      for (int k = k0; k < k1; ++k) {
        const float4 g = sv.rec[GG_REC * k];
        const float4 r = sv.rec[GG_REC * k + 2];
#pragma unroll
        for (int q = 0; q < GG_ITEMS; ++q) {
          const float e = gg_residual_scaled<FORM>(g, r, p[q], pp[q], bias[q]);
          prod[q] *= fmaf(e, e, 1.0f);
        }
        if (++cnt == 3) {
#pragma unroll
          for (int q = 0; q < GG_ITEMS; ++q) {
            acc[q] += __builtin_amdgcn_logf(prod[q]);  // log2: ln 2 applied once below
            prod[q] = 1.0f;
          }
          cnt = 0;
        }
      }
    });
    const float hn = -0.5f * (s.nu + 1.0f) * 0.6931471805599453f;
    if (cnt)  // uniform: every thread sees the same satellite count
#pragma unroll
      for (int q = 0; q < GG_ITEMS; ++q) acc[q] += __builtin_amdgcn_logf(prod[q]);
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) out[q] = hn * acc[q];
  } else if (s.lik == GG_LIK_GMM) {
    // per satellite log(w N(e; 0, sigma_los) + (1 - w) N(e; mu, sigma_nlos)) up
    // to the common -log sqrt(2 pi), formed as the reference does (max plus
    // log of the exponential sum); m.y = -w_s / (2 sigma_los^2), the NLOS
    // factor is m.y sigma_los^2 / sigma_nlos^2.  The log constants come from
    // the host (binary64 logs here would run on every likelihood's path).
    const float c_los = s.gmm_c_los;
    const float c_nlos = s.gmm_c_nlos;
    const float ratio = s.gmm_ratio;
    const float mu = s.gmm_mu_nlos;
    float lw[GG_ITEMS] = {0.0f, 0.0f, 0.0f, 0.0f};
    gg_for_groups<NISB>(sv, [&](auto grp, int k0, int k1) {
      float bias[GG_ITEMS];
#pragma unroll
      for (int q = 0; q < GG_ITEMS; ++q) bias[q] = gg_group_bias<NISB, decltype(grp)::value>(p[q]);
      for (int k = k0; k < k1; ++k) {
        const float4 g = sv.rec[GG_REC * k];
        const float4 m = sv.rec[GG_REC * k + 1];
        const float kn = m.y * ratio;
#pragma unroll
        for (int q = 0; q < GG_ITEMS; ++q) {
          const float e = gg_residual<FORM>(g, m, p[q], pp[q], bias[q]);
          const float en = e - mu;
          const float a = fmaf(m.y * e, e, c_los);
          const float b = fmaf(kn * en, en, c_nlos);
          const float mx = fmaxf(a, b);
          lw[q] += mx + gg_lnf(__expf(a - mx) + __expf(b - mx));
        }
      }
    });
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) out[q] = lw[q];
  } else if (s.lik == GG_LIK_HUBER) {
    const float hk = s.huber_k;
    const float isig = 1.0f / s.sigma_pr;
    float lw[GG_ITEMS] = {0.0f, 0.0f, 0.0f, 0.0f};
    gg_for_groups<NISB>(sv, [&](auto grp, int k0, int k1) {
      float bias[GG_ITEMS];
#pragma unroll
      for (int q = 0; q < GG_ITEMS; ++q) bias[q] = gg_group_bias<NISB, decltype(grp)::value>(p[q]);
      for (int k = k0; k < k1; ++k) {
        const float4 g = sv.rec[GG_REC * k];
        const float4 m = sv.rec[GG_REC * k + 1];
#pragma unroll
        for (int q = 0; q < GG_ITEMS; ++q) {
          const float e = gg_residual<FORM>(g, m, p[q], pp[q], bias[q]);
          const float z = fabsf(e) * isig;
          const float rho = (z < hk) ? 0.5f * z * z : hk * z - 0.5f * hk * hk;
          lw[q] -= m.y * rho;
        }
      }
    });
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) out[q] = lw[q];
  } else {
    float lw[GG_ITEMS] = {0.0f, 0.0f, 0.0f, 0.0f};
    gg_for_groups<NISB>(sv, [&](auto grp, int k0, int k1) {
      float bias[GG_ITEMS];
#pragma unroll
      for (int q = 0; q < GG_ITEMS; ++q) bias[q] = gg_group_bias<NISB, decltype(grp)::value>(p[q]);
      for (int k = k0; k < k1; ++k) {
        const float4 g = sv.rec[GG_REC * k];
        const float4 r = sv.rec[GG_REC * k + 2];
#pragma unroll
        for (int q = 0; q < GG_ITEMS; ++q) {
          const float e = gg_residual_scaled<FORM>(g, r, p[q], pp[q], bias[q]);
          lw[q] = fmaf(-e, e, lw[q]);
        }
      }
    });
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) out[q] = lw[q];
  }
}

template <int NISB>
__device__ __forceinline__ void gg_loglik4(const gg_sat_view& sv, int ns,
                                           const float (&p)[GG_ITEMS][4 + NISB],
                                           const gg_step_args& s, float (&out)[GG_ITEMS]) {
  if (s.gate_thr > 0.0f) {
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) out[q] = gg_loglik<NISB>(sv, ns, p[q], s);
    return;
  }
  float pp[GG_ITEMS];
  const int form = gg_pp4(sv, p, pp);
  if (form == GG_FORM_LIN)
    gg_loglik4_form<NISB, GG_FORM_LIN>(sv, p, pp, s, out);
  else if (form == GG_FORM_QUAD)
    gg_loglik4_form<NISB, GG_FORM_QUAD>(sv, p, pp, s, out);
  else
    gg_loglik4_form<NISB, GG_FORM_EXACT>(sv, p, pp, s, out);
}

// Learned likelihood, per-satellite part: features (elevation and azimuth
// from the staged unit vector, C/N0 and line-of-sight flag of the epoch slot,
// satellite count, ESS fraction of the last decision) and the first-layer
// bias c_s = b1 + W1[:, 1:] f of every record, into sv.cs.
__device__ __forceinline__ void gg_stage_learned(const gg_arena& a, const gg_filter& F,
                                                 const gg_dyn* dy, int slot, int ns,
                                                 const gg_mlp_w& w, gg_sat_view& sv) {
  const float* sf = a.satf + (int64_t)slot * 2 * a.sat_rows;
  const float essf = (F.n > 0) ? (float)(dy->ess / (double)F.n) : 1.0f;
  for (int i = threadIdx.x; i < ns * GG_MLP_H; i += blockDim.x) {
    const int k = i / GG_MLP_H;
    const int u = i % GG_MLP_H;
    const float4 g = sv.rec[GG_REC * k];
    const int64_t r = F.sat_off + __float_as_int(sv.rec[GG_REC * k + 1].z);
    float f[GG_MLP_NF - 1];
    gg_mlp_features(-0.5f * g.x, -0.5f * g.y, -0.5f * g.z, sf[r], sf[a.sat_rows + r], ns, essf,
                    f);
    sv.cs[k * GG_MLP_H + u] = gg_mlp_c(w, u, f);
  }
  __syncthreads();
}

// Learned likelihood through the per-satellite code tables (gg_k_mlp_table):
// the network output for a residual depends only on the bf16 code of its
// scaled value, so a lookup returns exactly what the per-pair network returns.
// The residuals and the per-particle sum follow gg_loglik_learned (record
// order, same range form), so both give identical log-weights.
template <int NISB, int FORM>
__device__ __forceinline__ void gg_loglik_table_form(const gg_sat_view& sv, int64_t row0,
                                                     const float2* __restrict__ tab,
                                                     const float (&p)[GG_ITEMS][4 + NISB],
                                                     const float (&pp)[GG_ITEMS],
                                                     float (&out)[GG_ITEMS]) {
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) out[q] = 0.0f;
  gg_for_groups<NISB>(sv, [&](auto grp, int k0, int k1) {
    float bias[GG_ITEMS];
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) bias[q] = gg_group_bias<NISB, decltype(grp)::value>(p[q]);
    for (int k = k0; k < k1; ++k) {
      const float4 g = sv.rec[GG_REC * k];
      const float4 m = sv.rec[GG_REC * k + 1];
      const float2* tk = tab + (row0 + __float_as_int(m.z)) * GG_MLP_CODES;
#pragma unroll
      for (int q = 0; q < GG_ITEMS; ++q) {
        const float e = gg_residual<FORM>(g, m, p[q], pp[q], bias[q]);
        const float2 z = tk[gg_mlp_code_index(gg_mlp_xe(e))];
        out[q] += gg_mlp_ll(e, z.x, z.y);
      }
    }
  });
}

template <int NISB>
__device__ __forceinline__ void gg_loglik_table(const gg_sat_view& sv, int64_t row0,
                                                const float2* __restrict__ tab,
                                                const float (&p)[GG_ITEMS][4 + NISB],
                                                float (&out)[GG_ITEMS]) {
  float pp[GG_ITEMS];
  const int form = gg_pp4(sv, p, pp);
  if (form == GG_FORM_LIN)
    gg_loglik_table_form<NISB, GG_FORM_LIN>(sv, row0, tab, p, pp, out);
  else if (form == GG_FORM_QUAD)
    gg_loglik_table_form<NISB, GG_FORM_QUAD>(sv, row0, tab, p, pp, out);
  else
    gg_loglik_table_form<NISB, GG_FORM_EXACT>(sv, row0, tab, p, pp, out);
}

// Learned likelihood of the workgroup's tile: for every satellite the 1024
// residuals go through LDS to the matrix-core layout (16 particles per pass
// and wavefront), the pair log-likelihoods are summed per particle in LDS.
// se and slw: GG_TILE floats of LDS each.  The loop bounds are uniform over
// the workgroup (the barriers inside); the range form is chosen per thread.
template <int NISB>
__device__ __forceinline__ void gg_loglik_learned(const gg_sat_view& sv,
                                                  const float (&p)[GG_ITEMS][4 + NISB],
                                                  const gg_mlp_w& w, float* se, float* slw,
                                                  float (&out)[GG_ITEMS]) {
  const int t = threadIdx.x;
  const int l = t & 63;
  const int base = (t >> 6) * 256;
  float pp[GG_ITEMS];
  const int form = gg_pp4(sv, p, pp);
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) slw[t * GG_ITEMS + q] = 0.0f;
  gg_mlp_regs wr;
  gg_mlp_load_regs(w, wr);
  gg_for_groups<NISB>(sv, [&](auto grp, int k0, int k1) {
    float bias[GG_ITEMS];
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) bias[q] = gg_group_bias<NISB, decltype(grp)::value>(p[q]);
    for (int k = k0; k < k1; ++k) {
      const float4 g = sv.rec[GG_REC * k];
      const float4 m = sv.rec[GG_REC * k + 1];
#pragma unroll
      for (int q = 0; q < GG_ITEMS; ++q)
        se[t * GG_ITEMS + q] =
            (form == GG_FORM_LIN)    ? gg_residual<GG_FORM_LIN>(g, m, p[q], pp[q], bias[q])
            : (form == GG_FORM_QUAD) ? gg_residual<GG_FORM_QUAD>(g, m, p[q], pp[q], bias[q])
                                     : gg_residual<GG_FORM_EXACT>(g, m, p[q], pp[q], bias[q]);
      __syncthreads();
      gg_bf16x8 a1[4];
      gg_mlp_a1(w.w_e, sv.cs + k * GG_MLP_H, a1);
      for (int cb = 0; cb < 16; ++cb) {
        const int i = base + 16 * cb + (l & 15);
        const float e = se[i];
        gg_mlp_act act;
        gg_mlp_fwd16(wr, a1, gg_mlp_xe(e), act);
        if (l < 16) slw[i] += gg_mlp_ll(e, act.z30, act.z31);
      }
      __syncthreads();
    }
  });
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) out[q] = slw[t * GG_ITEMS + q];
}

// ---------------------------------------------------------------------------
// Tile partials: m_b = max lw, S_b = sum e, S2_b = sum e^2, sum_d = sum e p_d
// with e = exp32(lw - m_b) (binary32, like lw itself).  S and S2 accumulate
// in binary64 in a fixed order: per thread left to right over its 4
// particles, then gg_block_scan256's total.  The weighted state sums (used
// only for the estimate) accumulate in binary32 inside each half-wavefront
// and in binary64 across the 8 half-wavefronts, also in a fixed order.
// ---------------------------------------------------------------------------
// Totals of NV doubles per thread over the block, each in the order of
// gg_block_scan256's total: inside a wavefront the pairwise tree that the
// Hillis-Steele scan leaves in lane 63 (lanes 2i and 2i+1, then pairs of
// pairs, ...), formed here by butterfly exchanges (addition is commutative,
// so both partners compute the same bits), then ((T0 + T1) + T2) + T3 over
// the wavefront totals T_w = (lanes 0..31) + (lanes 32..63).  The totals
// are formed in thread 0 only (the other threads get zeros); the call
// synchronises the block on entry to the combine and on return.
template <int NV>
__device__ __forceinline__ void gg_block_totals_n(const double (&v)[NV], double* lds, double* out) {
#pragma clang fp contract(off)
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double x = v[q];
    GG_BUTTERFLY32(f64, x, gg_add_f64);
    if ((lane & 31) == 0) lds[q * 8 + w * 2 + (lane >> 5)] = x;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) out[q] = 0.0;
  if (t == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const double* l = lds + q * 8;
      out[q] = (((l[0] + l[1]) + (l[2] + l[3])) + (l[4] + l[5])) + (l[6] + l[7]);
    }
  }
  __syncthreads();
}

// gg_block_scan256 of v (the same operations, so the same excl and total bit
// for bit) together with gg_block_totals_n of vv (the same operations, totals
// in thread 0, zeros elsewhere), sharing the scan's barriers.  ltot holds
// NV * 8 doubles of LDS.
template <int NV>
__device__ __forceinline__ void gg_block_scan256_totals(double v, const double (&vv)[NV],
                                                        double* lds, double* ltot, double& excl,
                                                        double& total, double (&tot)[NV]) {
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
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double y = vv[q];
    GG_BUTTERFLY32(f64, y, gg_add_f64);
    if ((lane & 31) == 0) ltot[q * 8 + w * 2 + (lane >> 5)] = y;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) tot[q] = 0.0;
  if (t == 0) {
    const double o1 = lds[0];
    const double o2 = o1 + lds[1];
    const double o3 = o2 + lds[2];
    lds[4] = o1;
    lds[5] = o2;
    lds[6] = o3;
    lds[7] = o3 + lds[3];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const double* l = ltot + q * 8;
      tot[q] = (((l[0] + l[1]) + (l[2] + l[3])) + (l[4] + l[5])) + (l[6] + l[7]);
    }
  }
  __syncthreads();
  if (w > 0) x = lds[3 + w] + x;
  lds[8 + t] = x;
  __syncthreads();
  excl = (t == 0) ? 0.0 : lds[8 + t - 1];
  total = lds[7];
  __syncthreads();
}

// Each value is reduced on its own (same operations whatever the count).
template <int D>
__device__ __forceinline__ void gg_block_totals(const double (&v)[D + 2], double* lds, double* out) {
  gg_block_totals_n<D + 2>(v, lds, out);
}

// Scale of a tile: its weights are e_i 2^k c relative to exp(lw_i) up to a
// common factor, with k = floor(m_b / ln 2) and c = exp64(m_b - k ln 2) in
// [1, 2) (Cody-Waite split of ln 2, as gg_exp64).  Tiles then combine through
// exact powers of two (2^(k_b - K)), so the combine needs no exponential.  An
// empty tile (m_b = -inf) gets k = -2^30, c = 0; a non-finite maximum k = 0,
// c = 1 (its NaN or infinite sums propagate as before).
__device__ __forceinline__ void gg_tile_scale(double mb, double& k, double& c) {
#pragma clang fp contract(off)
  if (!(mb > -__builtin_huge_val())) {
    k = -1073741824.0;
    c = 0.0;
  } else if (!(mb < __builtin_huge_val())) {
    k = 0.0;
    c = 1.0;
  } else {
    k = floor(mb * 0x1.71547652b82fep+0);
    c = gg_exp64((mb - k * 0x1.62e42fee00000p-1) - k * 0x1.a39ef35793c76p-33);
  }
}

// 2^(k_b - K) x v for a tile at most 1000 binary orders below the largest
// (0 below that): the result stays a normal number, so the scaling is exact.
__device__ __forceinline__ double gg_tile_shift(double v, double dk) {
  return (dk >= -1000.0) ? __builtin_ldexp(v, (int)dk) : 0.0;
}

template <int D>
__device__ __forceinline__ void gg_tile_partials(const float (&lw)[GG_ITEMS],
                                                 const bool (&valid)[GG_ITEMS],
                                                 const float (&p)[GG_ITEMS][D], double* lds,
                                                 double* part_out, double* p4_out) {
#pragma clang fp contract(off)
  float m = -__builtin_huge_valf();
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q)
    if (valid[q]) m = fmaxf(m, lw[q]);
  const float mb = gg_block_max256f(m, lds);
  float e[GG_ITEMS];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q)
    e[q] = (valid[q] && mb > -__builtin_huge_valf()) ? gg_exp32(lw[q] - mb) : 0.0f;
  // S and S2 in binary64 (they set the CDF tile offsets and the ESS)
  double v[2];
  v[0] = (double)e[0];
  v[1] = (double)e[0] * (double)e[0];
#pragma unroll
  for (int q = 1; q < GG_ITEMS; ++q) {
    v[0] = v[0] + (double)e[q];
    v[1] = __builtin_fma((double)e[q], (double)e[q], v[1]);  // the ESS only: no twin
  }
  // weighted state sums (the estimate): binary32 inside the wavefront,
  // binary64 across the 8 half-wavefront partials
  float ms[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    ms[d] = e[0] * p[0][d];
#pragma unroll
    for (int q = 1; q < GG_ITEMS; ++q) ms[d] = __builtin_fmaf(e[q], p[q][d], ms[d]);
  }
  const int t = threadIdx.x;
  const int lane = t & 63;
  float* lf = reinterpret_cast<float*>(lds + 16);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    float x = ms[d];
    GG_BUTTERFLY32(f32, x, gg_add_f32);
    if ((lane & 31) == 0) lf[d * 8 + (t >> 5)] = x;
  }
  double tot[2];
  gg_block_totals_n<2>(v, lds, tot);  // synchronises: lf is complete here
  if (t == 0) {
    double kb, cb;
    gg_tile_scale((double)mb, kb, cb);
    const double Sc = tot[0] * cb;
    const double S2c = tot[1] * (cb * cb);
    part_out[0] = (double)mb;
    part_out[1] = Sc;
    part_out[2] = S2c;
    part_out[GG_NPART - 2] = kb;
    part_out[GG_NPART - 1] = cb;
    *reinterpret_cast<double4*>(p4_out) = make_double4((double)mb, Sc, S2c, kb);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float* l = lf + d * 8;
      part_out[3 + d] = ((((double)l[0] + (double)l[1]) + ((double)l[2] + (double)l[3])) +
                         (((double)l[4] + (double)l[5]) + ((double)l[6] + (double)l[7]))) *
                        cb;
    }
  }
}

// systematic offset u0 (reference formula, int32 product included).
__device__ __forceinline__ double gg_u0(const gg_step_args& s, uint64_t seed, int n,
                                        int resample_count) {
  if (s.u0_mode == 2) return s.u0_value;
  if (s.u0_mode == 1) {
    const gg_u32x4 r = gg_draw(seed, 0u, (uint32_t)resample_count, GG_PURPOSE_U0, 0u);
    return gg_u01d(r.v[0], r.v[1]) - 1.1102230246251565e-16;  // [0, 1)
  }
  const int32_t den = (int32_t)(uint32_t)((uint64_t)1000000u * (uint64_t)(uint32_t)n);
  return (double)(seed % 1000000ull) / (double)den;
}

// #{ j in [0, n) : (u0 + j) / n <= c } with the targets rounded exactly as
// the reference computes them (pf_device.cu:1197); NaN counts nothing.  The
// estimate floor(c n - u0) is corrected by walking to the exact boundary.
__device__ __forceinline__ int gg_kcount(double c, double u0, int n) {
#pragma clang fp contract(off)
  const double dn = (double)n;
  if (c != c) return 0;
  const double x = c * dn - u0;
  const double est = floor(x);
  // x is within 2^-52 (n + |x|) of c n - u0, and a target's rounded quotient
  // fl(fl(u0 + j) / n) within 2^-52 of (u0 + j) / n, so when x is farther
  // than 1e-14 (n + |x|) from every integer the count is floor(x) + 1 exactly
  // (clamped to [0, n]); otherwise the exact comparisons below decide.
  const double fr = x - est;
  const double eps = 1e-14 * (dn + fabs(x));
  if (fr > eps && fr < 1.0 - eps) return est < 0.0 ? 0 : (est >= dn ? n : (int)est + 1);
  int j = (est < -1.0) ? -1 : (est > (double)(n - 1) ? n - 1 : (int)est);
  while (j + 1 < n && (u0 + (double)(j + 1)) / dn <= c) ++j;
  while (j >= 0 && (u0 + (double)j) / dn > c) --j;
  return j + 1;
}

// Copies of a filter record through the device-coherent level (device-scope
// atomic loads / stores, one 32-bit word per thread), for a record that one
// workgroup writes and another reads within the same launch.  A store copy
// completes (s_waitcnt) before the copying threads go on.
__device__ __forceinline__ void gg_dyn_load_coherent(gg_dyn* dst, const gg_dyn* src) {
  constexpr int NW = (int)(sizeof(gg_dyn) / 4);
  if (threadIdx.x < NW)
    reinterpret_cast<uint32_t*>(dst)[threadIdx.x] =
        __hip_atomic_load(reinterpret_cast<const uint32_t*>(src) + threadIdx.x, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void gg_dyn_store_coherent(gg_dyn* dst, const gg_dyn* src) {
  constexpr int NW = (int)(sizeof(gg_dyn) / 4);
  if (threadIdx.x < NW)
    __hip_atomic_store(reinterpret_cast<uint32_t*>(dst) + threadIdx.x,
                       reinterpret_cast<const uint32_t*>(src)[threadIdx.x], __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// Copy a filter record between global memory and LDS, one 32-bit word per
// thread (the record is a multiple of 4 bytes).
__device__ __forceinline__ void gg_dyn_copy(gg_dyn* dst, const gg_dyn* src) {
  constexpr int NW = (int)(sizeof(gg_dyn) / 4);
  static_assert(sizeof(gg_dyn) % 4 == 0, "gg_dyn must be a whole number of words");
  static_assert(NW <= GG_BLOCK, "gg_dyn larger than one word per thread");
  if (threadIdx.x < NW)
    reinterpret_cast<uint32_t*>(dst)[threadIdx.x] =
        reinterpret_cast<const uint32_t*>(src)[threadIdx.x];
}

// Frame of one step.  Predict moves the anchor by the uniform velocity and
// re-centres the offsets on the last posterior mean (shift = mean as binary32,
// added to the binary64 anchors), so the offsets stay at metre scale.  Every
// kernel of the step derives the same frame from the same record.
struct gg_frame {
  double anchor[3];
  double cb_anchor;
  double isb_anchor[GG_MAX_ISB];
  float shift[GG_MAX_D];
  gg_pred_plan pl;
};

template <int D>
__device__ __forceinline__ void gg_step_frame(const gg_dyn* dy, const gg_step_args& s,
                                              const double* vge, gg_frame& fr) {
#pragma clang fp contract(off)
  fr.pl = gg_make_plan(dy, s, vge);
  double vr[3] = {fr.pl.v_ref[0] * s.dt, fr.pl.v_ref[1] * s.dt, fr.pl.v_ref[2] * s.dt};
  double ve[3];
  gg_rot_ecef(dy->R, vr, ve);
#pragma unroll
  for (int d = 0; d < GG_MAX_D; ++d) fr.shift[d] = (s.recenter && d < D) ? (float)dy->mean[d] : 0.0f;
  double sp[3] = {(double)fr.shift[0], (double)fr.shift[1], (double)fr.shift[2]};
  double se[3];
  gg_rot_ecef(dy->R, sp, se);
  for (int k = 0; k < 3; ++k) fr.anchor[k] = (dy->anchor[k] + ve[k]) + se[k];
  fr.cb_anchor = dy->cb_anchor + (double)fr.shift[3];
  for (int k = 0; k < GG_MAX_ISB; ++k) fr.isb_anchor[k] = dy->isb_anchor[k] + (double)fr.shift[4 + k];
}

// Commit a step's frame to the record (one thread).  The old posterior mean is
// carried into the new frame.
template <int D>
__device__ __forceinline__ void gg_commit_frame(const gg_step_args& s, gg_dyn* dy,
                                                const gg_frame& fr) {
  for (int k = 0; k < 3; ++k) {
    dy->anchor[k] = fr.anchor[k];
    dy->vel_u[k] = fr.pl.v_new_u[k];
  }
  dy->cb_anchor = fr.cb_anchor;
  for (int k = 0; k < GG_MAX_ISB; ++k) dy->isb_anchor[k] = fr.isb_anchor[k];
  for (int d = 0; d < GG_MAX_D; ++d) dy->shift[d] = fr.shift[d];
  for (int d = 0; d < D; ++d) dy->mean[d] = dy->mean[d] - (double)fr.shift[d];
  dy->vel_mode = fr.pl.mode_out;
  if (s.velocity_kf && fr.pl.mode_out < 2)
    for (int k = 0; k < 6; ++k) dy->vcov_u[k] = fr.pl.vcov_new[k];
  dy->est_valid = 0;
}

// Result of combining a filter's tile partials.
struct gg_comb {
  double M, W, W2;
  double K;     // max tile exponent k_b (g_b = 2^(k_b - K) c_b)
  double excl;  // this thread's exclusive prefix in the chunked tile scan (pass 3)
  double sum[GG_MAX_D];
};

// Combine a filter's tile partials (one workgroup per filter, or redundantly
// in every workgroup that needs the result; the operations and their order
// are the same wherever it runs):
//   pass 1 (strided): M = max m_b and K = max k_b (exact, any order);
//   pass 2 (strided): T_b = 2^(k_b - K) c_b S_b staged in LDS (sT), the ESS
//           sum and (FULL) the weighted-mean sums scaled the same way, all
//           exact power-of-two shifts (gg_tile_scale), accumulated in the
//           strided order; g_b = 2^(k_b - K) c_b;
//   pass 3: W and the CDF tile offsets P_b in the fixed chunked order (thread
//           t owns tiles [t*c, t*c+c), sums them left to right from 0.0, the
//           per-thread values go through gg_block_scan256's operations, then
//           P_b = excl_t + (T_{t*c} + ... + T_{b-1})); the ESS and mean totals
//           share the scan's barriers (gg_block_scan256_totals).
// ALL writes g_b and P_b of every tile.  sT needs 8 * n_tiles bytes of LDS.
// Beyond 4 tiles per thread the strided passes load U tiles' records before
// using them (U round trips in flight instead of one); the order of every
// accumulation is the same for any U.
template <int D, bool FULL, bool ALL, int U = 1>
__device__ void gg_combine(const gg_arena& a, const gg_filter& F, double* lds, double* sT,
                           gg_comb& r) {
#pragma clang fp contract(off)
  constexpr int NV = FULL ? D + 2 : 1;
  __shared__ double ltot[NV * 8];
  const int nt = F.n_tiles;
  const int t = threadIdx.x;
  const double* part = a.part + (int64_t)F.tile_off * GG_NPART;
  const double* p4 = a.p4 + (int64_t)F.tile_off * 4;  // m, cS, c^2 S2, k packed: coalesced reads
  double vv[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) vv[d] = 0.0;
  // pass 2 for tile b given K (scaled S, S2 and the tile's exponent from p4;
  // gl and sums[0 .. D) the partial record's last word and state sums)
  auto pass2v = [&](int b, double Sc, double S2c, double kb, double K, double gl,
                    const double* sums) {
    const double dk = kb - K;
    if (ALL) a.g[F.tile_off + b] = gg_tile_shift(gl, dk);
    sT[b] = gg_tile_shift(Sc, dk);
    vv[0] = vv[0] + gg_tile_shift(S2c, 2.0 * dk);
    if (FULL) {
#pragma unroll
      for (int d = 0; d < D; ++d)
        vv[(FULL ? 2 : 0) + d] = vv[(FULL ? 2 : 0) + d] + gg_tile_shift(sums[d], dk);
    }
  };
  auto pass2 = [&](int b, double Sc, double S2c, double kb, double K) {
    const double* pb = part + (int64_t)b * GG_NPART;
    pass2v(b, Sc, S2c, kb, K, ALL ? pb[GG_NPART - 1] : 0.0, pb + 3);
  };
  constexpr int R = 4;  // tiles per thread held in registers across the maxima
  double M, K;
  if (nt <= R * GG_BLOCK) {
    double pm[R], pS[R], pS2[R], pk[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int b = t + k * GG_BLOCK;
      const double4 q = *reinterpret_cast<const double4*>(p4 + (int64_t)min(b, nt - 1) * 4);
      pm[k] = q.x;
      pS[k] = q.y;
      pS2[k] = q.z;
      pk[k] = q.w;
    }
    double m = -__builtin_huge_val(), km = -__builtin_huge_val();
#pragma unroll
    for (int k = 0; k < R; ++k)
      if (t + k * GG_BLOCK < nt) {
        m = fmax(m, pm[k]);
        km = fmax(km, pk[k]);
      }
    gg_block_max256x2(m, km, lds, M, K);
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int b = t + k * GG_BLOCK;
      if (b < nt) pass2(b, pS[k], pS2[k], pk[k], K);
    }
  } else if (U == 1) {
    double m = -__builtin_huge_val(), km = -__builtin_huge_val();
    for (int b = t; b < nt; b += GG_BLOCK) {
      m = fmax(m, p4[(int64_t)b * 4]);
      km = fmax(km, p4[(int64_t)b * 4 + 3]);
    }
    gg_block_max256x2(m, km, lds, M, K);
    for (int b = t; b < nt; b += GG_BLOCK) {
      const double* q = p4 + (int64_t)b * 4;
      pass2(b, q[1], q[2], q[3], K);
    }
  } else {
    // U tiles per round: their loads first (indices clamped to the last tile;
    // a repeated tile does not change a maximum and is not used in pass 2)
    double m = -__builtin_huge_val(), km = -__builtin_huge_val();
    for (int bb = t; bb < nt; bb += U * GG_BLOCK) {
      double2 q[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int b = min(bb + u * GG_BLOCK, nt - 1);
        q[u] = make_double2(p4[(int64_t)b * 4], p4[(int64_t)b * 4 + 3]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        m = fmax(m, q[u].x);
        km = fmax(km, q[u].y);
      }
    }
    gg_block_max256x2(m, km, lds, M, K);
    constexpr int NW = FULL ? D + 1 : 1;
    for (int bb = t; bb < nt; bb += U * GG_BLOCK) {
      double4 q[U];
      double pv[U][NW];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int b = min(bb + u * GG_BLOCK, nt - 1);
        q[u] = *reinterpret_cast<const double4*>(p4 + (int64_t)b * 4);
        const double* pb = part + (int64_t)b * GG_NPART;
        pv[u][0] = ALL ? pb[GG_NPART - 1] : 0.0;
        if (FULL)
#pragma unroll
          for (int d = 0; d < D; ++d) pv[u][(FULL ? 1 : 0) + d] = pb[3 + d];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (bb + u * GG_BLOCK < nt)
          pass2v(bb + u * GG_BLOCK, q[u].y, q[u].z, q[u].w, K, pv[u][0], &pv[u][FULL ? 1 : 0]);
    }
  }
  __syncthreads();
  const int c = (nt + GG_BLOCK - 1) / GG_BLOCK;
  const int b0 = min(t * c, nt);
  const int b1 = min(b0 + c, nt);
  double T = 0.0;
  for (int b = b0; b < b1; ++b) T = T + sT[b];
  double excl, W;
  double tot[NV];
  gg_block_scan256_totals<NV>(T, vv, lds, ltot, excl, W, tot);
  r.excl = excl;
  r.K = K;
  if (ALL) {
    double run = 0.0;
    for (int b = b0; b < b1; ++b) {
      a.Pb[F.tile_off + b] = excl + run;
      run = run + sT[b];
    }
  }
  r.M = M;
  r.W = W;
  r.W2 = tot[0];
  if (FULL)
    for (int d = 0; d < D; ++d) r.sum[d] = tot[(FULL ? 2 : 0) + d];
}

// Decision part of a combine, written to the (LDS) record by one thread.
template <int D>
__device__ __forceinline__ void gg_store_decision(const gg_step_args& s, const gg_filter& F,
                                                  const gg_comb& r, double u0, gg_dyn* sd) {
  sd->M = r.M;
  sd->W = r.W;
  sd->W2 = r.W2;
  const double ess = (r.W2 > 0.0) ? (r.W * r.W) / r.W2 : 0.0;
  sd->ess = ess;
  for (int d = 0; d < D; ++d) sd->mean[d] = r.sum[d] / r.W;
  if (s.auto_flag) sd->flag = (ess < s.ess_frac * (double)F.n) ? 1 : 0;
  sd->u0 = u0;
}

// Absolute estimate from the local mean (thread 0, after the anchor is final).
template <int D>
__device__ __forceinline__ void gg_make_estimate(gg_dyn* sd) {
  double mp[3] = {sd->mean[0], sd->mean[1], sd->mean[2]};
  double me[3];
  gg_rot_ecef(sd->R, mp, me);
  sd->est[0] = sd->anchor[0] + me[0];
  sd->est[1] = sd->anchor[1] + me[1];
  sd->est[2] = sd->anchor[2] + me[2];
  sd->est[3] = sd->cb_anchor + sd->mean[3];
  for (int d = 4; d < D; ++d) sd->est[d] = sd->isb_anchor[d - 4] + sd->mean[d];
  sd->est_valid = 1;
}

// ---------------------------------------------------------------------------
// Load/store helpers for the 4 consecutive particles of one thread.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void gg_load4(const gg_arena& a, int buf, int64_t base,
                                         float (&p)[GG_ITEMS][D]) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const float4 v = *reinterpret_cast<const float4*>(gg_comp(a, buf, d) + base);
    p[0][d] = v.x;
    p[1][d] = v.y;
    p[2][d] = v.z;
    p[3][d] = v.w;
  }
}

// Packed-record slot of particle slot i (filters start on tile boundaries):
// inside each tile the thread-major order 4t + q becomes q * 256 + t, so the
// records that a wavefront stores for one item q lie in one 2 KB span.
__device__ __forceinline__ int64_t gg_aos_slot(int64_t i) {
  const int64_t j = i & (GG_TILE - 1);
  return (i - j) + ((j & (GG_ITEMS - 1)) << 8) + (j >> 2);
}
static_assert(GG_TILE == 1024 && GG_ITEMS == 4, "gg_aos_slot: 256 threads x 4 items per tile");

// The thread's four particles (slots base .. base + 3) as packed 32-byte
// records (aos): D components, zero above.
template <int D>
__device__ __forceinline__ void gg_store_aos4(const gg_arena& a, int64_t base,
                                              const float (&p)[GG_ITEMS][D]) {
  float4* ao = reinterpret_cast<float4*>(a.aos);
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
    float r[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) r[d] = (d < D) ? p[q][d] : 0.0f;
    const int64_t o = 2 * gg_aos_slot(base + q);
    ao[o] = make_float4(r[0], r[1], r[2], r[3]);
    ao[o + 1] = make_float4(r[4], r[5], r[6], r[7]);
  }
}

template <int D>
__device__ __forceinline__ void gg_store4(const gg_arena& a, int buf, int64_t base,
                                          const float (&p)[GG_ITEMS][D]) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    float4 v;
    v.x = p[0][d];
    v.y = p[1][d];
    v.z = p[2][d];
    v.w = p[3][d];
    *reinterpret_cast<float4*>(gg_comp(a, buf, d) + base) = v;
  }
}
