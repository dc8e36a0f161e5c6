// Particle-filter kernels for gfx950 (CDNA4, 64-lane wavefronts).
//
// One workgroup = 256 threads = 4 wavefronts; each thread owns GG_ITEMS (4)
// consecutive particles, so a workgroup covers one tile of GG_TILE (1024)
// particles and every SoA load/store is a 16-byte vector access that the
// wavefront coalesces into one contiguous 1 KiB transaction.
//
// The systematic step for a batch of filters is four launches, each over all
// filters:
//   gg_k_weight    predict (optional) + pseudorange likelihood + per-tile
//                  partials (max, sums of exp32(lw - m_b), ESS and mean sums).
//   gg_k_cdf       every workgroup combines its filter's partials (max, total,
//                  ESS, CDF tile offset) itself and takes the resampling
//                  decision; the first tile of each filter writes the decision
//                  and estimate; when resampling, per-tile binary64 CDF and
//                  output ends.
//   gg_k_gather    merge-path partition of (ends, outputs), ancestor indices,
//                  state gather.
//   gg_k_finalize  mean of the resampled particles, buffer swap.
// Kernel boundaries are the only cross-workgroup synchronisation: the L2s of
// the 8 XCDs are not coherent, and an agent-scope fence in every workgroup
// (L2 writeback + invalidate) costs more than a launch inside a graph.
// Megopolis: gg_k_weight, gg_k_decide (one workgroup per filter), gg_k_mego
// (or B launches of gg_k_mego_ref), the ancestor-driven gather, finalize.
#pragma once

#include "gg_device.h"
#include "gg_mlp.h"

struct gg_arena {
  float* state;           // [2][ncomp][P]
  float* lw;              // [P]
  int32_t* E;             // [P] tile-local prefix max of the output ends
  int32_t* anc;           // [2][P]
  double* part;           // [n_tiles_total][GG_NPART]
  double* g;              // [n_tiles_total]
  double* Pb;             // [n_tiles_total]
  int32_t* emax;          // [n_tiles_total] tile maxima of E
  double* csum;           // [n_chunks_total][GG_MAX_D]
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
  int32_t pad;
};

__device__ __forceinline__ float* gg_comp(const gg_arena& a, int buf, int comp) {
  return a.state + ((int64_t)buf * a.ncomp + comp) * a.P;
}

// Satellite geometry staged in LDS as binary32 values relative to the anchor.
struct gg_sat_view {
  float* ux;
  float* uy;
  float* uz;
  float* irho;  // 1 / |sat - anchor|
  float* c;     // pr - |sat - anchor| - cb_anchor - isb_anchor
  float* w;
  int* isb;     // index into the ISB offsets, -1 for the reference system
  float* cs;    // learned likelihood: first-layer bias per satellite [cap][64]
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

// Predict one particle in place (binary32 offsets).
template <int NISB>
__device__ __forceinline__ void gg_predict_one(const gg_arena& a, int cur, int64_t slot,
                                               uint64_t seed, uint32_t lidx, uint32_t step,
                                               const gg_step_args& s, const gg_pred_plan& pl,
                                               float* p /*[4+NISB]*/) {
  float z0[4];
  gg_normal4(seed, lidx, step, GG_PURPOSE_PREDICT, 0u, z0);
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
    if (pl.per_particle_noise) {
      float zv[4];
      gg_normal4(seed, lidx, step, GG_PURPOSE_PREDICT, 1u, zv);
      for (int k = 0; k < 3; ++k) v[k] += zv[k] * s.sigma_vel;
    }
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
  if (NISB > 0) {
    float zi[4];
    gg_normal4(seed, lidx, step, GG_PURPOSE_PREDICT, 2u, zi);
#pragma unroll
    for (int k = 0; k < NISB; ++k) p[4 + k] += zi[k] * s.sigma_isb;
  }
}

// Predict the thread's four particles.  With a uniform velocity carried by the
// anchor (plan mode 0) and no velocity filter, the noise draws of the four
// particles are independent straight-line code, so their Philox chains
// interleave; every other mode goes particle by particle.  Both branches give
// the same values for the same particle.
template <int NISB>
__device__ __forceinline__ void gg_predict4(const gg_arena& a, int cur, int64_t base, int n,
                                            uint64_t seed, int i0, uint32_t step,
                                            const gg_step_args& s, const gg_pred_plan& pl,
                                            float (&p)[GG_ITEMS][4 + NISB]) {
  if (pl.mode_out == 0 && !s.velocity_kf) {
    float z[GG_ITEMS][4];
    float zi[GG_ITEMS][4];
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) {
      gg_normal4(seed, (uint32_t)(i0 + q), step, GG_PURPOSE_PREDICT, 0u, z[q]);
      if (NISB > 0) gg_normal4(seed, (uint32_t)(i0 + q), step, GG_PURPOSE_PREDICT, 2u, zi[q]);
    }
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) {
      const float n0 = z[q][0] * s.sigma_pos;
      const float n1 = z[q][1] * s.sigma_pos;
      const float n2 = z[q][2] * s.sigma_pos;
      p[q][0] += 0.0f + n0;
      p[q][1] += 0.0f + n1;
      p[q][2] += 0.0f + n2;
      p[q][3] += z[q][3] * s.sigma_cb;
#pragma unroll
      for (int k = 0; k < NISB; ++k) p[q][4 + k] += zi[q][k] * s.sigma_isb;
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q)
    if (i0 + q < n) gg_predict_one<NISB>(a, cur, base + q, seed, (uint32_t)(i0 + q), step, s, pl, p[q]);
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

// Stage the filter's satellites for this epoch slot in LDS.
// anchor[3], cb_anchor, isb_anchor[3]: the step frame's binary64 origins
// (usually in LDS).
__device__ __forceinline__ int gg_stage_sats(const gg_arena& a, int f, const gg_filter& F,
                                             const gg_dyn* dy, const double* anchor,
                                             const double* cb_anchor, const double* isb_anchor,
                                             int slot, gg_sat_view& sv, char* smem) {
  const double* sd = a.satd + (int64_t)slot * 5 * a.sat_rows;
  const int32_t* si = a.sati + (int64_t)slot * (a.sat_rows + a.n_filters);
  const int ns = si[a.sat_rows + f];
  const int cap = F.sat_cap;
  sv.ux = (float*)smem;
  sv.uy = sv.ux + cap;
  sv.uz = sv.uy + cap;
  sv.irho = sv.uz + cap;
  sv.c = sv.irho + cap;
  sv.w = sv.c + cap;
  sv.isb = (int*)(sv.w + cap);
  sv.cs = (float*)(sv.isb + cap);
  for (int k = threadIdx.x; k < ns; k += blockDim.x) {
    const int64_t r = F.sat_off + k;
    double d[3], rho, inv_rho;
    gg_sat_local(sd, dy, anchor, r, d, rho, inv_rho);
    const int sys = si[r];
    const int ib = (sys >= 1 && sys <= a.n_isb) ? sys - 1 : -1;
    double c = sd[3 * a.sat_rows + r] - rho - *cb_anchor;
    if (ib >= 0) c -= isb_anchor[ib];
    sv.ux[k] = (float)(d[0] * inv_rho);
    sv.uy[k] = (float)(d[1] * inv_rho);
    sv.uz[k] = (float)(d[2] * inv_rho);
    sv.irho[k] = (float)inv_rho;
    sv.c[k] = (float)c;
    sv.w[k] = (float)sd[4 * a.sat_rows + r];
    sv.isb[k] = ib;
  }
  __syncthreads();
  return ns;
}

// Pseudorange residual pr - (|sat - x| + cb) for a binary32 offset particle,
// using |sat-x| - rho = q / (1 + sqrt(1 + q/rho)), q = |p|^2/rho - 2 u.p,
// which never forms the 2e7 m range in binary32.
template <int NISB>
__device__ __forceinline__ float gg_residual(const gg_sat_view& sv, int k, const float* p,
                                             float pp) {
  const float ud = sv.ux[k] * p[0] + sv.uy[k] * p[1] + sv.uz[k] * p[2];
  const float ir = sv.irho[k];
  const float q = pp * ir - 2.0f * ud;
  const float dr = q * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_sqrtf(1.0f + q * ir));
  float e = sv.c[k] - dr - p[3];
  if (NISB > 0) {
    const int ib = sv.isb[k];
#pragma unroll
    for (int j = 0; j < NISB; ++j)
      if (ib == j) e -= p[4 + j];
  }
  return e;
}

// natural log on the hardware log2 unit (relative error ~1e-7 for the
// products of (1 + r^2/(nu s^2)) factors used here)
__device__ __forceinline__ float gg_lnf(float x) {
  return __builtin_amdgcn_logf(x) * 0.6931471805599453f;
}

template <int NISB>
__device__ __forceinline__ float gg_loglik(const gg_sat_view& sv, int ns, const float* p,
                                           const gg_step_args& s) {
  const float pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
  const float sig = s.sigma_pr;
  const float inv_s2 = 1.0f / (sig * sig);
  bool gate = s.gate_thr > 0.0f;
  if (gate) {
    int kept = 0;
    for (int k = 0; k < ns; ++k) {
      const float e = gg_residual<NISB>(sv, k, p, pp);
      kept += (fabsf(e) <= s.gate_thr) ? 1 : 0;
    }
    const int min_kept = ns < 4 ? ns : 4;
    if (kept < min_kept) gate = false;
  }
  float lw = 0.0f;
  if (s.lik == GG_LIK_STUDENT) {
    const float half_nup1 = 0.5f * (s.nu + 1.0f);
    const float inv_nu_s2 = 1.0f / (s.nu * sig * sig);
    // sum of logs = log of the product; the product is taken over groups of
    // three factors so that it stays far from binary32 overflow.
    float acc = 0.0f;
    float prod = 1.0f;
    int cnt = 0;
    for (int k = 0; k < ns; ++k) {
      const float e = gg_residual<NISB>(sv, k, p, pp);
      if (gate && fabsf(e) > s.gate_thr) continue;
      prod *= 1.0f + sv.w[k] * e * e * inv_nu_s2;
      if (++cnt == 3) {
        acc += gg_lnf(prod);
        prod = 1.0f;
        cnt = 0;
      }
    }
    acc += gg_lnf(prod);
    lw = -half_nup1 * acc;
  } else if (s.lik == GG_LIK_HUBER) {
    const float hk = s.huber_k;
    for (int k = 0; k < ns; ++k) {
      const float e = gg_residual<NISB>(sv, k, p, pp);
      if (gate && fabsf(e) > s.gate_thr) continue;
      const float z = fabsf(e) / sig;
      const float rho = (z < hk) ? 0.5f * z * z : hk * z - 0.5f * hk * hk;
      lw -= sv.w[k] * rho;
    }
  } else {
    for (int k = 0; k < ns; ++k) {
      const float e = gg_residual<NISB>(sv, k, p, pp);
      if (gate && fabsf(e) > s.gate_thr) continue;
      lw += -0.5f * sv.w[k] * e * e * inv_s2;
    }
  }
  return lw;
}

// Likelihood of the thread's four particles with the satellites in the outer
// loop, so that the four residual chains interleave.  Per particle the
// arithmetic and its order are those of gg_loglik; with the NLOS gate on the
// kept set differs per particle and gg_loglik is used as is.
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
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) pp[q] = p[q][0] * p[q][0] + p[q][1] * p[q][1] + p[q][2] * p[q][2];
  const float sig = s.sigma_pr;
  if (s.lik == GG_LIK_STUDENT) {
    const float half_nup1 = 0.5f * (s.nu + 1.0f);
    const float inv_nu_s2 = 1.0f / (s.nu * sig * sig);
    float acc[GG_ITEMS], prod[GG_ITEMS];
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) {
      acc[q] = 0.0f;
      prod[q] = 1.0f;
    }
    int cnt = 0;
    for (int k = 0; k < ns; ++k) {
      const float wk = sv.w[k];
#pragma unroll
      for (int q = 0; q < GG_ITEMS; ++q) {
        const float e = gg_residual<NISB>(sv, k, p[q], pp[q]);
        prod[q] *= 1.0f + wk * e * e * inv_nu_s2;
      }
      if (++cnt == 3) {
#pragma unroll
        for (int q = 0; q < GG_ITEMS; ++q) {
          acc[q] += gg_lnf(prod[q]);
          prod[q] = 1.0f;
        }
        cnt = 0;
      }
    }
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) {
      acc[q] += gg_lnf(prod[q]);
      out[q] = -half_nup1 * acc[q];
    }
  } else if (s.lik == GG_LIK_HUBER) {
    const float hk = s.huber_k;
    float lw[GG_ITEMS] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k = 0; k < ns; ++k) {
      const float wk = sv.w[k];
#pragma unroll
      for (int q = 0; q < GG_ITEMS; ++q) {
        const float e = gg_residual<NISB>(sv, k, p[q], pp[q]);
        const float z = fabsf(e) / sig;
        const float rho = (z < hk) ? 0.5f * z * z : hk * z - 0.5f * hk * hk;
        lw[q] -= wk * rho;
      }
    }
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) out[q] = lw[q];
  } else {
    const float inv_s2 = 1.0f / (sig * sig);
    float lw[GG_ITEMS] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k = 0; k < ns; ++k) {
      const float wk = sv.w[k];
#pragma unroll
      for (int q = 0; q < GG_ITEMS; ++q) {
        const float e = gg_residual<NISB>(sv, k, p[q], pp[q]);
        lw[q] += -0.5f * wk * e * e * inv_s2;
      }
    }
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) out[q] = lw[q];
  }
}

// Learned likelihood, per-satellite part: features (elevation and azimuth
// from the staged unit vector, C/N0 and line-of-sight flag of the epoch slot,
// satellite count, ESS fraction of the last decision) and the first-layer
// bias c_s = b1 + W1[:, 1:] f of every satellite, into sv.cs.
__device__ __forceinline__ void gg_stage_learned(const gg_arena& a, const gg_filter& F,
                                                 const gg_dyn* dy, int slot, int ns,
                                                 const gg_mlp_w& w, gg_sat_view& sv) {
  const float* sf = a.satf + (int64_t)slot * 2 * a.sat_rows;
  const float essf = (F.n > 0) ? (float)(dy->ess / (double)F.n) : 1.0f;
  for (int i = threadIdx.x; i < ns * GG_MLP_H; i += blockDim.x) {
    const int k = i / GG_MLP_H;
    const int u = i % GG_MLP_H;
    float f[GG_MLP_NF - 1];
    gg_mlp_features(sv.ux[k], sv.uy[k], sv.uz[k], sf[F.sat_off + k],
                    sf[a.sat_rows + F.sat_off + k], ns, essf, f);
    sv.cs[k * GG_MLP_H + u] = gg_mlp_c(w, u, f);
  }
  __syncthreads();
}

// Learned likelihood through the per-satellite code tables (gg_k_mlp_table):
// the network output for a residual depends only on the bf16 code of its
// scaled value, so a lookup returns exactly what the per-pair network returns.
// The per-particle sum runs over the satellites in the same order as
// gg_loglik_learned, so both give identical log-weights.
template <int NISB>
__device__ __forceinline__ void gg_loglik_table(const gg_sat_view& sv, int ns, int64_t row0,
                                                const float2* __restrict__ tab,
                                                const float (&p)[GG_ITEMS][4 + NISB],
                                                float (&out)[GG_ITEMS]) {
  float pp[GG_ITEMS];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
    pp[q] = p[q][0] * p[q][0] + p[q][1] * p[q][1] + p[q][2] * p[q][2];
    out[q] = 0.0f;
  }
  for (int k = 0; k < ns; ++k) {
    const float2* tk = tab + (row0 + k) * GG_MLP_CODES;
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) {
      const float e = gg_residual<NISB>(sv, k, p[q], pp[q]);
      const float2 z = tk[gg_mlp_code_index(gg_mlp_xe(e))];
      out[q] += gg_mlp_ll(e, z.x, z.y);
    }
  }
}

// Learned likelihood of the workgroup's tile: for every satellite the 1024
// residuals go through LDS to the matrix-core layout (16 particles per pass
// and wavefront), the pair log-likelihoods are summed per particle in LDS.
// se and slw: GG_TILE floats of LDS each.
template <int NISB>
__device__ __forceinline__ void gg_loglik_learned(const gg_sat_view& sv, int ns,
                                                  const float (&p)[GG_ITEMS][4 + NISB],
                                                  const gg_mlp_w& w, float* se, float* slw,
                                                  float (&out)[GG_ITEMS]) {
  const int t = threadIdx.x;
  const int l = t & 63;
  const int base = (t >> 6) * 256;
  float pp[GG_ITEMS];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
    pp[q] = p[q][0] * p[q][0] + p[q][1] * p[q][1] + p[q][2] * p[q][2];
    slw[t * GG_ITEMS + q] = 0.0f;
  }
  gg_mlp_regs wr;
  gg_mlp_load_regs(w, wr);
  for (int k = 0; k < ns; ++k) {
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) se[t * GG_ITEMS + q] = gg_residual<NISB>(sv, k, p[q], pp[q]);
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
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) out[q] = slw[t * GG_ITEMS + q];
}

// ---------------------------------------------------------------------------
// Tile partials: m_b = max lw, S_b = sum e, S2_b = sum e^2, sum_d = sum e p_d
// with e = exp32(lw - m_b) (binary32, like lw itself) accumulated in binary64,
// each sum in the fixed order: per-thread left to
// right over its 4 particles, then gg_block_scan256's total.
// ---------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void gg_block_totals_n(const double (&v)[NV], double* lds, double* out) {
#pragma clang fp contract(off)
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double x = v[q];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const double y = __shfl_up(x, d, 64);
      if (lane >= d) x = y + x;
    }
    if (lane == 63) lds[q * 4 + w] = x;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q)
    out[q] = ((lds[q * 4 + 0] + lds[q * 4 + 1]) + lds[q * 4 + 2]) + lds[q * 4 + 3];
  __syncthreads();
}

// Each value is reduced on its own (same operations whatever the count).
template <int D>
__device__ __forceinline__ void gg_block_totals(const double (&v)[D + 2], double* lds, double* out) {
  gg_block_totals_n<D + 2>(v, lds, out);
}

template <int D>
__device__ __forceinline__ void gg_tile_partials(const float (&lw)[GG_ITEMS],
                                                 const bool (&valid)[GG_ITEMS],
                                                 const float (&p)[GG_ITEMS][D], double* lds,
                                                 double* part_out) {
#pragma clang fp contract(off)
  float m = -__builtin_huge_valf();
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q)
    if (valid[q]) m = fmaxf(m, lw[q]);
  const float mb = (float)gg_block_max256((double)m, lds);
  double e[GG_ITEMS];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q)
    e[q] = (valid[q] && mb > -__builtin_huge_valf()) ? (double)gg_exp32(lw[q] - mb) : 0.0;
  double v[D + 2];
  v[0] = e[0];
  v[1] = e[0] * e[0];
#pragma unroll
  for (int d = 0; d < D; ++d) v[2 + d] = e[0] * (double)p[0][d];
#pragma unroll
  for (int q = 1; q < GG_ITEMS; ++q) {
    v[0] = v[0] + e[q];
    v[1] = v[1] + e[q] * e[q];
#pragma unroll
    for (int d = 0; d < D; ++d) v[2 + d] = v[2 + d] + e[q] * (double)p[q][d];
  }
  double tot[D + 2];
  gg_block_totals<D>(v, lds, tot);
  if (threadIdx.x == 0) {
    part_out[0] = (double)mb;
#pragma unroll
    for (int q = 0; q < D + 2; ++q) part_out[1 + q] = tot[q];
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
  const double est = floor(c * dn - u0);
  int j = (est < -1.0) ? -1 : (est > (double)(n - 1) ? n - 1 : (int)est);
  while (j + 1 < n && (u0 + (double)(j + 1)) / dn <= c) ++j;
  while (j >= 0 && (u0 + (double)j) / dn > c) --j;
  return j + 1;
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
  double M, W, W2, g_own, P_own;
  double sum[GG_MAX_D];
};

// Combine a filter's tile partials (one workgroup per filter, or redundantly
// in every workgroup that needs the result; the operations and their order
// are the same wherever it runs):
//   pass 1 (strided): M = max m_b;
//   pass 2 (strided): g_b = exp64(m_b - M), T_b = S_b g_b staged in LDS (sT),
//           S2 and (FULL) weighted-mean sums accumulated in the strided order;
//   pass 3: W and the CDF tile offsets P_b in the fixed chunked order (thread
//           t owns tiles [t*c, t*c+c), sums them left to right from 0.0, the
//           per-thread values go through gg_block_scan256, then
//           P_b = excl_t + (T_{t*c} + ... + T_{b-1})).
// ALL writes g_b and P_b of every tile; otherwise only tile `own`'s g and P
// are returned.  sT needs 8 * n_tiles bytes of LDS, sx two doubles.
template <int D, bool FULL, bool ALL>
__device__ void gg_combine(const gg_arena& a, const gg_filter& F, int own, double* lds,
                           double* sT, double* sx, gg_comb& r) {
#pragma clang fp contract(off)
  constexpr int NV = FULL ? D + 2 : 1;
  const int nt = F.n_tiles;
  const int t = threadIdx.x;
  const double* part = a.part + (int64_t)F.tile_off * GG_NPART;
  double m = -__builtin_huge_val();
  for (int b = t; b < nt; b += GG_BLOCK) m = fmax(m, part[(int64_t)b * GG_NPART]);
  const double M = gg_block_max256(m, lds);
  double vv[NV];
#pragma unroll
  for (int d = 0; d < NV; ++d) vv[d] = 0.0;
  for (int b = t; b < nt; b += GG_BLOCK) {
    const double* pb = part + (int64_t)b * GG_NPART;
    const double mb = pb[0];
    const double gb = (mb > -__builtin_huge_val()) ? gg_exp64(mb - M) : 0.0;
    if (ALL) a.g[F.tile_off + b] = gb;
    if (!ALL && b == own) sx[1] = gb;
    sT[b] = pb[1] * gb;
    vv[0] = vv[0] + pb[2] * (gb * gb);
    if (FULL) {
#pragma unroll
      for (int d = 0; d < D; ++d) vv[(FULL ? 2 : 0) + d] = vv[(FULL ? 2 : 0) + d] + pb[3 + d] * gb;
    }
  }
  __syncthreads();
  const int c = (nt + GG_BLOCK - 1) / GG_BLOCK;
  const int b0 = min(t * c, nt);
  const int b1 = min(b0 + c, nt);
  double T = 0.0;
  for (int b = b0; b < b1; ++b) T = T + sT[b];
  double excl, W;
  gg_block_scan256(T, lds, excl, W);
  if (ALL) {
    double run = 0.0;
    for (int b = b0; b < b1; ++b) {
      a.Pb[F.tile_off + b] = excl + run;
      run = run + sT[b];
    }
  } else if (own >= b0 && own < b1) {
    double run = 0.0;
    for (int b = b0; b < own; ++b) run = run + sT[b];
    sx[0] = excl + run;
  }
  double tot[NV];
  gg_block_totals_n<NV>(vv, lds, tot);  // synchronises: sx is visible below
  r.M = M;
  r.W = W;
  r.W2 = tot[0];
  if (!ALL) {
    r.P_own = sx[0];
    r.g_own = sx[1];
  }
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
