// Data layout shared by the host library and the kernels.
//
// A handle holds one *batch* of independent particle filters (a single
// ParticleFilterDevice is a batch of one).  All filters of a batch share one
// structure-of-arrays arena in HBM:
//
//   state[2][NCOMP][P]   binary32, double-buffered for resampling.
//                        comp 0..2  position offset (local ENU frame) [m]
//                        comp 3     clock-bias offset [m]
//                        comp 4..   inter-system-bias offsets [m] (n_isb)
//                        then, if velocity storage is on, 3 velocity comps
//                        and 6 velocity-covariance comps (c00 c01 c02 c11 c12 c22)
//   lw[P]                binary32 log-weights
//   E[P]                 int32 systematic-resampling output ends (see gg_pf.hip)
//   anc[2][P]            int32 ancestor indices (anc[0] is the public copy)
//
// Filter f owns slots [p_off, p_off + n) with p_off a multiple of GG_TILE, so
// every tile (GG_TILE consecutive slots) belongs to exactly one filter.
// Absolute values are anchor (binary64, per filter) + R^T * offset.
#pragma once

#include <stdint.h>

#define GG_BLOCK 256
#define GG_ITEMS 4
#define GG_TILE (GG_BLOCK * GG_ITEMS)
#define GG_QCHUNK (2 * GG_TILE)  // outputs per resampling work item (at most)
#define GG_MAX_PARTICLES (7680 * GG_TILE)  // per filter (per-tile values kept in LDS)
#define GG_MAX_ISB 3
#define GG_MAX_D (4 + GG_MAX_ISB)
#define GG_NVEL 9
// doubles per tile partial: m, c S, c^2 S2, c sum[GG_MAX_D], k, c (see gg_tile_scale)
#define GG_NPART 12
#define GG_REC 3  // float4 per staged satellite record (gg_sat_view)

#define GG_PURPOSE_INIT 1u
#define GG_PURPOSE_PREDICT 2u
#define GG_PURPOSE_MEGO_REF 3u
#define GG_PURPOSE_MEGO 4u
#define GG_PURPOSE_U0 5u
#define GG_PURPOSE_MEGO_U 6u

// likelihood kinds
#define GG_LIK_GAUSS 0
#define GG_LIK_STUDENT 1
#define GG_LIK_HUBER 2
#define GG_LIK_LEARNED 3  // per-pair MLP (gg_mlp), needs satellite C/N0 and LOS flags
#define GG_LIK_GMM 4      // LOS / NLOS Gaussian mixture (gmm_* below, sigma_pr the LOS sigma)

struct gg_filter {
  int64_t p_off;     // first particle slot in the arena
  int32_t n;         // live particles
  int32_t n_tiles;   // ceil(n / GG_TILE)
  int32_t tile_off;  // first global tile id
  int32_t n_chunks;  // resampling work items: n_tiles + ceil(n / GG_QCHUNK)
  int32_t chunk_off; // first global gather chunk id
  int32_t sat_off;   // first satellite row in an epoch slot
  int32_t sat_cap;   // satellite rows reserved for this filter
  int32_t pad0;
  uint64_t seed;
};

struct gg_dyn {
  double anchor[3];     // ECEF origin of the offsets [m]
  double cb_anchor;     // clock-bias origin [m]
  double isb_anchor[GG_MAX_ISB];
  double R[9];          // row-major rotation ECEF -> local
  double vel_u[3];      // uniform velocity (local frame) when vel_mode == 0
  double vcov_u[6];     // uniform velocity covariance (local) when vel_mode < 2
  double M;             // max log-weight of the last reduction
  double W;             // sum exp(lw - M) (fixed order)
  double W2;            // sum exp(lw - M)^2
  double ess;
  double u0;            // systematic offset of the last decision
  double mean[GG_MAX_D];  // weighted mean offset of the last reduction
  double est[GG_MAX_D];   // absolute estimate [x y z cb isb...]
  float shift[GG_MAX_D];  // re-centring shift applied by the next gather
  int32_t cur;          // state buffer holding the particles
  int32_t lw_zero;      // log-weights are logically all zero
  int32_t flag;         // resample requested by the last decision
  int32_t vel_mode;     // 0 uniform velocity, 1 per-particle velocity, 2 + covariance
  int32_t epoch;        // device step counter (graph-driven mode)
  int32_t resample_count;
  int32_t did_resample; // last step resampled (set by gather, cleared by weight)
  int32_t est_valid;
  uint32_t cnt_a;       // last-block counters
  uint32_t cnt_b;
  uint32_t cnt_c;
  int32_t aos_only;     // this Megopolis step keeps its predicted state in the packed
                        // records only (set by the likelihood launch)
};

// Per-launch scalar arguments shared by every filter of a batch.
struct gg_step_args {
  // predict
  double dt;
  float sigma_pos;
  float sigma_cb;
  float sigma_isb;
  float sigma_vel;
  float alpha;
  float qv;
  int32_t velocity_kf;
  int32_t step;          // Philox step when step_from_device == 0
  int32_t step_from_device;
  // weight
  float sigma_pr;
  float nu;
  float gate_thr;        // per-particle NLOS gate threshold [m] (<= 0: off)
  float huber_k;
  int32_t lik;           // GG_LIK_*
  int32_t slot_from_device;  // epoch slot = dyn.epoch % n_slots
  int32_t n_slots;
  // decision
  double ess_frac;       // resample when ess < ess_frac * n
  int32_t auto_flag;     // write the decision into dyn.flag
  int32_t u0_mode;       // 0: reference (seed % 1e6) / (1e6 * n); 1: fresh Philox;
                         // 2: u0_value
  int32_t recenter;      // gather re-centres offsets on the weighted mean
  int32_t learned_mode;  // GG_LIK_LEARNED: 0 network per pair, 1 per-satellite code table
  double u0_value;
  // GG_LIK_GMM: weight of the LOS component, mean and sigma of the NLOS one [m]
  float gmm_w_los;
  float gmm_mu_nlos;
  float gmm_sigma_nlos;
  // derived by the library before every launch (callers leave them as they are):
  // log(w_los) - log(sigma_pr), log(1 - w_los) - log(sigma_nlos) and
  // sigma_pr^2 / sigma_nlos^2, in binary64 from the binary32 fields, rounded once
  float gmm_c_los;
  float gmm_c_nlos;
  float gmm_ratio;
};
