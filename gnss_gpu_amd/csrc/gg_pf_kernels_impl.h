// Particle-filter kernels (see gg_pf_kernels.h for the launch plan).
#include "gg_bvh.h"
#include "gg_pf_kernels.h"

#include <climits>

// ---------------------------------------------------------------------------
// initialise: offsets = spread * N(0,1) around the anchor the host wrote.
// ---------------------------------------------------------------------------
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_init(gg_arena a, float spread_pos,
                                                      float spread_cb, float spread_isb,
                                                      float spread_vel) {
  constexpr int D = 4 + NISB;
  const int tile = blockIdx.x;
  const int f = a.tile_filter[tile];
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int lt = tile - F.tile_off;
  const int cur = dy->cur;
  const int i0 = lt * GG_TILE + threadIdx.x * GG_ITEMS;
  const int64_t base = F.p_off + i0;
  float p[GG_ITEMS][D];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
#pragma unroll
    for (int d = 0; d < D; ++d) p[q][d] = 0.0f;
    const int i = i0 + q;
    if (i >= F.n) continue;
    float z[4];
    gg_normal4(F.seed, (uint32_t)i, 0u, GG_PURPOSE_INIT, 0u, z);
    p[q][0] = z[0] * spread_pos;
    p[q][1] = z[1] * spread_pos;
    p[q][2] = z[2] * spread_pos;
    p[q][3] = z[3] * spread_cb;
    if (NISB > 0) {
      gg_normal4(F.seed, (uint32_t)i, 0u, GG_PURPOSE_INIT, 1u, z);
#pragma unroll
      for (int k = 0; k < NISB; ++k) p[q][4 + k] = z[k] * spread_isb;
    }
    if (dy->vel_mode >= 1) {
      gg_normal4(F.seed, (uint32_t)i, 0u, GG_PURPOSE_INIT, 2u, z);
      for (int k = 0; k < 3; ++k)
        gg_comp(a, cur, D + k)[base + q] = (float)dy->vel_u[k] + z[k] * spread_vel;
    }
  }
  gg_store4<D>(a, cur, base, p);
}

// ---------------------------------------------------------------------------
// predict only (drop-in path): in place on the current buffer, offsets
// re-centred on the step frame.  The frame is committed afterwards by
// gg_k_commit_predict (kernel boundaries are the only cross-workgroup
// synchronisation in this library: a device-scope fence costs an L2
// write-back per workgroup on gfx950).
// ---------------------------------------------------------------------------
__device__ __forceinline__ const double* gg_guide_ptr(const gg_arena& a, const gg_step_args& s,
                                                      int f, int epoch) {
  const int slot = s.slot_from_device ? (epoch % s.n_slots) : 0;
  return a.vel_guide + ((int64_t)slot * a.n_filters + f) * 3;
}

template <int D>
__device__ __forceinline__ void gg_apply_shift(const gg_frame& fr, float (&p)[GG_ITEMS][D]) {
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q)
#pragma unroll
    for (int d = 0; d < D; ++d) p[q][d] -= fr.shift[d];
}

template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_predict(gg_arena a, gg_step_args s) {
  constexpr int D = 4 + NISB;
  const int tile = blockIdx.x;
  const int f = a.tile_filter[tile];
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int lt = tile - F.tile_off;
  const int cur = dy->cur;
  const uint32_t step = s.step_from_device ? (uint32_t)(dy->epoch + 1) : (uint32_t)s.step;
  const double* vg = gg_guide_ptr(a, s, f, dy->epoch);
  double vge[3] = {vg[0], vg[1], vg[2]};
  gg_frame fr;
  gg_step_frame<D>(dy, s, vge, fr);
  const int i0 = lt * GG_TILE + threadIdx.x * GG_ITEMS;
  const int64_t base = F.p_off + i0;
  float p[GG_ITEMS][D];
  gg_load4<D>(a, cur, base, p);
  gg_predict4<NISB>(a, cur, base, F.n, F.seed, i0, step, s, fr.pl, p);
  gg_apply_shift<D>(fr, p);
  gg_store4<D>(a, cur, base, p);
}

// Commit of a predict's frame (one thread per filter), from the same record
// the predict blocks read.
template <int NISB>
__global__ void gg_k_commit_predict(gg_arena a, gg_step_args s) {
  constexpr int D = 4 + NISB;
  const int f = blockIdx.x;
  if (threadIdx.x != 0 || f >= a.n_filters) return;
  gg_dyn* dy = a.dyn + f;
  const double* vg = gg_guide_ptr(a, s, f, dy->epoch);
  double vge[3] = {vg[0], vg[1], vg[2]};
  gg_frame fr;
  gg_step_frame<D>(dy, s, vge, fr);
  gg_commit_frame<D>(s, dy, fr);
}

// ---------------------------------------------------------------------------
// predict (optional) + likelihood + per-tile partials.
// ---------------------------------------------------------------------------
// predict (PRED) + likelihood + tile partials.  LEARN: 0 the analytic
// likelihood of s.lik, 1 the learned network per pair, 2 the learned network
// through its per-satellite code tables.  The workgroup handles tiles
// [tb, te) in order; the step frame (computed by one thread into LDS, so its
// binary64 values do not occupy registers) and the satellite records are
// staged once per filter met.  PREFETCH issues the loads of tile k + 1 before
// tile k is computed, so the memory traffic of a multi-tile workgroup
// overlaps its arithmetic instead of waiting for it.  AOS (a Megopolis step,
// with a.aos_step set) also writes the predicted state as packed records.
template <int NISB, bool PRED, int LEARN, bool PREFETCH, bool AOS = false>
__device__ __forceinline__ void gg_weight_body(const gg_arena& a, const gg_step_args& s,
                                               char* smem, int tb, int te) {
  constexpr int D = 4 + NISB;
  __shared__ double lds[8 + GG_BLOCK];
  __shared__ gg_frame sfr;
  // the slot of particle 4t of tile `tile` (filter, buffer, index)
  auto slot_of = [&](int tile, int& f, int64_t& base, int& i0, int& cur) {
    f = a.tile_filter[tile];
    const gg_filter F = a.filt[f];
    i0 = (tile - F.tile_off) * GG_TILE + threadIdx.x * GG_ITEMS;
    base = F.p_off + i0;
    cur = a.dyn[f].cur;
  };
  float pn[GG_ITEMS][D];
  if (PREFETCH) {
    int f, i0, cur;
    int64_t base;
    slot_of(tb, f, base, i0, cur);
    gg_load4<D>(a, cur, base, pn);
  }
  int f_staged = -1;
  int ns = 0;
  gg_sat_view sv;
  for (int tile = tb; tile < te; ++tile) {
    int f, i0, cur;
    int64_t base;
    slot_of(tile, f, base, i0, cur);
    const gg_filter F = a.filt[f];
    const gg_dyn* dy = a.dyn + f;
    const int epoch = dy->epoch;
    const int slot = s.slot_from_device ? (epoch % s.n_slots) : 0;
    const uint32_t step = s.step_from_device ? (uint32_t)(epoch + 1) : (uint32_t)s.step;
    // A Megopolis step writes the predicted state as packed records, which the
    // chains read.  After a step that resampled it writes them only (the next
    // is likely to resample too; if it does not, gg_k_mego_w copies them back),
    // otherwise the state arrays as well.  Every workgroup of the filter
    // records the same choice.
    const bool aos_only = AOS && dy->did_resample;
    if (PRED && AOS && threadIdx.x == 0) a.dyn[f].aos_only = aos_only ? 1 : 0;
    float p[GG_ITEMS][D];
    if (PREFETCH) {
#pragma unroll
      for (int q = 0; q < GG_ITEMS; ++q)
#pragma unroll
        for (int d = 0; d < D; ++d) p[q][d] = pn[q][d];
      if (tile + 1 < te) {
        int f1, i1, cur1;
        int64_t base1;
        slot_of(tile + 1, f1, base1, i1, cur1);
        gg_load4<D>(a, cur1, base1, pn);
      }
    } else {
      gg_load4<D>(a, cur, base, p);
    }
    // the fast predict path (pl.mode_out == 0, no velocity filter) needs no
    // frame: its noise is drawn while the state loads are in flight and added
    // before the staging below
    const bool fast = PRED && dy->vel_mode == 0 && !s.velocity_kf && !(s.sigma_vel > 0.0f);
    if (fast) {
      float zn[GG_ITEMS][D];
      gg_noise4_fast<NISB>(F.seed, (uint32_t)i0 >> 2, step, zn);
      gg_apply_noise4_fast<D>(zn, s, p);
    }
    if (f != f_staged) {  // uniform over the workgroup
      if (f_staged >= 0) __syncthreads();  // the previous filter's records are no longer read
      if (threadIdx.x == 0) {
        if (PRED) {
          const double* vg = gg_guide_ptr(a, s, f, epoch);
          double vge[3] = {vg[0], vg[1], vg[2]};
          gg_step_frame<D>(dy, s, vge, sfr);
        } else {
          for (int k = 0; k < 3; ++k) sfr.anchor[k] = dy->anchor[k];
          sfr.cb_anchor = dy->cb_anchor;
          for (int k = 0; k < GG_MAX_ISB; ++k) sfr.isb_anchor[k] = dy->isb_anchor[k];
        }
      }
      __syncthreads();
      ns = gg_stage_sats(a, f, F, dy, sfr.anchor, &sfr.cb_anchor, sfr.isb_anchor, slot, s, sv,
                         smem);
      f_staged = f;
    }
    bool valid[GG_ITEMS];
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) valid[q] = (i0 + q) < F.n;
    if (PRED) {
      if (!fast) gg_predict4<NISB>(a, cur, base, F.n, F.seed, i0, step, s, sfr.pl, p);
      gg_apply_shift<D>(sfr, p);
      if (!aos_only) gg_store4<D>(a, cur, base, p);
    }
    if (AOS) gg_store_aos4<D>(a, base, p);
    float lw[GG_ITEMS];
    if constexpr (LEARN == 2) {
      gg_loglik_table<NISB>(sv, F.sat_off, a.mlp_table, p, lw);
    } else if constexpr (LEARN == 1) {
      __shared__ gg_mlp_w smlp;
      __shared__ float se[GG_TILE];
      __shared__ float slw[GG_TILE];
      {
        const uint4* src = reinterpret_cast<const uint4*>(a.mlp);
        uint4* dst = reinterpret_cast<uint4*>(&smlp);
        for (int i = threadIdx.x; i < (int)(sizeof(gg_mlp_w) / 16); i += GG_BLOCK) dst[i] = src[i];
      }
      __syncthreads();
      gg_stage_learned(a, F, dy, slot, ns, smlp, sv);
      gg_loglik_learned<NISB>(sv, p, smlp, se, slw, lw);
    } else {
      gg_loglik4<NISB>(sv, ns, p, s, lw);
    }
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) lw[q] = valid[q] ? lw[q] : 0.0f;
    *reinterpret_cast<float4*>(a.lw + base) = make_float4(lw[0], lw[1], lw[2], lw[3]);
    gg_tile_partials<D>(lw, valid, p, lds, a.part + (int64_t)tile * GG_NPART,
                        a.p4 + (int64_t)tile * 4);
  }
}

// One round of workgroups covers 1M particles only at 4 waves per SIMD, so the
// register budget of the plain likelihood is held to 128 VGPRs (one tile per
// workgroup).  The multi-tile form (gg_k_weight_mt) runs tpw tiles per
// workgroup with the next tile's state prefetched.
template <int NISB, bool PRED, bool AOS = false>
__global__ __launch_bounds__(GG_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
void gg_k_weight(gg_arena a, gg_step_args s) {
  extern __shared__ __align__(16) char smem[];
  gg_weight_body<NISB, PRED, 0, false, AOS>(a, s, smem, blockIdx.x, blockIdx.x + 1);
}

template <int NISB, bool PRED>
__global__ __launch_bounds__(GG_BLOCK) __attribute__((amdgpu_waves_per_eu(2)))
void gg_k_weight_mt(gg_arena a, gg_step_args s, int tpw) {
  extern __shared__ __align__(16) char smem[];
  const int tb = blockIdx.x * tpw;
  gg_weight_body<NISB, PRED, 0, true>(a, s, smem, tb, min(tb + tpw, a.n_tiles));
}

template <int NISB, bool PRED>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_weight_learned(gg_arena a, gg_step_args s) {
  extern __shared__ __align__(16) char smem[];
  gg_weight_body<NISB, PRED, 1, false>(a, s, smem, blockIdx.x, blockIdx.x + 1);
}

template <int NISB, bool PRED, bool AOS = false>
__global__ __launch_bounds__(GG_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
void gg_k_weight_table(gg_arena a, gg_step_args s) {
  extern __shared__ __align__(16) char smem[];
  gg_weight_body<NISB, PRED, 2, false, AOS>(a, s, smem, blockIdx.x, blockIdx.x + 1);
}

// Per-satellite code tables of the learned likelihood: for satellite row
// blockIdx.y of its filter, the network outputs (z3_0, z3_1) for every bf16
// code of the scaled residual, 1024 codes per workgroup (16 per wavefront
// pass), with the features the weight kernel would derive for that satellite
// in the same step (same frame, same staging arithmetic).
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_mlp_table(gg_arena a, gg_step_args s, int pred,
                                                           const int32_t* __restrict__ row_filter) {
  constexpr int D = 4 + NISB;
  __shared__ gg_frame sfr;
  __shared__ float sc[GG_MLP_H];
  __shared__ float su[3];
  const int row = blockIdx.y;
  const int f = row_filter[row];
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int k = row - F.sat_off;
  const int epoch = dy->epoch;
  const int slot = s.slot_from_device ? (epoch % s.n_slots) : 0;
  const int32_t* si = a.sati + (int64_t)slot * (a.sat_rows + a.n_filters);
  const int ns = si[a.sat_rows + f];
  if (k < 0 || k >= ns) return;
  const int t = threadIdx.x;
  if (t == 0) {
    if (pred) {
      const double* vg = gg_guide_ptr(a, s, f, epoch);
      double vge[3] = {vg[0], vg[1], vg[2]};
      gg_step_frame<D>(dy, s, vge, sfr);
    } else {
      for (int q = 0; q < 3; ++q) sfr.anchor[q] = dy->anchor[q];
    }
    const double* sd = a.satd + (int64_t)slot * 5 * a.sat_rows;
    double d[3], rho, inv_rho;
    gg_sat_local(sd, dy, sfr.anchor, row, d, rho, inv_rho);
    su[0] = (float)(d[0] * inv_rho);
    su[1] = (float)(d[1] * inv_rho);
    su[2] = (float)(d[2] * inv_rho);
  }
  __syncthreads();
  const gg_mlp_w& w = *a.mlp;
  if (t < GG_MLP_H) {
    const float* sf = a.satf + (int64_t)slot * 2 * a.sat_rows;
    const float essf = (F.n > 0) ? (float)(dy->ess / (double)F.n) : 1.0f;
    float fe[GG_MLP_NF - 1];
    gg_mlp_features(su[0], su[1], su[2], sf[row], sf[a.sat_rows + row], ns, essf, fe);
    sc[t] = gg_mlp_c(w, t, fe);
  }
  __syncthreads();
  gg_bf16x8 a1[4];
  gg_mlp_a1(w.w_e, sc, a1);
  gg_mlp_regs wr;
  gg_mlp_load_regs(w, wr);
  const int l = t & 63;
  float2* out = a.mlp_table + (int64_t)row * GG_MLP_CODES;
  for (int cb = 0; cb < 16; ++cb) {
    const int idx = blockIdx.x * 1024 + (t >> 6) * 256 + cb * 16 + (l & 15);
    gg_mlp_act act;
    gg_mlp_fwd16(wr, a1, gg_mlp_code_value(idx < GG_MLP_CODES ? idx : 0), act);
    if (l < 16 && idx < GG_MLP_CODES) out[idx] = make_float2(act.z30, act.z31);
  }
}

// ---------------------------------------------------------------------------
// Tile partials over the stored log-weights (no likelihood evaluation).
// ---------------------------------------------------------------------------
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_reduce(gg_arena a) {
  constexpr int D = 4 + NISB;
  __shared__ double lds[8 + GG_BLOCK];
  const int tile = blockIdx.x;
  const int f = a.tile_filter[tile];
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int lt = tile - F.tile_off;
  const int cur = dy->cur;
  const int i0 = lt * GG_TILE + threadIdx.x * GG_ITEMS;
  const int64_t base = F.p_off + i0;
  float p[GG_ITEMS][D];
  gg_load4<D>(a, cur, base, p);
  bool valid[GG_ITEMS];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) valid[q] = (i0 + q) < F.n;
  float lw[GG_ITEMS] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (!dy->lw_zero) {
    const float4 v = *reinterpret_cast<const float4*>(a.lw + base);
    lw[0] = v.x;
    lw[1] = v.y;
    lw[2] = v.z;
    lw[3] = v.w;
  }
  gg_tile_partials<D>(lw, valid, p, lds, a.part + (int64_t)tile * GG_NPART,
                      a.p4 + (int64_t)tile * 4);
}

// ---------------------------------------------------------------------------
// Decision: one workgroup per filter combines the tile partials.
// mode bit 0: commit the predict that preceded the weight launch;
// mode bit 1: the partials come from a fresh likelihood (clears lw_zero,
//             did_resample and advances the device epoch).
// ---------------------------------------------------------------------------
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_decide(gg_arena a, gg_step_args s, int mode) {
  constexpr int D = 4 + NISB;
  extern __shared__ __align__(16) char smem[];
  __shared__ double lds[8 + GG_BLOCK];
  __shared__ gg_dyn sd;
  __shared__ double svg[3];
  const int f = blockIdx.x;
  const gg_filter F = a.filt[f];
  double* sT = reinterpret_cast<double*>(smem);
  gg_dyn* dy = a.dyn + f;
  gg_dyn_copy(&sd, dy);
  __syncthreads();
  // the guide velocity load overlaps the combine
  double vg = 0.0;
  if ((mode & 1) && threadIdx.x < 3) vg = gg_guide_ptr(a, s, f, sd.epoch)[threadIdx.x];
  gg_comb r;
  gg_combine<D, true, true, 4>(a, F, lds, sT, r);
  if ((mode & 1) && threadIdx.x < 3) svg[threadIdx.x] = vg;
  __syncthreads();
  if (threadIdx.x == 0) {
    if (mode & 1) {
      gg_frame fr;
      gg_step_frame<D>(&sd, s, svg, fr);
      gg_commit_frame<D>(s, &sd, fr);
    }
    gg_store_decision<D>(s, F, r, gg_u0(s, F.seed, F.n, sd.resample_count), &sd);
    gg_make_estimate<D>(&sd);
    if (mode & 2) {
      sd.lw_zero = 0;
      sd.did_resample = 0;
      if (s.slot_from_device || s.step_from_device) sd.epoch = sd.epoch + 1;
    }
  }
  __syncthreads();
  gg_dyn_copy(dy, &sd);
}

// ---------------------------------------------------------------------------
// Resampling gather: copies the state of each output's ancestor (all loads of
// a thread in flight before its stores), then the item's finalize.
// anc_mode: 0 the ancestors the systematic resampler staged in LDS (sA),
// 1 ancestors from anc[src_buf], 2 identity, 3 ancestors from anc[src_buf]
// with the states read from the packed copy that gg_k_mego_w wrote (aos).
// ---------------------------------------------------------------------------

// End of a gather work item: the chunk's state sums go to its slot and, as
// fixed-point integers (units of 2^-24), into its arrival sub-group's totals
// (device-scope integer atomics), then the item counts itself in (gg_arrive_last); the item
// that arrives last reads the 32 sub-group totals in one round of loads (the
// same integers whichever item is last) and finalises the resampling: mean of
// the resampled particles as the estimate, buffer swap, weights reset,
// counters.  Only when a sum did not fit the fixed point does it add every
// chunk's binary64 sums in chunk order instead.  Nobody waits on anybody.
// Visibility without an L2 writeback: the sums are device-scope atomic stores
// and additions (performed at the device-coherent level), the wave waits for
// their completion before its counter increment (an atomic, device-coherent),
// and the last item reads them with device-scope atomic loads (from the
// coherent level).  A release/acquire pair would also write back and
// invalidate the whole L2 of every item's XCD, which costs more than the
// launch it replaces (gather 23 -> 52 us at 1M particles).
template <int D>
__device__ __forceinline__ void gg_gather_finish(const gg_arena& a, int f, const gg_filter& F,
                                                 int chunk, int n_items,
                                                 const double (&tot)[D + 2], double* lds,
                                                 int* ilds) {
  const int t = threadIdx.x;
  int64_t* acc = a.acc + (int64_t)f * GG_ARRIVE_SUBS * GG_ACC_W;
  if (t < 64) {
    // the item's sums (formed in thread 0), one component per lane: binary64 in
    // the chunk's slot, and rounded to fixed point (units of 2^-24) added into
    // the arrival sub-group's words; integer additions give the same totals in
    // any order.  A sum too large for the 63-bit totals (or not finite) raises
    // the sub-group's flag instead.
    double x = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double v = gg_readlane0_f64(tot[2 + d]);
      if (t == d) x = v;
    }
    int64_t* ag = acc + (int64_t)((chunk - F.chunk_off) % GG_ARRIVE_SUBS) * GG_ACC_W;
    const double lim = 0x1p62 / GG_ACC_SCALE / (double)n_items;
    const bool fits = !(t < D) || fabs(x) < lim;
    if (t < D) {
      __hip_atomic_store(a.csum + (int64_t)t * a.n_chunks + chunk, x, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      const int64_t q = fits ? (int64_t)__builtin_rint(x * GG_ACC_SCALE) : 0;
      if (q != 0) __hip_atomic_fetch_add(ag + t, q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (__builtin_amdgcn_ballot_w64(!fits) != 0 && t == 0)
      __hip_atomic_fetch_or(ag + GG_ACC_W - 1, (int64_t)1, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (t == 0)
      ilds[0] = gg_arrive_last(a.fin + (int64_t)f * GG_ARRIVE_WORDS, chunk - F.chunk_off,
                               n_items) ? 1 : 0;
  }
  __syncthreads();
  if (!ilds[0]) return;
  // The last item.  The record may have been written in this launch by another
  // workgroup (the decision of the systematic resampler's first item): it is
  // read and written through the device-coherent level.  Its loads and the
  // sub-groups' fixed-point words are one round of loads.
  gg_dyn* dy = a.dyn + f;
  constexpr int NW = (int)(sizeof(gg_dyn) / 4);
  static_assert(NW <= GG_BLOCK, "gg_dyn larger than one word per thread");
  static_assert(GG_ARRIVE_SUBS * GG_ACC_W == GG_BLOCK, "one fixed-point word per thread");
  uint32_t dw = 0u;
  if (t < NW)
    dw = __hip_atomic_load(reinterpret_cast<const uint32_t*>(dy) + t, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  const int64_t w = __hip_atomic_load(acc + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __shared__ gg_dyn sd;
  __shared__ unsigned long long sacc[GG_ACC_W];
  if (t < NW) reinterpret_cast<uint32_t*>(&sd)[t] = dw;
  if (t < GG_ACC_W) sacc[t] = 0ull;
  __syncthreads();
  // word t % GG_ACC_W of sub-group t / GG_ACC_W (two's complement sums, exact
  // and order-free)
  atomicAdd(&sacc[t % GG_ACC_W], (unsigned long long)w);
  // cleared for the next launch
  __hip_atomic_store(acc + t, (int64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const bool fixed = (sacc[GG_ACC_W - 1] == 0ull);  // uniform
  double sum[D + 2];
#pragma unroll
  for (int d = 0; d < D + 2; ++d) sum[d] = 0.0;
  if (fixed) {
    if (t == 0)
#pragma unroll
      for (int d = 0; d < D; ++d)
        sum[2 + d] = (double)(int64_t)sacc[d] * (1.0 / GG_ACC_SCALE);
  } else {
    // the chunks' binary64 sums, added in chunk order per thread, then over the block
    double v[D + 2];
#pragma unroll
    for (int d = 0; d < D + 2; ++d) v[d] = 0.0;
    // four chunks' sums in flight per thread before they are added;
    // component-major, so a wavefront's loads of one component are contiguous
    constexpr int U = 4;
    for (int c0 = t; c0 < n_items; c0 += U * GG_BLOCK) {
      double x[U][D];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int cc = min(c0 + u * GG_BLOCK, n_items - 1);
#pragma unroll
        for (int d = 0; d < D; ++d)
          x[u][d] = __hip_atomic_load(a.csum + (int64_t)d * a.n_chunks + F.chunk_off + cc,
                                      __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (c0 + u * GG_BLOCK < n_items)
#pragma unroll
          for (int d = 0; d < D; ++d) v[2 + d] = v[2 + d] + x[u][d];
    }
    gg_block_totals<D>(v, lds, sum);
  }
  if (t == 0) {
    const double n = (double)F.n;
    for (int d = 0; d < D; ++d) sd.mean[d] = sum[2 + d] / n;
    gg_make_estimate<D>(&sd);
    sd.cur = 1 - sd.cur;
    sd.lw_zero = 1;
    sd.flag = 0;
    sd.did_resample = 1;
    sd.resample_count += 1;
    sd.ess = n;
    sd.M = 0.0;
    sd.W = n;
    sd.W2 = n;
  }
  __syncthreads();
  gg_dyn_store_coherent(dy, &sd);
}

// Copy of one work item's outputs [jbase, jbase + nout) from their ancestors
// (sA for the systematic merge, anc[src_buf] or the identity otherwise; the
// packed states for anc_mode 3), four outputs per thread and round with every
// load of a thread issued before its stores, then the item's finalize.
template <int D>
__device__ __forceinline__ void gg_gather_copy(const gg_arena& a, int f, const gg_filter& F,
                                               int chunk, int n_items, const gg_dyn* dy, int cur,
                                               int jbase,
                                               int nout, int anc_mode, int src_buf,
                                               const int* sA, double* lds, int* ilds) {
  const int t = threadIdx.x;
  const int nxt = 1 - cur;
  const int32_t* asrc = a.anc + (int64_t)src_buf * a.P + F.p_off;
  double acc[D + 2];
#pragma unroll
  for (int d = 0; d < D + 2; ++d) acc[d] = 0.0;
  constexpr int U = 4;
  for (int k0 = 0; k0 < nout; k0 += U * GG_BLOCK) {
    int src[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * GG_BLOCK + t;
      ok[u] = k < nout;
      const int j = jbase + (ok[u] ? k : 0);
      src[u] = (anc_mode == 0) ? sA[ok[u] ? k : 0] : (anc_mode == 2 ? j : asrc[j]);
    }
    float v[U][D];
    if (anc_mode == 3) {
      const float4* ao = reinterpret_cast<const float4*>(a.aos) + 2 * F.p_off;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f), y = x;
        if (ok[u]) {
          const int64_t o = 2 * gg_aos_slot(src[u]);
          x = ao[o];
          y = ao[o + 1];
        }
        const float r[8] = {x.x, x.y, x.z, x.w, y.x, y.y, y.z, y.w};
#pragma unroll
        for (int d = 0; d < D; ++d) v[u][d] = r[d];
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int d = 0; d < D; ++d) v[u][d] = ok[u] ? gg_comp(a, cur, d)[F.p_off + src[u]] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
      const int j = jbase + k0 + u * GG_BLOCK + t;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        gg_comp(a, nxt, d)[F.p_off + j] = v[u][d];
        acc[2 + d] += (double)v[u][d];
      }
      a.anc[F.p_off + j] = src[u];
    }
  }
  const int vm = dy->vel_mode;
  if (vm >= 1) {
    const int nv = (vm == 2) ? 9 : 3;
    for (int k = t; k < nout; k += GG_BLOCK) {
      const int j = jbase + k;
      const int sj = a.anc[F.p_off + j];  // this thread wrote it above
      for (int k2 = 0; k2 < nv; ++k2)
        gg_comp(a, nxt, D + k2)[F.p_off + j] = gg_comp(a, cur, D + k2)[F.p_off + sj];
    }
  }
  double tot[D + 2];
  gg_block_totals<D>(acc, lds, tot);
  gg_gather_finish<D>(a, f, F, chunk, n_items, tot, lds, ilds);
}

// Gather by ancestor indices (anc_mode 1: anc[src_buf], 2: identity, 3:
// anc[src_buf] with the packed states); work item = GG_QCHUNK outputs.
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_gather(gg_arena a, gg_step_args s, int force,
                                                        int anc_mode, int src_buf) {
  constexpr int D = 4 + NISB;
  __shared__ int ilds[8];
  __shared__ double lds[8 + GG_BLOCK];
  const int chunk = blockIdx.x;
  const int f = a.chunk_filter[chunk];
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  if (!force && !dy->flag) return;
  const int cl = chunk - F.chunk_off;
  const int jbase = cl * GG_QCHUNK;
  const int nout = min(max(F.n - jbase, 0), GG_QCHUNK);
  gg_gather_copy<D>(a, f, F, chunk, F.n_chunks, dy, dy->cur, jbase, nout, anc_mode, src_buf,
                    nullptr, lds, ilds);
}

// ---------------------------------------------------------------------------
// Systematic resampling in one launch.  Work item = one input tile and at
// most GG_QCHUNK of its outputs (a tile with more outputs gets several items),
// so every item handles the same inputs whatever the weights and at most
// GG_QCHUNK outputs.
//
// with_decide (the device-driven step): every workgroup combines its filter's
// tile partials itself (gg_combine, identical in every workgroup) and takes
// the decision; the filter's first item commits the step frame and writes the
// decision and estimate to the record (through the device-coherent level,
// since the finalize below may rewrite the record in this launch).  Otherwise
// the decision, W, u0 and the tile offsets P_b and scales g_b come from the
// record and a.Pb / a.g (gg_k_decide).
//
// The CDF of particle i in tile b is cdf_i = (P_b + g_b L_i) / W, L the
// fixed-order inclusive scan of e_i = exp32(lw_i - m_b) in the tile, and its
// output end E_i = #{j : t_j <= cdf_i} clamped to the tile's end count
// K_{b+1} = #{j : t_j <= P_{b+1} / W} (P_nt = W); the tile's last particle
// takes K_{b+1} and the filter's last particle N (the hi = n-1 clamp of the
// reference's binary search, pf_device.cu:1200-1215).  E is then
// non-decreasing across the whole filter and tile b's outputs are exactly
// [K_b, K_{b+1}).  Every item computes K_b of all tiles (thread t owns tiles
// [t*c, t*c+c) of the chunked scan), the tiles' item counts
// ceil((K_{b+1} - K_b) / GG_QCHUNK) and their prefix sum, which gives its
// tile b and output range [j0, j1); then the ends of tile b alone.  The
// ancestor of output j is #{i : E_i <= j} = max{i : E_{i-1} <= j}: the input
// whose run of copies [E_{i-1}, E_i) meets [j0, j1) marks the output where
// that run starts (clamped to j0; no two inputs mark one output), and a block
// prefix max over the outputs fills the runs.  Then the gather and the
// finalize (items without outputs still count themselves in).
// Dynamic LDS: max(8 * max(n_tiles) + 4 * (max(n_tiles) + 1), 4 * GG_QCHUNK)
// bytes: the tile sums and end counts, then (once the item knows its tile)
// the ancestors of its outputs in the same memory.  Six workgroups per CU
// (at most 80 VGPRs) keep the n_tiles + n / GG_QCHUNK items of a 1M-particle
// filter in one round.
// ---------------------------------------------------------------------------
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) __attribute__((amdgpu_waves_per_eu(6)))
void gg_k_sysres(gg_arena a, gg_step_args s, int force, int with_decide) {
#pragma clang fp contract(off)
  constexpr int D = 4 + NISB;
  constexpr int RUN = GG_QCHUNK / GG_BLOCK;  // outputs per thread in the prefix max
  extern __shared__ __align__(16) char smem[];
  __shared__ int sE[GG_TILE];
  __shared__ int ilds[8];
  __shared__ int sflag;
  __shared__ double lds[8 + GG_BLOCK];
  __shared__ double sEx[GG_BLOCK + 1];  // exclusive prefix of each thread's tile chunk
  __shared__ double sTile[3];           // P_b, g_b, m_b of the item's tile
  __shared__ int sM[GG_BLOCK];          // per-thread prefix maxima of the ancestor markers
  __shared__ gg_dyn sd;
  const int chunk = blockIdx.x;
  const int f = a.chunk_filter[chunk];
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int cl = chunk - F.chunk_off;
  const int n = F.n;
  const int nt = F.n_tiles;
  double* sT = reinterpret_cast<double*>(smem);  // tile sums T_b, then CDF offsets P_b
  int* sA = reinterpret_cast<int*>(smem);        // ancestors of the outputs, once sT is read
  const int t = threadIdx.x;
  const int cur = dy->cur;
  // thread t owns tiles [b0, b1) of the chunked scan
  const int c = (nt + GG_BLOCK - 1) / GG_BLOCK;
  const int b0 = min(t * c, nt);
  const int b1 = min(b0 + c, nt);
  double W, u0, K = 0.0;
  bool lwz;
  if (with_decide) {
    const bool first = (cl == 0);
    if (first) gg_dyn_copy(&sd, dy);
    // the guide velocity as uniform loads (scalar registers across the combine)
    double vge[3] = {0.0, 0.0, 0.0};
    if (first) {
      const double* vg = gg_guide_ptr(a, s, f, dy->epoch);
      vge[0] = vg[0];
      vge[1] = vg[1];
      vge[2] = vg[2];
    }
    gg_comb r;
    if (first)
      gg_combine<D, true, false>(a, F, lds, sT, r);
    else
      gg_combine<D, false, false>(a, F, lds, sT, r);
    W = r.W;
    K = r.K;
    sEx[t] = r.excl;
    if (t == 0) sEx[GG_BLOCK] = r.W;
    u0 = gg_u0(s, F.seed, n, dy->resample_count);
    lwz = false;  // the step's weight launch wrote fresh log-weights
    if (t == 0) {
      const double ess = (r.W2 > 0.0) ? (r.W * r.W) / r.W2 : 0.0;
      sflag = s.auto_flag ? ((ess < s.ess_frac * (double)n) ? 1 : 0) : dy->flag;
    }
    __syncthreads();
    if (first) {
      if (t == 0) {
        gg_frame fr;
        gg_step_frame<D>(&sd, s, vge, fr);
        gg_commit_frame<D>(s, &sd, fr);
        gg_store_decision<D>(s, F, r, u0, &sd);
        gg_make_estimate<D>(&sd);
        sd.lw_zero = 0;
        sd.did_resample = 0;
        if (s.slot_from_device || s.step_from_device) sd.epoch = sd.epoch + 1;
      }
      __syncthreads();
      gg_dyn_store_coherent(a.dyn + f, &sd);
    }
    if (!force && !sflag) return;
  } else {
    if (!force && !dy->flag) return;
    W = dy->W;
    u0 = (s.u0_mode == 2) ? s.u0_value : dy->u0;
    lwz = dy->lw_zero != 0;
  }
  // CDF offsets P_b of the filter's tiles, in sT in place of the combine's
  // tile sums T_b: thread t turns its chunk into P_b = excl_t + (T_{tc} + ...
  // + T_{b-1}), the operations and order of gg_combine's scan (or from a.Pb)
  if (with_decide) {
    double run = 0.0;
    for (int b = b0; b < b1; ++b) {
      const double tsum = sT[b];
      sT[b] = sEx[t] + run;
      run = run + tsum;
    }
  } else {
    for (int b = t; b < nt; b += GG_BLOCK) sT[b] = a.Pb[F.tile_off + b];
  }
  __syncthreads();
  // K_b: outputs before tile b (K_0 = 0, K_nt = n)
  auto Kb = [&](int b) -> int {
    if (b <= 0) return 0;
    if (b >= nt) return n;
    return gg_kcount(sT[b] / W, u0, n);
  };
  // Work items: the intervals between consecutive elements of the merged
  // sequence of interior tile ends A_k = K_{k+1} (k < nt - 1) and interior
  // grid points G_k = (k + 1) QCHUNK (k < m - 1, m = ceil(n / QCHUNK)), a
  // tile end first on ties.  Item cl is the interval that follows merged
  // element cl - 1: it lies in one tile, the number of tile ends among those
  // cl elements, and holds at most QCHUNK outputs.  nt + m - 1 items cover
  // [0, n); the filter's last item (cl = nt + m - 1) and equal neighbours
  // give empty intervals.  The count of tile ends is the largest i in
  // [max(0, cl - (m - 1)), min(cl, nt - 1)] with i = 0 or A_{i-1} <= G_{cl-i}
  // (true, then false, as i grows), found by two rounds of 256 probes.
  const int nA = nt - 1;
  const int nG = (n + GG_QCHUNK - 1) / GG_QCHUNK - 1;
  auto grid = [&](int k) -> int { return (k >= nG) ? 0x7fffffff : (k + 1) * GG_QCHUNK; };
  int tb = -1, j0 = 0, j1 = 0, kb = 0, kn = 0;
  if (n > 0 && cl <= nA + nG) {  // uniform over the workgroup
    const int imax = min(cl, nA);
    int lo = max(0, cl - nG);
    int span = imax - lo + 1;
    while (span > 1) {  // uniform; one round per factor of 256 in the span
      const int stride = (span + GG_BLOCK - 1) / GG_BLOCK;
      const int i = lo + t * stride;
      int cand = -1;
      if (i < lo + span && (i == 0 || Kb(i) <= grid(cl - i))) cand = i;
      const int best = gg_block_max256_i(cand, ilds);  // >= lo: the probe at lo holds
      span = min(stride, imax - best + 1);
      lo = best;
    }
    if (t < 2) ilds[t] = Kb(lo + t);  // K_b and K_{b+1} of the item's tile
    __syncthreads();
    tb = lo;
    kb = ilds[0];
    kn = ilds[1];
    j0 = max(kb, (cl - tb) * GG_QCHUNK);
    j1 = min(kn, grid(cl - tb));
    __syncthreads();  // ilds is reused below
  }
  if (tb < 0 || j0 >= j1) {  // no outputs for this item: count it in with zero sums
    gg_gather_copy<D>(a, f, F, chunk, F.n_chunks, dy, cur, 0, 0, 0, 0, sA, lds, ilds);
    return;
  }
  const int nout = j1 - j0;
  // the tile's offset, scale and maximum, beside its log-weights
  const int i0 = tb * GG_TILE + t * GG_ITEMS;
  float4 lv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (!lwz && i0 < n) lv = *reinterpret_cast<const float4*>(a.lw + F.p_off + i0);
  if (t == 0) {
    const double* pb = a.part + (int64_t)(F.tile_off + tb) * GG_NPART;
    sTile[0] = sT[tb];
    sTile[1] = with_decide ? gg_tile_shift(pb[GG_NPART - 1], pb[GG_NPART - 2] - K)
                           : a.g[F.tile_off + tb];
    sTile[2] = pb[0];
  }
  __syncthreads();
  // ends of the tile's particles: in-tile scan of e, counts clamped to K_{b+1}
  int e4[GG_ITEMS];
  {
    const float lw[GG_ITEMS] = {lv.x, lv.y, lv.z, lv.w};
    const float mbf = (float)sTile[2];
    double l[GG_ITEMS];
    double run = 0.0;
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) {
      const bool valid = (i0 + q) < n;
      const double e =
          (valid && mbf > -__builtin_huge_valf()) ? (double)gg_exp32(lw[q] - mbf) : 0.0;
      run = (q == 0) ? e : run + e;
      l[q] = run;
    }
    double excl, total;
    gg_block_scan256(l[GG_ITEMS - 1], lds, excl, total);
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) {
      const int i = i0 + q;
      int e;
      if (i >= n) {
        e = 0x3fffffff;
      } else if (i >= n - 1) {
        e = n;
      } else if (t * GG_ITEMS + q == GG_TILE - 1) {
        e = kn;
      } else {
        const double L = excl + l[q];
        e = min(gg_kcount((sTile[0] + sTile[1] * L) / W, u0, n), kn);
      }
      e4[q] = e;
    }
  }
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) sE[t * GG_ITEMS + q] = e4[q];
  for (int k = t; k < GG_QCHUNK; k += GG_BLOCK) sA[k] = -1;
  __syncthreads();
  // markers: the input whose run [E_{i-1}, E_i) meets [j0, j1) marks its start
  {
    int prev = (t == 0) ? kb : sE[t * GG_ITEMS - 1];
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) {
      const int i = i0 + q;
      const int e = e4[q];
      if (i < n && e > prev && prev < j1 && e > j0) sA[prev > j0 ? prev - j0 : 0] = i;
      prev = e;
    }
  }
  __syncthreads();
  {
    int v[RUN];
    int run = -1;
#pragma unroll
    for (int q = 0; q < RUN; ++q) {
      run = max(run, sA[t * RUN + q]);
      v[q] = run;
    }
    const int incl = gg_block_prefix_max256_i(run, ilds);
    sM[t] = incl;
    __syncthreads();
    const int excl = (t == 0) ? -1 : sM[t - 1];
#pragma unroll
    for (int q = 0; q < RUN; ++q) sA[t * RUN + q] = max(excl, v[q]);
  }
  __syncthreads();
  gg_gather_copy<D>(a, f, F, chunk, F.n_chunks, dy, cur, j0, nout, 0, 0, sA, lds, ilds);
}

// ---------------------------------------------------------------------------
// Megopolis with the reference's semantics: slot weights lw[j] vs lw[i], a
// fresh proposal per particle and iteration, state copied through ancestors.
// Iteration `it` reads anc[it & 1] (identity for it == 0), writes anc[(it+1) & 1].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(GG_BLOCK) void gg_k_mego_ref(gg_arena a, int it, int force,
                                                          uint64_t seed_ovr, int use_ovr) {
#pragma clang fp contract(off)
  const int tile = blockIdx.x;
  const int f = a.tile_filter[tile];
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  if (!force && !dy->flag) return;
  const int lt = tile - F.tile_off;
  const int n = F.n;
  const uint64_t seed = use_ovr ? seed_ovr : F.seed;
  const bool lwz = dy->lw_zero != 0;
  const float* lw = a.lw + F.p_off;
  const int32_t* ar = a.anc + (int64_t)(it & 1) * a.P + F.p_off;
  int32_t* aw = a.anc + (int64_t)((it + 1) & 1) * a.P + F.p_off;
  const int i0 = lt * GG_TILE + threadIdx.x * GG_ITEMS;
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
    const int i = i0 + q;
    if (i >= n) continue;
    const gg_u32x4 r = gg_draw(seed, (uint32_t)i, (uint32_t)it, GG_PURPOSE_MEGO_REF, 0u);
    const double u1 = gg_u01d(r.v[0], r.v[1]);
    const double u2 = gg_u01d(r.v[2], r.v[3]);
    const int off = (int)(u1 * (double)(n - 1)) + 1;  // 1 <= off <= n - 1
    int j = i + off;
    if (j >= n) j -= n;
    const double lwi = lwz ? 0.0 : (double)lw[i];
    const double lwj = lwz ? 0.0 : (double)lw[j];
    const double alpha = fmin(1.0, gg_exp64(lwj - lwi));
    int src;
    if (u2 < alpha)
      src = (it == 0) ? j : ar[j];
    else
      src = (it == 0) ? i : ar[i];
    aw[i] = src;
  }
}

// ---------------------------------------------------------------------------
// Megopolis (Chesser et al. 2021): Metropolis chains on the original weights
// with one proposal offset per 256-particle segment and iteration.
// gg_k_mego_w first writes w_i = exp32(lw_i - M) (M: the decision's maximum;
// 1 for zeroed weights); a chain at k then moves to proposal j when
// u * w_k < w_j, u in (0, 1), i.e. with probability min(1, w_j / w_k) (a chain
// on a zero weight leaves for any positive one).  Wavefront w of a tile owns
// segment 4 * tile + w; lane l holds its particles l + 64 q, so every load and
// store of a wavefront is 64 consecutive values.  Writes anc[1].
// ---------------------------------------------------------------------------
// Megopolis weights w = exp32(lw - M), and the particles' D state components
// packed 8 floats per particle (aos): a chain's final ancestor is any particle,
// so the gather that follows reads one 32-byte record per output instead of D
// scattered words.  Inside a Megopolis step (a.aos_step) the likelihood launch
// has written the records already, and when it wrote them only (aos_only), a
// step that does not resample copies them into the current state buffer here.
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_mego_w(gg_arena a, int force) {
  constexpr int D = 4 + NISB;
  const int tile = blockIdx.x;
  const int f = a.tile_filter[tile];
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int lt = tile - F.tile_off;
  const int i0 = lt * GG_TILE + threadIdx.x * GG_ITEMS;
  if (!force && !dy->flag) {
    if (a.aos_step && dy->aos_only) {  // no resampling: only the records hold the state
      const float4* ao = reinterpret_cast<const float4*>(a.aos);
      float p[GG_ITEMS][D];
#pragma unroll
      for (int q = 0; q < GG_ITEMS; ++q) {
        const int64_t o = 2 * gg_aos_slot(F.p_off + i0 + q);
        const float4 x = ao[o], y = ao[o + 1];
        const float r[8] = {x.x, x.y, x.z, x.w, y.x, y.y, y.z, y.w};
#pragma unroll
        for (int d = 0; d < D; ++d) p[q][d] = r[d];
      }
      gg_store4<D>(a, dy->cur, F.p_off + i0, p);
    }
    return;
  }
  const bool lwz = dy->lw_zero != 0;
  const float M = (float)dy->M;
  const float4 lv = *reinterpret_cast<const float4*>(a.lw + F.p_off + i0);
  const float l4[GG_ITEMS] = {lv.x, lv.y, lv.z, lv.w};
  float w[GG_ITEMS];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q)
    w[q] = (i0 + q >= F.n) ? 0.0f : (lwz ? 1.0f : gg_exp32(l4[q] - M));
  *reinterpret_cast<float4*>(a.mw + F.p_off + i0) = make_float4(w[0], w[1], w[2], w[3]);
  if (a.aos_step) return;  // the step's likelihood launch wrote the records
  float p[GG_ITEMS][D];
  gg_load4<D>(a, dy->cur, F.p_off + i0, p);
  gg_store_aos4<D>(a, F.p_off + i0, p);
}

// DEPTH: iterations whose proposal weights are loaded before the first of
// them is decided (the proposals depend only on the particle and the
// iteration's offset, not on the chain's state), so DEPTH x 4 loads per lane
// are in flight instead of 4.  Results do not depend on DEPTH.
// The chains end in the resampling itself: each chain's particle receives its
// ancestor's state from the packed copy (aos, gg_k_mego_w), the ancestors go to
// anc[0], and the workgroup's state sums enter the gather's finalize
// (gg_gather_finish, one item per tile), so no gather launch follows.
static_assert(GG_QCHUNK == 2 * GG_TILE, "one gather item per tile: equal item counts");
template <int NISB, int DEPTH>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_mego(gg_arena a, int B, int force) {
#pragma clang fp contract(off)
  constexpr int D = 4 + NISB;
  __shared__ int ilds[8];
  __shared__ double lds[8 + GG_BLOCK];
  const int tile = blockIdx.x;
  const int f = a.tile_filter[tile];
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  if (!force && !dy->flag) return;
  const int lt = tile - F.tile_off;
  const int n = F.n;
  const uint32_t call = (uint32_t)dy->resample_count;
  const float* w = a.mw + F.p_off;
  const int lane = threadIdx.x & 63;
  const uint32_t seg = (uint32_t)(lt * (GG_TILE / 256) + (threadIdx.x >> 6));
  const int ib = (int)seg * 256 + lane;  // particle of item q: ib + 64 q
  // a chain past the filter's end holds wk = +inf: u wk = +inf is never below
  // a proposal's weight (whatever its slot holds), so it never moves, and its
  // proposal loads need no mask (j = (i + O) mod n stays inside the filter's
  // padded slots)
  int k[GG_ITEMS];
  float wk[GG_ITEMS];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
    const int i = ib + 64 * q;
    k[q] = i;
    wk[q] = (i < n) ? w[i] : __builtin_huge_valf();
  }
  // acceptance uniforms: one xoshiro128+ stream per chain, seeded by the
  // Philox block (seed, particle, call, GG_PURPOSE_MEGO_U, 0); iteration b
  // takes the stream's output b (also for chains past the filter's end)
  gg_xo xs[GG_ITEMS];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q)
    xs[q] = gg_xo_seed(F.seed, (uint32_t)(ib + 64 * q), call, GG_PURPOSE_MEGO_U, 0u);
  for (int b0 = 0; b0 < B; b0 += 64) {
    uint32_t myO = 0;
    if (b0 + lane < B) {
      const gg_u32x4 r = gg_draw(F.seed, seg, call, GG_PURPOSE_MEGO, (uint32_t)(b0 + lane));
      myO = (uint32_t)(((uint64_t)r.v[0] * (uint64_t)(uint32_t)n) >> 32);
    }
    const int bend = min(B - b0, 64);
    for (int c0 = 0; c0 < bend; c0 += DEPTH) {
      int jq[DEPTH][GG_ITEMS];
      float wq[DEPTH][GG_ITEMS];
#pragma unroll
      for (int cc = 0; cc < DEPTH; ++cc) {
        // past bend (uniform) the slot repeats iteration c0 and is not used
        const int O = __builtin_amdgcn_readlane((int)myO, (c0 + cc < bend) ? c0 + cc : c0);
#pragma unroll
        for (int q = 0; q < GG_ITEMS; ++q) {
          // (i + O) mod n for a chain of the filter (i, O < n): the smaller of
          // i + O and i + O - n as unsigned; past the end, i + O - n lies
          // below the filter's padded slot count
          const uint32_t j0 = (uint32_t)(ib + 64 * q + O);
          const int j = (int)min(j0, j0 - (uint32_t)n);
          jq[cc][q] = j;
          wq[cc][q] = w[j];
        }
      }
#pragma unroll
      for (int cc = 0; cc < DEPTH; ++cc) {
        if (c0 + cc >= bend) break;
#pragma unroll
        for (int q = 0; q < GG_ITEMS; ++q) {
          const float u = gg_u01f(gg_xo_next(xs[q]));
          const bool acc = u * wk[q] < wq[cc][q];
          k[q] = acc ? jq[cc][q] : k[q];
          wk[q] = acc ? wq[cc][q] : wk[q];
        }
      }
    }
  }
  // state of each chain's ancestor into the chain's particle (next buffer)
  const int cur = dy->cur;
  const int nxt = 1 - cur;
  const float4* ao = reinterpret_cast<const float4*>(a.aos) + 2 * F.p_off;
  double acc[D + 2];
#pragma unroll
  for (int d = 0; d < D + 2; ++d) acc[d] = 0.0;
  float4 x[GG_ITEMS], y[GG_ITEMS];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
    x[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    y[q] = x[q];
    if (ib + 64 * q < n) {
      const int64_t o = 2 * gg_aos_slot(k[q]);
      x[q] = ao[o];
      y[q] = ao[o + 1];
    }
  }
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
    const int i = ib + 64 * q;
    if (i >= n) continue;
    const float r[8] = {x[q].x, x[q].y, x[q].z, x[q].w, y[q].x, y[q].y, y[q].z, y[q].w};
#pragma unroll
    for (int d = 0; d < D; ++d) {
      gg_comp(a, nxt, d)[F.p_off + i] = r[d];
      acc[2 + d] += (double)r[d];
    }
    a.anc[F.p_off + i] = k[q];
  }
  const int vm = dy->vel_mode;
  if (vm >= 1) {
    const int nv = (vm == 2) ? 9 : 3;
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) {
      const int i = ib + 64 * q;
      if (i >= n) continue;
      for (int k2 = 0; k2 < nv; ++k2)
        gg_comp(a, nxt, D + k2)[F.p_off + i] = gg_comp(a, cur, D + k2)[F.p_off + k[q]];
    }
  }
  double tot[D + 2];
  gg_block_totals<D>(acc, lds, tot);
  gg_gather_finish<D>(a, f, F, F.chunk_off + lt, F.n_tiles, tot, lds, ilds);
}

// ---------------------------------------------------------------------------
// Weighted RMS distance of the positions to a centre: per-tile partials
// (m_b, S_b, sum e*d^2) and a one-workgroup-per-filter combine that writes
// scratch_f[f*16 + 0].
// ---------------------------------------------------------------------------
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_spread(gg_arena a, double cx, double cy,
                                                        double cz) {
#pragma clang fp contract(off)
  __shared__ double lds[8 + GG_BLOCK];
  const int tile = blockIdx.x;
  const int f = a.tile_filter[tile];
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int lt = tile - F.tile_off;
  const int cur = dy->cur;
  double cl[3];
  double ce[3] = {cx - dy->anchor[0], cy - dy->anchor[1], cz - dy->anchor[2]};
  gg_rot_local(dy->R, ce, cl);
  const int i0 = lt * GG_TILE + threadIdx.x * GG_ITEMS;
  const int64_t base = F.p_off + i0;
  float p[GG_ITEMS][3];
  gg_load4<3>(a, cur, base, p);
  float lw[GG_ITEMS] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (!dy->lw_zero) {
    const float4 v = *reinterpret_cast<const float4*>(a.lw + base);
    lw[0] = v.x;
    lw[1] = v.y;
    lw[2] = v.z;
    lw[3] = v.w;
  }
  bool valid[GG_ITEMS];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) valid[q] = (i0 + q) < F.n;
  double m = -__builtin_huge_val();
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q)
    if (valid[q]) m = fmax(m, (double)lw[q]);
  const double mb = gg_block_max256(m, lds);
  double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
    if (!valid[q]) continue;
    const double e = (mb > -__builtin_huge_val()) ? gg_exp64((double)lw[q] - mb) : 0.0;
    const double dx = (double)p[q][0] - cl[0];
    const double dyy = (double)p[q][1] - cl[1];
    const double dz = (double)p[q][2] - cl[2];
    v[0] = v[0] + e;
    v[2] = v[2] + e * (dx * dx + dyy * dyy + dz * dz);
  }
  double tot[3];
  gg_block_totals<1>(v, lds, tot);
  if (threadIdx.x == 0) {
    double* pb = a.part + (int64_t)tile * GG_NPART;
    pb[0] = mb;
    pb[1] = tot[0];
    pb[2] = tot[2];
  }
}

__global__ __launch_bounds__(GG_BLOCK) void gg_k_spread_combine(gg_arena a) {
#pragma clang fp contract(off)
  __shared__ double lds[8 + GG_BLOCK];
  const int f = blockIdx.x;
  const gg_filter F = a.filt[f];
  const int nt = F.n_tiles;
  const double* part = a.part + (int64_t)F.tile_off * GG_NPART;
  double mm = -__builtin_huge_val();
  for (int b = threadIdx.x; b < nt; b += GG_BLOCK) mm = fmax(mm, part[(int64_t)b * GG_NPART]);
  const double M = gg_block_max256(mm, lds);
  const int c = (nt + GG_BLOCK - 1) / GG_BLOCK;
  const int b0 = min((int)threadIdx.x * c, nt);
  const int b1 = min(b0 + c, nt);
  double w[3] = {0.0, 0.0, 0.0};
  for (int b = b0; b < b1; ++b) {
    const double* pb = part + (int64_t)b * GG_NPART;
    const double gb = (pb[0] > -__builtin_huge_val()) ? gg_exp64(pb[0] - M) : 0.0;
    w[0] = w[0] + pb[1] * gb;
    w[2] = w[2] + pb[2] * gb;
  }
  double tot[3];
  gg_block_totals<1>(w, lds, tot);
  if (threadIdx.x == 0) {
    double r = 0.0;
    if (tot[0] > 0.0 && isfinite(tot[2])) r = sqrt(fmax(tot[2] / tot[0], 0.0));
    a.scratch_f[(int64_t)f * 16 + 0] = r;
  }
}

// Log-weights as binary64 (0 while they are logically zero), for a history slot.
__global__ __launch_bounds__(GG_BLOCK) void gg_k_export_lw(gg_arena a, int f, double* out) {
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int i = blockIdx.x * GG_BLOCK + threadIdx.x;
  if (i >= F.n) return;
  out[i] = dy->lw_zero ? 0.0 : (double)a.lw[F.p_off + i];
}

// ---------------------------------------------------------------------------
// State export / import (binary64 absolute values, AoS rows).
// layout 0: [x y z cb]; 1: [x y z cb vx vy vz S00..S22] (16); 2: [x y z cb isb..]
// ---------------------------------------------------------------------------
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_export(gg_arena a, int f, int layout,
                                                        double* out) {
  constexpr int D = 4 + NISB;
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int i = blockIdx.x * GG_BLOCK + threadIdx.x;
  if (i >= F.n) return;
  const int cur = dy->cur;
  const int64_t s = F.p_off + i;
  double pl[3] = {(double)gg_comp(a, cur, 0)[s], (double)gg_comp(a, cur, 1)[s],
                  (double)gg_comp(a, cur, 2)[s]};
  double pe[3];
  gg_rot_ecef(dy->R, pl, pe);
  const int ncol = (layout == 0) ? 4 : (layout == 1 ? 16 : D);
  double* o = out + (int64_t)i * ncol;
  o[0] = dy->anchor[0] + pe[0];
  o[1] = dy->anchor[1] + pe[1];
  o[2] = dy->anchor[2] + pe[2];
  o[3] = dy->cb_anchor + (double)gg_comp(a, cur, 3)[s];
  if (layout == 2) {
    for (int k = 4; k < D; ++k) o[k] = dy->isb_anchor[k - 4] + (double)gg_comp(a, cur, k)[s];
  } else if (layout == 1) {
    double vl[3], ve[3];
    if (dy->vel_mode >= 1) {
      for (int k = 0; k < 3; ++k) vl[k] = (double)gg_comp(a, cur, D + k)[s];
    } else {
      for (int k = 0; k < 3; ++k) vl[k] = dy->vel_u[k];
    }
    gg_rot_ecef(dy->R, vl, ve);
    o[4] = ve[0];
    o[5] = ve[1];
    o[6] = ve[2];
    double c6[6];
    if (dy->vel_mode == 2) {
      for (int k = 0; k < 6; ++k) c6[k] = (double)gg_comp(a, cur, D + 3 + k)[s];
    } else {
      for (int k = 0; k < 6; ++k) c6[k] = dy->vcov_u[k];
    }
    const double Sl[9] = {c6[0], c6[1], c6[2], c6[1], c6[3], c6[4], c6[2], c6[4], c6[5]};
    // S_ecef = R^T S_local R
    const double* R = dy->R;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
        for (int p = 0; p < 3; ++p)
          for (int q = 0; q < 3; ++q) acc += R[p * 3 + r] * Sl[p * 3 + q] * R[q * 3 + c];
        o[7 + r * 3 + c] = acc;
      }
  }
}

// Import 16-column reference rows; the host has already set the anchor frame
// and vel_mode.
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_import(gg_arena a, int f, const double* in) {
  constexpr int D = 4 + NISB;
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int i = blockIdx.x * GG_BLOCK + threadIdx.x;
  if (i >= F.n) return;
  const int cur = dy->cur;
  const int64_t s = F.p_off + i;
  const double* r = in + (int64_t)i * 16;
  double de[3] = {r[0] - dy->anchor[0], r[1] - dy->anchor[1], r[2] - dy->anchor[2]};
  double dl[3];
  gg_rot_local(dy->R, de, dl);
  for (int k = 0; k < 3; ++k) gg_comp(a, cur, k)[s] = (float)dl[k];
  gg_comp(a, cur, 3)[s] = (float)(r[3] - dy->cb_anchor);
  for (int k = 4; k < D; ++k) gg_comp(a, cur, k)[s] = 0.0f;
  if (dy->vel_mode >= 1) {
    double vl[3];
    gg_rot_local(dy->R, r + 4, vl);
    for (int k = 0; k < 3; ++k) gg_comp(a, cur, D + k)[s] = (float)vl[k];
  }
  if (dy->vel_mode == 2) {
    // S_local = R S_ecef R^T
    const double* R = dy->R;
    double Sl[9];
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) {
        double acc = 0.0;
        for (int u = 0; u < 3; ++u)
          for (int v = 0; v < 3; ++v) acc += R[p * 3 + u] * r[7 + u * 3 + v] * R[q * 3 + v];
        Sl[p * 3 + q] = acc;
      }
    const double c6[6] = {Sl[0], Sl[1], Sl[2], Sl[4], Sl[5], Sl[8]};
    for (int k = 0; k < 6; ++k) gg_comp(a, cur, D + 3 + k)[s] = (float)c6[k];
  }
}

// Positions and clock biases as four binary64 arrays (structure of arrays),
// the layout of the legacy host-array functions.  Export: x, y, z, cb
// absolute; import: offsets from the record's anchor (ISB and velocity
// components untouched, the host resets the velocity mode).
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_export_soa(gg_arena a, int f, double* px,
                                                            double* py, double* pz, double* pcb) {
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int i = blockIdx.x * GG_BLOCK + threadIdx.x;
  if (i >= F.n) return;
  const int cur = dy->cur;
  const int64_t s = F.p_off + i;
  const double pl[3] = {(double)gg_comp(a, cur, 0)[s], (double)gg_comp(a, cur, 1)[s],
                        (double)gg_comp(a, cur, 2)[s]};
  double pe[3];
  gg_rot_ecef(dy->R, pl, pe);
  px[i] = dy->anchor[0] + pe[0];
  py[i] = dy->anchor[1] + pe[1];
  pz[i] = dy->anchor[2] + pe[2];
  pcb[i] = dy->cb_anchor + (double)gg_comp(a, cur, 3)[s];
}

template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_import_soa(gg_arena a, int f, const double* px,
                                                            const double* py, const double* pz,
                                                            const double* pcb) {
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int i = blockIdx.x * GG_BLOCK + threadIdx.x;
  if (i >= F.n) return;
  const int cur = dy->cur;
  const int64_t s = F.p_off + i;
  const double de[3] = {px[i] - dy->anchor[0], py[i] - dy->anchor[1], pz[i] - dy->anchor[2]};
  double dl[3];
  gg_rot_local(dy->R, de, dl);
  for (int k = 0; k < 3; ++k) gg_comp(a, cur, k)[s] = (float)dl[k];
  gg_comp(a, cur, 3)[s] = (float)(pcb[i] - dy->cb_anchor);
}

// O(1) frame edits: shift position (ECEF) and clock bias of every particle.
__global__ void gg_k_frame_shift(gg_dyn* dyn, int f, double dx, double dy_, double dz,
                                 double dcb) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  gg_dyn* d = dyn + f;
  d->anchor[0] += dx;
  d->anchor[1] += dy_;
  d->anchor[2] += dz;
  d->cb_anchor += dcb;
  d->est[0] += dx;
  d->est[1] += dy_;
  d->est[2] += dz;
  d->est[3] += dcb;
}

// Gaussian position-domain likelihood added to the log-weights.
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_position_update(gg_arena a, int f, double rx,
                                                                 double ry, double rz,
                                                                 double inv_s2) {
  const gg_filter F = a.filt[f];
  gg_dyn* dy = a.dyn + f;
  const int i = blockIdx.x * GG_BLOCK + threadIdx.x;
  if (i >= F.n) return;
  const int cur = dy->cur;
  const int64_t s = F.p_off + i;
  double de[3] = {rx - dy->anchor[0], ry - dy->anchor[1], rz - dy->anchor[2]};
  double rl[3];
  gg_rot_local(dy->R, de, rl);
  const double dx = (double)gg_comp(a, cur, 0)[s] - rl[0];
  const double dyy = (double)gg_comp(a, cur, 1)[s] - rl[1];
  const double dz = (double)gg_comp(a, cur, 2)[s] - rl[2];
  const float base = dy->lw_zero ? 0.0f : a.lw[s];
  a.lw[s] = base + (float)(-0.5 * (dx * dx + dyy * dyy + dz * dz) * inv_s2);
}

__global__ void gg_k_clear_lw_zero(gg_dyn* dyn, int f) {
  if (threadIdx.x == 0 && blockIdx.x == 0) dyn[f].lw_zero = 0;
}

// ---------------------------------------------------------------------------
// Double-differenced pseudorange and carrier-phase AFV likelihoods (single
// filter).  KIND 0: DD pseudorange (pf_device.cu:282-368), overwrites lw.
// KIND 1: undifferenced carrier AFV with the clock bias (:439-494), adds.
// KIND 2: DD carrier AFV with per-pair wavelengths (:502-604), adds.
// obs holds 11 binary64 per observation: sat_k xyz, ref xyz, observation
// (m or cycles), base_range_k, base_range_ref, weight, wavelength.
// Each workgroup stages per observation, in the filter's ENU frame, the unit
// vectors and inverse ranges from the anchor to both satellites and the
// binary64 part of the residual; ranges enter per particle only through
// r - rho = q / (1 + sqrt(1 + q / rho)), q = |p|^2 / rho - 2 u.p.  For the
// carrier kinds the binary64 part is reduced to its fractional cycle, which
// leaves x - rint(x) unchanged.  Dynamic LDS: 48 bytes per observation.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float gg_obs_delta(const float4& g, const float* p, float pp) {
  const float ud2 = fmaf(g.z, p[2], fmaf(g.y, p[1], g.x * p[0]));  // -2 u.p
  const float q = fmaf(pp, g.w, ud2);
  return q * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_sqrtf(fmaf(q, g.w, 1.0f)));
}

template <int KIND, int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_weight_obs(gg_arena a, const double* __restrict__ obs,
                                                            int n_obs, float sigma, float thr,
                                                            int huber, float huber_k) {
  constexpr int D = 4 + NISB;
  extern __shared__ __align__(16) char smem[];
  float4* rec = reinterpret_cast<float4*>(smem);  // [n_obs][3]
  const int tile = blockIdx.x;
  const int f = a.tile_filter[tile];
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int lt = tile - F.tile_off;
  const int cur = dy->cur;
  const int i0 = lt * GG_TILE + threadIdx.x * GG_ITEMS;
  const int64_t base = F.p_off + i0;
  float p[GG_ITEMS][D];
  gg_load4<D>(a, cur, base, p);
  for (int k = threadIdx.x; k < n_obs; k += blockDim.x) {
    const double* o = obs + (int64_t)k * 11;
    double dk[3], dr[3], rk, irk, rr, irr;
    gg_sat_local(o, dy, dy->anchor, 0, dk, rk, irk);
    gg_sat_local(o + 3, dy, dy->anchor, 0, dr, rr, irr);
    const double wl = o[10];
    double c;
    if (KIND == 1) {
      c = o[6] - (rk + dy->cb_anchor) * (1.0 / wl);
      c = c - rint(c);
    } else {
      c = o[6] - (((rk - rr) - o[7]) + o[8]) * (KIND == 2 ? 1.0 / wl : 1.0);
      if (KIND == 2) c = c - rint(c);
    }
    rec[3 * k + 0] = make_float4(-2.0f * (float)(dk[0] * irk), -2.0f * (float)(dk[1] * irk),
                                 -2.0f * (float)(dk[2] * irk), (float)irk);
    rec[3 * k + 1] = make_float4(-2.0f * (float)(dr[0] * irr), -2.0f * (float)(dr[1] * irr),
                                 -2.0f * (float)(dr[2] * irr), (float)irr);
    rec[3 * k + 2] = make_float4((float)c, (float)o[9], (float)(1.0 / wl), 0.0f);
  }
  __syncthreads();
  const float isig = 1.0f / sigma;
  const float inv_s2 = isig * isig;
  const bool use_huber = huber != 0 && huber_k > 0.0f;
  float lw[GG_ITEMS];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
    const float pp = fmaf(p[q][2], p[q][2], fmaf(p[q][1], p[q][1], p[q][0] * p[q][0]));
    // residual of observation k: metres (KIND 0) or the AFV fraction of a cycle
    auto resid = [&](int k) {
      const float4 m = rec[3 * k + 2];
      const float dk = gg_obs_delta(rec[3 * k + 0], p[q], pp);
      if (KIND == 1) {
        const float x = m.x - (dk + p[q][3]) * m.z;
        return x - rintf(x);
      }
      const float dd = dk - gg_obs_delta(rec[3 * k + 1], p[q], pp);
      if (KIND == 0) return m.x - dd;
      const float x = m.x - dd * m.z;
      return x - rintf(x);
    };
    bool gate = KIND != 1 && thr > 0.0f;
    if (gate) {
      int kept = 0;
      for (int k = 0; k < n_obs; ++k) kept += (fabsf(resid(k)) <= thr) ? 1 : 0;
      if (kept < (n_obs < 3 ? n_obs : 3)) gate = false;
    }
    float acc = 0.0f;
    for (int k = 0; k < n_obs; ++k) {
      const float r = resid(k);
      if (gate && fabsf(r) > thr) continue;
      const float w = rec[3 * k + 2].y;
      if (KIND != 1 && use_huber) {
        const float z = fabsf(r) * isig;
        acc -= w * ((z < huber_k) ? 0.5f * z * z : huber_k * z - 0.5f * huber_k * huber_k);
      } else {
        acc -= 0.5f * w * r * r * inv_s2;
      }
    }
    lw[q] = acc;
  }
  if (KIND != 0) {
    const bool lwz = dy->lw_zero != 0;
    const float4 old = *reinterpret_cast<const float4*>(a.lw + base);
    const float o4[GG_ITEMS] = {old.x, old.y, old.z, old.w};
#pragma unroll
    for (int q = 0; q < GG_ITEMS; ++q) lw[q] = (lwz ? 0.0f : o4[q]) + lw[q];
  }
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) lw[q] = (i0 + q < F.n) ? lw[q] : 0.0f;
  *reinterpret_cast<float4*>(a.lw + base) = make_float4(lw[0], lw[1], lw[2], lw[3]);
}

// ---------------------------------------------------------------------------
// Ray-traced NLOS weight on the filter's particles: the device-filter
// counterpart of the reference's pf_weight_3d_bvh (weight_3d_bvh.cu:127-220).
// Each particle is taken to ECEF in binary64 (anchor + R^T offset) and one
// ray per satellite walks the tree (gg_bvh.h); the LOS / NLOS mixture is
// evaluated in binary64 and stored as the binary32 log-weight (overwritten,
// as the reference does).  obs: 11 binary64 rows (stage_obs): ECEF position
// [0..2], pseudorange [6], system group [7], weight [9].  The clock bias and,
// for group g >= 1, the inter-system bias come from the particle's state.
// ---------------------------------------------------------------------------
template <int NISB>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_weight_3d_pf(gg_arena a,
                                                              const double* __restrict__ obs,
                                                              int n_obs,
                                                              const gg_bvh_node* __restrict__ nodes,
                                                              const double* __restrict__ tris,
                                                              int n_nodes, gg_rt::MixArgs mx) {
  constexpr int D = 4 + NISB;
  extern __shared__ __align__(16) char smem[];
  double* so = reinterpret_cast<double*>(smem);  // [n_obs][6]: x y z pr w group
  const int tile = blockIdx.x;
  const int f = a.tile_filter[tile];
  const gg_filter F = a.filt[f];
  const gg_dyn* dy = a.dyn + f;
  const int lt = tile - F.tile_off;
  const int i0 = lt * GG_TILE + threadIdx.x * GG_ITEMS;
  const int64_t base = F.p_off + i0;
  float p[GG_ITEMS][D];
  gg_load4<D>(a, dy->cur, base, p);
  for (int k = threadIdx.x; k < n_obs; k += blockDim.x) {
    const double* o = obs + (int64_t)k * 11;
    so[6 * k + 0] = o[0];
    so[6 * k + 1] = o[1];
    so[6 * k + 2] = o[2];
    so[6 * k + 3] = o[6];
    so[6 * k + 4] = o[9];
    so[6 * k + 5] = o[7];
  }
  __syncthreads();
  float lw[GG_ITEMS];
#pragma unroll
  for (int q = 0; q < GG_ITEMS; ++q) {
    lw[q] = 0.0f;
    if (i0 + q >= F.n) continue;
    const double off[3] = {(double)p[q][0], (double)p[q][1], (double)p[q][2]};
    double x[3];
    gg_rot_ecef(dy->R, off, x);
#pragma unroll
    for (int j = 0; j < 3; ++j) x[j] += dy->anchor[j];
    const double cb = dy->cb_anchor + (double)p[q][3];
    double acc = 0.0;
    for (int k = 0; k < n_obs; ++k) {
      const double* sk = so + 6 * k;
      gg_rt::Ray r;
      r.o[0] = x[0];
      r.o[1] = x[1];
      r.o[2] = x[2];
      const double dx = sk[0] - x[0], dy0 = sk[1] - x[1], dz = sk[2] - x[2];
      const double dist = sqrt(dx * dx + dy0 * dy0 + dz * dz);
      const double inv = 1.0 / dist;
      r.d[0] = dx * inv;
      r.d[1] = dy0 * inv;
      r.d[2] = dz * inv;
      r.tmax = dist;
      gg_rt::make_inv(r);
      const bool nl = n_nodes > 0 && gg_rt::gg_bvh_blocked(r, 1e-6, nodes, tris);
      const int g = (int)sk[5];
      double bias = 0.0;
      if (NISB > 0 && g >= 1 && g <= NISB) bias = dy->isb_anchor[g - 1] + (double)p[q][3 + g];
      const double res = sk[3] - (dist + cb + bias);
      const double los = -0.5 * sk[4] * res * res * mx.inv_los2;
      const double rn = res > 0.0 ? res - mx.nlos_bias : res;
      const double nlos = -0.5 * sk[4] * rn * rn * mx.inv_nlos2;
      const int m = nl ? 1 : 0;
      acc += gg_rt::mixture(los, nlos, mx.mode[m], mx.lp_keep[m], mx.lp_nlos[m]);
    }
    lw[q] = (float)acc;
  }
  *reinterpret_cast<float4*>(a.lw + base) = make_float4(lw[0], lw[1], lw[2], lw[3]);
}

// ---------------------------------------------------------------------------
// Doppler velocity updates (single filter), binary64 per particle in the
// filter's ENU frame (range rates and the Kalman update are invariant under
// the rotation).  obs: 8 binary64 per satellite: ECEF position, ECEF
// velocity, Doppler [Hz], weight.  Observed range rate = -doppler * lambda.
// KF = false: pfd_weight_doppler_kernel (pf_device.cu:610-740): clock drift
//   as the weighted mean misfit, Gaussian log-likelihood added to lw, then
//   (gain > 0, >= 4 satellites) the velocity moves toward the particle's
//   4-state Doppler WLS solution, the step capped at max_dv.
// KF = true: pfd_doppler_kf_update_kernel (:742-880): sequential scalar
//   Kalman updates of the particle's velocity Gaussian with the line-of-sight
//   rows and observations centred on their weighted means (clock drift
//   eliminated); adds sum of -(innov^2 / S + log S) / 2 to lw.
// Velocities and covariances are read per particle when the filter holds
// them (vel_mode >= 1 / 2), otherwise from the filter record, and are written
// per particle; gg_k_set_vel_mode raises vel_mode afterwards.
// Dynamic LDS: 64 bytes per satellite.
// ---------------------------------------------------------------------------
__device__ bool gg_solve4(double (&A)[4][5], double (&x)[4]) {
#pragma unroll
  for (int col = 0; col < 4; ++col) {
    int piv = col;
    double best = fabs(A[col][col]);
#pragma unroll
    for (int r = col + 1; r < 4; ++r)
      if (fabs(A[r][col]) > best) {
        best = fabs(A[r][col]);
        piv = r;
      }
    // swap through selects: a run-time row index would put A in scratch
#pragma unroll
    for (int r = col + 1; r < 4; ++r) {
      const bool sw = (r == piv);
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const double t0 = A[col][k], t1 = A[r][k];
        A[col][k] = sw ? t1 : t0;
        A[r][k] = sw ? t0 : t1;
      }
    }
    if (fabs(A[col][col]) < 1e-12) return false;
    for (int r = col + 1; r < 4; ++r) {
      const double fct = A[r][col] / A[col][col];
      for (int k = col; k < 5; ++k) A[r][k] -= fct * A[col][k];
    }
  }
  for (int r = 3; r >= 0; --r) {
    double v = A[r][4];
    for (int c = r + 1; c < 4; ++c) v -= A[r][c] * x[c];
    x[r] = v / A[r][r];
  }
  return isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]) && isfinite(x[3]);
}

template <int NISB, bool KF>
__global__ __launch_bounds__(GG_BLOCK) void gg_k_doppler(gg_arena a, const double* __restrict__ obs,
                                                         int n_sat, double wl, double sigma,
                                                         double gain, double max_dv) {
  constexpr int D = 4 + NISB;
  extern __shared__ __align__(16) char smem[];
  double* sd = reinterpret_cast<double*>(smem);  // [n_sat][8]
  const gg_filter F = a.filt[0];
  const gg_dyn* dy = a.dyn;
  for (int k = threadIdx.x; k < n_sat; k += blockDim.x) {
    const double* o = obs + (int64_t)k * 8;
    const double dx[3] = {o[0] - dy->anchor[0], o[1] - dy->anchor[1], o[2] - dy->anchor[2]};
    gg_rot_local(dy->R, dx, sd + k * 8);
    gg_rot_local(dy->R, o + 3, sd + k * 8 + 3);
    sd[k * 8 + 6] = -o[6] * wl;
    sd[k * 8 + 7] = fmax(o[7], 0.0);
  }
  __syncthreads();
  const int i = blockIdx.x * GG_BLOCK + threadIdx.x;
  if (i >= F.n) return;
  const int cur = dy->cur;
  const int64_t s = F.p_off + i;
  const int vm = dy->vel_mode;
  const double x[3] = {(double)gg_comp(a, cur, 0)[s], (double)gg_comp(a, cur, 1)[s],
                       (double)gg_comp(a, cur, 2)[s]};
  double v[3];
  for (int k = 0; k < 3; ++k) v[k] = (vm >= 1) ? (double)gg_comp(a, cur, D + k)[s] : dy->vel_u[k];
  // line of sight from the particle to satellite k (unit), false below 1 m
  auto los = [&](int k, double (&l)[3]) {
    const double d0 = sd[k * 8 + 0] - x[0], d1 = sd[k * 8 + 1] - x[1], d2 = sd[k * 8 + 2] - x[2];
    const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    if (r < 1.0) return false;
    l[0] = d0 / r;
    l[1] = d1 / r;
    l[2] = d2 / r;
    return true;
  };
  auto radial = [&](int k, const double (&l)[3]) {
    return sd[k * 8 + 3] * l[0] + sd[k * 8 + 4] * l[1] + sd[k * 8 + 5] * l[2];
  };
  double lw_add = 0.0;
  if (!KF) {
    // one pass over the satellites: the clock-drift misfit q_k = obs_k - (sv_k - v).l_k
    // (the residual is q_k - cd with cd the weighted mean of q), and the normal
    // equations of the velocity WLS; sum w (q - cd)^2 = Sq2 - 2 cd Sq + cd^2 W
    double wsum = 0.0, sq = 0.0, sq2 = 0.0;
    double A[4][5];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 5; ++c) A[r][c] = 0.0;
    for (int k = 0; k < n_sat; ++k) {
      double l[3];
      if (!los(k, l)) continue;
      const double w = sd[k * 8 + 7];
      const double rhs = sd[k * 8 + 6] - radial(k, l);
      const double q = rhs + (v[0] * l[0] + v[1] * l[1] + v[2] * l[2]);
      wsum += w;
      sq += w * q;
      sq2 += w * q * q;
      const double h[4] = {-l[0], -l[1], -l[2], 1.0};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        A[r][4] += h[r] * w * rhs;
#pragma unroll
        for (int c = 0; c < 4; ++c) A[r][c] += h[r] * w * h[c];
      }
    }
    const bool any = wsum > 0.0;  // otherwise the particle is left as it is
    const double cd = any ? sq / wsum : 0.0;
    const double inv_s2 = 1.0 / (sigma * sigma);
    if (any) lw_add = -0.5 * fmax(sq2 - 2.0 * cd * sq + cd * cd * wsum, 0.0) * inv_s2;
    const double g = fmin(1.0, fmax(0.0, gain));
    if (any && g > 0.0 && n_sat >= 4) {
      for (int d = 0; d < 4; ++d) A[d][d] += 1e-9;
      double sol[4] = {0.0, 0.0, 0.0, 0.0};
      if (gg_solve4(A, sol)) {
        double dv[3] = {sol[0] - v[0], sol[1] - v[1], sol[2] - v[2]};
        const double nrm = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        if (max_dv > 0.0 && nrm > max_dv) {
          const double sc = max_dv / fmax(nrm, 1e-9);
          for (int k = 0; k < 3; ++k) dv[k] *= sc;
        }
        for (int k = 0; k < 3; ++k) v[k] += g * dv[k];
      }
    }
  } else {
    double P[9];
    {
      double c6[6];
      for (int k = 0; k < 6; ++k)
        c6[k] = (vm >= 2) ? (double)gg_comp(a, cur, D + 3 + k)[s] : dy->vcov_u[k];
      P[0] = c6[0]; P[1] = c6[1]; P[2] = c6[2];
      P[3] = c6[1]; P[4] = c6[3]; P[5] = c6[4];
      P[6] = c6[2]; P[7] = c6[4]; P[8] = c6[5];
    }
    double hm[3] = {0.0, 0.0, 0.0}, ym = 0.0, wsum = 0.0;
    for (int k = 0; k < n_sat; ++k) {
      const double w = sd[k * 8 + 7];
      if (w <= 0.0) continue;
      double l[3];
      if (!los(k, l)) continue;
      hm[0] += w * -l[0];
      hm[1] += w * -l[1];
      hm[2] += w * -l[2];
      ym += w * (sd[k * 8 + 6] - radial(k, l));
      wsum += w;
    }
    const bool any = wsum > 0.0;  // otherwise the particle is left as it is
    if (any) {
      for (int k = 0; k < 3; ++k) hm[k] /= wsum;
      ym /= wsum;
    }
    const double s2 = sigma * sigma;
    for (int k = 0; any && k < n_sat; ++k) {
      const double w = sd[k * 8 + 7];
      if (w <= 0.0) continue;
      double l[3];
      if (!los(k, l)) continue;
      const double h[3] = {-l[0] - hm[0], -l[1] - hm[1], -l[2] - hm[2]};
      const double yc = (sd[k * 8 + 6] - radial(k, l)) - ym;
      const double ph0 = P[0] * h[0] + P[1] * h[1] + P[2] * h[2];
      const double ph1 = P[3] * h[0] + P[4] * h[1] + P[5] * h[2];
      const double ph2 = P[6] * h[0] + P[7] * h[1] + P[8] * h[2];
      const double S = h[0] * ph0 + h[1] * ph1 + h[2] * ph2 + s2 / w;
      if (!(S > 1e-18) || !isfinite(S)) continue;
      const double innov = yc - (h[0] * v[0] + h[1] * v[1] + h[2] * v[2]);
      const double K[3] = {ph0 / S, ph1 / S, ph2 / S};
      for (int q = 0; q < 3; ++q) v[q] += K[q] * innov;
      const double hp0 = h[0] * P[0] + h[1] * P[3] + h[2] * P[6];
      const double hp1 = h[0] * P[1] + h[1] * P[4] + h[2] * P[7];
      const double hp2 = h[0] * P[2] + h[1] * P[5] + h[2] * P[8];
      for (int q = 0; q < 3; ++q) {
        P[q * 3 + 0] -= K[q] * hp0;
        P[q * 3 + 1] -= K[q] * hp1;
        P[q * 3 + 2] -= K[q] * hp2;
      }
      const double p01 = 0.5 * (P[1] + P[3]);
      const double p02 = 0.5 * (P[2] + P[6]);
      const double p12 = 0.5 * (P[5] + P[7]);
      P[1] = P[3] = isfinite(p01) ? p01 : 0.0;
      P[2] = P[6] = isfinite(p02) ? p02 : 0.0;
      P[5] = P[7] = isfinite(p12) ? p12 : 0.0;
      P[0] = (isfinite(P[0]) && P[0] > 1e-12) ? P[0] : 1e-12;
      P[4] = (isfinite(P[4]) && P[4] > 1e-12) ? P[4] : 1e-12;
      P[8] = (isfinite(P[8]) && P[8] > 1e-12) ? P[8] : 1e-12;
      lw_add += -0.5 * (innov * innov / S + log(S));
    }
    const double c6[6] = {P[0], P[1], P[2], P[4], P[5], P[8]};
    for (int k = 0; k < 6; ++k) gg_comp(a, cur, D + 3 + k)[s] = (float)c6[k];
  }
  for (int k = 0; k < 3; ++k) gg_comp(a, cur, D + k)[s] = (float)v[k];
  const float base = dy->lw_zero ? 0.0f : a.lw[s];
  a.lw[s] = base + (float)lw_add;
}

__global__ void gg_k_set_vel_mode(gg_dyn* dyn, int f, int mode) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && dyn[f].vel_mode < mode) dyn[f].vel_mode = mode;
}
