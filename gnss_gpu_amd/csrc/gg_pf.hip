// Host side of the particle-filter C ABI: handle, allocation, launches,
// hipGraph capture.  Kernels live in gg_pf_kernels_impl.h.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gg_common_host.h"
#include "gg_pf_kernels_impl.h"
#include "gnss_gpu_amd.h"

struct gg_pf {
  int32_t n_filters = 0;
  int32_t n_isb = 0;
  int32_t D = 4;
  int32_t ncomp = 4;
  int32_t has_vel = 0;
  int32_t n_slots = 1;
  int32_t sat_rows = 0;
  int64_t P = 0;
  int32_t n_tiles = 0;
  int32_t n_chunks = 0;
  std::vector<gg_filter> filt_h;
  std::vector<int32_t> tile_filter_h, chunk_filter_h;
  // device memory
  float* state = nullptr;
  float* lw = nullptr;
  int32_t* E = nullptr;
  float* aos = nullptr;
  int32_t* anc = nullptr;
  double* part = nullptr;
  double* p4 = nullptr;
  double* g = nullptr;
  double* Pb = nullptr;
  double* csum = nullptr;
  uint32_t* fin = nullptr;
  int64_t* acc = nullptr;
  double* scratch_f = nullptr;
  gg_filter* filt = nullptr;
  gg_dyn* dyn = nullptr;
  int32_t* tile_filter = nullptr;
  int32_t* chunk_filter = nullptr;
  double* satd = nullptr;
  int32_t* sati = nullptr;
  double* vel_guide = nullptr;
  float* satf = nullptr;        // [n_slots][2][rows] C/N0 | LOS flag (learned likelihood)
  const void* mlp_w = nullptr;  // packed network of the attached gg_mlp
  float2* mlp_table = nullptr;  // [table_rows][GG_MLP_CODES] (learned_mode 1)
  int32_t* row_filter = nullptr;
  int table_rows = 0;
  double* io_tmp = nullptr;
  size_t io_tmp_bytes = 0;
  double* obsd = nullptr;       // [obs_cap][11] DD / carrier observations (single filter)
  int obs_cap = 0;
  // pinned host staging
  char* pin = nullptr;
  size_t pin_bytes = 0;
  hipEvent_t pin_done = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // host-side validity of device results (single-filter API)
  bool decide_valid = false;  // M, W, g, Pb consistent with lw
  bool ess_valid = false;
  bool est_valid = false;
  bool initialized = false;
  std::vector<double> init_isb;
  uint64_t seed0 = 0;
  // launch shapes of the step's kernels (gg_pfb_set_tuning / gg_pfb_autotune):
  // tiles per workgroup of the likelihood kernel, Megopolis iterations whose
  // proposal loads are issued together
  int32_t weight_tpw = 1;
  int32_t mego_depth = 1;
  // kernel timing (gg_pfb_profile_steps): while kt_ev holds events, every
  // launch of a step stage takes the next pair, recorded at the kernel's own
  // start and end
  std::vector<hipEvent_t> kt_ev;
  size_t kt_next = 0;
};

namespace {

template <typename T>
int dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  GG_HIP(hipMalloc((void**)p, count * sizeof(T)));
  // the handle's streams do not synchronise with the null stream, so the
  // clear must be complete before any of them touches the buffer
  GG_HIP(hipMemsetAsync(*p, 0, count * sizeof(T), nullptr));
  GG_HIP(hipStreamSynchronize(nullptr));
  return 0;
}

gg_arena arena_of(const gg_pf* h) {
  gg_arena a;
  a.state = h->state;
  a.lw = h->lw;
  a.anc = h->anc;
  a.part = h->part;
  a.p4 = h->p4;
  a.g = h->g;
  a.Pb = h->Pb;
  a.E = h->E;
  a.mw = reinterpret_cast<float*>(h->E);  // Megopolis steps do not use the ends
  a.aos = h->aos;
  a.aos_step = 0;
  a.csum = h->csum;
  a.fin = h->fin;
  a.acc = h->acc;
  a.scratch_f = h->scratch_f;
  a.filt = h->filt;
  a.dyn = h->dyn;
  a.tile_filter = h->tile_filter;
  a.chunk_filter = h->chunk_filter;
  a.satd = h->satd;
  a.satf = h->satf;
  a.mlp = (const gg_mlp_w*)h->mlp_w;
  a.mlp_table = h->mlp_table;
  a.sati = h->sati;
  a.vel_guide = h->vel_guide;
  a.P = h->P;
  a.ncomp = h->ncomp;
  a.n_isb = h->n_isb;
  a.has_vel = h->has_vel;
  a.sat_rows = h->sat_rows;
  a.n_filters = h->n_filters;
  a.n_tiles = h->n_tiles;
  a.n_chunks = h->n_chunks;
  return a;
}

// LDS of the staged satellites: a record of GG_REC float4 and a group index
// per satellite (+ 64 first-layer biases per satellite for the learned
// likelihood)
size_t sat_smem(const gg_pf* h, bool learned = false) {
  int cap = 0;
  for (const auto& f : h->filt_h) cap = std::max(cap, f.sat_cap);
  return (size_t)std::max(cap, 1) * (4 * GG_REC + 1 + (learned ? GG_MLP_H : 0)) * sizeof(float);
}

int ensure_pinned(gg_pf* h, size_t bytes) {
  if (h->pin_bytes >= bytes) {
    GG_HIP(hipEventSynchronize(h->pin_done));
    return 0;
  }
  if (h->pin) {
    GG_HIP(hipEventSynchronize(h->pin_done));
    GG_HIP(hipHostFree(h->pin));
    h->pin = nullptr;
  }
  GG_HIP(hipHostMalloc((void**)&h->pin, bytes, hipHostMallocDefault));
  h->pin_bytes = bytes;
  return 0;
}

int ensure_io(gg_pf* h, size_t bytes) {
  if (h->io_tmp_bytes >= bytes) return 0;
  if (h->io_tmp) {
    GG_HIP(hipStreamSynchronize(h->stream));
    GG_HIP(hipFree(h->io_tmp));
    h->io_tmp = nullptr;
  }
  GG_HIP(hipMalloc((void**)&h->io_tmp, bytes));
  h->io_tmp_bytes = bytes;
  return 0;
}

int read_dyn(gg_pf* h, int f, gg_dyn* out) {
  if (ensure_pinned(h, std::max(h->pin_bytes, sizeof(gg_dyn)))) return -1;
  GG_HIP(hipMemcpyAsync(h->pin, h->dyn + f, sizeof(gg_dyn), hipMemcpyDeviceToHost, h->stream));
  GG_HIP(hipStreamSynchronize(h->stream));
  std::memcpy(out, h->pin, sizeof(gg_dyn));
  return 0;
}

int write_dyn(gg_pf* h, int f, const gg_dyn& d) {
  if (ensure_pinned(h, std::max(h->pin_bytes, sizeof(gg_dyn)))) return -1;
  std::memcpy(h->pin, &d, sizeof(gg_dyn));
  GG_HIP(hipMemcpyAsync(h->dyn + f, h->pin, sizeof(gg_dyn), hipMemcpyHostToDevice, h->stream));
  GG_HIP(hipEventRecord(h->pin_done, h->stream));
  return 0;
}

// ENU rotation at the anchor (geocentric latitude); identity near the origin.
void enu_rotation(const double* x, double* R) {
  const double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  if (!(r > 1.0)) {
    for (int k = 0; k < 9; ++k) R[k] = (k % 4 == 0) ? 1.0 : 0.0;
    return;
  }
  const double lon = std::atan2(x[1], x[0]);
  const double lat = std::atan2(x[2], std::hypot(x[0], x[1]));
  const double sl = std::sin(lon), cl = std::cos(lon), sp = std::sin(lat), cp = std::cos(lat);
  R[0] = -sl;
  R[1] = cl;
  R[2] = 0.0;
  R[3] = -sp * cl;
  R[4] = -sp * sl;
  R[5] = cp;
  R[6] = cp * cl;
  R[7] = cp * sl;
  R[8] = sp;
}

#define GG_DISPATCH_ISB(nisb, ...) \
  switch (nisb) {                     \
    case 0: {                         \
      constexpr int NI = 0;           \
      __VA_ARGS__;                    \
    } break;                          \
    case 1: {                         \
      constexpr int NI = 1;           \
      __VA_ARGS__;                    \
    } break;                          \
    case 2: {                         \
      constexpr int NI = 2;           \
      __VA_ARGS__;                    \
    } break;                          \
    default: {                        \
      constexpr int NI = 3;           \
      __VA_ARGS__;                    \
    } break;                          \
  }

int max_tiles(const gg_pf* h) {
  int nt = 1;
  for (const auto& f : h->filt_h) nt = std::max(nt, f.n_tiles);
  return nt;
}

size_t decide_smem(const gg_pf* h) { return (size_t)max_tiles(h) * sizeof(double); }

// likelihood (+ predict) over all tiles, then the per-filter decision
// Code tables of the learned likelihood: one row per satellite row of the
// arena (GG_MLP_CODES float2 each) and the row -> filter map.
int ensure_table(gg_pf* h) {
  if (h->table_rows == h->sat_rows && h->mlp_table) return 0;
  GG_HIP(hipStreamSynchronize(h->stream));
  if (h->mlp_table) GG_HIP(hipFree(h->mlp_table));
  if (h->row_filter) GG_HIP(hipFree(h->row_filter));
  h->mlp_table = nullptr;
  h->row_filter = nullptr;
  h->table_rows = 0;
  if (dev_alloc(&h->mlp_table, (size_t)h->sat_rows * GG_MLP_CODES)) return -1;
  std::vector<int32_t> rf((size_t)h->sat_rows, 0);
  for (int f = 0; f < h->n_filters; ++f)
    for (int k = 0; k < h->filt_h[f].sat_cap; ++k) rf[(size_t)h->filt_h[f].sat_off + k] = f;
  if (dev_alloc(&h->row_filter, rf.size())) return -1;
  GG_HIP(hipMemcpy(h->row_filter, rf.data(), sizeof(int32_t) * rf.size(), hipMemcpyHostToDevice));
  h->table_rows = h->sat_rows;
  return 0;
}

// Launch on the handle's stream; with kernel timing on, bracketed by the
// next start/stop event pair (hipExtLaunchKernelGGL records them at the
// kernel's own begin and end, without the launch gaps).
template <typename K, typename... A>
void gg_launch(gg_pf* h, K kernel, const dim3& grid, const dim3& block, size_t smem, A... args) {
  if (h->kt_next + 2 <= h->kt_ev.size()) {
    hipEvent_t e0 = h->kt_ev[h->kt_next], e1 = h->kt_ev[h->kt_next + 1];
    h->kt_next += 2;
    hipExtLaunchKernelGGL(kernel, grid, block, (uint32_t)smem, h->stream, e0, e1, 0u, args...);
  } else {
    kernel<<<grid, block, smem, h->stream>>>(args...);
  }
}

// The mixture likelihood's constants, from its binary32 parameters (binary64
// logs, rounded once); every weight launch carries them.
gg_step_args with_derived(const gg_step_args& in) {
  gg_step_args s = in;
  const double sl = (double)s.sigma_pr, sn = (double)s.gmm_sigma_nlos;
  const double wl = (double)s.gmm_w_los;
  s.gmm_c_los = (float)(std::log(wl) - std::log(sl));
  s.gmm_c_nlos = (float)(std::log(1.0 - wl) - std::log(sn));
  s.gmm_ratio = (float)((sl * sl) / (sn * sn));
  return s;
}

// aos (in: requested, out: done; the plain predict + likelihood and code-table
// kernels do it): the launch writes the predicted state to the packed records
// (only, or also to the state arrays: gg_dyn.aos_only), for the Megopolis
// launches that follow it in the same step (launch_mego with aos_ready), which
// leave the state arrays complete.
int launch_weight(gg_pf* h, const gg_step_args& s_in, bool pred, bool with_decide = true,
                  bool* aos = nullptr) {
  const gg_step_args s = with_derived(s_in);
  const bool learned = s.lik == GG_LIK_LEARNED;
  if (learned && !h->mlp_w) return gg_fail(-3, "learned likelihood requested but no gg_mlp attached");
  const bool table = learned && s.learned_mode == 1;
  if (table && ensure_table(h)) return -1;
  const bool want_aos = aos && *aos && pred && (table || !(learned || h->weight_tpw > 1));
  if (aos) *aos = want_aos;
  gg_arena a = arena_of(h);
  a.aos_step = want_aos ? 1 : 0;
  const size_t smem = sat_smem(h, learned && !table);
  if (table) {
    const dim3 tgrid((GG_MLP_CODES + 1023) / 1024, h->sat_rows);
    GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_mlp_table<NI>), tgrid, dim3(GG_BLOCK), 0,
                                                   a, s, pred ? 1 : 0, h->row_filter));
    GG_HIP(hipGetLastError());
    if (want_aos) {
      GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_weight_table<NI, true, true>),
                                           dim3(h->n_tiles), dim3(GG_BLOCK), smem, a, s));
    } else if (pred) {
      GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_weight_table<NI, true>), dim3(h->n_tiles),
                                                    dim3(GG_BLOCK), smem, a, s));
    } else {
      GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_weight_table<NI, false>), dim3(h->n_tiles),
                                                    dim3(GG_BLOCK), smem, a, s));
    }
  } else if (learned) {
    if (pred) {
      GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_weight_learned<NI, true>), dim3(h->n_tiles),
                                                    dim3(GG_BLOCK), smem, a, s));
    } else {
      GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_weight_learned<NI, false>), dim3(h->n_tiles),
                                                    dim3(GG_BLOCK), smem, a, s));
    }
  } else if (pred && h->weight_tpw > 1) {
    const int tpw = h->weight_tpw;
    const int grid = (h->n_tiles + tpw - 1) / tpw;
    GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_weight_mt<NI, true>), dim3(grid),
                                                  dim3(GG_BLOCK), smem, a, s, tpw));
  } else if (want_aos) {
    GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_weight<NI, true, true>), dim3(h->n_tiles),
                                         dim3(GG_BLOCK), smem, a, s));
  } else if (pred) {
    GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_weight<NI, true>), dim3(h->n_tiles),
                                                  dim3(GG_BLOCK), smem, a, s));
  } else {
    GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_weight<NI, false>), dim3(h->n_tiles),
                                                  dim3(GG_BLOCK), smem, a, s));
  }
  GG_HIP(hipGetLastError());
  if (with_decide) {
    const int mode = (pred ? 1 : 0) | 2;
    GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_decide<NI>), dim3(h->n_filters),
                                                  dim3(GG_BLOCK), decide_smem(h), a, s,
                                                  mode));
    GG_HIP(hipGetLastError());
  }
  return 0;
}

int launch_reduce(gg_pf* h, const gg_step_args& s) {
  const gg_arena a = arena_of(h);
  GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_reduce<NI>), dim3(h->n_tiles),
                                                dim3(GG_BLOCK), 0, a));
  GG_HIP(hipGetLastError());
  GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_decide<NI>), dim3(h->n_filters),
                                                dim3(GG_BLOCK), decide_smem(h), a, s,
                                                0));
  GG_HIP(hipGetLastError());
  return 0;
}

int launch_decide(gg_pf* h, const gg_step_args& s, int mode) {
  const gg_arena a = arena_of(h);
  GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_decide<NI>), dim3(h->n_filters),
                                                dim3(GG_BLOCK), decide_smem(h), a, s,
                                                mode));
  GG_HIP(hipGetLastError());
  return 0;
}

// systematic resampling in one launch: decision (with_decide) or the record's,
// work item search over the merged tile ends and grid points, CDF ends of
// the item's tile, gather and finalize
int launch_sysres(gg_pf* h, const gg_step_args& s, int force, int with_decide) {
  const gg_arena a = arena_of(h);
  // tile sums, then CDF offsets, of every tile; later the ancestors of an
  // item's outputs
  const size_t smem = std::max(decide_smem(h), sizeof(int32_t) * (size_t)GG_QCHUNK);
  GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_sysres<NI>), dim3(h->n_chunks), dim3(GG_BLOCK),
                                       smem, a, s, force, with_decide));
  GG_HIP(hipGetLastError());
  return 0;
}

// Resampling gather; its last work item per filter finalises the resampling
// (gg_gather_finish), so no launch follows it.
int launch_gather(gg_pf* h, const gg_step_args& s, int force, int anc_mode, int src_buf) {
  const gg_arena a = arena_of(h);
  const size_t smem = 0;
  GG_DISPATCH_ISB(h->n_isb, gg_launch(h, (gg_k_gather<NI>), dim3(h->n_chunks),
                                                dim3(GG_BLOCK), smem, a, s, force,
                                                anc_mode, src_buf));
  GG_HIP(hipGetLastError());
  return 0;
}

int launch_mego_ref(gg_pf* h, int n_iter, int force, uint64_t seed, int use_ovr,
                    int* final_buf) {
  const gg_arena a = arena_of(h);
  for (int it = 0; it < n_iter; ++it) {
    gg_launch(h, gg_k_mego_ref, dim3(h->n_tiles), dim3(GG_BLOCK), 0, a, it,
                       force, seed, use_ovr);
    GG_HIP(hipGetLastError());
  }
  *final_buf = n_iter & 1;
  return 0;
}

// aos_ready: the step's likelihood launch, enqueued just before with nothing
// between that reads or changes the state, wrote the predicted state to the
// packed records (and, unless gg_dyn.aos_only, to the state arrays).
int launch_mego(gg_pf* h, int n_iter, int force, bool aos_ready = false) {
  gg_arena a = arena_of(h);
  a.aos_step = aos_ready ? 1 : 0;
  GG_DISPATCH_ISB(h->n_isb, gg_launch(h, gg_k_mego_w<NI>, dim3(h->n_tiles), dim3(GG_BLOCK), 0, a,
                                       force));
  GG_HIP(hipGetLastError());
  // the chain kernel finalises the resampling itself (one gather item per
  // tile: the filter's gather item count equals its tile count)
  switch (h->mego_depth) {
    case 8:
      GG_DISPATCH_ISB(h->n_isb, gg_launch(h, gg_k_mego<NI, 8>, dim3(h->n_tiles), dim3(GG_BLOCK), 0,
                                           a, n_iter, force));
      break;
    case 4:
      GG_DISPATCH_ISB(h->n_isb, gg_launch(h, gg_k_mego<NI, 4>, dim3(h->n_tiles), dim3(GG_BLOCK), 0,
                                           a, n_iter, force));
      break;
    case 2:
      GG_DISPATCH_ISB(h->n_isb, gg_launch(h, gg_k_mego<NI, 2>, dim3(h->n_tiles), dim3(GG_BLOCK), 0,
                                           a, n_iter, force));
      break;
    default:
      GG_DISPATCH_ISB(h->n_isb, gg_launch(h, gg_k_mego<NI, 1>, dim3(h->n_tiles), dim3(GG_BLOCK), 0,
                                           a, n_iter, force));
  }
  GG_HIP(hipGetLastError());
  return 0;
}

// Megopolis part of a step, after the decision kernel.
int enqueue_resample_after_decide(gg_pf* h, const gg_step_args& s, int resampler, int n_iter,
                                  bool aos_ready = false) {
  if (resampler == 1) {
    return launch_mego(h, n_iter, 0, aos_ready);
  }
  int fb = 0;
  if (n_iter <= 0) return launch_gather(h, s, 0, 2, 0);
  if (launch_mego_ref(h, n_iter, 0, 0, 0, &fb)) return -1;
  return launch_gather(h, s, 0, 1, fb);
}

// Resampling part of a device-driven step, after the weight launch (the
// gather's last work item per filter finalises).  Systematic: the CDF kernel takes the decision in every
// workgroup (and commits the predict's frame), so no decision kernel runs.
int enqueue_resample(gg_pf* h, const gg_step_args& s, int resampler, int n_iter,
                     bool aos_ready = false) {
  if (resampler == 0) {
    return launch_sysres(h, s, 0, 1);
  }
  if (launch_decide(h, s, 3)) return -1;
  return enqueue_resample_after_decide(h, s, resampler, n_iter, aos_ready);
}

// One device-driven step in two parts (no host synchronisation): part 0 is the
// likelihood launch (predict + weight + tile partials), part 1 the decision
// and resampling.
int enqueue_step_part(gg_pf* h, const gg_step_args& s, int resampler, int n_iter, int part) {
  if (part == 0) return launch_weight(h, s, true, false);
  return enqueue_resample(h, s, resampler, n_iter);
}

// A whole step: in a Megopolis step the likelihood launch writes the predicted
// state as packed records, which the chains read, and after a resampling step
// only those.  (Separate parts keep the plain layout: a caller may read or
// change the state between them.)
int enqueue_step(gg_pf* h, const gg_step_args& s, int resampler, int n_iter) {
  bool aos = resampler == 1;
  if (launch_weight(h, s, true, false, &aos)) return -1;
  return enqueue_resample(h, s, resampler, n_iter, aos);
}

gg_step_args default_args() {
  gg_step_args s;
  std::memset(&s, 0, sizeof(s));
  s.dt = 1.0;
  s.sigma_pos = 1.0f;
  s.sigma_cb = 300.0f;
  s.sigma_pr = 5.0f;
  s.alpha = 1.0f;
  s.huber_k = 1.5f;
  s.n_slots = 1;
  s.ess_frac = 0.5;
  s.recenter = 1;
  s.gmm_w_los = 0.7f;
  s.gmm_mu_nlos = 15.0f;
  s.gmm_sigma_nlos = 30.0f;
  return s;
}

int check_single(gg_pf* h) {
  if (!h) return gg_fail(-2, "null handle");
  if (!h->initialized) return gg_fail(-3, "particle filter not initialized");
  return 0;
}

int ensure_decide(gg_pf* h) {
  if (h->decide_valid) return 0;
  gg_step_args s = default_args();
  s.auto_flag = 0;
  s.recenter = 1;
  if (launch_reduce(h, s)) return -1;
  h->decide_valid = true;
  h->ess_valid = true;
  h->est_valid = true;
  return 0;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int gg_struct_sizes(int32_t* step_args, int32_t* dyn, int32_t* filter) {
  if (step_args) *step_args = (int32_t)sizeof(gg_step_args);
  if (dyn) *dyn = (int32_t)sizeof(gg_dyn);
  if (filter) *filter = (int32_t)sizeof(gg_filter);
  return 0;
}

int gg_pf_create(int32_t n_filters, const int32_t* n_particles, const int32_t* sat_cap,
                 int32_t n_isb, int32_t n_slots, int32_t with_velocity, gg_pf** out) {
  if (!out) return gg_fail(-2, "out is null");
  *out = nullptr;
  if (n_filters <= 0) return gg_fail(-2, "n_filters must be positive");
  if (n_isb < 0 || n_isb > GG_MAX_ISB) return gg_fail(-2, "n_isb must be in [0, 3]");
  if (n_slots <= 0) return gg_fail(-2, "n_slots must be positive");
  gg_pf* h = new gg_pf();
  h->n_filters = n_filters;
  h->n_isb = n_isb;
  h->D = 4 + n_isb;
  h->has_vel = with_velocity ? 1 : 0;
  h->ncomp = h->D + (h->has_vel ? GG_NVEL : 0);
  h->n_slots = n_slots;
  int64_t p = 0;
  int32_t tiles = 0, chunks = 0, rows = 0;
  for (int f = 0; f < n_filters; ++f) {
    const int n = n_particles[f];
    if (n <= 0 || n > GG_MAX_PARTICLES) {
      delete h;
      return gg_fail(-2, "n_particles must be in [1, 7864320] per filter");
    }
    gg_filter F;
    std::memset(&F, 0, sizeof(F));
    F.p_off = p;
    F.n = n;
    F.n_tiles = (n + GG_TILE - 1) / GG_TILE;
    F.tile_off = tiles;
    // systematic resampling items: one per tile plus one per GG_QCHUNK outputs bounds
    // sum_b ceil(outputs_b / GG_QCHUNK)
    F.n_chunks = F.n_tiles + (int32_t)(((int64_t)n + GG_QCHUNK - 1) / GG_QCHUNK);
    F.chunk_off = chunks;
    F.sat_off = rows;
    F.sat_cap = std::max(sat_cap ? sat_cap[f] : 64, 1);
    F.seed = 0;
    h->filt_h.push_back(F);
    for (int t = 0; t < F.n_tiles; ++t) h->tile_filter_h.push_back(f);
    for (int c = 0; c < F.n_chunks; ++c) h->chunk_filter_h.push_back(f);
    p += (int64_t)F.n_tiles * GG_TILE;
    tiles += F.n_tiles;
    chunks += F.n_chunks;
    rows += F.sat_cap;
  }
  h->P = p;
  h->n_tiles = tiles;
  h->n_chunks = chunks;
  h->sat_rows = rows;
  int rc = 0;
  if (!rc) rc = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) == hipSuccess ? 0 : -1;
  h->own_stream = true;
  if (!rc) rc = dev_alloc(&h->state, (size_t)2 * h->ncomp * h->P);
  if (!rc) rc = dev_alloc(&h->lw, (size_t)h->P);
  if (!rc) rc = dev_alloc(&h->anc, (size_t)2 * h->P);
  if (!rc) rc = dev_alloc(&h->part, (size_t)tiles * GG_NPART);
  if (!rc) rc = dev_alloc(&h->p4, (size_t)tiles * 4);
  if (!rc) rc = dev_alloc(&h->g, (size_t)tiles);
  if (!rc) rc = dev_alloc(&h->Pb, (size_t)tiles);
  if (!rc) rc = dev_alloc(&h->E, (size_t)h->P);
  if (!rc) rc = dev_alloc(&h->aos, (size_t)h->P * 8);
  if (!rc) rc = dev_alloc(&h->csum, (size_t)chunks * GG_MAX_D);
  if (!rc) rc = dev_alloc(&h->fin, (size_t)n_filters * GG_ARRIVE_WORDS);
  if (!rc) rc = dev_alloc(&h->acc, (size_t)n_filters * GG_ARRIVE_SUBS * GG_ACC_W);
  if (!rc) rc = dev_alloc(&h->scratch_f, (size_t)n_filters * 16);
  if (!rc) rc = dev_alloc(&h->filt, (size_t)n_filters);
  if (!rc) rc = dev_alloc(&h->dyn, (size_t)n_filters);
  if (!rc) rc = dev_alloc(&h->tile_filter, (size_t)tiles);
  if (!rc) rc = dev_alloc(&h->chunk_filter, (size_t)chunks);
  if (!rc) rc = dev_alloc(&h->satd, (size_t)n_slots * 5 * rows);
  if (!rc) rc = dev_alloc(&h->sati, (size_t)n_slots * (rows + n_filters));
  if (!rc) rc = dev_alloc(&h->satf, (size_t)n_slots * 2 * rows);
  if (!rc) rc = dev_alloc(&h->vel_guide, (size_t)n_slots * n_filters * 3);
  if (!rc) rc = hipEventCreateWithFlags(&h->pin_done, hipEventDisableTiming) == hipSuccess ? 0 : -1;
  if (!rc) rc = hipEventCreate(&h->ev0) == hipSuccess ? 0 : -1;
  if (!rc) rc = hipEventCreate(&h->ev1) == hipSuccess ? 0 : -1;
  if (!rc) rc = ensure_pinned(h, std::max<size_t>(64 * 1024, sizeof(gg_dyn) * n_filters));
  if (!rc) {
    rc = hipMemcpy(h->filt, h->filt_h.data(), sizeof(gg_filter) * n_filters,
                   hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
  }
  if (!rc) {
    rc = hipMemcpy(h->tile_filter, h->tile_filter_h.data(), sizeof(int32_t) * tiles,
                   hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
  }
  if (!rc) {
    rc = hipMemcpy(h->chunk_filter, h->chunk_filter_h.data(), sizeof(int32_t) * chunks,
                   hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
  }
  if (rc) {
    std::string msg = gg_last_error_str();
    gg_pf_destroy(h);
    return gg_fail(-1, msg.empty() ? "gg_pf_create: HIP allocation failed" : msg.c_str());
  }
  *out = h;
  return 0;
}

int gg_pf_device_create(int32_t n_particles, int32_t n_isb, gg_pf** out) {
  const int32_t cap = 64;
  return gg_pf_create(1, &n_particles, &cap, n_isb, 1, 1, out);
}

int gg_pf_destroy(gg_pf* h) {
  if (!h) return 0;
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->graph_exec) hipGraphExecDestroy(h->graph_exec);
  if (h->graph) hipGraphDestroy(h->graph);
  void* bufs[] = {h->state, h->lw,    h->E,  h->aos,  h->anc,  h->part, h->p4, h->g, h->Pb,
                  h->csum, h->fin, h->acc, h->scratch_f, h->filt,
                  h->dyn,   h->satd,  h->sati,      h->tile_filter, h->chunk_filter, h->vel_guide,
                  h->satf,  h->mlp_table, h->row_filter,
                  h->io_tmp, h->obsd};
  for (void* b : bufs)
    if (b) hipFree(b);
  if (h->pin) hipHostFree(h->pin);
  if (h->pin_done) hipEventDestroy(h->pin_done);
  if (h->ev0) hipEventDestroy(h->ev0);
  if (h->ev1) hipEventDestroy(h->ev1);
  if (h->stream && h->own_stream) hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int gg_pf_set_stream(gg_pf* h, void* stream) {
  if (!h) return gg_fail(-2, "null handle");
  GG_HIP(hipStreamSynchronize(h->stream));
  if (stream) {
    if (h->own_stream) GG_HIP(hipStreamDestroy(h->stream));
    h->stream = (hipStream_t)stream;
    h->own_stream = false;
  } else if (!h->own_stream) {
    GG_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->own_stream = true;
  }
  return 0;
}

int gg_pf_device_sync(gg_pf* h) {
  if (!h) return 0;
  GG_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int gg_pf_info(gg_pf* h, int32_t* n_filters, int32_t* n_isb, int64_t* arena_slots) {
  if (!h) return gg_fail(-2, "null handle");
  if (n_filters) *n_filters = h->n_filters;
  if (n_isb) *n_isb = h->n_isb;
  if (arena_slots) *arena_slots = h->P;
  return 0;
}

// ---------------------------------------------------------------------------
// single-filter API
// ---------------------------------------------------------------------------
int gg_pf_device_initialize(gg_pf* h, double init_x, double init_y, double init_z,
                            double init_cb, double spread_pos, double spread_cb, uint64_t seed,
                            double init_vx, double init_vy, double init_vz, double spread_vel,
                            double init_vel_sigma, const double* init_isb, double spread_isb) {
  if (!h) return gg_fail(-2, "null handle");
  if (spread_vel > 0.0 && !h->has_vel)
    return gg_fail(-2, "spread_vel > 0 needs velocity storage (with_velocity)");
  GG_HIP(hipStreamSynchronize(h->stream));
  for (int f = 0; f < h->n_filters; ++f) h->filt_h[f].seed = seed;
  GG_HIP(hipMemcpy(h->filt, h->filt_h.data(), sizeof(gg_filter) * h->n_filters,
                   hipMemcpyHostToDevice));
  h->seed0 = seed;
  for (int f = 0; f < h->n_filters; ++f) {
    gg_dyn d;
    std::memset(&d, 0, sizeof(d));
    d.anchor[0] = init_x;
    d.anchor[1] = init_y;
    d.anchor[2] = init_z;
    d.cb_anchor = init_cb;
    for (int k = 0; k < h->n_isb; ++k) d.isb_anchor[k] = init_isb ? init_isb[k] : 0.0;
    enu_rotation(d.anchor, d.R);
    const double v[3] = {init_vx, init_vy, init_vz};
    d.vel_u[0] = d.R[0] * v[0] + d.R[1] * v[1] + d.R[2] * v[2];
    d.vel_u[1] = d.R[3] * v[0] + d.R[4] * v[1] + d.R[5] * v[2];
    d.vel_u[2] = d.R[6] * v[0] + d.R[7] * v[1] + d.R[8] * v[2];
    const double vv = init_vel_sigma * init_vel_sigma;
    d.vcov_u[0] = vv;
    d.vcov_u[3] = vv;
    d.vcov_u[5] = vv;
    d.vel_mode = (spread_vel > 0.0) ? 1 : 0;
    d.lw_zero = 1;
    d.ess = (double)h->filt_h[f].n;  // uniform weights
    GG_HIP(hipMemcpy(h->dyn + f, &d, sizeof(d), hipMemcpyHostToDevice));
  }
  const gg_arena a = arena_of(h);
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_init<NI>), dim3(h->n_tiles),
                                                dim3(GG_BLOCK), 0, h->stream, a,
                                                (float)spread_pos, (float)spread_cb,
                                                (float)spread_isb, (float)spread_vel));
  GG_HIP(hipGetLastError());
  GG_HIP(hipStreamSynchronize(h->stream));
  h->initialized = true;
  h->decide_valid = false;
  h->ess_valid = false;
  h->est_valid = false;
  return 0;
}

int gg_pf_device_predict(gg_pf* h, double vx, double vy, double vz, double dt, double sigma_pos,
                         double sigma_cb, uint64_t seed, int32_t step, double sigma_vel,
                         double velocity_guide_alpha, int32_t velocity_kf,
                         double velocity_process_noise, double sigma_isb) {
  if (int rc = check_single(h)) return rc;
  if (sigma_vel > 0.0 && !velocity_kf && !h->has_vel)
    return gg_fail(-2, "sigma_vel > 0 needs velocity storage (with_velocity)");
  if (seed != h->filt_h[0].seed) {
    GG_HIP(hipStreamSynchronize(h->stream));
    for (auto& F : h->filt_h) F.seed = seed;
    GG_HIP(hipMemcpy(h->filt, h->filt_h.data(), sizeof(gg_filter) * h->n_filters,
                     hipMemcpyHostToDevice));
  }
  if (ensure_pinned(h, h->pin_bytes)) return -1;
  double* pv = (double*)h->pin;
  for (int f = 0; f < h->n_filters; ++f) {
    pv[f * 3 + 0] = vx;
    pv[f * 3 + 1] = vy;
    pv[f * 3 + 2] = vz;
  }
  GG_HIP(hipMemcpyAsync(h->vel_guide, pv, sizeof(double) * 3 * h->n_filters,
                        hipMemcpyHostToDevice, h->stream));
  GG_HIP(hipEventRecord(h->pin_done, h->stream));
  gg_step_args s = default_args();
  s.dt = dt;
  s.sigma_pos = (float)sigma_pos;
  s.sigma_cb = (float)sigma_cb;
  s.sigma_isb = (float)sigma_isb;
  s.sigma_vel = (float)sigma_vel;
  s.alpha = (float)velocity_guide_alpha;
  s.velocity_kf = velocity_kf ? 1 : 0;
  s.qv = (float)velocity_process_noise;
  s.step = step;
  const gg_arena a = arena_of(h);
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_predict<NI>), dim3(h->n_tiles),
                                                dim3(GG_BLOCK), 0, h->stream, a, s));
  GG_HIP(hipGetLastError());
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_commit_predict<NI>), dim3(h->n_filters),
                                                dim3(64), 0, h->stream, a, s));
  GG_HIP(hipGetLastError());
  h->est_valid = false;
  return 0;
}

}  // extern "C"

namespace {

// Satellites of the single-filter API into epoch slot 0 (grows the table past
// its capacity like the reference, pf_device.cu:1533-1545).
int stage_single(gg_pf* h, const double* sat_ecef, const double* pseudoranges,
                 const double* weights_sat, const int32_t* sys, int32_t n_sat, const float* cn0,
                 const float* los_flag) {
  if (n_sat < 0) return gg_fail(-2, "n_sat must be >= 0");
  if (h->n_filters != 1) return gg_fail(-2, "single-filter call on a batch handle");
  if (n_sat > h->filt_h[0].sat_cap) {
    GG_HIP(hipStreamSynchronize(h->stream));
    const int cap = std::max(n_sat, 2 * h->filt_h[0].sat_cap);
    h->filt_h[0].sat_cap = cap;
    h->sat_rows = cap;
    GG_HIP(hipFree(h->satd));
    GG_HIP(hipFree(h->sati));
    GG_HIP(hipFree(h->satf));
    h->satd = nullptr;
    h->sati = nullptr;
    h->satf = nullptr;
    if (dev_alloc(&h->satd, (size_t)h->n_slots * 5 * cap)) return -1;
    if (dev_alloc(&h->sati, (size_t)h->n_slots * (cap + 1))) return -1;
    if (dev_alloc(&h->satf, (size_t)h->n_slots * 2 * cap)) return -1;
    GG_HIP(hipMemcpy(h->filt, h->filt_h.data(), sizeof(gg_filter), hipMemcpyHostToDevice));
  }
  const int rows = h->sat_rows;
  const size_t dbytes = sizeof(double) * 5 * rows;
  const size_t ibytes = sizeof(int32_t) * (rows + 1);
  const size_t fbytes = sizeof(float) * 2 * rows;
  if (ensure_pinned(h, std::max(h->pin_bytes, dbytes + ibytes + fbytes + 64))) return -1;
  double* pd = (double*)h->pin;
  std::memset(pd, 0, dbytes);
  for (int k = 0; k < n_sat; ++k) {
    pd[k * 3 + 0] = sat_ecef[k * 3 + 0];
    pd[k * 3 + 1] = sat_ecef[k * 3 + 1];
    pd[k * 3 + 2] = sat_ecef[k * 3 + 2];
    pd[3 * rows + k] = pseudoranges[k];
    pd[4 * rows + k] = weights_sat ? weights_sat[k] : 1.0;
  }
  int32_t* pi = (int32_t*)(h->pin + dbytes);
  for (int k = 0; k < rows; ++k) pi[k] = (sys && k < n_sat) ? sys[k] : 0;
  pi[rows] = n_sat;
  GG_HIP(hipMemcpyAsync(h->satd, pd, dbytes, hipMemcpyHostToDevice, h->stream));
  GG_HIP(hipMemcpyAsync(h->sati, pi, ibytes, hipMemcpyHostToDevice, h->stream));
  if (cn0 && los_flag) {
    float* pf = (float*)(h->pin + dbytes + ibytes);
    for (int k = 0; k < rows; ++k) {
      pf[k] = (k < n_sat) ? cn0[k] : 0.0f;
      pf[rows + k] = (k < n_sat) ? los_flag[k] : 0.0f;
    }
    GG_HIP(hipMemcpyAsync(h->satf, pf, fbytes, hipMemcpyHostToDevice, h->stream));
  }
  GG_HIP(hipEventRecord(h->pin_done, h->stream));
  return 0;
}

// Observations of a DD / carrier update, 11 binary64 per observation (see
// gg_k_weight_obs), through the pinned buffer to h->obsd.
constexpr int GG_MAX_OBS = 1024;

int stage_obs(gg_pf* h, int n, const double* sk, const double* sr, const double* val,
              const double* brk, const double* brr, const double* w, const double* wl,
              double wl_all) {
  if (n < 0 || n > GG_MAX_OBS) return gg_fail(-2, "observation count must be in [0, 1024]");
  if (n > h->obs_cap) {
    GG_HIP(hipStreamSynchronize(h->stream));
    if (h->obsd) GG_HIP(hipFree(h->obsd));
    h->obsd = nullptr;
    h->obs_cap = 0;
    if (dev_alloc(&h->obsd, (size_t)std::max(n, 16) * 11)) return -1;
    h->obs_cap = std::max(n, 16);
  }
  const size_t bytes = sizeof(double) * 11 * (size_t)std::max(n, 1);
  if (ensure_pinned(h, std::max(h->pin_bytes, bytes))) return -1;
  double* o = (double*)h->pin;
  for (int k = 0; k < n; ++k) {
    double* r = o + (size_t)k * 11;
    for (int c = 0; c < 3; ++c) {
      r[c] = sk[k * 3 + c];
      r[3 + c] = sr ? sr[k * 3 + c] : 0.0;
    }
    r[6] = val[k];
    r[7] = brk ? brk[k] : 0.0;
    r[8] = brr ? brr[k] : 0.0;
    r[9] = w ? w[k] : 1.0;
    r[10] = wl ? wl[k] : wl_all;
  }
  if (n > 0)
    GG_HIP(hipMemcpyAsync(h->obsd, o, sizeof(double) * 11 * n, hipMemcpyHostToDevice, h->stream));
  GG_HIP(hipEventRecord(h->pin_done, h->stream));
  return 0;
}

int launch_obs(gg_pf* h, int kind, int n, double sigma, double thr, int huber, double huber_k) {
  const gg_arena a = arena_of(h);
  const size_t smem = (size_t)std::max(n, 1) * 48;
  const float sg = (float)sigma, th = (float)thr, hk = (float)huber_k;
#define GG_OBS_LAUNCH(K)                                                                       \
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_weight_obs<K, NI>), dim3(h->n_tiles),      \
                                                dim3(GG_BLOCK), smem, h->stream, a, h->obsd, n, \
                                                sg, th, huber, hk))
  if (kind == 0) {
    GG_OBS_LAUNCH(0);
  } else if (kind == 1) {
    GG_OBS_LAUNCH(1);
  } else {
    GG_OBS_LAUNCH(2);
  }
#undef GG_OBS_LAUNCH
  GG_HIP(hipGetLastError());
  hipLaunchKernelGGL(gg_k_clear_lw_zero, dim3(1), dim3(64), 0, h->stream, h->dyn, 0);
  GG_HIP(hipGetLastError());
  h->decide_valid = false;
  h->ess_valid = false;
  h->est_valid = false;
  return 0;
}

// Ray-traced NLOS weight on a single filter's particles (gg_k_weight_3d_pf).
int launch_weight_3d(gg_pf* h, const gg_bvh* bvh, int n, const gg_rt::MixArgs& mx) {
  const gg_arena a = arena_of(h);
  const size_t smem = (size_t)std::max(n, 1) * 48;
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_weight_3d_pf<NI>), dim3(h->n_tiles),
                                                dim3(GG_BLOCK), smem, h->stream, a, h->obsd, n,
                                                bvh->nodes, bvh->tris, bvh->n_nodes, mx));
  GG_HIP(hipGetLastError());
  hipLaunchKernelGGL(gg_k_clear_lw_zero, dim3(1), dim3(64), 0, h->stream, h->dyn, 0);
  GG_HIP(hipGetLastError());
  h->decide_valid = false;
  h->ess_valid = false;
  h->est_valid = false;
  return 0;
}

// Doppler rows, 8 binary64 per satellite (see gg_k_doppler), to h->obsd.
int stage_doppler(gg_pf* h, int n, const double* sat, const double* vel, const double* dop,
                  const double* w) {
  if (n > 512) return gg_fail(-2, "at most 512 Doppler satellites");
  if (n > h->obs_cap) {
    GG_HIP(hipStreamSynchronize(h->stream));
    if (h->obsd) GG_HIP(hipFree(h->obsd));
    h->obsd = nullptr;
    h->obs_cap = 0;
    if (dev_alloc(&h->obsd, (size_t)std::max(n, 16) * 11)) return -1;
    h->obs_cap = std::max(n, 16);
  }
  const size_t bytes = sizeof(double) * 8 * (size_t)n;
  if (ensure_pinned(h, std::max(h->pin_bytes, bytes))) return -1;
  double* o = (double*)h->pin;
  for (int k = 0; k < n; ++k) {
    double* r = o + (size_t)k * 8;
    for (int c = 0; c < 3; ++c) {
      r[c] = sat[k * 3 + c];
      r[3 + c] = vel[k * 3 + c];
    }
    r[6] = dop[k];
    r[7] = w ? w[k] : 1.0;
  }
  GG_HIP(hipMemcpyAsync(h->obsd, o, bytes, hipMemcpyHostToDevice, h->stream));
  GG_HIP(hipEventRecord(h->pin_done, h->stream));
  return 0;
}

int launch_doppler(gg_pf* h, bool kf, int n, double wl, double sigma, double gain,
                   double max_dv) {
  const gg_arena a = arena_of(h);
  const int grid = (h->filt_h[0].n + GG_BLOCK - 1) / GG_BLOCK;
  const size_t smem = (size_t)n * 64;
  if (kf) {
    GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_doppler<NI, true>), dim3(grid),
                                                  dim3(GG_BLOCK), smem, h->stream, a, h->obsd, n,
                                                  wl, sigma, gain, max_dv));
  } else {
    GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_doppler<NI, false>), dim3(grid),
                                                  dim3(GG_BLOCK), smem, h->stream, a, h->obsd, n,
                                                  wl, sigma, gain, max_dv));
  }
  GG_HIP(hipGetLastError());
  hipLaunchKernelGGL(gg_k_set_vel_mode, dim3(1), dim3(64), 0, h->stream, h->dyn, 0, kf ? 2 : 1);
  GG_HIP(hipGetLastError());
  hipLaunchKernelGGL(gg_k_clear_lw_zero, dim3(1), dim3(64), 0, h->stream, h->dyn, 0);
  GG_HIP(hipGetLastError());
  h->decide_valid = false;
  h->ess_valid = false;
  h->est_valid = false;
  return 0;
}

}  // namespace

extern "C" {

int gg_pf_device_weight_doppler(gg_pf* h, const double* sat_ecef, const double* sat_vel,
                                const double* doppler_hz, const double* weights_sat,
                                int32_t n_sat, double wavelength_m, double sigma_mps,
                                double velocity_update_gain, double max_velocity_update_mps) {
  if (int rc = check_single(h)) return rc;
  if (n_sat <= 0 || !(sigma_mps > 0.0) || !(wavelength_m > 0.0)) return 0;
  if (!h->has_vel) return gg_fail(-2, "Doppler updates need velocity storage");
  if (!sat_ecef || !sat_vel || !doppler_hz) return gg_fail(-2, "null argument");
  if (stage_doppler(h, n_sat, sat_ecef, sat_vel, doppler_hz, weights_sat)) return -1;
  return launch_doppler(h, false, n_sat, wavelength_m, sigma_mps, velocity_update_gain,
                        max_velocity_update_mps);
}

int gg_pf_device_doppler_kf_update(gg_pf* h, const double* sat_ecef, const double* sat_vel,
                                   const double* doppler_hz, const double* weights_sat,
                                   int32_t n_sat, double wavelength_m, double sigma_mps) {
  if (int rc = check_single(h)) return rc;
  if (n_sat <= 0 || !(sigma_mps > 0.0) || !(wavelength_m > 0.0)) return 0;
  if (!h->has_vel) return gg_fail(-2, "Doppler updates need velocity storage");
  if (!sat_ecef || !sat_vel || !doppler_hz) return gg_fail(-2, "null argument");
  if (stage_doppler(h, n_sat, sat_ecef, sat_vel, doppler_hz, weights_sat)) return -1;
  return launch_doppler(h, true, n_sat, wavelength_m, sigma_mps, 0.0, 0.0);
}

int gg_pf_device_weight_dd_pseudorange(gg_pf* h, const double* sat_ecef_k, const double* ref_ecef,
                                       const double* dd_pseudorange, const double* base_range_k,
                                       const double* base_range_ref, const double* weights_dd,
                                       int32_t n_dd, double sigma_pr,
                                       double per_particle_nlos_threshold_m,
                                       int32_t per_particle_huber, double per_particle_huber_k) {
  if (int rc = check_single(h)) return rc;
  if (n_dd > 0 && (!sat_ecef_k || !ref_ecef || !dd_pseudorange || !base_range_k ||
                   !base_range_ref || !weights_dd))
    return gg_fail(-2, "null argument");
  if (stage_obs(h, n_dd, sat_ecef_k, ref_ecef, dd_pseudorange, base_range_k, base_range_ref,
                weights_dd, nullptr, 1.0))
    return -1;
  return launch_obs(h, 0, n_dd, sigma_pr, per_particle_nlos_threshold_m, per_particle_huber,
                    per_particle_huber_k);
}

int gg_pf_device_weight_3d_bvh(gg_pf* h, const gg_bvh* bvh, const double* sat_ecef,
                               const double* pseudoranges, const double* weights_sat,
                               const int32_t* system_ids, int32_t n_sat, double sigma_pr_los,
                               double sigma_pr_nlos, double nlos_bias, double blocked_nlos_prob,
                               double clear_nlos_prob) {
  if (int rc = check_single(h)) return rc;
  if (!bvh) return gg_fail(-2, "null BVH handle");
  if (n_sat > 0 && (!sat_ecef || !pseudoranges)) return gg_fail(-2, "null argument");
  std::vector<double> grp((size_t)std::max(n_sat, 1), 0.0);
  for (int k = 0; k < n_sat; ++k) {
    const int g = system_ids ? system_ids[k] : 0;
    grp[k] = (g >= 1 && g <= h->n_isb) ? (double)g : 0.0;
  }
  if (stage_obs(h, n_sat, sat_ecef, nullptr, pseudoranges, grp.data(), nullptr, weights_sat,
                nullptr, 0.0))
    return -1;
  gg_rt::MixArgs mx;
  gg_rt::make_mix(sigma_pr_los, sigma_pr_nlos, nlos_bias, blocked_nlos_prob, clear_nlos_prob, mx);
  return launch_weight_3d(h, bvh, n_sat, mx);
}

int gg_pf_device_weight_carrier_afv(gg_pf* h, const double* sat_ecef, const double* carrier_phase,
                                    const double* weights_sat, int32_t n_sat, double wavelength,
                                    double sigma_cycles) {
  if (int rc = check_single(h)) return rc;
  if (n_sat > 0 && (!sat_ecef || !carrier_phase || !weights_sat))
    return gg_fail(-2, "null argument");
  if (stage_obs(h, n_sat, sat_ecef, nullptr, carrier_phase, nullptr, nullptr, weights_sat,
                nullptr, wavelength))
    return -1;
  return launch_obs(h, 1, n_sat, sigma_cycles, 0.0, 0, 1.5);
}

int gg_pf_device_weight_dd_carrier_afv(gg_pf* h, const double* sat_ecef_k, const double* ref_ecef,
                                       const double* dd_carrier, const double* base_range_k,
                                       const double* base_range_ref, const double* weights_dd,
                                       const double* wavelengths_m, int32_t n_dd,
                                       double sigma_cycles,
                                       double per_particle_nlos_threshold_cycles,
                                       int32_t per_particle_huber, double per_particle_huber_k) {
  if (int rc = check_single(h)) return rc;
  if (n_dd > 0 && (!sat_ecef_k || !ref_ecef || !dd_carrier || !base_range_k ||
                   !base_range_ref || !weights_dd || !wavelengths_m))
    return gg_fail(-2, "null argument");
  if (stage_obs(h, n_dd, sat_ecef_k, ref_ecef, dd_carrier, base_range_k, base_range_ref,
                weights_dd, wavelengths_m, 1.0))
    return -1;
  return launch_obs(h, 2, n_dd, sigma_cycles, per_particle_nlos_threshold_cycles,
                    per_particle_huber, per_particle_huber_k);
}

int gg_pf_device_weight(gg_pf* h, const double* sat_ecef, const double* pseudoranges,
                        const double* weights_sat, const int32_t* sys, int32_t n_sat,
                        double sigma_pr, double nu, double per_particle_nlos_threshold_m,
                        int32_t per_particle_huber, double per_particle_huber_k) {
  if (int rc = check_single(h)) return rc;
  if (stage_single(h, sat_ecef, pseudoranges, weights_sat, sys, n_sat, nullptr, nullptr))
    return -1;
  gg_step_args s = default_args();
  s.sigma_pr = (float)sigma_pr;
  s.nu = (float)nu;
  s.gate_thr = (float)per_particle_nlos_threshold_m;
  s.huber_k = (float)per_particle_huber_k;
  const bool student = nu > 0.0;
  s.lik = student ? GG_LIK_STUDENT
                  : ((per_particle_huber && per_particle_huber_k > 0.0) ? GG_LIK_HUBER
                                                                        : GG_LIK_GAUSS);
  s.auto_flag = 0;
  if (launch_weight(h, s, false)) return -1;
  h->decide_valid = true;
  h->ess_valid = true;
  h->est_valid = true;
  return 0;
}

int gg_pf_device_weight_gmm(gg_pf* h, const double* sat_ecef, const double* pseudoranges,
                            const double* weights_sat, int32_t n_sat, double sigma_pr,
                            double w_los, double mu_nlos, double sigma_nlos) {
  if (int rc = check_single(h)) return rc;
  if (stage_single(h, sat_ecef, pseudoranges, weights_sat, nullptr, n_sat, nullptr, nullptr))
    return -1;
  gg_step_args s = default_args();
  s.sigma_pr = (float)sigma_pr;
  s.lik = GG_LIK_GMM;
  s.gmm_w_los = (float)w_los;
  s.gmm_mu_nlos = (float)mu_nlos;
  s.gmm_sigma_nlos = (float)sigma_nlos;
  s.auto_flag = 0;
  if (launch_weight(h, s, false)) return -1;
  h->decide_valid = true;
  h->ess_valid = true;
  h->est_valid = true;
  return 0;
}

int gg_pf_device_weight_learned(gg_pf* h, const double* sat_ecef, const double* pseudoranges,
                                const double* weights_sat, const int32_t* sys, int32_t n_sat,
                                const float* cn0, const float* los_flag, int32_t learned_mode) {
  if (int rc = check_single(h)) return rc;
  if (!cn0 || !los_flag) return gg_fail(-2, "learned likelihood needs cn0 and los_flag");
  if (stage_single(h, sat_ecef, pseudoranges, weights_sat, sys, n_sat, cn0, los_flag)) return -1;
  if (learned_mode < 0 || learned_mode > 1) return gg_fail(-2, "learned_mode must be 0 or 1");
  gg_step_args s = default_args();
  s.lik = GG_LIK_LEARNED;
  s.learned_mode = learned_mode;
  s.auto_flag = 0;
  if (launch_weight(h, s, false)) return -1;
  h->decide_valid = true;
  h->ess_valid = true;
  h->est_valid = true;
  return 0;
}

int gg_pf_set_mlp(gg_pf* h, gg_mlp* m) {
  if (!h) return gg_fail(-2, "null handle");
  h->mlp_w = gg_mlp_packed_weights(m);
  return 0;
}

int gg_pfb_load_features(gg_pf* h, int32_t n_slots, const float* cn0, const float* los_flag) {
  if (!h || !cn0 || !los_flag) return gg_fail(-2, "null argument");
  if (n_slots <= 0 || n_slots > h->n_slots) return gg_fail(-2, "n_slots out of range");
  const int rows = h->sat_rows;
  std::vector<float> buf((size_t)n_slots * 2 * rows);
  for (int t = 0; t < n_slots; ++t)
    for (int k = 0; k < rows; ++k) {
      buf[(size_t)t * 2 * rows + k] = cn0[(size_t)t * rows + k];
      buf[(size_t)t * 2 * rows + rows + k] = los_flag[(size_t)t * rows + k];
    }
  GG_HIP(hipStreamSynchronize(h->stream));
  GG_HIP(hipMemcpy(h->satf, buf.data(), sizeof(float) * buf.size(), hipMemcpyHostToDevice));
  return 0;
}

int gg_pf_device_ess(gg_pf* h, double* out) {
  if (int rc = check_single(h)) return rc;
  if (!h->ess_valid && ensure_decide(h)) return -1;
  gg_dyn d;
  if (read_dyn(h, 0, &d)) return -1;
  *out = d.ess;
  return 0;
}

int gg_pf_device_position_spread(gg_pf* h, double cx, double cy, double cz, double* out) {
  if (int rc = check_single(h)) return rc;
  const gg_arena a = arena_of(h);
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_spread<NI>), dim3(h->n_tiles),
                                                dim3(GG_BLOCK), 0, h->stream, a, cx, cy, cz));
  GG_HIP(hipGetLastError());
  hipLaunchKernelGGL(gg_k_spread_combine, dim3(h->n_filters), dim3(GG_BLOCK), 0, h->stream, a);
  GG_HIP(hipGetLastError());
  // the spread kernel reuses the tile-partial buffer
  h->decide_valid = false;
  GG_HIP(hipMemcpyAsync(h->pin, h->scratch_f, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  GG_HIP(hipStreamSynchronize(h->stream));
  *out = *(double*)h->pin;
  return 0;
}

int gg_pf_device_resample_systematic(gg_pf* h, uint64_t seed, int32_t u0_mode) {
  if (int rc = check_single(h)) return rc;
  if (ensure_decide(h)) return -1;
  gg_step_args s = default_args();
  if (u0_mode == 0) {
    s.u0_mode = 2;
    const int n = h->filt_h[0].n;
    const int32_t den = (int32_t)(uint32_t)((uint64_t)1000000u * (uint64_t)(uint32_t)n);
    s.u0_value = (double)(seed % 1000000ull) / (double)den;
  } else {
    s.u0_mode = 1;
  }
  if (s.u0_mode == 1) {
    // fresh draw from the filter's stream, evaluated on the host with the same
    // Philox as the device
    gg_dyn d;
    if (read_dyn(h, 0, &d)) return -1;
    const gg_u32x4 r = gg_draw(seed, 0u, (uint32_t)d.resample_count, GG_PURPOSE_U0, 0u);
    s.u0_mode = 2;
    s.u0_value = gg_u01d(r.v[0], r.v[1]) - 1.1102230246251565e-16;
  }
  if (launch_sysres(h, s, 1, 0)) return -1;
  h->decide_valid = false;
  h->ess_valid = true;
  h->est_valid = true;
  return 0;
}

int gg_pf_device_resample_megopolis(gg_pf* h, int32_t n_iterations, uint64_t seed) {
  if (int rc = check_single(h)) return rc;
  if (n_iterations < 0) return gg_fail(-2, "n_iterations must be >= 0");
  if (ensure_decide(h)) return -1;
  gg_step_args s = default_args();
  int fb = 0;
  if (n_iterations == 0) {
    if (launch_gather(h, s, 1, 2, 0)) return -1;
  } else {
    if (launch_mego_ref(h, n_iterations, 1, seed, 1, &fb)) return -1;
    if (launch_gather(h, s, 1, 1, fb)) return -1;
  }
  h->decide_valid = false;
  h->ess_valid = true;
  h->est_valid = true;
  return 0;
}

int gg_pf_device_resample_megopolis_chesser(gg_pf* h, int32_t n_iterations) {
  if (int rc = check_single(h)) return rc;
  if (n_iterations < 0) return gg_fail(-2, "n_iterations must be >= 0");
  if (ensure_decide(h)) return -1;
  gg_step_args s = default_args();
  if (launch_mego(h, n_iterations, 1)) return -1;
  h->decide_valid = false;
  h->ess_valid = true;
  h->est_valid = true;
  return 0;
}

int gg_pf_device_estimate_full(gg_pf* h, double* out) {
  if (int rc = check_single(h)) return rc;
  if (!h->est_valid && ensure_decide(h)) return -1;
  gg_dyn d;
  if (read_dyn(h, 0, &d)) return -1;
  for (int k = 0; k < h->D; ++k) out[k] = d.est[k];
  return 0;
}

int gg_pf_device_estimate(gg_pf* h, double* out) {
  double tmp[GG_MAX_D];
  if (int rc = gg_pf_device_estimate_full(h, tmp)) return rc;
  for (int k = 0; k < 4; ++k) out[k] = tmp[k];
  return 0;
}

static int export_rows(gg_pf* h, int f, int layout, double* out) {
  const int n = h->filt_h[f].n;
  const int ncol = (layout == 0) ? 4 : (layout == 1 ? 16 : h->D);
  const size_t bytes = sizeof(double) * (size_t)n * ncol;
  if (ensure_io(h, bytes)) return -1;
  const gg_arena a = arena_of(h);
  const int grid = (n + GG_BLOCK - 1) / GG_BLOCK;
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_export<NI>), dim3(grid), dim3(GG_BLOCK),
                                                0, h->stream, a, f, layout, h->io_tmp));
  GG_HIP(hipGetLastError());
  GG_HIP(hipMemcpyAsync(out, h->io_tmp, bytes, hipMemcpyDeviceToHost, h->stream));
  GG_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int gg_pf_device_record(gg_pf* h, double* states_dev, double* log_weights_dev) {
  if (int rc = check_single(h)) return rc;
  if (!states_dev || !log_weights_dev) return gg_fail(-2, "null device pointer");
  const int n = h->filt_h[0].n;
  const gg_arena a = arena_of(h);
  const int grid = (n + GG_BLOCK - 1) / GG_BLOCK;
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_export<NI>), dim3(grid), dim3(GG_BLOCK),
                                                0, h->stream, a, 0, 0, states_dev));
  GG_HIP(hipGetLastError());
  hipLaunchKernelGGL(gg_k_export_lw, dim3(grid), dim3(GG_BLOCK), 0, h->stream, a, 0,
                     log_weights_dev);
  GG_HIP(hipGetLastError());
  return 0;
}

int gg_pf_device_get_particles(gg_pf* h, double* out) {
  if (int rc = check_single(h)) return rc;
  return export_rows(h, 0, 0, out);
}

int gg_pf_device_get_particles_full(gg_pf* h, double* out) {
  if (int rc = check_single(h)) return rc;
  return export_rows(h, 0, 2, out);
}

int gg_pf_device_get_particle_states(gg_pf* h, double* out) {
  if (int rc = check_single(h)) return rc;
  return export_rows(h, 0, 1, out);
}

int gg_pf_device_set_particle_states(gg_pf* h, const double* in) {
  if (int rc = check_single(h)) return rc;
  const int n = h->filt_h[0].n;
  gg_dyn d;
  if (read_dyn(h, 0, &d)) return -1;
  // new anchor: mean position and clock bias (binary64, host)
  double m[4] = {0, 0, 0, 0};
  bool vel_uniform = true, cov_uniform = true;
  for (int i = 0; i < n; ++i) {
    const double* r = in + (size_t)i * 16;
    for (int k = 0; k < 4; ++k) m[k] += r[k];
    for (int k = 4; k < 7; ++k) vel_uniform = vel_uniform && (r[k] == in[k]);
    for (int k = 7; k < 16; ++k) cov_uniform = cov_uniform && (r[k] == in[k]);
  }
  for (int k = 0; k < 4; ++k) m[k] /= (double)n;
  if ((!vel_uniform || !cov_uniform) && !h->has_vel)
    return gg_fail(-2, "non-uniform velocity states need velocity storage (with_velocity)");
  d.anchor[0] = m[0];
  d.anchor[1] = m[1];
  d.anchor[2] = m[2];
  d.cb_anchor = m[3];
  for (int k = 0; k < GG_MAX_D; ++k) d.mean[k] = 0.0;  // the offsets are centred on the new anchor
  const double* R = d.R;
  d.vel_u[0] = R[0] * in[4] + R[1] * in[5] + R[2] * in[6];
  d.vel_u[1] = R[3] * in[4] + R[4] * in[5] + R[5] * in[6];
  d.vel_u[2] = R[6] * in[4] + R[7] * in[5] + R[8] * in[6];
  {
    double Sl[9];
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) {
        double acc = 0.0;
        for (int u = 0; u < 3; ++u)
          for (int v = 0; v < 3; ++v) acc += R[p * 3 + u] * in[7 + u * 3 + v] * R[q * 3 + v];
        Sl[p * 3 + q] = acc;
      }
    d.vcov_u[0] = Sl[0];
    d.vcov_u[1] = Sl[1];
    d.vcov_u[2] = Sl[2];
    d.vcov_u[3] = Sl[4];
    d.vcov_u[4] = Sl[5];
    d.vcov_u[5] = Sl[8];
  }
  d.vel_mode = !cov_uniform ? 2 : (!vel_uniform ? 1 : 0);
  d.lw_zero = 1;
  d.ess = (double)n;
  d.flag = 0;
  d.est_valid = 0;
  d.cnt_a = d.cnt_b = d.cnt_c = 0;
  d.aos_only = 0;
  if (write_dyn(h, 0, d)) return -1;
  const size_t bytes = sizeof(double) * (size_t)n * 16;
  if (ensure_io(h, bytes)) return -1;
  GG_HIP(hipMemcpyAsync(h->io_tmp, in, bytes, hipMemcpyHostToDevice, h->stream));
  const gg_arena a = arena_of(h);
  const int grid = (n + GG_BLOCK - 1) / GG_BLOCK;
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_import<NI>), dim3(grid), dim3(GG_BLOCK),
                                                0, h->stream, a, 0, h->io_tmp));
  GG_HIP(hipGetLastError());
  GG_HIP(hipStreamSynchronize(h->stream));
  h->decide_valid = false;
  h->ess_valid = false;
  h->est_valid = false;
  return 0;
}

int gg_pf_device_get_positions(gg_pf* h, double* px, double* py, double* pz, double* pcb) {
  if (int rc = check_single(h)) return rc;
  if (!px || !py || !pz || !pcb) return gg_fail(-2, "null array");
  const int n = h->filt_h[0].n;
  const size_t bytes = sizeof(double) * (size_t)n;
  if (ensure_io(h, 4 * bytes)) return -1;
  double* d = h->io_tmp;
  const gg_arena a = arena_of(h);
  const int grid = (n + GG_BLOCK - 1) / GG_BLOCK;
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_export_soa<NI>), dim3(grid), dim3(GG_BLOCK),
                                                0, h->stream, a, 0, d, d + n, d + 2 * (size_t)n,
                                                d + 3 * (size_t)n));
  GG_HIP(hipGetLastError());
  double* outs[4] = {px, py, pz, pcb};
  for (int k = 0; k < 4; ++k)
    GG_HIP(hipMemcpyAsync(outs[k], d + (size_t)k * n, bytes, hipMemcpyDeviceToHost, h->stream));
  GG_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int gg_pf_device_set_positions(gg_pf* h, const double* px, const double* py, const double* pz,
                               const double* pcb) {
  if (int rc = check_single(h)) return rc;
  if (!px || !py || !pz || !pcb) return gg_fail(-2, "null array");
  const int n = h->filt_h[0].n;
  gg_dyn d;
  if (read_dyn(h, 0, &d)) return -1;
  // new anchor: the mean of an evenly strided sample of at most 4096 particles
  // (binary64); the offsets from it keep binary32 resolution for any cloud the
  // filter tracks, without an O(N) host pass
  const int stride = std::max(1, n / 4096);
  double m[4] = {0, 0, 0, 0};
  int cnt = 0;
  for (int i = 0; i < n; i += stride, ++cnt) {
    m[0] += px[i];
    m[1] += py[i];
    m[2] += pz[i];
    m[3] += pcb[i];
  }
  for (int k = 0; k < 4; ++k) m[k] /= (double)cnt;
  d.anchor[0] = m[0];
  d.anchor[1] = m[1];
  d.anchor[2] = m[2];
  d.cb_anchor = m[3];
  for (int k = 0; k < GG_MAX_D; ++k) d.mean[k] = 0.0;
  d.vel_mode = 0;
  d.lw_zero = 1;
  d.ess = (double)n;
  d.flag = 0;
  d.est_valid = 0;
  if (write_dyn(h, 0, d)) return -1;
  const size_t bytes = sizeof(double) * (size_t)n;
  if (ensure_io(h, 4 * bytes)) return -1;
  double* dv = h->io_tmp;
  const double* ins[4] = {px, py, pz, pcb};
  for (int k = 0; k < 4; ++k)
    GG_HIP(hipMemcpyAsync(dv + (size_t)k * n, ins[k], bytes, hipMemcpyHostToDevice, h->stream));
  const gg_arena a = arena_of(h);
  const int grid = (n + GG_BLOCK - 1) / GG_BLOCK;
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_import_soa<NI>), dim3(grid), dim3(GG_BLOCK),
                                                0, h->stream, a, 0, dv, dv + n, dv + 2 * (size_t)n,
                                                dv + 3 * (size_t)n));
  GG_HIP(hipGetLastError());
  GG_HIP(hipStreamSynchronize(h->stream));
  h->decide_valid = false;
  h->ess_valid = false;
  h->est_valid = false;
  return 0;
}

int gg_pf_device_get_log_weights(gg_pf* h, double* out) {
  if (int rc = check_single(h)) return rc;
  return gg_pfb_get_log_weights(h, 0, out);
}

int gg_pf_device_set_log_weights(gg_pf* h, const double* in) {
  if (int rc = check_single(h)) return rc;
  const int n = h->filt_h[0].n;
  const size_t bytes = sizeof(float) * (size_t)n;
  if (ensure_pinned(h, std::max(h->pin_bytes, bytes))) return -1;
  float* pf = (float*)h->pin;
  for (int i = 0; i < n; ++i) pf[i] = (float)in[i];
  GG_HIP(hipMemcpyAsync(h->lw + h->filt_h[0].p_off, pf, bytes, hipMemcpyHostToDevice, h->stream));
  GG_HIP(hipEventRecord(h->pin_done, h->stream));
  hipLaunchKernelGGL(gg_k_clear_lw_zero, dim3(1), dim3(64), 0, h->stream, h->dyn, 0);
  GG_HIP(hipGetLastError());
  GG_HIP(hipStreamSynchronize(h->stream));
  h->decide_valid = false;
  h->ess_valid = false;
  h->est_valid = false;
  return 0;
}

int gg_pf_device_get_resample_ancestors(gg_pf* h, int32_t* out) {
  if (int rc = check_single(h)) return rc;
  GG_HIP(hipMemcpyAsync(out, h->anc + h->filt_h[0].p_off, sizeof(int32_t) * h->filt_h[0].n,
                        hipMemcpyDeviceToHost, h->stream));
  GG_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int gg_pf_device_shift_clock_bias(gg_pf* h, double shift) {
  if (int rc = check_single(h)) return rc;
  hipLaunchKernelGGL(gg_k_frame_shift, dim3(1), dim3(64), 0, h->stream, h->dyn, 0, 0.0, 0.0, 0.0,
                     shift);
  GG_HIP(hipGetLastError());
  return 0;
}

int gg_pf_device_shift_position(gg_pf* h, double dx, double dy, double dz) {
  if (int rc = check_single(h)) return rc;
  hipLaunchKernelGGL(gg_k_frame_shift, dim3(1), dim3(64), 0, h->stream, h->dyn, 0, dx, dy, dz,
                     0.0);
  GG_HIP(hipGetLastError());
  return 0;
}

int gg_pf_device_position_update(gg_pf* h, double ref_x, double ref_y, double ref_z,
                                 double sigma_pos) {
  if (int rc = check_single(h)) return rc;
  const int n = h->filt_h[0].n;
  const gg_arena a = arena_of(h);
  const int grid = (n + GG_BLOCK - 1) / GG_BLOCK;
  const double inv_s2 = 1.0 / (sigma_pos * sigma_pos);
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_position_update<NI>), dim3(grid),
                                                dim3(GG_BLOCK), 0, h->stream, a, 0, ref_x, ref_y,
                                                ref_z, inv_s2));
  GG_HIP(hipGetLastError());
  hipLaunchKernelGGL(gg_k_clear_lw_zero, dim3(1), dim3(64), 0, h->stream, h->dyn, 0);
  GG_HIP(hipGetLastError());
  h->decide_valid = false;
  h->ess_valid = false;
  h->est_valid = false;
  return 0;
}

// ---------------------------------------------------------------------------
// batched API
// ---------------------------------------------------------------------------
int gg_pfb_initialize(gg_pf* h, const double* init, const uint64_t* seeds, double spread_pos,
                      double spread_cb, double spread_isb) {
  if (!h) return gg_fail(-2, "null handle");
  GG_HIP(hipStreamSynchronize(h->stream));
  for (int f = 0; f < h->n_filters; ++f) h->filt_h[f].seed = seeds[f];
  GG_HIP(hipMemcpy(h->filt, h->filt_h.data(), sizeof(gg_filter) * h->n_filters,
                   hipMemcpyHostToDevice));
  std::vector<gg_dyn> dv(h->n_filters);
  for (int f = 0; f < h->n_filters; ++f) {
    gg_dyn& d = dv[f];
    std::memset(&d, 0, sizeof(d));
    const double* x = init + (size_t)f * h->D;
    d.anchor[0] = x[0];
    d.anchor[1] = x[1];
    d.anchor[2] = x[2];
    d.cb_anchor = x[3];
    for (int k = 0; k < h->n_isb; ++k) d.isb_anchor[k] = x[4 + k];
    enu_rotation(d.anchor, d.R);
    d.lw_zero = 1;
    d.ess = (double)h->filt_h[f].n;  // uniform weights
  }
  GG_HIP(hipMemcpy(h->dyn, dv.data(), sizeof(gg_dyn) * h->n_filters, hipMemcpyHostToDevice));
  const gg_arena a = arena_of(h);
  GG_DISPATCH_ISB(h->n_isb, hipLaunchKernelGGL((gg_k_init<NI>), dim3(h->n_tiles),
                                                dim3(GG_BLOCK), 0, h->stream, a,
                                                (float)spread_pos, (float)spread_cb,
                                                (float)spread_isb, 0.0f));
  GG_HIP(hipGetLastError());
  GG_HIP(hipStreamSynchronize(h->stream));
  h->initialized = true;
  h->decide_valid = h->ess_valid = h->est_valid = false;
  return 0;
}

int gg_pfb_load_epochs(gg_pf* h, int32_t n_slots, const double* sat_xyz, const double* pr,
                       const double* w, const int32_t* sys, const int32_t* n_sat,
                       const double* vel) {
  if (!h) return gg_fail(-2, "null handle");
  if (n_slots <= 0 || n_slots > h->n_slots) return gg_fail(-2, "n_slots exceeds the handle's slots");
  const int rows = h->sat_rows, F = h->n_filters;
  for (int s = 0; s < n_slots; ++s)
    for (int f = 0; f < F; ++f) {
      const int ns = n_sat[(size_t)s * F + f];
      if (ns < 0 || ns > h->filt_h[f].sat_cap)
        return gg_fail(-2, "n_sat exceeds the filter's sat_cap");
    }
  std::vector<double> hd((size_t)n_slots * 5 * rows, 0.0);
  std::vector<int32_t> hi((size_t)n_slots * (rows + F), 0);
  for (int s = 0; s < n_slots; ++s) {
    double* d = hd.data() + (size_t)s * 5 * rows;
    int32_t* ii = hi.data() + (size_t)s * (rows + F);
    for (int r = 0; r < rows; ++r) {
      const size_t src = (size_t)s * rows + r;
      d[r * 3 + 0] = sat_xyz[src * 3 + 0];
      d[r * 3 + 1] = sat_xyz[src * 3 + 1];
      d[r * 3 + 2] = sat_xyz[src * 3 + 2];
      d[3 * rows + r] = pr[src];
      d[4 * rows + r] = w ? w[src] : 1.0;
      ii[r] = sys ? sys[src] : 0;
    }
    for (int f = 0; f < F; ++f) ii[rows + f] = n_sat[(size_t)s * F + f];
  }
  GG_HIP(hipStreamSynchronize(h->stream));
  GG_HIP(hipMemcpy(h->satd, hd.data(), hd.size() * sizeof(double), hipMemcpyHostToDevice));
  GG_HIP(hipMemcpy(h->sati, hi.data(), hi.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  if (vel) {
    GG_HIP(hipMemcpy(h->vel_guide, vel, sizeof(double) * 3 * F * n_slots, hipMemcpyHostToDevice));
  } else {
    GG_HIP(hipMemsetAsync(h->vel_guide, 0, sizeof(double) * 3 * F * n_slots, h->stream));
  }
  return 0;
}

int gg_pfb_step(gg_pf* h, const gg_step_args* args, int32_t resampler, int32_t n_iter) {
  if (!h || !args) return gg_fail(-2, "null argument");
  if (!h->initialized) return gg_fail(-3, "particle filter not initialized");
  if (args->n_slots <= 0 || args->n_slots > h->n_slots) return gg_fail(-2, "bad n_slots");
  if (enqueue_step(h, *args, resampler, n_iter)) return -1;
  h->decide_valid = h->ess_valid = h->est_valid = false;
  return 0;
}

int gg_pfb_step_part(gg_pf* h, const gg_step_args* args, int32_t resampler, int32_t n_iter,
                     int32_t part) {
  if (!h || !args) return gg_fail(-2, "null argument");
  if (!h->initialized) return gg_fail(-3, "particle filter not initialized");
  if (args->n_slots <= 0 || args->n_slots > h->n_slots) return gg_fail(-2, "bad n_slots");
  if (part != 0 && part != 1) return gg_fail(-2, "part must be 0 or 1");
  if (enqueue_step_part(h, *args, resampler, n_iter, part)) return -1;
  h->decide_valid = h->ess_valid = h->est_valid = false;
  return 0;
}

}  // extern "C"

namespace {

// `steps` consecutive steps captured into *graph on a private stream, so that
// a caller's stream (gg_pf_set_stream) never holds a capture; the graph
// replays on whichever stream is in use.  The handle's stream is idle here.
int capture_steps(gg_pf* h, const gg_step_args& s, int resampler, int n_iter, int steps,
                  hipGraph_t* graph) {
  *graph = nullptr;
  // allocations cannot happen inside the capture
  if (s.lik == GG_LIK_LEARNED && s.learned_mode == 1 && ensure_table(h)) return -1;
  hipStream_t cs = nullptr;
  GG_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipStream_t run = h->stream;
  h->stream = cs;
  int rc = 0;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int k = 0; k < steps && !rc; ++k) rc = enqueue_step(h, s, resampler, n_iter);
    e = hipStreamEndCapture(cs, &g);
  }
  h->stream = run;
  (void)hipStreamDestroy(cs);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return -1;
  }
  GG_HIP(e);
  *graph = g;
  return 0;
}

}  // namespace

extern "C" {

int gg_pfb_capture(gg_pf* h, const gg_step_args* args, int32_t resampler, int32_t n_iter,
                   int32_t steps) {
  if (!h || !args) return gg_fail(-2, "null argument");
  if (!h->initialized) return gg_fail(-3, "particle filter not initialized");
  if (args->n_slots <= 0 || args->n_slots > h->n_slots) return gg_fail(-2, "bad n_slots");
  GG_HIP(hipStreamSynchronize(h->stream));
  if (h->graph_exec) {
    GG_HIP(hipGraphExecDestroy(h->graph_exec));
    h->graph_exec = nullptr;
  }
  if (h->graph) {
    GG_HIP(hipGraphDestroy(h->graph));
    h->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  if (capture_steps(h, *args, resampler, n_iter, steps, &g)) return -1;
  h->graph = g;
  GG_HIP(hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
  return 0;
}

int gg_pfb_replay(gg_pf* h, int32_t n_replays) {
  if (!h || !h->graph_exec) return gg_fail(-2, "no captured graph");
  for (int k = 0; k < n_replays; ++k) GG_HIP(hipGraphLaunch(h->graph_exec, h->stream));
  h->decide_valid = h->ess_valid = h->est_valid = false;
  return 0;
}

int gg_pfb_time_replays(gg_pf* h, int32_t n_replays, float* ms_total) {
  if (!h || !h->graph_exec) return gg_fail(-2, "no captured graph");
  GG_HIP(hipEventRecord(h->ev0, h->stream));
  for (int k = 0; k < n_replays; ++k) GG_HIP(hipGraphLaunch(h->graph_exec, h->stream));
  GG_HIP(hipEventRecord(h->ev1, h->stream));
  GG_HIP(hipEventSynchronize(h->ev1));
  GG_HIP(hipEventElapsedTime(ms_total, h->ev0, h->ev1));
  h->decide_valid = h->ess_valid = h->est_valid = false;
  return 0;
}

int gg_pfb_time_replays_each(gg_pf* h, int32_t n_replays, float* ms_each) {
  if (!h || !h->graph_exec) return gg_fail(-2, "no captured graph");
  if (n_replays <= 0 || !ms_each) return gg_fail(-2, "bad n_replays");
  std::vector<hipEvent_t> ev((size_t)n_replays + 1, nullptr);
  int rc = 0;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) rc = gg_fail(-1, "hipEventCreate failed");
  for (int k = 0; rc == 0 && k < n_replays; ++k) {
    if (hipEventRecord(ev[k], h->stream) != hipSuccess ||
        hipGraphLaunch(h->graph_exec, h->stream) != hipSuccess)
      rc = gg_fail(-1, "graph replay failed");
  }
  if (rc == 0 && hipEventRecord(ev[n_replays], h->stream) != hipSuccess)
    rc = gg_fail(-1, "hipEventRecord failed");
  if (rc == 0 && hipEventSynchronize(ev[n_replays]) != hipSuccess)
    rc = gg_fail(-1, "hipEventSynchronize failed");
  for (int k = 0; rc == 0 && k < n_replays; ++k)
    if (hipEventElapsedTime(&ms_each[k], ev[k], ev[k + 1]) != hipSuccess)
      rc = gg_fail(-1, "hipEventElapsedTime failed");
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  h->decide_valid = h->ess_valid = h->est_valid = false;
  return rc;
}

}  // extern "C"

namespace {

// One step in its launch groups: weight | systematic resampling (decision,
// CDF, gather, finalize in one launch) or the Megopolis decision | Megopolis
// chains and gather (none for systematic); ends[g] = the kernel-timing event
// index after group g.
int enqueue_staged_step(gg_pf* h, const gg_step_args& s, int resampler, int n_iter,
                        size_t (&ends)[3]) {
  bool aos = resampler == 1;
  int rc = launch_weight(h, s, true, false, &aos);
  ends[0] = h->kt_next;
  if (!rc) {
    if (resampler == 0) {
      rc = launch_sysres(h, s, 0, 1);
      ends[1] = h->kt_next;
    } else {
      rc = launch_decide(h, s, 3);
      ends[1] = h->kt_next;
      if (!rc) rc = enqueue_resample_after_decide(h, s, resampler, n_iter, aos);
    }
  }
  ends[2] = h->kt_next;
  return rc ? gg_fail(-1, "profiled step launch failed") : 0;
}

// n_steps eager steps with every kernel timed at its own start and end
// (hipExtLaunchKernelGGL events), each step waited for.  A group's time runs
// from its first kernel's start to its last kernel's end.  Per-group means
// over the steps in which some filter resampled (r = 0) and over the others.
int profile_staged(gg_pf* h, const gg_step_args& s, int resampler, int n_iter, int n_steps,
                   float* stage_ms, int32_t* counts) {
  // launches per step: weight (+ table), decision or CDF, up to n_iter + 2
  // resampling kernels and the gather
  const size_t n_ev = 2 * (size_t)(6 + std::max(n_iter, 0));
  std::vector<hipEvent_t> ev(n_ev, nullptr);
  int rc = 0;
  for (auto& e : ev)
    if (!rc && hipEventCreate(&e) != hipSuccess) rc = gg_fail(-1, "hipEventCreate failed");
  double acc[2][3] = {{0, 0, 0}, {0, 0, 0}};
  int cnt[2] = {0, 0};
  std::vector<gg_dyn> dv(h->n_filters);
  for (int k = 0; rc == 0 && k < n_steps; ++k) {
    size_t ends[3] = {0, 0, 0};
    h->kt_ev = ev;
    h->kt_next = 0;
    rc = enqueue_staged_step(h, s, resampler, n_iter, ends);
    h->kt_ev.clear();
    h->kt_next = 0;
    if (rc) break;
    if (ends[2] == 0 || ends[0] == 0 || ends[1] == ends[0] || ends[2] > n_ev) {
      rc = gg_fail(-1, "profiled step: a launch group ran without timing events");
      break;
    }
    if (hipEventSynchronize(ev[ends[2] - 1]) != hipSuccess ||
        hipMemcpy(dv.data(), h->dyn, sizeof(gg_dyn) * h->n_filters, hipMemcpyDeviceToHost) !=
            hipSuccess) {
      rc = gg_fail(-1, "profiled step failed");
      break;
    }
    int any = 0;
    for (const auto& d : dv) any |= d.did_resample;
    const int r = any ? 0 : 1;
    cnt[r] += 1;
    size_t begin = 0;
    for (int st = 0; st < 3 && !rc; ++st) {
      float ms = 0.0f;  // a group without launches (the systematic step's third) takes 0
      if (ends[st] > begin && hipEventElapsedTime(&ms, ev[begin], ev[ends[st] - 1]) != hipSuccess)
        rc = gg_fail(-1, "hipEventElapsedTime failed");
      acc[r][st] += ms;
      begin = ends[st];
    }
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  for (int r = 0; r < 2; ++r)
    for (int st = 0; st < 3; ++st)
      stage_ms[r * 3 + st] = cnt[r] ? (float)(acc[r][st] / cnt[r]) : 0.0f;
  counts[0] = cnt[0];
  counts[1] = cnt[1];
  h->decide_valid = h->ess_valid = h->est_valid = false;
  return rc;
}

}  // namespace

extern "C" {

int gg_pfb_profile_steps(gg_pf* h, const gg_step_args* args, int32_t resampler, int32_t n_iter,
                         int32_t n_steps, float* stage_ms, int32_t* counts) {
  if (!h || !args || !stage_ms || !counts) return gg_fail(-2, "null argument");
  if (!h->initialized) return gg_fail(-3, "particle filter not initialized");
  if (args->n_slots <= 0 || args->n_slots > h->n_slots) return gg_fail(-2, "bad n_slots");
  return profile_staged(h, *args, resampler, n_iter, n_steps, stage_ms, counts);
}

int gg_pfb_time_stage(gg_pf* h, const gg_step_args* args, int32_t stage, int32_t n_iter,
                      int32_t reps, float* ms_total) {
  if (!h || !args) return gg_fail(-2, "null argument");
  if (!h->initialized) return gg_fail(-3, "particle filter not initialized");
  GG_HIP(hipEventRecord(h->ev0, h->stream));
  for (int k = 0; k < reps; ++k) {
    int rc = 0;
    switch (stage) {
      case 0: rc = launch_weight(h, *args, true, false); break;
      case 1: rc = launch_weight(h, *args, false, false); break;
      case 2:
      case 3: rc = launch_sysres(h, *args, 1, 1); break;
      case 4: rc = launch_mego(h, n_iter, 1); break;
      case 5: rc = launch_gather(h, *args, 1, 3, 0); break;
      case 6: rc = launch_decide(h, *args, 0); break;
      default: return gg_fail(-2, "unknown stage");
    }
    if (rc) return rc;
  }
  GG_HIP(hipEventRecord(h->ev1, h->stream));
  GG_HIP(hipEventSynchronize(h->ev1));
  GG_HIP(hipEventElapsedTime(ms_total, h->ev0, h->ev1));
  h->decide_valid = h->ess_valid = h->est_valid = false;
  return 0;
}

int gg_pfb_read_dyn(gg_pf* h, gg_dyn* out) {
  if (!h) return gg_fail(-2, "null handle");
  GG_HIP(hipStreamSynchronize(h->stream));
  GG_HIP(hipMemcpy(out, h->dyn, sizeof(gg_dyn) * h->n_filters, hipMemcpyDeviceToHost));
  return 0;
}

int gg_pfb_set_epoch(gg_pf* h, int32_t epoch) {
  if (!h) return gg_fail(-2, "null handle");
  if (epoch < 0) return gg_fail(-2, "epoch must be >= 0");
  GG_HIP(hipStreamSynchronize(h->stream));
  std::vector<gg_dyn> dv(h->n_filters);
  GG_HIP(hipMemcpy(dv.data(), h->dyn, sizeof(gg_dyn) * h->n_filters, hipMemcpyDeviceToHost));
  for (auto& d : dv) d.epoch = epoch;
  GG_HIP(hipMemcpy(h->dyn, dv.data(), sizeof(gg_dyn) * h->n_filters, hipMemcpyHostToDevice));
  return 0;
}

int gg_pfb_get_particles(gg_pf* h, int32_t f, double* out) {
  if (!h) return gg_fail(-2, "null handle");
  if (f < 0 || f >= h->n_filters) return gg_fail(-2, "filter index out of range");
  return export_rows(h, f, 2, out);
}

int gg_pfb_get_log_weights(gg_pf* h, int32_t f, double* out) {
  if (!h) return gg_fail(-2, "null handle");
  if (f < 0 || f >= h->n_filters) return gg_fail(-2, "filter index out of range");
  gg_dyn d;
  if (read_dyn(h, f, &d)) return -1;
  const int n = h->filt_h[f].n;
  if (d.lw_zero) {
    for (int i = 0; i < n; ++i) out[i] = 0.0;
    return 0;
  }
  const size_t bytes = sizeof(float) * (size_t)n;
  if (ensure_pinned(h, std::max(h->pin_bytes, bytes))) return -1;
  GG_HIP(hipMemcpyAsync(h->pin, h->lw + h->filt_h[f].p_off, bytes, hipMemcpyDeviceToHost,
                        h->stream));
  GG_HIP(hipStreamSynchronize(h->stream));
  const float* pf = (const float*)h->pin;
  for (int i = 0; i < n; ++i) out[i] = (double)pf[i];
  return 0;
}

int gg_pfb_get_log_weights_raw(gg_pf* h, int32_t f, float* out) {
  if (!h || !out) return gg_fail(-2, "null argument");
  if (f < 0 || f >= h->n_filters) return gg_fail(-2, "filter index out of range");
  GG_HIP(hipMemcpyAsync(out, h->lw + h->filt_h[f].p_off, sizeof(float) * h->filt_h[f].n,
                        hipMemcpyDeviceToHost, h->stream));
  GG_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int gg_pfb_get_ancestors(gg_pf* h, int32_t f, int32_t* out) {
  if (!h || !out) return gg_fail(-2, "null argument");
  if (f < 0 || f >= h->n_filters) return gg_fail(-2, "filter index out of range");
  GG_HIP(hipMemcpyAsync(out, h->anc + h->filt_h[f].p_off, sizeof(int32_t) * h->filt_h[f].n,
                        hipMemcpyDeviceToHost, h->stream));
  GG_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

int gg_pfb_set_tuning(gg_pf* h, int32_t weight_tpw, int32_t mego_depth) {
  if (!h) return gg_fail(-2, "null handle");
  if (weight_tpw < 1 || weight_tpw > 64) return gg_fail(-2, "weight_tpw must be in [1, 64]");
  if (mego_depth != 1 && mego_depth != 2 && mego_depth != 4 && mego_depth != 8)
    return gg_fail(-2, "mego_depth must be 1, 2, 4 or 8");
  h->weight_tpw = weight_tpw;
  h->mego_depth = mego_depth;
  return 0;
}

int gg_pfb_get_tuning(gg_pf* h, int32_t* weight_tpw, int32_t* mego_depth) {
  if (!h || !weight_tpw || !mego_depth) return gg_fail(-2, "null argument");
  *weight_tpw = h->weight_tpw;
  *mego_depth = h->mego_depth;
  return 0;
}

int gg_pfb_autotune(gg_pf* h, const gg_step_args* args, int32_t resampler, int32_t n_iter,
                    int32_t steps, int32_t reps, int32_t max_cand, int32_t* cand, float* ms,
                    int32_t* n_cand) {
  if (!h || !args || !cand || !ms || !n_cand) return gg_fail(-2, "null argument");
  if (!h->initialized) return gg_fail(-3, "particle filter not initialized");
  if (args->n_slots <= 0 || args->n_slots > h->n_slots) return gg_fail(-2, "bad n_slots");
  if (steps < 1 || reps < 1) return gg_fail(-2, "steps and reps must be >= 1");
  // candidate shapes: tiles per likelihood workgroup (the analytic likelihood
  // with predict is the only launch that takes it) x Megopolis load depth
  // (Chesser's form only)
  std::vector<int> tpws = {h->weight_tpw};
  if (args->lik != GG_LIK_LEARNED) tpws = {1, 2, 4};
  std::vector<int> depths = {h->mego_depth};
  if (resampler == 1) depths = {1, 2, 4, 8};
  const int nc = (int)(tpws.size() * depths.size());
  if (max_cand < nc) return gg_fail(-2, "max_cand below the candidate count");
  GG_HIP(hipStreamSynchronize(h->stream));
  // every trial starts from the same particles, weights and records, and the
  // handle is left as it was found (the step's intermediate buffers aside)
  const size_t b_state = sizeof(float) * 2 * (size_t)h->ncomp * h->P;
  const size_t b_lw = sizeof(float) * (size_t)h->P;
  const size_t b_anc = sizeof(int32_t) * 2 * (size_t)h->P;
  const size_t b_dyn = sizeof(gg_dyn) * (size_t)h->n_filters;
  char* snap = nullptr;
  GG_HIP(hipMalloc((void**)&snap, b_state + b_lw + b_anc + b_dyn));
  struct Buf { void* dev; size_t off, bytes; };
  const Buf bufs[] = {{h->state, 0, b_state}, {h->lw, b_state, b_lw},
                      {h->anc, b_state + b_lw, b_anc}, {h->dyn, b_state + b_lw + b_anc, b_dyn}};
  auto copy_all = [&](bool save) {
    for (const Buf& b : bufs) {
      const hipError_t e = save ? hipMemcpyAsync(snap + b.off, b.dev, b.bytes,
                                                 hipMemcpyDeviceToDevice, h->stream)
                                : hipMemcpyAsync(b.dev, snap + b.off, b.bytes,
                                                 hipMemcpyDeviceToDevice, h->stream);
      if (e != hipSuccess) return gg_fail_hip(e, "autotune snapshot copy", __FILE__, __LINE__);
    }
    return 0;
  };
  const int tpw0 = h->weight_tpw, depth0 = h->mego_depth;
  int rc = copy_all(true);
  float best = 0.0f;
  int best_k = -1;
  int k = 0;
  for (int tpw : tpws) {
    for (int depth : depths) {
      if (rc) break;
      h->weight_tpw = tpw;
      h->mego_depth = depth;
      hipGraph_t g = nullptr;
      hipGraphExec_t x = nullptr;
      rc = capture_steps(h, *args, resampler, n_iter, steps, &g);
      if (!rc && hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess)
        rc = gg_fail(-1, "autotune graph instantiation failed");
      // one untimed replay, then the timed ones, each from the snapshot
      if (!rc) rc = copy_all(false);
      if (!rc && hipGraphLaunch(x, h->stream) != hipSuccess) rc = gg_fail(-1, "autotune replay");
      if (!rc) rc = copy_all(false);
      if (!rc && hipEventRecord(h->ev0, h->stream) != hipSuccess) rc = gg_fail(-1, "event");
      for (int r = 0; r < reps && !rc; ++r)
        if (hipGraphLaunch(x, h->stream) != hipSuccess) rc = gg_fail(-1, "autotune replay");
      float t = 0.0f;
      if (!rc && (hipEventRecord(h->ev1, h->stream) != hipSuccess ||
                  hipEventSynchronize(h->ev1) != hipSuccess ||
                  hipEventElapsedTime(&t, h->ev0, h->ev1) != hipSuccess))
        rc = gg_fail(-1, "autotune timing");
      if (x) (void)hipGraphExecDestroy(x);
      if (g) (void)hipGraphDestroy(g);
      if (rc) break;
      // reps graph launches of `steps` steps: ms per step
      const float per = t / (float)(reps * steps);
      cand[2 * k] = tpw;
      cand[2 * k + 1] = depth;
      ms[k] = per;
      if (best_k < 0 || per < best) {
        best = per;
        best_k = k;
      }
      ++k;
    }
  }
  if (!rc) rc = copy_all(false);
  if (hipStreamSynchronize(h->stream) != hipSuccess && !rc) rc = gg_fail(-1, "autotune sync");
  (void)hipFree(snap);
  if (rc || best_k < 0) {
    h->weight_tpw = tpw0;
    h->mego_depth = depth0;
    return rc ? rc : gg_fail(-1, "autotune found no candidate");
  }
  h->weight_tpw = cand[2 * best_k];
  h->mego_depth = cand[2 * best_k + 1];
  *n_cand = k;
  // the trials overwrote the per-tile partials behind the single-filter API's
  // cached decision
  h->decide_valid = false;
  return 0;
}

}  // extern "C"
