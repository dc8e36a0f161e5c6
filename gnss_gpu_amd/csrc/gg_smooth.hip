// Backward trajectory sampling over a stored particle history on gfx950:
// FFBSi index sampling and the genealogy trace (python/gnss_gpu/particle_ffbsi.py).
//
// One workgroup per trajectory walks t = T-1 .. 0 inside the kernel; the
// history (log-weights [T][N], states [T][N][4], binary64) stays in HBM and
// each step streams one epoch of it.  Thread k owns the contiguous particle
// range [k*C, (k+1)*C), C = ceil(N / 256), so the inverse-CDF search is a
// block scan of the 256 range sums followed by a walk through one range.
//
// A step's backward weights are w_i = exp(lp_i - max lp) with
// lp_i = lw[t][i] + log p(x_{t+1} | x_t[i]) (particle_ffbsi.py:81-102; the
// reference's log-normalisation of lw[t] shifts every lp by one constant and
// cancels).  Given the caller's uniform u, the step picks the first i whose
// running sum of w exceeds u * sum(w) -- numpy's Generator.choice(N, p=w):
// cdf = cumsum(p) / cdf[-1], searchsorted(cdf, u, 'right').  The sums run in
// a different order than numpy's cumsum, so an index can differ only when u
// lies within rounding (~1e-16 relative) of a CDF step.
//
// Degenerate steps (a NaN or +inf log-weight, or all -inf) are reported in
// status[] and the walk stops there: the reference draws an integer instead of
// a uniform at such a step, so the host redraws and passes the index in
// forced[] (see _sample in particle_ffbsi.py of this package).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "gg_common_host.h"
#include "gg_device.h"
#include "gnss_gpu_amd.h"

#define GG_FB_BLOCK 256

namespace {

// log p(x_next | x_t) of one predict step (particle_ffbsi.py:14-29)
__device__ __forceinline__ double trans_lp(const double xn[4], const double* __restrict__ xt,
                                           const double v[3], double dt, double isp2,
                                           double isc2) {
  const double2 a = reinterpret_cast<const double2*>(xt)[0];
  const double2 b = reinterpret_cast<const double2*>(xt)[1];
  const double d0 = xn[0] - (a.x + v[0] * dt);
  const double d1 = xn[1] - (a.y + v[1] * dt);
  const double d2 = xn[2] - (b.x + v[2] * dt);
  const double dc = xn[3] - b.y;
  return -0.5 * ((d0 * d0 + d1 * d1) + d2 * d2) * isp2 - 0.5 * (dc * dc) * isc2;
}

struct StepIn {
  double xn[4], v[3], dt, isp2;
  bool trans;
};

__device__ __forceinline__ double step_lp(const double* __restrict__ lw,
                                          const double* __restrict__ X, int64_t row, int i,
                                          const StepIn& s, double isc2) {
  double lp = lw[row + i];
  if (s.trans) lp += trans_lp(s.xn, X + (row + i) * 4, s.v, s.dt, s.isp2, isc2);
  return lp;
}

__global__ __launch_bounds__(GG_FB_BLOCK) void gg_k_ffbsi(
    const double* __restrict__ lw, const double* __restrict__ X, const double* __restrict__ vel,
    const double* __restrict__ dts, const double* __restrict__ sps, double isc2,
    const int32_t* __restrict__ anc, int T, int N, const double* __restrict__ u,
    const int32_t* __restrict__ forced, int64_t* __restrict__ idx, int32_t* __restrict__ status) {
  __shared__ double lds[8 + GG_FB_BLOCK];
  __shared__ int ilds[8];
  __shared__ int sel_sh;
  const int tid = threadIdx.x;
  const int64_t m = blockIdx.x;
  const int C = (N + GG_FB_BLOCK - 1) / GG_FB_BLOCK;
  const int i0 = min(N, tid * C), i1 = min(N, i0 + C);
  int64_t* out = idx + m * T;
  int nxt = 0;
  for (int t = T - 1; t >= 0; --t) {
    const int f = forced[m * T + t];
    if (f >= 0) {  // the caller's integer draw for a degenerate step
      if (tid == 0) out[t] = f;
      nxt = f;
      continue;
    }
    if (anc && t < T - 1) {  // genealogy: follow the resampling ancestor
      nxt = anc[(int64_t)t * N + nxt];
      if (tid == 0) out[t] = nxt;
      continue;
    }
    StepIn s;
    s.trans = (t < T - 1);
    if (s.trans) {
      const double* xr = X + ((int64_t)(t + 1) * N + nxt) * 4;
#pragma unroll
      for (int k = 0; k < 4; ++k) s.xn[k] = xr[k];
#pragma unroll
      for (int k = 0; k < 3; ++k) s.v[k] = vel[(t + 1) * 3 + k];
      s.dt = dts[t + 1];
      s.isp2 = 1.0 / (sps[t + 1] * sps[t + 1]);
    }
    const int64_t row = (int64_t)t * N;
    // pass 1: running max and rescaled sum of exp over this thread's range
    double mk = -HUGE_VAL, sk = 0.0;
    int bad = 0;
    for (int i = i0; i < i1; ++i) {
      const double lp = step_lp(lw, X, row, i, s, isc2);
      if (isnan(lp) || lp == HUGE_VAL) {
        bad = 1;
      } else if (lp > mk) {
        sk = sk * exp(mk - lp) + 1.0;
        mk = lp;
      } else if (lp > -HUGE_VAL) {
        sk += exp(lp - mk);
      }
    }
    bad = gg_block_sum256_i(bad, ilds);
    const double M = gg_block_max256(mk, lds);
    if (bad > 0 || !(M > -HUGE_VAL)) {
      if (tid == 0) status[m] = (t == T - 1) ? -2 : t;
      return;
    }
    const double ck = sk > 0.0 ? sk * exp(mk - M) : 0.0;
    double excl, S;
    const double incl = gg_block_scan256(ck, lds, excl, S);
    const double target = u[m * T + t] * S;
    if (tid == 0) sel_sh = -1;
    __syncthreads();
    if (ck > 0.0 && excl <= target && target < incl) sel_sh = tid;
    __syncthreads();
    int owner = sel_sh;
    if (owner < 0) owner = gg_block_max256_i(ck > 0.0 ? tid : -1, ilds);  // target >= sum
    __syncthreads();
    if (tid == owner) {  // pass 2: walk the owner's range to the crossing
      double run = excl;
      int pick = -1, last = -1;
      for (int i = i0; i < i1; ++i) {
        const double lp = step_lp(lw, X, row, i, s, isc2);
        const double w = lp > -HUGE_VAL ? exp(lp - M) : 0.0;
        if (w > 0.0) last = i;
        run += w;
        if (run > target) {
          pick = i;
          break;
        }
      }
      sel_sh = pick >= 0 ? pick : last;
    }
    __syncthreads();
    // the owner holds the block maximum's range or another with ck > 0, so its
    // walk meets a positive weight and sel_sh >= 0; the clamp keeps the next
    // step's state load inside the history regardless
    nxt = max(sel_sh, 0);
    if (tid == 0) out[t] = nxt;
    __syncthreads();
  }
  if (tid == 0) status[m] = -1;
}

__global__ __launch_bounds__(GG_FB_BLOCK) void gg_k_trans_lp(const double* __restrict__ xn,
                                                             const double* __restrict__ xt, int n,
                                                             const double* __restrict__ v,
                                                             double dt, double isp2, double isc2,
                                                             double* __restrict__ out) {
  const int i = blockIdx.x * GG_FB_BLOCK + threadIdx.x;
  if (i >= n) return;
  const double a[4] = {xn[0], xn[1], xn[2], xn[3]};
  const double vv[3] = {v[0], v[1], v[2]};
  out[i] = trans_lp(a, xt + (int64_t)i * 4, vv, dt, isp2, isc2);
}

}  // namespace

extern "C" {

int gg_ffbsi_sample(const double* log_weights, const double* states, const double* vel,
                    const double* dt, const double* sigma_pos, double sigma_cb,
                    const int32_t* ancestors, int32_t T, int32_t N, int32_t n_traj,
                    const double* u, const int32_t* forced, int64_t* indices, int32_t* status) {
  if (T <= 0 || N <= 0 || n_traj < 0) return gg_fail(-2, "T, N must be positive");
  if (n_traj == 0) return 0;
  if (!log_weights || !u || !forced || !indices || !status) return gg_fail(-2, "null array");
  if (!ancestors && (!states || !vel || !dt || !sigma_pos))
    return gg_fail(-2, "FFBSi needs states, vel, dt and sigma_pos");
  const size_t tn = (size_t)T * N;
  const bool ffbsi = ancestors == nullptr;
  size_t bytes = sizeof(double) * tn + 256;  // log-weights
  if (ffbsi) bytes += sizeof(double) * (tn * 4 + (size_t)T * 5) + 1024;
  else bytes += sizeof(int32_t) * tn + 256;
  bytes += (sizeof(double) + sizeof(int32_t) + sizeof(int64_t)) * (size_t)n_traj * T +
           sizeof(int32_t) * n_traj + 1024;
  char* buf = nullptr;
  GG_HIP(hipMalloc((void**)&buf, bytes));
  size_t off = 0;
  auto take = [&](size_t b) {
    char* p = buf + off;
    off += (b + 255) & ~(size_t)255;
    return p;
  };
  double* d_lw = (double*)take(sizeof(double) * tn);
  double *d_X = nullptr, *d_vel = nullptr, *d_dt = nullptr, *d_sp = nullptr;
  int32_t* d_anc = nullptr;
  if (ffbsi) {
    d_X = (double*)take(sizeof(double) * tn * 4);
    d_vel = (double*)take(sizeof(double) * T * 3);
    d_dt = (double*)take(sizeof(double) * T);
    d_sp = (double*)take(sizeof(double) * T);
  } else {
    d_anc = (int32_t*)take(sizeof(int32_t) * tn);
  }
  double* d_u = (double*)take(sizeof(double) * n_traj * T);
  int32_t* d_f = (int32_t*)take(sizeof(int32_t) * n_traj * T);
  int64_t* d_idx = (int64_t*)take(sizeof(int64_t) * n_traj * T);
  int32_t* d_st = (int32_t*)take(sizeof(int32_t) * n_traj);
  int rc = 0;
  auto chk = [&](hipError_t e, const char* what) {
    if (!rc && e != hipSuccess) rc = gg_fail_hip(e, what, __FILE__, __LINE__);
  };
  chk(hipMemcpy(d_lw, log_weights, sizeof(double) * tn, hipMemcpyHostToDevice), "H2D lw");
  if (ffbsi) {
    chk(hipMemcpy(d_X, states, sizeof(double) * tn * 4, hipMemcpyHostToDevice), "H2D X");
    chk(hipMemcpy(d_vel, vel, sizeof(double) * T * 3, hipMemcpyHostToDevice), "H2D vel");
    chk(hipMemcpy(d_dt, dt, sizeof(double) * T, hipMemcpyHostToDevice), "H2D dt");
    chk(hipMemcpy(d_sp, sigma_pos, sizeof(double) * T, hipMemcpyHostToDevice), "H2D sigma");
  } else {
    chk(hipMemcpy(d_anc, ancestors, sizeof(int32_t) * tn, hipMemcpyHostToDevice), "H2D anc");
  }
  chk(hipMemcpy(d_u, u, sizeof(double) * n_traj * T, hipMemcpyHostToDevice), "H2D u");
  chk(hipMemcpy(d_f, forced, sizeof(int32_t) * n_traj * T, hipMemcpyHostToDevice), "H2D forced");
  if (!rc) {
    hipLaunchKernelGGL(gg_k_ffbsi, dim3(n_traj), dim3(GG_FB_BLOCK), 0, nullptr, d_lw, d_X, d_vel,
                       d_dt, d_sp, 1.0 / (sigma_cb * sigma_cb), d_anc, T, N, d_u, d_f, d_idx,
                       d_st);
    chk(hipGetLastError(), "gg_k_ffbsi");
  }
  chk(hipMemcpy(indices, d_idx, sizeof(int64_t) * n_traj * T, hipMemcpyDeviceToHost), "D2H idx");
  chk(hipMemcpy(status, d_st, sizeof(int32_t) * n_traj, hipMemcpyDeviceToHost), "D2H status");
  const hipError_t fe = hipFree(buf);
  if (!rc && fe != hipSuccess) rc = gg_fail_hip(fe, "hipFree", __FILE__, __LINE__);
  return rc;
}

int gg_ffbsi_sample_device(const double* log_weights, const double* states, const double* vel,
                           const double* dt, const double* sigma_pos, double sigma_cb,
                           const int32_t* ancestors, int32_t T, int32_t N, int32_t n_traj,
                           const double* u, const int32_t* forced, int64_t* indices,
                           int32_t* status, void* hip_stream) {
  if (T <= 0 || N <= 0 || n_traj < 0) return gg_fail(-2, "T, N must be positive");
  if (n_traj == 0) return 0;
  if (!log_weights || !u || !forced || !indices || !status) return gg_fail(-2, "null array");
  if (!ancestors && (!states || !vel || !dt || !sigma_pos))
    return gg_fail(-2, "FFBSi needs states, vel, dt and sigma_pos");
  hipLaunchKernelGGL(gg_k_ffbsi, dim3(n_traj), dim3(GG_FB_BLOCK), 0, (hipStream_t)hip_stream,
                     log_weights, states, vel, dt, sigma_pos, 1.0 / (sigma_cb * sigma_cb),
                     ancestors, T, N, u, forced, indices, status);
  GG_HIP(hipGetLastError());
  return 0;
}

int gg_transition_logpdf(const double* x_next, const double* x_t, int32_t N, const double* vel,
                         double dt, double sigma_pos, double sigma_cb, double* out) {
  if (N < 0) return gg_fail(-2, "negative size");
  if (N == 0) return 0;
  const size_t bytes = sizeof(double) * ((size_t)N * 5 + 8) + 1024;
  char* buf = nullptr;
  GG_HIP(hipMalloc((void**)&buf, bytes));
  double* d_xt = (double*)buf;
  double* d_out = d_xt + (size_t)N * 4;
  double* d_xn = d_out + N;
  double* d_v = d_xn + 4;
  int rc = 0;
  auto chk = [&](hipError_t e, const char* what) {
    if (!rc && e != hipSuccess) rc = gg_fail_hip(e, what, __FILE__, __LINE__);
  };
  chk(hipMemcpy(d_xt, x_t, sizeof(double) * N * 4, hipMemcpyHostToDevice), "H2D x_t");
  chk(hipMemcpy(d_xn, x_next, sizeof(double) * 4, hipMemcpyHostToDevice), "H2D x_next");
  chk(hipMemcpy(d_v, vel, sizeof(double) * 3, hipMemcpyHostToDevice), "H2D vel");
  if (!rc) {
    hipLaunchKernelGGL(gg_k_trans_lp, dim3((N + GG_FB_BLOCK - 1) / GG_FB_BLOCK),
                       dim3(GG_FB_BLOCK), 0, nullptr, d_xn, d_xt, N, d_v, dt,
                       1.0 / (sigma_pos * sigma_pos), 1.0 / (sigma_cb * sigma_cb), d_out);
    chk(hipGetLastError(), "gg_k_trans_lp");
  }
  chk(hipMemcpy(out, d_out, sizeof(double) * N, hipMemcpyDeviceToHost), "D2H out");
  const hipError_t fe = hipFree(buf);
  if (!rc && fe != hipSuccess) rc = gg_fail_hip(fe, "hipFree", __FILE__, __LINE__);
  return rc;
}

}  // extern "C"
