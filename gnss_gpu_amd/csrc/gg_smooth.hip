// Backward trajectory sampling over a stored particle history on gfx950:
// FFBSi index sampling and the genealogy trace (python/gnss_gpu/particle_ffbsi.py).
//
// One workgroup per trajectory walks t = T-1 .. 0 inside the kernel; the
// history (log-weights [T][N], states [T][N][4], binary64) stays in HBM and
// each step streams one epoch of it.  Up to 16384 particles, gg_k_ffbsi_rows
// holds the epoch in registers as rows of 1024 particles (one per thread) and
// locates the crossing by row, wavefront and lane.  Larger populations and the
// genealogy trace use gg_k_ffbsi: thread k owns the contiguous particle range
// [k*C, (k+1)*C), C = ceil(N / 256), so the inverse-CDF search is a block
// scan of the 256 range sums followed by a walk through one range.
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

// log p(x_next | x_t) of one predict step (particle_ffbsi.py:14-29); the
// state x_t arrives as its two halves (x, y) and (z, clock)
__device__ __forceinline__ double trans_lp2(const double xn[4], double2 a, double2 b,
                                            const double v[3], double dt, double isp2,
                                            double isc2) {
  const double d0 = xn[0] - (a.x + v[0] * dt);
  const double d1 = xn[1] - (a.y + v[1] * dt);
  const double d2 = xn[2] - (b.x + v[2] * dt);
  const double dc = xn[3] - b.y;
  return -0.5 * ((d0 * d0 + d1 * d1) + d2 * d2) * isp2 - 0.5 * (dc * dc) * isc2;
}

__device__ __forceinline__ double trans_lp(const double xn[4], const double* __restrict__ xt,
                                           const double v[3], double dt, double isp2,
                                           double isc2) {
  const double2* p = reinterpret_cast<const double2*>(xt);
  return trans_lp2(xn, p[0], p[1], v, dt, isp2, isc2);
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
    if (f < -1 || f >= N) {  // not an index nor -1: reported, nothing out of range written
      if (tid == 0) status[m] = -3;
      return;
    }
    if (f >= 0) {  // the caller's integer draw for a degenerate step
      if (tid == 0) out[t] = f;
      nxt = f;
      continue;
    }
    if (anc && t < T - 1) {  // genealogy: follow the resampling ancestor
      const int a = anc[(int64_t)t * N + nxt];
      if (a < 0 || a >= N) {
        if (tid == 0) status[m] = -3;
        return;
      }
      if (tid == 0) out[t] = a;
      nxt = a;
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

// ---------------------------------------------------------------------------
// FFBSi in particle rows, for N <= GG_FR_BLOCK * GG_FR_MAX_ROWS.  One
// 1024-thread workgroup per trajectory; particle i = j * 1024 + k is element j
// of thread k, so every load of a step is coalesced and the loads of up to 8
// rows are in flight at once, and the step's log-probabilities (and, up to 8
// rows, its weights) stay in registers.  Per step:
//   1. lp_i = lw[t][i] + log p(x_{t+1} | x_t[i]); block max M (degenerate
//      steps reported as in gg_k_ffbsi);
//   2. w_i = exp(lp_i - M); per wavefront and row the sum W[j][w] in the
//      balanced pairwise order (all rows folded together, fold_rows), which
//      equals, bit for bit, the last lane of the Hillis-Steele scan of step 4;
//   3. wavefront 0: row totals R_j = (((W[j][0] + W[j][1]) + W[j][2]) + ...)
//      and row offsets P_{j+1} = P_j + R_j; target = u * P_R; row J is the
//      first with P_{J+1} > target and, inside it, wavefront w* the first
//      whose running end P_J + E_w exceeds the target (E_w the running sum of
//      W[J][.]);
//   4. wavefront w* scans its lanes (Hillis-Steele) and picks the first
//      positive-weight lane with P_J + O_w* + scan > target.
// The running end of w* equals P_{J+1} bit for bit, so the crossing lies in
// that wavefront; if the target is not below the total (u * S rounded up to S)
// the last positive weight is taken, as the sequential walk of gg_k_ffbsi
// does.  Only the order of the sums differs from gg_k_ffbsi.
// ---------------------------------------------------------------------------
#define GG_FR_BLOCK 1024
#define GG_FR_WAVES (GG_FR_BLOCK / 64)
#define GG_FR_MAX_ROWS 16
#define GG_FR_CHUNK 8  // up to 8 rows load at once; 16 rows load 4 at a time

__device__ __forceinline__ double wave_max_f64(double v) {
  GG_BUTTERFLY32(f64, v, gg_max_f64);
  return fmax(v, __shfl_xor(v, 32, 64));
}

// The value of lane ^ (1 << S) (DPP where the pattern exists, else the
// swizzle or permute network).
template <int S>
__device__ __forceinline__ double xor_lane(double v) {
  if constexpr (S == 0) return gg_dpp_f64<0xB1>(v);        // quad_perm [1,0,3,2]
  else if constexpr (S == 1) return gg_dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  else if constexpr (S == 3) return gg_dpp_f64<0x128>(v);  // row_ror 8
  else if constexpr (S == 5) return __shfl_xor(v, 32, 64);
  else {  // ds_swizzle bit mode, xor mask 4 or 16
    constexpr int OFF = 0x1F | ((1 << S) << 10);
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)u, OFF);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_swizzle((int)(uint32_t)(u >> 32), OFF);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
  }
}

// Wavefront sums of the N values a[] of every lane, all rows at once: at level
// S each lane keeps half of its values (by bit S of the lane) and adds the
// partner's copy of that half, so level S pairs the aligned lane blocks of size
// 2^S -- the balanced pairwise tree of the Hillis-Steele scan's last lane, for
// every row.  Afterwards lane L holds, in a[0], the total of row
// sum_s bit_s(L) * (N >> (s + 1)) (a bit reversal of L for L < N).
template <int S, int N>
__device__ __forceinline__ void fold_rows(double* a, int lane) {
  if constexpr (N > 1) {
    constexpr int H = N / 2;
    const bool b = (lane >> S) & 1;
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const double send = b ? a[i] : a[i + H];
      const double keep = b ? a[i + H] : a[i];
      a[i] = keep + xor_lane<S>(send);
    }
    fold_rows<S + 1, H>(a, lane);
  } else if constexpr (S < 6) {
    a[0] = a[0] + xor_lane<S>(a[0]);
    fold_rows<S + 1, 1>(a, lane);
  }
}

template <int N>
__device__ __forceinline__ int folded_row(int lane) {
  int j = 0;
#pragma unroll
  for (int s = 0; (1 << s) < N; ++s) j += ((lane >> s) & 1) * (N >> (s + 1));
  return j;
}

// inclusive scan over the wavefront: x[l] = x[l - d] + x[l] for d = 1 .. 32
__device__ __forceinline__ double wave_scan_f64(double x, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double y = __shfl_up(x, d, 64);
    if (lane >= d) x = y + x;
  }
  return x;
}

// Read-only per-step inputs through the constant address space, so uniform
// reads of them are scalar loads.
typedef const __attribute__((address_space(4))) double* const_f64_ptr;

__device__ __forceinline__ const_f64_ptr const_view(const double* p) { return (const_f64_ptr)p; }

__device__ __forceinline__ double row_weight(double lp, double M) {
  return lp > -HUGE_VAL ? exp(lp - M) : 0.0;
}

struct RowsLds {
  double M[GG_FR_WAVES];
  int bad[GG_FR_WAVES];
  double PJ, Os, target;
  int J, ws, last, sel;
};

// One backward step at epoch t given the pick nxt at t + 1 (ignored at the
// last epoch) and the caller's uniform ut: the pick at t, or -1 when the step
// is degenerate (the same in every thread).
template <int R>
__device__ __forceinline__ int rows_step(const double* __restrict__ lw,
                                         const double* __restrict__ X, const_f64_ptr vel,
                                         const_f64_ptr dts, const_f64_ptr sps, double isc2,
                                         int T, int N, int t, int nxt, double ut,
                                         double (*sW)[GG_FR_WAVES], RowsLds& sh) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  const bool trans = t < T - 1;
  const int64_t row = (int64_t)t * N;
  double xn[4] = {0.0, 0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0}, dtt = 0.0, isp2 = 0.0;
  // 1. log-probabilities and the block max.  Rows are loaded GG_FR_CHUNK at a
  //    time, every load of a chunk in flight at once; row j of the epoch starts
  //    at a uniform address.
  constexpr int CH = R <= GG_FR_CHUNK ? R : GG_FR_CHUNK / 2;
  double lp[R];
  double mloc = -HUGE_VAL;
  int bad = 0;
#pragma unroll
  for (int c0 = 0; c0 < R; c0 += CH) {
    double lwv[CH];
    double2 xa[CH], xb[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int j = c0 + c;
      const int64_t r0 = row + (int64_t)j * GG_FR_BLOCK;
      lwv[c] = 0.0;
      xa[c] = make_double2(0.0, 0.0);
      xb[c] = xa[c];
      if (j * GG_FR_BLOCK + tid < N) {
        lwv[c] = (lw + r0)[tid];
        if (trans) {
          const double2* p = reinterpret_cast<const double2*>(X + r0 * 4);
          xa[c] = p[2 * tid];
          xb[c] = p[2 * tid + 1];
        }
      }
    }
    if (c0 == 0 && trans) {  // the picked state of epoch t + 1, with the first rows
      const double* xr = X + ((int64_t)(t + 1) * N + nxt) * 4;
#pragma unroll
      for (int k = 0; k < 4; ++k) xn[k] = xr[k];
#pragma unroll
      for (int k = 0; k < 3; ++k) v[k] = vel[(t + 1) * 3 + k];
      dtt = dts[t + 1];
      isp2 = 1.0 / (sps[t + 1] * sps[t + 1]);
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int j = c0 + c;
      double x = -HUGE_VAL;
      if (j * GG_FR_BLOCK + tid < N) {
        x = lwv[c];
        if (trans) x += trans_lp2(xn, xa[c], xb[c], v, dtt, isp2, isc2);
        if (isnan(x) || x == HUGE_VAL)
          bad = 1;
        else
          mloc = fmax(mloc, x);
      }
      lp[j] = x;
    }
  }
  const double wm = wave_max_f64(mloc);
  const int wbad = __any(bad);
  if (lane == 0) {
    sh.M[wv] = wm;
    sh.bad[wv] = wbad;
  }
  __syncthreads();
  double M = sh.M[0];
  int anyb = sh.bad[0];
#pragma unroll
  for (int k = 1; k < GG_FR_WAVES; ++k) {
    M = fmax(M, sh.M[k]);
    anyb |= sh.bad[k];
  }
  if (anyb || !(M > -HUGE_VAL)) return -1;
  // 2. weights and their per-wavefront row sums W[j][w]
  // up to 8 rows keep their weights for step 4; 16 rows recompute row J's
  // (the same expression, so the same bits) to leave registers free
  constexpr bool KEEP = R <= GG_FR_CHUNK;
  double f[R], w[KEEP ? R : 1];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    f[j] = row_weight(lp[j], M);
    if constexpr (KEEP) w[j] = f[j];
  }
  fold_rows<0, R>(f, lane);
  if (lane < R) sW[folded_row<R>(lane)][wv] = f[0];
  __syncthreads();
  // 3. wavefront 0: row totals (lane j holds R_j), row offsets, the row J and
  //    the wavefront ws of the crossing
  if (wv == 0) {
    double rj = 0.0;
    if (lane < R) {
      rj = sW[lane][0];
      for (int k = 1; k < GG_FR_WAVES; ++k) rj = rj + sW[lane][k];
    }
    double P[R + 1];
    P[0] = 0.0;
#pragma unroll
    for (int j = 0; j < R; ++j) P[j + 1] = P[j] + __shfl(rj, j, 64);
    const double target = ut * P[R];
    int J = -1;
#pragma unroll
    for (int j = R - 1; j >= 0; --j)
      if (P[j + 1] > target) J = j;
    const bool last = J < 0;  // target >= total: the last positive weight
    if (last) {
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (__shfl(rj, j, 64) > 0.0) J = j;
    }
    J = max(J, 0);
    double PJ = 0.0;
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (j == J) PJ = P[j];
    const double Ww = (lane < GG_FR_WAVES) ? sW[J][lane] : 0.0;
    int ws = -1, wpos = -1;
    double Os = 0.0, Opos = 0.0, O = 0.0;
#pragma unroll
    for (int k = 0; k < GG_FR_WAVES; ++k) {
      const double Wk = __shfl(Ww, k, 64);
      const double E = (k == 0) ? Wk : O + Wk;
      if (ws < 0 && !last && PJ + E > target) {
        ws = k;
        Os = O;
      }
      if (Wk > 0.0) {
        wpos = k;
        Opos = O;
      }
      O = E;
    }
    if (ws < 0) {
      ws = max(wpos, 0);
      Os = Opos;
    }
    if (lane == 0) {
      sh.J = J;
      sh.ws = ws;
      sh.last = last;
      sh.PJ = PJ;
      sh.Os = Os;
      sh.target = target;
    }
  }
  __syncthreads();
  // 4. the crossing inside wavefront ws
  const int ws = sh.ws;
  if (wv == ws) {
    const int J = sh.J;
    const bool last = sh.last;
    const double PJ = sh.PJ, target = sh.target;
    double x = 0.0;
    if constexpr (KEEP) {
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (j == J) x = w[j];
    } else {
      double lpJ = -HUGE_VAL;
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (j == J) lpJ = lp[j];
      x = row_weight(lpJ, M);
    }
    double inc = wave_scan_f64(x, lane);
    if (ws > 0) inc = sh.Os + inc;
    const unsigned long long cand = __ballot(!last && x > 0.0 && PJ + inc > target);
    const unsigned long long pos = __ballot(x > 0.0);
    int pick = 0;
    if (cand)
      pick = __ffsll((long long)cand) - 1;
    else if (pos)
      pick = 63 - __clzll((long long)pos);
    if (lane == 0) sh.sel = min(J * GG_FR_BLOCK + ws * 64 + pick, N - 1);
  }
  __syncthreads();
  return sh.sel;
}

template <int R>
__global__ __launch_bounds__(GG_FR_BLOCK) void gg_k_ffbsi_rows(
    const double* __restrict__ lw, const double* __restrict__ X, const double* __restrict__ vel,
    const double* __restrict__ dts, const double* __restrict__ sps, double isc2, int T, int N,
    const double* __restrict__ u, const int32_t* __restrict__ forced, int64_t* __restrict__ idx,
    int32_t* __restrict__ status) {
  __shared__ double sW[R][GG_FR_WAVES];
  __shared__ RowsLds sh;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int64_t m = blockIdx.x;
  int64_t* out = idx + m * T;
  // The forced draws, uniforms and picks of 64 steps at a time, lane l of a
  // wavefront holding step t0 - l: a step reads its inputs with readlane
  // instead of a memory round trip, and wavefront 0 stores 64 picks at once.
  int fch = -1, och = 0, nxt = 0;
  double uch = 0.0;
  for (int t = T - 1; t >= 0; --t) {
    const int l = (T - 1 - t) & 63;
    if (l == 0) {
      const int tt = t - lane;
      fch = tt >= 0 ? forced[m * T + tt] : -1;
      uch = tt >= 0 ? u[m * T + tt] : 0.0;
    }
    const int f = __builtin_amdgcn_readlane(fch, l);
    if (f < -1 || f >= N) {  // not an index nor -1: reported, nothing out of range written
      if (threadIdx.x == 0) status[m] = -3;
      return;
    }
    int pick;
    if (f >= 0) {  // the caller's integer draw for a degenerate step
      pick = f;
      nxt = f;
    } else {
      const uint64_t ub = __builtin_bit_cast(uint64_t, uch);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(ub >> 32), l);
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)ub, l);
      const double ut = __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
      nxt = rows_step<R>(lw, X, const_view(vel), const_view(dts), const_view(sps), isc2, T, N,
                         t, nxt, ut, sW, sh);
      if (nxt < 0) {
        if (threadIdx.x == 0) status[m] = (t == T - 1) ? -2 : t;
        return;
      }
      pick = nxt;
    }
    if (lane == l) och = pick;
    if ((l == 63 || t == 0) && wv == 0 && lane <= l) out[t + l - lane] = och;
  }
  if (threadIdx.x == 0) status[m] = -1;
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

// FFBSi up to 16384 particles in rows; larger populations and the genealogy
// trace in the per-thread-range kernel
hipError_t launch_backward(const double* lw, const double* X, const double* vel, const double* dt,
                           const double* sp, double sigma_cb, const int32_t* anc, int T, int N,
                           int n_traj, const double* u, const int32_t* forced, int64_t* idx,
                           int32_t* status, hipStream_t st) {
  const double isc2 = 1.0 / (sigma_cb * sigma_cb);
  const int rows = (N + GG_FR_BLOCK - 1) / GG_FR_BLOCK;
  if (anc || rows > GG_FR_MAX_ROWS) {
    hipLaunchKernelGGL(gg_k_ffbsi, dim3(n_traj), dim3(GG_FB_BLOCK), 0, st, lw, X, vel, dt, sp,
                       isc2, anc, T, N, u, forced, idx, status);
  } else {
#define GG_FR_LAUNCH(RR)                                                                      \
  hipLaunchKernelGGL(gg_k_ffbsi_rows<RR>, dim3(n_traj), dim3(GG_FR_BLOCK), 0, st, lw, X, vel, \
                     dt, sp, isc2, T, N, u, forced, idx, status)
    if (rows <= 1)
      GG_FR_LAUNCH(1);
    else if (rows <= 2)
      GG_FR_LAUNCH(2);
    else if (rows <= 4)
      GG_FR_LAUNCH(4);
    else if (rows <= 8)
      GG_FR_LAUNCH(8);
    else
      GG_FR_LAUNCH(16);
#undef GG_FR_LAUNCH
  }
  return hipGetLastError();
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
  for (size_t k = 0; k < (size_t)n_traj * T; ++k)
    if (forced[k] < -1 || forced[k] >= N) return gg_fail(-2, "forced index outside [-1, N)");
  if (ancestors)
    for (size_t k = 0; k < (size_t)T * N; ++k)
      if (ancestors[k] < 0 || ancestors[k] >= N) return gg_fail(-2, "ancestor outside [0, N)");
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
  if (!rc)
    chk(launch_backward(d_lw, d_X, d_vel, d_dt, d_sp, sigma_cb, d_anc, T, N, n_traj, d_u, d_f,
                        d_idx, d_st, nullptr),
        "gg_k_ffbsi");
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
  GG_HIP(launch_backward(log_weights, states, vel, dt, sigma_pos, sigma_cb, ancestors, T, N,
                         n_traj, u, forced, indices, status, (hipStream_t)hip_stream));
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
