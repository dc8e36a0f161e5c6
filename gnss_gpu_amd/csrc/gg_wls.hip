// Batched Gauss-Newton WLS (4 states: x y z cb) on gfx950.
//
// One 32-lane group per epoch: lane k holds satellite k (rows beyond 32 are
// folded in by a strided loop), the 10 unique entries of H^T W H and the 4 of
// H^T W dy are reduced with a DPP butterfly inside the group, and every lane
// then solves the same 4x4 system in binary64, so the group stays in lock
// step without LDS.  Two epochs per wavefront.
//
// mode 0 follows wls_batch_kernel (src/positioning/wls.cu:126-241): Sagnac
// range model, Gaussian elimination that skips tiny pivots.
// mode 1 follows wls_position (wls.cu:10-123): no Sagnac, -1 for < 4
// satellites or a degenerate centroid, early return on a singular matrix.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "gg_common_host.h"
#include "gg_device.h"
#include "gnss_gpu_amd.h"

#define GG_WLS_GROUP 32
#define GG_WLS_BLOCK 256

namespace {

constexpr double kWgs84A = 6378137.0;
constexpr double kOmegaE = 7.2921151467e-5;
constexpr double kCLight = 299792458.0;

// sum over the 32 lanes of a group (DPP butterfly, every lane gets the total)
__device__ __forceinline__ double group_sum(double v) {
  GG_BUTTERFLY32(f64, v, gg_add_f64);
  return v;
}

// receiver-minus-satellite vector and range, optionally with the Sagnac
// rotation of the satellite by omega_E * range / c (range_model.py:11-40)
__device__ __forceinline__ void los(double sx, double sy, double sz, double x, double y, double z,
                                    bool sagnac, double& dx, double& dy, double& dz, double& r) {
  if (sagnac) {
    const double dx0 = x - sx, dy0 = y - sy, dz0 = z - sz;
    const double ra = sqrt(dx0 * dx0 + dy0 * dy0 + dz0 * dz0);
    const double th = kOmegaE * (ra / kCLight);
    double st, ct;
    sincos(th, &st, &ct);
    const double sxr = sx * ct + sy * st;
    const double syr = -sx * st + sy * ct;
    dx = x - sxr;
    dy = y - syr;
    dz = z - sz;
    r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r < 1e-12) r = 1e-12;
  } else {
    dx = x - sx;
    dy = y - sy;
    dz = z - sz;
    r = sqrt(dx * dx + dy * dy + dz * dz);
  }
}

// Gaussian elimination with partial pivoting on A[4][5].
// cpu_rule: a pivot below 1e-15 returns false (wls_position); otherwise the
// pivot is skipped and back substitution divides by it (wls_batch).
__device__ __forceinline__ bool solve4(double A[4][5], double* x, bool cpu_rule) {
#pragma unroll
  for (int col = 0; col < 4; ++col) {
    int mr = col;
    double best = fabs(A[col][col]);
#pragma unroll
    for (int row = col + 1; row < 4; ++row)
      if (fabs(A[row][col]) > best) {
        best = fabs(A[row][col]);
        mr = row;
      }
    // swap through selects: a run-time row index would put A in scratch
#pragma unroll
    for (int row = col + 1; row < 4; ++row) {
      const bool sw = (row == mr);
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const double a = A[col][k], b = A[row][k];
        A[col][k] = sw ? b : a;
        A[row][k] = sw ? a : b;
      }
    }
    if (fabs(A[col][col]) < 1e-15) {
      if (cpu_rule) return false;
      continue;
    }
    for (int row = col + 1; row < 4; ++row) {
      const double fac = A[row][col] / A[col][col];
      for (int k = col; k < 5; ++k) A[row][k] -= fac * A[col][k];
    }
  }
  for (int row = 3; row >= 0; --row) {
    double v = A[row][4];
    for (int c = row + 1; c < 4; ++c) v -= A[row][c] * x[c];
    x[row] = v / A[row][row];
  }
  return true;
}

__global__ __launch_bounds__(GG_WLS_BLOCK) void gg_k_wls(const double* __restrict__ sat,
                                                         const double* __restrict__ pr,
                                                         const double* __restrict__ w,
                                                         const int32_t* __restrict__ nsat,
                                                         int n_epoch, int s_max, int max_iter,
                                                         double tol, double* __restrict__ res,
                                                         int32_t* __restrict__ iters, int mode) {
  const int lane = threadIdx.x & (GG_WLS_GROUP - 1);
  const int64_t e = ((int64_t)blockIdx.x * GG_WLS_BLOCK + threadIdx.x) / GG_WLS_GROUP;
  if (e >= n_epoch) return;
  const int ns = nsat ? nsat[e] : s_max;
  const double* es = sat + e * (int64_t)s_max * 3;
  const double* ep = pr + e * (int64_t)s_max;
  const double* ew = w + e * (int64_t)s_max;
  const bool sagnac = (mode == 0);
  const bool cpu_rule = (mode == 1);
  if (ns < 4 || ns > s_max) {
    if (lane == 0) {
      const double fill = cpu_rule ? 0.0 : __builtin_nan("");
      for (int k = 0; k < 4; ++k) res[e * 4 + k] = fill;
      if (iters) iters[e] = -1;
    }
    return;
  }
  // centroid projected onto the WGS84 semi-major-axis sphere
  double cx = 0.0, cy = 0.0, cz = 0.0;
  for (int k = lane; k < ns; k += GG_WLS_GROUP) {
    cx += es[k * 3 + 0];
    cy += es[k * 3 + 1];
    cz += es[k * 3 + 2];
  }
  cx = group_sum(cx) / ns;
  cy = group_sum(cy) / ns;
  cz = group_sum(cz) / ns;
  const double cn = sqrt(cx * cx + cy * cy + cz * cz);
  if (cpu_rule && cn < 1e-6) {
    if (lane == 0) {
      for (int k = 0; k < 4; ++k) res[e * 4 + k] = 0.0;
      if (iters) iters[e] = -1;
    }
    return;
  }
  const double scale = kWgs84A / cn;
  double x = cx * scale, y = cy * scale, z = cz * scale;
  double cb = 0.0;
  for (int k = lane; k < ns; k += GG_WLS_GROUP) {
    double dx, dy, dz, r;
    los(es[k * 3 + 0], es[k * 3 + 1], es[k * 3 + 2], x, y, z, sagnac, dx, dy, dz, r);
    cb += ep[k] - r;
  }
  cb = group_sum(cb) / ns;
  int it;
  for (it = 0; it < max_iter; ++it) {
    double v[14];
    for (int q = 0; q < 14; ++q) v[q] = 0.0;
    for (int k = lane; k < ns; k += GG_WLS_GROUP) {
      double dx, dy, dz, r;
      los(es[k * 3 + 0], es[k * 3 + 1], es[k * 3 + 2], x, y, z, sagnac, dx, dy, dz, r);
      const double resid = ep[k] - (r + cb);
      const double wk = ew[k];
      const double H[4] = {dx / r, dy / r, dz / r, 1.0};
      int q = 0;
      for (int a = 0; a < 4; ++a)
        for (int b = a; b < 4; ++b) v[q++] += H[a] * wk * H[b];
      for (int a = 0; a < 4; ++a) v[10 + a] += H[a] * wk * resid;
    }
    for (int q = 0; q < 14; ++q) v[q] = group_sum(v[q]);
    double A[4][5];
    int q = 0;
    for (int a = 0; a < 4; ++a)
      for (int b = a; b < 4; ++b) {
        A[a][b] = v[q];
        A[b][a] = v[q];
        ++q;
      }
    for (int a = 0; a < 4; ++a) A[a][4] = v[10 + a];
    double d[4];
    if (!solve4(A, d, cpu_rule)) break;  // singular (wls_position): keep estimate
    x += d[0];
    y += d[1];
    z += d[2];
    cb += d[3];
    const double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + d[3] * d[3]);
    if (nrm < tol) {
      ++it;
      break;
    }
  }
  if (lane == 0) {
    res[e * 4 + 0] = x;
    res[e * 4 + 1] = y;
    res[e * 4 + 2] = z;
    res[e * 4 + 3] = cb;
    if (iters) iters[e] = it;
  }
}

int launch_wls(const double* sat, const double* pr, const double* w, const int32_t* nsat,
               int n_epoch, int s_max, int max_iter, double tol, double* res, int32_t* iters,
               int mode, hipStream_t st) {
  if (n_epoch <= 0) return 0;
  const int64_t threads = (int64_t)n_epoch * GG_WLS_GROUP;
  const int grid = (int)((threads + GG_WLS_BLOCK - 1) / GG_WLS_BLOCK);
  hipLaunchKernelGGL(gg_k_wls, dim3(grid), dim3(GG_WLS_BLOCK), 0, st, sat, pr, w, nsat, n_epoch,
                     s_max, max_iter, tol, res, iters, mode);
  GG_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Multi-GNSS WLS: position plus one clock bias per system, 3 + NS states.
//
// Same 32-lane group per epoch as gg_k_wls.  Every row of H = [u, e_sys] has
// a single clock entry, so H^T W H is fixed by the 6 geometry sums and, per
// system, the 3 sums of w*u and the weight sum; H^T W dy by 3 + NS sums.
// Those 9 + 5*NS values go through a DPP butterfly (instead of the
// (3+NS)(4+NS)/2 + 3+NS of a dense assembly) and every lane assembles the
// augmented matrix, with exact zeros between different clocks as the
// reference's dense sums give, and solves it under solve_NxN's rules
// (src/positioning/multi_gnss.cu:157-187): partial pivoting, a pivot below
// 1e-15 skips its column, back substitution zeroes the unknown of a pivot
// <= 1e-15 (a system without satellites keeps its clock).
//
// mode 0 follows wls_multi_gnss_batch_kernel (multi_gnss.cu:211-305): Sagnac
// ranges, no guards.  mode 1 follows the CPU wls_multi_gnss
// (multi_gnss.cu:50-153): plain ranges, zeros and -1 for n_sat < 3 + NS or a
// centroid below 1e-6 m.  A system id outside [0, NS) adds no clock column
// and is predicted with clock 0's bias (multi_gnss.cu:111, :120).
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void solve_nxn(double (&A)[N][N + 1], double (&x)[N]) {
#pragma unroll
  for (int col = 0; col < N; ++col) {
    int mr = col;
    double best = fabs(A[col][col]);
#pragma unroll
    for (int row = col + 1; row < N; ++row) {
      const double v = fabs(A[row][col]);
      if (v > best) {
        best = v;
        mr = row;
      }
    }
    // row swap through selects keeps A in registers (mr is a run-time index)
#pragma unroll
    for (int row = col + 1; row < N; ++row) {
      const bool sw = (row == mr);
#pragma unroll
      for (int k = 0; k < N + 1; ++k) {
        const double a = A[col][k], b = A[row][k];
        A[col][k] = sw ? b : a;
        A[row][k] = sw ? a : b;
      }
    }
    const double p = A[col][col];
    if (!(fabs(p) < 1e-15)) {
#pragma unroll
      for (int row = col + 1; row < N; ++row) {
        const double f = A[row][col] / p;
#pragma unroll
        for (int k = col; k < N + 1; ++k) A[row][k] -= f * A[col][k];
      }
    }
  }
#pragma unroll
  for (int row = N - 1; row >= 0; --row) {
    double v = A[row][N];
#pragma unroll
    for (int c = row + 1; c < N; ++c) v -= A[row][c] * x[c];
    x[row] = fabs(A[row][row]) > 1e-15 ? v / A[row][row] : 0.0;
  }
}

template <int NS>
__global__ __launch_bounds__(GG_WLS_BLOCK) void gg_k_wls_mg(
    const double* __restrict__ sat, const double* __restrict__ pr, const double* __restrict__ w,
    const int32_t* __restrict__ sys, int n_epoch, int n_sat, int max_iter, double tol,
    double* __restrict__ res, int32_t* __restrict__ iters, int mode) {
  constexpr int N = 3 + NS;
  constexpr int NB = NS > 0 ? NS : 1;
  const int lane = threadIdx.x & (GG_WLS_GROUP - 1);
  // whole groups share an epoch, so the butterflies below never mix a live
  // group with a retired one; a group past the end leaves as a whole
  const int64_t e = ((int64_t)blockIdx.x * GG_WLS_BLOCK + threadIdx.x) / GG_WLS_GROUP;
  if (e >= n_epoch) return;
  const double* es = sat + e * (int64_t)n_sat * 3;
  const double* ep = pr + e * (int64_t)n_sat;
  const double* ew = w + e * (int64_t)n_sat;
  const int32_t* eid = sys + e * (int64_t)n_sat;
  double* out = res + e * N;
  const bool sagnac = (mode == 0);
  const bool cpu_rule = (mode == 1);
  if (cpu_rule && n_sat < N) {
    if (lane == 0) {
      for (int k = 0; k < N; ++k) out[k] = 0.0;
      if (iters) iters[e] = -1;
    }
    return;
  }
  double cx = 0.0, cy = 0.0, cz = 0.0;
  for (int k = lane; k < n_sat; k += GG_WLS_GROUP) {
    cx += es[k * 3 + 0];
    cy += es[k * 3 + 1];
    cz += es[k * 3 + 2];
  }
  cx = group_sum(cx) / n_sat;
  cy = group_sum(cy) / n_sat;
  cz = group_sum(cz) / n_sat;
  const double cn = sqrt(cx * cx + cy * cy + cz * cz);
  if (cpu_rule && cn < 1e-6) {
    if (lane == 0) {
      for (int k = 0; k < N; ++k) out[k] = 0.0;
      if (iters) iters[e] = -1;
    }
    return;
  }
  const double scale = kWgs84A / cn;
  double x = cx * scale, y = cy * scale, z = cz * scale;
  // per-system clocks from the mean residual of that system's satellites
  double cb[NB], cnt[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) cb[j] = cnt[j] = 0.0;
  for (int k = lane; k < n_sat; k += GG_WLS_GROUP) {
    double dx, dy, dz, r;
    los(es[k * 3 + 0], es[k * 3 + 1], es[k * 3 + 2], x, y, z, sagnac, dx, dy, dz, r);
    const int sk = eid[k];
    const double rr = ep[k] - r;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      cb[j] += (sk == j) ? rr : 0.0;
      cnt[j] += (sk == j) ? 1.0 : 0.0;
    }
  }
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const double c = group_sum(cnt[j]);
    const double s = group_sum(cb[j]);
    cb[j] = c > 0.0 ? s / c : 0.0;
  }
  if (NS == 0) cb[0] = 0.0;
  int it;
  for (it = 0; it < max_iter; ++it) {
    double g[9], cu[NB][3], cw[NB], cr[NB];
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] = 0.0;
#pragma unroll
    for (int j = 0; j < NB; ++j) cu[j][0] = cu[j][1] = cu[j][2] = cw[j] = cr[j] = 0.0;
    for (int k = lane; k < n_sat; k += GG_WLS_GROUP) {
      double dx, dy, dz, r;
      los(es[k * 3 + 0], es[k * 3 + 1], es[k * 3 + 2], x, y, z, sagnac, dx, dy, dz, r);
      const int sk = eid[k];
      double cbk = cb[0];
#pragma unroll
      for (int j = 1; j < NS; ++j) cbk = (sk == j) ? cb[j] : cbk;
      const double resid = ep[k] - (r + cbk);
      const double wk = ew[k];
      const double u[3] = {dx / r, dy / r, dz / r};
      g[0] += u[0] * wk * u[0];
      g[1] += u[0] * wk * u[1];
      g[2] += u[0] * wk * u[2];
      g[3] += u[1] * wk * u[1];
      g[4] += u[1] * wk * u[2];
      g[5] += u[2] * wk * u[2];
#pragma unroll
      for (int a = 0; a < 3; ++a) g[6 + a] += u[a] * wk * resid;
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const double m = (sk == j) ? wk : 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) cu[j][a] += u[a] * m;
        cw[j] += m;
        cr[j] += m * resid;
      }
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] = group_sum(g[q]);
#pragma unroll
    for (int j = 0; j < NS; ++j) {
#pragma unroll
      for (int a = 0; a < 3; ++a) cu[j][a] = group_sum(cu[j][a]);
      cw[j] = group_sum(cw[j]);
      cr[j] = group_sum(cr[j]);
    }
    double A[N][N + 1];
    A[0][0] = g[0];
    A[0][1] = A[1][0] = g[1];
    A[0][2] = A[2][0] = g[2];
    A[1][1] = g[3];
    A[1][2] = A[2][1] = g[4];
    A[2][2] = g[5];
#pragma unroll
    for (int a = 0; a < 3; ++a) A[a][N] = g[6 + a];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
#pragma unroll
      for (int a = 0; a < 3; ++a) A[a][3 + j] = A[3 + j][a] = cu[j][a];
#pragma unroll
      for (int i = 0; i < NS; ++i) A[3 + j][3 + i] = (i == j) ? cw[j] : 0.0;
      A[3 + j][N] = cr[j];
    }
    double d[N];
    solve_nxn<N>(A, d);
    x += d[0];
    y += d[1];
    z += d[2];
#pragma unroll
    for (int j = 0; j < NS; ++j) cb[j] += d[3 + j];
    double nrm = 0.0;
#pragma unroll
    for (int q = 0; q < N; ++q) nrm += d[q] * d[q];
    if (sqrt(nrm) < tol) {
      ++it;
      break;
    }
  }
  if (lane == 0) {
    out[0] = x;
    out[1] = y;
    out[2] = z;
#pragma unroll
    for (int j = 0; j < NS; ++j) out[3 + j] = cb[j];
    if (iters) iters[e] = it;
  }
}

int launch_wls_mg(const double* sat, const double* pr, const double* w, const int32_t* sys,
                  int n_epoch, int n_sat, int n_systems, int max_iter, double tol, double* res,
                  int32_t* iters, int mode, hipStream_t st) {
  if (n_epoch <= 0) return 0;
  const int64_t threads = (int64_t)n_epoch * GG_WLS_GROUP;
  const int grid = (int)((threads + GG_WLS_BLOCK - 1) / GG_WLS_BLOCK);
#define GG_MG_LAUNCH(NS_)                                                                      \
  hipLaunchKernelGGL(gg_k_wls_mg<NS_>, dim3(grid), dim3(GG_WLS_BLOCK), 0, st, sat, pr, w, sys, \
                     n_epoch, n_sat, max_iter, tol, res, iters, mode)
  switch (n_systems) {
    case 0: GG_MG_LAUNCH(0); break;
    case 1: GG_MG_LAUNCH(1); break;
    case 2: GG_MG_LAUNCH(2); break;
    case 3: GG_MG_LAUNCH(3); break;
    case 4: GG_MG_LAUNCH(4); break;
    case 5: GG_MG_LAUNCH(5); break;
    default: return gg_fail(-2, "n_systems must be in [0, 5]");
  }
#undef GG_MG_LAUNCH
  GG_HIP(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int gg_wls_batch_device(const double* sat_ecef, const double* pseudoranges,
                        const double* weights, const int32_t* n_sat, int32_t n_epoch,
                        int32_t n_sat_max, int32_t max_iter, double tol, double* results,
                        int32_t* iters, int32_t mode, void* hip_stream) {
  if (n_epoch < 0 || n_sat_max < 0) return gg_fail(-2, "negative size");
  if (mode != 0 && mode != 1) return gg_fail(-2, "mode must be 0 or 1");
  return launch_wls(sat_ecef, pseudoranges, weights, n_sat, n_epoch, n_sat_max, max_iter, tol,
                    results, iters, mode, (hipStream_t)hip_stream);
}

// Host-array batch: one device allocation per call, like the reference's
// wls_batch host wrapper (wls.cu:243-278).
static int wls_host(const double* sat_ecef, const double* pseudoranges, const double* weights,
                    const int32_t* n_sat, int32_t n_epoch, int32_t n_sat_max, int32_t max_iter,
                    double tol, double* results, int32_t* iters, int mode) {
  if (n_epoch < 0 || n_sat_max < 0) return gg_fail(-2, "negative size");
  if (n_epoch == 0) return 0;
  const size_t ns = (size_t)n_epoch * n_sat_max;
  const size_t bytes = sizeof(double) * (ns * 5 + (size_t)n_epoch * 4) +
                       sizeof(int32_t) * (size_t)n_epoch * 2 + 256;
  char* buf = nullptr;
  GG_HIP(hipMalloc((void**)&buf, bytes));
  double* d_sat = (double*)buf;
  double* d_pr = d_sat + ns * 3;
  double* d_w = d_pr + ns;
  double* d_res = d_w + ns;
  int32_t* d_it = (int32_t*)(d_res + (size_t)n_epoch * 4);
  int32_t* d_ns = d_it + n_epoch;
  hipStream_t st = nullptr;
  int rc = 0;
  auto chk = [&](hipError_t e, const char* what) {
    if (!rc && e != hipSuccess) rc = gg_fail_hip(e, what, __FILE__, __LINE__);
  };
  chk(hipMemcpy(d_sat, sat_ecef, sizeof(double) * ns * 3, hipMemcpyHostToDevice), "H2D sat");
  chk(hipMemcpy(d_pr, pseudoranges, sizeof(double) * ns, hipMemcpyHostToDevice), "H2D pr");
  if (weights) {
    chk(hipMemcpy(d_w, weights, sizeof(double) * ns, hipMemcpyHostToDevice), "H2D w");
  } else {
    std::vector<double> ones(ns, 1.0);
    chk(hipMemcpy(d_w, ones.data(), sizeof(double) * ns, hipMemcpyHostToDevice), "H2D w");
  }
  if (n_sat) chk(hipMemcpy(d_ns, n_sat, sizeof(int32_t) * n_epoch, hipMemcpyHostToDevice), "H2D n");
  if (!rc)
    rc = launch_wls(d_sat, d_pr, d_w, n_sat ? d_ns : nullptr, n_epoch, n_sat_max, max_iter, tol,
                    d_res, d_it, mode, st);
  chk(hipMemcpy(results, d_res, sizeof(double) * n_epoch * 4, hipMemcpyDeviceToHost), "D2H res");
  if (iters)
    chk(hipMemcpy(iters, d_it, sizeof(int32_t) * n_epoch, hipMemcpyDeviceToHost), "D2H iters");
  const hipError_t fe = hipFree(buf);
  if (!rc && fe != hipSuccess) rc = gg_fail_hip(fe, "hipFree", __FILE__, __LINE__);
  return rc;
}

int gg_wls_batch(const double* sat_ecef, const double* pseudoranges, const double* weights,
                 const int32_t* n_sat, int32_t n_epoch, int32_t n_sat_max, int32_t max_iter,
                 double tol, double* results, int32_t* iters) {
  return wls_host(sat_ecef, pseudoranges, weights, n_sat, n_epoch, n_sat_max, max_iter, tol,
                  results, iters, 0);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free host-array batch (gg_wls_batch_noalloc): the inputs and
// outputs are staged in `scratch`, a file-scope device array that every GPU
// the library runs on holds until the process ends.  Host code never passes
// the array's name to a kernel or a copy: the device address comes from
// hipGetSymbolAddress, taken once per device (with that device current) on
// the entry point's first call there.
// ---------------------------------------------------------------------------
static __device__ __align__(256) char scratch[GG_WLS_NOALLOC_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;           // guards g_scratch_ptr and serialises the calls
char* g_scratch_ptr[kMaxDevices];  // device address of `scratch` per device (null: not taken)

int scratch_of_current_device(char** out) {
  int dev = 0;
  GG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return gg_fail(-2, "device index beyond 63");
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    GG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return gg_fail(-1, "hipGetSymbolAddress returned a null staging address");
    g_scratch_ptr[dev] = static_cast<char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

__global__ void gg_k_fill_f64(double* p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

int gg_wls_batch_noalloc(const double* sat_ecef, const double* pseudoranges,
                         const double* weights, const int32_t* n_sat, int32_t n_epoch,
                         int32_t n_sat_max, int32_t max_iter, double tol, double* results,
                         int32_t* iters) {
  if (n_epoch < 0 || n_sat_max < 0) return gg_fail(-2, "negative size");
  if (n_epoch == 0) return 0;
  if (!sat_ecef || !pseudoranges || !results) return gg_fail(-2, "null array");
  const size_t ns = (size_t)n_epoch * (size_t)n_sat_max;
  const size_t o_pr = round256(sizeof(double) * ns * 3);
  const size_t o_w = o_pr + round256(sizeof(double) * ns);
  const size_t o_res = o_w + round256(sizeof(double) * ns);
  const size_t o_it = o_res + round256(sizeof(double) * (size_t)n_epoch * 4);
  const size_t o_ns = o_it + round256(sizeof(int32_t) * (size_t)n_epoch);
  const size_t need = o_ns + round256(sizeof(int32_t) * (size_t)n_epoch);
  if (need > (size_t)GG_WLS_NOALLOC_BYTES)
    return gg_fail(-2, "gg_wls_batch_noalloc: n_epoch * n_sat_max exceeds the 16 MiB staging "
                       "area (at most 40 * n_epoch * n_sat_max + 40 * n_epoch bytes plus "
                       "alignment)");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  char* base = nullptr;
  if (scratch_of_current_device(&base)) return -1;
  double* d_sat = reinterpret_cast<double*>(base);
  double* d_pr = reinterpret_cast<double*>(base + o_pr);
  double* d_w = reinterpret_cast<double*>(base + o_w);
  double* d_res = reinterpret_cast<double*>(base + o_res);
  int32_t* d_it = reinterpret_cast<int32_t*>(base + o_it);
  int32_t* d_ns = reinterpret_cast<int32_t*>(base + o_ns);
  GG_HIP(hipMemcpy(d_sat, sat_ecef, sizeof(double) * ns * 3, hipMemcpyHostToDevice));
  GG_HIP(hipMemcpy(d_pr, pseudoranges, sizeof(double) * ns, hipMemcpyHostToDevice));
  if (weights) {
    GG_HIP(hipMemcpy(d_w, weights, sizeof(double) * ns, hipMemcpyHostToDevice));
  } else if (ns > 0) {
    // the previous call's contents stay in the area: the weights are written in full
    hipLaunchKernelGGL(gg_k_fill_f64, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, nullptr,
                       d_w, (int64_t)ns, 1.0);
    GG_HIP(hipGetLastError());
  }
  if (n_sat)
    GG_HIP(hipMemcpy(d_ns, n_sat, sizeof(int32_t) * (size_t)n_epoch, hipMemcpyHostToDevice));
  if (launch_wls(d_sat, d_pr, d_w, n_sat ? d_ns : nullptr, n_epoch, n_sat_max, max_iter, tol,
                 d_res, d_it, 0, nullptr))
    return -1;
  GG_HIP(hipMemcpy(results, d_res, sizeof(double) * (size_t)n_epoch * 4, hipMemcpyDeviceToHost));
  if (iters)
    GG_HIP(hipMemcpy(iters, d_it, sizeof(int32_t) * (size_t)n_epoch, hipMemcpyDeviceToHost));
  return 0;
}

int gg_wls_position(const double* sat_ecef, const double* pseudoranges, const double* weights,
                    int32_t n_sat, int32_t max_iter, double tol, double* result,
                    int32_t* iters) {
  int32_t it = 0;
  const int rc = wls_host(sat_ecef, pseudoranges, weights, nullptr, 1, n_sat, max_iter, tol,
                          result, &it, 1);
  if (iters) *iters = it;
  return rc;
}

int gg_wls_multi_gnss_batch_device(const double* sat_ecef, const double* pseudoranges,
                                   const double* weights, const int32_t* system_ids,
                                   int32_t n_epoch, int32_t n_sat, int32_t n_systems,
                                   int32_t max_iter, double tol, double* results, int32_t* iters,
                                   int32_t mode, void* hip_stream) {
  if (n_epoch < 0 || n_sat < 0) return gg_fail(-2, "negative size");
  if (n_systems < 0 || n_systems > GG_MAX_SYSTEMS) return gg_fail(-2, "n_systems must be in [0, 5]");
  if (mode != 0 && mode != 1) return gg_fail(-2, "mode must be 0 or 1");
  if (mode == 0 && n_sat < 3 + n_systems)
    return gg_fail(-2, "wls_multi_gnss_batch requires at least 3 + n_systems satellites");
  if (n_epoch > 0 && (!sat_ecef || !pseudoranges || !weights || !system_ids || !results))
    return gg_fail(-2, "null array");
  return launch_wls_mg(sat_ecef, pseudoranges, weights, system_ids, n_epoch, n_sat, n_systems,
                       max_iter, tol, results, iters, mode, (hipStream_t)hip_stream);
}

// Host-array multi-GNSS solve: one device allocation per call, like the
// reference's wls_multi_gnss_batch host wrapper (multi_gnss.cu:307-350).
static int wls_mg_host(const double* sat_ecef, const double* pseudoranges, const double* weights,
                       const int32_t* system_ids, int32_t n_epoch, int32_t n_sat,
                       int32_t n_systems, int32_t max_iter, double tol, double* results,
                       int32_t* iters, int mode) {
  if (n_epoch < 0 || n_sat < 0) return gg_fail(-2, "negative size");
  if (n_systems < 0 || n_systems > GG_MAX_SYSTEMS) return gg_fail(-2, "n_systems must be in [0, 5]");
  if (mode == 0 && n_sat < 3 + n_systems)
    return gg_fail(-2, "wls_multi_gnss_batch requires at least 3 + n_systems satellites");
  const int n_state = 3 + n_systems;
  if (n_epoch == 0) return 0;
  if (n_sat == 0) {  // only the single-epoch rule reaches here: too few satellites
    for (int64_t k = 0; k < (int64_t)n_epoch * n_state; ++k) results[k] = 0.0;
    if (iters)
      for (int32_t e = 0; e < n_epoch; ++e) iters[e] = -1;
    return 0;
  }
  const size_t ns = (size_t)n_epoch * n_sat;
  const size_t nr = (size_t)n_epoch * n_state;
  const size_t bytes = sizeof(double) * (ns * 5 + nr) + sizeof(int32_t) * (ns + n_epoch) + 256;
  char* buf = nullptr;
  GG_HIP(hipMalloc((void**)&buf, bytes));
  double* d_sat = (double*)buf;
  double* d_pr = d_sat + ns * 3;
  double* d_w = d_pr + ns;
  double* d_res = d_w + ns;
  int32_t* d_sys = (int32_t*)(d_res + nr);
  int32_t* d_it = d_sys + ns;
  int rc = 0;
  auto chk = [&](hipError_t e, const char* what) {
    if (!rc && e != hipSuccess) rc = gg_fail_hip(e, what, __FILE__, __LINE__);
  };
  chk(hipMemcpy(d_sat, sat_ecef, sizeof(double) * ns * 3, hipMemcpyHostToDevice), "H2D sat");
  chk(hipMemcpy(d_pr, pseudoranges, sizeof(double) * ns, hipMemcpyHostToDevice), "H2D pr");
  if (weights) {
    chk(hipMemcpy(d_w, weights, sizeof(double) * ns, hipMemcpyHostToDevice), "H2D w");
  } else {
    std::vector<double> ones(ns, 1.0);
    chk(hipMemcpy(d_w, ones.data(), sizeof(double) * ns, hipMemcpyHostToDevice), "H2D w");
  }
  chk(hipMemcpy(d_sys, system_ids, sizeof(int32_t) * ns, hipMemcpyHostToDevice), "H2D sys");
  if (!rc)
    rc = launch_wls_mg(d_sat, d_pr, d_w, d_sys, n_epoch, n_sat, n_systems, max_iter, tol, d_res,
                       d_it, mode, nullptr);
  chk(hipMemcpy(results, d_res, sizeof(double) * nr, hipMemcpyDeviceToHost), "D2H res");
  if (iters)
    chk(hipMemcpy(iters, d_it, sizeof(int32_t) * n_epoch, hipMemcpyDeviceToHost), "D2H iters");
  const hipError_t fe = hipFree(buf);
  if (!rc && fe != hipSuccess) rc = gg_fail_hip(fe, "hipFree", __FILE__, __LINE__);
  return rc;
}

int gg_wls_multi_gnss_batch(const double* sat_ecef, const double* pseudoranges,
                            const double* weights, const int32_t* system_ids, int32_t n_epoch,
                            int32_t n_sat, int32_t n_systems, int32_t max_iter, double tol,
                            double* results, int32_t* iters) {
  return wls_mg_host(sat_ecef, pseudoranges, weights, system_ids, n_epoch, n_sat, n_systems,
                     max_iter, tol, results, iters, 0);
}

int gg_wls_multi_gnss(const double* sat_ecef, const double* pseudoranges, const double* weights,
                      const int32_t* system_ids, int32_t n_sat, int32_t n_systems,
                      int32_t max_iter, double tol, double* result, int32_t* iters) {
  int32_t it = 0;
  const int rc = wls_mg_host(sat_ecef, pseudoranges, weights, system_ids, 1, n_sat, n_systems,
                             max_iter, tol, result, &it, 1);
  if (iters) *iters = it;
  return rc;
}

}  // extern "C"
