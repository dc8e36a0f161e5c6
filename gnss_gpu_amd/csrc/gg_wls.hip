// Batched Gauss-Newton WLS (4 states: x y z cb) on gfx950.
//
// One 32-lane group per epoch: lane k holds satellite k (rows beyond 32 are
// folded in by a strided loop), the 10 unique entries of H^T W H and the 4 of
// H^T W dy are reduced with a xor butterfly inside the group, and every lane
// then solves the same 4x4 system in binary64, so the group stays in lock
// step without LDS.  Two epochs per wavefront.
//
// mode 0 follows wls_batch_kernel (src/positioning/wls.cu:126-241): Sagnac
// range model, Gaussian elimination that skips tiny pivots.
// mode 1 follows wls_position (wls.cu:10-123): no Sagnac, -1 for < 4
// satellites or a degenerate centroid, early return on a singular matrix.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "gg_common_host.h"
#include "gnss_gpu_amd.h"

#define GG_WLS_GROUP 32
#define GG_WLS_BLOCK 256

namespace {

constexpr double kWgs84A = 6378137.0;
constexpr double kOmegaE = 7.2921151467e-5;
constexpr double kCLight = 299792458.0;

__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int d = GG_WLS_GROUP / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, GG_WLS_GROUP);
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
  for (int col = 0; col < 4; ++col) {
    int mr = col;
    for (int row = col + 1; row < 4; ++row)
      if (fabs(A[row][col]) > fabs(A[mr][col])) mr = row;
    if (mr != col)
      for (int k = 0; k < 5; ++k) {
        const double t = A[col][k];
        A[col][k] = A[mr][k];
        A[mr][k] = t;
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

int gg_wls_position(const double* sat_ecef, const double* pseudoranges, const double* weights,
                    int32_t n_sat, int32_t max_iter, double tol, double* result,
                    int32_t* iters) {
  int32_t it = 0;
  const int rc = wls_host(sat_ecef, pseudoranges, weights, nullptr, 1, n_sat, max_iter, tol,
                          result, &it, 1);
  if (iters) *iters = it;
  return rc;
}

}  // extern "C"
