/*
 * C restatement of the reference's Gauss-Newton WLS (src/positioning/wls.cu),
 * used as the timed CPU baseline of the batched WLS (TEST/MEASUREMENT
 * INFRASTRUCTURE ONLY; the product never links it).  One epoch per loop
 * iteration, OpenMP over epochs:
 *   initial guess  wls.cu:183-199   centroid projected to the WGS84 radius,
 *                                   clock bias = mean residual
 *   ranges         wls.cu:150-170   Sagnac rotation theta = omega_E r / c
 *                                   (sagnac = 0: plain ranges, wls.cu:10-123)
 *   iteration      wls.cu:200-236   4x4 normal equations, Gaussian elimination
 *                                   with partial pivoting, ||dx|| < tol
 */
#include <math.h>
#include <stdint.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define WGS84_A 6378137.0
#define OMEGA_E 7.2921151467e-5
#define C_LIGHT 299792458.0

static void los(const double* s, double x, double y, double z, int sagnac, double* dx,
                double* dy, double* dz, double* r) {
  double sx = s[0], sy = s[1];
  if (sagnac) {
    const double ax = x - s[0], ay = y - s[1], az = z - s[2];
    const double ra = sqrt(ax * ax + ay * ay + az * az);
    const double th = OMEGA_E * (ra / C_LIGHT);
    const double ct = cos(th), st = sin(th);
    sx = s[0] * ct + s[1] * st;
    sy = -s[0] * st + s[1] * ct;
  }
  *dx = x - sx;
  *dy = y - sy;
  *dz = z - s[2];
  *r = sqrt(*dx * *dx + *dy * *dy + *dz * *dz);
  if (sagnac && *r < 1e-12) *r = 1e-12;
}

/* Gaussian elimination with partial pivoting on A[4][5]; 0 if a pivot is < 1e-15. */
static int solve4(double A[4][5], double* x) {
  for (int c = 0; c < 4; ++c) {
    int p = c;
    for (int r = c + 1; r < 4; ++r)
      if (fabs(A[r][c]) > fabs(A[p][c])) p = r;
    if (p != c)
      for (int k = 0; k < 5; ++k) {
        const double t = A[c][k];
        A[c][k] = A[p][k];
        A[p][k] = t;
      }
    if (fabs(A[c][c]) < 1e-15) return 0;
    for (int r = c + 1; r < 4; ++r) {
      const double f = A[r][c] / A[c][c];
      for (int k = c; k < 5; ++k) A[r][k] -= f * A[c][k];
    }
  }
  for (int r = 3; r >= 0; --r) {
    double v = A[r][4];
    for (int k = r + 1; k < 4; ++k) v -= A[r][k] * x[k];
    x[r] = v / A[r][r];
  }
  return 1;
}

int cpu_wls_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* sat[E][S][3], pr[E][S], w[E][S], n_sat[E] (<= S); out[E][4], iters[E]. */
void cpu_wls_batch(const double* sat, const double* pr, const double* w, const int32_t* n_sat,
                   int E, int S, int max_iter, double tol, int sagnac, int threads, double* out,
                   int32_t* iters) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(static)
  for (int e = 0; e < E; ++e) {
    const double* se = sat + (int64_t)e * S * 3;
    const double* pe = pr + (int64_t)e * S;
    const double* we = w + (int64_t)e * S;
    const int ns = n_sat ? n_sat[e] : S;
    double* o = out + (int64_t)e * 4;
    if (ns < 4) {
      o[0] = o[1] = o[2] = o[3] = NAN;
      iters[e] = -1;
      continue;
    }
    double cx = 0, cy = 0, cz = 0;
    for (int s = 0; s < ns; ++s) {
      cx += se[s * 3];
      cy += se[s * 3 + 1];
      cz += se[s * 3 + 2];
    }
    cx /= ns;
    cy /= ns;
    cz /= ns;
    const double cn = sqrt(cx * cx + cy * cy + cz * cz);
    double x = cx * (WGS84_A / cn), y = cy * (WGS84_A / cn), z = cz * (WGS84_A / cn), cb = 0;
    for (int s = 0; s < ns; ++s) {
      double dx, dy, dz, r;
      los(se + s * 3, x, y, z, sagnac, &dx, &dy, &dz, &r);
      cb += pe[s] - r;
    }
    cb /= ns;
    int it = 0;
    for (int k = 0; k < max_iter; ++k) {
      double A[4][5] = {{0}};
      for (int s = 0; s < ns; ++s) {
        double dx, dy, dz, r;
        los(se + s * 3, x, y, z, sagnac, &dx, &dy, &dz, &r);
        const double res = pe[s] - (r + cb);
        const double H[4] = {dx / r, dy / r, dz / r, 1.0};
        for (int a = 0; a < 4; ++a) {
          A[a][4] += H[a] * we[s] * res;
          for (int b = 0; b < 4; ++b) A[a][b] += H[a] * we[s] * H[b];
        }
      }
      double d[4];
      if (!solve4(A, d)) {
        it = k;
        break;
      }
      x += d[0];
      y += d[1];
      z += d[2];
      cb += d[3];
      it = k + 1;
      if (sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + d[3] * d[3]) < tol) break;
    }
    o[0] = x;
    o[1] = y;
    o[2] = z;
    o[3] = cb;
    iters[e] = it;
  }
}

/*
 * Multi-GNSS WLS (src/positioning/multi_gnss.cu): state x, y, z and one clock
 * per system, dense (3+n)x(3+n) normal equations in the reference's order.
 *   sagnac = 1, guards = 0: wls_multi_gnss_batch_kernel (:211-305)
 *   sagnac = 0, guards = 1: wls_multi_gnss              (:50-153)
 */
#define MG_MAX_STATE 8

static void solve_nxn(double* A, double* d, int n) { /* multi_gnss.cu:15-47 */
  const int m = n + 1;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(A[r * m + c]) > fabs(A[p * m + c])) p = r;
    if (p != c)
      for (int k = 0; k < m; ++k) {
        const double t = A[c * m + k];
        A[c * m + k] = A[p * m + k];
        A[p * m + k] = t;
      }
    if (fabs(A[c * m + c]) < 1e-15) continue;
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * m + c] / A[c * m + c];
      for (int k = c; k < m; ++k) A[r * m + k] -= f * A[c * m + k];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    d[r] = A[r * m + n];
    for (int k = r + 1; k < n; ++k) d[r] -= A[r * m + k] * d[k];
    d[r] = fabs(A[r * m + r]) > 1e-15 ? d[r] / A[r * m + r] : 0.0;
  }
}

/* sat[E][S][3], pr/w[E][S], sys[E][S]; out[E][3+n_sys], iters[E]. */
void cpu_wls_mg_batch(const double* sat, const double* pr, const double* w, const int32_t* sys,
                      int E, int S, int n_sys, int max_iter, double tol, int sagnac, int guards,
                      int threads, double* out, int32_t* iters) {
  const int n = 3 + n_sys;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(static)
  for (int e = 0; e < E; ++e) {
    const double* se = sat + (int64_t)e * S * 3;
    const double* pe = pr + (int64_t)e * S;
    const double* we = w + (int64_t)e * S;
    const int32_t* ie = sys + (int64_t)e * S;
    double* o = out + (int64_t)e * n;
    if (guards && S < n) {
      for (int k = 0; k < n; ++k) o[k] = 0.0;
      iters[e] = -1;
      continue;
    }
    double cx = 0, cy = 0, cz = 0;
    for (int s = 0; s < S; ++s) {
      cx += se[s * 3];
      cy += se[s * 3 + 1];
      cz += se[s * 3 + 2];
    }
    cx /= S;
    cy /= S;
    cz /= S;
    const double cn = sqrt(cx * cx + cy * cy + cz * cz);
    if (guards && cn < 1e-6) {
      for (int k = 0; k < n; ++k) o[k] = 0.0;
      iters[e] = -1;
      continue;
    }
    double x = cx * (WGS84_A / cn), y = cy * (WGS84_A / cn), z = cz * (WGS84_A / cn);
    double cb[MG_MAX_STATE] = {0};
    int cnt[MG_MAX_STATE] = {0};
    for (int s = 0; s < S; ++s) {
      double dx, dy, dz, r;
      los(se + s * 3, x, y, z, sagnac, &dx, &dy, &dz, &r);
      const int k = ie[s];
      if (k >= 0 && k < n_sys) {
        cb[k] += pe[s] - r;
        cnt[k]++;
      }
    }
    for (int k = 0; k < n_sys; ++k)
      if (cnt[k] > 0) cb[k] /= cnt[k];
    int it;
    for (it = 0; it < max_iter; ++it) {
      double A[MG_MAX_STATE * (MG_MAX_STATE + 1)] = {0};
      for (int s = 0; s < S; ++s) {
        double dx, dy, dz, r;
        los(se + s * 3, x, y, z, sagnac, &dx, &dy, &dz, &r);
        const int k = ie[s];
        const int ok = k >= 0 && k < n_sys;
        const double res = pe[s] - (r + (ok ? cb[k] : cb[0]));
        double H[MG_MAX_STATE] = {0};
        H[0] = dx / r;
        H[1] = dy / r;
        H[2] = dz / r;
        if (ok) H[3 + k] = 1.0;
        for (int a = 0; a < n; ++a) {
          A[a * (n + 1) + n] += H[a] * we[s] * res;
          for (int b = 0; b < n; ++b) A[a * (n + 1) + b] += H[a] * we[s] * H[b];
        }
      }
      double d[MG_MAX_STATE];
      solve_nxn(A, d, n);
      x += d[0];
      y += d[1];
      z += d[2];
      for (int k = 0; k < n_sys; ++k) cb[k] += d[3 + k];
      double nrm = 0;
      for (int k = 0; k < n; ++k) nrm += d[k] * d[k];
      if (sqrt(nrm) < tol) {
        ++it;
        break;
      }
    }
    o[0] = x;
    o[1] = y;
    o[2] = z;
    for (int k = 0; k < n_sys; ++k) o[3 + k] = cb[k];
    iters[e] = it;
  }
}
