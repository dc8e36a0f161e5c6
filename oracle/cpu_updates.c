/*
 * CPU restatement of the reference's observation updates of the device
 * particle filter, used as the timed CPU baseline of the update kernels
 * (TEST/MEASUREMENT INFRASTRUCTURE ONLY; the product never links it).
 * binary64, row-major inputs, OpenMP over particles; the formulas are those of
 * the NumPy oracle (oracle/pf.py), which restates the reference:
 *
 *   cpu_dd_pr_loglik          pf_device.cu:282-368  DD pseudorange (gate, Huber)
 *   cpu_carrier_afv_loglik    pf_device.cu:439-494  undifferenced carrier AFV
 *   cpu_dd_carrier_afv_loglik pf_device.cu:502-604  DD carrier AFV (gate, Huber)
 *   cpu_doppler_update        pf_device.cu:610-740  Doppler log-weight and the
 *                                                   velocity nudge toward the
 *                                                   per-particle Doppler WLS
 *   cpu_doppler_kf_update     pf_device.cu:742-880  per-particle velocity Kalman
 *                                                   update (RBPF)
 *
 * Every function returns its per-particle log-weight increment in out[n] and
 * (Doppler kinds) updates vel[n*3] (and cov[n*9]) in place.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static void set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

static inline double dist3(const double* a, const double* b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

/* Gaussian or weighted-Huber cost of one residual (pf_device.cu:19-26). */
static inline double cost(double r, double w, double sigma, int huber, double k) {
  if (huber && k > 0.0) {
    const double z = fabs(r) / sigma;
    return w * ((z < k) ? 0.5 * z * z : k * z - 0.5 * k * k);
  }
  return 0.5 * w * r * r / (sigma * sigma);
}

/* Sum of the costs of m residuals with the DD kernels' gate: residuals beyond
 * thr are dropped unless fewer than min(m, 3) would remain. */
static double gated_sum(const double* res, const double* w, int m, double sigma, double thr,
                        int huber, double k) {
  int keep_all = 1;
  if (thr > 0.0) {
    int kept = 0;
    for (int j = 0; j < m; ++j) kept += fabs(res[j]) <= thr;
    keep_all = kept < (m < 3 ? m : 3);
  }
  double s = 0.0;
  for (int j = 0; j < m; ++j)
    if (keep_all || fabs(res[j]) <= thr) s -= cost(res[j], w[j], sigma, huber, k);
  return s;
}

#define MAXOBS 256

int cpu_updates_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void cpu_dd_pr_loglik(int n, const double* pos, int m, const double* sat_k, const double* sat_ref,
                      const double* dd, const double* br_k, const double* br_ref, const double* w,
                      double sigma, double thr, int huber, double huber_k, double* out,
                      int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double res[MAXOBS];
    const double* p = pos + 3 * (size_t)i;
    for (int j = 0; j < m; ++j)
      res[j] = dd[j] - (((dist3(p, sat_k + 3 * j) - dist3(p, sat_ref + 3 * j)) - br_k[j]) +
                        br_ref[j]);
    out[i] = gated_sum(res, w, m, sigma, thr, huber, huber_k);
  }
}

void cpu_carrier_afv_loglik(int n, const double* pos, const double* cb, int m, const double* sat,
                            const double* cp, const double* w, double wavelength,
                            double sigma_cycles, double* out, int threads) {
  set_threads(threads);
  const double iwl = 1.0 / wavelength;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const double* p = pos + 3 * (size_t)i;
    double s = 0.0;
    for (int j = 0; j < m; ++j) {
      const double x = cp[j] - (dist3(p, sat + 3 * j) + cb[i]) * iwl;
      const double a = x - rint(x);
      s += -0.5 * w[j] * a * a / (sigma_cycles * sigma_cycles);
    }
    out[i] = s;
  }
}

void cpu_dd_carrier_afv_loglik(int n, const double* pos, int m, const double* sat_k,
                               const double* sat_ref, const double* ddc, const double* br_k,
                               const double* br_ref, const double* w, const double* wl,
                               double sigma_cycles, double thr, int huber, double huber_k,
                               double* out, int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double res[MAXOBS];
    const double* p = pos + 3 * (size_t)i;
    for (int j = 0; j < m; ++j) {
      const double e = (((dist3(p, sat_k + 3 * j) - dist3(p, sat_ref + 3 * j)) - br_k[j]) +
                        br_ref[j]) * (1.0 / wl[j]);
      const double x = ddc[j] - e;
      res[j] = x - rint(x);
    }
    out[i] = gated_sum(res, w, m, sigma_cycles, thr, huber, huber_k);
  }
}

/* unit line of sight from p to the satellite; 0 when closer than 1 m */
static inline int los(const double* p, const double* s, double* l) {
  const double d[3] = {s[0] - p[0], s[1] - p[1], s[2] - p[2]};
  const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (!(r >= 1.0)) return 0;
  for (int k = 0; k < 3; ++k) l[k] = d[k] / r;
  return 1;
}

/* Solve the 4x4 system A x = b by Gaussian elimination with partial pivoting;
 * returns 0 for a singular or non-finite system. */
static int solve4(double A[4][5], double* x) {
  for (int c = 0; c < 4; ++c) {
    int p = c;
    for (int r = c + 1; r < 4; ++r)
      if (fabs(A[r][c]) > fabs(A[p][c])) p = r;
    if (!(fabs(A[p][c]) > 0.0)) return 0;
    if (p != c)
      for (int k = 0; k < 5; ++k) {
        const double t = A[c][k];
        A[c][k] = A[p][k];
        A[p][k] = t;
      }
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
  for (int k = 0; k < 4; ++k)
    if (!isfinite(x[k])) return 0;
  return 1;
}

void cpu_doppler_update(int n, const double* pos, double* vel, int m, const double* sat,
                        const double* sv, const double* dop, const double* w_in, double wl,
                        double sigma, double gain, double max_dv, double* out, int threads) {
  set_threads(threads);
  const double g = gain < 0.0 ? 0.0 : (gain > 1.0 ? 1.0 : gain);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const double* p = pos + 3 * (size_t)i;
    double* v = vel + 3 * (size_t)i;
    double l[MAXOBS][3], w[MAXOBS], obs[MAXOBS], svr[MAXOBS];
    double wsum = 0.0, csum = 0.0;
    for (int j = 0; j < m; ++j) {
      const int ok = los(p, sat + 3 * j, l[j]);
      w[j] = ok ? (w_in[j] > 0.0 ? w_in[j] : 0.0) : 0.0;
      if (!ok) l[j][0] = l[j][1] = l[j][2] = 0.0;
      obs[j] = -dop[j] * wl;
      svr[j] = l[j][0] * sv[3 * j] + l[j][1] * sv[3 * j + 1] + l[j][2] * sv[3 * j + 2];
      const double rxr = l[j][0] * v[0] + l[j][1] * v[1] + l[j][2] * v[2];
      wsum += w[j];
      csum += w[j] * (obs[j] - svr[j] + rxr);
    }
    if (!(wsum > 0.0)) {
      out[i] = 0.0;
      continue;
    }
    const double cd = csum / wsum;
    double add = 0.0;
    for (int j = 0; j < m; ++j) {
      const double rxr = l[j][0] * v[0] + l[j][1] * v[1] + l[j][2] * v[2];
      const double res = obs[j] - (svr[j] - rxr + cd);
      add += -0.5 * w[j] * res * res / (sigma * sigma);
    }
    out[i] = add;
    if (g > 0.0 && m >= 4) {
      double A[4][5];
      memset(A, 0, sizeof(A));
      for (int j = 0; j < m; ++j) {
        const double h[4] = {-l[j][0], -l[j][1], -l[j][2], 1.0};
        const double rhs = obs[j] - svr[j];
        for (int a = 0; a < 4; ++a) {
          for (int b = 0; b < 4; ++b) A[a][b] += w[j] * h[a] * h[b];
          A[a][4] += w[j] * h[a] * rhs;
        }
      }
      for (int a = 0; a < 4; ++a) A[a][a] += 1e-9;
      double x[4];
      if (!solve4(A, x)) continue;
      double dv[3] = {x[0] - v[0], x[1] - v[1], x[2] - v[2]};
      const double nrm = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
      if (max_dv > 0.0 && nrm > max_dv) {
        const double f = max_dv / (nrm > 1e-9 ? nrm : 1e-9);
        for (int k = 0; k < 3; ++k) dv[k] *= f;
      }
      for (int k = 0; k < 3; ++k) v[k] += g * dv[k];
    }
  }
}

void cpu_doppler_kf_update(int n, const double* pos, double* vel, double* cov, int m,
                           const double* sat, const double* sv, const double* dop,
                           const double* w_in, double wl, double sigma, double* out,
                           int threads) {
  set_threads(threads);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const double* p = pos + 3 * (size_t)i;
    double* mu = vel + 3 * (size_t)i;
    double* P = cov + 9 * (size_t)i;
    double l[MAXOBS][3], w[MAXOBS], y[MAXOBS];
    double wsum = 0.0, hm[3] = {0.0, 0.0, 0.0}, ym = 0.0;
    for (int j = 0; j < m; ++j) {
      const int ok = los(p, sat + 3 * j, l[j]);
      w[j] = (ok && w_in[j] > 0.0) ? w_in[j] : 0.0;
      const double svr = ok ? l[j][0] * sv[3 * j] + l[j][1] * sv[3 * j + 1] +
                                  l[j][2] * sv[3 * j + 2]
                            : 0.0;
      y[j] = -dop[j] * wl - svr;
      wsum += w[j];
      if (w[j] > 0.0) {
        for (int k = 0; k < 3; ++k) hm[k] += w[j] * -l[j][k];
        ym += w[j] * y[j];
      }
    }
    double add = 0.0;
    if (wsum > 0.0) {
      for (int k = 0; k < 3; ++k) hm[k] /= wsum;
      ym /= wsum;
      for (int j = 0; j < m; ++j) {
        if (!(w[j] > 0.0)) continue;
        const double h[3] = {-l[j][0] - hm[0], -l[j][1] - hm[1], -l[j][2] - hm[2]};
        const double yc = y[j] - ym;
        double ph[3];
        for (int a = 0; a < 3; ++a) ph[a] = P[3 * a] * h[0] + P[3 * a + 1] * h[1] + P[3 * a + 2] * h[2];
        const double S = h[0] * ph[0] + h[1] * ph[1] + h[2] * ph[2] + sigma * sigma / w[j];
        if (!(S > 1e-18 && isfinite(S))) continue;
        const double innov = yc - (h[0] * mu[0] + h[1] * mu[1] + h[2] * mu[2]);
        double K[3], hp[3];
        for (int a = 0; a < 3; ++a) K[a] = ph[a] / S;
        for (int b = 0; b < 3; ++b) hp[b] = h[0] * P[b] + h[1] * P[3 + b] + h[2] * P[6 + b];
        for (int a = 0; a < 3; ++a) mu[a] += K[a] * innov;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) P[3 * a + b] -= K[a] * hp[b];
        for (int a = 0; a < 3; ++a)
          for (int b = a + 1; b < 3; ++b) {
            const double s = 0.5 * (P[3 * a + b] + P[3 * b + a]);
            P[3 * a + b] = s;
            P[3 * b + a] = s;
          }
        for (int d = 0; d < 3; ++d) {
          const double v = P[4 * d];
          P[4 * d] = (isfinite(v) && v > 1e-12) ? v : 1e-12;
        }
        add += -0.5 * (innov * innov / S + log(S));
      }
    }
    out[i] = add;
  }
}
