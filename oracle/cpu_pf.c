/*
 * CPU restatement of the reference particle-filter step, used as the timed
 * CPU baseline (TEST/MEASUREMENT INFRASTRUCTURE ONLY; the product never links
 * it).  binary64 structure-of-arrays state [x y z cb isb...], like the
 * reference's device state (src/particle_filter/pf_device.cu), OpenMP over
 * particles.
 *
 *   predict   pf_device.cu:106-187   random walk, velocity guide (alpha = 1)
 *   weight    pf_device.cu:189-276   Gaussian (nu <= 0) or Student-t / Cauchy
 *   ESS       pf_device.cu:924-965 + 1722-1762
 *   systematic pf_device.cu:1813-1878 (normalise, inclusive scan, search)
 *   estimate  pf_device.cu:968-1023 + 1926-1969
 *
 * Normals come from the same Philox4x32-10 streams as the GPU library
 * (binary64 Box-Muller), so a step costs what the reference's cuRAND-based
 * step costs on a CPU.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  uint32_t v[4];
} u4;

static inline u4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                        uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  u4 o = {{c0, c1, c2, c3}};
  return o;
}

static inline double u01(uint32_t x) { return ((double)(x >> 9) + 0.5) * (1.0 / 8388608.0); }

static inline void normal4(uint64_t seed, uint32_t idx, uint32_t step, uint32_t call, double z[4]) {
  const u4 r = philox(idx, step, 2u, call, (uint32_t)seed, (uint32_t)(seed >> 32));
  for (int k = 0; k < 2; ++k) {
    const double a = u01(r.v[2 * k]), b = u01(r.v[2 * k + 1]);
    const double rad = sqrt(-2.0 * log(a));
    z[2 * k] = rad * cos(2.0 * M_PI * b);
    z[2 * k + 1] = rad * sin(2.0 * M_PI * b);
  }
}

typedef struct {
  int n, n_isb, n_threads;
  double *st[7], *tmp[7];  /* x y z cb isb0..2 */
  double *lw, *cdf;
  uint64_t seed;
} cpu_pf;

int cpu_pf_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

cpu_pf* cpu_pf_create(int n, int n_isb, uint64_t seed, int n_threads) {
  cpu_pf* h = (cpu_pf*)calloc(1, sizeof(cpu_pf));
  h->n = n;
  h->n_isb = n_isb;
  h->seed = seed;
  h->n_threads = n_threads;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  for (int k = 0; k < 4 + n_isb; ++k) {
    h->st[k] = (double*)malloc(sizeof(double) * n);
    h->tmp[k] = (double*)malloc(sizeof(double) * n);
  }
  h->lw = (double*)malloc(sizeof(double) * n);
  h->cdf = (double*)malloc(sizeof(double) * n);
  return h;
}

void cpu_pf_destroy(cpu_pf* h) {
  if (!h) return;
  for (int k = 0; k < 7; ++k) {
    free(h->st[k]);
    free(h->tmp[k]);
  }
  free(h->lw);
  free(h->cdf);
  free(h);
}

void cpu_pf_initialize(cpu_pf* h, const double* init /*4+n_isb*/, double spread_pos,
                       double spread_cb, double spread_isb) {
  const int D = 4 + h->n_isb;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < h->n; ++i) {
    double z[4];
    normal4(h->seed, (uint32_t)i, 0u, 0u, z);
    h->st[0][i] = init[0] + spread_pos * z[0];
    h->st[1][i] = init[1] + spread_pos * z[1];
    h->st[2][i] = init[2] + spread_pos * z[2];
    h->st[3][i] = init[3] + spread_cb * z[3];
    if (D > 4) {
      normal4(h->seed, (uint32_t)i, 0u, 1u, z);
      for (int k = 4; k < D; ++k) h->st[k][i] = init[k] + spread_isb * z[k - 4];
    }
    h->lw[i] = 0.0;
  }
}

/* One step: predict, weight, ESS, systematic resample if ESS < frac*N, estimate.
 * Returns 1 if it resampled; est[4+n_isb] gets the weighted mean. */
int cpu_pf_step(cpu_pf* h, const double* vel, double dt, double sigma_pos, double sigma_cb,
                double sigma_isb, int step, const double* sat /*S*3*/, const double* pr,
                const int* sys, int n_sat, double sigma_pr, double nu, double ess_frac,
                double* est) {
  const int n = h->n, D = 4 + h->n_isb;
  double** s = h->st;
  /* predict + weight (one pass over the particles) */
  const double inv_s2 = 1.0 / (sigma_pr * sigma_pr);
  const double half_nup1 = 0.5 * (nu + 1.0);
  const double inv_nu_s2 = nu > 0.0 ? 1.0 / (nu * sigma_pr * sigma_pr) : 0.0;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double z[4];
    normal4(h->seed, (uint32_t)i, (uint32_t)step, 0u, z);
    s[0][i] += vel[0] * dt + sigma_pos * z[0];
    s[1][i] += vel[1] * dt + sigma_pos * z[1];
    s[2][i] += vel[2] * dt + sigma_pos * z[2];
    s[3][i] += sigma_cb * z[3];
    if (D > 4) {
      normal4(h->seed, (uint32_t)i, (uint32_t)step, 2u, z);
      for (int k = 4; k < D; ++k) s[k][i] += sigma_isb * z[k - 4];
    }
    double lw = 0.0;
    for (int q = 0; q < n_sat; ++q) {
      const double dx = s[0][i] - sat[3 * q], dy = s[1][i] - sat[3 * q + 1],
                   dz = s[2][i] - sat[3 * q + 2];
      double pred = sqrt(dx * dx + dy * dy + dz * dz) + s[3][i];
      if (sys && sys[q] >= 1 && sys[q] <= h->n_isb) pred += s[3 + sys[q]][i];
      const double r = pr[q] - pred;
      if (nu > 0.0)
        lw -= half_nup1 * log(1.0 + r * r * inv_nu_s2);
      else
        lw -= 0.5 * r * r * inv_s2;
    }
    h->lw[i] = lw;
  }
  /* ESS */
  double m = -INFINITY;
#pragma omp parallel for reduction(max : m)
  for (int i = 0; i < n; ++i) m = h->lw[i] > m ? h->lw[i] : m;
  double sw = 0.0, sw2 = 0.0;
#pragma omp parallel for reduction(+ : sw, sw2)
  for (int i = 0; i < n; ++i) {
    const double w = exp(h->lw[i] - m);
    h->cdf[i] = w;
    sw += w;
    sw2 += w * w;
  }
  const double ess = sw * sw / sw2;
  int resampled = 0;
  if (ess < ess_frac * n) {
    /* normalise + inclusive scan (two-pass parallel prefix) */
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    double* part = (double*)calloc((size_t)nt + 1, sizeof(double));
#pragma omp parallel
    {
      int t = 0;
#ifdef _OPENMP
      t = omp_get_thread_num();
#endif
      const int lo = (int)((int64_t)n * t / nt), hi = (int)((int64_t)n * (t + 1) / nt);
      double run = 0.0;
      for (int i = lo; i < hi; ++i) {
        run += h->cdf[i] / sw;
        h->cdf[i] = run;
      }
      part[t + 1] = run;
#pragma omp barrier
#pragma omp single
      for (int k = 1; k <= nt; ++k) part[k] += part[k - 1];
      for (int i = lo; i < hi; ++i) h->cdf[i] += part[t];
    }
    free(part);
    const int32_t den = (int32_t)(uint32_t)((uint64_t)1000000u * (uint64_t)(uint32_t)n);
    const double u0 = (double)(h->seed % 1000000ull) / (double)den;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j) {
      const double t = (u0 + (double)j) / (double)n;
      int lo = 0, hi = n - 1;
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (h->cdf[mid] < t)
          lo = mid + 1;
        else
          hi = mid;
      }
      for (int k = 0; k < D; ++k) h->tmp[k][j] = s[k][lo];
    }
    for (int k = 0; k < D; ++k) {
      double* t = s[k];
      s[k] = h->tmp[k];
      h->tmp[k] = t;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) h->lw[i] = 0.0;
    resampled = 1;
    m = 0.0;
  }
  /* estimate */
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0, a6 = 0, aw = 0;
#pragma omp parallel for reduction(+ : a0, a1, a2, a3, a4, a5, a6, aw)
  for (int i = 0; i < n; ++i) {
    const double w = exp(h->lw[i] - m);
    aw += w;
    a0 += w * s[0][i];
    a1 += w * s[1][i];
    a2 += w * s[2][i];
    a3 += w * s[3][i];
    if (D > 4) a4 += w * s[4][i];
    if (D > 5) a5 += w * s[5][i];
    if (D > 6) a6 += w * s[6][i];
  }
  acc[0] = a0;
  acc[1] = a1;
  acc[2] = a2;
  acc[3] = a3;
  acc[4] = a4;
  acc[5] = a5;
  acc[6] = a6;
  for (int k = 0; k < D; ++k) est[k] = acc[k] / aw;
  return resampled;
}
