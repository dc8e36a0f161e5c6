/* C/OpenMP restatement of the ray-traced NLOS weight over a BVH (TEST AND
 * MEASUREMENT INFRASTRUCTURE ONLY: the CPU baseline of the ray-traced weight and
 * a second check of the library's tree walk).  Binary64 throughout.
 *
 * Tree: the reference's flat node rows [n_nodes][10] = min xyz, max xyz, left,
 * right, tri_start, tri_count (left == -1 marks a leaf), triangles [n_tri][3][3]
 * in the tree's order (src/raytrace/bvh.cu, BVHNode).
 *
 * cpu_weight_3d_bvh follows src/particle_filter/weight_3d_bvh.cu:
 *   slab test            :20-33  (tmin 0, tmax = range to the satellite)
 *   Moller-Trumbore      :35-57  (hit when 1e-6 < t < range)
 *   mixture              :59-72
 *   any-hit tree walk    :79-112 (a popped node is tested against its own box;
 *                                 both children pushed while the stack has room)
 *   per-particle weight  :157-219 (satellites summed in order)
 * The edge vectors e1 = v1 - v0 and e2 = v2 - v0 are formed once per triangle
 * instead of once per test: the same binary64 values.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>

#define RT_STACK 64

int cpu_raytrace_threads(void) { return omp_get_max_threads(); }

static int slab(const double o[3], const double inv[3], double tmax, const double* row) {
  double t0 = 0.0, t1 = tmax;
  for (int k = 0; k < 3; ++k) {
    double a = (row[k] - o[k]) * inv[k];
    double b = (row[3 + k] - o[k]) * inv[k];
    if (a > b) {
      const double t = a;
      a = b;
      b = t;
    }
    if (a > t0) t0 = a;
    if (b < t1) t1 = b;
    if (t0 > t1) return 0;
  }
  return 1;
}

/* tri: v0, e1, e2 (9 doubles) */
static int mt_blocks(const double o[3], const double d[3], const double* tri, double tmax) {
  const double* v0 = tri;
  const double* e1 = tri + 3;
  const double* e2 = tri + 6;
  const double h0 = d[1] * e2[2] - d[2] * e2[1];
  const double h1 = d[2] * e2[0] - d[0] * e2[2];
  const double h2 = d[0] * e2[1] - d[1] * e2[0];
  const double a = e1[0] * h0 + e1[1] * h1 + e1[2] * h2;
  if (fabs(a) < 1e-12) return 0;
  const double f = 1.0 / a;
  const double s0 = o[0] - v0[0], s1 = o[1] - v0[1], s2 = o[2] - v0[2];
  const double u = f * (s0 * h0 + s1 * h1 + s2 * h2);
  if (u < 0.0 || u > 1.0) return 0;
  const double q0 = s1 * e1[2] - s2 * e1[1];
  const double q1 = s2 * e1[0] - s0 * e1[2];
  const double q2 = s0 * e1[1] - s1 * e1[0];
  const double v = f * (d[0] * q0 + d[1] * q1 + d[2] * q2);
  if (v < 0.0 || u + v > 1.0) return 0;
  const double t = f * (e2[0] * q0 + e2[1] * q1 + e2[2] * q2);
  return t > 1e-6 && t < tmax;
}

static int blocked(const double o[3], const double d[3], const double inv[3], double dist,
                   const double* nodes, const int32_t* link, const double* tri) {
  int stack[RT_STACK];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const int i = stack[--sp];
    if (!slab(o, inv, dist, nodes + (size_t)i * 10)) continue;
    const int32_t* l = link + (size_t)i * 4;
    if (l[0] == -1) {
      for (int k = 0; k < l[3]; ++k)
        if (mt_blocks(o, d, tri + (size_t)(l[2] + k) * 9, dist)) return 1;
    } else if (sp < RT_STACK - 1) {
      stack[sp++] = l[0];
      stack[sp++] = l[1];
    }
  }
  return 0;
}

static double mixture(double los, double nlos, double p) {
  if (p <= 0.0) return los;
  if (p >= 1.0) return nlos;
  const double a = log(1.0 - p) + los, b = log(p) + nlos;
  const double m = a > b ? a : b;
  return m + log(exp(a - m) + exp(b - m));
}

/* Log-weights lw[n] on `threads` OpenMP threads (0: the runtime's default).
 * Returns 0, or -1 for a malformed tree (a child or leaf range outside the
 * arrays) or an allocation failure. */
int cpu_weight_3d_bvh(const double* px, const double* py, const double* pz, const double* pcb,
                      int n, const double* sat, const double* pr, const double* ws, int n_sat,
                      const double* nodes, int n_nodes, const double* tris, int n_tri,
                      double sigma_los, double sigma_nlos, double nlos_bias,
                      double blocked_nlos_prob, double clear_nlos_prob, double* lw,
                      int threads) {
  int32_t* link = NULL;
  double* tri = NULL;
  if (n_nodes > 0) {
    link = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)n_nodes);
    tri = (double*)malloc(sizeof(double) * 9 * (size_t)(n_tri > 0 ? n_tri : 1));
    if (!link || !tri) {
      free(link);
      free(tri);
      return -1;
    }
    for (int i = 0; i < n_nodes; ++i) {
      const double* r = nodes + (size_t)i * 10;
      for (int k = 0; k < 4; ++k) link[(size_t)i * 4 + k] = (int32_t)r[6 + k];
      const int32_t* l = link + (size_t)i * 4;
      const int bad = l[0] == -1
          ? (l[2] < 0 || l[3] < 0 || (int64_t)l[2] + l[3] > n_tri)
          : (l[0] < 0 || l[0] >= n_nodes || l[1] < 0 || l[1] >= n_nodes);
      if (bad) {
        free(link);
        free(tri);
        return -1;
      }
    }
    for (int i = 0; i < n_tri; ++i) {
      const double* v = tris + (size_t)i * 9;
      double* o = tri + (size_t)i * 9;
      for (int k = 0; k < 3; ++k) {
        o[k] = v[k];
        o[3 + k] = v[3 + k] - v[k];
        o[6 + k] = v[6 + k] - v[k];
      }
    }
  }
  const double il2 = 1.0 / (sigma_los * sigma_los);
  const double in2 = 1.0 / (sigma_nlos * sigma_nlos);
  const int nthr = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthr)
  for (int i = 0; i < n; ++i) {
    const double o[3] = {px[i], py[i], pz[i]};
    double acc = 0.0;
    for (int s = 0; s < n_sat; ++s) {
      const double dx = sat[3 * s] - o[0], dy = sat[3 * s + 1] - o[1], dz = sat[3 * s + 2] - o[2];
      const double dist = sqrt(dx * dx + dy * dy + dz * dz);
      const double id = 1.0 / dist;
      const double d[3] = {dx * id, dy * id, dz * id};
      double inv[3];
      for (int k = 0; k < 3; ++k)
        inv[k] = fabs(d[k]) > 1e-20 ? 1.0 / d[k] : (d[k] >= 0.0 ? 1e20 : -1e20);
      const int nl = n_nodes > 0 && blocked(o, d, inv, dist, nodes, link, tri);
      const double res = pr[s] - (dist + pcb[i]);
      const double los = -0.5 * ws[s] * res * res * il2;
      const double rn = res > 0.0 ? res - nlos_bias : res;
      const double nlos = -0.5 * ws[s] * rn * rn * in2;
      acc += mixture(los, nlos, nl ? blocked_nlos_prob : clear_nlos_prob);
    }
    lw[i] = acc;
  }
  free(link);
  free(tri);
  return 0;
}
