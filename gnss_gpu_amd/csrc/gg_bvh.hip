// Ray casting through a building mesh on gfx950: BVH build (host), LOS and
// first-order multipath queries, and the ray-traced NLOS pseudorange weight.
//
// Boundary.  Trees cross the C ABI in the reference's flat node format
// (python/gnss_gpu/_bvh_bindings.cpp:12-49): [n_nodes][10] binary64 rows
// min_xyz, max_xyz, left, right, tri_start, tri_count; left == -1 marks a
// leaf; node 0 is the root; triangles are given in the tree's order.  Any
// valid tree is accepted (this library's gg_bvh_build or another builder);
// it is checked on the host (a proper tree reachable from the root, leaf
// ranges inside the triangle array, depth <= 62) before anything reaches the
// GPU, so a malformed tree is an error, never a fault or an endless walk.
//
// Device layout and the ray tests: gg_bvh.h.
//
// Arithmetic is binary64 throughout, as in the reference kernels
// (src/raytrace/bvh.cu, src/particle_filter/weight_3d_bvh.cu): in ECEF the
// coordinates are ~6.4e6 m, where binary32 would resolve only ~0.5 m.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gg_bvh.h"
#include "gg_common_host.h"
#include "gnss_gpu_amd.h"

namespace {

using namespace gg_rt;

// ---------------------------------------------------------------- build --
struct Prim {
  double lo[3], hi[3], c[3];
  int32_t id;
};

struct Flat {
  double lo[3], hi[3];
  int32_t left, right, start, count;
};

constexpr int kLeafMax = 4;
constexpr int kBins = 16;
constexpr int kSahDepth = 32;  // below this depth the split is the median by count

double half_area(const double lo[3], const double hi[3]) {
  const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  return dx * dy + dy * dz + dz * dx;
}

void grow(double lo[3], double hi[3], const double plo[3], const double phi[3]) {
  for (int k = 0; k < 3; ++k) {
    lo[k] = std::min(lo[k], plo[k]);
    hi[k] = std::max(hi[k], phi[k]);
  }
}

int bin_of(double c, double clo, double scale) {
  int b = (int)((c - clo) * scale);
  return b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
}

// Top-down binned SAH (16 bins, all three axes) on triangle centroids; past
// depth 32, or when no bin boundary separates the centroids, a median split
// along the widest centroid axis keeps the depth within 32 + log2(n / 4).
int build_node(std::vector<Prim>& p, int s, int e, int depth, std::vector<Flat>& out) {
  Flat nd;
  double clo[3], chi[3];
  for (int k = 0; k < 3; ++k) {
    nd.lo[k] = clo[k] = HUGE_VAL;
    nd.hi[k] = chi[k] = -HUGE_VAL;
  }
  for (int i = s; i < e; ++i) {
    grow(nd.lo, nd.hi, p[i].lo, p[i].hi);
    grow(clo, chi, p[i].c, p[i].c);
  }
  const int idx = (int)out.size();
  out.push_back(nd);
  const int n = e - s;
  if (n <= kLeafMax) {
    out[idx].left = out[idx].right = -1;
    out[idx].start = s;
    out[idx].count = n;
    return idx;
  }
  int mid = -1;
  if (depth < kSahDepth) {
    double best = HUGE_VAL;
    int best_axis = -1, best_split = 0;
    for (int a = 0; a < 3; ++a) {
      const double ext = chi[a] - clo[a];
      if (!(ext > 0.0)) continue;
      const double scale = kBins / ext;
      int cnt[kBins] = {0};
      double blo[kBins][3], bhi[kBins][3];
      for (int b = 0; b < kBins; ++b)
        for (int k = 0; k < 3; ++k) {
          blo[b][k] = HUGE_VAL;
          bhi[b][k] = -HUGE_VAL;
        }
      for (int i = s; i < e; ++i) {
        const int b = bin_of(p[i].c[a], clo[a], scale);
        ++cnt[b];
        grow(blo[b], bhi[b], p[i].lo, p[i].hi);
      }
      double rcost[kBins];
      double rlo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, rhi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
      int rc = 0;
      for (int b = kBins - 1; b >= 1; --b) {
        rc += cnt[b];
        grow(rlo, rhi, blo[b], bhi[b]);
        rcost[b] = rc ? half_area(rlo, rhi) * rc : 0.0;
      }
      double llo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, lhi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
      int lc = 0;
      for (int b = 1; b < kBins; ++b) {  // split before bin b
        lc += cnt[b - 1];
        grow(llo, lhi, blo[b - 1], bhi[b - 1]);
        if (lc == 0 || lc == n) continue;
        const double cost = half_area(llo, lhi) * lc + rcost[b];
        if (cost < best) {
          best = cost;
          best_axis = a;
          best_split = b;
        }
      }
    }
    if (best_axis >= 0) {
      const double scale = kBins / (chi[best_axis] - clo[best_axis]);
      const double c0 = clo[best_axis];
      mid = (int)(std::partition(p.begin() + s, p.begin() + e,
                                 [&](const Prim& q) {
                                   return bin_of(q.c[best_axis], c0, scale) < best_split;
                                 }) -
                  p.begin());
    }
  }
  if (mid <= s || mid >= e) {
    int a = 0;
    for (int k = 1; k < 3; ++k)
      if (chi[k] - clo[k] > chi[a] - clo[a]) a = k;
    mid = s + n / 2;
    std::nth_element(p.begin() + s, p.begin() + mid, p.begin() + e,
                     [a](const Prim& x, const Prim& y) {
                       return x.c[a] < y.c[a] || (x.c[a] == y.c[a] && x.id < y.id);
                     });
  }
  const int l = build_node(p, s, mid, depth + 1, out);
  const int r = build_node(p, mid, e, depth + 1, out);
  out[idx].left = l;
  out[idx].right = r;
  out[idx].start = 0;
  out[idx].count = 0;
  return idx;
}

// ----------------------------------------------------- flat -> device --
int parse_tree(const double* flat, int n_nodes, int n_tri, std::vector<gg_bvh_node>& dev) {
  dev.clear();
  if (n_nodes <= 0) return 0;
  std::vector<Flat> f((size_t)n_nodes);
  for (int i = 0; i < n_nodes; ++i) {
    const double* r = flat + (size_t)i * 10;
    for (int k = 0; k < 3; ++k) {
      f[i].lo[k] = r[k];
      f[i].hi[k] = r[3 + k];
    }
    for (int k = 6; k < 10; ++k)
      if (!(std::fabs(r[k]) < 2147483647.0) || r[k] != std::trunc(r[k]))
        return gg_fail(-2, "BVH node indices must be integers");
    f[i].left = (int32_t)r[6];
    f[i].right = (int32_t)r[7];
    f[i].start = (int32_t)r[8];
    f[i].count = (int32_t)r[9];
  }
  auto leaf_ok = [&](const Flat& x) {
    return x.start >= 0 && x.count >= 0 && (int64_t)x.start + x.count <= n_tri;
  };
  // depth-first walk from the root: every node reached once, depth bounded
  std::vector<int32_t> map((size_t)n_nodes, -1);
  std::vector<char> seen((size_t)n_nodes, 0);
  std::vector<std::pair<int, int>> todo;
  todo.push_back({0, 0});
  seen[0] = 1;
  int n_internal = 0;
  while (!todo.empty()) {
    const auto [i, d] = todo.back();
    todo.pop_back();
    if (d > GG_BVH_MAX_DEPTH) return gg_fail(-2, "BVH deeper than 62 levels");
    if (f[i].left == -1) {
      if (!leaf_ok(f[i])) return gg_fail(-2, "BVH leaf range outside the triangle array");
      continue;
    }
    map[i] = n_internal++;
    for (int ch : {f[i].left, f[i].right}) {
      if (ch < 0 || ch >= n_nodes || seen[ch]) return gg_fail(-2, "BVH child index invalid");
      seen[ch] = 1;
      todo.push_back({ch, d + 1});
    }
  }
  auto set_child = [&](gg_bvh_node& dn, int side, const Flat& ch, int ci) {
    for (int k = 0; k < 3; ++k) {
      dn.b[side * 6 + k] = ch.lo[k];
      dn.b[side * 6 + 3 + k] = ch.hi[k];
    }
    if (ch.left == -1) {
      dn.c[side] = ch.count > 0 ? ch.start : -1;
      dn.c[2 + side] = ch.count;
    } else {
      dn.c[side] = map[ci];
      dn.c[2 + side] = 0;
    }
  };
  if (f[0].left == -1) {  // the whole tree is one leaf: a root holding it
    gg_bvh_node dn;
    std::memset(&dn, 0, sizeof(dn));
    set_child(dn, 0, f[0], 0);
    dn.c[1] = -1;
    dn.c[3] = 0;
    dev.push_back(dn);
    return 0;
  }
  dev.assign((size_t)n_internal, gg_bvh_node{});
  for (int i = 0; i < n_nodes; ++i) {
    if (map[i] < 0) continue;
    gg_bvh_node& dn = dev[map[i]];
    set_child(dn, 0, f[f[i].left], f[i].left);
    set_child(dn, 1, f[f[i].right], f[i].right);
  }
  return 0;
}

void pack_tris(const double* tris, int n_tri, std::vector<double>& out) {
  out.assign((size_t)n_tri * 10, 0.0);
  for (int i = 0; i < n_tri; ++i) {
    const double* t = tris + (size_t)i * 9;
    double* o = out.data() + (size_t)i * 10;
    for (int k = 0; k < 3; ++k) {
      o[k] = t[k];
      o[3 + k] = t[3 + k] - t[k];
      o[6 + k] = t[6 + k] - t[k];
    }
  }
}

// LOS flags (bvh.cu:293-354, 517-585): one thread per (epoch, satellite).
// check_finite: a non-finite coordinate gives 0 (the batch query's rule).
__global__ __launch_bounds__(GG_BVH_BLOCK) void gg_k_los(const double* __restrict__ rx,
                                                         const double* __restrict__ sat,
                                                         int n_epoch, int n_sat,
                                                         const gg_bvh_node* __restrict__ nodes,
                                                         const double* __restrict__ tris,
                                                         int n_nodes, int check_finite,
                                                         int32_t* __restrict__ is_los) {
  const int64_t tid = (int64_t)blockIdx.x * GG_BVH_BLOCK + threadIdx.x;
  if (tid >= (int64_t)n_epoch * n_sat) return;
  const int64_t e = tid / n_sat;
  Ray r;
  double s[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    r.o[k] = rx[e * 3 + k];
    s[k] = sat[tid * 3 + k];
  }
  if (check_finite) {
    bool fin = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) fin = fin && isfinite(r.o[k]) && isfinite(s[k]);
    if (!fin) {
      is_los[tid] = 0;
      return;
    }
  }
  double d0 = s[0] - r.o[0], d1 = s[1] - r.o[1], d2 = s[2] - r.o[2];
  const double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  if (dist < 1e-15 || n_nodes == 0) {
    is_los[tid] = 1;
    return;
  }
  r.d[0] = d0 / dist;
  r.d[1] = d1 / dist;
  r.d[2] = d2 / dist;
  r.tmax = dist;
  make_inv(r);
  is_los[tid] = gg_rt::gg_bvh_blocked(r, 1e-12, nodes, tris) ? 0 : 1;
}

// First-order specular reflection with the smallest positive excess delay
// over every triangle (bvh.cu:362-462, 587-691): the receiver is mirrored in
// the triangle's plane and the mirror-to-satellite ray must cross the
// triangle.  One thread per (epoch, satellite); triangles in tree order.
__global__ __launch_bounds__(GG_BVH_BLOCK) void gg_k_multipath(
    const double* __restrict__ rx, const double* __restrict__ sat, int n_epoch, int n_sat,
    const double* __restrict__ tris, int n_tri, int check_finite, double* __restrict__ refl_out,
    double* __restrict__ delay_out) {
  const int64_t tid = (int64_t)blockIdx.x * GG_BVH_BLOCK + threadIdx.x;
  if (tid >= (int64_t)n_epoch * n_sat) return;
  const int64_t e = tid / n_sat;
  double p[3], s[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p[k] = rx[e * 3 + k];
    s[k] = sat[tid * 3 + k];
  }
  double best = 0.0, br[3] = {0.0, 0.0, 0.0};
  bool fin = true;
  if (check_finite) {
#pragma unroll
    for (int k = 0; k < 3; ++k) fin = fin && isfinite(p[k]) && isfinite(s[k]);
  }
  if (fin) {
    const double d_rs = sqrt((p[0] - s[0]) * (p[0] - s[0]) + (p[1] - s[1]) * (p[1] - s[1]) +
                             (p[2] - s[2]) * (p[2] - s[2]));
    bool found = false;
    for (int i = 0; i < n_tri; ++i) {
      double tr[10];
      load_tri(tris, i, tr);
      const double* v0 = tr;
      const double* e1 = tr + 3;
      const double* e2 = tr + 6;
      double n0 = e1[1] * e2[2] - e1[2] * e2[1];
      double n1 = e1[2] * e2[0] - e1[0] * e2[2];
      double n2 = e1[0] * e2[1] - e1[1] * e2[0];
      const double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      if (nn < 1e-15) continue;
      n0 /= nn;
      n1 /= nn;
      n2 /= nn;
      const double h = (p[0] - v0[0]) * n0 + (p[1] - v0[1]) * n1 + (p[2] - v0[2]) * n2;
      const double m[3] = {p[0] - 2.0 * h * n0, p[1] - 2.0 * h * n1, p[2] - 2.0 * h * n2};
      double d[3] = {s[0] - m[0], s[1] - m[1], s[2] - m[2]};
      const double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (dl < 1e-15) continue;
      d[0] /= dl;
      d[1] /= dl;
      d[2] /= dl;
      double t;
      if (!mt(m, d, tr, t) || !(t > 1e-12)) continue;
      const double q[3] = {m[0] + t * d[0], m[1] + t * d[1], m[2] + t * d[2]};
      const double a = sqrt((p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) +
                            (p[2] - q[2]) * (p[2] - q[2]));
      const double b = sqrt((q[0] - s[0]) * (q[0] - s[0]) + (q[1] - s[1]) * (q[1] - s[1]) +
                            (q[2] - s[2]) * (q[2] - s[2]));
      const double ex = (a + b) - d_rs;
      if (ex > 0.0 && (!found || ex < best)) {
        best = ex;
        br[0] = q[0];
        br[1] = q[1];
        br[2] = q[2];
        found = true;
      }
    }
  }
  delay_out[tid] = best;
#pragma unroll
  for (int k = 0; k < 3; ++k) refl_out[tid * 3 + k] = br[k];
}

// Ray-traced NLOS weight (weight_3d_bvh.cu:127-220): one thread per
// particle, satellites in the outer loop (their rows are wave-uniform, so
// they come through the scalar cache); log_weights are overwritten.
__global__ __launch_bounds__(GG_BVH_BLOCK) void gg_k_weight_3d(
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    const double* __restrict__ pcb, const double* __restrict__ sat,
    const double* __restrict__ pr, const double* __restrict__ ws,
    const gg_bvh_node* __restrict__ nodes, const double* __restrict__ tris, int n_nodes, int n,
    int n_sat, MixArgs mx, double* __restrict__ lw) {
  const int i = blockIdx.x * GG_BVH_BLOCK + threadIdx.x;
  if (i >= n) return;
  const double x = px[i], y = py[i], z = pz[i], cb = pcb[i];
  double acc = 0.0;
  for (int s = 0; s < n_sat; ++s) {
    Ray r;
    r.o[0] = x;
    r.o[1] = y;
    r.o[2] = z;
    const double dx = sat[s * 3 + 0] - x, dy = sat[s * 3 + 1] - y, dz = sat[s * 3 + 2] - z;
    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
    const double inv = 1.0 / dist;
    r.d[0] = dx * inv;
    r.d[1] = dy * inv;
    r.d[2] = dz * inv;
    r.tmax = dist;
    make_inv(r);
    const bool nl = n_nodes > 0 && gg_rt::gg_bvh_blocked(r, 1e-6, nodes, tris);
    const double res = pr[s] - (dist + cb);
    const double los = -0.5 * ws[s] * res * res * mx.inv_los2;
    const double rn = res > 0.0 ? res - mx.nlos_bias : res;
    const double nlos = -0.5 * ws[s] * rn * rn * mx.inv_nlos2;
    const int k = nl ? 1 : 0;
    acc += mixture(los, nlos, mx.mode[k], mx.lp_keep[k], mx.lp_nlos[k]);
  }
  lw[i] = acc;
}

int upload(const double* nodes_flat, int n_nodes, const double* tris, int n_tri, gg_bvh* b) {
  if (n_nodes < 0 || n_tri < 0) return gg_fail(-2, "negative BVH size");
  if ((n_nodes > 0 && !nodes_flat) || (n_tri > 0 && !tris)) return gg_fail(-2, "null BVH array");
  std::vector<gg_bvh_node> dev;
  if (int rc = parse_tree(nodes_flat, n_nodes, n_tri, dev)) return rc;
  std::vector<double> pt;
  pack_tris(tris, n_tri, pt);
  b->n_nodes = (int32_t)dev.size();
  b->n_tri = n_tri;
  if (!dev.empty()) {
    GG_HIP(hipMalloc((void**)&b->nodes, dev.size() * sizeof(gg_bvh_node)));
    GG_HIP(hipMemcpy(b->nodes, dev.data(), dev.size() * sizeof(gg_bvh_node),
                     hipMemcpyHostToDevice));
  }
  if (n_tri > 0) {
    GG_HIP(hipMalloc((void**)&b->tris, pt.size() * sizeof(double)));
    GG_HIP(hipMemcpy(b->tris, pt.data(), pt.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return 0;
}

void release(gg_bvh* b) {
  if (b->nodes) (void)hipFree(b->nodes);
  if (b->tris) (void)hipFree(b->tris);
  b->nodes = nullptr;
  b->tris = nullptr;
}

int grid_for(int64_t n) { return (int)((n + GG_BVH_BLOCK - 1) / GG_BVH_BLOCK); }

// Device buffer holding up to 16 host arrays for one call, each 256-aligned.
struct Scratch {
  char* base = nullptr;
  size_t off = 0;
  ~Scratch() {
    if (base) (void)hipFree(base);
  }
  int alloc(size_t bytes) {
    GG_HIP(hipMalloc((void**)&base, bytes + 256 * 16));
    return 0;
  }
  template <class T>
  T* take(size_t count) {
    T* p = (T*)(base + off);
    off += (count * sizeof(T) + 255) & ~(size_t)255;
    return p;
  }
};

int los_query(const double* rx, const double* sat, int n_epoch, int n_sat,
              const double* nodes_flat, int n_nodes, const double* tris, int n_tri,
              int check_finite, int32_t* is_los) {
  if (n_epoch < 0 || n_sat < 0) return gg_fail(-2, "negative size");
  if ((int64_t)n_epoch * n_sat == 0) return 0;
  gg_bvh b;
  int rc = upload(nodes_flat, n_nodes, tris, n_tri, &b);
  if (!rc) {
    const size_t nr = (size_t)n_epoch * n_sat;
    Scratch sc;
    rc = sc.alloc(sizeof(double) * (n_epoch * 3 + nr * 3) + sizeof(int32_t) * nr);
    if (!rc) {
      double* d_rx = sc.take<double>((size_t)n_epoch * 3);
      double* d_sat = sc.take<double>(nr * 3);
      int32_t* d_los = sc.take<int32_t>(nr);
      rc = hipMemcpy(d_rx, rx, sizeof(double) * n_epoch * 3, hipMemcpyHostToDevice) ||
                   hipMemcpy(d_sat, sat, sizeof(double) * nr * 3, hipMemcpyHostToDevice)
               ? gg_fail(-3, "host to device copy failed")
               : 0;
      if (!rc) {
        hipLaunchKernelGGL(gg_k_los, dim3(grid_for((int64_t)nr)), dim3(GG_BVH_BLOCK), 0, nullptr,
                           d_rx, d_sat, n_epoch, n_sat, b.nodes, b.tris, b.n_nodes,
                           check_finite, d_los);
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) rc = gg_fail_hip(le, "gg_k_los", __FILE__, __LINE__);
      }
      if (!rc) {
        const hipError_t ce = hipMemcpy(is_los, d_los, sizeof(int32_t) * nr, hipMemcpyDeviceToHost);
        if (ce != hipSuccess) rc = gg_fail_hip(ce, "D2H is_los", __FILE__, __LINE__);
      }
    }
  }
  release(&b);
  return rc;
}

int multipath_query(const double* rx, const double* sat, int n_epoch, int n_sat,
                    const double* nodes_flat, int n_nodes, const double* tris, int n_tri,
                    int check_finite, double* refl, double* delay) {
  if (n_epoch < 0 || n_sat < 0) return gg_fail(-2, "negative size");
  if ((int64_t)n_epoch * n_sat == 0) return 0;
  gg_bvh b;
  int rc = upload(nodes_flat, n_nodes, tris, n_tri, &b);
  if (!rc) {
    const size_t nr = (size_t)n_epoch * n_sat;
    Scratch sc;
    rc = sc.alloc(sizeof(double) * (n_epoch * 3 + nr * 7));
    if (!rc) {
      double* d_rx = sc.take<double>((size_t)n_epoch * 3);
      double* d_sat = sc.take<double>(nr * 3);
      double* d_refl = sc.take<double>(nr * 3);
      double* d_delay = sc.take<double>(nr);
      rc = hipMemcpy(d_rx, rx, sizeof(double) * n_epoch * 3, hipMemcpyHostToDevice) ||
                   hipMemcpy(d_sat, sat, sizeof(double) * nr * 3, hipMemcpyHostToDevice)
               ? gg_fail(-3, "host to device copy failed")
               : 0;
      if (!rc) {
        hipLaunchKernelGGL(gg_k_multipath, dim3(grid_for((int64_t)nr)), dim3(GG_BVH_BLOCK), 0,
                           nullptr, d_rx, d_sat, n_epoch, n_sat, b.tris, b.n_tri, check_finite,
                           d_refl, d_delay);
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) rc = gg_fail_hip(le, "gg_k_multipath", __FILE__, __LINE__);
      }
      if (!rc && (hipMemcpy(refl, d_refl, sizeof(double) * nr * 3, hipMemcpyDeviceToHost) ||
                  hipMemcpy(delay, d_delay, sizeof(double) * nr, hipMemcpyDeviceToHost)))
        rc = gg_fail(-3, "device to host copy failed");
    }
  }
  release(&b);
  return rc;
}

}  // namespace

extern "C" {

int gg_bvh_build(const double* triangles, int32_t n_tri, double* nodes_flat, int32_t* n_nodes,
                 int32_t* sorted_tri_indices) {
  if (n_tri < 0 || !n_nodes) return gg_fail(-2, "invalid arguments");
  *n_nodes = 0;
  if (n_tri == 0) return 0;
  if (!triangles || !nodes_flat || !sorted_tri_indices) return gg_fail(-2, "null array");
  std::vector<Prim> p((size_t)n_tri);
  for (int i = 0; i < n_tri; ++i) {
    const double* t = triangles + (size_t)i * 9;
    for (int k = 0; k < 3; ++k) {
      p[i].lo[k] = std::min(t[k], std::min(t[3 + k], t[6 + k]));
      p[i].hi[k] = std::max(t[k], std::max(t[3 + k], t[6 + k]));
      p[i].c[k] = (t[k] + t[3 + k] + t[6 + k]) / 3.0;
      if (!std::isfinite(p[i].c[k])) return gg_fail(-2, "non-finite triangle vertex");
    }
    p[i].id = i;
  }
  std::vector<Flat> out;
  out.reserve((size_t)2 * n_tri);
  build_node(p, 0, n_tri, 0, out);
  for (size_t i = 0; i < out.size(); ++i) {
    double* r = nodes_flat + i * 10;
    for (int k = 0; k < 3; ++k) {
      r[k] = out[i].lo[k];
      r[3 + k] = out[i].hi[k];
    }
    r[6] = out[i].left;
    r[7] = out[i].right;
    r[8] = out[i].start;
    r[9] = out[i].count;
  }
  for (int i = 0; i < n_tri; ++i) sorted_tri_indices[i] = p[i].id;
  *n_nodes = (int32_t)out.size();
  return 0;
}

int gg_bvh_create(const double* nodes_flat, int32_t n_nodes, const double* sorted_tris,
                  int32_t n_tri, gg_bvh** out) {
  if (!out) return gg_fail(-2, "null handle pointer");
  *out = nullptr;
  gg_bvh* b = new gg_bvh();
  const int rc = upload(nodes_flat, n_nodes, sorted_tris, n_tri, b);
  if (rc) {
    release(b);
    delete b;
    return rc;
  }
  *out = b;
  return 0;
}

int gg_bvh_destroy(gg_bvh* b) {
  if (!b) return 0;
  release(b);
  delete b;
  return 0;
}

int gg_raytrace_los_check_bvh(const double* rx_ecef, const double* sat_ecef,
                              const double* nodes_flat, int32_t n_nodes,
                              const double* sorted_tris, int32_t n_tri, int32_t* is_los,
                              int32_t n_sat) {
  return los_query(rx_ecef, sat_ecef, 1, n_sat, nodes_flat, n_nodes, sorted_tris, n_tri, 0,
                   is_los);
}

int gg_raytrace_los_check_bvh_batch(const double* rx_ecef, const double* sat_ecef,
                                    const double* nodes_flat, int32_t n_nodes,
                                    const double* sorted_tris, int32_t n_tri, int32_t* is_los,
                                    int32_t n_epoch, int32_t n_sat) {
  return los_query(rx_ecef, sat_ecef, n_epoch, n_sat, nodes_flat, n_nodes, sorted_tris, n_tri,
                   1, is_los);
}

int gg_raytrace_multipath_bvh(const double* rx_ecef, const double* sat_ecef,
                              const double* nodes_flat, int32_t n_nodes,
                              const double* sorted_tris, int32_t n_tri,
                              double* reflection_points, double* excess_delays, int32_t n_sat) {
  return multipath_query(rx_ecef, sat_ecef, 1, n_sat, nodes_flat, n_nodes, sorted_tris, n_tri,
                         0, reflection_points, excess_delays);
}

int gg_raytrace_multipath_bvh_batch(const double* rx_ecef, const double* sat_ecef,
                                    const double* nodes_flat, int32_t n_nodes,
                                    const double* sorted_tris, int32_t n_tri,
                                    double* reflection_points, double* excess_delays,
                                    int32_t n_epoch, int32_t n_sat) {
  return multipath_query(rx_ecef, sat_ecef, n_epoch, n_sat, nodes_flat, n_nodes, sorted_tris,
                         n_tri, 1, reflection_points, excess_delays);
}

int gg_pf_weight_3d_bvh_device(const gg_bvh* bvh, const double* px, const double* py,
                               const double* pz, const double* pcb, const double* sat_ecef,
                               const double* pseudoranges, const double* weights_sat,
                               double* log_weights, int32_t n_particles, int32_t n_sat,
                               double sigma_pr_los, double sigma_pr_nlos, double nlos_bias,
                               double blocked_nlos_prob, double clear_nlos_prob,
                               void* hip_stream) {
  if (!bvh) return gg_fail(-2, "null BVH handle");
  if (n_particles < 0 || n_sat < 0) return gg_fail(-2, "negative size");
  if (n_particles == 0) return 0;
  MixArgs mx;
  make_mix(sigma_pr_los, sigma_pr_nlos, nlos_bias, blocked_nlos_prob, clear_nlos_prob, mx);
  hipLaunchKernelGGL(gg_k_weight_3d, dim3(grid_for(n_particles)), dim3(GG_BVH_BLOCK), 0,
                     (hipStream_t)hip_stream, px, py, pz, pcb, sat_ecef, pseudoranges,
                     weights_sat, bvh->nodes, bvh->tris, bvh->n_nodes, n_particles, n_sat, mx,
                     log_weights);
  GG_HIP(hipGetLastError());
  return 0;
}

int gg_pf_weight_3d_bvh(const double* px, const double* py, const double* pz, const double* pcb,
                        const double* sat_ecef, const double* pseudoranges,
                        const double* weights_sat, const double* nodes_flat, int32_t n_nodes,
                        const double* sorted_tris, int32_t n_tri, double* log_weights,
                        int32_t n_particles, int32_t n_sat, double sigma_pr_los,
                        double sigma_pr_nlos, double nlos_bias, double blocked_nlos_prob,
                        double clear_nlos_prob) {
  if (n_particles < 0 || n_sat < 0) return gg_fail(-2, "negative size");
  if (n_particles == 0) return 0;
  gg_bvh b;
  int rc = upload(nodes_flat, n_nodes, sorted_tris, n_tri, &b);
  if (!rc) {
    const size_t n = (size_t)n_particles, ns = (size_t)std::max(n_sat, 1);
    Scratch sc;
    rc = sc.alloc(sizeof(double) * (n * 5 + ns * 5));
    if (!rc) {
      double* d_p[4];
      for (int k = 0; k < 4; ++k) d_p[k] = sc.take<double>(n);
      double* d_lw = sc.take<double>(n);
      double* d_sat = sc.take<double>(ns * 3);
      double* d_pr = sc.take<double>(ns);
      double* d_ws = sc.take<double>(ns);
      const double* hp[4] = {px, py, pz, pcb};
      for (int k = 0; k < 4 && !rc; ++k)
        if (hipMemcpy(d_p[k], hp[k], sizeof(double) * n, hipMemcpyHostToDevice))
          rc = gg_fail(-3, "host to device copy failed");
      if (!rc && n_sat > 0 &&
          (hipMemcpy(d_sat, sat_ecef, sizeof(double) * n_sat * 3, hipMemcpyHostToDevice) ||
           hipMemcpy(d_pr, pseudoranges, sizeof(double) * n_sat, hipMemcpyHostToDevice) ||
           hipMemcpy(d_ws, weights_sat, sizeof(double) * n_sat, hipMemcpyHostToDevice)))
        rc = gg_fail(-3, "host to device copy failed");
      if (!rc)
        rc = gg_pf_weight_3d_bvh_device(&b, d_p[0], d_p[1], d_p[2], d_p[3], d_sat, d_pr, d_ws,
                                        d_lw, n_particles, n_sat, sigma_pr_los, sigma_pr_nlos,
                                        nlos_bias, blocked_nlos_prob, clear_nlos_prob, nullptr);
      if (!rc) {
        const hipError_t ce =
            hipMemcpy(log_weights, d_lw, sizeof(double) * n, hipMemcpyDeviceToHost);
        if (ce != hipSuccess) rc = gg_fail_hip(ce, "D2H log_weights", __FILE__, __LINE__);
      }
    }
  }
  release(&b);
  return rc;
}

// ------------------------------------------- unaccelerated counterparts --
// The reference's linear-scan entry points test every triangle; here they
// build a tree over the triangles first and run the same kernels (the box
// tests only skip triangles that the ray cannot hit).
static int build_sorted(const double* tris, int n_tri, std::vector<double>& flat, int* n_nodes,
                        std::vector<double>& sorted) {
  flat.assign((size_t)std::max(2 * n_tri, 1) * 10, 0.0);
  std::vector<int32_t> order((size_t)std::max(n_tri, 1));
  *n_nodes = 0;
  if (int rc = gg_bvh_build(tris, n_tri, flat.data(), n_nodes, order.data())) return rc;
  sorted.resize((size_t)n_tri * 9);
  for (int i = 0; i < n_tri; ++i)
    std::memcpy(&sorted[(size_t)i * 9], tris + (size_t)order[i] * 9, 9 * sizeof(double));
  return 0;
}

int gg_raytrace_los_check(const double* rx_ecef, const double* sat_ecef, const double* triangles,
                          int32_t n_tri, int32_t* is_los, int32_t n_sat) {
  if (n_tri < 0) return gg_fail(-2, "negative size");
  std::vector<double> flat, sorted;
  int n_nodes = 0;
  if (int rc = build_sorted(triangles, n_tri, flat, &n_nodes, sorted)) return rc;
  return los_query(rx_ecef, sat_ecef, 1, n_sat, flat.data(), n_nodes, sorted.data(), n_tri, 0,
                   is_los);
}

int gg_raytrace_multipath(const double* rx_ecef, const double* sat_ecef, const double* triangles,
                          int32_t n_tri, double* reflection_points, double* excess_delays,
                          int32_t n_sat) {
  if (n_tri < 0) return gg_fail(-2, "negative size");
  std::vector<double> flat, sorted;
  int n_nodes = 0;
  if (int rc = build_sorted(triangles, n_tri, flat, &n_nodes, sorted)) return rc;
  return multipath_query(rx_ecef, sat_ecef, 1, n_sat, flat.data(), n_nodes, sorted.data(), n_tri,
                         0, reflection_points, excess_delays);
}

int gg_pf_weight_3d(const double* px, const double* py, const double* pz, const double* pcb,
                    const double* sat_ecef, const double* pseudoranges, const double* weights_sat,
                    const double* triangles, int32_t n_tri, double* log_weights,
                    int32_t n_particles, int32_t n_sat, double sigma_pr_los,
                    double sigma_pr_nlos, double nlos_bias, double blocked_nlos_prob,
                    double clear_nlos_prob) {
  if (n_tri < 0) return gg_fail(-2, "negative size");
  std::vector<double> flat, sorted;
  int n_nodes = 0;
  if (int rc = build_sorted(triangles, n_tri, flat, &n_nodes, sorted)) return rc;
  return gg_pf_weight_3d_bvh(px, py, pz, pcb, sat_ecef, pseudoranges, weights_sat, flat.data(),
                             n_nodes, sorted.data(), n_tri, log_weights, n_particles, n_sat,
                             sigma_pr_los, sigma_pr_nlos, nlos_bias, blocked_nlos_prob,
                             clear_nlos_prob);
}

}  // extern "C"
