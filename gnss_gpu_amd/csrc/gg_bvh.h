// BVH layout and ray tests shared by the ray-casting kernels (gg_bvh.hip) and
// the particle filter's ray-traced weight (gg_pf_kernels_impl.h).
//
// Device layout: each internal node carries the boxes of both children
// (112 B: 12 doubles + 4 ints, seven 16-byte loads), so one fetch decides
// both descents; a leaf child is stored in its parent as (tri_start, count).
// Triangles are stored as v0, e1 = v1 - v0, e2 = v2 - v0 (80 B, five
// 16-byte loads): the same binary64 values the Moller-Trumbore test computes
// from the vertices.  Trees reach the device only after the host has checked
// them (gg_bvh.hip: a proper tree from node 0, leaf ranges inside the
// triangle array, depth <= 62), which bounds the walk's stack.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#define GG_BVH_STACK 64
#define GG_BVH_MAX_DEPTH 62
#define GG_BVH_BLOCK 256

struct alignas(16) gg_bvh_node {
  double b[12];  // left child min xyz, max xyz, right child min xyz, max xyz
  int32_t c[4];  // left code, right code, left count, right count
};
// child code/count: count > 0 -> leaf of triangles [code, code + count);
// count == 0 and code >= 0 -> internal node index; code < 0 -> no child.

struct gg_bvh {
  gg_bvh_node* nodes = nullptr;
  double* tris = nullptr;  // [n_tri][10]: v0, e1, e2, pad
  int32_t n_nodes = 0;     // device nodes
  int32_t n_tri = 0;
};

namespace gg_rt {

struct Ray {
  double o[3], d[3], inv[3], tmax;
};

__device__ __forceinline__ void make_inv(Ray& r) {
#pragma unroll
  for (int k = 0; k < 3; ++k)
    r.inv[k] = fabs(r.d[k]) > 1e-20 ? 1.0 / r.d[k] : (r.d[k] >= 0.0 ? 1e20 : -1e20);
}

// slab test of [0, tmax] against a box (bvh.cu:276-289)
__device__ __forceinline__ bool slab(const Ray& r, const double* b) {
  double t0 = 0.0, t1 = r.tmax;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double a = (b[k] - r.o[k]) * r.inv[k];
    double c = (b[3 + k] - r.o[k]) * r.inv[k];
    if (a > c) {
      const double t = a;
      a = c;
      c = t;
    }
    if (a > t0) t0 = a;
    if (c < t1) t1 = c;
    if (t0 > t1) return false;
  }
  return true;
}

// Moller-Trumbore (bvh.cu:250-272) on (v0, e1, e2); hit parameter in t
__device__ __forceinline__ bool mt(const double o[3], const double d[3], const double* tr,
                                   double& t) {
  const double* v0 = tr;
  const double* e1 = tr + 3;
  const double* e2 = tr + 6;
  const double h0 = d[1] * e2[2] - d[2] * e2[1];
  const double h1 = d[2] * e2[0] - d[0] * e2[2];
  const double h2 = d[0] * e2[1] - d[1] * e2[0];
  const double a = e1[0] * h0 + e1[1] * h1 + e1[2] * h2;
  if (fabs(a) < 1e-12) return false;
  const double f = 1.0 / a;
  const double s0 = o[0] - v0[0], s1 = o[1] - v0[1], s2 = o[2] - v0[2];
  const double u = f * (s0 * h0 + s1 * h1 + s2 * h2);
  if (u < 0.0 || u > 1.0) return false;
  const double q0 = s1 * e1[2] - s2 * e1[1];
  const double q1 = s2 * e1[0] - s0 * e1[2];
  const double q2 = s0 * e1[1] - s1 * e1[0];
  const double v = f * (d[0] * q0 + d[1] * q1 + d[2] * q2);
  if (v < 0.0 || u + v > 1.0) return false;
  t = f * (e2[0] * q0 + e2[1] * q1 + e2[2] * q2);
  return true;
}

__device__ __forceinline__ void load_tri(const double* __restrict__ tris, int i, double tr[10]) {
  const double2* p = reinterpret_cast<const double2*>(tris + (size_t)i * 10);
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const double2 v = p[k];
    tr[2 * k] = v.x;
    tr[2 * k + 1] = v.y;
  }
}

__device__ __forceinline__ void load_node(const gg_bvh_node* __restrict__ nodes, int i,
                                          double b[12], int c[4]) {
  const double2* p = reinterpret_cast<const double2*>(nodes + i);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double2 v = p[k];
    b[2 * k] = v.x;
    b[2 * k + 1] = v.y;
  }
  const int4 q = reinterpret_cast<const int4*>(nodes + i)[6];
  c[0] = q.x;
  c[1] = q.y;
  c[2] = q.z;
  c[3] = q.w;
}

// Any hit with tmin < t < r.tmax among the triangles of the tree.  Children
// hit by the ray are visited depth first; leaves are tested as soon as their
// box is hit.  The stack holds at most one sibling per level (depth <= 62).
__device__ inline bool gg_bvh_blocked(const Ray& r, double tmin, const gg_bvh_node* __restrict__ nodes,
                        const double* __restrict__ tris) {
  int stack[GG_BVH_STACK];
  int sp = 0;
  int cur = 0;
  for (;;) {
    double b[12];
    int c[4];
    load_node(nodes, cur, b, c);
    int next = -1;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int code = c[side], cnt = c[2 + side];
      if (code < 0 || !slab(r, b + 6 * side)) continue;
      if (cnt > 0) {
        for (int i = 0; i < cnt; ++i) {
          double tr[10], t;
          load_tri(tris, code + i, tr);
          if (mt(r.o, r.d, tr, t) && t > tmin && t < r.tmax) return true;
        }
      } else if (next < 0) {
        next = code;
      } else if (sp < GG_BVH_STACK) {  // sp <= tree depth <= 61: never full
        stack[sp++] = code;
      }
    }
    if (next < 0) {
      if (sp == 0) return false;
      next = stack[--sp];
    }
    cur = next;
  }
}

// LOS / NLOS mixture of the ray-traced weight (weight_3d_bvh.cu:59-72, 199-216)
struct MixArgs {
  double inv_los2, inv_nlos2, nlos_bias;
  double lp_keep[2], lp_nlos[2];  // [clear, blocked]: log(1 - p), log(p)
  int mode[2];                    // 0: LOS term, 1: NLOS term, 2: mixture
};

__device__ __forceinline__ double mixture(double los, double nlos, int mode, double l0,
                                          double l1) {
  if (mode == 0) return los;
  if (mode == 1) return nlos;
  const double a = l0 + los, b = l1 + nlos;
  const double m = a > b ? a : b;
  return m + log(exp(a - m) + exp(b - m));
}

inline int make_mix(double sigma_los, double sigma_nlos, double nlos_bias, double p_blocked,
             double p_clear, MixArgs& mx) {
  mx.inv_los2 = 1.0 / (sigma_los * sigma_los);
  mx.inv_nlos2 = 1.0 / (sigma_nlos * sigma_nlos);
  mx.nlos_bias = nlos_bias;
  const double p[2] = {p_clear, p_blocked};
  for (int k = 0; k < 2; ++k) {
    mx.mode[k] = p[k] <= 0.0 ? 0 : (p[k] >= 1.0 ? 1 : 2);
    mx.lp_keep[k] = mx.mode[k] == 2 ? std::log(1.0 - p[k]) : 0.0;
    mx.lp_nlos[k] = mx.mode[k] == 2 ? std::log(p[k]) : 0.0;
  }
  return 0;
}

}  // namespace gg_rt
