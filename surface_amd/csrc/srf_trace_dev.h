// srf_trace_dev.h -- device helpers shared by the traversal kernels
// (srf_trace.hip) and the two-level scene kernels (srf_scene.hip).
#pragma once

#include <float.h>
#include <math.h>

#include "srf_internal.h"

namespace srf {
namespace {

constexpr int kTraceBlock = 128;
constexpr int kTraceStack = kMaxBranchLevels;  // see above
constexpr uint32_t kDone = 0xffffffffu;

struct RaySetup {
  float o[3], d[3], inv[3];
  int sign[3];
  uint32_t sbits;  // sign[k] << k, so the near child needs no indexed load
};

__device__ __forceinline__ RaySetup setup_ray(const srf_ray *rays, uint32_t i, uint32_t &flags,
                                              uint32_t &prev) {
  const float4 *r4 = reinterpret_cast<const float4 *>(rays + i);
  float4 a = r4[0], b = r4[1];
  RaySetup s;
  s.o[0] = a.x;
  s.o[1] = a.y;
  s.o[2] = a.z;
  s.d[0] = b.x;
  s.d[1] = b.y;
  s.d[2] = b.z;
  flags = __float_as_uint(a.w);
  prev = __float_as_uint(b.w);
  for (int k = 0; k < 3; k++) {
    s.sign[k] = s.d[k] < 0.0f ? 1 : 0;
    // (float)(1.0 / dir): correctly rounded float division gives the same bits
    s.inv[k] = 1.0f / s.d[k];
  }
  s.sbits = (uint32_t)s.sign[0] | ((uint32_t)s.sign[1] << 1) | ((uint32_t)s.sign[2] << 2);
  return s;
}

// IntersectRayAABB, render/render_accel_triangle.cc:1815-1861 (two-box
// nodes, tmax clamped to maxT) and render/render_accel_tetra.cc:1934-1978
// (own box, tmin <= maxT).
template <bool kClampTmax>
__device__ __forceinline__ bool ray_aabb(const RaySetup &s, float maxT, float bx0, float by0,
                                         float bz0, float bx1, float by1, float bz1) {
  float mnx = s.sign[0] ? bx1 : bx0, mxx = s.sign[0] ? bx0 : bx1;
  float mny = s.sign[1] ? by1 : by0, mxy = s.sign[1] ? by0 : by1;
  float mnz = s.sign[2] ? bz1 : bz0, mxz = s.sign[2] ? bz0 : bz1;
  float tminx = (mnx - s.o[0]) * s.inv[0];
  float tmaxx = (mxx - s.o[0]) * s.inv[0];
  float tminy = (mny - s.o[1]) * s.inv[1];
  float tmaxy = (mxy - s.o[1]) * s.inv[1];
  float tmin = (tminx > tminy) ? tminx : tminy;
  float tmax = (tmaxx < tmaxy) ? tmaxx : tmaxy;
  float tminz = (mnz - s.o[2]) * s.inv[2];
  float tmaxz = (mxz - s.o[2]) * s.inv[2];
  tmin = (tmin > tminz) ? tmin : tminz;
  tmax = (tmax < tmaxz) ? tmax : tmaxz;
  if (kClampTmax) {
    tmax = (maxT < tmax) ? maxT : tmax;
    return (tmax > 0.0f) && (tmin <= tmax);
  }
  return (tmax > 0.0f) && (tmin <= tmax) && (tmin <= maxT);
}

// The same test for a wavefront whose rays all have the direction signs
// `sb` (sign[k] << k, wavefront-uniform) and a node loaded through the
// scalar unit: the near and far planes are chosen once for the wavefront
// (scalar selects) instead of per lane; the arithmetic is ray_aabb's.
template <bool kClampTmax>
__device__ __forceinline__ bool ray_aabb_wave(const RaySetup &s, uint32_t sb, float maxT,
                                              float bx0, float by0, float bz0, float bx1,
                                              float by1, float bz1) {
  const float mnx = (sb & 1u) ? bx1 : bx0, mxx = (sb & 1u) ? bx0 : bx1;
  const float mny = (sb & 2u) ? by1 : by0, mxy = (sb & 2u) ? by0 : by1;
  const float mnz = (sb & 4u) ? bz1 : bz0, mxz = (sb & 4u) ? bz0 : bz1;
  float tminx = (mnx - s.o[0]) * s.inv[0];
  float tmaxx = (mxx - s.o[0]) * s.inv[0];
  float tminy = (mny - s.o[1]) * s.inv[1];
  float tmaxy = (mxy - s.o[1]) * s.inv[1];
  float tmin = (tminx > tminy) ? tminx : tminy;
  float tmax = (tmaxx < tmaxy) ? tmaxx : tmaxy;
  float tminz = (mnz - s.o[2]) * s.inv[2];
  float tmaxz = (mxz - s.o[2]) * s.inv[2];
  tmin = (tmin > tminz) ? tmin : tminz;
  tmax = (tmax < tmaxz) ? tmax : tmaxz;
  if (kClampTmax) {
    tmax = (maxT < tmax) ? maxT : tmax;
    return (tmax > 0.0f) && (tmin <= tmax);
  }
  return (tmax > 0.0f) && (tmin <= tmax) && (tmin <= maxT);
}

// ---------------------------------------------------------------- triangles

// TriangleIsect (Moller-Trumbore), render/render_accel_triangle.cc:1887-1938,
// on a packed triangle (p0, e1 = p1 - p0, e2 = p2 - p0). u, v and t are the
// reference's expressions; the estimates ue/ve/te (numerator x hardware
// reciprocal, within ~4e-7 relative of the exact values) only reject
// triangles the exact comparisons would reject as well. kEPS >= 1e-30 keeps
// det normal, so the reciprocal estimate is accurate.
__device__ __forceinline__ bool tri_isect(const RaySetup &s, float4 A, float4 B, float4 C,
                                          bool double_sided, float kEPS, bool filter,
                                          float tin, float &tout, float &uout, float &vout) {
  const float e1x = B.x, e1y = B.y, e1z = B.z, e2x = C.x, e2y = C.y, e2z = C.z;
  const float dx = s.d[0], dy = s.d[1], dz = s.d[2];
  float px = dy * e2z - dz * e2y;
  float py = dz * e2x - dx * e2z;
  float pz = dx * e2y - dy * e2x;
  float det = (e1x * px + e1y * py) + e1z * pz;
  if (double_sided) {
    if (fabsf(det) < kEPS) return false;
  } else {
    if (det < kEPS) return false;
  }
  float sx = s.o[0] - A.x, sy = s.o[1] - A.y, sz = s.o[2] - A.z;
  float un = (sx * px + sy * py) + sz * pz;
  const float rc = __builtin_amdgcn_rcpf(det);
  const float ue = un * rc;
  if (filter && (ue < -1e-5f || ue > 1.00001f)) return false;
  float qx = sy * e1z - sz * e1y;
  float qy = sz * e1x - sx * e1z;
  float qz = sx * e1y - sy * e1x;
  float vn = (qx * dx + qy * dy) + qz * dz;
  const float ve = vn * rc;
  if (filter && (ve < -1e-5f || ue + ve > 1.00001f)) return false;
  float tn = (e2x * qx + e2y * qy) + e2z * qz;
  const float te = tn * rc;
  if (filter && (te < -1e-30f || te > tin * 1.00001f)) return false;
  float inv_det = 1.0f / det;  // correctly rounded: equals (float)(1.0 / det)
  float u = un * inv_det;
  float v = vn * inv_det;
  float t = tn * inv_det;
  if (u < 0.0f || u > 1.0f) return false;
  if (v < 0.0f || u + v > 1.0f) return false;
  if (t < 0.0f || t > tin) return false;
  tout = t;
  uout = u;
  vout = v;
  return true;
}

// Empty asm that takes the three values as operands: their loads are issued
// here, together, instead of being sunk past the branch that follows (one
// memory round trip instead of two).
__device__ __forceinline__ void keep_loaded(const float4 &a, const float4 &b, const float4 &c) {
  asm volatile("" ::"v"(a.x), "v"(a.y), "v"(a.z), "v"(a.w), "v"(b.x), "v"(b.y), "v"(b.z),
               "v"(b.w), "v"(c.x), "v"(c.y), "v"(c.z), "v"(c.w));
}

// The same for values loaded through the scalar unit (wavefront-uniform
// addresses): the loads are issued together and waited for once, and the
// values stay in scalar registers.
__device__ __forceinline__ void keep_loaded_s(const float4 &a, const float4 &b, const float4 &c) {
  asm volatile("" ::"s"(a.x), "s"(a.y), "s"(a.z), "s"(a.w), "s"(b.x), "s"(b.y), "s"(b.z),
               "s"(b.w), "s"(c.x), "s"(c.y), "s"(c.z), "s"(c.w));
}

__device__ __forceinline__ void keep_loaded_s2(const float4 &a, const float4 &b) {
  asm volatile("" ::"s"(a.x), "s"(a.y), "s"(a.z), "s"(a.w), "s"(b.x), "s"(b.y), "s"(b.z),
               "s"(b.w));
}

// A lane's pending far children in LDS, kStride entries apart (one column
// per lane, so a wavefront's accesses never share a bank row twice).
template <int kStride>
struct Stack32 {
  uint32_t *w;
  __device__ __forceinline__ void put(int k, uint32_t x) const { w[k * kStride] = x; }
  __device__ __forceinline__ uint32_t get(int k) const { return w[k * kStride]; }
};
// The same for node indices below 2^24: the low 16 and the high 8 bits in two
// planes, 3 bytes per entry instead of 4, so that 8 wavefronts per SIMD fit in
// the LDS (measured on the heightfield: 1.57 -> 1.48 ms).
template <int kStride>
struct Stack24 {
  uint16_t *lo;
  uint8_t *hi;
  __device__ __forceinline__ void put(int k, uint32_t x) const {
    lo[k * kStride] = (uint16_t)x;
    hi[k * kStride] = (uint8_t)(x >> 16);
  }
  __device__ __forceinline__ uint32_t get(int k) const {
    return (uint32_t)lo[k * kStride] | ((uint32_t)hi[k * kStride] << 16);
  }
};

// Closest hit (kAny: any hit closer than tlimit) of one ray against a
// triangle accelerator, TriangleAccel::Traverse
// (render/render_accel_triangle.cc:2511-2669). `st` is this lane's stack
// (Stack32 or Stack24). Returns whether an occluder was found (kAny);
// t/u/v/prim receive the closest hit (t = FLT_MAX, prim = SRF_NO_PRIM on a
// miss).
template <bool kAny, class Stk>
__device__ __forceinline__ bool tri_traverse(const float4 *__restrict__ nodes,
                                             const float4 *__restrict__ prims,
                                             uint32_t max_nodes, uint32_t nprims,
                                             const RaySetup &s, bool ds, float kEPS,
                                             uint32_t prev, float tlimit, const Stk &st,
                                             float &t, float &u, float &v, uint32_t &prim) {
  const bool filter = kEPS >= 1e-30f;
  t = kAny ? tlimit : FLT_MAX;
  u = 0.0f;
  v = 0.0f;
  prim = SRF_NO_PRIM;
  int sp = 0;
  uint32_t cur = 0;
  uint4 meta;
  for (;;) {
    // descend until this lane holds a leaf (or has nothing left)
    for (;;) {
      if (cur >= max_nodes) {
        cur = kDone;
        break;
      }
      const float4 *nd = nodes + 4 * (size_t)cur;
      // the whole 64-byte node in one round trip (a leaf ignores the boxes)
      const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
      meta = reinterpret_cast<const uint4 *>(nd)[3];
      keep_loaded(n0, n1, n2);
      if (meta.x != 0) break;
      const bool h0 = ray_aabb<true>(s, t, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x);
      const bool h1 = ray_aabb<true>(s, t, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w);
      const int nearc = (s.sbits >> meta.y) & 1u;
      const bool hn = nearc ? h1 : h0, hf = nearc ? h0 : h1;
      const uint32_t cn = nearc ? meta.w : meta.z, cf = nearc ? meta.z : meta.w;
      if (hn) {
        if (hf && sp < kTraceStack) st.put(sp++, cf);
        cur = cn;
      } else if (hf) {
        cur = cf;
      } else if (sp == 0) {
        cur = kDone;
        break;
      } else {
        cur = st.get(--sp);
      }
    }
    if (cur == kDone) break;
    // leaf: TestLeafNode, render/render_accel_triangle.cc:2386-2433; the
    // next triangle's 48 bytes are in flight while this one is tested
    const uint32_t first = meta.w, end = min(meta.w + meta.z, nprims);
    if (first < end) {
      const float4 *P = prims + 3 * (size_t)first;
      float4 A = P[0], B = P[1], C = P[2];
      for (uint32_t q = first; q < end; q++) {
        float4 nA = A, nB = B, nC = C;
        if (q + 1 < end) {
          P += 3;
          nA = P[0];
          nB = P[1];
          nC = P[2];
        }
        keep_loaded(nA, nB, nC);
        if (__float_as_uint(A.w) != prev) {
          float tt, uu, vv;
          if (tri_isect(s, A, B, C, ds, kEPS, filter, t, tt, uu, vv)) {
            if (kAny) {
              if (tt < tlimit) return true;
            } else {
              t = tt;
              u = uu;
              v = vv;
              prim = __float_as_uint(A.w);
            }
          }
        }
        A = nA;
        B = nB;
        C = nC;
      }
    }
    if (sp == 0) break;
    cur = st.get(--sp);
  }
  return false;
}

}  // namespace
}  // namespace srf
