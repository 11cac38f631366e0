// srf_trace.hip -- closest-hit and any-hit traversal on gfx950.
//
// One lane per ray; 8x8 pixel tiles per wavefront when the caller lays rays
// out that way. The next node lives in a register; the far children wait in
// a per-lane LDS stack laid out column-major (lane stride), so stack traffic
// never bank-conflicts. A far child is pushed by its parent branch and each
// root-to-node path holds at most 24 branches (root + depths 0..22), so 24
// slots always suffice. Each wavefront descends through branches until
// every active lane holds a leaf (or is done) and only then runs the leaf
// loops, so leaf work is not interleaved with branch work lane by lane.
//
// Work the reference performs and then rejects is skipped when its outcome
// is certain, never otherwise:
//  * a triangle whose u, v or t, estimated with the hardware reciprocal, lies
//    outside the accepting range by a margin (1e-5 relative, ~30x the
//    estimate's error) is rejected before the exact IEEE division;
//  * a tetra child box the ray misses (independent of the current hit
//    distance) is never pushed; the distance test is made at pop time with
//    the stored entry distance, exactly as the reference tests the popped
//    node's own box;
//  * a tetra whose padded float bounding box the ray misses, or which lies
//    wholly behind the current hit by a margin, is not handed to the FP64
//    Pluecker test (an intersection it reports lies inside the tetra to
//    within double rounding, far inside the padding).
//
// Visit order, box tests and primitive tests reproduce the reference exactly
// (hit ids bit-exact, t/u/v bit-exact):
//   triangles  TriangleAccel::Traverse  render/render_accel_triangle.cc:2511-2669
//   tetras     TetraAccel::Traverse     render/render_accel_tetra.cc:2200-2289
//   spheres    ParticleAccel::Traverse  render/render_accel_particle.cc:3053-3218
#include <algorithm>

#include "srf_trace_dev.h"

namespace srf {
namespace {

// kSmall: node indices below 2^24, the 3-byte stack
template <bool kAny, bool kSmall>
__global__ void __launch_bounds__(kTraceBlock)
k_trace_tri(const float4 *__restrict__ nodes, const float4 *__restrict__ prims,
            const BuildParams *__restrict__ params, const srf_ray *__restrict__ rays, uint32_t n,
            srf_hit *__restrict__ hits, const float *__restrict__ tmax_in,
            uint32_t *__restrict__ occluded, uint32_t max_nodes, uint32_t nprims) {
  __shared__ uint32_t stack[kSmall ? 1 : kTraceStack * kTraceBlock];
  __shared__ uint16_t stack_lo[kSmall ? kTraceStack * kTraceBlock : 1];
  __shared__ uint8_t stack_hi[kSmall ? kTraceStack * kTraceBlock : 1];
  const uint32_t i = xcd_block() * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t flags, prev;
  const RaySetup s = setup_ray(rays, i, flags, prev);
  const bool ds = (flags & SRF_RAY_DOUBLE_SIDED) != 0;
  const float kEPS = FLT_EPSILON * (float)params->eps_scale;
  const float tlimit = kAny ? (tmax_in ? tmax_in[i] : FLT_MAX) : FLT_MAX;
  float t, u, v;
  uint32_t prim;
  bool occ;
  if (kSmall)
    occ = tri_traverse<kAny>(nodes, prims, max_nodes, nprims, s, ds, kEPS, prev, tlimit,
                             Stack24<kTraceBlock>{stack_lo + threadIdx.x, stack_hi + threadIdx.x},
                             t, u, v, prim);
  else
    occ = tri_traverse<kAny>(nodes, prims, max_nodes, nprims, s, ds, kEPS, prev, tlimit,
                             Stack32<kTraceBlock>{stack + threadIdx.x}, t, u, v, prim);
  if (kAny) {
    occluded[i] = occ ? 1u : 0u;
  } else {
    reinterpret_cast<float4 *>(hits)[i] = make_float4(t, u, v, __uint_as_float(prim));
  }
}

// ------------------------------------------------------------------ spheres

// SphereIsect, render/render_accel_particle.cc:2423-2512
__device__ __forceinline__ bool sphere_isect(const RaySetup &s, float4 c, float tin,
                                             double prevT, const float *prevN, bool self,
                                             float &tout) {
  const float ocx = s.o[0] - c.x, ocy = s.o[1] - c.y, ocz = s.o[2] - c.z;
  const float dx = s.d[0], dy = s.d[1], dz = s.d[2];
  const float radius = c.w;
  float a = (dx * dx + dy * dy) + dz * dz;
  float b = (float)(2.0 * (double)((dx * ocx + dy * ocy) + dz * ocz));
  float cc = ((ocx * ocx + ocy * ocy) + ocz * ocz) - radius * radius;
  float disc = (float)((double)(b * b) - (4.0 * (double)a) * (double)cc);
  float t0, t1;
  if (disc < 0.0f) return false;
  if (disc == 0.0f) {
    t0 = t1 = (float)(-0.5 * (double)(b / a));
  } else {
    float dsq = sqrtf(disc);
    float q;
    if (b < 0.0f)
      q = (float)((double)(-b - dsq) / 2.0);
    else
      q = (float)((double)(-b + dsq) / 2.0);
    t0 = q / a;
    t1 = cc / q;
  }
  if (t0 > t1) {
    float tmp = t0;
    t0 = t1;
    t1 = tmp;
  }
  if (t1 < 0.0f) return false;
  if (self) {
    float t = t1;
    float nx = (s.o[0] + t * dx) - c.x, ny = (s.o[1] + t * dy) - c.y,
          nz = (s.o[2] + t * dz) - c.z;
    float nd = (nx * prevN[0] + ny * prevN[1]) + nz * prevN[2];
    if ((t < tin) && ((double)t < prevT) && (nd <= 0.0f)) {
      tout = t;
      return true;
    }
    return false;
  }
  float t = (t0 < 0.0f) ? t1 : t0;
  if (t < tin) {
    tout = t;
    return true;
  }
  return false;
}

// Entry distance bound of a secondary-tree box for the sphere test: false
// when no sphere inside can report a hit (for any t); else `lo` is at most
// the distance any sphere inside can report. The box holds the spheres'
// extents. SphereIsect's float arithmetic moves its roots away from the exact
// ones: with u = 2^-24, its discriminant carries an error of at most about
// 90 u |d|^2 (|o - c|^2 + r^2), so a reported root lies no nearer than the
// entry into the ball of radius r + 1.2e-3 sqrt(|o - c|^2 + r^2) around the
// centre (plus terms of order u |o - c|). The box is grown by 2e-3 of the
// farthest distance from the origin to it, which bounds that for every
// sphere inside. A ray whose origin is inside the grown box gets lo < 0.
__device__ __forceinline__ bool sub_box(const RaySetup &s, float x0, float y0, float z0,
                                        float x1, float y1, float z1, float &lo) {
  const float fx = fmaxf(fabsf(s.o[0] - x0), fabsf(s.o[0] - x1));
  const float fy = fmaxf(fabsf(s.o[1] - y0), fabsf(s.o[1] - y1));
  const float fz = fmaxf(fabsf(s.o[2] - z0), fabsf(s.o[2] - z1));
  const float pad = 2e-3f * sqrtf((fx * fx + fy * fy) + fz * fz);
  const float b0[3] = {x0 - pad, y0 - pad, z0 - pad}, b1[3] = {x1 + pad, y1 + pad, z1 + pad};
  float tn = -FLT_MAX, tf = FLT_MAX;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (!(fabsf(s.inv[k]) <= FLT_MAX)) {
      // direction component zero (or too small for a finite reciprocal):
      // the ray stays in its plane
      if (s.o[k] < b0[k] || s.o[k] > b1[k]) return false;
    } else {
      const float a = (b0[k] - s.o[k]) * s.inv[k], b = (b1[k] - s.o[k]) * s.inv[k];
      tn = fmaxf(tn, fminf(a, b));
      tf = fminf(tf, fmaxf(a, b));
    }
  }
  lo = tn;
  return tn <= tf && tf >= 0.0f;
}

// One big leaf through its secondary tree (the leaf's spheres, built as a
// tree of their own). The reference scans the leaf in order and a sphere
// replaces the kept one only when strictly closer, so the leaf yields the
// smallest (t, position) among the spheres whose test reports t below the
// distance on entry; the secondary tree finds that pair in any order,
// skipping boxes whose bound lies beyond the kept distance (kAny: at or
// beyond tlimit). `st` holds kTraceStack pending nodes (stride kStride).
template <bool kAny, int kStride>
__device__ __forceinline__ bool sph_big_leaf(const SphSub &sub, uint32_t nbase, uint32_t pbase,
                                             uint32_t first, uint32_t count,
                                             const uint32_t *__restrict__ sorted_index,
                                             const RaySetup &s, uint32_t prev, double prevT,
                                             const float prevN[3], float tlimit, uint32_t *st,
                                             float &t, uint32_t &best) {
  bool kept = false;  // `best` is one of this leaf's spheres
  int sp = 0;
  uint32_t cur = 0;
  const uint32_t nnodes = 2 * count - 1;
  for (;;) {
    uint4 meta;
    for (;;) {
      if (cur >= nnodes || nbase + cur >= sub.nnodes) {
        cur = kDone;
        break;
      }
      const float4 *nd = sub.nodes + 4 * (size_t)(nbase + cur);
      const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
      meta = reinterpret_cast<const uint4 *>(nd)[3];
      keep_loaded(n0, n1, n2);
      if (meta.x != 0) break;
      float l0, l1;
      bool h0 = sub_box(s, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x, l0);
      bool h1 = sub_box(s, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w, l1);
      h0 = h0 && (kAny ? l0 < tlimit : l0 <= t);
      h1 = h1 && (kAny ? l1 < tlimit : l1 <= t);
      const bool one_first = h1 && (!h0 || l1 < l0);
      const uint32_t cn = one_first ? meta.w : meta.z, cf = one_first ? meta.z : meta.w;
      if (h0 && h1) {
        if (sp < kTraceStack) st[(sp++) * kStride] = cf;
        cur = cn;
      } else if (h0 || h1) {
        cur = cn;
      } else if (sp == 0) {
        cur = kDone;
        break;
      } else {
        cur = st[(--sp) * kStride];
      }
    }
    if (cur == kDone) break;
    const uint32_t f0 = meta.w, fend = min(meta.w + meta.z, count);
    for (uint32_t k = f0; k < fend; k++) {
      const uint32_t j = pbase + k;
      if (j >= sub.nprims) break;
      const float4 c = sub.prims[j];
      const uint32_t loc = sub.sidx[j];
      if (loc >= count) continue;
      const uint32_t q = first + loc;
      const bool self = (prev != SRF_NO_PRIM) && (sorted_index[q] == prev);
      float tt;
      if (!sphere_isect(s, c, INFINITY, prevT, prevN, self, tt)) continue;
      if (kAny) {
        if (tt < tlimit) return true;
      } else if (tt < t || (tt == t && kept && q < best)) {
        t = tt;
        best = q;
        kept = true;
      }
    }
    if (sp == 0) break;
    cur = st[(--sp) * kStride];
  }
  return false;
}

// Closest (kAny: any closer than tlimit) sphere of one ray,
// ParticleAccel::Traverse (render/render_accel_particle.cc:3053-3218).
// `best` receives the sorted position of the hit sphere (SRF_NO_PRIM on a
// miss), t its distance. `st` holds 2 kTraceStack entries (stride kStride):
// the tree's pending nodes, then a big leaf's.
template <bool kAny, int kStride>
__device__ __forceinline__ bool sph_traverse(const float4 *__restrict__ nodes,
                                             const float4 *__restrict__ prims,
                                             const uint32_t *__restrict__ sorted_index,
                                             uint32_t max_nodes, uint32_t nprims,
                                             const SphSub &sub, const RaySetup &s, uint32_t prev,
                                             double prevT, const float prevN[3], float tlimit,
                                             uint32_t *st, float &t, uint32_t &best) {
  t = kAny ? tlimit : FLT_MAX;
  best = SRF_NO_PRIM;
  int sp = 0;
  uint32_t cur = 0;
  uint4 meta;
  for (;;) {
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
        if (hf && sp < kTraceStack) st[(sp++) * kStride] = cf;
        cur = cn;
      } else if (hf) {
        cur = cf;
      } else if (sp == 0) {
        cur = kDone;
        break;
      } else {
        cur = st[(--sp) * kStride];
      }
    }
    if (cur == kDone) break;
    if (meta.z > kBigLeaf && sub.nodes) {
      // words 3 and 4 of a big leaf: its secondary tree's node and sphere offsets
      const float4 *nd = nodes + 4 * (size_t)cur;
      const uint32_t nbase = __float_as_uint(nd[0].w), pbase = __float_as_uint(nd[1].x);
      if (meta.w < nprims && meta.z <= nprims - meta.w &&
          sph_big_leaf<kAny, kStride>(sub, nbase, pbase, meta.w, meta.z, sorted_index, s, prev,
                                      prevT, prevN, tlimit, st + kTraceStack * kStride, t,
                                      best))
        return true;
      if (sp == 0) break;
      cur = st[(--sp) * kStride];
      continue;
    }
    const uint32_t first = meta.w, end = min(meta.w + meta.z, nprims);
    for (uint32_t q = first; q < end; q++) {
      const float4 c = prims[q];
      const bool self = (prev != SRF_NO_PRIM) && (sorted_index[q] == prev);
      float tt;
      if (sphere_isect(s, c, t, prevT, prevN, self, tt)) {
        if (kAny) {
          if (tt < tlimit) return true;
        } else {
          t = tt;
          best = q;
        }
      }
    }
    if (sp == 0) break;
    cur = st[(--sp) * kStride];
  }
  return false;
}

template <bool kAny>
__global__ void __launch_bounds__(kTraceBlock)
k_trace_sph(const float4 *__restrict__ nodes, const float4 *__restrict__ prims,
            const uint32_t *__restrict__ sorted_index, const srf_ray *__restrict__ rays,
            const srf_sphere_prev *__restrict__ prevs, uint32_t n, srf_hit *__restrict__ hits,
            const float *__restrict__ tmax_in, uint32_t *__restrict__ occluded,
            uint32_t max_nodes, uint32_t nprims, SphSub sub) {
  __shared__ uint32_t stack[2 * kTraceStack * kTraceBlock];
  const uint32_t i = xcd_block() * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t flags, prev;
  const RaySetup s = setup_ray(rays, i, flags, prev);
  double prevT = 0.0;
  float prevN[3] = {0.0f, 0.0f, 0.0f};
  if (prevs) {
    prevT = prevs[i].prev_hit_t;
    for (int k = 0; k < 3; k++) prevN[k] = prevs[i].prev_hit_normal[k];
  }
  const float tlimit = kAny ? (tmax_in ? tmax_in[i] : FLT_MAX) : FLT_MAX;
  float t;
  uint32_t best;
  const bool occ = sph_traverse<kAny, kTraceBlock>(nodes, prims, sorted_index, max_nodes, nprims,
                                                   sub, s, prev, prevT, prevN, tlimit,
                                                   stack + threadIdx.x, t, best);
  if (kAny) {
    occluded[i] = occ ? 1u : 0u;
    return;
  }
  float u = 0.0f, v = 0.0f;
  uint32_t prim = SRF_NO_PRIM;
  if (best != SRF_NO_PRIM) {
    prim = sorted_index[best];
    // BuildIntersection (render/render_accel_particle.cc:2968-3049) with
    // vnormalize (:2041-2056); u, v through double atan2/acos.
    const float4 c = prims[best];
    float cx = (s.o[0] + t * s.d[0]) - c.x;
    float cy = (s.o[1] + t * s.d[1]) - c.y;
    float cz = (s.o[2] + t * s.d[2]) - c.z;
    const float kEPS = (float)(FLT_EPSILON * 16.0);
    float len = sqrtf((cx * cx + cy * cy) + cz * cz);
    float nx = cx, ny = cy, nz = cz;
    if (len > kEPS) {
      float il = 1.0f / len;
      nx = cx * il;
      ny = cy * il;
      nz = cz * il;
    }
    u = (float)((atan2((double)nx, (double)nz) + M_PI) * 0.5 * (1.0 / M_PI));
    v = (float)(acos((double)ny) / M_PI);
  }
  reinterpret_cast<float4 *>(hits)[i] = make_float4(t, u, v, __uint_as_float(prim));
}

// ------------------------------------------------------------------- tetras

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 d3sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 d3add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 d3scale(double f, D3 v) { return {v.x * f, v.y * f, v.z * f}; }
__device__ __forceinline__ D3 d3cross(D3 a, D3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double d3dot(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ int fsign(double x) {
  // render/render_accel_tetra.cc:89-93
  return x > DBL_EPSILON ? 1 : (x < -DBL_EPSILON ? -1 : 0);
}
struct Pl {
  D3 d, c;
};
__device__ __forceinline__ Pl mk_pl(D3 v0, D3 v1) { return {d3sub(v1, v0), d3cross(v1, v0)}; }
__device__ __forceinline__ double pl_dot(const Pl &a, const Pl &b) {
  return d3dot(a.d, b.c) + d3dot(b.d, a.c);
}
__device__ __forceinline__ D3 bary3(double a, double b, D3 A, D3 B, D3 C) {
  return d3add(d3add(d3scale((1 - a) - b, A), d3scale(a, B)), d3scale(b, C));
}
__device__ __forceinline__ void param_dist(D3 o, D3 d, D3 ep, D3 lp, double &te, double &tl) {
  // ComputeParametricDist, render/render_accel_tetra.cc:123-139
  if (d.x != 0.0) {
    double inv = 1.0 / d.x;
    te = (ep.x - o.x) * inv;
    tl = (lp.x - o.x) * inv;
  } else if (d.y != 0.0) {
    double inv = 1.0 / d.y;
    te = (ep.y - o.y) * inv;
    tl = (lp.y - o.y) * inv;
  } else {
    double inv = 1.0 / d.z;
    te = (ep.z - o.z) * inv;
    tl = (lp.z - o.z) * inv;
  }
}

// RayTetraPluecker (Platis & Theoharis), render/render_accel_tetra.cc:243-567.
// Outputs the entry face, uEnter1, uLeave1 (the call site's argument order
// puts these into isect.u and isect.v, :2103-2111) and both distances.
__device__ bool ray_tet(D3 orig, D3 dir, const Pl &plRay, const D3 vert[4], int &enterFace,
                        double &uE1, double &uL1, double &tE, double &tL) {
  enterFace = -1;
  double uE2 = 0, uL2 = 0;
  uE1 = 0;
  uL1 = 0;
  double uAB = 0, uAC = 0, uDB = 0, uDC = 0, uBC = 0, uAD = 0;
  int signAB = -2, signAC = -2, signDB = -2, signDC = -2, signBC = -2, signAD = -2;
  int nextSign = 0;
  D3 ep = {0, 0, 0}, lp = {0, 0, 0};

  uAB = pl_dot(plRay, mk_pl(vert[0], vert[1]));
  signAB = fsign(uAB);
  uAC = pl_dot(plRay, mk_pl(vert[0], vert[2]));
  signAC = fsign(uAC);

  if ((signAC == -signAB) || (signAC == 0) || (signAB == 0)) {
    uBC = pl_dot(plRay, mk_pl(vert[1], vert[2]));
    signBC = fsign(uBC);
    int signABC = signAB;
    if (signABC == 0) {
      signABC = -signAC;
      if (signABC == 0) signABC = signBC;
    }
    if ((signABC != 0) && ((signBC == signABC) || (signBC == 0))) {
      double invVolABC = 1.0 / ((uAB + uBC) - uAC);
      if (signABC == 1) {
        enterFace = 3;
        uE1 = -uAC * invVolABC;
        uE2 = uAB * invVolABC;
        ep = bary3(uE1, uE2, vert[0], vert[1], vert[2]);
        nextSign = -1;
      } else {
        uL1 = -uAC * invVolABC;
        uL2 = uAB * invVolABC;
        lp = bary3(uL1, uL2, vert[0], vert[1], vert[2]);
        nextSign = 1;
      }
      uAD = pl_dot(plRay, mk_pl(vert[0], vert[3]));
      signAD = fsign(uAD);
      if ((signAD == nextSign) || (signAD == 0)) {
        uDB = pl_dot(plRay, mk_pl(vert[3], vert[1]));
        signDB = fsign(uDB);
        if ((signDB == nextSign) || ((signDB == 0) && ((signAD != 0) || (signAB != 0)))) {
          double invVolBAD = 1.0 / ((uAD + uDB) - uAB);
          if (nextSign == 1) {
            enterFace = 2;
            uE1 = uDB * invVolBAD;
            uE2 = -uAB * invVolBAD;
            ep = bary3(uE1, uE2, vert[1], vert[0], vert[3]);
          } else {
            uL1 = uDB * invVolBAD;
            uL2 = -uAB * invVolBAD;
            lp = bary3(uL1, uL2, vert[1], vert[0], vert[3]);
          }
          param_dist(orig, dir, ep, lp, tE, tL);
          return true;
        }
      }
      uDC = pl_dot(plRay, mk_pl(vert[3], vert[2]));
      signDC = fsign(uDC);
      if ((signDC == -nextSign) || ((signDC == 0) && ((signAD != 0) || (signAC != 0)))) {
        double invVolCDA = 1.0 / ((uAC - uDC) - uAD);
        if (nextSign == 1) {
          enterFace = 1;
          uE1 = uAC * invVolCDA;
          uE2 = -uDC * invVolCDA;
          ep = bary3(uE1, uE2, vert[2], vert[3], vert[0]);
        } else {
          uL1 = uAC * invVolCDA;
          uL2 = -uDC * invVolCDA;
          lp = bary3(uL1, uL2, vert[2], vert[3], vert[0]);
        }
        param_dist(orig, dir, ep, lp, tE, tL);
        return true;
      } else {
        if (signDB == -2) uDB = pl_dot(plRay, mk_pl(vert[3], vert[1]));
        double invVolDCB = 1.0 / ((uDC - uBC) - uDB);
        if (nextSign == 1) {
          enterFace = 0;
          uE1 = -uDB * invVolDCB;
          uE2 = uDC * invVolDCB;
          ep = bary3(uE1, uE2, vert[3], vert[2], vert[1]);
        } else {
          uL1 = -uDB * invVolDCB;
          uL2 = uDC * invVolDCB;
          lp = bary3(uL1, uL2, vert[3], vert[2], vert[1]);
        }
        param_dist(orig, dir, ep, lp, tE, tL);
        return true;
      }
    }
  }

  uAD = pl_dot(plRay, mk_pl(vert[0], vert[3]));
  signAD = fsign(uAD);
  if ((signAD == -signAB) || (signAB == 0) || (signAD == 0)) {
    uDB = pl_dot(plRay, mk_pl(vert[3], vert[1]));
    signDB = fsign(uDB);
    int signBAD = -signAB;
    if (signBAD == 0) {
      signBAD = signAD;
      if (signBAD == 0) signBAD = signDB;
    }
    if ((signBAD != 0) && ((signDB == signBAD) || (signDB == 0))) {
      double invVolBAD = 1.0 / ((uAD + uDB) - uAB);
      if (signBAD == 1) {
        enterFace = 2;
        uE1 = uDB * invVolBAD;
        uE2 = -uAB * invVolBAD;
        ep = bary3(uE1, uE2, vert[1], vert[0], vert[3]);
        nextSign = -1;
      } else {
        uL1 = uDB * invVolBAD;
        uL2 = -uAB * invVolBAD;
        lp = bary3(uL1, uL2, vert[1], vert[0], vert[3]);
        nextSign = 1;
      }
      uDC = pl_dot(plRay, mk_pl(vert[3], vert[2]));
      signDC = fsign(uDC);
      if ((signDC == -nextSign) || ((signDC == 0) && ((signAD != 0) || (signAC != 0)))) {
        double invVolCDA = 1.0 / ((uAC - uDC) - uAD);
        if (nextSign == 1) {
          enterFace = 1;
          uE1 = uAC * invVolCDA;
          uE2 = -uDC * invVolCDA;
          ep = bary3(uE1, uE2, vert[2], vert[3], vert[0]);
        } else {
          uL1 = uAC * invVolCDA;
          uL2 = -uDC * invVolCDA;
          lp = bary3(uL1, uL2, vert[2], vert[3], vert[0]);
        }
        param_dist(orig, dir, ep, lp, tE, tL);
        return true;
      } else {
        if (signBC == -2) uBC = pl_dot(plRay, mk_pl(vert[1], vert[2]));
        double invVolDCB = 1.0 / ((uDC - uBC) - uDB);
        if (nextSign == 1) {
          enterFace = 0;
          uE1 = -uDB * invVolDCB;
          uE2 = uDC * invVolDCB;
          ep = bary3(uE1, uE2, vert[3], vert[2], vert[1]);
        } else {
          uL1 = -uDB * invVolDCB;
          uL2 = uDC * invVolDCB;
          lp = bary3(uL1, uL2, vert[3], vert[2], vert[1]);
        }
        param_dist(orig, dir, ep, lp, tE, tL);
        return true;
      }
    }
  }

  if ((-signAD == signAC) || (signAC == 0) || (signAD == 0)) {
    uDC = pl_dot(plRay, mk_pl(vert[3], vert[2]));
    signDC = fsign(uDC);
    int signCDA = signAC;
    if (signCDA == 0) {
      signCDA = -signAD;
      if (signCDA == 0) signCDA = -signDC;
    }
    if ((signCDA != 0) && ((-signDC == signCDA) || (signDC == 0))) {
      double invVolCDA = 1.0 / ((uAC - uDC) - uAD);
      if (signBC == -2) uBC = pl_dot(plRay, mk_pl(vert[1], vert[2]));
      if (signDB == -2) uDB = pl_dot(plRay, mk_pl(vert[3], vert[1]));
      double invVolDCB = 1.0 / ((uDC - uBC) - uDB);
      if (signCDA == 1) {
        enterFace = 1;
        uE1 = uAC * invVolCDA;
        uE2 = -uDC * invVolCDA;
        ep = bary3(uE1, uE2, vert[2], vert[3], vert[0]);
        uL1 = -uDB * invVolDCB;
        uL2 = uDC * invVolDCB;
        lp = bary3(uL1, uL2, vert[3], vert[2], vert[1]);
      } else {
        uL1 = uAC * invVolCDA;
        uL2 = -uDC * invVolCDA;
        lp = bary3(uL1, uL2, vert[2], vert[3], vert[0]);
        enterFace = 0;
        uE1 = -uDB * invVolDCB;
        uE2 = uDC * invVolDCB;
        ep = bary3(uE1, uE2, vert[3], vert[2], vert[1]);
      }
      param_dist(orig, dir, ep, lp, tE, tL);
      return true;
    }
  }
  return false;
}

// InterpolateTetra, interpolate_mode 0 (render/render_accel_tetra.cc:148-241),
// in float with the real3 operators' evaluation order.
struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 f3sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 f3add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 f3mul(F3 a, float f) { return {a.x * f, a.y * f, a.z * f}; }
__device__ __forceinline__ F3 f3div(F3 a, float f) { return {a.x / f, a.y / f, a.z / f}; }
__device__ __forceinline__ F3 f3cross(F3 l, F3 r) {
  return {l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x};
}
__device__ __forceinline__ float f3dot(F3 l, F3 r) { return (l.x * r.x + l.y * r.y) + l.z * r.z; }
__device__ __forceinline__ float f3len(F3 a) { return sqrtf((a.x * a.x + a.y * a.y) + a.z * a.z); }
__device__ __forceinline__ bool contains3(int v, int a, int b, int c) {
  return v == a || v == b || v == c;
}

__device__ void interpolate_tetra(float d_out[4], F3 pt, const F3 verts[4]) {
  const float kEPS = (float)0.001;
  F3 edges[6];
  edges[0] = f3sub(verts[1], verts[0]);
  edges[1] = f3sub(verts[2], verts[1]);
  edges[2] = f3sub(verts[0], verts[2]);
  edges[3] = f3sub(verts[3], verts[0]);
  edges[4] = f3sub(verts[3], verts[1]);
  edges[5] = f3sub(verts[3], verts[2]);
  for (int v = 0; v < 4; v++) {
    F3 rayorg = verts[v];
    F3 raydir = f3sub(pt, verts[v]);
    F3 raypc = f3cross(raydir, rayorg);
    float d = f3len(f3sub(pt, rayorg));
    float ws[6];
    bool cw0[6], cw1[6];
    for (int i = 0; i < 6; i++) {
      ws[i] = f3dot(raydir, f3cross(edges[i], verts[i % 3])) + f3dot(edges[i], raypc);
      cw0[i] = ws[i] >= -kEPS;
      cw1[i] = ws[i] <= kEPS;
    }
    F3 cp = {(float)1e16, (float)1e16, (float)1e16};
    // face vertex lists {0,2,1}, {1,2,3}, {0,3,2}, {0,1,3}
    if (cw0[0] && cw0[1] && cw0[2] && !contains3(v, 0, 2, 1))
      cp = f3div(f3add(f3add(f3mul(verts[2], ws[0]), f3mul(verts[0], ws[1])),
                       f3mul(verts[1], ws[2])),
                 (ws[0] + ws[1]) + ws[2]);
    else if (cw1[1] && cw0[4] && cw1[5] && !contains3(v, 1, 2, 3))
      cp = f3div(f3sub(f3add(f3mul(verts[3], ws[1]), f3mul(verts[1], ws[5])),
                       f3mul(verts[2], ws[4])),
                 (ws[5] + ws[1]) - ws[4]);
    else if (cw1[2] && cw1[3] && cw0[5] && !contains3(v, 0, 3, 2))
      cp = f3div(f3sub(f3add(f3mul(verts[3], ws[2]), f3mul(verts[2], ws[3])),
                       f3mul(verts[0], ws[5])),
                 (ws[2] + ws[3]) - ws[5]);
    else if (cw1[0] && cw0[3] && cw1[4] && !contains3(v, 0, 1, 3))
      cp = f3div(f3sub(f3add(f3mul(verts[3], ws[0]), f3mul(verts[0], ws[4])),
                       f3mul(verts[1], ws[3])),
                 (ws[0] + ws[4]) - ws[3]);
    float dv = d / f3len(f3sub(cp, rayorg));
    if (dv < 0.0f)
      dv = 0.0f;
    else if (!(dv < 1.0f) && !(dv > 0.0f))
      dv = 1.0f;
    d_out[v] = 1.0f - dv;
  }
  float ds = 0.0f, w = 0.0f, dd[4];
  for (int i = 0; i < 4; i++) ds += f3len(f3sub(pt, verts[i]));
  for (int i = 0; i < 4; i++) dd[i] = ds / f3len(f3sub(pt, verts[i]));
  for (int i = 0; i < 4; i++) {
    d_out[i] *= dd[i];
    w += d_out[i];
  }
  for (int i = 0; i < 4; i++) d_out[i] = d_out[i] / w;
}

// Entry distance of a box the ray crosses (t-independent part of
// IntersectRayAABB, render/render_accel_tetra.cc:1934-1978): returns false
// when tmax <= 0 or tmin > tmax; the caller compares tmin with the current
// hit distance when the node is popped, as the reference does.
__device__ __forceinline__ bool box_span(const RaySetup &s, float bx0, float by0, float bz0,
                                         float bx1, float by1, float bz1, float &tmin_out) {
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
  tmin_out = tmin;
  return (tmax > 0.0f) && (tmin <= tmax);
}

template <bool kAny>
__global__ void __launch_bounds__(kTraceBlock)
k_trace_tet(const float4 *__restrict__ nodes, const float4 *__restrict__ prims,
            const uint4 *__restrict__ sorted_faces, const BuildParams *__restrict__ params,
            const srf_ray *__restrict__ rays, uint32_t n, const float *__restrict__ scalars,
            srf_hit *__restrict__ hits, srf_tet_hit_ext *__restrict__ ext,
            const float *__restrict__ tmax_in, uint32_t *__restrict__ occluded,
            uint32_t max_nodes, uint32_t nprims) {
  // pending far children as (child, entry distance): the popped child is
  // checked against the current hit distance without reloading its parent
  // (24 KB of LDS per block holds the kernel at 3 wavefronts per SIMD, which
  // measured faster than 4 with a parent reload per pop: the traversal is
  // bound by the vector memory path, not by latency)
  __shared__ uint2 st_node[kTraceStack * kTraceBlock];
  const uint32_t i = xcd_block() * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t flags, prev;
  const RaySetup s = setup_ray(rays, i, flags, prev);
  const float tlimit = kAny ? (tmax_in ? tmax_in[i] : FLT_MAX) : FLT_MAX;
  float t = kAny ? tlimit : FLT_MAX, u = 0.0f, v = 0.0f;
  uint32_t best = SRF_NO_PRIM, subface = 0;
  double tleave = 0.0;
  const D3 ro = {s.o[0], s.o[1], s.o[2]}, rd = {s.d[0], s.d[1], s.d[2]};
  const Pl plRay = mk_pl(ro, d3add(ro, rd));
  // prefilter scales: float slab rounding is ~6e-8 of the coordinates
  // involved, the Pluecker entry distance's error ~1e-16 of them over the
  // direction component ComputeParametricDist divides by
  const BuildParams &P = *params;
  float cext = 0.0f;
  for (int k = 0; k < 3; k++)
    cext = fmaxf(cext, fmaxf(fabsf(P.root_bmin[k]), fabsf(P.root_bmax[k])));
  const float omax = fmaxf(fabsf(s.o[0]), fmaxf(fabsf(s.o[1]), fabsf(s.o[2])));
  const float pad = 1e-5f * (2.0f * cext + omax);
  const float dfirst = s.d[0] != 0.0f ? s.d[0] : (s.d[1] != 0.0f ? s.d[1] : s.d[2]);
  const float tslack = 1e-6f * (cext + omax) / fabsf(dfirst);
  uint2 *sn = st_node + threadIdx.x;
  int sp = 0;
  // pops entries until one whose box the ray enters no later than t
  auto pop_next = [&]() -> uint32_t {
    while (sp > 0) {
      const uint2 e = sn[(--sp) * kTraceBlock];
      if (__uint_as_float(e.y) <= t) return e.x;
    }
    return kDone;
  };
  uint32_t cur = kDone;
  {
    // the root's own box is tested first (TetraAccel::Traverse pops node 0)
    float tm;
    if (box_span(s, P.root_bmin[0], P.root_bmin[1], P.root_bmin[2], P.root_bmax[0],
                 P.root_bmax[1], P.root_bmax[2], tm) &&
        tm <= t)
      cur = 0;
  }
  uint4 meta;
  while (cur != kDone) {
    // descend: `cur` passed its own box test
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
      float tm0, tm1;
      const bool g0 = box_span(s, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x, tm0);
      const bool g1 = box_span(s, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w, tm1);
      const int nearc = (s.sbits >> meta.y) & 1u;
      const bool gn = nearc ? g1 : g0, gf = nearc ? g0 : g1;
      const float tn = nearc ? tm1 : tm0, tf = nearc ? tm0 : tm1;
      const uint32_t cn = nearc ? meta.w : meta.z, cf = nearc ? meta.z : meta.w;
      // the reference pushes far then near and pops near first
      if (gn && tn <= t) {
        if (gf && sp < kTraceStack) sn[(sp++) * kTraceBlock] = make_uint2(cf, __float_as_uint(tf));
        cur = cn;
        continue;
      }
      if (gf && tf <= t) {
        cur = cf;
        continue;
      }
      cur = pop_next();
      if (cur == kDone) break;
    }
    if (cur == kDone) break;
    // leaf: TestLeafNode, render/render_accel_tetra.cc:1980-2122
    const uint32_t count = meta.z, first = meta.w;
    for (uint32_t q = first; q < min(first + count, nprims); q++) {
      const float4 *Pp = prims + 4 * (size_t)q;
      const float4 p0 = Pp[0], p1 = Pp[1], p2 = Pp[2], p3 = Pp[3];  // one round trip
      keep_loaded(p1, p2, p3);
      const uint32_t id = __float_as_uint(p0.w);
      if (id == prev) continue;
      {
        const float bx0 = fminf(fminf(p0.x, p1.x), fminf(p2.x, p3.x)) - pad;
        const float by0 = fminf(fminf(p0.y, p1.y), fminf(p2.y, p3.y)) - pad;
        const float bz0 = fminf(fminf(p0.z, p1.z), fminf(p2.z, p3.z)) - pad;
        const float bx1 = fmaxf(fmaxf(p0.x, p1.x), fmaxf(p2.x, p3.x)) + pad;
        const float by1 = fmaxf(fmaxf(p0.y, p1.y), fmaxf(p2.y, p3.y)) + pad;
        const float bz1 = fmaxf(fmaxf(p0.z, p1.z), fmaxf(p2.z, p3.z)) + pad;
        float tm;
        if (!box_span(s, bx0, by0, bz0, bx1, by1, bz1, tm)) continue;
        if (tm > t + 1e-5f * t + tslack) continue;
      }
      const D3 vtx[4] = {{p0.x, p0.y, p0.z}, {p1.x, p1.y, p1.z}, {p2.x, p2.y, p2.z},
                         {p3.x, p3.y, p3.z}};
      int ef;
      double eu, lu, et = 0.0, lt = 0.0;
      if (ray_tet(ro, rd, plRay, vtx, ef, eu, lu, et, lt)) {
        if ((et >= 0.0) && (et < (double)t)) {
          if (kAny) {
            if ((float)et < tlimit) {
              occluded[i] = 1u;
              return;
            }
          } else {
            t = (float)et;
            u = (float)eu;
            v = (float)lu;
            best = q;
            subface = (uint32_t)ef;
            tleave = lt;
          }
        }
      }
    }
    cur = pop_next();
  }
  if (kAny) {
    occluded[i] = 0u;
    return;
  }
  uint32_t prim = SRF_NO_PRIM;
  srf_tet_hit_ext e = {};
  if (best != SRF_NO_PRIM) {
    const float4 *Pp = prims + 4 * (size_t)best;
    prim = __float_as_uint(Pp[0].w);
    if (ext) {
      F3 verts[4];
      for (int k = 0; k < 4; k++) verts[k] = {Pp[k].x, Pp[k].y, Pp[k].z};
      F3 pos = {s.o[0] + s.d[0] * t, s.o[1] + s.d[1] * t, s.o[2] + s.d[2] * t};
      interpolate_tetra(e.d, pos, verts);
      e.subface = subface;
      e.t_leave = (float)tleave;
      if (scalars) {
        const uint4 f = sorted_faces[best];
        e.scalar = ((scalars[f.x] * e.d[0] + scalars[f.y] * e.d[1]) + scalars[f.z] * e.d[2]) +
                   scalars[f.w] * e.d[3];
      }
    }
  }
  reinterpret_cast<float4 *>(hits)[i] = make_float4(t, u, v, __uint_as_float(prim));
  if (ext) {
    float4 *e4 = reinterpret_cast<float4 *>(ext + i);
    e4[0] = make_float4(e.d[0], e.d[1], e.d[2], e.d[3]);
    e4[1] = make_float4(e.scalar, __uint_as_float(e.subface), e.t_leave, 0.0f);
  }
}

// ------------------------------------------------- sorted / persistent any-hit

// Occlusion for rays taken in `order` (order[j] is the j-th ray to trace,
// the flag goes back to its own slot). With `counter` the grid is
// persistent: each wavefront takes 64 consecutive positions of the order at
// a time until none is left.
template <int K>
__global__ void __launch_bounds__(kTraceBlock)
k_any_ordered(const float4 *__restrict__ nodes, const float4 *__restrict__ prims,
              const uint32_t *__restrict__ sorted_index, const BuildParams *__restrict__ params,
              const srf_ray *__restrict__ rays, const float *__restrict__ tmax_in,
              const uint32_t *__restrict__ order, uint32_t n, uint32_t *__restrict__ occluded,
              uint32_t max_nodes, uint32_t nprims, uint32_t *counter, SphSub sub) {
  __shared__ uint32_t stack[(K == kSph ? 2 : 1) * kTraceStack * kTraceBlock];
  uint32_t *st = stack + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t base = (counter ? blockIdx.x : xcd_block()) * kTraceBlock + (threadIdx.x & ~63u);
  for (;;) {
    if (counter) {
      uint32_t b = 0;
      if (lane == 0) b = atomicAdd(counter, 64u);
      base = __shfl(b, 0, 64);
    }
    if (base >= n) break;
    const uint32_t j = base + lane;
    if (j < n) {
      const uint32_t i = order ? order[j] : j;
      uint32_t flags, prev;
      const RaySetup s = setup_ray(rays, i, flags, prev);
      const float tlimit = tmax_in ? tmax_in[i] : FLT_MAX;
      bool occ;
      if (K == kTri) {
        const float kEPS = FLT_EPSILON * (float)params->eps_scale;
        float t, u, v;
        uint32_t prim;
        occ = tri_traverse<true>(nodes, prims, max_nodes, nprims, s,
                                 (flags & SRF_RAY_DOUBLE_SIDED) != 0, kEPS, prev,
                                 tlimit, Stack32<kTraceBlock>{st}, t, u, v, prim);
      } else {
        const float zero[3] = {0.0f, 0.0f, 0.0f};
        float t;
        uint32_t best;
        occ = sph_traverse<true, kTraceBlock>(nodes, prims, sorted_index, max_nodes, nprims,
                                              sub, s, prev, 0.0, zero, tlimit, st, t, best);
      }
      occluded[i] = occ ? 1u : 0u;
    }
    if (!counter) break;
  }
}

// Sort key of a ray: direction octant over the 3D Morton code of its origin
// in the accelerator's scene box (origins outside clamp to the box), the top
// key_bits of those 30 bits.
__global__ void __launch_bounds__(256) k_ray_keys(const srf_ray *__restrict__ rays, uint32_t n,
                                                 const BuildParams *__restrict__ params,
                                                 uint32_t key_bits, uint32_t *__restrict__ keys) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 *r4 = reinterpret_cast<const float4 *>(rays + i);
  const float4 a = r4[0], b = r4[1];
  const float o[3] = {a.x, a.y, a.z};
  uint32_t q[3];
  for (int k = 0; k < 3; k++) {
    float x = (o[k] - params->bmin[k]) * params->inv[k];
    x = x > 0.0f ? (x < 1023.0f ? x : 1023.0f) : 0.0f;  // NaN goes to 0
    q[k] = (uint32_t)x;
  }
  const uint32_t m = (part_by2_30(q[0]) << 2) | (part_by2_30(q[1]) << 1) | part_by2_30(q[2]);
  const uint32_t oct = (b.x < 0.0f ? 4u : 0u) | (b.y < 0.0f ? 2u : 0u) | (b.z < 0.0f ? 1u : 0u);
  const uint32_t key30 = (oct << 27) | (m >> 3);
  keys[i] = key_bits >= 30 ? key30 : (key_bits == 0 ? 0u : key30 >> (30 - key_bits));
}

__global__ void k_zero_word(uint32_t *c) { *c = 0u; }

// PCG32 (O'Neill 2014, pcg32_srandom / pcg32_random): one stream per pixel.
struct Pcg32 {
  uint64_t state, inc;
  __device__ __forceinline__ uint32_t next() {
    const uint64_t old = state;
    state = old * 6364136223846793005ull + inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
  }
  __device__ __forceinline__ void seed(uint64_t initstate, uint64_t initseq) {
    state = 0u;
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
  }
  __device__ __forceinline__ float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

// Ambient-occlusion rays from sphere hits: for each primary ray i that hit
// sphere c, `per_hit` cosine-weighted directions about the outward normal
// n = normalize(p - c) at p = o + t d, from p + eps r n, with t_max = dist
// and prev_prim_id = the sphere (never reported again). Output slots are
// handed out per wavefront (the order of wavefronts is not fixed; every
// ray carries its pixel in out_pixel).
__global__ void __launch_bounds__(256)
k_ao_rays_sph(const srf_ray *__restrict__ primary, const srf_hit *__restrict__ hits, uint32_t n,
              const float *__restrict__ P, const float *__restrict__ R, float rconst,
              uint32_t per_hit, float eps, float dist, uint64_t seed, srf_ray *__restrict__ out,
              float *__restrict__ out_tmax, uint32_t *__restrict__ out_pixel, uint32_t *count) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t prim = SRF_NO_PRIM;
  float t = 0.0f;
  if (i < n) {
    const float4 h = reinterpret_cast<const float4 *>(hits)[i];
    prim = __float_as_uint(h.w);
    t = h.x;
  }
  const bool hit = prim != SRF_NO_PRIM;
  const uint64_t mask = __ballot(hit);
  uint32_t base = 0;
  if (lane == 0 && mask) base = atomicAdd(count, (uint32_t)__popcll(mask) * per_hit);
  base = __shfl(base, 0, 64);
  if (!hit) return;
  const uint64_t below = lane ? (mask & (~0ull >> (64 - lane))) : 0ull;
  const uint32_t slot = base + (uint32_t)__popcll(below) * per_hit;
  const float4 *r4 = reinterpret_cast<const float4 *>(primary + i);
  const float4 ra = r4[0], rb = r4[1];
  const float px = ra.x + t * rb.x, py = ra.y + t * rb.y, pz = ra.z + t * rb.z;
  const float cx = P[3 * (size_t)prim], cy = P[3 * (size_t)prim + 1], cz = P[3 * (size_t)prim + 2];
  const float r = R ? R[prim] : rconst;
  float nx = px - cx, ny = py - cy, nz = pz - cz;
  const float len = sqrtf((nx * nx + ny * ny) + nz * nz);
  if (len > 0.0f) {
    nx = nx / len;
    ny = ny / len;
    nz = nz / len;
  } else {
    nx = 0.0f;
    ny = 0.0f;
    nz = 1.0f;
  }
  // orthonormal basis about n (Duff et al. 2017)
  const float sg = nz >= 0.0f ? 1.0f : -1.0f;
  const float ia = -1.0f / (sg + nz);
  const float bb = nx * ny * ia;
  const float b1x = 1.0f + sg * nx * nx * ia, b1y = sg * bb, b1z = -sg * nx;
  const float b2x = bb, b2y = sg + ny * ny * ia, b2z = -ny;
  const float ox = px + eps * r * nx, oy = py + eps * r * ny, oz = pz + eps * r * nz;
  Pcg32 rng;
  rng.seed(seed, (uint64_t)i);
  for (uint32_t k = 0; k < per_hit; k++) {
    const float u1 = rng.uniform(), u2 = rng.uniform();
    const float rr = sqrtf(u1), phi = 6.2831853071795864769f * u2;
    const float lx = rr * cosf(phi), ly = rr * sinf(phi);
    const float lz = sqrtf(fmaxf(0.0f, 1.0f - u1));
    const float dx = (lx * b1x + ly * b2x) + lz * nx;
    const float dy = (lx * b1y + ly * b2y) + lz * ny;
    const float dz = (lx * b1z + ly * b2z) + lz * nz;
    float4 *o4 = reinterpret_cast<float4 *>(out + slot + k);
    o4[0] = make_float4(ox, oy, oz, __uint_as_float(0u));
    o4[1] = make_float4(dx, dy, dz, __uint_as_float(prim));
    out_tmax[slot + k] = dist;
    out_pixel[slot + k] = i;
  }
}

// blocks for n rays, padded to a multiple of kXcds (see xcd_block)
inline uint32_t trace_blocks(uint64_t n) {
  const uint64_t b = (n + kTraceBlock - 1) / kTraceBlock;
  return (uint32_t)((b + kXcdPad - 1) / kXcdPad * kXcdPad);
}
inline uint32_t max_nodes(const Accel *a) { return (uint32_t)(2 * a->n - 1); }
// every node index fits the 3-byte stack entries
inline bool small_nodes(const Accel *a) { return 2 * a->n - 1 <= (1u << 24); }
inline SphSub sph_sub(const Accel *a) {
  SphSub s{};
  if (a->kind == kSph && a->nbig) {
    s.nodes = (const float4 *)a->sub_nodes;
    s.prims = (const float4 *)a->sub_prims;
    s.sidx = a->sub_sidx;
    s.nnodes = (uint32_t)a->sub_nodes_used;
    s.nprims = (uint32_t)a->sub_prims_used;
  }
  return s;
}

}  // namespace

srf_status trace_tri(const Accel *a, const srf_ray *rays, uint64_t n, srf_hit *hits,
                     hipStream_t s) {
  if (n == 0) return SRF_OK;
  if (small_nodes(a))
    hipLaunchKernelGGL((k_trace_tri<false, true>), dim3(trace_blocks(n)), dim3(kTraceBlock), 0, s,
                       (const float4 *)a->nodes, (const float4 *)a->prims, a->params, rays,
                       (uint32_t)n, hits, nullptr, nullptr, max_nodes(a), (uint32_t)a->n);
  else
    hipLaunchKernelGGL((k_trace_tri<false, false>), dim3(trace_blocks(n)), dim3(kTraceBlock), 0,
                       s, (const float4 *)a->nodes, (const float4 *)a->prims, a->params, rays,
                       (uint32_t)n, hits, nullptr, nullptr, max_nodes(a), (uint32_t)a->n);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

srf_status trace_tet(const Accel *a, const srf_ray *rays, uint64_t n, const float *scalars,
                     srf_hit *hits, srf_tet_hit_ext *ext, hipStream_t s) {
  if (n == 0) return SRF_OK;
  hipLaunchKernelGGL(k_trace_tet<false>, dim3(trace_blocks(n)), dim3(kTraceBlock), 0, s,
                     (const float4 *)a->nodes, (const float4 *)a->prims,
                     (const uint4 *)a->sorted_faces, a->params, rays, (uint32_t)n, scalars, hits,
                     ext, nullptr, nullptr, max_nodes(a), (uint32_t)a->n);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

srf_status trace_sph(const Accel *a, const srf_ray *rays, const srf_sphere_prev *prev,
                     uint64_t n, srf_hit *hits, hipStream_t s) {
  if (n == 0) return SRF_OK;
  hipLaunchKernelGGL(k_trace_sph<false>, dim3(trace_blocks(n)), dim3(kTraceBlock), 0, s,
                     (const float4 *)a->nodes, (const float4 *)a->prims, a->sorted_index, rays,
                     prev, (uint32_t)n, hits, nullptr, nullptr, max_nodes(a), (uint32_t)a->n,
                     sph_sub(a));
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

srf_status trace_any(const Accel *a, const srf_ray *rays, const float *tmax, uint64_t n,
                     uint32_t *occluded, hipStream_t s) {
  if (n == 0) return SRF_OK;
  const dim3 g(trace_blocks(n)), b(kTraceBlock);
  if (a->kind == kTri && small_nodes(a))
    hipLaunchKernelGGL((k_trace_tri<true, true>), g, b, 0, s, (const float4 *)a->nodes,
                       (const float4 *)a->prims, a->params, rays, (uint32_t)n, nullptr, tmax,
                       occluded, max_nodes(a), (uint32_t)a->n);
  else if (a->kind == kTri)
    hipLaunchKernelGGL((k_trace_tri<true, false>), g, b, 0, s, (const float4 *)a->nodes,
                       (const float4 *)a->prims, a->params, rays, (uint32_t)n, nullptr, tmax,
                       occluded, max_nodes(a), (uint32_t)a->n);
  else if (a->kind == kTet)
    hipLaunchKernelGGL(k_trace_tet<true>, g, b, 0, s, (const float4 *)a->nodes,
                       (const float4 *)a->prims, (const uint4 *)a->sorted_faces, a->params, rays,
                       (uint32_t)n, nullptr, nullptr, nullptr, tmax, occluded, max_nodes(a),
                       (uint32_t)a->n);
  else
    hipLaunchKernelGGL(k_trace_sph<true>, g, b, 0, s, (const float4 *)a->nodes,
                       (const float4 *)a->prims, a->sorted_index, rays, nullptr, (uint32_t)n,
                       nullptr, tmax, occluded, max_nodes(a), (uint32_t)a->n, sph_sub(a));
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

srf_status trace_any_sorted(Ctx *ctx, const Accel *a, const srf_ray *rays, const float *tmax,
                            uint64_t n, uint32_t key_bits, int mode, uint32_t *occluded,
                            hipStream_t s) {
  if (n == 0) return SRF_OK;
  if (a->kind == kTet)
    return set_error(SRF_ERR_UNSUPPORTED, "sorted any-hit takes triangles or spheres");
  if (key_bits > 30) return set_error(SRF_ERR_INVALID_ARG, "key_bits must be <= 30");
  const uint32_t nn = (uint32_t)n;
  // scratch: [counter][keys][sorted keys][order] then the sort's own slices
  const size_t head = 256 + 3 * (((size_t)4 * nn + 255) & ~(size_t)255);
  const uint32_t *order = nullptr;
  uint32_t *counter = nullptr;
  if (key_bits > 0 || mode == SRF_LAUNCH_PERSISTENT) {
    srf_status st = ensure_scratch(ctx, head + sort_pairs30_bytes(nn), s);
    if (st != SRF_OK) return st;
    Arena ar{(char *)ctx->scratch, ctx->scratch_bytes, 0};
    counter = ar.take<uint32_t>(64);
    uint32_t *keys = ar.take<uint32_t>(nn);
    uint32_t *skeys = ar.take<uint32_t>(nn);
    uint32_t *ord = ar.take<uint32_t>(nn);
    if (key_bits > 0) {
      hipLaunchKernelGGL(k_ray_keys, dim3((nn + 255) / 256), dim3(256), 0, s, rays, nn, a->params,
                         key_bits, keys);
      SRF_HIP(hipGetLastError());
      st = sort_pairs30_at(ctx, head, keys, nn, skeys, ord, s);
      if (st != SRF_OK) return st;
      order = ord;
    }
    if (mode != SRF_LAUNCH_PERSISTENT) counter = nullptr;
  }
  uint32_t blocks = trace_blocks(n);
  if (counter) {
    int dev = 0, cus = 0;
    SRF_HIP(hipGetDevice(&dev));
    SRF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    blocks = std::min<uint32_t>(blocks, (uint32_t)cus * 16u);
    hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, s, counter);
  }
  if (a->kind == kTri)
    hipLaunchKernelGGL(k_any_ordered<kTri>, dim3(blocks), dim3(kTraceBlock), 0, s,
                       (const float4 *)a->nodes, (const float4 *)a->prims, a->sorted_index,
                       a->params, rays, tmax, order, nn, occluded, max_nodes(a), (uint32_t)a->n,
                       counter, SphSub{});
  else
    hipLaunchKernelGGL(k_any_ordered<kSph>, dim3(blocks), dim3(kTraceBlock), 0, s,
                       (const float4 *)a->nodes, (const float4 *)a->prims, a->sorted_index,
                       a->params, rays, tmax, order, nn, occluded, max_nodes(a), (uint32_t)a->n,
                       counter, sph_sub(a));
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

srf_status ao_rays_spheres(const srf_ray *primary, const srf_hit *hits, uint64_t n,
                           const float *positions, const float *radius, float rconst,
                           uint32_t per_hit, float eps, float dist, uint64_t seed,
                           srf_ray *out, float *out_tmax, uint32_t *out_pixel,
                           uint32_t *count, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, s, count);
  if (n == 0) return SRF_OK;
  hipLaunchKernelGGL(k_ao_rays_sph, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, primary,
                     hits, (uint32_t)n, positions, radius, rconst, per_hit, eps, dist, seed, out,
                     out_tmax, out_pixel, count);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

}  // namespace srf
