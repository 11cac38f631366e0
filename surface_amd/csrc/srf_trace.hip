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
//   solids     SolidAccel::Traverse     render/render_accel_solid.cc:2345-2435
//   lines      LineAccel::Traverse      render/render_accel_line.cc:704-774
#include <algorithm>
#include <mutex>

#include "srf_trace_dev.h"

namespace srf {
namespace {

// A node read by a whole wavefront (uniform index: scalar loads).
struct WaveNode {
  float4 n0, n1, n2;
  uint4 meta;
};
__device__ __forceinline__ WaveNode load_wave_node(const float4 *__restrict__ nodes,
                                                   uint32_t idx) {
  const float4 *p = nodes + 4 * (size_t)idx;
  WaveNode r;
  r.n0 = p[0];
  r.n1 = p[1];
  r.n2 = p[2];
  r.meta = reinterpret_cast<const uint4 *>(p)[3];
  return r;
}

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
  __shared__ uint32_t wnode[kAny ? kTraceBlock / 64 : 1][kTraceStack];
  __shared__ uint64_t wmask[kAny ? kTraceBlock / 64 : 1][kTraceStack];
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
  const uint32_t sb = __builtin_amdgcn_readfirstlane(s.sbits);
  if (kAny && __all(s.sbits == sb)) {
    // Occlusion once per wavefront when the lanes' direction signs agree: at
    // the fixed search distance tlimit the nodes and triangles a lane's
    // traversal tests, and whether one occludes, do not depend on the order;
    // the wavefront walks the union of its lanes' nodes and a lane leaves
    // once occluded.
    const bool filter = kEPS >= 1e-30f;
    const int wv = threadIdx.x >> 6;
    const uint64_t me = 1ull << (threadIdx.x & 63u);
    uint64_t mask = __ballot(1);
    uint32_t cur = 0;
    int wsp = 0;
    occ = false;
    for (;;) {
      uint4 meta = make_uint4(0u, 0u, 0u, 0u);
      bool leaf = false;
      while (cur < max_nodes) {
        cur = __builtin_amdgcn_readfirstlane(cur);
        const float4 *nd = nodes + 4 * (size_t)cur;
        const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
        meta = reinterpret_cast<const uint4 *>(nd)[3];
        keep_loaded_s(n0, n1, n2);
        if (meta.x != 0) {
          leaf = true;
          break;
        }
        const bool in = (mask & me) != 0 && !occ;
        const bool h0 =
            in && ray_aabb_wave<true>(s, sb, tlimit, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x);
        const bool h1 =
            in && ray_aabb_wave<true>(s, sb, tlimit, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w);
        const int nearc = (sb >> meta.y) & 1u;
        const uint64_t nm = __ballot(nearc ? h1 : h0), fm = __ballot(nearc ? h0 : h1);
        const uint32_t cn = nearc ? meta.w : meta.z, cf = nearc ? meta.z : meta.w;
        if (nm) {
          if (fm && wsp < kTraceStack) {
            wnode[wv][wsp] = cf;
            wmask[wv][wsp] = fm;
            wsp++;
          }
          cur = cn;
          mask = nm;
        } else if (fm) {
          cur = cf;
          mask = fm;
        } else {
          break;
        }
      }
      if (leaf) {
        const bool in = (mask & me) != 0;
        const uint32_t q0 = __builtin_amdgcn_readfirstlane(meta.w);
        const uint32_t q1 = __builtin_amdgcn_readfirstlane(min(meta.w + meta.z, nprims));
        for (uint32_t q = q0; q < q1; q++) {
          const float4 *P = prims + 3 * (size_t)q;
          const float4 A = P[0], B = P[1], C = P[2];
          if (!in || occ || __float_as_uint(A.w) == prev) continue;
          float tt, uu, vv;
          if (tri_isect(s, A, B, C, ds, kEPS, filter, tlimit, tt, uu, vv) && tt < tlimit)
            occ = true;
        }
      }
      if (__ballot(!occ) == 0 || wsp == 0) break;
      --wsp;
      cur = __builtin_amdgcn_readfirstlane(wnode[wv][wsp]);
      mask = wmask[wv][wsp];
    }
  } else if (kSmall)
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

// The wavefront traversal of k_trace_tri_wave for rays whose direction signs
// are all SB (sign[k] << k): with the signs fixed at compile time the box
// tests choose their near and far planes without per-lane selects.
template <uint32_t SB>
__device__ __forceinline__ void tri_wave_closest(const float4 *__restrict__ nodes,
                                              const float4 *__restrict__ prims,
                                              uint32_t max_nodes, uint32_t nprims,
                                              const RaySetup &s, bool ds, float kEPS,
                                              uint32_t prev, uint32_t *wnode, uint64_t *wmask,
                                              float &t, float &u, float &v, uint32_t &prim) {
  constexpr uint32_t sb = SB;
  const bool filter = kEPS >= 1e-30f;
  const uint64_t me = 1ull << (threadIdx.x & 63u);
  uint64_t mask = __ballot(1);
  uint32_t cur = 0;
  int sp = 0;
  for (;;) {
    // descend while some lane's traversal goes on below this node
    uint4 meta = make_uint4(0u, 0u, 0u, 0u);
    bool leaf = false;
    while (cur < max_nodes) {
      cur = __builtin_amdgcn_readfirstlane(cur);
      const float4 *nd = nodes + 4 * (size_t)cur;
      const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
      meta = reinterpret_cast<const uint4 *>(nd)[3];
      keep_loaded_s(n0, n1, n2);
      if (meta.x != 0) {
        leaf = true;
        break;
      }
      const bool in = (mask & me) != 0;
      const bool h0 = in && ray_aabb_wave<true>(s, SB, t, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x);
      const bool h1 = in && ray_aabb_wave<true>(s, SB, t, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w);
      const int nearc = (sb >> meta.y) & 1u;
      const uint64_t nm = __ballot(nearc ? h1 : h0), fm = __ballot(nearc ? h0 : h1);
      const uint32_t cn = nearc ? meta.w : meta.z, cf = nearc ? meta.z : meta.w;
      if (nm) {
        // the lanes hitting both boxes pushed the far child; those hitting
        // only the far one go there next, after nothing of theirs
        if (fm && sp < kTraceStack) {
          wnode[sp] = cf;
          wmask[sp] = fm;
          sp++;
        }
        cur = cn;
        mask = nm;
      } else if (fm) {
        cur = cf;
        mask = fm;
      } else {
        break;
      }
    }
    if (leaf) {
      // TestLeafNode, render/render_accel_triangle.cc:2386-2433
      const bool in = (mask & me) != 0;
      const uint32_t first = __builtin_amdgcn_readfirstlane(meta.w),
                     end = __builtin_amdgcn_readfirstlane(min(meta.w + meta.z, nprims));
      for (uint32_t q = first; q < end; q += 4) {
        // four triangles' records in one round of scalar loads (past the
        // leaf's end the last one again, unused), tested in order
        float4 A[4], B[4], C[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const float4 *P = prims + 3 * (size_t)min(q + (uint32_t)j, end - 1);
          A[j] = P[0];
          B[j] = P[1];
          C[j] = P[2];
        }
#pragma unroll
        for (int j = 0; j < 4; j++) keep_loaded_s(A[j], B[j], C[j]);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (q + (uint32_t)j >= end) break;
          if (in && __float_as_uint(A[j].w) != prev) {
            float tt, uu, vv;
            if (tri_isect(s, A[j], B[j], C[j], ds, kEPS, filter, t, tt, uu, vv)) {
              t = tt;
              u = uu;
              v = vv;
              prim = __float_as_uint(A[j].w);
            }
          }
        }
      }
    }
    if (sp == 0) break;
    --sp;
    cur = __builtin_amdgcn_readfirstlane(wnode[sp]);
    mask = wmask[sp];
  }
}

// Closest hit with one traversal per wavefront. When every ray of the
// wavefront has the same direction signs, the reference's per-ray order
// (near child by the split axis's sign, far child pushed when both boxes are
// hit, no re-test at the pop) is the same static pre-order for all of them,
// so the wavefront walks the union of their paths in that order with one
// shared stack of (node, lane mask): a lane takes part in a node exactly when
// its own traversal would visit it, with the distance its own traversal
// would hold there (between two of its nodes it is masked off and changes
// nothing), so every lane tests the same triangles in the same order as the
// per-ray kernel and ends with the same hit. Nodes and triangles are loaded
// once per wavefront instead of once per lane (the node and triangle
// indices are made wavefront-uniform, so the loads go through the scalar
// unit). A wavefront with mixed signs runs the per-ray traversal.
template <bool kSmall>
__global__ void __launch_bounds__(kTraceBlock)
k_trace_tri_wave(const float4 *__restrict__ nodes, const float4 *__restrict__ prims,
                 const BuildParams *__restrict__ params, const srf_ray *__restrict__ rays,
                 uint32_t n, srf_hit *__restrict__ hits, uint32_t max_nodes, uint32_t nprims) {
  __shared__ uint32_t stack[kSmall ? 1 : kTraceStack * kTraceBlock];
  __shared__ uint16_t stack_lo[kSmall ? kTraceStack * kTraceBlock : 1];
  __shared__ uint8_t stack_hi[kSmall ? kTraceStack * kTraceBlock : 1];
  __shared__ uint32_t wnode[kTraceBlock / 64][kTraceStack];
  __shared__ uint64_t wmask[kTraceBlock / 64][kTraceStack];
  const uint32_t i = xcd_block() * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t flags, prev;
  const RaySetup s = setup_ray(rays, i, flags, prev);
  const bool ds = (flags & SRF_RAY_DOUBLE_SIDED) != 0;
  const float kEPS = FLT_EPSILON * (float)params->eps_scale;
  float t = FLT_MAX, u = 0.0f, v = 0.0f;
  uint32_t prim = SRF_NO_PRIM;
  const uint32_t sb = __builtin_amdgcn_readfirstlane(s.sbits);
  if (!__all(s.sbits == sb)) {
    if (kSmall)
      tri_traverse<false>(nodes, prims, max_nodes, nprims, s, ds, kEPS, prev, FLT_MAX,
                          Stack24<kTraceBlock>{stack_lo + threadIdx.x, stack_hi + threadIdx.x},
                          t, u, v, prim);
    else
      tri_traverse<false>(nodes, prims, max_nodes, nprims, s, ds, kEPS, prev, FLT_MAX,
                          Stack32<kTraceBlock>{stack + threadIdx.x}, t, u, v, prim);
  } else {
    uint32_t *wn = wnode[threadIdx.x >> 6];
    uint64_t *wm = wmask[threadIdx.x >> 6];
#define SRF_TRI_WAVE(B)                                                                    \
  case B:                                                                                  \
    tri_wave_closest<B>(nodes, prims, max_nodes, nprims, s, ds, kEPS, prev, wn, wm, t, u, v, \
                        prim);                                                             \
    break;
    switch (sb) {
      SRF_TRI_WAVE(0)
      SRF_TRI_WAVE(1)
      SRF_TRI_WAVE(2)
      SRF_TRI_WAVE(3)
      SRF_TRI_WAVE(4)
      SRF_TRI_WAVE(5)
      SRF_TRI_WAVE(6)
      SRF_TRI_WAVE(7)
      default:
        break;
    }
#undef SRF_TRI_WAVE
    if (prim == SRF_NO_PRIM) t = FLT_MAX;
  }
  reinterpret_cast<float4 *>(hits)[i] = make_float4(t, u, v, __uint_as_float(prim));
}

// ------------------------------------------------------------------ spheres

// SphereIsect, render/render_accel_particle.cc:2423-2512
// `tcut`: the caller keeps a reported t only if it is at most tcut, so a
// sphere whose two roots both certainly lie beyond it is rejected before the
// exact roots are formed: q is computed exactly as below, and q / a and
// cc / q are estimated with hardware reciprocals (within 3 ulp of the
// correctly rounded quotients; the margin is 2e-5 relative). q itself may
// carry cancellation error (its formula is the reference's), which the
// estimates share with the exact roots, so it costs no accuracy here.
__device__ __forceinline__ bool sphere_isect(const RaySetup &s, float4 c, float tin,
                                             double prevT, const float *prevN, bool self,
                                             float tcut, float &tout) {
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
    const float dsq = sqrtf(disc);
    {
      // q below, exactly (halving is exact above FLT_MIN); only the two
      // divisions are estimated
      const float qe = (b < 0.0f) ? (-b - dsq) * 0.5f : (-b + dsq) * 0.5f;
      const float lim = tcut * 1.00002f;
      if (fabsf(qe) >= FLT_MIN && a >= FLT_MIN && qe * __builtin_amdgcn_rcpf(a) > lim &&
          cc * __builtin_amdgcn_rcpf(qe) > lim)
        return false;
    }
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
// ones. With u = 2^-24, s = |o - c| and A = |d|^2: oc = fl(o - c) is exact
// for a centre c' within u s of c; against that centre the dot product is
// within 3u |d| s, |oc|^2 and a within 3u relative, r^2 and the subtraction
// within u, b * b within u, the double 4 a cc exact, so the computed
// discriminant differs from the exact one of (c', r) by at most
// 56 u A s^2 + 4 u A r^2. A root of it is the entry into the ball around c'
// of radius sqrt(r^2 (1 + u) + 14 u s^2) <= r + 9.2e-4 s, and the remaining
// roundings (sqrt, the root formula, the divisions) move it by O(10 u s)
// along the ray. So a reported root lies no nearer than the entry into the
// ball of radius r + 1.0e-3 s around c. The box is grown by 1.25e-3 of the
// farthest distance from the origin to it (>= s + r for every sphere
// inside), which bounds that with a margin of 25 %. A ray whose origin is
// inside the grown box gets lo < 0.
//
// The padding is computed once per big leaf (sub_pad) from the root's two
// boxes, which contain every box below them: the farthest distance from the
// origin to them is at least the farthest distance to any node's box, so the
// padding is at least the one the bound asks for at every node.
__device__ __forceinline__ float sub_pad(const RaySetup &s, float4 n0, float4 n1, float4 n2) {
  const float x0 = fminf(n0.x, n0.w), y0 = fminf(n0.y, n1.x), z0 = fminf(n0.z, n1.y);
  const float x1 = fmaxf(n1.z, n2.y), y1 = fmaxf(n1.w, n2.z), z1 = fmaxf(n2.x, n2.w);
  const float fx = fmaxf(fabsf(s.o[0] - x0), fabsf(s.o[0] - x1));
  const float fy = fmaxf(fabsf(s.o[1] - y0), fabsf(s.o[1] - y1));
  const float fz = fmaxf(fabsf(s.o[2] - z0), fabsf(s.o[2] - z1));
  // (the hardware square root is within 1 ulp; the bound has a 25 % margin)
  return 1.25e-3f * __builtin_amdgcn_sqrtf((fx * fx + fy * fy) + fz * fz);
}

// The slab test of a box grown by `pad` on every side (sub_pad's value).
__device__ __forceinline__ bool sub_box(const RaySetup &s, float pad, float x0, float y0,
                                        float z0, float x1, float y1, float z1, float &lo) {
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
  float pad = 0.0f;
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
      if (cur == 0) pad = sub_pad(s, n0, n1, n2);  // the root, visited first
      float l0, l1;
      bool h0 = sub_box(s, pad, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x, l0);
      bool h1 = sub_box(s, pad, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w, l1);
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
      if (!sphere_isect(s, c, INFINITY, prevT, prevN, self, kAny ? tlimit : t, tt)) continue;
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

// The same for all lanes of a wavefront at once (called with uniform
// control flow; `in` says whether this lane takes part). The lanes share
// one walk over the secondary tree -- node and sphere indices are made
// wavefront-uniform, so their loads go through the scalar unit -- in the
// order of their common direction signs (near child by the node's split
// axis), with a shared stack of (node, lane mask): a lane takes part in a
// node when its own bound test passes at its own kept distance. Whether a
// box can be skipped depends only on the lane's kept distance, and the
// leaf's result (the smallest (t, position) below the distance on entry;
// kAny: any t below tlimit) does not depend on the order the spheres are
// tested in, so every lane ends with what sph_big_leaf gives it. sb: the
// wavefront's sign bits; wn / wm: kTraceStack stack entries of this
// wavefront.
template <bool kAny>
__device__ __forceinline__ void sph_big_leaf_wave(const SphSub &sub, uint32_t nbase,
                                                  uint32_t pbase, uint32_t first, uint32_t count,
                                                  const uint32_t *__restrict__ sorted_index,
                                                  const RaySetup &s, uint32_t sb, uint32_t prev,
                                                  double prevT, const float prevN[3],
                                                  float tlimit, bool in, uint32_t *wn,
                                                  uint64_t *wm, float &t, uint32_t &best,
                                                  bool &occ) {
  const uint64_t me = 1ull << (threadIdx.x & 63u);
  const uint64_t all = __ballot(in);
  uint64_t mask = all;
  bool kept = false;  // `best` is one of this leaf's spheres
  const uint32_t nnodes = 2 * count - 1;
  // a sphere's own id matters only to a lane with a previous hit
  const bool any_prev = __ballot(in && prev != SRF_NO_PRIM) != 0;
  uint32_t cur = 0;
  int wsp = 0;
  float pad = 0.0f;
  WaveNode nd;
  while (mask) {
    bool leaf = false, bad = false;
    for (;;) {
      cur = __builtin_amdgcn_readfirstlane(cur);
      if (cur >= nnodes || nbase + cur >= sub.nnodes) {
        bad = true;
        break;
      }
      nd = load_wave_node(sub.nodes, nbase + cur);
      if (nd.meta.x != 0) {
        leaf = true;
        break;
      }
      const uint32_t c0 = nd.meta.z, c1 = nd.meta.w;
      const bool act = (mask & me) != 0 && !(kAny && occ);
      float l0 = 0.0f, l1 = 0.0f;
      if (cur == 0) pad = sub_pad(s, nd.n0, nd.n1, nd.n2);  // the root, visited first
      bool h0 = act && sub_box(s, pad, nd.n0.x, nd.n0.y, nd.n0.z, nd.n1.z, nd.n1.w, nd.n2.x, l0);
      bool h1 = act && sub_box(s, pad, nd.n0.w, nd.n1.x, nd.n1.y, nd.n2.y, nd.n2.z, nd.n2.w, l1);
      h0 = h0 && (kAny ? l0 < tlimit : l0 <= t);
      h1 = h1 && (kAny ? l1 < tlimit : l1 <= t);
      const int nearc = (sb >> nd.meta.y) & 1u;
      const uint64_t nm = __ballot(nearc ? h1 : h0), fm = __ballot(nearc ? h0 : h1);
      const uint32_t cn = nearc ? c1 : c0, cf = nearc ? c0 : c1;
      if (nm) {
        if (fm && wsp < kTraceStack) {
          wn[wsp] = cf;
          wm[wsp] = fm;
          wsp++;
        }
        cur = cn;
        mask = nm;
      } else if (fm) {
        cur = cf;
        mask = fm;
      } else {
        break;
      }
    }
    if (bad) break;
    if (leaf) {
      const bool act = (mask & me) != 0;
      const uint32_t k0 = __builtin_amdgcn_readfirstlane(nd.meta.w);
      const uint32_t k1 = __builtin_amdgcn_readfirstlane(min(nd.meta.w + nd.meta.z, count));
      // the leaf's spheres two at a time (order and rules unchanged)
      for (uint32_t k = k0; k < k1; k += 2) {
        float4 c[2];
        uint32_t loc[2];
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
          // past the leaf's end (or the array's) the last entry is loaded again, unused
          uint32_t j = pbase + min(k + (uint32_t)jj, k1 - 1);
          j = j < sub.nprims ? j : sub.nprims - 1;
          c[jj] = sub.prims[j];
          loc[jj] = sub.sidx[j];
        }
        keep_loaded_s2(c[0], c[1]);
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
          const uint32_t kk = k + (uint32_t)jj;
          if (kk >= k1 || pbase + kk >= sub.nprims) break;
          if (loc[jj] >= count) continue;
          const uint32_t q = first + loc[jj];
          const uint32_t sid = any_prev ? sorted_index[q] : SRF_NO_PRIM;
          if (!act || (kAny && occ)) continue;
          const bool self = (prev != SRF_NO_PRIM) && (sid == prev);
          float tt;
          if (!sphere_isect(s, c[jj], INFINITY, prevT, prevN, self, kAny ? tlimit : t, tt))
            continue;
          if (kAny) {
            if (tt < tlimit) occ = true;
          } else if (tt < t || (tt == t && kept && q < best)) {
            t = tt;
            best = q;
            kept = true;
          }
        }
      }
    }
    if (kAny && (__ballot(occ) & all) == all) break;
    if (wsp == 0) break;
    --wsp;
    cur = wn[wsp];
    mask = wm[wsp];
  }
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
      if (sphere_isect(s, c, t, prevT, prevN, self, kAny ? tlimit : t, tt)) {
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

// Closest hit (and plain occlusion) of spheres. kPart selects the
// wavefronts a launch handles: 0 all, 1 those whose lanes share direction
// signs (the wavefront traversal; no per-lane stack, so the launch holds
// more wavefronts per CU), 2 the others (the per-lane traversal). Part 1
// followed by part 2 on the same rays gives every ray once.
template <bool kAny, int kPart>
__global__ void __launch_bounds__(kTraceBlock)
k_trace_sph(const float4 *__restrict__ nodes, const float4 *__restrict__ prims,
            const uint32_t *__restrict__ sorted_index, const srf_ray *__restrict__ rays,
            const srf_sphere_prev *__restrict__ prevs, uint32_t n, srf_hit *__restrict__ hits,
            const float *__restrict__ tmax_in, uint32_t *__restrict__ occluded,
            uint32_t max_nodes, uint32_t nprims, SphSub sub) {
  constexpr bool kWavePath = !kAny && kPart != 2;
  constexpr bool kLanePath = kPart != 1;
  __shared__ uint32_t stack[kLanePath ? 2 * kTraceStack * kTraceBlock : 1];
  __shared__ uint32_t wnode[kWavePath ? kTraceBlock / 64 : 1][2 * kTraceStack];
  __shared__ uint64_t wmask[kWavePath ? kTraceBlock / 64 : 1][2 * kTraceStack];
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
  float t = FLT_MAX;
  uint32_t best = SRF_NO_PRIM;
  bool occ = false;
  const uint32_t sb = __builtin_amdgcn_readfirstlane(s.sbits);
  const bool uniform = __all(s.sbits == sb);
  if (kPart == 1 && !uniform) return;  // left to part 2
  if (kPart == 2 && !kAny && uniform) return;  // done by part 1
  if (kWavePath && uniform) {
    // Closest hit once per wavefront, as k_trace_tri_wave: with the lanes'
    // direction signs equal, the reference's order (near child by the split
    // axis's sign, far child pushed, no re-test at the pop) is one static
    // pre-order, and each lane takes part exactly in the nodes its own
    // traversal visits, holding the distance that traversal would hold
    // there. Small leaves are scanned once per wavefront in the reference's
    // order; a big leaf's secondary tree, whose result does not depend on
    // its order, runs per lane.
    t = FLT_MAX;
    best = SRF_NO_PRIM;
    const int wv = threadIdx.x >> 6;
    const uint64_t me = 1ull << (threadIdx.x & 63u);
    uint64_t mask = __ballot(1);
    uint32_t cur = 0;
    int wsp = 0;
    // the descent's node in scalar registers, both children requested before
    // its boxes are tested, so the next node of the descent is in registers
    // by the time the tests have picked it
    WaveNode nd;
    bool have = false;
    for (;;) {
      bool leaf = false;
      while (cur < max_nodes) {
        cur = __builtin_amdgcn_readfirstlane(cur);
        if (!have) nd = load_wave_node(nodes, cur);
        have = false;
        if (nd.meta.x != 0) {
          leaf = true;
          break;
        }
        const uint32_t c0 = nd.meta.z, c1 = nd.meta.w;
        const WaveNode p0 = load_wave_node(nodes, c0 < max_nodes ? c0 : 0u);
        const WaveNode p1 = load_wave_node(nodes, c1 < max_nodes ? c1 : 0u);
        const bool in = (mask & me) != 0;
        const bool h0 = in && ray_aabb_wave<true>(s, sb, t, nd.n0.x, nd.n0.y, nd.n0.z, nd.n1.z,
                                                  nd.n1.w, nd.n2.x);
        const bool h1 = in && ray_aabb_wave<true>(s, sb, t, nd.n0.w, nd.n1.x, nd.n1.y, nd.n2.y,
                                                  nd.n2.z, nd.n2.w);
        const int nearc = (sb >> nd.meta.y) & 1u;
        const uint64_t nm = __ballot(nearc ? h1 : h0), fm = __ballot(nearc ? h0 : h1);
        const uint32_t cn = nearc ? c1 : c0, cf = nearc ? c0 : c1;
        if (nm) {
          if (fm && wsp < kTraceStack) {
            wnode[wv][wsp] = cf;
            wmask[wv][wsp] = fm;
            wsp++;
          }
          cur = cn;
          mask = nm;
        } else if (fm) {
          cur = cf;
          mask = fm;
        } else {
          break;
        }
        nd = cur == c0 ? p0 : p1;
        have = cur < max_nodes;
      }
      const uint4 meta = nd.meta;
      if (leaf) {
        const bool in = (mask & me) != 0;
        if (meta.z > kBigLeaf && sub.nodes) {
          const uint32_t nbase = __builtin_amdgcn_readfirstlane(__float_as_uint(nd.n0.w));
          const uint32_t pbase = __builtin_amdgcn_readfirstlane(__float_as_uint(nd.n1.x));
          if (meta.w < nprims && meta.z <= nprims - meta.w) {
            bool unused = false;
            sph_big_leaf_wave<false>(sub, nbase, pbase, meta.w, meta.z, sorted_index, s, sb, prev,
                                     prevT, prevN, FLT_MAX, in, &wnode[wv][kTraceStack],
                                     &wmask[wv][kTraceStack], t, best, unused);
          }
        } else {
          // the leaf's spheres in order, loaded two at a time
          const uint32_t q0 = __builtin_amdgcn_readfirstlane(meta.w);
          const uint32_t q1 = __builtin_amdgcn_readfirstlane(min(meta.w + meta.z, nprims));
          for (uint32_t q = q0; q < q1; q += 2) {
            float4 c[2];
            uint32_t sid[2];
#pragma unroll
            for (int j = 0; j < 2; j++) {
              const uint32_t qq = min(q + (uint32_t)j, q1 - 1);  // past the end: unused
              c[j] = prims[qq];
              sid[j] = sorted_index[qq];
            }
            keep_loaded_s2(c[0], c[1]);
#pragma unroll
            for (int j = 0; j < 2; j++) {
              if (q + (uint32_t)j >= q1) break;
              if (!in) continue;
              const bool self = (prev != SRF_NO_PRIM) && (sid[j] == prev);
              float tt;
              if (sphere_isect(s, c[j], t, prevT, prevN, self, t, tt)) {
                t = tt;
                best = q + (uint32_t)j;
              }
            }
          }
        }
      }
      if (wsp == 0) break;
      --wsp;
      cur = __builtin_amdgcn_readfirstlane(wnode[wv][wsp]);
      mask = wmask[wv][wsp];
      have = false;
    }
  } else if constexpr (kLanePath) {
    occ = sph_traverse<kAny, kTraceBlock>(nodes, prims, sorted_index, max_nodes, nprims, sub, s,
                                          prev, prevT, prevN, tlimit, stack + threadIdx.x, t,
                                          best);
  }
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

// kSmall (node indices below 2^24): the stack in 16-bit planes, 6 bytes an
// entry, so that the LDS holds 4 wavefronts per SIMD; the register cap of 4
// wavefronts per SIMD (128 VGPRs, the closest-hit kernel spills 4) lets
// them run (1024^2 rays on the 64^3 tetra grid: 0.53 -> 0.47 ms at 3 -> 4).
template <bool kAny, bool kSmall>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_trace_tet(const float4 *__restrict__ nodes, const float4 *__restrict__ prims,
            const uint4 *__restrict__ sorted_faces, const BuildParams *__restrict__ params,
            const srf_ray *__restrict__ rays, uint32_t n, const float *__restrict__ scalars,
            srf_hit *__restrict__ hits, srf_tet_hit_ext *__restrict__ ext,
            const float *__restrict__ tmax_in, uint32_t *__restrict__ occluded,
            uint32_t max_nodes, uint32_t nprims) {
  // Closest hit in two passes over the same tree. Pass A visits the nearer
  // child first by entry distance and finds the closest crossing's distance
  // tA with little work. Pass B then runs the reference's own traversal
  // (near child by split axis, box test at pop, TestLeafNode's acceptance)
  // from a running distance T0 = tA (1 + 1e-5) + slack instead of FLT_MAX.
  // The reference's sequence of accepted tetras ends below T0; up to its
  // first acceptance below T0 it visits a superset of pass B's nodes and
  // accepts nothing pass B needs (what it accepts lies above T0, and a node
  // pass B culls has its entry beyond T0, far above float/double rounding of
  // any crossing inside it); from that acceptance on both hold the same
  // distance and make the same decisions, so hit, t, u, v, weights and tie
  // winner are the reference's. Occlusion (kAny) keeps the single pass.
  //
  // pending far children as (child, entry distance): the popped child is
  // checked against the current hit distance without reloading its parent
  // (24 KB of LDS per block holds the kernel at 3 wavefronts per SIMD).
  // kSmall keeps 6-byte entries in 16-bit planes (18 KB): pass A stores the
  // child and its entry distance rounded down to 24 bits (a pop may then
  // visit a node exact culling would skip, which pass A tolerates); pass B
  // stores the parent and the child's side and recomputes the child's exact
  // entry distance at the pop, as its order and culling must be the
  // reference's (occlusion's traversal likewise).
  __shared__ uint2 st_node[kSmall ? 1 : kTraceStack * kTraceBlock];
  __shared__ uint16_t st_p0[kSmall ? kTraceStack * kTraceBlock : 1];
  __shared__ uint16_t st_p1[kSmall ? kTraceStack * kTraceBlock : 1];
  __shared__ uint16_t st_p2[kSmall ? kTraceStack * kTraceBlock : 1];
  // kSmall closest hit: pass A runs once per wavefront over a stack of
  // (parent, side, lane mask) entries
  constexpr bool kWave = kSmall && !kAny;
  __shared__ uint32_t wv_par[kSmall ? kTraceBlock / 64 : 1][kTraceStack];
  __shared__ uint64_t wv_msk[kSmall ? kTraceBlock / 64 : 1][kTraceStack];
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
  uint2 *sn = st_node + (kSmall ? 0 : threadIdx.x);
  uint16_t *q0 = st_p0 + (kSmall ? threadIdx.x : 0), *q1 = st_p1 + (kSmall ? threadIdx.x : 0),
           *q2 = st_p2 + (kSmall ? threadIdx.x : 0);
  int sp = 0;
  // pass A's entries: child and entry distance (kSmall: rounded down)
  auto push_a = [&](uint32_t child, float e) {
    if constexpr (kSmall) {
      const uint32_t e24 = e > 0.0f ? __float_as_uint(e) >> 8 : 0u;
      q0[sp * kTraceBlock] = (uint16_t)child;
      q1[sp * kTraceBlock] = (uint16_t)((child >> 16) | ((e24 & 255u) << 8));
      q2[sp * kTraceBlock] = (uint16_t)(e24 >> 8);
    } else {
      sn[sp * kTraceBlock] = make_uint2(child, __float_as_uint(e));
    }
    sp++;
  };
  auto pop_a = [&](float &e) -> uint32_t {
    --sp;
    if constexpr (kSmall) {
      const uint32_t a = q0[sp * kTraceBlock], b = q1[sp * kTraceBlock],
                     c = q2[sp * kTraceBlock];
      e = __uint_as_float(((b >> 8) | (c << 8)) << 8);
      return a | ((b & 255u) << 16);
    } else {
      const uint2 x = sn[sp * kTraceBlock];
      e = __uint_as_float(x.y);
      return x.x;
    }
  };
  // the reference-order traversals' entries: the far child with its entry
  // distance (kSmall: its parent and side)
  auto push_b = [&](uint32_t parent, uint32_t side, uint32_t child, float e) {
    if constexpr (kSmall) {
      q0[sp * kTraceBlock] = (uint16_t)parent;
      q1[sp * kTraceBlock] = (uint16_t)((parent >> 16) | (side << 8));
    } else {
      sn[sp * kTraceBlock] = make_uint2(child, __float_as_uint(e));
    }
    sp++;
  };
  // pops entries until one whose box the ray enters no later than t
  auto pop_next = [&]() -> uint32_t {
    while (sp > 0) {
      --sp;
      if constexpr (kSmall) {
        const uint32_t b = q1[sp * kTraceBlock];
        const uint32_t parent = (uint32_t)q0[sp * kTraceBlock] | ((b & 255u) << 16);
        const float4 *nd = nodes + 4 * (size_t)parent;
        const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
        const uint4 meta = reinterpret_cast<const uint4 *>(nd)[3];
        float tm;
        if (b >> 8)
          box_span(s, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w, tm);
        else
          box_span(s, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x, tm);
        if (tm <= t) return (b >> 8) ? meta.w : meta.z;
      } else {
        const uint2 e = sn[sp * kTraceBlock];
        if (__uint_as_float(e.y) <= t) return e.x;
      }
    }
    return kDone;
  };
  // the Pluecker test of one tetra against the running distance tt; returns
  // true when it is accepted (TestLeafNode's rule), with its outputs. A
  // crossing that is not accepted leaves et >= 0; et < 0 otherwise (a tetra
  // the prefilter skips is not crossed before tt (1 + 1e-5) + tslack).
  auto test_tet = [&](uint32_t q, float tt, int &ef, double &eu, double &lu, double &et,
                      double &lt) -> bool {
    const float4 *Pp = prims + 4 * (size_t)q;
    const float4 p0 = Pp[0], p1 = Pp[1], p2 = Pp[2], p3 = Pp[3];  // one round trip
    keep_loaded(p1, p2, p3);
    et = -1.0;
    const uint32_t id = __float_as_uint(p0.w);
    if (id == prev) return false;
    {
      const float bx0 = fminf(fminf(p0.x, p1.x), fminf(p2.x, p3.x)) - pad;
      const float by0 = fminf(fminf(p0.y, p1.y), fminf(p2.y, p3.y)) - pad;
      const float bz0 = fminf(fminf(p0.z, p1.z), fminf(p2.z, p3.z)) - pad;
      const float bx1 = fmaxf(fmaxf(p0.x, p1.x), fmaxf(p2.x, p3.x)) + pad;
      const float by1 = fmaxf(fmaxf(p0.y, p1.y), fmaxf(p2.y, p3.y)) + pad;
      const float bz1 = fmaxf(fmaxf(p0.z, p1.z), fmaxf(p2.z, p3.z)) + pad;
      float tm;
      if (!box_span(s, bx0, by0, bz0, bx1, by1, bz1, tm)) return false;
      if (tm > tt + 1e-5f * tt + tslack) return false;
    }
    const D3 vtx[4] = {{p0.x, p0.y, p0.z}, {p1.x, p1.y, p1.z}, {p2.x, p2.y, p2.z},
                       {p3.x, p3.y, p3.z}};
    et = 0.0;
    lt = 0.0;
    if (!ray_tet(ro, rd, plRay, vtx, ef, eu, lu, et, lt)) {
      et = -1.0;
      return false;
    }
    return (et >= 0.0) && (et < (double)tt);
  };
  // root box (TetraAccel::Traverse pops node 0 first)
  float troot;
  const bool root_hit = box_span(s, P.root_bmin[0], P.root_bmin[1], P.root_bmin[2],
                                 P.root_bmax[0], P.root_bmax[1], P.root_bmax[2], troot);
  // pass A: the closest crossing's distance tA with little work, visiting
  // nearer children first. s2 is the nearest other crossing seen, ec the
  // nearest entry of a box left out because it lies beyond the running
  // distance; any visiting order gives a valid tA, s2 and ec.
  float tA = FLT_MAX, s2 = FLT_MAX, ec = FLT_MAX;
  bool dropped = false;
  if constexpr (kWave) {
    // One traversal per wavefront: a lane takes part in a node while the
    // node's box is entered no later than its own tA; a node's children go
    // in the order most of its lanes would take (nearer entry first). Nodes
    // and tetras are loaded once per wavefront. At a pop every lane of the
    // entry recomputes the child's entry distance from the parent node.
    const int wv = threadIdx.x >> 6;
    const uint64_t me = 1ull << (threadIdx.x & 63u);
    uint64_t mask = __ballot(root_hit && troot <= t);
    uint32_t cur = mask ? 0u : kDone;
    int wsp = 0;
    while (cur != kDone) {
      while (cur < max_nodes) {
        cur = __builtin_amdgcn_readfirstlane(cur);
        const float4 *nd = nodes + 4 * (size_t)cur;
        const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
        const uint4 meta = reinterpret_cast<const uint4 *>(nd)[3];
        keep_loaded(n0, n1, n2);
        const bool in = (mask & me) != 0;
        if (meta.x != 0) {
          const uint32_t q0 = __builtin_amdgcn_readfirstlane(meta.w);
          const uint32_t q1 = __builtin_amdgcn_readfirstlane(min(meta.w + meta.z, nprims));
          for (uint32_t q = q0; q < q1; q++) {
            if (!in) continue;
            int ef;
            double eu, lu, et, lt;
            if (test_tet(q, tA, ef, eu, lu, et, lt)) {
              s2 = fminf(s2, tA);
              tA = (float)et;
              u = (float)eu;
              v = (float)lu;
              best = q;
              subface = (uint32_t)ef;
              tleave = lt;
            } else if (et >= 0.0) {
              s2 = fminf(s2, (float)et);
            }
          }
          break;
        }
        float tm0 = 0.0f, tm1 = 0.0f;
        const bool b0 = in && box_span(s, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x, tm0);
        const bool b1 = in && box_span(s, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w, tm1);
        const bool g0 = b0 && tm0 <= tA, g1 = b1 && tm1 <= tA;
        ec = fminf(ec, fminf(b0 && !g0 ? tm0 : FLT_MAX, b1 && !g1 ? tm1 : FLT_MAX));
        const uint64_t m0 = __ballot(g0), m1 = __ballot(g1);
        const uint64_t pref0 = __ballot(g0 && g1 && tm0 <= tm1);
        const int nboth = __popcll(m0 & m1);
        const bool first0 = m1 == 0 || (m0 != 0 && 2 * __popcll(pref0) >= nboth);
        const uint64_t mf = first0 ? m0 : m1, ms = first0 ? m1 : m0;
        if (mf == 0) break;
        if (ms != 0) {
          if (wsp < kTraceStack) {
            wv_par[wv][wsp] = cur | ((first0 ? 1u : 0u) << 31);
            wv_msk[wv][wsp] = ms;
            wsp++;
          } else if (ms & me) {
            dropped = true;
          }
        }
        cur = first0 ? meta.z : meta.w;
        mask = mf;
      }
      cur = kDone;
      while (wsp > 0) {
        --wsp;
        const uint32_t e = __builtin_amdgcn_readfirstlane(wv_par[wv][wsp]);
        const uint64_t em = wv_msk[wv][wsp];
        const uint32_t parent = e & 0x7fffffffu, side = e >> 31;
        const float4 *nd = nodes + 4 * (size_t)parent;
        const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
        const uint4 meta = reinterpret_cast<const uint4 *>(nd)[3];
        float tm = 0.0f;
        const bool in = (em & me) != 0;
        if (in) {
          if (side)
            box_span(s, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w, tm);
          else
            box_span(s, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x, tm);
        }
        const bool keep = in && tm <= tA;
        if (in && !keep) ec = fminf(ec, tm);
        mask = __ballot(keep);
        if (mask) {
          cur = side ? meta.w : meta.z;
          break;
        }
      }
    }
  }
  if (!kAny && root_hit && troot <= t) {
    uint32_t cur = kWave ? kDone : 0u;
    sp = 0;
    while (cur != kDone) {
      for (;;) {
        if (cur >= max_nodes) {
          cur = kDone;
          break;
        }
        const float4 *nd = nodes + 4 * (size_t)cur;
        const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
        const uint4 meta = reinterpret_cast<const uint4 *>(nd)[3];
        keep_loaded(n0, n1, n2);
        if (meta.x != 0) {
          for (uint32_t q = meta.w; q < min(meta.w + meta.z, nprims); q++) {
            int ef;
            double eu, lu, et, lt;
            if (test_tet(q, tA, ef, eu, lu, et, lt)) {
              s2 = fminf(s2, tA);
              tA = (float)et;
              u = (float)eu;
              v = (float)lu;
              best = q;
              subface = (uint32_t)ef;
              tleave = lt;
            } else if (et >= 0.0) {
              s2 = fminf(s2, (float)et);
            }
          }
          break;
        }
        float tm0, tm1;
        const bool b0 = box_span(s, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x, tm0);
        const bool b1 = box_span(s, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w, tm1);
        const bool g0 = b0 && tm0 <= tA, g1 = b1 && tm1 <= tA;
        ec = fminf(ec, fminf(b0 && !g0 ? tm0 : FLT_MAX, b1 && !g1 ? tm1 : FLT_MAX));
        if (g0 && g1) {
          const bool first0 = tm0 <= tm1;
          if (sp < kTraceStack)
            push_a(first0 ? meta.w : meta.z, first0 ? tm1 : tm0);
          else
            dropped = true;
          cur = first0 ? meta.z : meta.w;
          continue;
        }
        if (g0 || g1) {
          cur = g0 ? meta.z : meta.w;
          continue;
        }
        break;
      }
      cur = kDone;
      while (sp > 0) {
        float e;
        const uint32_t x = pop_a(e);
        if (e <= tA) {
          cur = x;
          break;
        }
        ec = fminf(ec, e);
      }
    }
    // When no other crossing and no box left out lies at or below
    // Tc = tA (1 + 2e-6) + tslack / 10, pass A's hit is the reference's:
    // every other tetra is crossed beyond Tc less rounding (a box entry's
    // float slab arithmetic is within 3 float ulps of its value, the double
    // Pluecker distance far closer), so the reference's running distance,
    // the float of an accepted crossing, stays above the hit's distance and
    // its boxes' entries until it reaches the hit's tetra, which it then
    // accepts, and nothing after it is accepted. Otherwise pass B settles
    // the result.
    const float T0 = tA + 1e-5f * tA + tslack;
    const float Tc = tA + 2e-6f * tA + 0.1f * tslack;
    const bool certified = !dropped && s2 > Tc && ec > Tc;
    if (tA < FLT_MAX && certified) t = tA;
    if (!certified) {
      best = SRF_NO_PRIM;
      u = 0.0f;
      v = 0.0f;
    }
    // pass B: the reference's traversal and acceptance rule from a running
    // distance just above tA instead of FLT_MAX. Should it accept nothing,
    // the reference's answer lies beyond T0 (below it, the reference would
    // have visited and accepted the same tetra), which only rounding far
    // outside the margin could cause; the traversal from FLT_MAX settles it.
    for (float t0 = T0; tA < FLT_MAX && !certified; t0 = FLT_MAX) {
      t = t0;
      sp = 0;
      cur = troot <= t ? 0u : kDone;
      uint4 meta;
      while (cur != kDone) {
        for (;;) {
          if (cur >= max_nodes) {
            cur = kDone;
            break;
          }
          const float4 *nd = nodes + 4 * (size_t)cur;
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
          if (gn && tn <= t) {
            if (gf && sp < kTraceStack) push_b(cur, nearc ? 0u : 1u, cf, tf);
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
        for (uint32_t q = meta.w; q < min(meta.w + meta.z, nprims); q++) {
          int ef;
          double eu, lu, et, lt;
          if (test_tet(q, t, ef, eu, lu, et, lt)) {
            t = (float)et;
            u = (float)eu;
            v = (float)lu;
            best = q;
            subface = (uint32_t)ef;
            tleave = lt;
          }
        }
        cur = pop_next();
      }
      if (best != SRF_NO_PRIM || t0 == FLT_MAX) break;
    }
    if (best == SRF_NO_PRIM) t = FLT_MAX;
  }
  if constexpr (kAny && kSmall) {
    const uint32_t sb = __builtin_amdgcn_readfirstlane(s.sbits);
    if (__all(s.sbits == sb)) {
      // Occlusion once per wavefront when the lanes' direction signs agree:
      // at the fixed search distance tlimit the nodes and tetras a lane's
      // traversal tests, and whether one of them occludes, do not depend on
      // the order, so the wavefront walks the union of its lanes' nodes (the
      // lanes' common near children first) and a lane leaves once occluded.
      const int wv = threadIdx.x >> 6;
      const uint64_t me = 1ull << (threadIdx.x & 63u);
      float tr;
      const bool rin = box_span(s, P.root_bmin[0], P.root_bmin[1], P.root_bmin[2],
                                P.root_bmax[0], P.root_bmax[1], P.root_bmax[2], tr) &&
                       tr <= t;
      uint64_t mask = __ballot(rin);
      bool occ = false;
      uint32_t cur = mask ? 0u : kDone;
      int wsp = 0;
      while (cur != kDone) {
        uint4 meta = make_uint4(0u, 0u, 0u, 0u);
        bool leaf = false;
        while (cur < max_nodes) {
          cur = __builtin_amdgcn_readfirstlane(cur);
          const float4 *nd = nodes + 4 * (size_t)cur;
          const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
          meta = reinterpret_cast<const uint4 *>(nd)[3];
          keep_loaded(n0, n1, n2);
          if (meta.x != 0) {
            leaf = true;
            break;
          }
          const bool in = (mask & me) != 0 && !occ;
          float tm0 = 0.0f, tm1 = 0.0f;
          const bool g0 = in && box_span(s, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x, tm0) && tm0 <= t;
          const bool g1 = in && box_span(s, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w, tm1) && tm1 <= t;
          const int nearc = (sb >> meta.y) & 1u;
          const uint64_t nm = __ballot(nearc ? g1 : g0), fm = __ballot(nearc ? g0 : g1);
          const uint32_t cn = nearc ? meta.w : meta.z, cf = nearc ? meta.z : meta.w;
          if (nm) {
            if (fm && wsp < kTraceStack) {
              wv_par[wv][wsp] = cf;
              wv_msk[wv][wsp] = fm;
              wsp++;
            }
            cur = cn;
            mask = nm;
          } else if (fm) {
            cur = cf;
            mask = fm;
          } else {
            break;
          }
        }
        if (leaf) {
          const bool in = (mask & me) != 0;
          const uint32_t q0 = __builtin_amdgcn_readfirstlane(meta.w);
          const uint32_t q1 = __builtin_amdgcn_readfirstlane(min(meta.w + meta.z, nprims));
          for (uint32_t q = q0; q < q1; q++) {
            if (!in || occ) continue;
            int ef;
            double eu, lu, et, lt;
            if (test_tet(q, t, ef, eu, lu, et, lt) && (float)et < tlimit) occ = true;
          }
        }
        if (__ballot(!occ) == 0) break;
        cur = kDone;
        if (wsp > 0) {
          --wsp;
          cur = __builtin_amdgcn_readfirstlane(wv_par[wv][wsp]);
          mask = wv_msk[wv][wsp];
        }
      }
      occluded[i] = occ ? 1u : 0u;
      return;
    }
  }
  if (kAny) {
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
        if (gf && sp < kTraceStack) push_b(cur, nearc ? 0u : 1u, cf, tf);
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

// ------------------------------------------------------------------- solids
// SolidAccel (pyramids, prisms, hexahedra), render/render_accel_solid.cc.
// TestLeafNode always returns false there (:2206-2333, its `hit` is shadowed
// inside the loop), so Traverse never narrows its search: every node whose
// own box the ray enters is visited, nearer child first, and the closest
// crossing wins (a solid's entry before its exit, strict <, first in visit
// order). A crossing's point need not lie inside the solid's box (a warped
// quad face is crossed on the plane of three of its corners), so no node or
// solid can be skipped by distance; the kernel visits exactly the
// reference's nodes in the reference's order.

// Edge i of a solid runs from vertex kSolA[i] (= i % 4 for pyramids, i % nv
// otherwise) to kSolB[i]: GetEdges, render/render_accel_solid.cc:119-154.
template <int NV>
struct SolTopo;
template <>
struct SolTopo<5> {
  static constexpr int kEdges = 8;
  static __device__ __forceinline__ int a(int i) { return i & 3; }
  static __device__ __forceinline__ int b(int i) { return i < 4 ? ((i + 1) & 3) : 4; }
};
template <>
struct SolTopo<6> {
  static constexpr int kEdges = 9;
  static __device__ __forceinline__ int a(int i) { return i < 6 ? i : i - 6; }
  static __device__ __forceinline__ int b(int i) {
    return i < 3 ? (i + 1) % 3 : (i < 6 ? 3 + (i - 2) % 3 : i - 3);
  }
};
template <>
struct SolTopo<8> {
  static constexpr int kEdges = 12;
  static __device__ __forceinline__ int a(int i) { return i < 8 ? i : i - 8; }
  static __device__ __forceinline__ int b(int i) {
    return i < 4 ? (i + 1) & 3 : (i < 8 ? 4 + ((i + 1) & 3) : i - 4);
  }
};

__device__ __forceinline__ D3 f4d(const float4 &p) { return {p.x, p.y, p.z}; }
__device__ __forceinline__ double d3c(const D3 &v, int j) {
  return j == 0 ? v.x : (j == 1 ? v.y : v.z);
}

// Face f of the crossing each pass reports, from the side masks m0 (edges
// with side >= 0) and m1 (side <= 0); -1 = none, and the solid is missed.
// The pass's own mask A and the other B play the roles of the reference's
// cw_f[i] / cw_f[n]:
//   pyramid   Intersect{Pyramid}D  :455-505 (isects[n], i = 1 - n)
//   prism     IntersectPrismD      :559-611 (isects[i], n = 1 - i)
//   hexahedron IntersectHexaD      :664-716 (isects[i], n = 1 - i)
template <int NV>
__device__ __forceinline__ int sol_face(uint32_t A, uint32_t B) {
  auto has = [](uint32_t m, uint32_t bits) { return (m & bits) == bits; };
  if (NV == 5) {
    const uint32_t Mi = B, Mn = A;
    if (has(Mi, 0xfu)) return 0;
    if (has(Mn, 0x21u) && has(Mi, 0x10u)) return 1;
    if (has(Mn, 0x42u) && has(Mi, 0x20u)) return 2;
    if (has(Mn, 0x84u) && has(Mi, 0x40u)) return 3;
    if (has(Mn, 0x18u) && has(Mi, 0x80u)) return 4;
    return -1;
  } else if (NV == 6) {
    const uint32_t Mi = A, Mn = B;
    if (has(Mn, 0x7u)) return 0;
    if (has(Mi, 0x38u)) return 1;
    if (has(Mi, 0x44u) && has(Mn, 0x120u)) return 2;
    if (has(Mi, 0x81u) && has(Mn, 0x48u)) return 3;
    if (has(Mi, 0x102u) && has(Mn, 0x90u)) return 4;
    return -1;
  } else {
    const uint32_t Mi = A, Mn = B;
    if (has(Mn, 0xfu)) return 0;
    if (has(Mi, 0xf0u)) return 1;
    if (has(Mi, 0x201u) && has(Mn, 0x110u)) return 2;
    if (has(Mi, 0x402u) && has(Mn, 0x220u)) return 3;
    if (has(Mi, 0x804u) && has(Mn, 0x440u)) return 4;
    if (has(Mi, 0x108u) && has(Mn, 0x880u)) return 5;
    return -1;
  }
}

// How face f's crossing point is formed: 0 = SetCrossPoint_Sq (:438-444,
// all double, rounded to float at the end) over corners (v0, v1, v2), side
// products (w0, w1, w2, w3) and edges (e2, e3); 1 = the pyramid's side
// triangles (v4 * -wa + vb * wb + vc * -wc) / (-wa + wb - wc); 2 = the
// prism's end triangles (va * w0 + vb * w1 + vc * w2) / (w0 + w1 + w2); for
// 1 and 2 the double numerator and denominator round to float before the
// float division (real3::operator/(real)). nrm = the edges whose cross
// product is the face normal. Negative w entries mean -ws[~w].
struct SolFace {
  int kind, v0, v1, v2, w0, w1, w2, w3, e2, e3, na, nb;
};
template <int NV>
__device__ __forceinline__ SolFace sol_face_desc(int f) {
  if (NV == 5) {
    switch (f) {
      case 0: return {0, 0, 1, 2, 0, 1, 2, 3, 2, 3, 1, 0};
      case 1: return {1, 4, 1, 0, 0, 4, 5, 0, 0, 0, 4, 0};
      case 2: return {1, 4, 2, 1, 1, 5, 6, 0, 0, 0, 5, 1};
      case 3: return {1, 4, 3, 2, 2, 6, 7, 0, 0, 0, 6, 2};
      default: return {1, 4, 0, 3, 3, 7, 4, 0, 0, 0, 7, 3};
    }
  } else if (NV == 6) {
    switch (f) {
      case 0: return {2, 2, 0, 1, 0, 1, 2, 0, 0, 0, 0, 1};
      case 1: return {2, 5, 3, 4, 3, 4, 5, 0, 0, 0, 3, 4};
      case 2: return {0, 2, 0, 3, 2, 6, ~5, ~8, 5, 8, 2, 6};
      case 3: return {0, 0, 1, 4, 0, 7, ~3, ~6, 3, 6, 0, 7};
      default: return {0, 1, 2, 5, 1, 8, ~4, ~7, 4, 7, 1, 8};
    }
  } else {
    switch (f) {
      case 0: return {0, 0, 1, 2, 0, 1, 2, 3, 2, 3, 0, 1};
      case 1: return {0, 4, 5, 6, 4, 5, 6, 7, 6, 7, 4, 5};
      case 2: return {0, 0, 1, 5, 0, 9, ~4, ~8, 4, 8, 0, 9};
      case 3: return {0, 1, 2, 6, 1, 10, ~5, ~9, 5, 9, 1, 10};
      case 4: return {0, 2, 3, 7, 2, 11, ~6, ~10, 6, 10, 2, 11};
      default: return {0, 3, 0, 4, 3, 8, ~7, ~11, 7, 11, 3, 8};
    }
  }
}

template <int NV>
__device__ __forceinline__ D3 sol_edge(const float4 *P, int i) {
  return d3sub(f4d(P[SolTopo<NV>::b(i)]), f4d(P[SolTopo<NV>::a(i)]));
}
__device__ __forceinline__ double sol_w(const double *ws, int w) { return w < 0 ? -ws[~w] : ws[w]; }

// Coordinates of face F's crossing point: component j only (mask 1 << j)
// during the search, all three for the reported hit.
template <int NV>
__device__ __forceinline__ void sol_point_of(const float4 *P, const double *ws, const D3 &rd,
                                             const SolFace F, int jmask, float out[3]) {
  const D3 v0 = f4d(P[F.v0]), v1 = f4d(P[F.v1]), v2 = f4d(P[F.v2]);
  const double w0 = sol_w(ws, F.w0), w1 = sol_w(ws, F.w1), w2 = sol_w(ws, F.w2);
  if (F.kind == 0) {
    const double w3 = sol_w(ws, F.w3);
    const double w = (w2 + w3) + d3dot(d3cross(sol_edge<NV>(P, F.e2), sol_edge<NV>(P, F.e3)), rd);
    const double den = (w0 + w1) + w;
    for (int j = 0; j < 3; j++)
      if (jmask & (1 << j))
        out[j] = (float)((((d3c(v2, j) * w0) + (d3c(v0, j) * w1)) + (d3c(v1, j) * w)) / den);
  } else if (F.kind == 1) {
    // (v4 * -wa + vb * wb + vc * -wc) / (-wa + wb - wc) with v4 = v0 here
    const float den = (float)(((-w0) + w1) - w2);
    for (int j = 0; j < 3; j++)
      if (jmask & (1 << j))
        out[j] = (float)(((d3c(v0, j) * (-w0)) + (d3c(v1, j) * w1)) + (d3c(v2, j) * (-w2))) / den;
  } else {
    const float den = (float)((w0 + w1) + w2);
    for (int j = 0; j < 3; j++)
      if (jmask & (1 << j))
        out[j] = (float)(((d3c(v0, j) * w0) + (d3c(v1, j) * w1)) + (d3c(v2, j) * w2)) / den;
  }
}

// Face f's point / normal edges with every index a compile-time constant
// (one case per face), so the side products and vertices stay in registers.
template <int NV>
__device__ __forceinline__ void sol_point(const float4 *P, const double *ws, const D3 &rd, int f,
                                          int jmask, float out[3]) {
  switch (f) {
    case 0: sol_point_of<NV>(P, ws, rd, sol_face_desc<NV>(0), jmask, out); break;
    case 1: sol_point_of<NV>(P, ws, rd, sol_face_desc<NV>(1), jmask, out); break;
    case 2: sol_point_of<NV>(P, ws, rd, sol_face_desc<NV>(2), jmask, out); break;
    case 3: sol_point_of<NV>(P, ws, rd, sol_face_desc<NV>(3), jmask, out); break;
    case 4: sol_point_of<NV>(P, ws, rd, sol_face_desc<NV>(4), jmask, out); break;
    default: sol_point_of<NV>(P, ws, rd, sol_face_desc<NV>(5), jmask, out); break;
  }
}
template <int NV>
__device__ __forceinline__ D3 sol_normal_of(const float4 *P, const SolFace F) {
  return d3cross(sol_edge<NV>(P, F.na), sol_edge<NV>(P, F.nb));
}
template <int NV>
__device__ __forceinline__ D3 sol_normal(const float4 *P, int f) {
  switch (f) {
    case 0: return sol_normal_of<NV>(P, sol_face_desc<NV>(0));
    case 1: return sol_normal_of<NV>(P, sol_face_desc<NV>(1));
    case 2: return sol_normal_of<NV>(P, sol_face_desc<NV>(2));
    case 3: return sol_normal_of<NV>(P, sol_face_desc<NV>(3));
    case 4: return sol_normal_of<NV>(P, sol_face_desc<NV>(4));
    default: return sol_normal_of<NV>(P, sol_face_desc<NV>(5));
  }
}

// Side products of the ray with every edge: dot(d, e x v_a) + dot(e, d x o),
// render/render_accel_solid.cc:465-474 (and the prism / hexahedron copies).
template <int NV>
__device__ __forceinline__ void sol_sides(const float4 *P, const D3 &rd, const D3 &raypc,
                                          double *ws, uint32_t &m0, uint32_t &m1) {
  m0 = m1 = 0;
#pragma unroll
  for (int i = 0; i < SolTopo<NV>::kEdges; i++) {
    const D3 e = sol_edge<NV>(P, i);
    ws[i] = d3dot(rd, d3cross(e, f4d(P[SolTopo<NV>::a(i)]))) + d3dot(e, raypc);
    m0 |= (ws[i] >= 0 ? 1u : 0u) << i;
    m1 |= (ws[i] <= 0 ? 1u : 0u) << i;
  }
}

// The float real3 side of interpolate() (mode 0), render/render_accel_solid.cc:
// 247-435: vertex 0 moved to the origin, then per vertex the ray from it
// through p and the face it leaves by. The reference never initialises
// the accumulator of its tolerance kEPS; it starts from 0 here.
__device__ __forceinline__ F3 sol_cp_sq(F3 v0, F3 v1, F3 v2, double ws0, double ws1, double ws2,
                                        double ws3, F3 e2, F3 e3, F3 rd) {
  // getCP_sq, render/render_accel_solid.cc:238-244
  const double w = (ws2 + ws3) + (double)f3dot(f3cross(e2, e3), rd);
  const F3 num = f3add(f3add(f3mul(v2, (float)ws0), f3mul(v0, (float)ws1)), f3mul(v1, (float)w));
  return f3div(num, (float)((ws0 + ws1) + w));
}
__device__ __forceinline__ F3 sol_cp_pyr(F3 vx, F3 vy, F3 vz, float wa, float wb, float wc) {
  return f3div(f3add(f3add(f3mul(vx, -wa), f3mul(vy, wb)), f3mul(vz, -wc)), ((-wa) + wb) - wc);
}
__device__ __forceinline__ F3 sol_cp_tri(F3 va, F3 vb, F3 vc, float w0, float w1, float w2) {
  return f3div(f3add(f3add(f3mul(va, w0), f3mul(vb, w1)), f3mul(vc, w2)), (w0 + w1) + w2);
}
__device__ __forceinline__ bool has4(int v, int a, int b, int c, int d) {
  return v == a || v == b || v == c || v == d;
}

template <int NV>
__device__ void sol_interpolate(float d_out[8], const float p[3], const float4 *P) {
  constexpr int NE = SolTopo<NV>::kEdges;
  F3 V[NV], E[NE], EPC[NE];
  for (int i = 0; i < NV; i++) V[i] = {P[i].x, P[i].y, P[i].z};
  F3 pt = f3sub(F3{p[0], p[1], p[2]}, V[0]);
  for (int i = NV - 1; i >= 0; i--) V[i] = f3sub(V[i], V[0]);
  float keps = 0.0f;
  for (int i = 0; i < NE; i++) {
    E[i] = f3sub(V[SolTopo<NV>::b(i)], V[SolTopo<NV>::a(i)]);
    keps += f3len(E[i]);
  }
  keps = (float)(((double)keps * 1e-6) / NE);
  for (int i = 0; i < NE; i++) EPC[i] = f3cross(E[i], V[SolTopo<NV>::a(i)]);
  for (int v = 0; v < NV; v++) {
    const F3 ro = V[v], rd = f3sub(pt, ro), rpc = f3cross(rd, ro);
    const float d = f3len(f3sub(pt, ro));
    float w[NE];
    uint32_t c0 = 0, c1 = 0;
    for (int i = 0; i < NE; i++) {
      w[i] = f3dot(rd, EPC[i]) + f3dot(E[i], rpc);
      c0 |= (w[i] >= -keps ? 1u : 0u) << i;
      c1 |= (w[i] <= keps ? 1u : 0u) << i;
    }
    auto all = [](uint32_t m, uint32_t bits) { return (m & bits) == bits; };
    F3 cp;
    bool found = true;
    if (NV == 5) {
      if (all(c1, 0xfu) && !has4(v, 0, 3, 2, 1))
        cp = sol_cp_sq(V[0], V[1], V[2], w[0], w[1], w[2], w[3], E[2], E[3], rd);
      else if (all(c0, 0x21u) && all(c1, 0x10u) && !has4(v, 0, 1, 4, -1))
        cp = sol_cp_pyr(V[4], V[1], V[0], w[0], w[4], w[5]);
      else if (all(c0, 0x42u) && all(c1, 0x20u) && !has4(v, 1, 2, 4, -1))
        cp = sol_cp_pyr(V[4], V[2], V[1], w[1], w[5], w[6]);
      else if (all(c0, 0x84u) && all(c1, 0x40u) && !has4(v, 2, 3, 4, -1))
        cp = sol_cp_pyr(V[4], V[3], V[2], w[2], w[6], w[7]);
      else if (all(c0, 0x18u) && all(c1, 0x80u) && !has4(v, 0, 4, 3, -1))
        cp = sol_cp_pyr(V[4], V[0], V[3], w[3], w[7], w[4]);
      else
        found = false;
    } else if (NV == 6) {
      if (all(c0, 0x7u) && !has4(v, 0, 2, 1, -1))
        cp = sol_cp_tri(V[2], V[0], V[1], w[0], w[1], w[2]);
      else if (all(c1, 0x38u) && !has4(v, 3, 4, 5, -1))
        cp = sol_cp_tri(V[5], V[3], V[4], w[3], w[4], w[5]);
      else if (all(c1, 0x44u) && all(c0, 0x120u) && !has4(v, 0, 3, 5, 2))
        cp = sol_cp_sq(V[2], V[0], V[3], w[2], w[6], -w[5], -w[8], E[5], E[8], rd);
      else if (all(c1, 0x81u) && all(c0, 0x48u) && !has4(v, 0, 1, 4, 3))
        cp = sol_cp_sq(V[0], V[1], V[4], w[0], w[7], -w[3], -w[6], E[3], E[6], rd);
      else if (all(c1, 0x102u) && all(c0, 0x90u) && !has4(v, 1, 2, 5, 4))
        cp = sol_cp_sq(V[1], V[2], V[5], w[1], w[8], -w[4], -w[7], E[4], E[7], rd);
      else
        found = false;
    } else {
      if (all(c0, 0xfu) && !has4(v, 0, 3, 2, 1))
        cp = sol_cp_sq(V[0], V[1], V[2], w[0], w[1], w[2], w[3], E[2], E[3], rd);
      else if (all(c1, 0xf0u) && !has4(v, 4, 5, 6, 7))
        cp = sol_cp_sq(V[4], V[5], V[6], w[4], w[5], w[6], w[7], E[6], E[7], rd);
      else if (all(c1, 0x201u) && all(c0, 0x110u) && !has4(v, 0, 1, 5, 4))
        cp = sol_cp_sq(V[0], V[1], V[5], w[0], w[9], -w[4], -w[8], E[4], E[8], rd);
      else if (all(c1, 0x402u) && all(c0, 0x220u) && !has4(v, 1, 2, 6, 5))
        cp = sol_cp_sq(V[1], V[2], V[6], w[1], w[10], -w[5], -w[9], E[5], E[9], rd);
      else if (all(c1, 0x804u) && all(c0, 0x440u) && !has4(v, 2, 3, 7, 6))
        cp = sol_cp_sq(V[2], V[3], V[7], w[2], w[11], -w[6], -w[10], E[6], E[10], rd);
      else if (all(c1, 0x108u) && all(c0, 0x880u) && !has4(v, 0, 4, 7, 3))
        cp = sol_cp_sq(V[3], V[0], V[4], w[3], w[8], -w[7], -w[11], E[7], E[11], rd);
      else
        found = false;
    }
    if (!found) {
      d_out[v] = 0.0f;
      continue;
    }
    float dv = d / f3len(f3sub(cp, ro));
    if (dv < 0.0f)
      dv = 0.0f;
    else if (!(dv < 1.0f) && !(dv > 0.0f))
      dv = 1.0f;
    d_out[v] = 1.0f - dv;
  }
  float ds = 0.0f, wsum = 0.0f, dd[NV];
  for (int i = 0; i < NV; i++) ds += f3len(f3sub(pt, V[i]));
  for (int i = 0; i < NV; i++) dd[i] = ds / f3len(f3sub(pt, V[i]));
  for (int i = 0; i < NV; i++) {
    d_out[i] *= dd[i];
    wsum += d_out[i];
  }
  for (int i = 0; i < NV; i++) d_out[i] = d_out[i] / wsum;
  for (int i = NV; i < 8; i++) d_out[i] = 0.0f;
}

// Pyramids (142 VGPRs uncapped) run at 4 wavefronts per SIMD with 14 VGPRs
// spilled (1024^2 rays on 32^3 cells of six pyramids: 8.55 -> 8.0-8.3 ms);
// prisms and hexahedra (162 and 206 VGPRs) spill heavily under any cap and
// are slower with one, so they are left uncapped.
template <int NV, bool kAny>
__global__ void __launch_bounds__(kTraceBlock)
    __attribute__((amdgpu_waves_per_eu(NV == 5 ? 4 : 1)))
k_trace_sol(const float4 *__restrict__ nodes, const float4 *__restrict__ prims,
            const BuildParams *__restrict__ params, const srf_ray *__restrict__ rays, uint32_t n,
            const float *__restrict__ scalars, srf_hit *__restrict__ hits,
            srf_solid_hit_ext *__restrict__ ext, const float *__restrict__ tmax_in,
            uint32_t *__restrict__ occluded, uint32_t max_nodes, uint32_t nprims) {
  constexpr int NE = SolTopo<NV>::kEdges;
  __shared__ uint32_t st_node[kTraceStack * kTraceBlock];
  __shared__ uint32_t wnode[kAny ? 1 : kTraceBlock / 64][kTraceStack];
  __shared__ uint64_t wmask[kAny ? 1 : kTraceBlock / 64][kTraceStack];
  const uint32_t i = xcd_block() * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t flags, prev;
  const RaySetup s = setup_ray(rays, i, flags, prev);
  // t_max for occlusion: the closest crossing is < t_max iff some crossing is
  const float tlimit = kAny ? fminf(tmax_in ? tmax_in[i] : FLT_MAX, FLT_MAX) : FLT_MAX;
  float t = FLT_MAX;
  uint32_t best = SRF_NO_PRIM, side = 0;
  const D3 ro = {s.o[0], s.o[1], s.o[2]}, rd = {s.d[0], s.d[1], s.d[2]};
  const D3 raypc = d3cross(rd, ro);
  // t comes from the crossing point along the last axis with |dir| > 0.1
  // (render/render_accel_solid.cc:496-500); with none the reference's t is
  // undefined and no solid is hit
  int axis = -1;
  for (int j = 0; j < 3; j++)
    if ((double)s.d[j] > 0.1 || (double)s.d[j] < -0.1) axis = j;
  const BuildParams &P = *params;
  uint32_t *sn = st_node + threadIdx.x;
  int sp = 0;
  uint32_t cur = kDone;
  {
    // node 0's own box is tested first (SolidAccel::Traverse pops node 0)
    float tm;
    if (box_span(s, P.root_bmin[0], P.root_bmin[1], P.root_bmin[2], P.root_bmax[0],
                 P.root_bmax[1], P.root_bmax[2], tm) &&
        tm <= FLT_MAX && axis >= 0)
      cur = 0;
  }
  const double oj = axis == 0 ? ro.x : (axis == 1 ? ro.y : ro.z);
  const double dj = axis == 0 ? rd.x : (axis == 1 ? rd.y : rd.z);
  uint4 meta;
  const uint32_t sb = __builtin_amdgcn_readfirstlane(s.sbits);
  if (!kAny && __all(s.sbits == sb)) {
    // Closest hit once per wavefront when the lanes' direction signs agree:
    // boxes are culled by entry alone, so a ray's visits do not depend on
    // its running hit and every lane's pre-order is one static order (the
    // line kernel's shared-stack walk); each lane tests its own leaves'
    // solids in its own order, the solids loaded through the scalar unit.
    const int wv = threadIdx.x >> 6;
    const uint64_t me = 1ull << (threadIdx.x & 63u);
    uint64_t mask = __ballot(cur == 0);
    uint32_t c = 0;
    int wsp = 0;
    while (mask) {
      uint4 m = make_uint4(0u, 0u, 0u, 0u);
      bool leaf = false;
      while (c < max_nodes) {
        c = __builtin_amdgcn_readfirstlane(c);
        const float4 *nd = nodes + 4 * (size_t)c;
        const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
        m = reinterpret_cast<const uint4 *>(nd)[3];
        keep_loaded_s(n0, n1, n2);
        if (m.x != 0) {
          leaf = true;
          break;
        }
        const bool in = (mask & me) != 0;
        float tm0, tm1;
        const bool g0 =
            in && box_span(s, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x, tm0) && tm0 <= FLT_MAX;
        const bool g1 =
            in && box_span(s, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w, tm1) && tm1 <= FLT_MAX;
        const uint32_t nearc = (sb >> m.y) & 1u;
        const uint64_t nm = __ballot(nearc ? g1 : g0), fm = __ballot(nearc ? g0 : g1);
        const uint32_t cn = nearc ? m.w : m.z, cf = nearc ? m.z : m.w;
        if (nm) {
          if (fm && wsp < kTraceStack) {
            wnode[wv][wsp] = cf;
            wmask[wv][wsp] = fm;
            wsp++;
          }
          c = cn;
          mask = nm;
        } else if (fm) {
          c = cf;
          mask = fm;
        } else {
          break;
        }
      }
      if (leaf && (mask & me) != 0) {
        // TestLeafNode, render/render_accel_solid.cc:2206-2333
        const uint32_t q0 = __builtin_amdgcn_readfirstlane(m.w);
        const uint32_t q1 = __builtin_amdgcn_readfirstlane(min(m.w + m.z, nprims));
        for (uint32_t q = q0; q < q1; q++) {
          float4 V[NV];
          const float4 *Pq = prims + (size_t)NV * q;
#pragma unroll
          for (int k = 0; k < NV; k++) V[k] = Pq[k];
          const uint32_t id = __float_as_uint(V[0].w);
          if (id == prev) continue;
          double ws[NE];
          uint32_t m0, m1;
          sol_sides<NV>(V, rd, raypc, ws, m0, m1);
          const int f0 = sol_face<NV>(m0, m1);
          if (f0 < 0) continue;
          const int f1 = sol_face<NV>(m1, m0);
          if (f1 < 0) continue;
          float c0[3], c1[3];
          sol_point<NV>(V, ws, rd, f0, 1 << axis, c0);
          sol_point<NV>(V, ws, rd, f1, 1 << axis, c1);
          const float tc[2] = {(float)(((double)c0[axis] - oj) / dj),
                               (float)(((double)c1[axis] - oj) / dj)};
          for (int k = 0; k < 2; k++) {
            if ((tc[k] >= 0.0f) && (tc[k] < t)) {
              t = tc[k];
              best = q;
              side = (uint32_t)k;
            }
          }
        }
      }
      if (wsp == 0) break;
      --wsp;
      c = __builtin_amdgcn_readfirstlane(wnode[wv][wsp]);
      mask = wmask[wv][wsp];
    }
    cur = kDone;
  }
  while (cur != kDone) {
    for (;;) {
      if (cur >= max_nodes) {
        cur = kDone;
        break;
      }
      const float4 *nd = nodes + 4 * (size_t)cur;
      const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
      meta = reinterpret_cast<const uint4 *>(nd)[3];
      keep_loaded(n0, n1, n2);
      if (meta.x != 0) break;
      float tm0, tm1;
      const bool g0 = box_span(s, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x, tm0) && tm0 <= FLT_MAX;
      const bool g1 = box_span(s, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w, tm1) && tm1 <= FLT_MAX;
      const int nearc = (s.sbits >> meta.y) & 1u;
      const bool gn = nearc ? g1 : g0, gf = nearc ? g0 : g1;
      const uint32_t cn = nearc ? meta.w : meta.z, cf = nearc ? meta.z : meta.w;
      if (gn) {
        if (gf && sp < kTraceStack) sn[(sp++) * kTraceBlock] = cf;
        cur = cn;
        continue;
      }
      if (gf) {
        cur = cf;
        continue;
      }
      cur = sp > 0 ? sn[(--sp) * kTraceBlock] : kDone;
      if (cur == kDone) break;
    }
    if (cur == kDone) break;
    // leaf: TestLeafNode, render/render_accel_solid.cc:2206-2333
    const uint32_t count = meta.z, first = meta.w;
    for (uint32_t q = first; q < min(first + count, nprims); q++) {
      float4 V[NV];
      const float4 *Pq = prims + (size_t)NV * q;
#pragma unroll
      for (int k = 0; k < NV; k++) V[k] = Pq[k];
      const uint32_t id = __float_as_uint(V[0].w);
      if (id == prev) continue;
      double ws[NE];
      uint32_t m0, m1;
      sol_sides<NV>(V, rd, raypc, ws, m0, m1);
      const int f0 = sol_face<NV>(m0, m1);
      if (f0 < 0) continue;
      const int f1 = sol_face<NV>(m1, m0);
      if (f1 < 0) continue;
      float c0[3], c1[3];
      sol_point<NV>(V, ws, rd, f0, 1 << axis, c0);
      sol_point<NV>(V, ws, rd, f1, 1 << axis, c1);
      const float tc[2] = {(float)(((double)c0[axis] - oj) / dj),
                           (float)(((double)c1[axis] - oj) / dj)};
      for (int k = 0; k < 2; k++) {
        if ((tc[k] >= 0.0f) && (tc[k] < t)) {
          if (kAny) {
            if (tc[k] < tlimit) {
              occluded[i] = 1u;
              return;
            }
          }
          t = tc[k];
          best = q;
          side = (uint32_t)k;
        }
      }
    }
    cur = sp > 0 ? sn[(--sp) * kTraceBlock] : kDone;
  }
  if (kAny) {
    occluded[i] = 0u;
    return;
  }
  uint32_t prim = SRF_NO_PRIM;
  if (best != SRF_NO_PRIM) {
    float4 V[NV];
    const float4 *Pq = prims + (size_t)NV * best;
    for (int k = 0; k < NV; k++) V[k] = Pq[k];
    prim = __float_as_uint(V[0].w);
    if (ext) {
      // the reported crossing again, in full: point, normal, weights
      double ws[NE];
      uint32_t m0, m1;
      sol_sides<NV>(V, rd, raypc, ws, m0, m1);
      const int f = side ? sol_face<NV>(m1, m0) : sol_face<NV>(m0, m1);
      srf_solid_hit_ext e = {};
      if (f >= 0) {
        sol_point<NV>(V, ws, rd, f, 7, e.position);
        // toreal3(normalize(vcrossd(ea, eb))), render/render_accel_solid.cc:69-85
        const D3 c = sol_normal<NV>(V, f);
        const double len = sqrt((c.x * c.x + c.y * c.y) + c.z * c.z);
        const float sg = side ? -1.0f : 1.0f;
        e.normal[0] = sg * (float)(c.x / len);
        e.normal[1] = sg * (float)(c.y / len);
        e.normal[2] = sg * (float)(c.z / len);
        sol_interpolate<NV>(e.d, e.position, V);
        e.side = side;
        if (scalars) {
          // f_k = nv * prim + k (:2284-2291), summed left to right
          // (gles/gles_common.cc:1078-1140)
          const float *sc = scalars + (size_t)NV * prim;
          float acc = sc[0] * e.d[0];
          for (int k = 1; k < NV; k++) acc = acc + sc[k] * e.d[k];
          e.scalar = acc;
        }
      }
      float4 *e4 = reinterpret_cast<float4 *>(ext + i);
      e4[0] = make_float4(e.d[0], e.d[1], e.d[2], e.d[3]);
      e4[1] = make_float4(e.d[4], e.d[5], e.d[6], e.d[7]);
      e4[2] = make_float4(e.position[0], e.position[1], e.position[2], e.scalar);
      e4[3] = make_float4(e.normal[0], e.normal[1], e.normal[2], __uint_as_float(e.side));
    }
  } else if (ext) {
    float4 *e4 = reinterpret_cast<float4 *>(ext + i);
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    e4[0] = z;
    e4[1] = z;
    e4[2] = z;
    e4[3] = z;
  }
  reinterpret_cast<float4 *>(hits)[i] = make_float4(t, 0.0f, 0.0f, __uint_as_float(prim));
}

// -------------------------------------------------------------------- lines
// LineAccel::Traverse (render/render_accel_line.cc:704-774) over the LBVH.
// Every node whose box the ray enters is visited, nearer child first, with
// no narrowing by the hit distance: the crossing a float quadratic reports
// for a grazing ray may lie outside its padded box, so the closest crossing
// over all lines is found. Running t as TestLeafNode keeps it: a cap
// replaces it only when strictly nearer, the side when not farther.

__device__ __forceinline__ float l3dot(float ax, float ay, float az, float bx, float by,
                                       float bz) {
  return (ax * bx + ay * by) + az * bz;
}
// vnormalize, render/render_accel_line.cc:439-454
__device__ __forceinline__ void l3norm(float &x, float &y, float &z) {
  const float len = sqrtf(l3dot(x, y, z, x, y, z));
  if (len > FLT_EPSILON * 16.0f) {
    const float inv = (float)(1.0 / (double)len);
    x *= inv;
    y *= inv;
    z *= inv;
  }
}

// CylinderIsect, render/render_accel_line.cc:488-577 (solve2e :456-486):
// t / u / v / cap updated in place on a hit.
__device__ __forceinline__ bool line_isect(float &t, float &u, float &v, bool &hit_cap,
                                           const float4 &P0, const float4 &P1, const RaySetup &s,
                                           bool cap) {
  const float tmax = t, rr = P1.w;
  const float dx = P1.x - P0.x, dy = P1.y - P0.y, dz = P1.z - P0.z;
  const float mx = s.o[0] - P0.x, my = s.o[1] - P0.y, mz = s.o[2] - P0.z;
  const float md = l3dot(mx, my, mz, dx, dy, dz);
  const float nd = l3dot(s.d[0], s.d[1], s.d[2], dx, dy, dz);
  const float dd = l3dot(dx, dy, dz, dx, dy, dz);
  hit_cap = false;
  float capT = FLT_MAX;
  bool hit = false;
  if (cap) {
    float n0x = P0.x - P1.x, n0y = P0.y - P1.y, n0z = P0.z - P1.z;
    l3norm(n0x, n0y, n0z);
    const float n1x = -n0x, n1y = -n0y, n1z = -n0z;
    float rx = s.d[0], ry = s.d[1], rz = s.d[2];
    l3norm(rx, ry, rz);
    if (fabsf(l3dot(s.d[0], s.d[1], s.d[2], n0x, n0y, n0z)) > 1.0e-6f) {
      const float p0D = -l3dot(P0.x, P0.y, P0.z, n0x, n0y, n0z);
      const float p1D = -l3dot(P1.x, P1.y, P1.z, n1x, n1y, n1z);
      const float p0T = -(l3dot(s.o[0], s.o[1], s.o[2], n0x, n0y, n0z) + p0D) /
                        l3dot(rx, ry, rz, n0x, n0y, n0z);
      const float p1T = -(l3dot(s.o[0], s.o[1], s.o[2], n1x, n1y, n1z) + p1D) /
                        l3dot(rx, ry, rz, n1x, n1y, n1z);
      const float e0x = (s.o[0] + p0T * rx) - P0.x, e0y = (s.o[1] + p0T * ry) - P0.y,
                  e0z = (s.o[2] + p0T * rz) - P0.z;
      const float e1x = (s.o[0] + p1T * rx) - P1.x, e1y = (s.o[1] + p1T * ry) - P1.y,
                  e1z = (s.o[2] + p1T * rz) - P1.z;
      const float q0 = l3dot(e0x, e0y, e0z, e0x, e0y, e0z);
      const float q1 = l3dot(e1x, e1y, e1z, e1x, e1y, e1z);
      if (p0T > 0.0f && p0T < tmax && (q0 < rr * rr)) {
        hit_cap = hit = true;
        capT = p0T;
        t = capT;
        u = sqrtf(q0);
        v = 0.0f;
      }
      if (p1T > 0.0f && p1T < tmax && p1T < capT && (q1 < rr * rr)) {
        hit_cap = hit = true;
        capT = p1T;
        t = capT;
        u = sqrtf(q1);
        v = 1.0f;
      }
    }
  }
  if (md <= 0.0f && nd <= 0.0f) return hit;
  if (md >= dd && nd >= 0.0f) return hit;
  const float nn = l3dot(s.d[0], s.d[1], s.d[2], s.d[0], s.d[1], s.d[2]);
  const float mn = l3dot(mx, my, mz, s.d[0], s.d[1], s.d[2]);
  const float A = dd * nn - nd * nd;
  const float kk = l3dot(mx, my, mz, mx, my, mz) - rr * rr;
  const float C = dd * kk - md * md;
  const float B = dd * mn - nd * md;
  float root;
  bool ok = true;
  if (fabsf(A) <= FLT_EPSILON) {
    root = -C / B;
  } else {
    const float D = B * B - A * C;
    if (D < 0.0f) {
      ok = false;
      root = 0.0f;
    } else if (D == 0.0f) {
      root = -B / A;
    } else {
      // fabs and sqrt there are the double C library functions: the sum and
      // the division run in double
      float x1 = (float)((fabs((double)B) + sqrt((double)D)) / (double)A);
      if (B >= 0.0f) x1 = -x1;
      const float x2 = C / (A * x1);
      root = x1 > x2 ? x2 : x1;
    }
  }
  if (ok && 0.0f <= root && root <= tmax && root <= capT) {
    float sa = md + root * nd;
    sa /= dd;
    if (0.0f <= sa && sa <= 1.0f) {
      hit_cap = false;
      t = root;
      u = 0.0f;
      v = sa;
      return true;
    }
  }
  return hit;
}

template <bool kAny>
__global__ void __launch_bounds__(kTraceBlock)
k_trace_lin(const float4 *__restrict__ nodes, const float4 *__restrict__ prims,
            const BuildParams *__restrict__ params, const srf_ray *__restrict__ rays, uint32_t n,
            srf_hit *__restrict__ hits, srf_line_hit_ext *__restrict__ ext,
            const float *__restrict__ tmax_in, uint32_t *__restrict__ occluded, uint32_t max_nodes,
            uint32_t nprims, uint32_t cap) {
  __shared__ uint32_t st_node[kTraceStack * kTraceBlock];
  __shared__ uint32_t wnode[kAny ? 1 : kTraceBlock / 64][kTraceStack];
  __shared__ uint64_t wmask[kAny ? 1 : kTraceBlock / 64][kTraceStack];
  const uint32_t i = xcd_block() * kTraceBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t flags, prev;
  const RaySetup s = setup_ray(rays, i, flags, prev);
  const float tlimit = kAny ? fminf(tmax_in ? tmax_in[i] : FLT_MAX, FLT_MAX) : FLT_MAX;
  float t = FLT_MAX, u = 0.0f, v = 0.0f;
  uint32_t best = SRF_NO_PRIM, best_q = 0;
  bool best_cap = false;
  const BuildParams &P = *params;
  uint32_t *sn = st_node + threadIdx.x;
  int sp = 0;
  uint32_t cur = kDone;
  {
    float tm;
    if (box_span(s, P.root_bmin[0], P.root_bmin[1], P.root_bmin[2], P.root_bmax[0],
                 P.root_bmax[1], P.root_bmax[2], tm) &&
        tm <= FLT_MAX)
      cur = 0;
  }
  uint4 meta;
  const uint32_t sb = __builtin_amdgcn_readfirstlane(s.sbits);
  if (!kAny && __all(s.sbits == sb)) {
    // Closest hit once per wavefront when the lanes' direction signs agree.
    // A node's boxes are culled by entry alone (no distance), so the nodes a
    // ray visits do not depend on its running hit, and with shared signs
    // every lane's pre-order (near child by the split axis' sign, far child
    // pushed when both boxes are entered) is one static order: the
    // wavefront walks the union of the lanes' visits in that order with one
    // stack of (node, lane mask), and each lane tests its own leaves' lines
    // in its own order, as in the per-ray loop below.
    const int wv = threadIdx.x >> 6;
    const uint64_t me = 1ull << (threadIdx.x & 63u);
    uint64_t mask = __ballot(cur == 0);
    uint32_t c = 0;
    int wsp = 0;
    while (mask) {
      uint4 m = make_uint4(0u, 0u, 0u, 0u);
      bool leaf = false;
      while (c < max_nodes) {
        c = __builtin_amdgcn_readfirstlane(c);
        const float4 *nd = nodes + 4 * (size_t)c;
        const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
        m = reinterpret_cast<const uint4 *>(nd)[3];
        keep_loaded_s(n0, n1, n2);
        if (m.x != 0) {
          leaf = true;
          break;
        }
        const bool in = (mask & me) != 0;
        float tm0, tm1;
        const bool g0 =
            in && box_span(s, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x, tm0) && tm0 <= FLT_MAX;
        const bool g1 =
            in && box_span(s, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w, tm1) && tm1 <= FLT_MAX;
        const uint32_t nearc = (sb >> m.y) & 1u;
        const uint64_t nm = __ballot(nearc ? g1 : g0), fm = __ballot(nearc ? g0 : g1);
        const uint32_t cn = nearc ? m.w : m.z, cf = nearc ? m.z : m.w;
        if (nm) {
          if (fm && wsp < kTraceStack) {
            wnode[wv][wsp] = cf;
            wmask[wv][wsp] = fm;
            wsp++;
          }
          c = cn;
          mask = nm;
        } else if (fm) {
          c = cf;
          mask = fm;
        } else {
          break;
        }
      }
      if (leaf) {
        // TestLeafNode, render/render_accel_line.cc:579-633
        const bool in = (mask & me) != 0;
        const uint32_t q0 = __builtin_amdgcn_readfirstlane(m.w);
        const uint32_t q1 = __builtin_amdgcn_readfirstlane(min(m.w + m.z, nprims));
        for (uint32_t q = q0; q < q1; q++) {
          const float4 P0 = prims[2 * (size_t)q], P1 = prims[2 * (size_t)q + 1];
          keep_loaded_s2(P0, P1);
          if (!in) continue;
          bool hc;
          if (line_isect(t, u, v, hc, P0, P1, s, cap != 0)) {
            best = __float_as_uint(P0.w);
            best_q = q;
            best_cap = hc;
          }
        }
      }
      if (wsp == 0) break;
      --wsp;
      c = __builtin_amdgcn_readfirstlane(wnode[wv][wsp]);
      mask = wmask[wv][wsp];
    }
    cur = kDone;
  }
  while (cur != kDone) {
    for (;;) {
      if (cur >= max_nodes) {
        cur = kDone;
        break;
      }
      const float4 *nd = nodes + 4 * (size_t)cur;
      const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
      meta = reinterpret_cast<const uint4 *>(nd)[3];
      keep_loaded(n0, n1, n2);
      if (meta.x != 0) break;
      float tm0, tm1;
      const bool g0 = box_span(s, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x, tm0) && tm0 <= FLT_MAX;
      const bool g1 = box_span(s, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w, tm1) && tm1 <= FLT_MAX;
      const int nearc = (s.sbits >> meta.y) & 1u;
      const bool gn = nearc ? g1 : g0, gf = nearc ? g0 : g1;
      const uint32_t cn = nearc ? meta.w : meta.z, cf = nearc ? meta.z : meta.w;
      if (gn) {
        if (gf && sp < kTraceStack) sn[(sp++) * kTraceBlock] = cf;
        cur = cn;
        continue;
      }
      if (gf) {
        cur = cf;
        continue;
      }
      cur = sp > 0 ? sn[(--sp) * kTraceBlock] : kDone;
      if (cur == kDone) break;
    }
    if (cur == kDone) break;
    // leaf: TestLeafNode, render/render_accel_line.cc:579-633
    const uint32_t count = meta.z, first = meta.w;
    for (uint32_t q = first; q < min(first + count, nprims); q++) {
      const float4 P0 = prims[2 * (size_t)q], P1 = prims[2 * (size_t)q + 1];
      bool hc;
      if (line_isect(t, u, v, hc, P0, P1, s, cap != 0)) {
        if (kAny && t < tlimit) {
          occluded[i] = 1u;
          return;
        }
        best = __float_as_uint(P0.w);
        best_q = q;
        best_cap = hc;
      }
    }
    cur = sp > 0 ? sn[(--sp) * kTraceBlock] : kDone;
  }
  if (kAny) {
    occluded[i] = 0u;
    return;
  }
  if (best == SRF_NO_PRIM) u = v = 0.0f;
  if (ext) {
    // BuildIntersection, render/render_accel_line.cc:654-702
    float4 e0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), e1 = e0;
    if (best != SRF_NO_PRIM) {
      const float4 P0 = prims[2 * (size_t)best_q], P1 = prims[2 * (size_t)best_q + 1];
      const float px = s.o[0] + t * s.d[0], py = s.o[1] + t * s.d[1], pz = s.o[2] + t * s.d[2];
      const float dx = P1.x - P0.x, dy = P1.y - P0.y, dz = P1.z - P0.z;
      float nx, ny, nz;
      if (best_cap) {
        const float cx = 0.5f * dx + P0.x, cy = 0.5f * dy + P0.y, cz = 0.5f * dz + P0.z;
        nx = dx;
        ny = dy;
        nz = dz;
        l3norm(nx, ny, nz);
        if (!(l3dot(px - cx, py - cy, pz - cz, nx, ny, nz) > 0.0f)) {
          nx = -nx;
          ny = -ny;
          nz = -nz;
        }
      } else {
        nx = px - (P0.x + v * dx);
        ny = py - (P0.y + v * dy);
        nz = pz - (P0.z + v * dz);
        l3norm(nx, ny, nz);
      }
      e0 = make_float4(px, py, pz, __uint_as_float(best_cap ? 1u : 0u));
      e1 = make_float4(nx, ny, nz, 0.0f);
    }
    reinterpret_cast<float4 *>(ext + i)[0] = e0;
    reinterpret_cast<float4 *>(ext + i)[1] = e1;
  }
  reinterpret_cast<float4 *>(hits)[i] = make_float4(t, u, v, __uint_as_float(best));
}

// ------------------------------------------------- sorted / persistent any-hit

// Occlusion for rays taken in `order` (order[j] is the j-th ray to trace,
// the flag goes back to its own slot). With `counter` the grid is
// persistent: each wavefront takes 64 consecutive positions of the order at
// a time until none is left.
// For spheres kPart selects the 64-ray groups a launch handles: 0 all, 1
// those whose rays share direction signs (the wavefront traversal, without a
// per-lane stack, so more wavefronts fit per CU), 2 the others, one group per
// wavefront of a plain grid. Part 1 followed by part 2 gives every ray once.
template <int K, int kPart>
__global__ void __launch_bounds__(kTraceBlock)
k_any_ordered(const float4 *__restrict__ nodes, const float4 *__restrict__ prims,
              const uint32_t *__restrict__ sorted_index, const BuildParams *__restrict__ params,
              const srf_ray *__restrict__ rays, const float *__restrict__ tmax_in,
              const uint32_t *__restrict__ order, uint32_t n, uint32_t *__restrict__ occluded,
              uint32_t max_nodes, uint32_t nprims, uint32_t *counter, SphSub sub) {
  constexpr bool kLanePath = kPart != 1;
  __shared__ uint32_t stack[kLanePath ? (K == kSph ? 2 : 1) * kTraceStack * kTraceBlock : 1];
  __shared__ uint32_t wnode[K == kSph && kPart != 2 ? kTraceBlock / 64 : 1][2 * kTraceStack];
  __shared__ uint64_t wmask[K == kSph && kPart != 2 ? kTraceBlock / 64 : 1][2 * kTraceStack];
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
      bool occ = false;
      const uint32_t sb = __builtin_amdgcn_readfirstlane(s.sbits);
      const bool uniform = K == kSph && __all(s.sbits == sb);
      bool skip = false;
      if (kPart == 1 && !uniform) skip = true;  // left to part 2
      if (kPart == 2 && uniform) skip = true;   // done by part 1
      if (skip) {
      } else if (K == kTri) {
        if constexpr (kLanePath) {
          const float kEPS = FLT_EPSILON * (float)params->eps_scale;
          float t, u, v;
          uint32_t prim;
          occ = tri_traverse<true>(nodes, prims, max_nodes, nprims, s,
                                   (flags & SRF_RAY_DOUBLE_SIDED) != 0, kEPS, prev,
                                   tlimit, Stack32<kTraceBlock>{st}, t, u, v, prim);
        }
      } else if (!uniform) {
        // directions of mixed signs (rays not sorted): one traversal per lane
        if constexpr (kLanePath) {
          const float zero[3] = {0.0f, 0.0f, 0.0f};
          float t;
          uint32_t best;
          occ = sph_traverse<true, kTraceBlock>(nodes, prims, sorted_index, max_nodes, nprims,
                                                sub, s, prev, 0.0, zero, tlimit, st, t, best);
        }
      } else if constexpr (kPart != 2) {
        // Once per wavefront: with the search distance fixed at tlimit, the
        // set of nodes and spheres a lane's traversal tests does not depend
        // on their order, and neither does whether one of them occludes, so
        // the wavefront walks the union of its lanes' nodes in one order (the
        // lanes' common near children) and a lane leaves once occluded.
        const float zero[3] = {0.0f, 0.0f, 0.0f};
        float t = tlimit;
        uint32_t best = SRF_NO_PRIM;
        const int wv = threadIdx.x >> 6;
        const uint64_t me = 1ull << lane;
        occ = false;
        uint64_t mask = __ballot(1);
        uint32_t cur = 0;
        int wsp = 0;
        for (;;) {
          uint4 meta = make_uint4(0u, 0u, 0u, 0u);
          bool leaf = false;
          while (cur < max_nodes) {
            cur = __builtin_amdgcn_readfirstlane(cur);
            const float4 *nd = nodes + 4 * (size_t)cur;
            const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2];
            meta = reinterpret_cast<const uint4 *>(nd)[3];
            keep_loaded_s(n0, n1, n2);
            if (meta.x != 0) {
              leaf = true;
              break;
            }
            const bool in = (mask & me) != 0 && !occ;
            const bool h0 =
                in && ray_aabb_wave<true>(s, sb, t, n0.x, n0.y, n0.z, n1.z, n1.w, n2.x);
            const bool h1 =
                in && ray_aabb_wave<true>(s, sb, t, n0.w, n1.x, n1.y, n2.y, n2.z, n2.w);
            const int nearc = (sb >> meta.y) & 1u;
            const uint64_t nm = __ballot(nearc ? h1 : h0), fm = __ballot(nearc ? h0 : h1);
            const uint32_t cn = nearc ? meta.w : meta.z, cf = nearc ? meta.z : meta.w;
            if (nm) {
              if (fm && wsp < kTraceStack) {
                wnode[wv][wsp] = cf;
                wmask[wv][wsp] = fm;
                wsp++;
              }
              cur = cn;
              mask = nm;
            } else if (fm) {
              cur = cf;
              mask = fm;
            } else {
              break;
            }
          }
          if (leaf) {
            const bool in = (mask & me) != 0 && !occ;
            if (meta.z > kBigLeaf && sub.nodes) {
              const float4 *nd = nodes + 4 * (size_t)cur;
              const uint32_t nbase = __builtin_amdgcn_readfirstlane(__float_as_uint(nd[0].w));
              const uint32_t pbase = __builtin_amdgcn_readfirstlane(__float_as_uint(nd[1].x));
              if (meta.w < nprims && meta.z <= nprims - meta.w)
                sph_big_leaf_wave<true>(sub, nbase, pbase, meta.w, meta.z, sorted_index, s, sb,
                                        prev, 0.0, zero, tlimit, in, &wnode[wv][kTraceStack],
                                        &wmask[wv][kTraceStack], t, best, occ);
            } else {
              const uint32_t q0 = __builtin_amdgcn_readfirstlane(meta.w);
              const uint32_t q1 = __builtin_amdgcn_readfirstlane(min(meta.w + meta.z, nprims));
              for (uint32_t q = q0; q < q1; q++) {
                const float4 c = prims[q];
                const uint32_t sid = sorted_index[q];
                if (!in || occ) continue;
                const bool self = (prev != SRF_NO_PRIM) && (sid == prev);
                float tt;
                if (sphere_isect(s, c, t, 0.0, zero, self, tlimit, tt) && tt < tlimit) occ = true;
              }
            }
          }
          if (__ballot(!occ) == 0) break;
          if (wsp == 0) break;
          --wsp;
          cur = __builtin_amdgcn_readfirstlane(wnode[wv][wsp]);
          mask = wmask[wv][wsp];
        }
      }
      if (!skip) occluded[i] = occ ? 1u : 0u;
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
inline bool small_nodes(const Accel *a) { return !a->wide_stack && 2 * a->n - 1 <= (1u << 24); }
// every node index fits the tetra kernel's 24-bit entries
inline bool small_tet(const Accel *a) { return !a->wide_stack && max_nodes(a) < (1u << 24); }
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
    hipLaunchKernelGGL((k_trace_tri_wave<true>), dim3(trace_blocks(n)), dim3(kTraceBlock), 0, s,
                       (const float4 *)a->nodes, (const float4 *)a->prims, a->params, rays,
                       (uint32_t)n, hits, max_nodes(a), (uint32_t)a->n);
  else
    hipLaunchKernelGGL((k_trace_tri_wave<false>), dim3(trace_blocks(n)), dim3(kTraceBlock), 0,
                       s, (const float4 *)a->nodes, (const float4 *)a->prims, a->params, rays,
                       (uint32_t)n, hits, max_nodes(a), (uint32_t)a->n);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

srf_status trace_tet(const Accel *a, const srf_ray *rays, uint64_t n, const float *scalars,
                     srf_hit *hits, srf_tet_hit_ext *ext, hipStream_t s) {
  if (n == 0) return SRF_OK;
  if (small_tet(a))
    hipLaunchKernelGGL((k_trace_tet<false, true>), dim3(trace_blocks(n)), dim3(kTraceBlock), 0,
                       s, (const float4 *)a->nodes, (const float4 *)a->prims,
                       (const uint4 *)a->sorted_faces, a->params, rays, (uint32_t)n, scalars,
                       hits, ext, nullptr, nullptr, max_nodes(a), (uint32_t)a->n);
  else
    hipLaunchKernelGGL((k_trace_tet<false, false>), dim3(trace_blocks(n)), dim3(kTraceBlock), 0,
                       s, (const float4 *)a->nodes, (const float4 *)a->prims,
                       (const uint4 *)a->sorted_faces, a->params, rays, (uint32_t)n, scalars,
                       hits, ext, nullptr, nullptr, max_nodes(a), (uint32_t)a->n);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

template <bool kAny>
srf_status launch_sol(const Accel *a, const srf_ray *rays, uint64_t n, const float *scalars,
                      srf_hit *hits, srf_solid_hit_ext *ext, const float *tmax,
                      uint32_t *occluded, hipStream_t s) {
  const dim3 g(trace_blocks(n)), b(kTraceBlock);
  const float4 *nd = (const float4 *)a->nodes, *pr = (const float4 *)a->prims;
  if (a->nvps == 5)
    hipLaunchKernelGGL((k_trace_sol<5, kAny>), g, b, 0, s, nd, pr, a->params, rays, (uint32_t)n,
                       scalars, hits, ext, tmax, occluded, max_nodes(a), (uint32_t)a->n);
  else if (a->nvps == 6)
    hipLaunchKernelGGL((k_trace_sol<6, kAny>), g, b, 0, s, nd, pr, a->params, rays, (uint32_t)n,
                       scalars, hits, ext, tmax, occluded, max_nodes(a), (uint32_t)a->n);
  else if (a->nvps == 8)
    hipLaunchKernelGGL((k_trace_sol<8, kAny>), g, b, 0, s, nd, pr, a->params, rays, (uint32_t)n,
                       scalars, hits, ext, tmax, occluded, max_nodes(a), (uint32_t)a->n);
  else
    return set_error(SRF_ERR_INTERNAL, "solid accelerator with %u vertices per solid", a->nvps);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

srf_status trace_lin(const Accel *a, const srf_ray *rays, uint64_t n, srf_hit *hits,
                     srf_line_hit_ext *ext, hipStream_t s) {
  if (n == 0) return SRF_OK;
  hipLaunchKernelGGL(k_trace_lin<false>, dim3(trace_blocks(n)), dim3(kTraceBlock), 0, s,
                     (const float4 *)a->nodes, (const float4 *)a->prims, a->params, rays,
                     (uint32_t)n, hits, ext, nullptr, nullptr, max_nodes(a), (uint32_t)a->n,
                     a->line_cap);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

srf_status trace_sol(const Accel *a, const srf_ray *rays, uint64_t n, const float *scalars,
                     srf_hit *hits, srf_solid_hit_ext *ext, hipStream_t s) {
  if (n == 0) return SRF_OK;
  return launch_sol<false>(a, rays, n, scalars, hits, ext, nullptr, nullptr, s);
}

srf_status trace_sph(const Accel *a, const srf_ray *rays, const srf_sphere_prev *prev,
                     uint64_t n, srf_hit *hits, hipStream_t s) {
  if (n == 0) return SRF_OK;
  // the wavefront traversal for wavefronts of equal direction signs, then
  // the per-lane one for the rest
  hipLaunchKernelGGL((k_trace_sph<false, 1>), dim3(trace_blocks(n)), dim3(kTraceBlock), 0, s,
                     (const float4 *)a->nodes, (const float4 *)a->prims, a->sorted_index, rays,
                     prev, (uint32_t)n, hits, nullptr, nullptr, max_nodes(a), (uint32_t)a->n,
                     sph_sub(a));
  SRF_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_trace_sph<false, 2>), dim3(trace_blocks(n)), dim3(kTraceBlock), 0, s,
                     (const float4 *)a->nodes, (const float4 *)a->prims, a->sorted_index, rays,
                     prev, (uint32_t)n, hits, nullptr, nullptr, max_nodes(a), (uint32_t)a->n,
                     sph_sub(a));
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

srf_status trace_any(const Accel *a, const srf_ray *rays, const float *tmax, uint64_t n,
                     uint32_t *occluded, hipStream_t s) {
  if (n == 0) return SRF_OK;
  if (a->kind == kSol) return launch_sol<true>(a, rays, n, nullptr, nullptr, nullptr, tmax, occluded, s);
  if (a->kind == kLin) {
    hipLaunchKernelGGL(k_trace_lin<true>, dim3(trace_blocks(n)), dim3(kTraceBlock), 0, s,
                       (const float4 *)a->nodes, (const float4 *)a->prims, a->params, rays,
                       (uint32_t)n, nullptr, nullptr, tmax, occluded, max_nodes(a),
                       (uint32_t)a->n, a->line_cap);
    SRF_HIP(hipGetLastError());
    return SRF_OK;
  }
  const dim3 g(trace_blocks(n)), b(kTraceBlock);
  if (a->kind == kTri && small_nodes(a))
    hipLaunchKernelGGL((k_trace_tri<true, true>), g, b, 0, s, (const float4 *)a->nodes,
                       (const float4 *)a->prims, a->params, rays, (uint32_t)n, nullptr, tmax,
                       occluded, max_nodes(a), (uint32_t)a->n);
  else if (a->kind == kTri)
    hipLaunchKernelGGL((k_trace_tri<true, false>), g, b, 0, s, (const float4 *)a->nodes,
                       (const float4 *)a->prims, a->params, rays, (uint32_t)n, nullptr, tmax,
                       occluded, max_nodes(a), (uint32_t)a->n);
  else if (a->kind == kTet) {
    if (small_tet(a))
      hipLaunchKernelGGL((k_trace_tet<true, true>), g, b, 0, s, (const float4 *)a->nodes,
                         (const float4 *)a->prims, (const uint4 *)a->sorted_faces, a->params,
                         rays, (uint32_t)n, nullptr, nullptr, nullptr, tmax, occluded,
                         max_nodes(a), (uint32_t)a->n);
    else
      hipLaunchKernelGGL((k_trace_tet<true, false>), g, b, 0, s, (const float4 *)a->nodes,
                         (const float4 *)a->prims, (const uint4 *)a->sorted_faces, a->params,
                         rays, (uint32_t)n, nullptr, nullptr, nullptr, tmax, occluded,
                         max_nodes(a), (uint32_t)a->n);
  } else
    hipLaunchKernelGGL((k_trace_sph<true, 0>), g, b, 0, s, (const float4 *)a->nodes,
                       (const float4 *)a->prims, a->sorted_index, rays, nullptr, (uint32_t)n,
                       nullptr, tmax, occluded, max_nodes(a), (uint32_t)a->n, sph_sub(a));
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

namespace {

// Scratch of a sorted occlusion launch: [counter][keys][sorted keys][order]
// then the sort's own slices.
size_t any_sorted_head(uint32_t n) { return 256 + 3 * (((size_t)4 * n + 255) & ~(size_t)255); }
size_t any_sorted_bytes(uint32_t n) { return any_sorted_head(n) + sort_pairs30_bytes(n); }

srf_status check_any_sorted(const Accel *a, uint32_t key_bits) {
  if (a->kind == kTet || a->kind == kSol || a->kind == kLin)
    return set_error(SRF_ERR_UNSUPPORTED, "sorted any-hit takes triangles or spheres");
  if (key_bits > 30) return set_error(SRF_ERR_INVALID_ARG, "key_bits must be <= 30");
  return SRF_OK;
}

// The sorted occlusion launch in the buffer [base, base + size) (at least
// any_sorted_bytes(n) when rays are sorted or the launch is persistent);
// allocates nothing.
srf_status trace_any_sorted_in(char *base, size_t size, const Accel *a, const srf_ray *rays,
                               const float *tmax, uint32_t nn, uint32_t key_bits, int mode,
                               uint32_t *occluded, hipStream_t s) {
  const uint32_t *order = nullptr;
  uint32_t *counter = nullptr;
  if (key_bits > 0 || mode == SRF_LAUNCH_PERSISTENT) {
    Arena ar{base, size, 0};
    counter = ar.take<uint32_t>(64);
    uint32_t *keys = ar.take<uint32_t>(nn);
    uint32_t *skeys = ar.take<uint32_t>(nn);
    uint32_t *ord = ar.take<uint32_t>(nn);
    if (key_bits > 0) {
      hipLaunchKernelGGL(k_ray_keys, dim3((nn + 255) / 256), dim3(256), 0, s, rays, nn, a->params,
                         key_bits, keys);
      SRF_HIP(hipGetLastError());
      srf_status st = sort_pairs30_in(base, size, any_sorted_head(nn), keys, nn, skeys, ord, s);
      if (st != SRF_OK) return st;
      order = ord;
    }
    if (mode != SRF_LAUNCH_PERSISTENT) counter = nullptr;
  }
  uint32_t blocks = trace_blocks(nn);
  if (counter) {
    int dev = 0, cus = 0;
    SRF_HIP(hipGetDevice(&dev));
    SRF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    blocks = std::min<uint32_t>(blocks, (uint32_t)cus * 16u);
    hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, s, counter);
  }
  if (a->kind == kTri) {
    hipLaunchKernelGGL((k_any_ordered<kTri, 0>), dim3(blocks), dim3(kTraceBlock), 0, s,
                       (const float4 *)a->nodes, (const float4 *)a->prims, a->sorted_index,
                       a->params, rays, tmax, order, nn, occluded, max_nodes(a), (uint32_t)a->n,
                       counter, SphSub{});
  } else {
    // groups of equal direction signs by the wavefront traversal, then the
    // others per lane (one group per wavefront of a plain grid)
    hipLaunchKernelGGL((k_any_ordered<kSph, 1>), dim3(blocks), dim3(kTraceBlock), 0, s,
                       (const float4 *)a->nodes, (const float4 *)a->prims, a->sorted_index,
                       a->params, rays, tmax, order, nn, occluded, max_nodes(a), (uint32_t)a->n,
                       counter, sph_sub(a));
    SRF_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_any_ordered<kSph, 2>), dim3(trace_blocks(nn)), dim3(kTraceBlock), 0, s,
                       (const float4 *)a->nodes, (const float4 *)a->prims, a->sorted_index,
                       a->params, rays, tmax, order, nn, occluded, max_nodes(a), (uint32_t)a->n,
                       nullptr, sph_sub(a));
  }
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

}  // namespace

srf_status trace_any_sorted(Ctx *ctx, const Accel *a, const srf_ray *rays, const float *tmax,
                            uint64_t n, uint32_t key_bits, int mode, uint32_t *occluded,
                            hipStream_t s) {
  if (n == 0) return SRF_OK;
  srf_status st = check_any_sorted(a, key_bits);
  if (st != SRF_OK) return st;
  const uint32_t nn = (uint32_t)n;
  if (key_bits > 0 || mode == SRF_LAUNCH_PERSISTENT) {
    st = ensure_scratch(ctx, any_sorted_bytes(nn), s);
    if (st != SRF_OK) return st;
  }
  return trace_any_sorted_in((char *)ctx->scratch, ctx->scratch_bytes, a, rays, tmax, nn,
                             key_bits, mode, occluded, s);
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

// ---------------------------------------------------------------------------
// Sorted occlusion that allocates nothing (srf_trace_any_sorted_noalloc): the
// keys, the order, the work counter and the sort's slices live in `scratch`,
// a static device array each GPU holds for the life of the process. Host code
// never passes the array's name to a kernel: it looks up the array's device
// address once per device (hipGetSymbolAddress, with that device current)
// and hands that pointer on. Every call on one GPU shares the array, so calls
// that overlap on one GPU (on different streams, or from different threads)
// race; the counter and the sort's histograms and tile flags are cleared by a
// kernel at the start of each call.
namespace srf {
namespace {

constexpr uint32_t kNoAllocMaxRays = 1u << 20;
constexpr size_t kNoAllocScratchBytes = (size_t)26 << 20;  // any_sorted_bytes(2^20) + padding
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocMaxDevices];

// The device address of this device's copy of `scratch` (dev is current).
srf_status noalloc_scratch(int dev, char **out) {
  if (dev < 0 || dev >= kNoAllocMaxDevices)
    return set_error(SRF_ERR_UNSUPPORTED, "device %d beyond the %d the library tracks", dev,
                     kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_error(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<char *>(g_scratch_ptr[dev]);
  return SRF_OK;
}

}  // namespace
}  // namespace srf

extern "C" srf_status srf_trace_any_sorted_noalloc(const srf_accel *accel, const srf_ray *rays,
                                                   const float *t_max, uint64_t n,
                                                   uint32_t key_bits, int mode,
                                                   uint32_t *occluded, void *stream) {
  using namespace srf;
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!a || (n && (!rays || !occluded))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n > kNoAllocMaxRays)
    return set_error(SRF_ERR_INVALID_ARG, "%llu rays: at most %u per call without allocation",
                     (unsigned long long)n, kNoAllocMaxRays);
  if (mode != SRF_LAUNCH_PLAIN && mode != SRF_LAUNCH_PERSISTENT)
    return set_error(SRF_ERR_INVALID_ARG, "mode must be SRF_LAUNCH_PLAIN or _PERSISTENT");
  srf_status st = check_any_sorted(a, key_bits);
  if (st != SRF_OK) return st;
  if (n == 0) return SRF_OK;
  const uint32_t nn = (uint32_t)n;
  if (any_sorted_bytes(nn) + 8 * 256 > kNoAllocScratchBytes)
    return set_error(SRF_ERR_INTERNAL, "no-allocation scratch too small for %u rays", nn);
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) prev = -1;
  if (prev != a->device && hipSetDevice(a->device) != hipSuccess)
    return set_error(SRF_ERR_HIP, "hipSetDevice(%d) failed", a->device);
  char *base = nullptr;
  st = noalloc_scratch(a->device, &base);
  if (st == SRF_OK)
    st = trace_any_sorted_in(base, kNoAllocScratchBytes, a, rays, t_max, nn, key_bits, mode,
                             occluded, reinterpret_cast<hipStream_t>(stream));
  if (prev >= 0 && prev != a->device) (void)hipSetDevice(prev);
  return st;
}
