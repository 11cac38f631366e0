// srf_scene.hip -- two-level scenes: many accelerators placed by transforms,
// a top-level BVH over their world boxes, and the trace through both levels
// (one kernel when every element is a triangle accelerator; element passes
// when the scene mixes primitive kinds, see the note above scene_trace_mixed).
//
// The top level follows the reference's RenderGraph exactly:
//   world box    TransformBoundingBox          gles/gles_render_graph.cc:19-79
//   top BVH      BVHTree::BuildTree / ConstructTree
//                                               render/render_bvh_tree.cc:319-502
//   trace        BVHTree::Trace (double slabs)  render/render_bvh_tree.cc:504-586
//                RenderGraph::Trace             gles/gles_render_graph.cc:145-314
// The top-level build runs on the host (a handful of boxes, as in the
// reference); the trace is one lane per ray on the GPU:
//   1. walk the top BVH in the reference's order and pick, among the leaves
//      whose box the ray enters, the next one by (entry t, visit order) --
//      the order std::sort leaves them in for up to 16 entries (insertion
//      sort, stable); the walk is repeated per element, so no per-ray list
//      is kept;
//   2. stop once the nearest hit so far lies before that entry t (the
//      reference skips every later element in that case);
//   3. move the ray into the element's space with its invT / invT33 (double),
//      trace the element's LBVH with the single-level traversal, and keep
//      the hit if its world distance |T (o + t d) - o| is nearer.
// With many more than 16 entries the reference's introsort may order equal
// entry distances differently; only exact ties of entry t are affected.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <new>
#include <vector>

#include "srf_trace_dev.h"

namespace srf {

// 64-byte top-level node (BVHNode, render/render_bvh_tree.h): box in double,
// children (0 = none), split axis, leaf flag, element id.
struct alignas(16) TlasNode {
  double bmin[3];
  double bmax[3];
  int32_t child[2];
  uint32_t axis_leaf;  // axis | (is_leaf << 8)
  uint32_t id;
};
static_assert(sizeof(TlasNode) == 64, "TlasNode is 64 bytes");

// Per element on the device: its accelerator and the rows 0..3, columns
// 0..2 of T, invT and invT33 (MultV reads nothing else).
struct alignas(16) SceneInst {
  const float4 *nodes;
  const float4 *prims;
  uint32_t max_nodes;
  uint32_t nprims;
  float kEPS;
  uint32_t pad;
  double T[12];
  double invT[12];
  double invT33[12];
};

// Scratch of the element passes of a mixed scene's trace (grown on demand).
struct MixedScratch {
  std::mutex mu;                 // one mixed trace per scene at a time
  void *buf = nullptr;
  size_t bytes = 0;
  std::vector<uint32_t> counts;  // rays per element (host copy)
  std::vector<uint32_t> off;     // element list offsets
};

struct Scene {
  int device;
  uint32_t count;       // elements
  uint32_t num_nodes;   // top-level nodes
  uint32_t depth;       // most branches on a root-to-leaf path
  double bmin[3], bmax[3];
  TlasNode *d_nodes;
  SceneInst *d_inst;
  uint32_t *d_counter;  // work counter of the persistent launch
  uint32_t *d_leaf;     // element -> its top-level leaf node
  std::vector<TlasNode> h_nodes;
  int cus;              // compute units of the device
  srf_launch_config tuned;
  bool mixed;           // some element is not a triangle accelerator
  std::vector<const srf_accel *> accels;
  MixedScratch *mx;
};

namespace {

constexpr uint32_t kChunk = 64;  // rays a wavefront takes per grab

// ------------------------------------------------------------------ host

// Matrixd::MultV (render/render_matrix.h:165-183), row vector times m.
void mult_v_host(const double m[16], const double v[3], double out[3]) {
  for (int k = 0; k < 3; k++)
    out[k] = ((m[k] * v[0] + m[4 + k] * v[1]) + m[8 + k] * v[2]) + m[12 + k];
}

double dmin_ref(double a, double b) { return (b < a) ? b : a; }
double dmax_ref(double a, double b) { return (a < b) ? b : a; }

// TransformBoundingBox: corners x fastest, min/max starting from corner 0.
void world_box(const double bmin[3], const double bmax[3], const double T[16], double mn[3],
               double mx[3]) {
  for (int c = 0; c < 8; c++) {
    const double p[3] = {(c & 1) ? bmax[0] : bmin[0], (c & 2) ? bmax[1] : bmin[1],
                         (c & 4) ? bmax[2] : bmin[2]};
    double q[3];
    mult_v_host(T, p, q);
    for (int k = 0; k < 3; k++) {
      mn[k] = c ? dmin_ref(q[k], mn[k]) : q[k];
      mx[k] = c ? dmax_ref(q[k], mx[k]) : q[k];
    }
  }
}

struct Item {  // BVHData
  double bmin[3], bmax[3];
  uint32_t id;
};

class TlasBuilder {
 public:
  explicit TlasBuilder(std::vector<Item> items) : items_(std::move(items)) {}

  std::vector<TlasNode> build() {
    nodes_.assign(2 * items_.size(), TlasNode{});
    count_ = 1;
    span_box(0, items_.size(), nodes_[0].bmin, nodes_[0].bmax);  // root: unpadded
    construct(0, 0, items_.size());
    nodes_.resize(count_);
    return nodes_;
  }

 private:
  // CalculateBoundingBox over [lo, hi)
  void span_box(size_t lo, size_t hi, double mn[3], double mx[3]) const {
    for (int k = 0; k < 3; k++) {
      mn[k] = items_[lo].bmin[k];
      mx[k] = items_[lo].bmax[k];
    }
    for (size_t i = lo + 1; i < hi; i++)
      for (int k = 0; k < 3; k++) {
        mn[k] = dmin_ref(mn[k], items_[i].bmin[k]);
        mx[k] = dmax_ref(mx[k], items_[i].bmax[k]);
      }
  }

  // PartitionNodes: items after the first whose bmax on the axis lies below
  // the split are swapped forward; the first item trades places with the
  // last of them; the returned split counts every item below the split.
  size_t partition(size_t lo, size_t hi, double pos, int axis) {
    const Item head = items_[lo];
    size_t pivot = lo;
    for (size_t i = lo + 1; i < hi; i++)
      if (items_[i].bmax[axis] < pos) std::swap(items_[++pivot], items_[i]);
    items_[lo] = items_[pivot];
    items_[pivot] = head;
    size_t split = lo;
    for (size_t i = lo; i < hi; i++)
      if (items_[i].bmax[axis] < pos) split++;
    return split;
  }

  uint32_t add_child(size_t lo, size_t hi) {
    const uint32_t c = count_++;
    TlasNode &nd = nodes_[c];
    span_box(lo, hi, nd.bmin, nd.bmax);
    for (int k = 0; k < 3; k++) {  // 1e-6 of the extent on each side
      const double ext = nd.bmax[k] - nd.bmin[k];
      nd.bmin[k] = nd.bmin[k] - ext * 1.0e-6;
      nd.bmax[k] = nd.bmax[k] + ext * 1.0e-6;
    }
    return c;
  }

  // ConstructTree: spatial median along the longest axis of the node box,
  // then the other two axes in turn; object median if none separates.
  void construct(uint32_t node, size_t lo, size_t hi) {
    static const int kOrder[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
    const size_t n = hi - lo;
    if (n == 1) {
      nodes_[node].axis_leaf = 1u << 8;
      nodes_[node].child[0] = nodes_[node].child[1] = 0;
      nodes_[node].id = items_[lo].id;
      return;
    }
    double size[3], base[3];
    for (int k = 0; k < 3; k++) {
      size[k] = nodes_[node].bmax[k] - nodes_[node].bmin[k];
      base[k] = nodes_[node].bmin[k];
    }
    int first = 0;
    double longest = size[0];
    if (longest < size[1]) {
      longest = size[1];
      first = 1;
    }
    if (longest < size[2]) first = 2;
    size_t split = lo, nleft = 0, nright = n;
    int axis = 0;
    for (int i = 0; i < 3; i++) {
      axis = kOrder[first][i];
      split = partition(lo, hi, base[axis] + size[axis] * 0.5, axis);
      if (split != lo && split != hi) {
        nleft = split - lo;
        nright = n - nleft;
        break;
      }
    }
    if (split == lo || split == hi) {
      nleft = n / 2;
      nright = n - nleft;
      split = lo + (hi - lo) / 2;
    }
    nodes_[node].axis_leaf = (uint32_t)axis;
    if (nleft > 0) {
      const uint32_t c = add_child(lo, split);
      nodes_[node].child[0] = (int32_t)c;
      construct(c, lo, split);
    }
    if (nright > 0) {
      const uint32_t c = add_child(split, hi);
      nodes_[node].child[1] = (int32_t)c;
      construct(c, split, hi);
    }
  }

  std::vector<Item> items_;
  std::vector<TlasNode> nodes_;
  uint32_t count_ = 0;
};

uint32_t branch_depth(const std::vector<TlasNode> &nodes, uint32_t node) {
  const TlasNode &nd = nodes[node];
  if (nd.axis_leaf >> 8) return 0;
  uint32_t d = 0;
  for (int c = 0; c < 2; c++)
    if (nd.child[c] > 0) d = std::max(d, branch_depth(nodes, (uint32_t)nd.child[c]));
  return d + 1;
}

// ---------------------------------------------------------------- device

struct TlasRay {
  double o[3], inv[3];
  int sign[3];
  uint32_t sbits;  // sign[k] << k: the near child without an indexed load
};

// SetupRay, render/render_bvh_tree.cc:40-71 (directions below 2e-15 get
// +-FLT_MAX as inverse)
__device__ __forceinline__ void tlas_setup(const double o[3], const double d[3], TlasRay &r) {
  for (int k = 0; k < 3; k++) {
    r.o[k] = o[k];
    r.sign[k] = d[k] < 0.0 ? 1 : 0;
    r.inv[k] = fabs(d[k]) > 2.0e-15 ? 1.0 / d[k] : (d[k] < 0.0 ? -FLT_MAX : FLT_MAX);
  }
  r.sbits = (uint32_t)r.sign[0] | ((uint32_t)r.sign[1] << 1) | ((uint32_t)r.sign[2] << 2);
}

// TestRayAABB, render/render_bvh_tree.cc:198-255
__device__ __forceinline__ bool tlas_box(const TlasRay &r, const TlasNode &nd, double &tmin) {
  double lo[3], hi[3];
  for (int k = 0; k < 3; k++) {
    lo[k] = r.sign[k] ? nd.bmax[k] : nd.bmin[k];
    hi[k] = r.sign[k] ? nd.bmin[k] : nd.bmax[k];
  }
  const double tminx = (lo[0] - r.o[0]) * r.inv[0], tmaxx = (hi[0] - r.o[0]) * r.inv[0];
  const double tminy = (lo[1] - r.o[1]) * r.inv[1], tmaxy = (hi[1] - r.o[1]) * r.inv[1];
  double t0 = (tminx > tminy) ? tminx : tminy;
  double t1 = (tmaxx < tmaxy) ? tmaxx : tmaxy;
  const double tminz = (lo[2] - r.o[2]) * r.inv[2], tmaxz = (hi[2] - r.o[2]) * r.inv[2];
  t0 = (t0 > tminz) ? t0 : tminz;
  t1 = (t1 < tmaxz) ? t1 : tmaxz;
  tmin = t0;
  return (t1 > 0.0) && (t0 <= t1);
}

__device__ __forceinline__ TlasNode load_tlas(const TlasNode *tl, uint32_t i) {
  const double2 *p = reinterpret_cast<const double2 *>(tl + i);
  const double2 a = p[0], b = p[1], c = p[2];
  const uint4 m = reinterpret_cast<const uint4 *>(tl + i)[3];
  TlasNode nd;
  nd.bmin[0] = a.x;
  nd.bmin[1] = a.y;
  nd.bmin[2] = b.x;
  nd.bmax[0] = b.y;
  nd.bmax[1] = c.x;
  nd.bmax[2] = c.y;
  nd.child[0] = (int32_t)m.x;
  nd.child[1] = (int32_t)m.y;
  nd.axis_leaf = m.z;
  nd.id = m.w;
  return nd;
}

// One pass over the top BVH in the reference's visit order: the leaf with
// the smallest key (entry t, visit rank) above the previous key. Returns
// false when none is left.
template <int kBlock>
__device__ __forceinline__ bool next_element(const TlasNode *__restrict__ tl, uint32_t nnodes,
                                             const TlasRay &r, uint32_t *tst, uint32_t tdepth,
                                             bool have_prev, double prev_t, uint32_t prev_rank,
                                             double &next_t, uint32_t &next_rank,
                                             uint32_t &next_id) {
  bool found = false;
  uint32_t rank = 0, sp = 0, node = 0;
  for (;;) {
    if (node >= nnodes) break;
    const TlasNode nd = load_tlas(tl, node);
    if (nd.axis_leaf >> 8) {
      double t0;
      if (tlas_box(r, nd, t0)) {
        const bool after = !have_prev || t0 > prev_t || (t0 == prev_t && rank > prev_rank);
        const bool before = !found || t0 < next_t || (t0 == next_t && rank < next_rank);
        if (after && before) {
          found = true;
          next_t = t0;
          next_rank = rank;
          next_id = nd.id;
        }
        rank++;
      }
      if (sp == 0) break;
      node = tst[(--sp) * kBlock];
    } else {
      double a, b;
      const bool hl = nd.child[0] > 0 && (uint32_t)nd.child[0] < nnodes &&
                      tlas_box(r, load_tlas(tl, (uint32_t)nd.child[0]), a);
      const bool hr = nd.child[1] > 0 && (uint32_t)nd.child[1] < nnodes &&
                      tlas_box(r, load_tlas(tl, (uint32_t)nd.child[1]), b);
      if (hl && hr) {
        const int o = (int)((r.sbits >> (nd.axis_leaf & 3u)) & 1u);
        const uint32_t c0 = (uint32_t)nd.child[0], c1 = (uint32_t)nd.child[1];
        if (sp < tdepth) tst[(sp++) * kBlock] = o ? c0 : c1;
        node = o ? c1 : c0;
      } else if (hl) {
        node = (uint32_t)nd.child[0];
      } else if (hr) {
        node = (uint32_t)nd.child[1];
      } else {
        if (sp == 0) break;
        node = tst[(--sp) * kBlock];
      }
    }
  }
  return found;
}

__device__ __forceinline__ void mult_rows(const double *m, const double v[3], double out[3]) {
  for (int k = 0; k < 3; k++)
    out[k] = ((m[k] * v[0] + m[3 + k] * v[1]) + m[6 + k] * v[2]) + m[9 + k];
}

// RenderGraph::Trace for one ray.
template <int kBlock>
__device__ __forceinline__ void scene_ray(const TlasNode *__restrict__ tl, uint32_t nnodes,
                                          const SceneInst *__restrict__ inst, uint32_t ninst,
                                          const srf_ray *__restrict__ rays, uint32_t i,
                                          uint32_t *bst, uint32_t *tst, uint32_t tdepth,
                                          srf_hit *__restrict__ hits,
                                          uint32_t *__restrict__ inst_out) {
  const float4 *r4 = reinterpret_cast<const float4 *>(rays + i);
  const float4 ra = r4[0], rb = r4[1];
  const float wo[3] = {ra.x, ra.y, ra.z};
  const bool ds = (__float_as_uint(ra.w) & SRF_RAY_DOUBLE_SIDED) != 0;
  const double ro[3] = {ra.x, ra.y, ra.z}, rd[3] = {rb.x, rb.y, rb.z};
  TlasRay tr;
  tlas_setup(ro, rd, tr);
  double nearest = DBL_MAX;
  float bt = FLT_MAX, bu = 0.0f, bv = 0.0f;
  uint32_t bprim = SRF_NO_PRIM, binst = SRF_NO_PRIM;
  bool have_prev = false;
  double prev_t = 0.0;
  uint32_t prev_rank = 0;
  for (uint32_t pass = 0; pass < ninst; pass++) {
    double nt = 0.0;
    uint32_t nr = 0, nid = 0;
    if (!next_element<kBlock>(tl, nnodes, tr, tst, tdepth, have_prev, prev_t, prev_rank, nt, nr,
                              nid))
      break;
    if (nearest < nt) break;  // every later element is skipped as well
    have_prev = true;
    prev_t = nt;
    prev_rank = nr;
    if (nid >= ninst) continue;
    const SceneInst *e = inst + nid;
    double lo[3], ld[3];
    mult_rows(e->invT, ro, lo);
    mult_rows(e->invT33, rd, ld);
    RaySetup s;
    for (int k = 0; k < 3; k++) {
      s.o[k] = (float)lo[k];
      s.d[k] = (float)ld[k];
      s.sign[k] = s.d[k] < 0.0f ? 1 : 0;
      s.inv[k] = 1.0f / s.d[k];
    }
    s.sbits = (uint32_t)s.sign[0] | ((uint32_t)s.sign[1] << 1) | ((uint32_t)s.sign[2] << 2);
    float t, u, v;
    uint32_t prim;
    tri_traverse<false>(e->nodes, e->prims, e->max_nodes, e->nprims, s, ds, e->kEPS,
                        SRF_NO_PRIM, FLT_MAX, Stack32<kBlock>{bst}, t, u, v, prim);
    if (prim == SRF_NO_PRIM) continue;
    // world distance of the local hit point (real3 arithmetic in float)
    double lp[3], wp[3];
    for (int k = 0; k < 3; k++) lp[k] = (double)(float)(lo[k] + (double)t * ld[k]);
    mult_rows(e->T, lp, wp);
    const float px = (float)wp[0] - wo[0], py = (float)wp[1] - wo[1], pz = (float)wp[2] - wo[2];
    const float tw = sqrtf((px * px + py * py) + pz * pz);
    if ((double)tw < nearest) {
      nearest = tw;
      bt = tw;
      bu = u;
      bv = v;
      bprim = prim;
      binst = nid;
    }
  }
  reinterpret_cast<float4 *>(hits)[i] = make_float4(bt, bu, bv, __uint_as_float(bprim));
  if (inst_out) inst_out[i] = binst;
}

// One lane per ray; with `counter` the launch is persistent: each wavefront
// takes kChunk consecutive rays (one 8x8 tile) at a time until none is left.
template <int kBlock>
__global__ void __launch_bounds__(kBlock)
k_scene_tri(const TlasNode *__restrict__ tl, uint32_t nnodes, const SceneInst *__restrict__ inst,
            uint32_t ninst, const srf_ray *__restrict__ rays, uint32_t n,
            srf_hit *__restrict__ hits, uint32_t *__restrict__ inst_out, uint32_t tdepth,
            uint32_t *counter) {
  extern __shared__ uint32_t smem[];
  uint32_t *bst = smem + threadIdx.x;
  uint32_t *tst = smem + kTraceStack * kBlock + threadIdx.x;
  if (!counter) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i < n) scene_ray<kBlock>(tl, nnodes, inst, ninst, rays, i, bst, tst, tdepth, hits, inst_out);
    return;
  }
  const uint32_t lane = threadIdx.x & 63u;
  for (;;) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(counter, kChunk);
    base = __shfl(base, 0, 64);
    if (base >= n) break;
    const uint32_t i = base + lane;
    if (i < n) scene_ray<kBlock>(tl, nnodes, inst, ninst, rays, i, bst, tst, tdepth, hits, inst_out);
  }
}

__global__ void k_zero_counter(uint32_t *c) { *c = 0u; }

__global__ void __launch_bounds__(256) k_zero_words(uint32_t *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = 0u;
}

// The element passes of a mixed scene. For every ray and element, the test
// the element walk applies to the element's top-level leaf (the ray enters
// the element's world box). The counting pass adds the entering rays per
// element; the filling pass (kFill) gives each such pair slot j of the
// element's list (cursor[e]), writes the ray moved into the element's space
// there as the element walk forms it (invT / invT33 in double, then float),
// and records inv[e * n + i] = j (~0 for the other pairs).
template <bool kFill>
__global__ void __launch_bounds__(256)
k_elem_select(const TlasNode *__restrict__ tl, const uint32_t *__restrict__ leaf,
              const SceneInst *__restrict__ inst, uint32_t ninst,
              const srf_ray *__restrict__ rays, uint32_t n, uint32_t *counts,
              const uint32_t *__restrict__ off, srf_ray *__restrict__ lrays,
              uint32_t *__restrict__ inv) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u;
  const bool live = i < n;
  double ro[3] = {0.0, 0.0, 0.0}, rd[3] = {0.0, 0.0, 1.0};
  uint32_t flags = 0;
  if (live) {
    const float4 *r4 = reinterpret_cast<const float4 *>(rays + i);
    const float4 ra = r4[0], rb = r4[1];
    ro[0] = ra.x, ro[1] = ra.y, ro[2] = ra.z;
    rd[0] = rb.x, rd[1] = rb.y, rd[2] = rb.z;
    flags = __float_as_uint(ra.w) & SRF_RAY_DOUBLE_SIDED;
  }
  TlasRay tr;
  tlas_setup(ro, rd, tr);
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (uint32_t e = 0; e < ninst; e++) {
    bool sel = false;
    if (live) {
      double t0;
      sel = tlas_box(tr, load_tlas(tl, leaf[e]), t0);
    }
    const uint64_t m = __ballot(sel);
    uint32_t base = 0;
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      if ((int)lane == leader) base = atomicAdd(&counts[e], (uint32_t)__popcll(m));
      base = (uint32_t)__shfl((int)base, leader, 64);
    }
    if (!kFill) continue;
    const uint32_t j = base + (uint32_t)__popcll(m & below);
    if (live) inv[(size_t)e * n + i] = sel ? j : 0xffffffffu;
    if (sel) {
      const SceneInst *el = inst + e;
      double lo[3], ld[3];
      mult_rows(el->invT, ro, lo);
      mult_rows(el->invT33, rd, ld);
      float4 *w = reinterpret_cast<float4 *>(lrays + off[e] + j);
      w[0] = make_float4((float)lo[0], (float)lo[1], (float)lo[2], __uint_as_float(flags));
      w[1] = make_float4((float)ld[0], (float)ld[1], (float)ld[2], __uint_as_float(SRF_NO_PRIM));
    }
  }
}

// RenderGraph::Trace from the elements' own traces: the element walk of
// scene_ray, with each element's bottom-level hit read from its list
// (lhits[off[e] + inv[e * n + i]]) instead of traced here. The bottom-level
// trace of an element does not depend on the hits before it (it starts from
// t_max = FLT_MAX), so the walk's result is the same.
template <int kBlock>
__global__ void __launch_bounds__(kBlock)
k_scene_resolve(const TlasNode *__restrict__ tl, uint32_t nnodes,
                const SceneInst *__restrict__ inst, uint32_t ninst,
                const srf_ray *__restrict__ rays, uint32_t n, const uint32_t *__restrict__ inv,
                const uint32_t *__restrict__ off, const srf_hit *__restrict__ lhits,
                srf_hit *__restrict__ hits, uint32_t *__restrict__ inst_out, uint32_t tdepth) {
  extern __shared__ uint32_t smem[];
  uint32_t *tst = smem + threadIdx.x;
  const uint32_t i = xcd_block() * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 *r4 = reinterpret_cast<const float4 *>(rays + i);
  const float4 ra = r4[0], rb = r4[1];
  const float wo[3] = {ra.x, ra.y, ra.z};
  const double ro[3] = {ra.x, ra.y, ra.z}, rd[3] = {rb.x, rb.y, rb.z};
  TlasRay tr;
  tlas_setup(ro, rd, tr);
  double nearest = DBL_MAX;
  float bt = FLT_MAX, bu = 0.0f, bv = 0.0f;
  uint32_t bprim = SRF_NO_PRIM, binst = SRF_NO_PRIM;
  bool have_prev = false;
  double prev_t = 0.0;
  uint32_t prev_rank = 0;
  for (uint32_t pass = 0; pass < ninst; pass++) {
    double nt = 0.0;
    uint32_t nr = 0, nid = 0;
    if (!next_element<kBlock>(tl, nnodes, tr, tst, tdepth, have_prev, prev_t, prev_rank, nt, nr,
                              nid))
      break;
    if (nearest < nt) break;  // every later element is skipped as well
    have_prev = true;
    prev_t = nt;
    prev_rank = nr;
    if (nid >= ninst) continue;
    const uint32_t j = inv[(size_t)nid * n + i];
    if (j == 0xffffffffu) continue;
    const float4 h = reinterpret_cast<const float4 *>(lhits)[off[nid] + j];
    const uint32_t prim = __float_as_uint(h.w);
    if (prim == SRF_NO_PRIM) continue;
    const SceneInst *e = inst + nid;
    double lo[3], ld[3];
    mult_rows(e->invT, ro, lo);
    mult_rows(e->invT33, rd, ld);
    // world distance of the local hit point (real3 arithmetic in float)
    double lp[3], wp[3];
    for (int k = 0; k < 3; k++) lp[k] = (double)(float)(lo[k] + (double)h.x * ld[k]);
    mult_rows(e->T, lp, wp);
    const float px = (float)wp[0] - wo[0], py = (float)wp[1] - wo[1], pz = (float)wp[2] - wo[2];
    const float tw = sqrtf((px * px + py * py) + pz * pz);
    if ((double)tw < nearest) {
      nearest = tw;
      bt = tw;
      bu = h.y;
      bv = h.z;
      bprim = prim;
      binst = nid;
    }
  }
  reinterpret_cast<float4 *>(hits)[i] = make_float4(bt, bu, bv, __uint_as_float(bprim));
  if (inst_out) inst_out[i] = binst;
}

}  // namespace

// Matrixd::Inverse (render/render_matrix.h:40-134): the cofactor inverse by
// Cramer's rule of Intel's published SSE note, in the reference's operation
// order (products of transposed-source pairs, each cofactor a sum of three
// products minus a sum of three, the determinant from row 0, then every
// entry times 1 / det), so the inverses are the reference's bit for bit.
// Its last cofactor repeats the source's slip (element 0 where the note
// reads element 10); MultV never reads that entry.
void cramer_inverse(const double in[16], double m[16]) {
  double s[16];  // transposed source
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) s[i + 4 * j] = in[4 * i + j];
  auto pairs = [&](const int a[12][2], double p[12]) {
    for (int k = 0; k < 12; k++) p[k] = s[a[k][0]] * s[a[k][1]];
  };
  static const int lo[12][2] = {{10, 15}, {11, 14}, {9, 15}, {11, 13}, {9, 14}, {10, 13},
                                {8, 15},  {11, 12}, {8, 14}, {10, 12}, {8, 13}, {9, 12}};
  static const int hi[12][2] = {{2, 7}, {3, 6}, {1, 7}, {3, 5}, {1, 6}, {2, 5},
                                {0, 7}, {3, 4}, {0, 6}, {2, 4}, {0, 5}, {1, 4}};
  // each cofactor: (p[a0] s[b0] + p[a1] s[b1] + p[a2] s[b2])
  //              - (p[c0] s[d0] + p[c1] s[d1] + p[c2] s[d2])
  struct Cof {
    int8_t a[3], b[3], c[3], d[3];
  };
  static const Cof first[8] = {
      {{0, 3, 4}, {5, 6, 7}, {1, 2, 5}, {5, 6, 7}},
      {{1, 6, 9}, {4, 6, 7}, {0, 7, 8}, {4, 6, 7}},
      {{2, 7, 10}, {4, 5, 7}, {3, 6, 11}, {4, 5, 7}},
      {{5, 8, 11}, {4, 5, 6}, {4, 9, 10}, {4, 5, 6}},
      {{1, 2, 5}, {1, 2, 3}, {0, 3, 4}, {1, 2, 3}},
      {{0, 7, 8}, {0, 2, 3}, {1, 6, 9}, {0, 2, 3}},
      {{3, 6, 11}, {0, 1, 3}, {2, 7, 10}, {0, 1, 3}},
      {{4, 9, 10}, {0, 1, 2}, {5, 8, 11}, {0, 1, 2}}};
  static const Cof second[8] = {
      {{0, 3, 4}, {13, 14, 15}, {1, 2, 5}, {13, 14, 15}},
      {{1, 6, 9}, {12, 14, 15}, {0, 7, 8}, {12, 14, 15}},
      {{2, 7, 10}, {12, 13, 15}, {3, 6, 11}, {12, 13, 15}},
      {{5, 8, 11}, {12, 13, 14}, {4, 9, 10}, {12, 13, 14}},
      {{2, 5, 1}, {10, 11, 9}, {4, 0, 3}, {11, 9, 10}},
      {{8, 0, 7}, {11, 8, 10}, {6, 9, 1}, {10, 11, 8}},
      {{6, 11, 3}, {9, 11, 8}, {10, 2, 7}, {11, 8, 9}},
      {{10, 4, 9}, {10, 8, 9}, {8, 11, 5}, {9, 0, 8}}};
  double p[12];
  auto cof = [&](const Cof &c) {
    double v = p[c.a[0]] * s[c.b[0]] + p[c.a[1]] * s[c.b[1]] + p[c.a[2]] * s[c.b[2]];
    v -= p[c.c[0]] * s[c.d[0]] + p[c.c[1]] * s[c.d[1]] + p[c.c[2]] * s[c.d[2]];
    return v;
  };
  pairs(lo, p);
  for (int k = 0; k < 8; k++) m[k] = cof(first[k]);
  pairs(hi, p);
  for (int k = 0; k < 8; k++) m[8 + k] = cof(second[k]);
  double det = s[0] * m[0] + s[1] * m[1] + s[2] * m[2] + s[3] * m[3];
  det = 1.0 / det;
  for (int k = 0; k < 16; k++) m[k] *= det;
}

// ------------------------------------------------------------------ API

srf_status scene_create(Ctx *ctx, uint32_t n, const srf_instance *inst, Scene **out);
srf_status scene_trace(const Scene *sc, const srf_ray *rays, uint64_t n, srf_hit *hits,
                       uint32_t *inst_out, int mode, hipStream_t stream);
void scene_free(Scene *sc);
int scene_device(const Scene *sc);
srf_status scene_query(const Scene *sc, double bmin[3], double bmax[3], srf_tlas_node *nodes,
                       uint32_t *count);
srf_status fetch_params_for(const Accel *a, BuildParams *p);

srf_status scene_create(Ctx *ctx, uint32_t n, const srf_instance *inst, Scene **out) {
  std::vector<Item> items(n);
  std::vector<SceneInst> hinst(n);
  for (uint32_t i = 0; i < n; i++) {
    const Accel *a = reinterpret_cast<const Accel *>(inst[i].accel);
    if (!a) return set_error(SRF_ERR_INVALID_ARG, "instance %u: null accelerator", i);
    if (a->device != ctx->device)
      return set_error(SRF_ERR_INVALID_ARG, "instance %u: accelerator on another device", i);
    BuildParams p;
    srf_status st = fetch_params_for(a, &p);
    if (st != SRF_OK) return st;
    // BoundingBox() of the element, then its world box (RenderElement::Update)
    const double bmin[3] = {p.root_bmin[0], p.root_bmin[1], p.root_bmin[2]};
    const double bmax[3] = {p.root_bmax[0], p.root_bmax[1], p.root_bmax[2]};
    world_box(bmin, bmax, inst[i].T, items[i].bmin, items[i].bmax);
    items[i].id = i;
    SceneInst &d = hinst[i];
    d.nodes = reinterpret_cast<const float4 *>(a->nodes);
    d.prims = reinterpret_cast<const float4 *>(a->prims);
    d.max_nodes = (uint32_t)(2 * a->n - 1);
    d.nprims = (uint32_t)a->n;
    d.kEPS = FLT_EPSILON * (float)p.eps_scale;
    d.pad = 0;
    for (int r = 0; r < 4; r++)
      for (int c = 0; c < 3; c++) {
        d.T[3 * r + c] = inst[i].T[4 * r + c];
        d.invT[3 * r + c] = inst[i].invT[4 * r + c];
        d.invT33[3 * r + c] = inst[i].invT33[4 * r + c];
      }
  }
  Scene *sc = new (std::nothrow) Scene();
  if (!sc) return set_error(SRF_ERR_OUT_OF_MEMORY, "host allocation failed");
  sc->device = ctx->device;
  sc->count = n;
  sc->mixed = false;
  sc->mx = nullptr;
  sc->d_leaf = nullptr;
  for (uint32_t i = 0; i < n; i++) {
    sc->accels.push_back(inst[i].accel);
    if (reinterpret_cast<const Accel *>(inst[i].accel)->kind != kTri) sc->mixed = true;
  }
  sc->h_nodes = TlasBuilder(items).build();
  sc->num_nodes = (uint32_t)sc->h_nodes.size();
  sc->depth = branch_depth(sc->h_nodes, 0);
  for (int k = 0; k < 3; k++) {
    sc->bmin[k] = sc->h_nodes[0].bmin[k];
    sc->bmax[k] = sc->h_nodes[0].bmax[k];
  }
  sc->d_nodes = nullptr;
  sc->d_inst = nullptr;
  sc->d_counter = nullptr;
  sc->cus = 256;
  sc->tuned = srf_launch_config{128u, 1u, 8u, 0.0f};
  hipError_t e = hipDeviceGetAttribute(&sc->cus, hipDeviceAttributeMultiprocessorCount,
                                       ctx->device);
  if (e == hipSuccess) e = hipMalloc(&sc->d_nodes, sizeof(TlasNode) * sc->num_nodes);
  if (e == hipSuccess) e = hipMalloc(&sc->d_inst, sizeof(SceneInst) * n);
  if (e == hipSuccess) e = hipMalloc(&sc->d_counter, 256);
  // each element's top-level leaf (the element passes test it directly)
  std::vector<uint32_t> leaf(n, 0u);
  for (uint32_t k = 0; k < sc->num_nodes; k++)
    if ((sc->h_nodes[k].axis_leaf >> 8) && sc->h_nodes[k].id < n) leaf[sc->h_nodes[k].id] = k;
  if (e == hipSuccess) e = hipMalloc(&sc->d_leaf, sizeof(uint32_t) * n);
  if (e == hipSuccess)
    e = hipMemcpy(sc->d_leaf, leaf.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess && sc->mixed) {
    sc->mx = new (std::nothrow) MixedScratch();
    if (!sc->mx) e = hipErrorOutOfMemory;
  }
  if (e == hipSuccess)
    e = hipMemcpy(sc->d_nodes, sc->h_nodes.data(), sizeof(TlasNode) * sc->num_nodes,
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(sc->d_inst, hinst.data(), sizeof(SceneInst) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    scene_free(sc);
    return hip_error(e, "scene upload");
  }
  *out = sc;
  return SRF_OK;
}

int scene_device(const Scene *sc) { return sc->device; }

srf_status scene_query(const Scene *sc, double bmin[3], double bmax[3], srf_tlas_node *nodes,
                       uint32_t *count) {
  if (bmin && bmax)
    for (int k = 0; k < 3; k++) {
      bmin[k] = sc->bmin[k];
      bmax[k] = sc->bmax[k];
    }
  if (count) *count = sc->num_nodes;
  if (nodes)
    for (uint32_t i = 0; i < sc->num_nodes; i++) {
      const TlasNode &a = sc->h_nodes[i];
      srf_tlas_node &b = nodes[i];
      for (int k = 0; k < 3; k++) {
        b.bmin[k] = a.bmin[k];
        b.bmax[k] = a.bmax[k];
      }
      b.child[0] = a.child[0];
      b.child[1] = a.child[1];
      b.is_leaf = (int32_t)(a.axis_leaf >> 8);
      b.axis = b.is_leaf ? 0 : (int32_t)(a.axis_leaf & 3u);
      b.node_id = b.is_leaf ? a.id : 0u;
      b.pad = 0;
    }
  return SRF_OK;
}

void scene_free(Scene *sc) {
  if (!sc) return;
  if (sc->d_nodes) (void)hipFree(sc->d_nodes);
  if (sc->d_inst) (void)hipFree(sc->d_inst);
  if (sc->d_counter) (void)hipFree(sc->d_counter);
  if (sc->d_leaf) (void)hipFree(sc->d_leaf);
  if (sc->mx) {
    if (sc->mx->buf) (void)hipFree(sc->mx->buf);
    delete sc->mx;
  }
  delete sc;
}

srf_status scene_launch(const Scene *sc, const srf_launch_config &cfg, const srf_ray *rays,
                        uint64_t n, srf_hit *hits, uint32_t *inst_out, hipStream_t stream) {
  if (n == 0) return SRF_OK;
  const uint32_t block = cfg.block_size;
  if (block != 64 && block != 128 && block != 256)
    return set_error(SRF_ERR_INVALID_ARG, "block size %u (64, 128 or 256)", block);
  const uint32_t tdepth = std::max<uint32_t>(sc->depth, 1u);
  const size_t lds = sizeof(uint32_t) * (size_t)(kTraceStack + tdepth) * block;
  if (lds > 64 * 1024)
    return set_error(SRF_ERR_UNSUPPORTED, "top-level BVH %u levels deep exceeds the LDS stack",
                     sc->depth);
  uint32_t *counter = nullptr;
  uint32_t blocks = (uint32_t)((n + block - 1) / block);
  blocks = (blocks + kXcdPad - 1) / kXcdPad * kXcdPad;  // see xcd_block
  if (cfg.persistent) {
    const uint32_t waves = (uint32_t)((n + kChunk - 1) / kChunk);
    const uint32_t per_cu = std::max<uint32_t>(cfg.blocks_per_cu, 1u);
    blocks = std::min<uint32_t>(blocks, (uint32_t)sc->cus * per_cu);
    blocks = std::min<uint32_t>(blocks, (waves + (block / 64) - 1) / (block / 64));
    counter = sc->d_counter;
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, stream, counter);
  }
#define SRF_SCENE_LAUNCH(B)                                                               \
  hipLaunchKernelGGL((k_scene_tri<B>), dim3(blocks), dim3(B), lds, stream, sc->d_nodes,    \
                     sc->num_nodes, sc->d_inst, sc->count, rays, (uint32_t)n, hits, inst_out, \
                     tdepth, counter)
  if (block == 64)
    SRF_SCENE_LAUNCH(64);
  else if (block == 128)
    SRF_SCENE_LAUNCH(128);
  else
    SRF_SCENE_LAUNCH(256);
#undef SRF_SCENE_LAUNCH
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

// A scene that mixes primitive kinds is traced in element passes (the one
// kernel above inlines the triangle traversal only): the rays entering each
// element's world box are counted (one readback, so this path synchronises
// the stream once per chunk of rays), listed per element in the element's
// space, traced by that element's own closest-hit launch (srf_trace_closest:
// triangles, tetrahedra, spheres, lines or solids, exactly as a single-level
// trace of that accelerator), and k_scene_resolve walks the top level per
// ray in RenderGraph::Trace's order over those results. Rays go in chunks of
// at most 2^26 / elements so the (element, ray) slot map stays within
// 256 MB. The scratch is the scene's own; traces of one mixed scene are
// serialised.
srf_status scene_trace_mixed(const Scene *sc, const srf_ray *rays, uint64_t n, srf_hit *hits,
                             uint32_t *inst_out, hipStream_t stream) {
  MixedScratch &mx = *sc->mx;
  std::lock_guard<std::mutex> lk(mx.mu);
  const uint32_t ne = sc->count;
  const uint32_t tdepth = std::max<uint32_t>(sc->depth, 1u);
  constexpr int kResolveBlock = 128;
  const size_t lds = sizeof(uint32_t) * (size_t)tdepth * kResolveBlock;
  if (lds > 64 * 1024)
    return set_error(SRF_ERR_UNSUPPORTED, "top-level BVH %u levels deep exceeds the LDS stack",
                     sc->depth);
  const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(n, (1ull << 26) / ne));
  mx.counts.assign(ne, 0u);
  mx.off.assign(ne + 1, 0u);
  for (uint64_t c0 = 0; c0 < n; c0 += chunk) {
    const uint32_t cn = (uint32_t)std::min<uint64_t>(chunk, n - c0);
    const srf_ray *cr = rays + c0;
    const uint32_t sblocks = (cn + 255) / 256;
    // counting pass: the cursors live in the first words of the scratch
    size_t need = 256 + 4 * (size_t)ne * 2;
    if (need > mx.bytes) {
      if (mx.buf) (void)hipFree(mx.buf);
      mx.buf = nullptr;
      mx.bytes = 0;
      SRF_HIP(hipMalloc(&mx.buf, need));
      mx.bytes = need;
    }
    uint32_t *d_counts = reinterpret_cast<uint32_t *>(mx.buf);
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(256), 0, stream, d_counts, ne);
    hipLaunchKernelGGL((k_elem_select<false>), dim3(sblocks), dim3(256), 0, stream, sc->d_nodes,
                       sc->d_leaf, sc->d_inst, ne, cr, cn, d_counts, nullptr, nullptr, nullptr);
    SRF_HIP(hipGetLastError());
    SRF_HIP(hipMemcpyAsync(mx.counts.data(), d_counts, 4 * (size_t)ne, hipMemcpyDeviceToHost,
                           stream));
    SRF_HIP(hipStreamSynchronize(stream));
    uint64_t total = 0;
    for (uint32_t e = 0; e < ne; e++) {
      mx.off[e] = (uint32_t)total;
      total += mx.counts[e];
    }
    mx.off[ne] = (uint32_t)total;
    // layout: cursors, offsets, the (element, ray) slot map, local rays, hits
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_off = up(4 * (size_t)ne), o_inv = o_off + up(4 * ((size_t)ne + 1));
    const size_t o_ray = o_inv + up(4 * (size_t)ne * cn), o_hit = o_ray + up(32 * total);
    need = o_hit + up(16 * total) + 256;
    if (need > mx.bytes) {
      (void)hipFree(mx.buf);
      mx.buf = nullptr;
      mx.bytes = 0;
      SRF_HIP(hipMalloc(&mx.buf, need));
      mx.bytes = need;
    }
    char *b = reinterpret_cast<char *>(mx.buf);
    uint32_t *cursor = reinterpret_cast<uint32_t *>(b);
    uint32_t *d_off = reinterpret_cast<uint32_t *>(b + o_off);
    uint32_t *inv = reinterpret_cast<uint32_t *>(b + o_inv);
    srf_ray *lrays = reinterpret_cast<srf_ray *>(b + o_ray);
    srf_hit *lhits = reinterpret_cast<srf_hit *>(b + o_hit);
    SRF_HIP(hipMemcpyAsync(d_off, mx.off.data(), 4 * ((size_t)ne + 1), hipMemcpyHostToDevice,
                           stream));
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(256), 0, stream, cursor, ne);
    hipLaunchKernelGGL((k_elem_select<true>), dim3(sblocks), dim3(256), 0, stream, sc->d_nodes,
                       sc->d_leaf, sc->d_inst, ne, cr, cn, cursor, d_off, lrays, inv);
    SRF_HIP(hipGetLastError());
    for (uint32_t e = 0; e < ne; e++) {
      if (!mx.counts[e]) continue;
      srf_status st = srf_trace_closest(sc->accels[e], lrays + mx.off[e], mx.counts[e],
                                        lhits + mx.off[e], stream);
      if (st != SRF_OK) return st;
    }
    uint32_t rblocks = (cn + kResolveBlock - 1) / kResolveBlock;
    rblocks = (rblocks + kXcdPad - 1) / kXcdPad * kXcdPad;  // see xcd_block
    hipLaunchKernelGGL((k_scene_resolve<kResolveBlock>), dim3(rblocks), dim3(kResolveBlock), lds,
                       stream, sc->d_nodes, sc->num_nodes, sc->d_inst, ne, cr, cn, inv, d_off,
                       lhits, hits + c0, inst_out ? inst_out + c0 : nullptr, tdepth);
    SRF_HIP(hipGetLastError());
    // the next chunk reuses the scratch
    if (c0 + chunk < n) SRF_HIP(hipStreamSynchronize(stream));
  }
  return SRF_OK;
}

srf_status scene_trace(const Scene *sc, const srf_ray *rays, uint64_t n, srf_hit *hits,
                       uint32_t *inst_out, int mode, hipStream_t stream) {
  if (sc->mixed) {
    if (mode != SRF_LAUNCH_TUNED && mode != SRF_LAUNCH_PLAIN && mode != SRF_LAUNCH_PERSISTENT)
      return set_error(SRF_ERR_INVALID_ARG, "launch mode %d", mode);
    if (n == 0) return SRF_OK;
    return scene_trace_mixed(sc, rays, n, hits, inst_out, stream);
  }
  srf_launch_config cfg = sc->tuned;
  if (mode == SRF_LAUNCH_PLAIN) {
    cfg.persistent = 0;
  } else if (mode == SRF_LAUNCH_PERSISTENT && !cfg.persistent) {
    cfg.persistent = 1;
    cfg.blocks_per_cu = 8;
  } else if (mode != SRF_LAUNCH_TUNED && mode != SRF_LAUNCH_PERSISTENT) {
    return set_error(SRF_ERR_INVALID_ARG, "launch mode %d", mode);
  }
  return scene_launch(sc, cfg, rays, n, hits, inst_out, stream);
}

// Times every launch configuration on the caller's rays (HIP events on the
// stream, best of two runs after a warm-up) and keeps the fastest for
// SRF_LAUNCH_TUNED.
srf_status scene_autotune(Scene *sc, const srf_ray *rays, uint64_t n, srf_hit *hits,
                          uint32_t *inst_out, hipStream_t stream, srf_launch_config *best) {
  if (sc->mixed) {
    // the element passes have no launch shape to choose: one timed trace
    const auto t0 = std::chrono::steady_clock::now();
    srf_status st = scene_trace_mixed(sc, rays, n, hits, inst_out, stream);
    if (st != SRF_OK) return st;
    SRF_HIP(hipStreamSynchronize(stream));
    const auto t1 = std::chrono::steady_clock::now();
    srf_launch_config win{128u, 0u, 0u,
                          (float)std::chrono::duration<double, std::milli>(t1 - t0).count()};
    sc->tuned = win;
    if (best) *best = win;
    return SRF_OK;
  }
  hipEvent_t e0, e1;
  SRF_HIP(hipEventCreate(&e0));
  hipError_t he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hip_error(he, "hipEventCreate");
  }
  srf_launch_config win{};
  win.ms = FLT_MAX;
  srf_status st = SRF_OK;
  for (uint32_t block : {64u, 128u, 256u}) {
    for (uint32_t per_cu : {0u, 2u, 4u, 8u, 16u}) {
      srf_launch_config cfg{block, per_cu ? 1u : 0u, per_cu, 0.0f};
      if (per_cu * block > 2048u) continue;  // more lanes than a CU holds
      st = scene_launch(sc, cfg, rays, n, hits, inst_out, stream);  // warm-up
      if (st != SRF_OK) break;
      float t = FLT_MAX;
      for (int rep = 0; rep < 2 && st == SRF_OK; rep++) {
        he = hipEventRecord(e0, stream);
        if (he == hipSuccess) st = scene_launch(sc, cfg, rays, n, hits, inst_out, stream);
        if (he == hipSuccess && st == SRF_OK) he = hipEventRecord(e1, stream);
        if (he == hipSuccess && st == SRF_OK) he = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (he == hipSuccess && st == SRF_OK) he = hipEventElapsedTime(&ms, e0, e1);
        if (he != hipSuccess) st = hip_error(he, "autotune timing");
        t = std::min(t, ms);
      }
      if (st != SRF_OK) break;
      if (t < win.ms) {
        win = cfg;
        win.ms = t;
      }
    }
    if (st != SRF_OK) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != SRF_OK) return st;
  sc->tuned = win;
  if (best) *best = win;
  return SRF_OK;
}

}  // namespace srf

// ----------------------------------------------------------- sparse volume
// SparseVolumeAccel (render/render_accel_volume.cc:35-139): a BVHTree over
// the blocks' integer boxes (the same top-level tree as above, built on the
// host) and, per position, BVHTree::Locate (render/render_bvh_tree.cc:
// 588-654: the first leaf in its visit order whose own box contains the
// point) and the nearest voxel of that block. As in the reference, the local
// offset is scaled by size[0] / extent[0] on every axis.
namespace srf {

struct SparseBlock {  // device copy of one srf_volume_block
  int32_t offset[3], extent[3], size[3], id;
  const void *data;
};

struct SparseVolume {
  int device;
  uint32_t count;
  uint32_t num_nodes;
  int components;  // min(components, 4)
  int format;      // 1 byte, 8 float, 16 double
  int global_dim[3];
  TlasNode *d_nodes;
  SparseBlock *d_blocks;
};

namespace {

constexpr int kLocateStack = 64;

// BVHTree::Locate's first hit (TestPointAABB / TestPointBVHNode, :173-209)
__device__ __forceinline__ int sparse_locate(const TlasNode *__restrict__ tl, uint32_t nnodes,
                                             const double p[3]) {
  uint32_t stack[kLocateStack];
  int sp = 0;
  uint32_t node = 0;
  auto inside = [&](const TlasNode &nd) {
    return p[0] >= nd.bmin[0] && p[0] <= nd.bmax[0] && p[1] >= nd.bmin[1] &&
           p[1] <= nd.bmax[1] && p[2] >= nd.bmin[2] && p[2] <= nd.bmax[2];
  };
  for (;;) {
    if (node >= nnodes) return -1;
    const TlasNode nd = load_tlas(tl, node);
    if (nd.axis_leaf >> 8) {
      if (inside(nd)) return (int)nd.id;
      if (sp < 1) return -1;
      node = stack[--sp];
      continue;
    }
    const bool l = nd.child[0] > 0 && inside(load_tlas(tl, (uint32_t)nd.child[0]));
    const bool r = nd.child[1] > 0 && inside(load_tlas(tl, (uint32_t)nd.child[1]));
    if (l && r) {
      if (sp < kLocateStack) stack[sp++] = (uint32_t)nd.child[1];
      node = (uint32_t)nd.child[0];
    } else if (l) {
      node = (uint32_t)nd.child[0];
    } else if (r) {
      node = (uint32_t)nd.child[1];
    } else {
      if (sp < 1) return -1;
      node = stack[--sp];
    }
  }
}

// SparseVolumeAccel::Sample (render/render_accel_volume.cc:59-139). With uvw
// the position is Texture::Fetch's u * globalDim (float) per axis
// (gles/gles_common.h:353-363) and the result goes out as float RGBA.
__global__ void __launch_bounds__(256)
k_sparse_sample(const TlasNode *__restrict__ tl, uint32_t nnodes,
                const SparseBlock *__restrict__ blocks, int comps, int format,
                const double *__restrict__ pos, const float *__restrict__ uvw, int gx, int gy,
                int gz, uint32_t n, double4 *__restrict__ out_d, float4 *__restrict__ out_f) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double p[3];
  if (uvw) {
    p[0] = (double)(uvw[3 * (size_t)i] * (float)gx);
    p[1] = (double)(uvw[3 * (size_t)i + 1] * (float)gy);
    p[2] = (double)(uvw[3 * (size_t)i + 2] * (float)gz);
  } else {
    p[0] = pos[3 * (size_t)i];
    p[1] = pos[3 * (size_t)i + 1];
    p[2] = pos[3 * (size_t)i + 2];
  }
  double v[4] = {0.0, 0.0, 0.0, 1.0};
  const int b = sparse_locate(tl, nnodes, p);
  if (b < 0) {
    v[3] = 0.0;
  } else {
    const SparseBlock &bk = blocks[b];
    double x = p[0] - (double)bk.offset[0];
    double y = p[1] - (double)bk.offset[1];
    double z = p[2] - (double)bk.offset[2];
    const double sx = (double)bk.size[0] / (double)bk.extent[0];
    x = x * sx;
    y = y * sx;
    z = z * sx;
    const int ix = max(min(bk.size[0] - 1, (int)x), 0);
    const int iy = max(min(bk.size[1] - 1, (int)y), 0);
    const int iz = max(min(bk.size[2] - 1, (int)z), 0);
    const size_t idx = (size_t)(iz * bk.size[0] * bk.size[1] + iy * bk.size[0] + ix);
    for (int c = 0; c < comps; c++) {
      const size_t e = (size_t)comps * idx + c;
      if (format == SRF_TEX_BYTE)
        v[c] = (double)((float)((const unsigned char *)bk.data)[e] / 255.0f);
      else if (format == SRF_TEX_FLOAT)
        v[c] = (double)((const float *)bk.data)[e];
      else
        v[c] = (double)(float)((const double *)bk.data)[e];
    }
  }
  if (out_d) out_d[i] = make_double4(v[0], v[1], v[2], v[3]);
  if (out_f) out_f[i] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
}

}  // namespace

srf_status sparse_create(Ctx *ctx, const srf_volume_block *blocks, uint32_t count,
                         const int global_dim[3], int components, int format,
                         SparseVolume **out) {
  std::vector<Item> items(count);
  for (uint32_t i = 0; i < count; i++) {
    const srf_volume_block &b = blocks[i];
    for (int k = 0; k < 3; k++) {
      if (b.extent[k] <= 0 || b.size[k] <= 0)
        return set_error(SRF_ERR_INVALID_ARG, "block %u: extent and size must be > 0", i);
      items[i].bmin[k] = (double)b.offset[k];
      items[i].bmax[k] = (double)(b.offset[k] + b.extent[k]);
    }
    if (!b.data) return set_error(SRF_ERR_INVALID_ARG, "block %u: null data", i);
    items[i].id = i;
  }
  std::vector<TlasNode> nodes = TlasBuilder(std::move(items)).build();
  if (branch_depth(nodes, 0) >= (uint32_t)kLocateStack)
    return set_error(SRF_ERR_UNSUPPORTED, "block tree deeper than %d levels", kLocateStack);
  std::vector<SparseBlock> dev(count);
  for (uint32_t i = 0; i < count; i++) {
    for (int k = 0; k < 3; k++) {
      dev[i].offset[k] = blocks[i].offset[k];
      dev[i].extent[k] = blocks[i].extent[k];
      dev[i].size[k] = blocks[i].size[k];
    }
    dev[i].id = blocks[i].id;
    dev[i].data = blocks[i].data;
  }
  SparseVolume *sv = new (std::nothrow) SparseVolume();
  if (!sv) return set_error(SRF_ERR_OUT_OF_MEMORY, "host allocation failed");
  sv->device = ctx->device;
  sv->count = count;
  sv->num_nodes = (uint32_t)nodes.size();
  sv->components = components >= 4 ? 4 : components;
  sv->format = format;
  for (int k = 0; k < 3; k++) sv->global_dim[k] = global_dim ? global_dim[k] : 0;
  sv->d_nodes = nullptr;
  sv->d_blocks = nullptr;
  hipError_t e = hipMalloc((void **)&sv->d_nodes, sizeof(TlasNode) * nodes.size());
  if (e == hipSuccess) e = hipMalloc((void **)&sv->d_blocks, sizeof(SparseBlock) * count);
  if (e == hipSuccess)
    e = hipMemcpy(sv->d_nodes, nodes.data(), sizeof(TlasNode) * nodes.size(),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(sv->d_blocks, dev.data(), sizeof(SparseBlock) * count, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(sv->d_nodes);
    (void)hipFree(sv->d_blocks);
    delete sv;
    return hip_error(e, "sparse volume upload");
  }
  *out = sv;
  return SRF_OK;
}

void sparse_free(SparseVolume *sv) {
  if (!sv) return;
  (void)hipFree(sv->d_nodes);
  (void)hipFree(sv->d_blocks);
  delete sv;
}

srf_status sparse_sample(const SparseVolume *sv, const double *pos, const float *uvw, uint64_t n,
                         double *out_d, float *out_f, hipStream_t s) {
  if (n == 0) return SRF_OK;
  hipLaunchKernelGGL(k_sparse_sample, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s,
                     sv->d_nodes, sv->num_nodes, sv->d_blocks, sv->components, sv->format, pos,
                     uvw, sv->global_dim[0], sv->global_dim[1], sv->global_dim[2], (uint32_t)n,
                     reinterpret_cast<double4 *>(out_d), reinterpret_cast<float4 *>(out_f));
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

}  // namespace srf
