// srf_scene.hip -- two-level scenes: many triangle accelerators placed by
// transforms, a top-level BVH over their world boxes, and one kernel that
// traces rays through both levels.
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

struct Scene {
  int device;
  uint32_t count;       // elements
  uint32_t num_nodes;   // top-level nodes
  uint32_t depth;       // most branches on a root-to-leaf path
  double bmin[3], bmax[3];
  TlasNode *d_nodes;
  SceneInst *d_inst;
  uint32_t *d_counter;  // work counter of the persistent launch
  std::vector<TlasNode> h_nodes;
  int cus;              // compute units of the device
  srf_launch_config tuned;
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

}  // namespace

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
    if (a->kind != kTri)
      return set_error(SRF_ERR_UNSUPPORTED, "instance %u: scenes hold triangle accelerators", i);
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

srf_status scene_trace(const Scene *sc, const srf_ray *rays, uint64_t n, srf_hit *hits,
                       uint32_t *inst_out, int mode, hipStream_t stream) {
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
