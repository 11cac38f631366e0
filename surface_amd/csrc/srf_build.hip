// srf_build.hip -- LBVH construction on gfx950, stream-ordered, no host sync.
//
// Pipeline (one launch each unless noted), for n primitives:
//   bounds   fused min/max over every primitive's vertices, one box per
//            block; a second launch merges each mesh's block boxes and
//            derives inv/kEPS/epsScale (no device-scope fences); it also
//            clears the Morton histogram and the sort's tile counters.
//   morton   centroid -> 30-bit code + identity index, and all four 8-bit
//            digit histograms of the codes (LDS, then one global atomic per
//            bin per block).
//   sort     4 onesweep passes: LDS-staged tiles ranked with wave64 ballots,
//            tile digit totals chained by decoupled look-back, scattered
//            through LDS so each digit run leaves as one coalesced burst.
//   karras   one thread per internal node (Karras 2012): split, leaf flags,
//            range end, parent pointers.
//   classify top-down from the roots: branch or collapsed (range <= leaf
//            limit or depth > 22), the first levels in one workgroup, then
//            three levels per launch; marks leaf starts and lists each
//            launch's heads by level.
//   scan     exclusive scan of the leaf starts (3 launches).
//   emit     one thread per branch: writes the node at its depth-first
//            preorder position pos = ancestors + 2*leaves_left - right_turns,
//            which is where the reference's serial builder puts it, and the
//            leaf records of its leaf children.
//   leaves   one thread per leaf: leaf box + packed primitives, written to
//            the leaf's own slot and its parent's slot for it; leaves of more
//            than kBigLeaf primitives by whole blocks.
//   refit    bottom-up, three branch levels per launch, then the first
//            levels in one workgroup: each branch merges its two child boxes
//            into its parent's slot.
// A batch of meshes runs the same launches over all meshes at once (see the
// segment notes below).
#include <float.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "srf_internal.h"

namespace srf {
namespace {

constexpr int kBoundsBlocks = 1024;
constexpr int kSortBlock = 256;   // one thread per digit
constexpr int kSortItems = 15;    // keys per thread
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 26;
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

// (uint32_t)(float): x86-64 truncates through int64; for |x| < 2^31 and NaN
// (both give 0) this equals the 32-bit truncation.
__device__ __forceinline__ uint32_t f2u_trunc(float x) { return (uint32_t)(int)x; }

__device__ __forceinline__ uint32_t morton30(float px, float py, float pz,
                                             const BuildParams &p) {
  // MortionCode30, render/render_prefix_tree_util.h:143-157
  uint32_t ix = f2u_trunc((px - p.bmin[0]) * p.inv[0]);
  uint32_t iy = f2u_trunc((py - p.bmin[1]) * p.inv[1]);
  uint32_t iz = f2u_trunc((pz - p.bmin[2]) * p.inv[2]);
  return (part_by2_30(ix) << 2) | (part_by2_30(iy) << 1) | part_by2_30(iz);
}

__device__ __forceinline__ int clz32(uint32_t x) { return x == 0 ? 32 : __clz(x); }

// std::min / std::max as the reference calls them: keep a unless b is
// strictly smaller / larger.
__device__ __forceinline__ float fmin_ref(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmax_ref(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ float3 load_v3(const float *V, uint32_t i) {
  const float *p = V + 3 * (size_t)i;
  return make_float3(p[0], p[1], p[2]);
}

struct Box {
  float mn[3], mx[3];
};

__device__ __forceinline__ void box_add(Box &b, float x, float y, float z, float e) {
  b.mn[0] = fmin_ref(b.mn[0], x - e);
  b.mn[1] = fmin_ref(b.mn[1], y - e);
  b.mn[2] = fmin_ref(b.mn[2], z - e);
  b.mx[0] = fmax_ref(b.mx[0], x + e);
  b.mx[1] = fmax_ref(b.mx[1], y + e);
  b.mx[2] = fmax_ref(b.mx[2], z + e);
}

__device__ __forceinline__ void box_merge(Box &a, const Box &b) {
  for (int k = 0; k < 3; k++) {
    a.mn[k] = fmin_ref(a.mn[k], b.mn[k]);
    a.mx[k] = fmax_ref(a.mx[k], b.mx[k]);
  }
}

__device__ __forceinline__ void box_init(Box &b) {
  for (int k = 0; k < 3; k++) {
    b.mn[k] = FLT_MAX;
    b.mx[k] = -FLT_MAX;
  }
}

struct PrimSrc {
  const float *V;
  const uint32_t *F;
  const float *R;
  float rconst;
  uint32_t n;
  // spheres: floats from one centre (radius) to the next -- 3 (1) for the
  // caller's arrays, 4 (4) for packed (centre, radius) records
  uint32_t vstride, rstride;
  uint32_t nv;  // solids: vertices per solid (5, 6, 8)
};

__device__ __forceinline__ float3 sph_pos(const PrimSrc &s, uint32_t i) {
  const float *p = s.V + (size_t)s.vstride * i;
  return make_float3(p[0], p[1], p[2]);
}
__device__ __forceinline__ float sph_rad(const PrimSrc &s, uint32_t i) {
  return s.R ? s.R[(size_t)s.rstride * i] : s.rconst;
}

// ---- segments: many meshes in one launch sequence ----
// A build covers nseg meshes laid end to end in one primitive index space
// (codes, sort keys, Karras links, per-leaf and per-node scratch). Mesh m
// owns primitives [prim_off, prim_off + n), sort tiles [tile_off, ...),
// bounds/Morton blocks [bblk_off, ...) / [mblk_off, ...), global leaves
// [leaf_base, ...), and per-node scratch [2 prim_off, 2 prim_off + 2n - 1).
// Every kernel maps its global index to (mesh, local index) and works in
// local indices, so each mesh gets exactly the tree a build of it alone
// gives. M = false is the one-mesh build: the segment travels in the kernel
// arguments and all offsets are zero.
struct Seg {
  PrimSrc src;
  void *nodes;
  void *prims;
  uint32_t *sorted_index;
  uint32_t *sorted_faces;
  BuildParams *params;
};

struct SegView {
  Seg one;                      // M = false
  const Seg *seg;               // M = true: nseg entries
  const uint32_t *prim_off;     // nseg + 1 prefix arrays
  const uint32_t *tile_off;
  const uint32_t *bblk_off;
  const uint32_t *mblk_off;
  const uint32_t *leaf_base;    // written by k_seg_leaves
  // M = true: the mesh of every primitive-space index (written by k_morton)
  // and of every global leaf (written by k_emit), so the per-node and
  // per-leaf kernels read one word instead of searching the prefix arrays
  uint32_t *seg_of;
  uint32_t *leaf_seg;
  uint32_t nseg;
};

// largest m with off[m] <= x (off[0] = 0, off ascending)
template <bool M>
__device__ __forceinline__ uint32_t seg_find(const uint32_t *off, uint32_t nseg, uint32_t x) {
  if (!M) return 0;
  uint32_t lo = 0, hi = nseg;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <bool M>
__device__ __forceinline__ Seg seg_get(const SegView &v, uint32_t m) {
  return M ? v.seg[m] : v.one;
}

template <bool M>
__device__ __forceinline__ uint32_t seg_off(const uint32_t *off, uint32_t m) {
  return M ? off[m] : 0u;
}

// A line's box: radius max(r0, r1) around both endpoints (the cylinder and
// end discs CylinderIsect tests, render/render_accel_line.cc:488-577), each
// bound widened by 16 FLT_EPSILON of its magnitude. The reference's own
// GetBoundingBoxOfLine (:83-121) pads p0 by r0 and p1 by r1 only.
__device__ __forceinline__ void line_box(const PrimSrc &s, uint32_t i, float mn[3], float mx[3]) {
  const float kE = FLT_EPSILON * 16.0f;
  const uint32_t a = s.F[2 * (size_t)i], b = s.F[2 * (size_t)i + 1];
  const float r0 = s.R ? s.R[a] : s.rconst, r1 = s.R ? s.R[b] : s.rconst;
  const float rr = r0 > r1 ? r0 : r1;
  const float3 p = load_v3(s.V, a), q = load_v3(s.V, b);
  const float pc[3] = {p.x, p.y, p.z}, qc[3] = {q.x, q.y, q.z};
  for (int k = 0; k < 3; k++) {
    float lo = pc[k] - rr, hi = pc[k] + rr;
    const float lo2 = qc[k] - rr, hi2 = qc[k] + rr;
    if (lo2 < lo) lo = lo2;
    if (hi2 > hi) hi = hi2;
    mn[k] = lo - fabsf(lo) * kE;
    mx[k] = hi + fabsf(hi) * kE;
  }
}

// Per-primitive contribution to the scene box:
//  tri  ComputeBoundingBoxOMP, epsScale 0  render/render_accel_triangle.cc:711-797,1430-1443
//  tet  kEPS = 16*FLT_EPSILON               render/render_accel_tetra.cc:971-1055
//  sph  pos -+ r -+ 16*FLT_EPSILON          render/render_accel_particle.cc:303-378
//  sol  kEPS = 16*FLT_EPSILON               render/render_accel_solid.cc:1172-1262
template <int K>
__device__ __forceinline__ void scene_add(Box &b, const PrimSrc &s, uint32_t i) {
  if (K == kLin) {
    float mn[3], mx[3];
    line_box(s, i, mn, mx);
    for (int k = 0; k < 3; k++) {
      b.mn[k] = fmin_ref(b.mn[k], mn[k]);
      b.mx[k] = fmax_ref(b.mx[k], mx[k]);
    }
  } else if (K == kSol) {
    for (uint32_t j = 0; j < s.nv; j++) {
      float3 v = load_v3(s.V, s.F[(size_t)s.nv * i + j]);
      box_add(b, v.x, v.y, v.z, FLT_EPSILON * 16);
    }
  } else if (K == kTri) {
    for (int j = 0; j < 3; j++) {
      float3 v = load_v3(s.V, s.F[3 * (size_t)i + j]);
      box_add(b, v.x, v.y, v.z, FLT_EPSILON * 0.0f);
    }
  } else if (K == kTet) {
    for (int j = 0; j < 4; j++) {
      float3 v = load_v3(s.V, s.F[4 * (size_t)i + j]);
      box_add(b, v.x, v.y, v.z, FLT_EPSILON * 16);
    }
  } else {
    float3 p = sph_pos(s, i);
    float r = sph_rad(s, i);
    const float e = (float)(FLT_EPSILON * 16.0);
    b.mn[0] = fmin_ref(b.mn[0], (p.x - r) - e);
    b.mn[1] = fmin_ref(b.mn[1], (p.y - r) - e);
    b.mn[2] = fmin_ref(b.mn[2], (p.z - r) - e);
    b.mx[0] = fmax_ref(b.mx[0], (p.x + r) + e);
    b.mx[1] = fmax_ref(b.mx[1], (p.y + r) + e);
    b.mx[2] = fmax_ref(b.mx[2], (p.z + r) + e);
  }
}

// Triangle and tetra centroids, as the Morton codes use them, kept from the
// bounds pass so that the code pass reads 12 contiguous bytes per primitive
// instead of gathering its vertices again (scratch that the sort has not yet
// claimed; all null for the other kinds).
struct CenCache {
  float *x, *y, *z;
};

// scene_add for triangles and tetras, also giving the centroid expression of
// CalculateMortonCodes{Triangle,Tetra}Float30 (the same operations as
// prim_code, so the codes are bit-identical).
template <int K>
__device__ __forceinline__ float3 scene_add_cen(Box &b, const PrimSrc &s, uint32_t i) {
  if (K == kTri) {
    const float one_third = 1.0f / 3.0f;
    float3 a = load_v3(s.V, s.F[3 * (size_t)i]);
    float3 bb = load_v3(s.V, s.F[3 * (size_t)i + 1]);
    float3 c = load_v3(s.V, s.F[3 * (size_t)i + 2]);
    box_add(b, a.x, a.y, a.z, FLT_EPSILON * 0.0f);
    box_add(b, bb.x, bb.y, bb.z, FLT_EPSILON * 0.0f);
    box_add(b, c.x, c.y, c.z, FLT_EPSILON * 0.0f);
    return make_float3(one_third * ((a.x + bb.x) + c.x), one_third * ((a.y + bb.y) + c.y),
                       one_third * ((a.z + bb.z) + c.z));
  } else {
    const float q = 1.0f / 4.0f;
    float3 a = load_v3(s.V, s.F[4 * (size_t)i]);
    float3 bb = load_v3(s.V, s.F[4 * (size_t)i + 1]);
    float3 c = load_v3(s.V, s.F[4 * (size_t)i + 2]);
    float3 d = load_v3(s.V, s.F[4 * (size_t)i + 3]);
    box_add(b, a.x, a.y, a.z, FLT_EPSILON * 16);
    box_add(b, bb.x, bb.y, bb.z, FLT_EPSILON * 16);
    box_add(b, c.x, c.y, c.z, FLT_EPSILON * 16);
    box_add(b, d.x, d.y, d.z, FLT_EPSILON * 16);
    return make_float3(q * (((a.x + bb.x) + c.x) + d.x), q * (((a.y + bb.y) + c.y) + d.y),
                       q * (((a.z + bb.z) + c.z) + d.z));
  }
}

__device__ __forceinline__ void box_wave_reduce(Box &b) {
  for (int off = 32; off >= 1; off >>= 1) {
    Box o;
    for (int k = 0; k < 3; k++) {
      o.mn[k] = __shfl_xor(b.mn[k], off, 64);
      o.mx[k] = __shfl_xor(b.mx[k], off, 64);
    }
    box_merge(b, o);
  }
}

// The scene box's parameters: Morton scale, epsScale and kEPS; the
// counters and flags the later stages accumulate into start at zero.
template <int K>
__device__ __forceinline__ void write_params(BuildParams *params, const Box &r) {
  BuildParams p;
  for (int k = 0; k < 3; k++) {
    p.bmin[k] = r.mn[k];
    p.bmax[k] = r.mx[k];
    // kDIV / (bmax - bmin), render/render_prefix_tree_util.cc:242-245
    p.inv[k] = (float)1024 / (r.mx[k] - r.mn[k]);
  }
  // epsScale = max extent, render/render_accel_triangle.cc:1451-1463
  double bs0 = (double)(r.mx[0] - r.mn[0]);
  double bs1 = (double)(r.mx[1] - r.mn[1]);
  double bs2 = (double)(r.mx[2] - r.mn[2]);
  double es = bs0;
  if (es < bs1) es = bs1;
  if (es < bs2) es = bs2;
  p.eps_scale = es;
  if (K == kTri)
    p.kEPS = FLT_EPSILON * (float)es;  // ComputeBoundingBox30's real kEPS
  else if (K == kTet || K == kSol || K == kLin)
    p.kEPS = FLT_EPSILON * 16;
  else
    p.kEPS = (float)(FLT_EPSILON * 16.0);
  for (int k = 0; k < 3; k++) {
    p.root_bmin[k] = FLT_MAX;
    p.root_bmax[k] = -FLT_MAX;
  }
  p.num_leaves = 0;
  p.num_nodes = 0;
  p.max_steps = 0;
  p.error = 0;
  p.ticket = 0;
  p.pad[0] = p.pad[1] = p.pad[2] = 0;
  *params = p;
}

// The block's box of its mesh's primitives. A mesh reduced by one block
// writes its parameters here; a mesh of several blocks leaves per-block
// boxes for k_bounds_final (no device-scope fence or ticket: on MI355X each
// such fence writes back the L2 of its XCD).
template <int K, bool M>
__global__ void __launch_bounds__(256) k_bounds(SegView v, Box *partials,
                                               uint32_t *clear, uint32_t nclear,
                                               CenCache cen) {
  // blocks [bblk_off[m], bblk_off[m+1]) reduce mesh m
  const uint32_t m = seg_find<M>(v.bblk_off, v.nseg, blockIdx.x);
  const uint32_t b0 = seg_off<M>(v.bblk_off, m);
  const uint32_t nblk = M ? v.bblk_off[m + 1] - b0 : gridDim.x;
  const Seg sg = seg_get<M>(v, m);
  const PrimSrc &s = sg.src;
  Box b;
  box_init(b);
  if ((K == kTri || K == kTet) && cen.x) {
    const uint32_t base = seg_off<M>(v.prim_off, m);
    for (uint32_t i = (blockIdx.x - b0) * blockDim.x + threadIdx.x; i < s.n;
         i += nblk * blockDim.x) {
      const float3 c = scene_add_cen<K>(b, s, i);
      cen.x[base + i] = c.x;
      cen.y[base + i] = c.y;
      cen.z[base + i] = c.z;
    }
  } else {
    for (uint32_t i = (blockIdx.x - b0) * blockDim.x + threadIdx.x; i < s.n;
         i += nblk * blockDim.x)
      scene_add<K>(b, s, i);
  }
  box_wave_reduce(b);
  __shared__ Box wbox[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wbox[wave] = b;
  // clear the histogram / tile counters the next stages accumulate into
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nclear;
       i += gridDim.x * blockDim.x)
    clear[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    Box r = wbox[0];
    for (int w = 1; w < 4; w++) box_merge(r, wbox[w]);
    if (nblk > 1)
      partials[blockIdx.x] = r;
    else
      write_params<K>(sg.params, r);
  }
}

// One block per mesh of several bounds blocks: their boxes merged, then the
// mesh's parameters.
template <int K, bool M>
__global__ void __launch_bounds__(256) k_bounds_final(SegView v, const Box *partials,
                                                     uint32_t nblk_one) {
  const uint32_t m = blockIdx.x;
  if (M && m >= v.nseg) return;
  const uint32_t b0 = seg_off<M>(v.bblk_off, m);
  const uint32_t nblk = M ? v.bblk_off[m + 1] - b0 : nblk_one;
  if (nblk <= 1) return;
  Box b;
  box_init(b);
  for (uint32_t i = b0 + threadIdx.x; i < b0 + nblk; i += blockDim.x) box_merge(b, partials[i]);
  box_wave_reduce(b);
  __shared__ Box wbox[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wbox[wave] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    Box r = wbox[0];
    for (int w = 1; w < 4; w++) box_merge(r, wbox[w]);
    write_params<K>(seg_get<M>(v, m).params, r);
  }
}

// ---- Morton codes + digit histograms ----
// tri CalculateMortonCodesTriangleFloat30 render/render_prefix_tree_util.cc:230-265
// tet CalculateMortonCodesTetraFloat30    render/render_prefix_tree_util.cc:267-304
// sph CalculateMortonCodes30SIMD          render/render_prefix_tree_util.cc:176-228
// sol CalculateMortonCodesSolidFloat30    render/render_prefix_tree_util.cc:347-382
template <int K>
__device__ __forceinline__ uint32_t prim_code(const PrimSrc &s, uint32_t i,
                                              const BuildParams &p) {
  if (K == kLin) {
    // the segment's midpoint (p0 + p1) * 0.5f
    const float3 a = load_v3(s.V, s.F[2 * (size_t)i]), b = load_v3(s.V, s.F[2 * (size_t)i + 1]);
    return morton30((a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f, p);
  } else if (K == kSol) {
    // the vertices summed from (0, 0, 0) in index order, times 1 / nv
    const float inv_div = 1.0f / (float)s.nv;
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    for (uint32_t j = 0; j < s.nv; j++) {
      float3 a = load_v3(s.V, s.F[(size_t)s.nv * i + j]);
      cx = cx + a.x;
      cy = cy + a.y;
      cz = cz + a.z;
    }
    return morton30(cx * inv_div, cy * inv_div, cz * inv_div, p);
  } else if (K == kTri) {
    const float one_third = 1.0f / 3.0f;
    float3 a = load_v3(s.V, s.F[3 * (size_t)i]);
    float3 b = load_v3(s.V, s.F[3 * (size_t)i + 1]);
    float3 c = load_v3(s.V, s.F[3 * (size_t)i + 2]);
    return morton30(one_third * ((a.x + b.x) + c.x), one_third * ((a.y + b.y) + c.y),
                    one_third * ((a.z + b.z) + c.z), p);
  } else if (K == kTet) {
    const float q = 1.0f / 4.0f;
    float3 a = load_v3(s.V, s.F[4 * (size_t)i]);
    float3 b = load_v3(s.V, s.F[4 * (size_t)i + 1]);
    float3 c = load_v3(s.V, s.F[4 * (size_t)i + 2]);
    float3 d = load_v3(s.V, s.F[4 * (size_t)i + 3]);
    return morton30(q * (((a.x + b.x) + c.x) + d.x), q * (((a.y + b.y) + c.y) + d.y),
                    q * (((a.z + b.z) + c.z) + d.z), p);
  } else {
    float3 a = sph_pos(s, i);
    return morton30(a.x, a.y, a.z, p);
  }
}

// Counts digit d of every valid lane into the LDS histogram row h: the lanes
// that share the first valid lane's digit add with one atomic, the others
// each with their own. Neighbouring primitives (and rays in sorted order)
// share their high digits, and atomics from many lanes to one LDS address
// serialise.
__device__ __forceinline__ void hist_add(uint32_t *h, uint32_t d, bool valid) {
  const uint64_t act = __ballot(valid);
  if (act == 0) return;
  const int leader = __ffsll((long long)act) - 1;
  const uint32_t v = __shfl(d, leader, 64);
  const uint64_t same = __ballot(valid && d == v);
  if ((int)(threadIdx.x & 63u) == leader) atomicAdd(&h[v], (uint32_t)__popcll(same));
  if (valid && d != v) atomicAdd(&h[d], 1u);
}

template <int K, bool M>
__global__ void __launch_bounds__(256) k_morton(SegView v, uint32_t *codes, uint32_t *hist,
                                               CenCache cen) {
  // blocks [mblk_off[m], mblk_off[m+1]) code mesh m; histograms per mesh
  const uint32_t m = seg_find<M>(v.mblk_off, v.nseg, blockIdx.x);
  const uint32_t b0 = seg_off<M>(v.mblk_off, m);
  const uint32_t nblk = M ? v.mblk_off[m + 1] - b0 : gridDim.x;
  const Seg sg = seg_get<M>(v, m);
  const PrimSrc &s = sg.src;
  const uint32_t pbase = seg_off<M>(v.prim_off, m);
  codes += pbase;
  hist += 1024 * (size_t)m;
  const bool cached = (K == kTri || K == kTet) && cen.x;
  __shared__ uint32_t h[4][256];
  for (int i = threadIdx.x; i < 4 * 256; i += blockDim.x) (&h[0][0])[i] = 0;
  __syncthreads();
  const BuildParams p = *sg.params;
  // whole wavefronts step together (hist_add takes ballots)
  for (uint32_t i0 = (blockIdx.x - b0) * blockDim.x + (threadIdx.x & ~63u); i0 < s.n;
       i0 += nblk * blockDim.x) {
    const uint32_t i = i0 + (threadIdx.x & 63u);
    const bool valid = i < s.n;
    uint32_t c = 0;
    if (valid) {
      c = cached ? morton30(cen.x[pbase + i], cen.y[pbase + i], cen.z[pbase + i], p)
                 : prim_code<K>(s, i, p);
      codes[i] = c;
      if (M) v.seg_of[pbase + i] = m;
    }
    hist_add(h[0], c & 255, valid);
    hist_add(h[1], (c >> 8) & 255, valid);
    hist_add(h[2], (c >> 16) & 255, valid);
    hist_add(h[3], c >> 24, valid);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 4 * 256; i += blockDim.x) {
    uint32_t v = (&h[0][0])[i];
    if (v) atomicAdd(&hist[i], v);
  }
}

// Morton codes for an explicit scene box (the stage entry point).
template <int K>
__global__ void __launch_bounds__(256) k_morton_box(PrimSrc s, BuildParams p, uint32_t *codes) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n;
       i += gridDim.x * blockDim.x)
    codes[i] = prim_code<K>(s, i, p);
}

__global__ void __launch_bounds__(256) k_histogram(const uint32_t *codes, uint32_t n,
                                                  uint32_t *hist) {
  __shared__ uint32_t h[4][256];
  for (int i = threadIdx.x; i < 4 * 256; i += blockDim.x) (&h[0][0])[i] = 0;
  __syncthreads();
  for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < n;
       i0 += gridDim.x * blockDim.x) {
    const uint32_t i = i0 + (threadIdx.x & 63u);
    const bool valid = i < n;
    const uint32_t c = valid ? codes[i] : 0u;
    hist_add(h[0], c & 255, valid);
    hist_add(h[1], (c >> 8) & 255, valid);
    hist_add(h[2], (c >> 16) & 255, valid);
    hist_add(h[3], c >> 24, valid);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 4 * 256; i += blockDim.x) {
    uint32_t v = (&h[0][0])[i];
    if (v) atomicAdd(&hist[i], v);
  }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---- onesweep radix pass ----
// Block-wide exclusive scan of one value per thread (256 threads).
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t *tmp,
                                                        uint32_t *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) tmp[wave] = x;
  __syncthreads();
  uint32_t base = 0, all = 0;
  for (int w = 0; w < 4; w++) {
    uint32_t t = tmp[w];
    if (w < wave) base += t;
    all += t;
  }
  __syncthreads();
  if (total) *total = all;
  return base + x - v;
}

// One LSD pass over 8 bits of the keys. With M, tiles [tile_off[m],
// tile_off[m+1]) sort mesh m's keys among themselves (digit totals from that
// mesh's histogram, look-back stopping at its first tile); values are local
// primitive indices.
template <bool M>
__global__ void __launch_bounds__(kSortBlock)
k_onesweep(const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
           uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out, uint32_t n_all,
           int shift, int pass, const uint32_t *__restrict__ hist_all,
           uint32_t *status /* [tiles][256] */, uint32_t *tile_counter,
           uint32_t *error, SegView v) {
  __shared__ uint32_t s_keys[kSortTile];
  __shared__ uint32_t s_vals[kSortTile];
  __shared__ uint32_t s_whist[4][256];   // per-wave digit counts, then offsets
  __shared__ uint32_t s_gbase[256];      // global destination of digit run
  __shared__ uint32_t s_texcl[256];      // tile-exclusive digit prefix
  __shared__ uint32_t s_tmp[8];
  __shared__ uint32_t s_tile;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_tile = atomicAdd(tile_counter, 1u);
  for (int i = threadIdx.x; i < 4 * 256; i += kSortBlock) (&s_whist[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t m = seg_find<M>(v.tile_off, v.nseg, tile);
  const uint32_t t0 = seg_off<M>(v.tile_off, m);
  const uint32_t base = seg_off<M>(v.prim_off, m);
  const uint32_t n = M ? v.prim_off[m + 1] - base : n_all;
  const uint32_t *hist = hist_all + 1024 * (size_t)m + 256 * pass;
  keys_in += base;
  keys_out += base;
  if (vals_in) vals_in += base;
  vals_out += base;
  const uint32_t ltile = tile - t0;
  const uint32_t tile_base = ltile * kSortTile;
  const uint32_t seg = tile_base + wave * (kSortItems * 64);

  uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    uint32_t idx = seg + k * 64 + lane;
    bool ok = idx < n;
    key[k] = ok ? keys_in[idx] : 0xffffffffu;
    val[k] = ok ? (vals_in ? vals_in[idx] : idx) : 0u;
  }
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    uint32_t idx = seg + k * 64 + lane;
    bool ok = idx < n;
    uint32_t d = (key[k] >> shift) & 255u;
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t bm = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bm : ~bm;
    }
    uint32_t before = __popcll(peers & lt_mask);
    uint32_t cnt = __popcll(peers);
    uint32_t prev = ok ? s_whist[wave][d] : 0u;
    rank[k] = prev + before;
    // the highest peer publishes the new running count
    bool top = ok && (before == cnt - 1);
    if (top) s_whist[wave][d] = prev + cnt;
  }
  __syncthreads();

  // per-digit wave offsets and tile totals (thread = digit)
  const uint32_t d = threadIdx.x;
  uint32_t run = 0;
  for (int w = 0; w < 4; w++) {
    uint32_t c = s_whist[w][d];
    s_whist[w][d] = run;
    run += c;
  }
  const uint32_t tile_count = run;
  // publish this tile's aggregate, then look back for the prefix
  uint32_t *my = status + (size_t)tile * 256 + d;
  if (ltile == 0) {
    __hip_atomic_store(my, kFlagInc | tile_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_store(my, kFlagAgg | tile_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  uint32_t excl = 0;
  if (ltile > 0) {
    int32_t j = (int32_t)tile - 1;
    uint32_t spins = 0;
    while (j >= (int32_t)t0) {
      uint32_t s = __hip_atomic_load(status + (size_t)j * 256 + d, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
      uint32_t flag = s & ~kValueMask;
      if (flag == 0) {
        if (++spins > kSpinLimit) {
          atomicOr(error, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      excl += s & kValueMask;
      if (flag == kFlagInc) break;
      j--;
    }
    __hip_atomic_store(my, kFlagInc | ((excl + tile_count) & kValueMask), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  // global exclusive prefix of the digit over the whole pass
  uint32_t gpre = block_excl_scan256(hist[d], s_tmp, nullptr);
  uint32_t texcl = block_excl_scan256(tile_count, s_tmp, nullptr);
  s_gbase[d] = gpre + excl;
  s_texcl[d] = texcl;
  __syncthreads();

  // stage the tile in digit order through LDS
#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    uint32_t idx = seg + k * 64 + lane;
    if (idx < n) {
      uint32_t dd = (key[k] >> shift) & 255u;
      uint32_t pos = s_texcl[dd] + s_whist[wave][dd] + rank[k];
      s_keys[pos] = key[k];
      s_vals[pos] = val[k];
    }
  }
  __syncthreads();
  const uint32_t valid = (n - tile_base) < (uint32_t)kSortTile ? (n - tile_base) : kSortTile;
  for (uint32_t i = threadIdx.x; i < valid; i += kSortBlock) {
    uint32_t kk = s_keys[i];
    uint32_t dd = (kk >> shift) & 255u;
    uint32_t dst = s_gbase[dd] + (i - s_texcl[dd]);
    if (dst < n) {
      keys_out[dst] = kk;
      vals_out[dst] = s_vals[i];
    }
  }
}

// A batched build's meshes of one sort tile (the secondary sphere trees'
// big leaves hold about a thousand spheres each) are sorted by one workgroup
// each, all four digit passes in LDS, and own no onesweep tiles: the
// same stable LSD ranking as k_onesweep's (wave, item, lane order), so the
// permutation is identical, without four passes over mostly empty tiles.
// Keys are read from and the sorted pairs written to where the four passes
// leave them.
__global__ void __launch_bounds__(kSortBlock) k_sort_local(uint32_t *keys, uint32_t *vals,
                                                          SegView v) {
  __shared__ uint32_t s_keys[kSortTile];
  __shared__ uint32_t s_vals[kSortTile];
  __shared__ uint32_t s_whist[4][256];
  __shared__ uint32_t s_texcl[256];
  __shared__ uint32_t s_tmp[8];
  const uint32_t m = blockIdx.x;
  if (m >= v.nseg) return;
  const uint32_t base = v.prim_off[m], n = v.prim_off[m + 1] - base;
  if (n > (uint32_t)kSortTile) return;  // the onesweep passes sort this mesh
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t seg = wave * (kSortItems * 64);
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    const uint32_t idx = seg + k * 64 + lane;
    const bool ok = idx < n;
    key[k] = ok ? keys[base + idx] : 0xffffffffu;
    val[k] = ok ? idx : 0u;
  }
  for (int pass = 0; pass < 4; pass++) {
    const int shift = 8 * pass;
    for (int i = threadIdx.x; i < 4 * 256; i += kSortBlock) (&s_whist[0][0])[i] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
      const uint32_t idx = seg + k * 64 + lane;
      const bool ok = idx < n;
      const uint32_t d = (key[k] >> shift) & 255u;
      uint64_t peers = __ballot(ok);
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const uint64_t bm = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? bm : ~bm;
      }
      const uint32_t before = __popcll(peers & lt_mask), cnt = __popcll(peers);
      const uint32_t prev = ok ? s_whist[wave][d] : 0u;
      rank[k] = prev + before;
      if (ok && before == cnt - 1) s_whist[wave][d] = prev + cnt;
    }
    __syncthreads();
    // per-digit wave offsets and the digit's start in the tile (thread = digit)
    const uint32_t d = threadIdx.x;
    uint32_t run = 0;
    for (int w = 0; w < 4; w++) {
      const uint32_t c = s_whist[w][d];
      s_whist[w][d] = run;
      run += c;
    }
    s_texcl[d] = block_excl_scan256(run, s_tmp, nullptr);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
      const uint32_t idx = seg + k * 64 + lane;
      if (idx < n) {
        const uint32_t dd = (key[k] >> shift) & 255u;
        const uint32_t pos = s_texcl[dd] + s_whist[wave][dd] + rank[k];
        s_keys[pos] = key[k];
        s_vals[pos] = val[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
      const uint32_t idx = seg + k * 64 + lane;
      if (idx < n) {
        key[k] = s_keys[idx];
        val[k] = s_vals[idx];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    const uint32_t idx = seg + k * 64 + lane;
    if (idx < n) {
      keys[base + idx] = key[k];
      vals[base + idx] = val[k];
    }
  }
}

// ---- Karras radix tree ----
// Delta30, render/render_prefix_tree_util.cc:501-526
__device__ __forceinline__ int delta30(const uint32_t *c, int i, int j, int n) {
  if (j < 0 || j >= n) return -1;
  uint32_t ki = c[i], kj = c[j];
  int off = 0;
  if (ki == kj) {
    ki = (uint32_t)i;
    kj = (uint32_t)j;
    off = 32;
  }
  return clz32(ki ^ kj) + off;
}

// ConstructBinaryRadixTree30, render/render_prefix_tree_util.cc:562-649
__device__ __forceinline__ void radix_node30(const uint32_t *c, int i, int n, uint32_t &child,
                                             uint32_t &leafbits, int &jout) {
  int di = delta30(c, i, i + 1, n) - delta30(c, i, i - 1, n);
  int d = di >= 0 ? 1 : -1;
  int dmin = delta30(c, i, i - d, n);
  int lmax = 2;
  while (delta30(c, i, i + lmax * d, n) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta30(c, i, i + (l + t) * d, n) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta30(c, i, j, n);
  int s = 0;
  int k = 2;
  int t = (l + k - 1) / k;  // (int)ceil((double)l / k)
  if (t > 0) {
    for (;;) {
      if (delta30(c, i, i + (s + t) * d, n) > dnode) s += t;
      if (t == 1) break;
      k *= 2;
      t = (l + k - 1) / k;
    }
  }
  int r = i + s * d + (d < 0 ? d : 0);
  child = (uint32_t)r;
  int lo = i < j ? i : j, hi = i < j ? j : i;
  leafbits = (lo == r ? 1u : 0u) | (hi == r + 1 ? 2u : 0u);
  jout = j;
}

// Maps a global primitive-space index to (mesh, local index, mesh size).
template <bool M>
__device__ __forceinline__ void seg_local(const SegView &v, uint32_t n_all, uint32_t x,
                                          uint32_t &m, uint32_t &base, uint32_t &n) {
  m = M ? min(v.seg_of[x], v.nseg - 1) : 0u;
  base = seg_off<M>(v.prim_off, m);
  n = M ? v.prim_off[m + 1] - base : n_all;
}

// Also clears what the classification accumulates into: leaf_start and cls
// (one word per primitive) and the 64 level counters.
template <bool M>
__global__ void __launch_bounds__(256) k_karras(const uint32_t *codes, uint32_t n_all,
                                               uint32_t *kchild, uint32_t *kleaf,
                                               uint32_t *kother, uint32_t *leaf_start,
                                               uint32_t *cls, uint32_t *level_counts,
                                               SegView v) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < 64) level_counts[x] = 0;
  if (x >= n_all) return;
  leaf_start[x] = 0;
  cls[x] = 0;
  uint32_t m, base, n;
  seg_local<M>(v, n_all, x, m, base, n);
  const uint32_t i = x - base;
  if (i + 1 >= n) return;
  uint32_t child, leafbits;
  int j;
  radix_node30(codes + base, (int)i, (int)n, child, leafbits, j);
  kchild[x] = child;
  kleaf[x] = leafbits;
  kother[x] = (uint32_t)j;
}

__global__ void __launch_bounds__(256) k_karras_nodeinfo(const uint32_t *codes, uint32_t n,
                                                        srf_node_info *out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n) return;
  uint32_t child, leafbits;
  int j;
  radix_node30(codes, (int)i, (int)n, child, leafbits, j);
  out[i].child_index = child;
  out[i].left_leaf = leafbits & 1u;
  out[i].right_leaf = (leafbits >> 1) & 1u;
}

// ---- collapse: which Karras nodes become branches ----
// BuildTreeRecursive32, render/render_accel_triangle.cc:895-992: a child range
// [l, r] becomes a leaf iff it is a Karras leaf, r - l <= MAX_LEAF, or its
// depth (root's children = 0) exceeds 22. The branches are found top down,
// one launch per depth: a branch at depth L (L ancestors) records its class
// cls = 1 | L << 1 | right_turns << 8, marks where its leaf children start,
// and hands its internal children that are branches too to depth L + 1.
// Work is proportional to the branches, whatever the Karras tree's depth.
template <bool M>
__global__ void __launch_bounds__(64) k_classify_roots(SegView v, uint2 *front,
                                                      uint32_t *counts) {
  for (uint32_t m = threadIdx.x; m < v.nseg; m += blockDim.x)
    front[m] = make_uint2(seg_off<M>(v.prim_off, m), m << 8);
  if (threadIdx.x == 0) counts[0] = v.nseg;
}

// A counter another workgroup may have changed during this launch.
__device__ __forceinline__ uint32_t count_load(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Start of level `level`'s branches in the level-ordered branch list: the
// levels are stored one after another, level 0 (the roots) first.
__device__ __forceinline__ uint32_t level_start(const uint32_t *counts, uint32_t level) {
  uint32_t s = 0;
  for (uint32_t k = 0; k < level; k++) s += counts[k];
  return s;
}

constexpr uint32_t kClassifyBlock = 256;
// branch levels per classification and refit launch below the top levels
constexpr int kLevelSpan = 3;

// Levels `level` .. `level` + S - 1 in one launch: each thread classifies
// one branch of level `level` and then its branch descendants down to level
// + S - 1, and only the branches of level + S are appended (the next
// launch's heads); the levels in between need no lists of their own,
// because their branches are reached from their heads here and in the
// refit below. Each node's decision is the collapse rule (a child is a leaf
// iff it is a Karras leaf, holds at most max_leaf + 1 primitives, or lies
// below the depth cap), so the tree does not depend on how the levels are
// grouped into launches. The workgroup steps through the heads as a whole,
// so the appends take one atomic per workgroup and step (per-wavefront
// atomics on the one counter serialised the big levels).
template <bool M>
__device__ __forceinline__ void classify_node(uint32_t x, uint32_t base, uint32_t n,
                                              uint32_t level, uint32_t turns, int max_leaf,
                                              const uint32_t *kchild, const uint32_t *kleaf,
                                              const uint32_t *kother, uint32_t *cls,
                                              uint32_t *leaf_start, BuildParams *params,
                                              uint32_t &cl, uint32_t &cr, bool &bl, bool &br) {
  bl = br = false;
  cl = cr = 0;
  const uint32_t i = x - base;
  if (i + 1 >= n) {
    atomicOr(&params->error, 8u);  // a leaf handed over as a branch: never expected
    return;
  }
  cls[x] = 1u | (level << 1) | (turns << 8);
  const bool deep = level > (uint32_t)kMaxDepth32;
  const uint32_t jo = kother[x];
  const uint32_t l = i < jo ? i : jo, r = i < jo ? jo : i;
  const uint32_t g = kchild[x], lb = kleaf[x];
  const bool left_leaf = (lb & 1u) || (g - l <= (uint32_t)max_leaf) || deep;
  const bool right_leaf = (lb & 2u) || (r - (g + 1) <= (uint32_t)max_leaf) || deep;
  if (left_leaf)
    leaf_start[base + l] = 1;
  else
    bl = true, cl = base + g;
  if (right_leaf)
    leaf_start[base + g + 1] = 1;
  else
    br = true, cr = base + g + 1;
}

// the deepest branch level per mesh: one atomic per wavefront and mesh
__device__ __forceinline__ void note_depth(BuildParams *params, uint32_t level) {
  const uint32_t lane = threadIdx.x & 63u;
  const bool has = params != nullptr;
  uint64_t pending = __ballot(has);
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const uint64_t pv = (uint64_t)__shfl((long long)(uintptr_t)params, leader, 64);
    const bool mine = has && (uint64_t)(uintptr_t)params == pv;
    // the largest level among this mesh's lanes
    uint32_t top = mine ? level : 0u;
    for (int off = 32; off >= 1; off >>= 1) top = max(top, (uint32_t)__shfl_xor((int)top, off, 64));
    if ((int)lane == leader && count_load(&params->max_steps) < top)
      atomicMax(&params->max_steps, top);
    pending &= ~__ballot(mine);
  }
}

template <bool M, int S>
__global__ void __launch_bounds__(kClassifyBlock) k_classify_span(
    uint32_t n_all, int max_leaf, uint32_t level, const uint32_t *kchild, const uint32_t *kleaf,
    const uint32_t *kother, uint2 *levels, uint32_t *counts, uint32_t *cls,
    uint32_t *leaf_start, SegView v) {
  constexpr int kOut = 1 << S;  // branches of level + S one head can reach
  const uint32_t count = counts[level];
  const uint32_t in_off = level_start(counts, level);
  const uint2 *front_in = levels + in_off;
  // the levels in between keep empty lists, so level + S's list follows
  uint2 *front_out = levels + in_off + count;
  const uint32_t out_cap = n_all - in_off - count;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  __shared__ uint32_t s_tot[kClassifyBlock / 64], s_base;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t j0 = blockIdx.x * blockDim.x; j0 < count; j0 += stride) {
    const uint32_t j = j0 + threadIdx.x;
    // the current layer's branches (Karras index, turns), widest 2^S
    uint32_t lx[kOut], lt[kOut];
    bool lp[kOut];
#pragma unroll
    for (int q = 0; q < kOut; q++) lx[q] = lt[q] = 0, lp[q] = false;
    BuildParams *params = nullptr;
    uint32_t m = 0, deepest = level;
    if (j < count) {
      const uint2 e = front_in[j];
      lx[0] = e.x;
      lt[0] = e.y & 255u;
      lp[0] = true;
      m = e.y >> 8;
      const uint32_t base = seg_off<M>(v.prim_off, m);
      const uint32_t n = M ? v.prim_off[m + 1] - base : n_all;
      params = seg_get<M>(v, m).params;
#pragma unroll
      for (int d = 0; d < S; d++) {
        // layer d (level + d, at most 2^d branches) -> layer d + 1
        uint32_t nx[kOut], nt[kOut];
        bool np[kOut];
#pragma unroll
        for (int q = 0; q < kOut; q++) nx[q] = nt[q] = 0, np[q] = false;
        if (level + (uint32_t)d < (uint32_t)kMaxBranchLevels) {
#pragma unroll
          for (int q = 0; q < (1 << d); q++) {
            if (!lp[q]) continue;
            deepest = level + (uint32_t)d;
            classify_node<M>(lx[q], base, n, level + (uint32_t)d, lt[q], max_leaf, kchild,
                             kleaf, kother, cls, leaf_start, params, nx[2 * q], nx[2 * q + 1],
                             np[2 * q], np[2 * q + 1]);
            nt[2 * q] = lt[q];
            nt[2 * q + 1] = lt[q] + 1;
          }
        }
#pragma unroll
        for (int q = 0; q < kOut; q++) lx[q] = nx[q], lt[q] = nt[q], lp[q] = np[q];
      }
    }
    note_depth(params, deepest);
    uint64_t mk[kOut];
    uint32_t total = 0;
#pragma unroll
    for (int q = 0; q < kOut; q++) {
      mk[q] = __ballot(lp[q]);
      total += (uint32_t)__popcll(mk[q]);
    }
    if (lane == 0) s_tot[wave] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t sum = 0;
      for (uint32_t w = 0; w < kClassifyBlock / 64; w++) sum += s_tot[w];
      s_base = sum ? atomicAdd(&counts[level + S], sum) : 0u;
    }
    __syncthreads();
    uint32_t at = s_base;
    for (uint32_t w = 0; w < wave; w++) at += s_tot[w];
    __syncthreads();  // s_tot and s_base are rewritten by the next step
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
#pragma unroll
    for (int q = 0; q < kOut; q++) {
      const uint32_t o = at + (uint32_t)__popcll(mk[q] & below);
      if (lp[q] && o < out_cap) front_out[o] = make_uint2(lx[q], lt[q] | (m << 8));
      at += (uint32_t)__popcll(mk[q]);
    }
  }
}

// Levels 0 .. top-1 in one workgroup of 1024 threads. The host picks top so
// that none of these levels holds more than 1024 branches (level L holds at
// most nseg * 2^L): one branch per thread, the level's list kept in LDS and
// counted with LDS atomics, so a level costs one round of Karras-link loads
// instead of a launch. The lists and counts still go to global memory for
// the later levels and the refit.
template <bool M>
__global__ void __launch_bounds__(1024) k_classify_top(uint32_t n_all, int max_leaf, uint32_t top,
                                                      const uint32_t *kchild,
                                                      const uint32_t *kleaf,
                                                      const uint32_t *kother, uint2 *levels,
                                                      uint32_t *counts, uint32_t *cls,
                                                      uint32_t *leaf_start, SegView v) {
  __shared__ uint2 buf[2][1024];
  __shared__ uint32_t cnt[2];
  const uint32_t lane = threadIdx.x & 63u, j = threadIdx.x;
  // the roots (one per mesh; the host runs this kernel only for at most 1024
  // meshes), written here rather than by k_classify_roots
  uint32_t count = v.nseg;
  if (j < count) {
    const uint2 e = make_uint2(seg_off<M>(v.prim_off, j), j << 8);
    buf[0][j] = e;
    levels[j] = e;
  }
  if (j == 0) counts[0] = count;
  if (j < 2) cnt[j] = 0;
  __syncthreads();
  uint32_t in_off = 0;
  for (uint32_t level = 0; level < top; level++) {
    const int cur = level & 1, nxt = cur ^ 1;
    const bool last = level + 1 == top;
    const bool deep = level > (uint32_t)kMaxDepth32;
    const uint32_t out_cap = n_all - in_off - count;
    uint32_t push_l = 0, push_r = 0, cl = 0, cr = 0, turns = 0, m = 0;
    BuildParams *params = nullptr;
    if (j < count) {
      const uint2 e = buf[cur][j];
      const uint32_t x = e.x;
      turns = e.y & 255u;
      m = e.y >> 8;
      const uint32_t base = seg_off<M>(v.prim_off, m);
      const uint32_t n = M ? v.prim_off[m + 1] - base : n_all;
      const uint32_t i = x - base;
      params = seg_get<M>(v, m).params;
      if (i + 1 < n) {
        cls[x] = 1u | (level << 1) | (turns << 8);
        const uint32_t jo = kother[x];
        const uint32_t l = i < jo ? i : jo, r = i < jo ? jo : i;
        const uint32_t g = kchild[x], lb = kleaf[x];
        const bool left_leaf = (lb & 1u) || (g - l <= (uint32_t)max_leaf) || deep;
        const bool right_leaf = (lb & 2u) || (r - (g + 1) <= (uint32_t)max_leaf) || deep;
        if (left_leaf)
          leaf_start[base + l] = 1;
        else
          push_l = 1, cl = base + g;
        if (right_leaf)
          leaf_start[base + g + 1] = 1;
        else
          push_r = 1, cr = base + g + 1;
      } else {
        atomicOr(&params->error, 8u);  // a leaf handed over as a branch: never expected
      }
    }
    {
      const bool has = params != nullptr;
      uint64_t pending = __ballot(has);
      while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const uint64_t pv = (uint64_t)__shfl((long long)(uintptr_t)params, leader, 64);
        const bool mine = has && (uint64_t)(uintptr_t)params == pv;
        if ((int)lane == leader) atomicMax(&params->max_steps, level);
        pending &= ~__ballot(mine);
      }
    }
    const uint64_t ml = __ballot(push_l), mr = __ballot(push_r);
    const uint32_t total = (uint32_t)(__popcll(ml) + __popcll(mr));
    uint32_t base_out = 0;
    if (lane == 0 && total) base_out = atomicAdd(&cnt[nxt], total);
    base_out = __shfl(base_out, 0, 64);
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint32_t ol = base_out + (uint32_t)__popcll(ml & below);
    const uint32_t orr = base_out + (uint32_t)__popcll(ml) + (uint32_t)__popcll(mr & below);
    uint2 *front_out = levels + in_off + count;
    if (push_l) {
      const uint2 c = make_uint2(cl, turns | (m << 8));
      if (ol < out_cap) front_out[ol] = c;
      if (!last && ol < 1024u) buf[nxt][ol] = c;
    }
    if (push_r) {
      const uint2 c = make_uint2(cr, (turns + 1) | (m << 8));
      if (orr < out_cap) front_out[orr] = c;
      if (!last && orr < 1024u) buf[nxt][orr] = c;
    }
    __syncthreads();
    const uint32_t next = cnt[nxt];
    in_off += count;
    count = next;
    __syncthreads();
    if (j == 0) {
      cnt[cur] = 0;
      counts[level + 1] = next;
    }
    __syncthreads();
  }
}

// ---- exclusive scan (3 launches) ----
__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const uint32_t *in, uint32_t n,
                                                           uint32_t *partial) {
  uint32_t base = blockIdx.x * kScanTile;
  uint32_t sum = 0;
  for (int k = 0; k < kScanItems; k++) {
    uint32_t idx = base + k * kScanBlock + threadIdx.x;
    if (idx < n) sum += in[idx];
  }
  for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, 64);
  __shared__ uint32_t w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = w[0] + w[1] + w[2] + w[3];
}

__global__ void __launch_bounds__(1024) k_scan_partials(uint32_t *partial, uint32_t m) {
  // single block: exclusive scan of m partials in place
  __shared__ uint32_t s[1024];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < m; base += 1024) {
    uint32_t idx = base + threadIdx.x;
    uint32_t v = idx < m ? partial[idx] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      uint32_t t = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    uint32_t incl = s[threadIdx.x];
    if (idx < m) partial[idx] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += incl;
    __syncthreads();
  }
}

// Exclusive scan of `in` into `out`; the total goes to *total and, for a
// one-mesh build (params != null), to num_leaves / num_nodes.
__global__ void __launch_bounds__(kScanBlock) k_scan_downsweep(const uint32_t *in, uint32_t n,
                                                              const uint32_t *partial,
                                                              uint32_t *out, uint32_t *total_out,
                                                              BuildParams *params) {
  // blocked arrangement: thread t owns items [t*kScanItems, (t+1)*kScanItems)
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
  for (int k = 0; k < kScanItems; k++) {
    uint32_t idx = base + k;
    v[k] = idx < n ? in[idx] : 0;
    sum += v[k];
  }
  __shared__ uint32_t tmp[8];
  uint32_t total;
  uint32_t pre = block_excl_scan256(sum, tmp, &total) + partial[blockIdx.x];
  for (int k = 0; k < kScanItems; k++) {
    uint32_t idx = base + k;
    if (idx < n) out[idx] = pre;
    pre += v[k];
    if (idx == n - 1) {
      *total_out = pre;
      if (params) {
        params->num_leaves = pre;
        params->num_nodes = 2 * pre - 1;
      }
    }
  }
}

// k_scan_downsweep for at most kScanFuse tiles: each block adds up the
// partial sums of the tiles before it itself (at most kScanFuse / 256 loads
// per thread), so k_scan_partials is not launched.
constexpr uint32_t kScanFuse = 1024;
__global__ void __launch_bounds__(kScanBlock) k_scan_downsweep_fused(
    const uint32_t *in, uint32_t n, const uint32_t *partial, uint32_t *out,
    uint32_t *total_out, BuildParams *params) {
  __shared__ uint32_t tmp[8];
  uint32_t mine = 0, before = 0;
  for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kScanBlock) mine += partial[b];
  (void)block_excl_scan256(mine, tmp, &before);  // the block-wide total
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
  for (int k = 0; k < kScanItems; k++) {
    uint32_t idx = base + k;
    v[k] = idx < n ? in[idx] : 0;
    sum += v[k];
  }
  uint32_t total;
  uint32_t pre = block_excl_scan256(sum, tmp, &total) + before;
  for (int k = 0; k < kScanItems; k++) {
    uint32_t idx = base + k;
    if (idx < n) out[idx] = pre;
    pre += v[k];
    if (idx == n - 1) {
      *total_out = pre;
      if (params) {
        params->num_leaves = pre;
        params->num_nodes = 2 * pre - 1;
      }
    }
  }
}

// Per-mesh leaf counts of a many-mesh build: mesh m's leaves are the global
// leaves [leaf_rank[prim_off[m]], leaf_rank[prim_off[m+1]]).
__global__ void __launch_bounds__(256) k_seg_leaves(SegView v, const uint32_t *leaf_rank,
                                                   const uint32_t *total, uint32_t *leaf_base) {
  for (uint32_t m = threadIdx.x; m < v.nseg; m += blockDim.x) {
    const uint32_t b = leaf_rank[v.prim_off[m]];
    const uint32_t e = (m + 1 < v.nseg) ? leaf_rank[v.prim_off[m + 1]] : *total;
    leaf_base[m] = b;
    if (m + 1 == v.nseg) leaf_base[v.nseg] = *total;
    BuildParams *p = v.seg[m].params;
    p->num_leaves = e - b;
    p->num_nodes = 2 * (e - b) - 1;
  }
}

// ---- node emission ----
struct NodeOut {
  uint32_t *parent_link;  // per node: parent pos << 1 | side
  uint32_t *branch_pos;   // per Karras internal node that is a branch: its node
  uint32_t *leaf_pos;     // per leaf: node position
  uint32_t *leaf_first;   // per leaf: first sorted primitive
  uint32_t *leaf_count;
};

__device__ __forceinline__ int split_axis(uint32_t code) {
  // GetSplitAxis, render/render_accel_triangle.cc:854-863 (asserts compiled out)
  return (clz32(code) - 2) % 3;
}

__device__ __forceinline__ void write_meta(void *nodes, uint32_t pos, bool leaf, int axis,
                                           uint32_t d0, uint32_t d1) {
  // words 12..15 of the 64-byte node: flag, axis, data[0], data[1]
  uint4 *w = reinterpret_cast<uint4 *>(nodes) + 4 * (size_t)pos + 3;
  *w = make_uint4(leaf ? 1u : 0u, (uint32_t)axis, d0, d1);
}

template <int K, bool M>
__global__ void __launch_bounds__(256) k_emit(uint32_t n_all, int max_leaf,
                                             const uint32_t *codes, const uint32_t *kchild,
                                             const uint32_t *kleaf, const uint32_t *kother,
                                             const uint32_t *cls, const uint32_t *leaf_rank,
                                             NodeOut o, SegView v) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n_all) return;
  uint32_t c = cls[x];
  if (!(c & 1u)) return;
  uint32_t m, base, n;
  seg_local<M>(v, n_all, x, m, base, n);
  const uint32_t i = x - base;
  if (i + 1 >= n) return;
  void *nodes = seg_get<M>(v, m).nodes;
  const uint32_t lbase = seg_off<M>(v.leaf_base, m);
  uint32_t *parent_link = o.parent_link + 2 * (size_t)base;
  uint32_t steps = (c >> 1) & 127u, turns = c >> 8;
  uint32_t j = kother[x];
  uint32_t l = i < j ? i : j, r = i < j ? j : i;
  uint32_t g = kchild[x], lb = kleaf[x];
  if (g + 1 >= n) return;
  // global leaf ids; preorder positions use the mesh-local leaf ranks
  const uint32_t yl = leaf_rank[base + l], yr = leaf_rank[base + g + 1];
  uint32_t pos = steps + 2 * (yl - lbase) - turns;
  uint32_t posL = pos + 1;
  uint32_t posR = steps + 2 * (yr - lbase) - turns;
  const uint32_t max_nodes = 2 * n - 1;
  if (pos >= max_nodes || posL >= max_nodes || posR >= max_nodes) return;
  write_meta(nodes, pos, false, split_axis(codes[base + i]), posL, posR);
  o.branch_pos[x] = pos;
  parent_link[posL] = pos << 1;
  parent_link[posR] = (pos << 1) | 1u;
  bool deep = steps > (uint32_t)kMaxDepth32;
  if ((lb & 1u) || (g - l <= (uint32_t)max_leaf) || deep) {
    if (M) v.leaf_seg[yl] = m;
    o.leaf_pos[yl] = posL;
    o.leaf_first[yl] = l;
    o.leaf_count[yl] = g + 1 - l;
    write_meta(nodes, posL, true, 0, g + 1 - l, l);
  }
  if ((lb & 2u) || (r - (g + 1) <= (uint32_t)max_leaf) || deep) {
    if (M) v.leaf_seg[yr] = m;
    o.leaf_pos[yr] = posR;
    o.leaf_first[yr] = g + 1;
    o.leaf_count[yr] = r - g;
    write_meta(nodes, posR, true, 0, r - g, g + 1);
  }
}

// ---- leaves: box, packing, atomic refit ----
template <int K>
__device__ __forceinline__ void leaf_box_pack(const PrimSrc &s, const uint32_t *sidx,
                                              uint32_t first, uint32_t count, float e,
                                              void *prims, uint32_t *sorted_faces, Box &b) {
  // tri ComputeBoundingBox30 render/render_accel_triangle.cc:660-709
  // tet ComputeBoundingBox30 render/render_accel_tetra.cc:921-969
  // sph ComputeBoundingBox30 render/render_accel_particle.cc:417-462
  // sol ComputeBoundingBox30 render/render_accel_solid.cc:1120-1169
  if (K == kTri || K == kTet) {
    // four primitives at a time: their ids, vertex indices and vertices are
    // loaded in three rounds instead of three dependent loads per primitive,
    // then boxed and packed in order
    constexpr int NVT = K == kTet ? 4 : 3;
    for (uint32_t q0 = first; q0 < first + count; q0 += 4) {
      const uint32_t nq = min(4u, first + count - q0);
      uint32_t ids[4], f[4][NVT];
      float3 vv[4][NVT];
#pragma unroll
      for (int j = 0; j < 4; j++) ids[j] = (uint32_t)j < nq ? sidx[q0 + j] : 0u;
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int k = 0; k < NVT; k++)
          f[j][k] = (uint32_t)j < nq ? s.F[(size_t)NVT * ids[j] + k] : 0u;
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int k = 0; k < NVT; k++)
          vv[j][k] = (uint32_t)j < nq ? load_v3(s.V, f[j][k]) : make_float3(0.0f, 0.0f, 0.0f);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if ((uint32_t)j >= nq) break;
        const uint32_t q = q0 + j, id = ids[j];
        const float3 *v = vv[j];
        if (q == first) {
          b.mn[0] = v[0].x - e; b.mn[1] = v[0].y - e; b.mn[2] = v[0].z - e;
          b.mx[0] = v[0].x + e; b.mx[1] = v[0].y + e; b.mx[2] = v[0].z + e;
        }
        for (int k = 0; k < NVT; k++) box_add(b, v[k].x, v[k].y, v[k].z, e);
        if (K == kTri) {
          float4 *P = reinterpret_cast<float4 *>(prims) + 3 * (size_t)q;
          P[0] = make_float4(v[0].x, v[0].y, v[0].z, __uint_as_float(id));
          P[1] = make_float4(v[1].x - v[0].x, v[1].y - v[0].y, v[1].z - v[0].z, 0.0f);
          P[2] = make_float4(v[2].x - v[0].x, v[2].y - v[0].y, v[2].z - v[0].z, 0.0f);
        } else {
          float4 *P = reinterpret_cast<float4 *>(prims) + 4 * (size_t)q;
          P[0] = make_float4(v[0].x, v[0].y, v[0].z, __uint_as_float(id));
          P[1] = make_float4(v[1].x, v[1].y, v[1].z, 0.0f);
          P[2] = make_float4(v[2].x, v[2].y, v[2].z, 0.0f);
          P[3] = make_float4(v[NVT - 1].x, v[NVT - 1].y, v[NVT - 1].z, 0.0f);
          reinterpret_cast<uint4 *>(sorted_faces)[q] =
              make_uint4(f[j][0], f[j][1], f[j][2], f[j][NVT - 1]);
        }
      }
    }
    return;
  }
  if (K == kSph) {
    // eight spheres at a time: ids, then centres and radii, in two rounds
    for (uint32_t q0 = first; q0 < first + count; q0 += 8) {
      const uint32_t nq = min(8u, first + count - q0);
      uint32_t ids[8];
      float3 pp[8];
      float rr[8];
#pragma unroll
      for (int j = 0; j < 8; j++) ids[j] = (uint32_t)j < nq ? sidx[q0 + j] : 0u;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        pp[j] = (uint32_t)j < nq ? sph_pos(s, ids[j]) : make_float3(0.0f, 0.0f, 0.0f);
        rr[j] = (uint32_t)j < nq ? sph_rad(s, ids[j]) : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 8; j++) {
        if ((uint32_t)j >= nq) break;
        const uint32_t q = q0 + j;
        const float3 p = pp[j];
        const float r = rr[j];
        const float lo0 = (p.x - r) - e, lo1 = (p.y - r) - e, lo2 = (p.z - r) - e;
        const float hi0 = (p.x + r) + e, hi1 = (p.y + r) + e, hi2 = (p.z + r) + e;
        if (q == first) {
          b.mn[0] = lo0; b.mn[1] = lo1; b.mn[2] = lo2;
          b.mx[0] = hi0; b.mx[1] = hi1; b.mx[2] = hi2;
        }
        b.mn[0] = fmin_ref(b.mn[0], lo0); b.mn[1] = fmin_ref(b.mn[1], lo1);
        b.mn[2] = fmin_ref(b.mn[2], lo2);
        b.mx[0] = fmax_ref(b.mx[0], hi0); b.mx[1] = fmax_ref(b.mx[1], hi1);
        b.mx[2] = fmax_ref(b.mx[2], hi2);
        reinterpret_cast<float4 *>(prims)[q] = make_float4(p.x, p.y, p.z, r);
      }
    }
    return;
  }
  for (uint32_t q = first; q < first + count; q++) {
    uint32_t id = sidx[q];
    if (K == kLin) {
      // two records per line: (p0, id) and (p1, max radius)
      float mn[3], mx[3];
      line_box(s, id, mn, mx);
      if (q == first) {
        for (int k = 0; k < 3; k++) {
          b.mn[k] = mn[k];
          b.mx[k] = mx[k];
        }
      }
      for (int k = 0; k < 3; k++) {
        b.mn[k] = fmin_ref(b.mn[k], mn[k]);
        b.mx[k] = fmax_ref(b.mx[k], mx[k]);
      }
      const uint32_t a = s.F[2 * (size_t)id], c = s.F[2 * (size_t)id + 1];
      const float r0 = s.R ? s.R[a] : s.rconst, r1 = s.R ? s.R[c] : s.rconst;
      const float3 p = load_v3(s.V, a), pq = load_v3(s.V, c);
      float4 *P = reinterpret_cast<float4 *>(prims) + 2 * (size_t)q;
      P[0] = make_float4(p.x, p.y, p.z, __uint_as_float(id));
      P[1] = make_float4(pq.x, pq.y, pq.z, r0 > r1 ? r0 : r1);
    } else if (K == kSol) {
      // nv records per solid: vertex k in xyz, the solid's id in w of the first
      float4 *P = reinterpret_cast<float4 *>(prims) + (size_t)s.nv * q;
      for (uint32_t k = 0; k < s.nv; k++) {
        float3 v = load_v3(s.V, s.F[(size_t)s.nv * id + k]);
        if (q == first && k == 0) {
          b.mn[0] = v.x - e; b.mn[1] = v.y - e; b.mn[2] = v.z - e;
          b.mx[0] = v.x + e; b.mx[1] = v.y + e; b.mx[2] = v.z + e;
        }
        box_add(b, v.x, v.y, v.z, e);
        P[k] = make_float4(v.x, v.y, v.z, k == 0 ? __uint_as_float(id) : 0.0f);
      }
    } else {
      float3 p = sph_pos(s, id);
      float r = sph_rad(s, id);
      float lo0 = (p.x - r) - e, lo1 = (p.y - r) - e, lo2 = (p.z - r) - e;
      float hi0 = (p.x + r) + e, hi1 = (p.y + r) + e, hi2 = (p.z + r) + e;
      if (q == first) {
        b.mn[0] = lo0; b.mn[1] = lo1; b.mn[2] = lo2;
        b.mx[0] = hi0; b.mx[1] = hi1; b.mx[2] = hi2;
      }
      b.mn[0] = fmin_ref(b.mn[0], lo0); b.mn[1] = fmin_ref(b.mn[1], lo1);
      b.mn[2] = fmin_ref(b.mn[2], lo2);
      b.mx[0] = fmax_ref(b.mx[0], hi0); b.mx[1] = fmax_ref(b.mx[1], hi1);
      b.mx[2] = fmax_ref(b.mx[2], hi2);
      reinterpret_cast<float4 *>(prims)[q] = make_float4(p.x, p.y, p.z, r);
    }
  }
}

__device__ __forceinline__ void store_box_slot(float *node_words, int slot, const Box &b) {
  // TriangleNode: bmin[slot][k] at word 3*slot+k, bmax[slot][k] at 6+3*slot+k
  for (int k = 0; k < 3; k++) {
    node_words[3 * slot + k] = b.mn[k];
    node_words[6 + 3 * slot + k] = b.mx[k];
  }
}

// MergeBoundingBox(left, right), render/render_accel_triangle.cc:804-814
__device__ __forceinline__ Box merge_lr(const Box &l, const Box &r) {
  Box m;
  for (int k = 0; k < 3; k++) {
    m.mn[k] = fmin_ref(l.mn[k], r.mn[k]);
    m.mx[k] = fmax_ref(l.mx[k], r.mx[k]);
  }
  return m;
}

// One thread per leaf (global leaf id y): leaf box + packed primitives, the
// leaf's own slot 0 and its parent's slot for it. With M the sorted index
// comes from the shared scratch and is copied into the mesh's accelerator.
template <int K, bool M>
__global__ void __launch_bounds__(256) k_leaves(uint32_t n_all, const uint32_t *sidx_all,
                                               NodeOut o, const uint32_t *total_leaves,
                                               uint32_t *big, uint32_t *big_count,
                                               SegView v) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= *total_leaves || y >= n_all) return;
  const uint32_t m = M ? min(v.leaf_seg[y], v.nseg - 1) : 0u;
  const Seg sg = seg_get<M>(v, m);
  const PrimSrc &s = sg.src;
  const uint32_t base = seg_off<M>(v.prim_off, m);
  BuildParams *params = sg.params;
  const float e = params->kEPS;
  uint32_t first = o.leaf_first[y], count = o.leaf_count[y];
  if (first >= s.n || count > s.n - first) {
    atomicOr(&params->error, 2u);
    return;
  }
  if (count > kBigLeaf) {  // a whole block does this one (k_leaves_big)
    big[atomicAdd(big_count, 1u)] = y;
    return;
  }
  const uint32_t *sidx = M ? sidx_all + base : sg.sorted_index;
  if (M) {
    for (uint32_t q = first; q < first + count; q++) sg.sorted_index[q] = sidx[q];
  }
  Box b;
  box_init(b);
  leaf_box_pack<K>(s, sidx, first, count, e, sg.prims, sg.sorted_faces, b);
  uint32_t node = o.leaf_pos[y];
  float *words = reinterpret_cast<float *>(sg.nodes);
  {
    // a leaf keeps its own box in slot 0 and leaves slot 1 unused
    // (MakeLeaf32, render/render_accel_triangle.cc:865-890)
    float4 *w4 = reinterpret_cast<float4 *>(sg.nodes) + 4 * (size_t)node;
    w4[0] = make_float4(b.mn[0], b.mn[1], b.mn[2], 0.0f);
    w4[1] = make_float4(0.0f, 0.0f, b.mx[0], b.mx[1]);
    w4[2] = make_float4(b.mx[2], 0.0f, 0.0f, 0.0f);
  }
  // every branch keeps both child boxes (TriangleNode / ParticleNode layout;
  // a tetra node's own box is the one its parent keeps for it)
  if (node == 0 || node >= 2 * s.n - 1) {
    atomicOr(&params->error, 4u);
    return;
  }
  uint32_t link = o.parent_link[2 * (size_t)base + node];
  uint32_t p = link >> 1;
  if (p >= 2 * s.n - 1) {
    atomicOr(&params->error, 4u);
    return;
  }
  store_box_slot(words + 16 * (size_t)p, (int)(link & 1u), b);
}

// One primitive's packed record and its box candidates: vertices (tri 3,
// tet 4) or the sphere, each -+ e, in the order the sequential leaf box
// visits them.
template <int K>
__device__ __forceinline__ int pack_prim(const PrimSrc &s, const uint32_t *sidx, uint32_t q,
                                         float e, void *prims, uint32_t *sorted_faces,
                                         float lo[8][3], float hi[8][3]) {
  const uint32_t id = sidx[q];
  if (K == kLin) {
    line_box(s, id, lo[0], hi[0]);
    const uint32_t a = s.F[2 * (size_t)id], c = s.F[2 * (size_t)id + 1];
    const float r0 = s.R ? s.R[a] : s.rconst, r1 = s.R ? s.R[c] : s.rconst;
    const float3 p = load_v3(s.V, a), pq = load_v3(s.V, c);
    float4 *P = reinterpret_cast<float4 *>(prims) + 2 * (size_t)q;
    P[0] = make_float4(p.x, p.y, p.z, __uint_as_float(id));
    P[1] = make_float4(pq.x, pq.y, pq.z, r0 > r1 ? r0 : r1);
    return 1;
  } else if (K == kSol) {
    float4 *P = reinterpret_cast<float4 *>(prims) + (size_t)s.nv * q;
    const int nv = s.nv < 8u ? (int)s.nv : 8;
    for (int k = 0; k < nv; k++) {
      float3 v = load_v3(s.V, s.F[(size_t)s.nv * id + k]);
      P[k] = make_float4(v.x, v.y, v.z, k == 0 ? __uint_as_float(id) : 0.0f);
      lo[k][0] = v.x - e; lo[k][1] = v.y - e; lo[k][2] = v.z - e;
      hi[k][0] = v.x + e; hi[k][1] = v.y + e; hi[k][2] = v.z + e;
    }
    return nv;
  } else if (K == kTri) {
    float3 v[3];
    for (int k = 0; k < 3; k++) v[k] = load_v3(s.V, s.F[3 * (size_t)id + k]);
    float4 *P = reinterpret_cast<float4 *>(prims) + 3 * (size_t)q;
    P[0] = make_float4(v[0].x, v[0].y, v[0].z, __uint_as_float(id));
    P[1] = make_float4(v[1].x - v[0].x, v[1].y - v[0].y, v[1].z - v[0].z, 0.0f);
    P[2] = make_float4(v[2].x - v[0].x, v[2].y - v[0].y, v[2].z - v[0].z, 0.0f);
    for (int k = 0; k < 3; k++) {
      lo[k][0] = v[k].x - e; lo[k][1] = v[k].y - e; lo[k][2] = v[k].z - e;
      hi[k][0] = v[k].x + e; hi[k][1] = v[k].y + e; hi[k][2] = v[k].z + e;
    }
    return 3;
  } else if (K == kTet) {
    uint32_t f[4];
    float3 v[4];
    for (int k = 0; k < 4; k++) {
      f[k] = s.F[4 * (size_t)id + k];
      v[k] = load_v3(s.V, f[k]);
    }
    float4 *P = reinterpret_cast<float4 *>(prims) + 4 * (size_t)q;
    P[0] = make_float4(v[0].x, v[0].y, v[0].z, __uint_as_float(id));
    P[1] = make_float4(v[1].x, v[1].y, v[1].z, 0.0f);
    P[2] = make_float4(v[2].x, v[2].y, v[2].z, 0.0f);
    P[3] = make_float4(v[3].x, v[3].y, v[3].z, 0.0f);
    reinterpret_cast<uint4 *>(sorted_faces)[q] = make_uint4(f[0], f[1], f[2], f[3]);
    for (int k = 0; k < 4; k++) {
      lo[k][0] = v[k].x - e; lo[k][1] = v[k].y - e; lo[k][2] = v[k].z - e;
      hi[k][0] = v[k].x + e; hi[k][1] = v[k].y + e; hi[k][2] = v[k].z + e;
    }
    return 4;
  } else {
    const float3 p = sph_pos(s, id);
    const float r = sph_rad(s, id);
    reinterpret_cast<float4 *>(prims)[q] = make_float4(p.x, p.y, p.z, r);
    lo[0][0] = (p.x - r) - e; lo[0][1] = (p.y - r) - e; lo[0][2] = (p.z - r) - e;
    hi[0][0] = (p.x + r) + e; hi[0][1] = (p.y + r) + e; hi[0][2] = (p.z + r) + e;
    return 1;
  }
}

// A running extreme that remembers where it was seen: the sequential box
// keeps the first candidate on ties (std::min / std::max), so a parallel
// reduction breaks ties by the lower candidate index. Only +0 / -0 can tie
// with different bits.
struct Ext {
  float v;
  uint32_t at;
};
__device__ __forceinline__ Ext ext_min(Ext a, Ext b) {
  return (b.v < a.v || (b.v == a.v && b.at < a.at)) ? b : a;
}
__device__ __forceinline__ Ext ext_max(Ext a, Ext b) {
  return (b.v > a.v || (b.v == a.v && b.at < a.at)) ? b : a;
}

// One block per big leaf: primitives packed in parallel, the leaf box
// reduced with the sequential tie rule, then the leaf's own slot and its
// parent's slot written as k_leaves does for small leaves.
template <int K, bool M>
__global__ void __launch_bounds__(256) k_leaves_big(const uint32_t *sidx_all, NodeOut o,
                                                   const uint32_t *big, const uint32_t *big_count,
                                                   SegView v) {
  __shared__ Ext red[6][4];
  const uint32_t nbig = *big_count;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t j = blockIdx.x; j < nbig; j += gridDim.x) {
    const uint32_t y = big[j];
    const uint32_t m = M ? min(v.leaf_seg[y], v.nseg - 1) : 0u;
    const Seg sg = seg_get<M>(v, m);
    const PrimSrc &s = sg.src;
    const uint32_t base = seg_off<M>(v.prim_off, m);
    const float e = sg.params->kEPS;
    const uint32_t first = o.leaf_first[y], count = o.leaf_count[y];
    const uint32_t *sidx = M ? sidx_all + base : sg.sorted_index;
    Ext ex[6];
    for (int c = 0; c < 3; c++) {
      ex[c] = Ext{FLT_MAX, 0xffffffffu};
      ex[3 + c] = Ext{-FLT_MAX, 0xffffffffu};
    }
    for (uint32_t q = first + threadIdx.x; q < first + count; q += blockDim.x) {
      if (M) sg.sorted_index[q] = sidx[q];
      float lo[8][3], hi[8][3];
      const int nc = pack_prim<K>(s, sidx, q, e, sg.prims, sg.sorted_faces, lo, hi);
      for (int k = 0; k < nc; k++) {
        const uint32_t at = (q - first) * 8u + (uint32_t)k;
        for (int c = 0; c < 3; c++) {
          ex[c] = ext_min(ex[c], Ext{lo[k][c], at});
          ex[3 + c] = ext_max(ex[3 + c], Ext{hi[k][c], at});
        }
      }
    }
    for (int c = 0; c < 6; c++) {
      for (int off = 32; off >= 1; off >>= 1) {
        Ext o2{__shfl_xor(ex[c].v, off, 64), (uint32_t)__shfl_xor((int)ex[c].at, off, 64)};
        ex[c] = c < 3 ? ext_min(ex[c], o2) : ext_max(ex[c], o2);
      }
      if (lane == 0) red[c][wave] = ex[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      Box b;
      for (int c = 0; c < 3; c++) {
        Ext mn = red[c][0], mx = red[3 + c][0];
        for (int w = 1; w < (int)(blockDim.x / 64); w++) {
          mn = ext_min(mn, red[c][w]);
          mx = ext_max(mx, red[3 + c][w]);
        }
        b.mn[c] = mn.v;
        b.mx[c] = mx.v;
      }
      const uint32_t node = o.leaf_pos[y];
      float4 *w4 = reinterpret_cast<float4 *>(sg.nodes) + 4 * (size_t)node;
      w4[0] = make_float4(b.mn[0], b.mn[1], b.mn[2], 0.0f);
      w4[1] = make_float4(0.0f, 0.0f, b.mx[0], b.mx[1]);
      w4[2] = make_float4(b.mx[2], 0.0f, 0.0f, 0.0f);
      const uint32_t link = o.parent_link[2 * (size_t)base + node];
      const uint32_t p = link >> 1;
      if (node == 0 || node >= 2 * s.n - 1 || p >= 2 * s.n - 1)
        atomicOr(&sg.params->error, 4u);
      else
        store_box_slot(reinterpret_cast<float *>(sg.nodes) + 16 * (size_t)p, (int)(link & 1u), b);
    }
    __syncthreads();
  }
}

// The big leaves of a one-mesh sphere build in chunks of kBigChunk spheres:
// the leaves of depth-capped clusters hold up to 10^5 spheres, and one block
// per leaf left the whole launch waiting on the largest. k_big_chunks
// numbers the (leaf, chunk) units (one workgroup scans the chunk counts),
// k_leaves_big_part packs each unit's spheres and reduces its candidates to
// (value, index) extremes, and k_leaves_big_combine reduces each leaf's
// units. Ext's order is total on (value, index), so the extremes, and with
// them the box bits, are the single block's whatever the grouping.
constexpr uint32_t kBigChunk = 2048;

__global__ void __launch_bounds__(1024) k_big_chunks(const uint32_t *big,
                                                     const uint32_t *big_count, NodeOut o,
                                                     uint32_t *unit_off) {
  __shared__ uint32_t s[1024];
  __shared__ uint32_t carry;
  const uint32_t nbig = *big_count;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < nbig; b0 += 1024) {
    const uint32_t j = b0 + threadIdx.x;
    const uint32_t c = j < nbig ? (o.leaf_count[big[j]] + kBigChunk - 1) / kBigChunk : 0u;
    s[threadIdx.x] = c;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {  // inclusive scan in LDS
      const uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (j < nbig) unit_off[j] = carry + s[threadIdx.x] - c;
    __syncthreads();
    if (threadIdx.x == 1023) carry += s[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) unit_off[nbig] = carry;
}

template <int K>
__global__ void __launch_bounds__(256) k_leaves_big_part(const uint32_t *big,
                                                        const uint32_t *big_count, NodeOut o,
                                                        const uint32_t *unit_off, Ext *part,
                                                        SegView v) {
  __shared__ Ext red[6][4];
  const uint32_t nbig = *big_count, units = unit_off[nbig];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const Seg sg = seg_get<false>(v, 0);
  const PrimSrc &s = sg.src;
  const float e = sg.params->kEPS;
  for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
    // the leaf whose units hold u: the last j with unit_off[j] <= u
    uint32_t lo = 0, hi = nbig;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (unit_off[mid] <= u) lo = mid; else hi = mid;
    }
    const uint32_t y = big[lo];
    const uint32_t first = o.leaf_first[y], count = o.leaf_count[y];
    const uint32_t c0 = (u - unit_off[lo]) * kBigChunk;
    const uint32_t c1 = min(c0 + kBigChunk, count);
    Ext ex[6];
    for (int c = 0; c < 3; c++) {
      ex[c] = Ext{FLT_MAX, 0xffffffffu};
      ex[3 + c] = Ext{-FLT_MAX, 0xffffffffu};
    }
    for (uint32_t q = first + c0 + threadIdx.x; q < first + c1; q += blockDim.x) {
      float lo3[8][3], hi3[8][3];
      const int nc = pack_prim<K>(s, sg.sorted_index, q, e, sg.prims, sg.sorted_faces, lo3, hi3);
      for (int k = 0; k < nc; k++) {
        const uint32_t at = (q - first) * 8u + (uint32_t)k;
        for (int c = 0; c < 3; c++) {
          ex[c] = ext_min(ex[c], Ext{lo3[k][c], at});
          ex[3 + c] = ext_max(ex[3 + c], Ext{hi3[k][c], at});
        }
      }
    }
    for (int c = 0; c < 6; c++) {
      for (int off = 32; off >= 1; off >>= 1) {
        Ext o2{__shfl_xor(ex[c].v, off, 64), (uint32_t)__shfl_xor((int)ex[c].at, off, 64)};
        ex[c] = c < 3 ? ext_min(ex[c], o2) : ext_max(ex[c], o2);
      }
      if (lane == 0) red[c][wave] = ex[c];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
      const int c = threadIdx.x;
      Ext r = red[c][0];
      for (int w = 1; w < (int)(blockDim.x / 64); w++)
        r = c < 3 ? ext_min(r, red[c][w]) : ext_max(r, red[c][w]);
      part[6 * (size_t)u + c] = r;
    }
    __syncthreads();
  }
}

// One thread per big leaf: its units' extremes, then the leaf's own slot and
// its parent's slot (as k_leaves_big writes them).
__global__ void __launch_bounds__(256) k_leaves_big_combine(const uint32_t *big,
                                                           const uint32_t *big_count,
                                                           NodeOut o, const uint32_t *unit_off,
                                                           const Ext *part, SegView v) {
  const uint32_t nbig = *big_count;
  const Seg sg = seg_get<false>(v, 0);
  const PrimSrc &s = sg.src;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nbig;
       j += gridDim.x * blockDim.x) {
    const uint32_t y = big[j];
    Ext ex[6];
    for (int c = 0; c < 3; c++) {
      ex[c] = Ext{FLT_MAX, 0xffffffffu};
      ex[3 + c] = Ext{-FLT_MAX, 0xffffffffu};
    }
    for (uint32_t u = unit_off[j]; u < unit_off[j + 1]; u++)
      for (int c = 0; c < 6; c++)
        ex[c] = c < 3 ? ext_min(ex[c], part[6 * (size_t)u + c])
                      : ext_max(ex[c], part[6 * (size_t)u + c]);
    Box b;
    for (int c = 0; c < 3; c++) {
      b.mn[c] = ex[c].v;
      b.mx[c] = ex[3 + c].v;
    }
    const uint32_t node = o.leaf_pos[y];
    if (node == 0 || node >= 2 * s.n - 1) {
      atomicOr(&sg.params->error, 4u);
      continue;
    }
    float4 *w4 = reinterpret_cast<float4 *>(sg.nodes) + 4 * (size_t)node;
    w4[0] = make_float4(b.mn[0], b.mn[1], b.mn[2], 0.0f);
    w4[1] = make_float4(0.0f, 0.0f, b.mx[0], b.mx[1]);
    w4[2] = make_float4(b.mx[2], 0.0f, 0.0f, 0.0f);
    const uint32_t link = o.parent_link[node];
    const uint32_t p = link >> 1;
    if (p >= 2 * s.n - 1)
      atomicOr(&sg.params->error, 4u);
    else
      store_box_slot(reinterpret_cast<float *>(sg.nodes) + 16 * (size_t)p, (int)(link & 1u), b);
  }
}

// The big leaves of a one-mesh sphere build as (node, first, count, 0); the
// count goes to out_count (the host checks it against cap).
__global__ void __launch_bounds__(256) k_big_export(const uint32_t *big,
                                                   const uint32_t *big_count, NodeOut o,
                                                   uint4 *out, uint32_t *out_count,
                                                   uint32_t cap) {
  const uint32_t nbig = *big_count;
  const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t j = i0; j < nbig && j < cap; j += gridDim.x * blockDim.x) {
    const uint32_t y = big[j];
    out[j] = make_uint4(o.leaf_pos[y], o.leaf_first[y], o.leaf_count[y], 0u);
  }
  if (i0 == 0) *out_count = nbig;
}

// Links each big sphere leaf to its secondary tree: words 3 and 4 of the
// leaf (its unused second box slot) take the offsets of the tree's nodes
// and spheres; the secondary builds' error flags join the main one.
__global__ void __launch_bounds__(256) k_sub_link(float4 *nodes, uint32_t nnodes,
                                                 const uint4 *link, uint32_t nbig,
                                                 const BuildParams *sub_params,
                                                 BuildParams *params) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nbig;
       j += gridDim.x * blockDim.x) {
    const uint4 l = link[j];
    if (sub_params[j].error) atomicOr(&params->error, 64u);
    if (l.x >= nnodes) {
      atomicOr(&params->error, 32u);
      continue;
    }
    float *w = reinterpret_cast<float *>(nodes + 4 * (size_t)l.x);
    w[3] = __uint_as_float(l.y);
    w[4] = __uint_as_float(l.z);
  }
}

// The bottom-up refit: each branch merges its two slots,
// MergeBoundingBox(left, right) (render/render_accel_triangle.cc:804-814),
// into its parent's slot; the root's merge is BoundingBox(). Levels
// `level` + S - 1 .. `level` in one launch (the counterpart of
// k_classify_span): each thread takes one branch of level `level`, loads
// its branch descendants down to level + S - 1 (their own slots below that
// are complete: leaves, or branches of level + S refit by the previous
// launch), merges them bottom-up into their parents' slots, in registers
// and in memory, and finally merges the head into its parent's slot, or the
// root box. The merges and their order (left, right) are the same whichever
// launch does them, so the boxes do not depend on the grouping.
__device__ __forceinline__ void load_slots(const float4 *w4, Box &l, Box &r, uint4 &meta) {
  const float4 a = w4[0], b = w4[1], c = w4[2];
  meta = reinterpret_cast<const uint4 *>(w4)[3];
  l.mn[0] = a.x; l.mn[1] = a.y; l.mn[2] = a.z;
  r.mn[0] = a.w; r.mn[1] = b.x; r.mn[2] = b.y;
  l.mx[0] = b.z; l.mx[1] = b.w; l.mx[2] = c.x;
  r.mx[0] = c.y; r.mx[1] = c.z; r.mx[2] = c.w;
}

template <bool M, int S>
__global__ void __launch_bounds__(256) k_refit_span(uint32_t n_all, uint32_t level,
                                                   const uint2 *levels, const uint32_t *counts,
                                                   NodeOut o, SegView v) {
  constexpr int kNodes = (1 << S) - 1;  // the head's subtree down to level + S - 1
  const uint32_t count = counts[level];
  const uint2 *list = levels + level_start(counts, level);
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < count;
       j += gridDim.x * blockDim.x) {
    const uint2 e = list[j];
    const uint32_t x = e.x, m = e.y >> 8;
    if (x >= n_all || (M && m >= v.nseg)) continue;
    const uint32_t base = seg_off<M>(v.prim_off, m);
    const uint32_t n = M ? v.prim_off[m + 1] - base : n_all;
    if (x < base || x - base + 1 >= n) continue;
    const Seg sg = seg_get<M>(v, m);
    const uint32_t pos = o.branch_pos[x];
    const uint32_t max_nodes = 2 * n - 1;
    if (pos >= max_nodes) {
      atomicOr(&sg.params->error, 16u);
      continue;
    }
    // the subtree in heap order (node k's children 2k + 1, 2k + 2): node
    // positions, whether each is a branch, and its two slots
    uint32_t np[kNodes];
    bool br[kNodes];
    Box sl[kNodes][2];
    np[0] = pos;
    br[0] = true;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < kNodes; k++) {
      if (k > 0) br[k] = br[(k - 1) >> 1];  // a child of a branch: load it
      if (!br[k]) continue;
      uint4 meta;
      load_slots(reinterpret_cast<const float4 *>(sg.nodes) + 4 * (size_t)np[k], sl[k][0],
                 sl[k][1], meta);
      if (k > 0 && meta.x != 0) {  // a leaf: its box is already in its parent's slot
        br[k] = false;
        continue;
      }
      if (2 * k + 2 < kNodes) {
#pragma unroll
        for (int side = 0; side < 2; side++) {
          const uint32_t cp = side ? meta.w : meta.z;
          if (cp >= max_nodes) bad = true;
          np[2 * k + 1 + side] = cp < max_nodes ? cp : 0u;
        }
      }
    }
    if (bad) {
      atomicOr(&sg.params->error, 16u);
      continue;
    }
    // bottom-up: each branch below the head merges its slots into its
    // parent's slot for it (in registers and in memory)
#pragma unroll
    for (int k = kNodes - 1; k > 0; k--) {
      if (!br[k]) continue;
      const int par = (k - 1) >> 1, side = (k - 1) & 1;
      sl[par][side] = merge_lr(sl[k][0], sl[k][1]);
      store_box_slot(reinterpret_cast<float *>(sg.nodes) + 16 * (size_t)np[par], side,
                     sl[par][side]);
    }
    const Box mb = merge_lr(sl[0][0], sl[0][1]);
    if (pos == 0) {
      for (int k = 0; k < 3; k++) {
        sg.params->root_bmin[k] = mb.mn[k];
        sg.params->root_bmax[k] = mb.mx[k];
      }
      continue;
    }
    const uint32_t link = o.parent_link[2 * (size_t)base + pos];
    const uint32_t p = link >> 1;
    if (p >= max_nodes) {
      atomicOr(&sg.params->error, 16u);
      continue;
    }
    store_box_slot(reinterpret_cast<float *>(sg.nodes) + 16 * (size_t)p, (int)(link & 1u), mb);
  }
}

// Levels top-1 .. 0 in one workgroup (the counterpart of k_classify_top):
// every branch's node position and parent link are loaded into LDS first,
// then each level costs one read of the branch's node and one store into
// its parent; the barrier orders a level's stores before the next level's
// reads.
template <bool M>
__global__ void __launch_bounds__(1024) k_refit_top(uint32_t n_all, uint32_t top,
                                                   const uint2 *levels, const uint32_t *counts,
                                                   NodeOut o, SegView v) {
  __shared__ uint32_t s_pos[2048], s_link[2048], s_m[2048];
  uint32_t total = 0;
  for (uint32_t l = 0; l < top; l++) total += counts[l];
  if (total > 2048u) total = 2048u;  // top is chosen so that this never clips
  for (uint32_t k = threadIdx.x; k < total; k += blockDim.x) {
    const uint2 e = levels[k];
    const uint32_t x = e.x, m = e.y >> 8;
    uint32_t pos = 0xffffffffu, link = 0xffffffffu;
    if (x < n_all && !(M && m >= v.nseg)) {
      const uint32_t base = seg_off<M>(v.prim_off, m);
      const uint32_t n = M ? v.prim_off[m + 1] - base : n_all;
      if (x >= base && x - base + 1 < n) {
        const Seg sg = seg_get<M>(v, m);
        pos = o.branch_pos[x];
        if (pos >= 2 * n - 1) {
          atomicOr(&sg.params->error, 16u);
          pos = 0xffffffffu;
        } else if (pos != 0) {
          link = o.parent_link[2 * (size_t)base + pos];
          if ((link >> 1) >= 2 * n - 1) {
            atomicOr(&sg.params->error, 16u);
            pos = 0xffffffffu;
          }
        }
      }
    }
    s_pos[k] = pos;
    s_link[k] = link;
    s_m[k] = m;
  }
  __syncthreads();
  uint32_t end = total;
  for (int level = (int)top - 1; level >= 0; level--) {
    const uint32_t count = counts[level];
    const uint32_t begin = end - count;
    for (uint32_t k = begin + threadIdx.x; k < end; k += blockDim.x) {
      const uint32_t pos = s_pos[k];
      if (pos == 0xffffffffu) continue;
      const Seg sg = seg_get<M>(v, s_m[k]);
      const float4 *w4 = reinterpret_cast<const float4 *>(sg.nodes) + 4 * (size_t)pos;
      const float4 a = w4[0], b = w4[1], cc = w4[2];
      Box l, r;
      l.mn[0] = a.x; l.mn[1] = a.y; l.mn[2] = a.z;
      r.mn[0] = a.w; r.mn[1] = b.x; r.mn[2] = b.y;
      l.mx[0] = b.z; l.mx[1] = b.w; l.mx[2] = cc.x;
      r.mx[0] = cc.y; r.mx[1] = cc.z; r.mx[2] = cc.w;
      const Box mb = merge_lr(l, r);
      if (pos == 0) {
        for (int q = 0; q < 3; q++) {
          sg.params->root_bmin[q] = mb.mn[q];
          sg.params->root_bmax[q] = mb.mx[q];
        }
        continue;
      }
      const uint32_t link = s_link[k];
      store_box_slot(reinterpret_cast<float *>(sg.nodes) + 16 * (size_t)(link >> 1),
                     (int)(link & 1u), mb);
    }
    end = begin;
    __syncthreads();
  }
}

// Levels 0 .. top-1 hold at most nseg * 2^level branches each (a branch has
// two children); those of up to 1024 branches (one per thread of a
// 1024-thread workgroup; 2047 in all) are classified and refitted by a
// single workgroup.
inline uint32_t top_levels(uint32_t nseg) {
  uint32_t top = 0;
  while (top < (uint32_t)kMaxBranchLevels && ((uint64_t)nseg << top) <= 1024u) top++;
  return top;
}

inline uint32_t blocks_for(uint64_t n, int block, uint32_t cap = 0x7fffffffu) {
  uint64_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  return (uint32_t)(b < cap ? b : cap);
}

PrimSrc make_src(Kind kind, const void *desc) {
  PrimSrc s{};
  if (kind == kTri) {
    const srf_tri_desc *d = (const srf_tri_desc *)desc;
    s.V = d->vertices;
    s.F = d->faces;
    s.n = (uint32_t)d->num_faces;
  } else if (kind == kTet) {
    const srf_tet_desc *d = (const srf_tet_desc *)desc;
    s.V = d->vertices;
    s.F = d->indices;
    s.n = (uint32_t)d->num_tetras;
  } else if (kind == kSol) {
    const srf_solid_desc *d = (const srf_solid_desc *)desc;
    s.V = d->vertices;
    s.F = d->indices;
    s.n = (uint32_t)d->num_solids;
    s.nv = d->vertices_per_solid;
  } else if (kind == kLin) {
    const srf_line_desc *d = (const srf_line_desc *)desc;
    s.V = d->positions;
    s.F = d->segments;
    s.R = d->radius;
    s.rconst = d->constant_radius;
    s.n = (uint32_t)d->num_lines;
  } else {
    const srf_sphere_desc *d = (const srf_sphere_desc *)desc;
    s.V = d->positions;
    s.R = d->radius;
    s.rconst = d->constant_radius;
    s.n = (uint32_t)d->num_particles;
  }
  s.vstride = 3;
  s.rstride = 1;
  return s;
}

uint32_t sort_tiles(uint32_t n) { return (n + kSortTile - 1) / kSortTile; }

struct SortScratch {
  uint32_t *keys_b, *vals_b, *vals_a, *status, *counters, *hist;
};

// Runs the 4 onesweep passes over `tiles` tiles; keys start in keys_a (vals
// implicit local identity), the sorted pairs end in (keys_a, vals_a).
// hist/counters/status must be zero before the histogram was accumulated.
template <bool M>
srf_status run_sort_passes(uint32_t *keys_a, uint32_t n, uint32_t tiles, const SortScratch &ss,
                           const SegView &v, hipStream_t stream) {
  if (tiles == 0) return SRF_OK;  // a batch of single-tile meshes (k_sort_local)
  uint32_t *src_k = keys_a, *src_v = nullptr, *dst_k = ss.keys_b, *dst_v = ss.vals_b;
  for (int pass = 0; pass < 4; pass++) {
    hipLaunchKernelGGL(k_onesweep<M>, dim3(tiles), dim3(kSortBlock), 0, stream, src_k, src_v,
                       dst_k, dst_v, n, pass * 8, pass, ss.hist,
                       ss.status + (size_t)pass * tiles * 256, ss.counters + pass,
                       ss.counters + 4, v);
    SRF_HIP(hipGetLastError());
    uint32_t *tk = src_k, *tv = src_v;
    src_k = dst_k;
    src_v = dst_v;
    dst_k = tk;
    dst_v = (pass == 0) ? ss.vals_a : tv;
  }
  return SRF_OK;
}

// Layout of one build's scratch (one-mesh or many-mesh) in the context arena.
struct BuildScratch {
  uint32_t *codes;
  SortScratch ss;
  uint32_t *kchild, *kleaf, *kother, *cls, *leaf_start, *leaf_rank, *partial;
  uint2 *levels;
  uint32_t *level_counts;  // [0, 25): branches per depth; [32]: big leaves
  uint32_t *big;
  uint32_t *total;
  NodeOut o;
  Box *partials;
  uint32_t *big_units;  // big leaves' unit offsets (one-mesh sphere builds)
  void *big_part;       // per unit: 6 (value, index) extremes
  // many-mesh only
  Seg *seg;
  uint32_t *offs;       // prim_off, tile_off, bblk_off, mblk_off (nseg + 1 each)
  uint32_t *leaf_base;  // nseg + 1
  uint32_t *sidx;       // sorted local indices before they move to each mesh
  uint32_t *seg_of;     // n: mesh of each primitive-space index
  uint32_t *leaf_seg;   // n: mesh of each global leaf
};

// Takes every slice from the arena; returns the bytes used (the arena may be
// a null sizing pass).
size_t layout_build(Arena &ar, uint32_t n, uint32_t tiles, uint32_t bblocks, uint32_t nseg,
                    bool multi, BuildScratch &b) {
  b.codes = ar.take<uint32_t>(n);
  b.ss.keys_b = ar.take<uint32_t>(n);
  b.ss.vals_b = ar.take<uint32_t>(n);
  b.ss.vals_a = nullptr;
  b.ss.status = ar.take<uint32_t>((size_t)4 * tiles * 256);
  b.ss.counters = ar.take<uint32_t>(8);
  b.ss.hist = ar.take<uint32_t>((size_t)1024 * nseg);
  b.kchild = ar.take<uint32_t>(n);
  b.kleaf = ar.take<uint32_t>(n);
  b.kother = ar.take<uint32_t>(n);
  b.levels = ar.take<uint2>(n);
  b.level_counts = ar.take<uint32_t>(64);
  b.big = ar.take<uint32_t>(n);
  b.cls = ar.take<uint32_t>(n);
  b.leaf_start = ar.take<uint32_t>(n);
  b.leaf_rank = ar.take<uint32_t>(n);
  b.partial = ar.take<uint32_t>(blocks_for(n, kScanTile) + 1);
  b.total = ar.take<uint32_t>(64);
  b.o.parent_link = ar.take<uint32_t>(2 * (size_t)n);
  b.o.branch_pos = ar.take<uint32_t>(n);
  b.o.leaf_pos = ar.take<uint32_t>(n);
  b.o.leaf_first = ar.take<uint32_t>(n);
  b.o.leaf_count = ar.take<uint32_t>(n);
  b.partials = ar.take<Box>(bblocks);
  // a big leaf holds more than kBigLeaf primitives, a unit at most kBigChunk
  const size_t max_big = (size_t)n / (kBigLeaf + 1) + 2;
  b.big_units = ar.take<uint32_t>(max_big + 1);
  b.big_part = ar.take<uint64_t>(6 * (max_big + (size_t)n / kBigChunk + 2));
  b.seg = nullptr;
  b.offs = b.leaf_base = b.sidx = b.seg_of = b.leaf_seg = nullptr;
  if (multi) {
    b.seg = ar.take<Seg>(nseg);
    b.offs = ar.take<uint32_t>(4 * ((size_t)nseg + 1));
    b.leaf_base = ar.take<uint32_t>((size_t)nseg + 1);
    b.sidx = ar.take<uint32_t>(n);
    b.seg_of = ar.take<uint32_t>(n);
    b.leaf_seg = ar.take<uint32_t>(n);
  }
  return ar.off;
}

// ------------------------------------------------------------------ host

// Builds one tree per segment (segs[m].src into segs[m]'s buffers) with one
// launch per stage. One segment is the plain Build32; more upload the
// segment table first (through the context's pinned staging buffer) and
// give every mesh its own tree. With big_out (one-segment sphere builds)
// the leaves of more than kBigLeaf spheres are listed in
// big_out->big_info / big_count.
srf_status build_segs(Ctx *ctx, Kind kind, std::vector<Seg> &segs, hipStream_t stream,
                      Accel *big_out) {
  const uint32_t nseg = (uint32_t)segs.size();
  if (nseg == 0) return SRF_OK;
  const bool multi = nseg > 1;
  std::vector<uint32_t> offs(4 * ((size_t)nseg + 1), 0u);
  uint32_t *prim_off = offs.data(), *tile_off = prim_off + nseg + 1,
           *bblk_off = tile_off + nseg + 1, *mblk_off = bblk_off + nseg + 1;
  uint64_t total = 0;
  for (uint32_t m = 0; m < nseg; m++) {
    const uint32_t n = segs[m].src.n;
    total += n;
    if (total >= (1ull << 30))
      return set_error(SRF_ERR_UNSUPPORTED, "batch of %llu primitives exceeds 2^30",
                       (unsigned long long)total);
    prim_off[m + 1] = prim_off[m] + n;
    // a batch's meshes of one tile are sorted by k_sort_local and own no
    // onesweep tiles
    tile_off[m + 1] = tile_off[m] + ((multi && n <= (uint32_t)kSortTile) ? 0u : sort_tiles(n));
    // batches: a block per 4096 primitives, so most small meshes take one
    bblk_off[m + 1] = bblk_off[m] + blocks_for(n, multi ? 256 * 16 : 256, kBoundsBlocks);
    mblk_off[m + 1] = mblk_off[m] + blocks_for(n, 256, 1024);
  }
  const uint32_t N = prim_off[nseg], tiles = tile_off[nseg];
  const uint32_t bblocks = bblk_off[nseg], mblocks = mblk_off[nseg];
  BuildScratch b;
  Arena sizing{nullptr, 0, 0};
  const size_t need = layout_build(sizing, N, tiles, bblocks, nseg, multi, b) + 256;
  srf_status st = ensure_scratch(ctx, need, stream);
  if (st != SRF_OK) return st;
  Arena ar{(char *)ctx->scratch, ctx->scratch_bytes, 0};
  layout_build(ar, N, tiles, bblocks, nseg, multi, b);
  if (ar.off > ar.size) return set_error(SRF_ERR_INTERNAL, "scratch layout overflow");

  SegView v{};
  v.nseg = nseg;
  v.one = segs[0];
  if (multi) {
    // table bytes: Seg[nseg] then the four prefix arrays
    const size_t seg_bytes = sizeof(Seg) * nseg, off_bytes = 4 * offs.size();
    st = ensure_host_stage(ctx, seg_bytes + off_bytes);
    if (st != SRF_OK) return st;
    memcpy(ctx->host_stage, segs.data(), seg_bytes);
    memcpy((char *)ctx->host_stage + seg_bytes, offs.data(), off_bytes);
    SRF_HIP(hipMemcpyAsync(b.seg, ctx->host_stage, seg_bytes, hipMemcpyHostToDevice, stream));
    SRF_HIP(hipMemcpyAsync(b.offs, (char *)ctx->host_stage + seg_bytes, off_bytes,
                           hipMemcpyHostToDevice, stream));
    SRF_HIP(hipEventRecord(ctx->stage_done, stream));
    v.seg = b.seg;
    v.prim_off = b.offs;
    v.tile_off = b.offs + (nseg + 1);
    v.bblk_off = b.offs + 2 * (nseg + 1);
    v.mblk_off = b.offs + 3 * (nseg + 1);
    v.leaf_base = b.leaf_base;
    v.seg_of = b.seg_of;
    v.leaf_seg = b.leaf_seg;
    b.ss.vals_a = b.sidx;
  } else {
    b.ss.vals_a = segs[0].sorted_index;  // the sorted index lands in the accel
  }
  // SolidAccel's MAX_LEAF_ELEMENTS is the tetra's 4 (render/render_accel_solid.cc:34)
  // lines: the same 4 (LineAccel's own tree stops below 4 lines per leaf,
  // render/render_accel_line.h:59)
  int max_leaf = kind == kTri ? kTriMaxLeaf
                              : ((kind == kTet || kind == kSol || kind == kLin) ? kTetMaxLeaf
                                                                                : kSphMaxLeaf);

  // the region [status, hist] is contiguous: clear it inside k_bounds
  uint32_t *clear_begin = b.ss.status;
  uint32_t nclear = (uint32_t)((b.ss.hist + 1024 * (size_t)nseg) - clear_begin);
  const uint32_t nblocks = blocks_for(N, 256);

#define SRF_DISPATCH(KERNEL, ...)                                                      \
  do {                                                                                 \
    if (multi) {                                                                       \
      if (kind == kTri)                                                                \
        hipLaunchKernelGGL((KERNEL<kTri, true>), __VA_ARGS__);                           \
      else if (kind == kTet)                                                           \
        hipLaunchKernelGGL((KERNEL<kTet, true>), __VA_ARGS__);                           \
      else if (kind == kSol)                                                           \
        hipLaunchKernelGGL((KERNEL<kSol, true>), __VA_ARGS__);                           \
      else if (kind == kLin)                                                           \
        hipLaunchKernelGGL((KERNEL<kLin, true>), __VA_ARGS__);                           \
      else                                                                             \
        hipLaunchKernelGGL((KERNEL<kSph, true>), __VA_ARGS__);                           \
    } else {                                                                           \
      if (kind == kTri)                                                                \
        hipLaunchKernelGGL((KERNEL<kTri, false>), __VA_ARGS__);                          \
      else if (kind == kTet)                                                           \
        hipLaunchKernelGGL((KERNEL<kTet, false>), __VA_ARGS__);                          \
      else if (kind == kSol)                                                           \
        hipLaunchKernelGGL((KERNEL<kSol, false>), __VA_ARGS__);                           \
      else if (kind == kLin)                                                           \
        hipLaunchKernelGGL((KERNEL<kLin, false>), __VA_ARGS__);                          \
      else                                                                             \
        hipLaunchKernelGGL((KERNEL<kSph, false>), __VA_ARGS__);                          \
    }                                                                                  \
    SRF_HIP(hipGetLastError());                                                        \
  } while (0)
#define SRF_DISPATCH_M(KERNEL, ...)                                                    \
  do {                                                                                 \
    if (multi)                                                                         \
      hipLaunchKernelGGL(KERNEL<true>, __VA_ARGS__);                                   \
    else                                                                               \
      hipLaunchKernelGGL(KERNEL<false>, __VA_ARGS__);                                  \
    SRF_HIP(hipGetLastError());                                                        \
  } while (0)

  const CenCache cen =
      (kind == kTri || kind == kTet)
          ? CenCache{reinterpret_cast<float *>(b.ss.keys_b), reinterpret_cast<float *>(b.ss.vals_b),
                     reinterpret_cast<float *>(b.kchild)}
          : CenCache{nullptr, nullptr, nullptr};
  SRF_DISPATCH(k_bounds, dim3(bblocks), dim3(256), 0, stream, v, b.partials, clear_begin,
               nclear, cen);
  SRF_DISPATCH(k_bounds_final, dim3(multi ? nseg : 1u), dim3(256), 0, stream, v, b.partials,
               bblocks);
  SRF_DISPATCH(k_morton, dim3(mblocks), dim3(256), 0, stream, v, b.codes, b.ss.hist, cen);
  if (multi) {
    hipLaunchKernelGGL(k_sort_local, dim3(nseg), dim3(kSortBlock), 0, stream, b.codes,
                       b.ss.vals_a, v);
    SRF_HIP(hipGetLastError());
  }
  st = multi ? run_sort_passes<true>(b.codes, N, tiles, b.ss, v, stream)
             : run_sort_passes<false>(b.codes, N, tiles, b.ss, v, stream);
  if (st != SRF_OK) return st;
  // sorted codes in `codes`, sorted local indices in ss.vals_a
  SRF_DISPATCH_M(k_karras, dim3(nblocks), dim3(256), 0, stream, b.codes, N, b.kchild, b.kleaf,
                 b.kother, b.leaf_start, b.cls, b.level_counts, v);
  // top-down branch classification, one launch per depth
  // the first levels, which hold at most nseg * 2^level branches each, in
  // one workgroup (which also lists the roots); the rest in spans below
  const uint32_t top = top_levels(nseg);
  if (top > 0)
    SRF_DISPATCH_M(k_classify_top, dim3(1), dim3(1024), 0, stream, N, max_leaf, top, b.kchild,
                   b.kleaf, b.kother, b.levels, b.level_counts, b.cls, b.leaf_start, v);
  else
    SRF_DISPATCH_M(k_classify_roots, dim3(1), dim3(64), 0, stream, v, b.levels,
                   b.level_counts);
  // kLevelSpan levels per launch from `top` on: heads top, top + span, ...
  // (the lists of the levels between stay empty)
  const uint32_t cblocks = blocks_for(N, 256, 1024);
  for (uint32_t level = top; level < (uint32_t)kMaxBranchLevels; level += kLevelSpan) {
    if (multi)
      hipLaunchKernelGGL((k_classify_span<true, kLevelSpan>), dim3(cblocks), dim3(256), 0,
                         stream, N, max_leaf, level, b.kchild, b.kleaf, b.kother, b.levels,
                         b.level_counts, b.cls, b.leaf_start, v);
    else
      hipLaunchKernelGGL((k_classify_span<false, kLevelSpan>), dim3(cblocks), dim3(256), 0,
                         stream, N, max_leaf, level, b.kchild, b.kleaf, b.kother, b.levels,
                         b.level_counts, b.cls, b.leaf_start, v);
    SRF_HIP(hipGetLastError());
  }
  const uint32_t scan_blocks = blocks_for(N, kScanTile);
  hipLaunchKernelGGL(k_scan_reduce, dim3(scan_blocks), dim3(kScanBlock), 0, stream,
                     b.leaf_start, N, b.partial);
  if (scan_blocks <= kScanFuse) {
    hipLaunchKernelGGL(k_scan_downsweep_fused, dim3(scan_blocks), dim3(kScanBlock), 0, stream,
                       b.leaf_start, N, b.partial, b.leaf_rank, b.total,
                       multi ? nullptr : segs[0].params);
  } else {
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, stream, b.partial, scan_blocks);
    hipLaunchKernelGGL(k_scan_downsweep, dim3(scan_blocks), dim3(kScanBlock), 0, stream,
                       b.leaf_start, N, b.partial, b.leaf_rank, b.total,
                       multi ? nullptr : segs[0].params);
  }
  SRF_HIP(hipGetLastError());
  if (multi) {
    hipLaunchKernelGGL(k_seg_leaves, dim3(1), dim3(256), 0, stream, v, b.leaf_rank, b.total,
                       b.leaf_base);
    SRF_HIP(hipGetLastError());
  }
  SRF_DISPATCH(k_emit, dim3(nblocks), dim3(256), 0, stream, N, max_leaf, b.codes, b.kchild,
               b.kleaf, b.kother, b.cls, b.leaf_rank, b.o, v);
  SRF_DISPATCH(k_leaves, dim3(nblocks), dim3(256), 0, stream, N, b.ss.vals_a, b.o, b.total,
               b.big, b.level_counts + 32, v);
  if (kind == kSph && !multi) {
    // big leaves in chunks (depth-capped clusters hold up to 10^5 spheres)
    Ext *part = reinterpret_cast<Ext *>(b.big_part);
    hipLaunchKernelGGL(k_big_chunks, dim3(1), dim3(1024), 0, stream, b.big,
                       b.level_counts + 32, b.o, b.big_units);
    hipLaunchKernelGGL(k_leaves_big_part<kSph>, dim3(blocks_for(N, 256 * (int)kBigLeaf, 2048)),
                       dim3(256), 0, stream, b.big, b.level_counts + 32, b.o, b.big_units, part,
                       v);
    hipLaunchKernelGGL(k_leaves_big_combine, dim3(blocks_for(N / kBigLeaf + 1, 256, 256)),
                       dim3(256), 0, stream, b.big, b.level_counts + 32, b.o, b.big_units, part,
                       v);
    SRF_HIP(hipGetLastError());
  } else {
    SRF_DISPATCH(k_leaves_big, dim3(blocks_for(N, 256 * (int)kBigLeaf, 1024)), dim3(256), 0,
                 stream, b.ss.vals_a, b.o, b.big, b.level_counts + 32, v);
  }
  // bottom-up refit, one level per launch (levels are branch steps 23..top),
  // then levels top-1..0 in one workgroup
  const uint32_t rblocks = blocks_for(N, 256, 2048);
  int last_head = (int)top;
  while (last_head + kLevelSpan < kMaxBranchLevels) last_head += kLevelSpan;
  for (int level = last_head; level >= (int)top; level -= kLevelSpan) {
    if (multi)
      hipLaunchKernelGGL((k_refit_span<true, kLevelSpan>), dim3(rblocks), dim3(256), 0, stream,
                         N, (uint32_t)level, b.levels, b.level_counts, b.o, v);
    else
      hipLaunchKernelGGL((k_refit_span<false, kLevelSpan>), dim3(rblocks), dim3(256), 0, stream,
                         N, (uint32_t)level, b.levels, b.level_counts, b.o, v);
    SRF_HIP(hipGetLastError());
  }
  if (top > 0)
    SRF_DISPATCH_M(k_refit_top, dim3(1), dim3(1024), 0, stream, N, top, b.levels,
                   b.level_counts, b.o, v);
  if (big_out && !multi && kind == kSph) {
    hipLaunchKernelGGL(k_big_export, dim3(blocks_for(big_out->big_cap, 256, 1024)), dim3(256), 0,
                       stream, b.big, b.level_counts + 32, b.o, big_out->big_info,
                       big_out->big_count, (uint32_t)big_out->big_cap);
    SRF_HIP(hipGetLastError());
  }
#undef SRF_DISPATCH
#undef SRF_DISPATCH_M
  return SRF_OK;
}

// Secondary trees over the big leaves of a finished sphere build: the big
// leaves' records come to the host (the build's only synchronisation), each
// leaf becomes one segment of a batched sphere build over its packed
// spheres, and each big leaf's unused second box slot receives the offsets
// of its tree's nodes and spheres.
srf_status build_sph_subtrees(Ctx *ctx, Accel *a, hipStream_t stream) {
  a->nbig = 0;
  a->sub_nodes_used = a->sub_prims_used = 0;
  SRF_HIP(hipStreamSynchronize(stream));
  uint32_t nbig = 0;
  SRF_HIP(hipMemcpy(&nbig, a->big_count, 4, hipMemcpyDeviceToHost));
  if (nbig == 0) return SRF_OK;
  if (nbig > a->big_cap) return set_error(SRF_ERR_INTERNAL, "big-leaf list overflow");
  std::vector<uint4> info(nbig);
  SRF_HIP(hipMemcpy(info.data(), a->big_info, 16 * (size_t)nbig, hipMemcpyDeviceToHost));
  // a fixed order (by first sphere) makes the layout independent of the
  // order the build's atomics listed the leaves in
  std::sort(info.begin(), info.end(), [](const uint4 &x, const uint4 &y) { return x.y < y.y; });
  uint64_t nodes_total = 0, prims_total = 0;
  std::vector<uint4> link(nbig);
  for (uint32_t m = 0; m < nbig; m++) {
    const uint4 r = info[m];
    if (r.z <= kBigLeaf || (uint64_t)r.y + r.z > a->n)
      return set_error(SRF_ERR_INTERNAL, "inconsistent big-leaf record");
    link[m] = make_uint4(r.x, (uint32_t)nodes_total, (uint32_t)prims_total, r.z);
    nodes_total += 2 * (uint64_t)r.z - 1;
    prims_total += r.z;
  }
  if (nodes_total >= (1ull << 32))
    return set_error(SRF_ERR_UNSUPPORTED, "secondary trees exceed 2^32 nodes");
  // buffers grow only
  if (prims_total > a->sub_cap_prims || nodes_total > a->sub_cap_nodes || nbig > a->sub_cap_trees) {
    free_sub_buffers(a);
    SRF_HIP(hipMalloc(&a->sub_nodes, 64 * nodes_total));
    a->sub_cap_nodes = nodes_total;
    SRF_HIP(hipMalloc(&a->sub_prims, 16 * prims_total));
    SRF_HIP(hipMalloc((void **)&a->sub_sidx, 4 * prims_total));
    a->sub_cap_prims = prims_total;
    SRF_HIP(hipMalloc((void **)&a->sub_params, sizeof(BuildParams) * (size_t)nbig));
    a->sub_cap_trees = nbig;
  }
  std::vector<Seg> segs(nbig);
  const float *packed = reinterpret_cast<const float *>(a->prims);
  for (uint32_t m = 0; m < nbig; m++) {
    const uint32_t first = info[m].y, count = info[m].z;
    PrimSrc s{};
    s.V = packed + 4 * (size_t)first;
    s.R = s.V + 3;
    s.n = count;
    s.vstride = 4;
    s.rstride = 4;
    segs[m] = Seg{s, (char *)a->sub_nodes + 64 * (size_t)link[m].y,
                  (char *)a->sub_prims + 16 * (size_t)link[m].z, a->sub_sidx + link[m].z,
                  nullptr, a->sub_params + m};
  }
  srf_status st = build_segs(ctx, kSph, segs, stream, nullptr);
  if (st != SRF_OK) return st;
  // the link table travels through big_info (no longer needed as such)
  st = ensure_host_stage(ctx, 16 * (size_t)nbig);
  if (st != SRF_OK) return st;
  memcpy(ctx->host_stage, link.data(), 16 * (size_t)nbig);
  SRF_HIP(hipMemcpyAsync(a->big_info, ctx->host_stage, 16 * (size_t)nbig, hipMemcpyHostToDevice,
                         stream));
  SRF_HIP(hipEventRecord(ctx->stage_done, stream));
  hipLaunchKernelGGL(k_sub_link, dim3(blocks_for(nbig, 256, 1024)), dim3(256), 0, stream,
                     reinterpret_cast<float4 *>(a->nodes), (uint32_t)(2 * a->n - 1), a->big_info,
                     nbig, a->sub_params, a->params);
  SRF_HIP(hipGetLastError());
  a->nbig = nbig;
  a->sub_nodes_used = nodes_total;
  a->sub_prims_used = prims_total;
  return SRF_OK;
}

}  // namespace

srf_status build_lbvh_segments(Ctx *ctx, Kind kind, const void *const *descs,
                               Accel *const *accs, uint32_t nseg, hipStream_t stream) {
  std::vector<Seg> segs(nseg);
  for (uint32_t m = 0; m < nseg; m++) {
    Accel *a = accs[m];
    segs[m] = Seg{make_src(kind, descs[m]), a->nodes, a->prims, a->sorted_index,
                  a->sorted_faces, a->params};
  }
  srf_status st = build_segs(ctx, kind, segs, stream,
                             (nseg == 1 && kind == kSph) ? accs[0] : nullptr);
  if (st != SRF_OK || nseg != 1 || kind != kSph) return st;
  return build_sph_subtrees(ctx, accs[0], stream);
}

srf_status build_lbvh(Ctx *ctx, Kind kind, const void *desc, hipStream_t stream, Accel *acc) {
  return build_lbvh_segments(ctx, kind, &desc, &acc, 1, stream);
}

srf_status launch_morton(Kind kind, const void *desc, const float bmin[3],
                         const float bmax[3], uint32_t *codes, hipStream_t stream) {
  PrimSrc s = make_src(kind, desc);
  BuildParams p{};
  for (int k = 0; k < 3; k++) {
    p.bmin[k] = bmin[k];
    p.bmax[k] = bmax[k];
    p.inv[k] = (float)1024 / (bmax[k] - bmin[k]);
  }
  uint32_t blocks = blocks_for(s.n, 256, 4096);
  if (kind == kTri)
    hipLaunchKernelGGL(k_morton_box<kTri>, dim3(blocks), dim3(256), 0, stream, s, p, codes);
  else if (kind == kTet)
    hipLaunchKernelGGL(k_morton_box<kTet>, dim3(blocks), dim3(256), 0, stream, s, p, codes);
  else if (kind == kSol)
    hipLaunchKernelGGL(k_morton_box<kSol>, dim3(blocks), dim3(256), 0, stream, s, p, codes);
  else if (kind == kLin)
    hipLaunchKernelGGL(k_morton_box<kLin>, dim3(blocks), dim3(256), 0, stream, s, p, codes);
  else
    hipLaunchKernelGGL(k_morton_box<kSph>, dim3(blocks), dim3(256), 0, stream, s, p, codes);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

size_t sort_pairs30_bytes(uint32_t n) {
  const uint32_t tiles = sort_tiles(n);
  return 4 * (3 * (size_t)n + 4 * (size_t)tiles * 256 + 8 + 1024) + 8 * 256;
}

srf_status sort_pairs30_at(Ctx *ctx, size_t arena_off, const uint32_t *codes_in, uint32_t n,
                           uint32_t *codes_out, uint32_t *index_out, hipStream_t stream) {
  srf_status st = ensure_scratch(ctx, arena_off + sort_pairs30_bytes(n), stream);
  if (st != SRF_OK) return st;
  return sort_pairs30_in((char *)ctx->scratch, ctx->scratch_bytes, arena_off, codes_in, n,
                         codes_out, index_out, stream);
}

srf_status sort_pairs30_in(char *base, size_t size, size_t arena_off, const uint32_t *codes_in,
                           uint32_t n, uint32_t *codes_out, uint32_t *index_out,
                           hipStream_t stream) {
  const uint32_t tiles = sort_tiles(n);
  Arena ar{base, size, arena_off};
  SortScratch ss;
  ss.keys_b = ar.take<uint32_t>(n);
  ss.vals_b = ar.take<uint32_t>(n);
  ss.vals_a = index_out;
  ss.status = ar.take<uint32_t>((size_t)4 * tiles * 256);
  ss.counters = ar.take<uint32_t>(8);
  ss.hist = ar.take<uint32_t>(1024);
  if (ar.off > ar.size) return set_error(SRF_ERR_INTERNAL, "scratch layout overflow");
  size_t nclear = (size_t)((ss.hist + 1024) - ss.status);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(nclear, 256, 2048)), dim3(256), 0, stream,
                     ss.status, 0u, nclear);
  if (codes_out != codes_in)
    SRF_HIP(hipMemcpyAsync(codes_out, codes_in, 4 * (size_t)n, hipMemcpyDeviceToDevice, stream));
  hipLaunchKernelGGL(k_histogram, dim3(blocks_for(n, 256, 4096)), dim3(256), 0, stream,
                     codes_out, n, ss.hist);
  SRF_HIP(hipGetLastError());
  SegView v{};
  v.nseg = 1;
  return run_sort_passes<false>(codes_out, n, tiles, ss, v, stream);
}

srf_status sort_pairs30(Ctx *ctx, const uint32_t *codes_in, uint32_t n, uint32_t *codes_out,
                        uint32_t *index_out, hipStream_t stream) {
  return sort_pairs30_at(ctx, 0, codes_in, n, codes_out, index_out, stream);
}

srf_status launch_radix_tree(const uint32_t *sorted_codes, uint32_t n, srf_node_info *out,
                             hipStream_t stream) {
  hipLaunchKernelGGL(k_karras_nodeinfo, dim3(blocks_for(n, 256)), dim3(256), 0, stream,
                     sorted_codes, n, out);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

}  // namespace srf
