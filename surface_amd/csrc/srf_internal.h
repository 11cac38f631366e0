// srf_internal.h -- shared definitions of the HIP implementation of srf.h.
//
// Device layouts (all in HBM, 16-byte aligned):
//   node (64 B, every primitive kind)  the reference's TriangleNode word
//       order (render/render_accel_triangle.h:23-42): bmin[2][3], bmax[2][3],
//       flag, axis, data[2], fetched as four 16-byte loads. A branch holds
//       both child boxes; a leaf holds its own box in slot 0. ParticleNode
//       (render/render_accel_particle.h:24-44) has the same fields; TetraNode
//       (render/render_accel_tetra.h:24-43) keeps one box per node, which is
//       the box its parent holds for it (the root's is BoundingBox()), so a
//       tetra branch can cull both children before they are pushed.
//       srf_accel_export converts to the reference layouts.
//   packed primitives, in sorted (leaf) order:
//       triangle 48 B  (p0.xyz, id) (e1.xyz, 0) (e2.xyz, 0), e1 = p1 - p0,
//                      e2 = p2 - p0 -- the same float subtractions
//                      TriangleIsect performs, so results are unchanged;
//       tetra    64 B  (v0.xyz, id) (v1.xyz, 0) (v2.xyz, 0) (v3.xyz, 0);
//       sphere   16 B  (centre.xyz, radius).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/srf.h"

namespace srf {

constexpr int kTriMaxLeaf = 16;  // render/render_accel_triangle.cc:34
constexpr int kTetMaxLeaf = 4;   // render/render_accel_tetra.cc:34
constexpr int kSphMaxLeaf = 16;  // render/render_accel_particle.cc:51
constexpr int kMaxDepth32 = 22;  // render/render_accel_triangle.cc:35
// Branch levels are bounded: the root, then depths 0..22 (a child deeper
// than 22 is a leaf), 24 branch nodes on any root-to-leaf path. A traversal
// pushes at most one entry per branch node on its current path (the node it
// descends into stays in registers), so a stack of 24 entries never fills.
constexpr int kMaxBranchLevels = 24;
constexpr int kStackSize = 32;
// Leaves with more primitives than this are packed by a whole block during
// the build; for spheres each such leaf (depth-capped leaves of clustered
// particles hold 10^5 and more) also gets a secondary tree over its spheres.
constexpr uint32_t kBigLeaf = 64;

// kSol: pyramids, prisms and hexahedra (SolidAccel), Accel::nvps vertices
// per solid; kLin: capped cylinders along line segments (LineAccel)
enum Kind : uint32_t { kTri = 0, kTet = 1, kSph = 2, kSol = 3, kLin = 4 };

// PartBy2_30 (render/render_prefix_tree_util.h:113-120): the low 10 bits of
// n spread to every third bit.
__device__ __forceinline__ uint32_t part_by2_30(uint32_t n) {
  n &= 0x000003ffu;
  n = (n ^ (n << 16)) & 0xff0000ffu;
  n = (n ^ (n << 8)) & 0x0300f00fu;
  n = (n ^ (n << 4)) & 0x030c30c3u;
  n = (n ^ (n << 2)) & 0x09249249u;
  return n;
}

// Workgroups go to the 8 XCDs round-robin (block b runs on XCD b % 8), each
// XCD with its own L2. Renumbering the blocks so that consecutive blocks run
// on the same XCD keeps neighbouring ray tiles, which touch the same nodes or
// voxels, on the same L2.
constexpr uint32_t kXcds = 8;
// Runs of SRF_XCD_GROUP consecutive blocks go to one XCD in turn: locality
// within a run, balance across the grid (0 = one contiguous eighth per XCD,
// 1 = the hardware's own round robin). Measured on 1024^2 ray tiles (128
// rays per block): 128 gave the tetra trace 7.2 -> 5.4 ms and the volume
// march 2.4 -> 2.05 ms, the triangle trace within 2 %; 0 lost 25 % on
// triangles to imbalance between the eighths.
#ifndef SRF_XCD_GROUP
#define SRF_XCD_GROUP 128
#endif
// grids using xcd_block are padded to a multiple of this
constexpr uint32_t kXcdPad = kXcds * (SRF_XCD_GROUP > 0 ? SRF_XCD_GROUP : 1);
template <uint32_t kGroup = SRF_XCD_GROUP>
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t x = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
  if (kGroup == 0) return x * (gridDim.x / kXcds) + k;
  return (k / kGroup) * (kXcds * kGroup) + x * kGroup + (k % kGroup);
}

// Per-build scalars produced on the device and read by later kernels.
struct alignas(16) BuildParams {
  float bmin[3];       // scene box used for Morton codes
  float bmax[3];
  float inv[3];        // (float)1024 / (bmax - bmin)
  float kEPS;          // leaf-box padding (tri: FLT_EPSILON*epsScale)
  double eps_scale;    // triangles: max scene extent
  float root_bmin[3];  // BoundingBox() of the finished tree
  float root_bmax[3];
  uint32_t num_leaves;
  uint32_t num_nodes;
  uint32_t max_steps;  // deepest branch (root = 0 steps)
  uint32_t error;      // nonzero: a bounded wait gave up (never expected)
  uint32_t ticket;     // unused (kept zero; the block keeps its 128-byte layout)
  uint32_t pad[3];
};

struct Accel {
  Kind kind;
  int device;
  uint64_t n;            // primitives
  uint64_t nv;           // vertices (tets: for scalar lookups)
  uint32_t nvps;         // solids: vertices per solid (5, 6, 8)
  uint32_t line_cap;     // lines: LineBuildOptions::cap (end discs tested)
  uint32_t wide_stack;   // traverse with the stack layouts of trees of 2^24 or
                         // more nodes (srf_debug_force_wide_stack)
  uint64_t capacity;     // primitives the buffers can hold
  void *nodes;           // TriNode / TetNode / SphNode
  void *prims;           // packed primitives (sorted order)
  uint32_t *sorted_index;
  uint32_t *sorted_faces;  // tets: original 4 vertex ids per sorted tet
  BuildParams *params;   // device
  hipEvent_t built;      // recorded on the build stream
  // spheres: the leaves of more than kBigLeaf spheres and their secondary
  // trees (one Build32 tree per such leaf, built over its packed spheres;
  // the leaf's unused second box slot links to it)
  uint4 *big_info;       // per big leaf: (node, first, count, 0) after the build
  uint32_t *big_count;   // device counter, written by the build
  uint64_t big_cap;      // big_info entries
  uint32_t nbig;         // big leaves of the last build (host)
  void *sub_nodes;       // 64-byte nodes of all secondary trees
  void *sub_prims;       // their packed spheres (16 B)
  uint32_t *sub_sidx;    // a secondary tree's sphere -> position in its leaf
  BuildParams *sub_params;
  uint64_t sub_cap_prims, sub_cap_nodes, sub_cap_trees;
  uint64_t sub_nodes_used, sub_prims_used;
  // The geometry of the last build in device memory, which BuildIntersection
  // reads in srf_trace_closest_host (the reference keeps the caller's Mesh
  // pointer likewise): the caller's arrays after a device build, copies the
  // accelerator owns (geom_own, geom_bytes) after a host build.
  const void *geom[3];  // vertices / positions, faces / indices, radius
  void *geom_own[3];
  size_t geom_bytes[3];
};

// Records the device arrays of a build's descriptor in a->geom.
void record_geometry(Accel *a, const void *desc);
void free_geometry(Accel *a);
// BuildIntersection launchers (srf_isect.hip) for an accelerator's recorded
// geometry.
srf_status launch_isect(const Accel *a, const srf_ray *rays, const srf_hit *hits,
                        const srf_tet_hit_ext *ext, uint32_t n, srf_isect *out, hipStream_t s);

// Device view of the secondary sphere trees handed to the traversal.
struct SphSub {
  const float4 *nodes;
  const float4 *prims;
  const uint32_t *sidx;
  uint32_t nnodes, nprims;  // valid extents of nodes / prims
};

struct Ctx {
  int device;
  void *scratch;
  size_t scratch_bytes;
  void *host_stage;       // pinned staging for segment tables (many-mesh builds)
  size_t host_stage_bytes;
  hipEvent_t stage_done;  // the last copy out of host_stage
};

// Bump allocator over the context scratch (256-byte aligned slices).
struct Arena {
  char *base;
  size_t size, off;
  template <class T>
  T *take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T *p = reinterpret_cast<T *>(base + off);
    off += sizeof(T) * count;
    return p;
  }
};

// Error plumbing shared by the translation units.
srf_status set_error(srf_status s, const char *fmt, ...);
srf_status hip_error(hipError_t e, const char *what);
srf_status ensure_scratch(Ctx *ctx, size_t bytes, hipStream_t stream);
// Waits for the previous copy out of the staging buffer, then grows it.
srf_status ensure_host_stage(Ctx *ctx, size_t bytes);

#define SRF_HIP(call)                                  \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return srf::hip_error(e_, #call); \
  } while (0)

// Build-stage launchers (srf_build.hip).
srf_status build_lbvh(Ctx *ctx, Kind kind, const void *desc, hipStream_t stream,
                      Accel *acc);
srf_status build_lbvh_segments(Ctx *ctx, Kind kind, const void *const *descs,
                               Accel *const *accs, uint32_t nseg, hipStream_t stream);
// Frees a sphere accelerator's secondary-tree buffers (srf_api.hip).
void free_sub_buffers(Accel *a);
srf_status launch_morton(Kind kind, const void *desc, const float bmin[3],
                         const float bmax[3], uint32_t *codes, hipStream_t s);
srf_status sort_pairs30(Ctx *ctx, const uint32_t *codes_in, uint32_t n,
                        uint32_t *codes_out, uint32_t *index_out, hipStream_t s);
srf_status launch_radix_tree(const uint32_t *sorted_codes, uint32_t n,
                             srf_node_info *out, hipStream_t s);

// Traversal launchers (srf_trace.hip).
srf_status trace_tri(const Accel *a, const srf_ray *rays, uint64_t n,
                     srf_hit *hits, hipStream_t s);
srf_status trace_tet(const Accel *a, const srf_ray *rays, uint64_t n,
                     const float *scalars, srf_hit *hits, srf_tet_hit_ext *ext,
                     hipStream_t s);
srf_status trace_sol(const Accel *a, const srf_ray *rays, uint64_t n, const float *scalars,
                     srf_hit *hits, srf_solid_hit_ext *ext, hipStream_t s);
srf_status trace_lin(const Accel *a, const srf_ray *rays, uint64_t n, srf_hit *hits,
                     srf_line_hit_ext *ext, hipStream_t s);
srf_status trace_sph(const Accel *a, const srf_ray *rays,
                     const srf_sphere_prev *prev, uint64_t n, srf_hit *hits,
                     hipStream_t s);
srf_status trace_any(const Accel *a, const srf_ray *rays, const float *tmax,
                     uint64_t n, uint32_t *occluded, hipStream_t s);
srf_status trace_any_sorted(Ctx *ctx, const Accel *a, const srf_ray *rays, const float *tmax,
                            uint64_t n, uint32_t key_bits, int mode, uint32_t *occluded,
                            hipStream_t s);
srf_status ao_rays_spheres(const srf_ray *primary, const srf_hit *hits, uint64_t n,
                           const float *positions, const float *radius, float rconst,
                           uint32_t per_hit, float eps, float dist, uint64_t seed,
                           srf_ray *out, float *out_tmax, uint32_t *out_pixel,
                           uint32_t *count, hipStream_t s);
// Stable sort of 30-bit codes with their indices, using the arena from
// `arena_off` on (the caller's own slices come before it).
srf_status sort_pairs30_at(Ctx *ctx, size_t arena_off, const uint32_t *codes_in, uint32_t n,
                           uint32_t *codes_out, uint32_t *index_out, hipStream_t stream);
size_t sort_pairs30_bytes(uint32_t n);
// The same in a caller-provided buffer [base, base + size) from arena_off on;
// allocates nothing.
srf_status sort_pairs30_in(char *base, size_t size, size_t arena_off, const uint32_t *codes_in,
                           uint32_t n, uint32_t *codes_out, uint32_t *index_out,
                           hipStream_t stream);

}  // namespace srf
