/*
 * srf_oracle.h -- CPU restatement of SURFACE's LBVH build, closest-hit
 * traversal and 3D-texture sampling, used ONLY as the parity checker.
 *
 * TEST INFRASTRUCTURE. Nothing in the product (surface_amd/, include/) may
 * include, link or call this code; only tests/, the smoke check and the
 * bench's cpu_baseline leg do.
 *
 * Layouts below are byte-identical to the reference's node structs so that a
 * tree emitted here can be compared field by field with one written by the
 * reference's own Dump():
 *   or_tri_node  == TriangleNode  render/render_accel_triangle.h:23-42 (64 B)
 *   or_tet_node  == TetraNode     render/render_accel_tetra.h:24-43    (40 B)
 *                   and SolidNode render/render_accel_solid.h:24-43    (40 B)
 *   or_sph_node  == ParticleNode  render/render_accel_particle.h:24-44 (60 B)
 *   or_nodeinfo  == NodeInfo32    render/render_prefix_tree_util.h:51-61
 */
#ifndef SRF_ORACLE_H_
#define SRF_ORACLE_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  float bmin[2][3];
  float bmax[2][3];
  int32_t flag; /* 1 = leaf */
  int32_t axis;
  uint32_t data[2]; /* leaf: {count, first}; branch: {left, right} */
} or_tri_node;

typedef struct {
  float bmin[3];
  float bmax[3];
  int32_t flag;
  int32_t axis;
  uint32_t data[2];
} or_tet_node;

typedef struct {
  float bmin[2][3];
  float bmax[2][3];
  uint32_t data[2];
  int16_t flag;
  int16_t axis;
} or_sph_node;

typedef struct {
  uint32_t child;
  int32_t left_type;  /* 1 = leaf */
  int32_t right_type; /* 1 = leaf */
} or_nodeinfo;

/* One ray as handed to the tracer: origin, direction, double-sided flag and
 * the primitive to skip (prev_prim_id, ~0 = none). For spheres the
 * self-intersection rule also needs prev_hit_t / prev_hit_normal. */
typedef struct {
  float org[3];
  uint32_t flags; /* bit0: double sided */
  float dir[3];
  uint32_t prev_prim_id;
} or_ray;

typedef struct {
  float t, u, v;
  uint32_t prim_id;
} or_hit;

/* Tetra extras: InterpolateTetra weights d0..d3 at the entry point, the
 * interpolated per-vertex scalar, the entry face and the exit distance. */
typedef struct {
  float d[4];
  float scalar;
  uint32_t subface;
  float t_leave;
  float pad;
} or_tet_ext;

/* Solid extras at the reported crossing: interpolate() weights d0..d7
 * (render/render_accel_solid.cc:247-435; unused entries 0), the crossing's
 * Intersection::position and normal (the exit crossing's normal negated),
 * the varying sum of scalar[nv * prim + k] * d_k (gles/gles_common.cc:
 * 1078-1140) and which crossing it is (0 entry, 1 exit). */
typedef struct {
  float d[8];
  float pos[3];
  float scalar;
  float nrm[3];
  uint32_t side;
} or_sol_ext;

/* Line extras: Intersection::position (org + t dir) and normal
 * (render/render_accel_line.cc:654-702) and whether an end cap was hit. */
typedef struct {
  float pos[3];
  uint32_t hit_cap;
  float nrm[3];
  float pad;
} or_lin_ext;

/* ---- LBVH stages ---- */
void or_scene_bounds_tri(const float *V, const uint32_t *F, int64_t n,
                         float bmin[3], float bmax[3]);
void or_scene_bounds_tet(const float *V, const uint32_t *F, int64_t n,
                         float bmin[3], float bmax[3]);
void or_scene_bounds_sph(const float *P, const float *R, float rconst,
                         int64_t n, float bmin[3], float bmax[3]);
void or_morton_tri(const float *V, const uint32_t *F, int64_t n,
                   const float bmin[3], const float bmax[3], uint32_t *codes);
void or_morton_tet(const float *V, const uint32_t *F, int64_t n,
                   const float bmin[3], const float bmax[3], uint32_t *codes);
void or_scene_bounds_sol(const float *V, const uint32_t *F, int64_t n, int nv,
                         float bmin[3], float bmax[3]);
void or_morton_sol(const float *V, const uint32_t *F, int64_t n, int nv,
                   const float bmin[3], const float bmax[3], uint32_t *codes);
void or_morton_pts(const float *P, int64_t n, const float bmin[3],
                   const float bmax[3], uint32_t *codes);
uint32_t or_morton30_point(const float p[3], const float bmin[3],
                           float invx, float invy, float invz);
/* Stable sort of (code, index=0..n-1) by code. */
void or_sort_keys(const uint32_t *codes, int64_t n, uint32_t *sorted_codes,
                  uint32_t *sorted_index);
void or_radix_tree30(const uint32_t *sorted_codes, int64_t n,
                     or_nodeinfo *out /* n-1 */);

/* ---- Full Build32 sequences. Return the node count; nodes/indices must be
 * sized 2n and n. bounds[6] = root bmin/bmax as stored by BoundingBox().
 * eps_scale (triangles only) receives the scene-size epsilon scale. ---- */
int64_t or_build_tri(const float *V, const uint32_t *F, int64_t n,
                     or_tri_node *nodes, uint32_t *indices, double bounds[6],
                     double *eps_scale);
int64_t or_build_tet(const float *V, const uint32_t *F, int64_t n,
                     or_tet_node *nodes, uint32_t *indices, double bounds[6]);
int64_t or_build_sph(const float *P, const float *R, float rconst, int64_t n,
                     or_sph_node *nodes, uint32_t *indices, double bounds[6]);
/* lines: segments F (2 vertex ids per line), radius per vertex R or rconst;
 * the LBVH over segment midpoints (the reference has only its top-down
 * LineAccel::Build) */
void or_scene_bounds_lin(const float *V, const uint32_t *F, const float *R, float rconst,
                         int64_t n, float bmin[3], float bmax[3]);
void or_morton_lin(const float *V, const uint32_t *F, int64_t n, const float bmin[3],
                   const float bmax[3], uint32_t *codes);
int64_t or_build_lin(const float *V, const uint32_t *F, const float *R, float rconst, int64_t n,
                     or_tet_node *nodes, uint32_t *indices, double bounds[6]);
/* solids of nv = 5 (pyramid), 6 (prism) or 8 (hexahedron) vertices */
int64_t or_build_sol(const float *V, const uint32_t *F, int64_t n, int nv,
                     or_tet_node *nodes, uint32_t *indices, double bounds[6]);

/* ---- Closest-hit traversal (one ray after another, or OpenMP over rays
 * when threads > 1). stats (optional) receives, per ray, {nodes visited,
 * primitives tested}. ---- */
void or_trace_tri(const or_tri_node *nodes, const uint32_t *indices,
                  const float *V, const uint32_t *F, double eps_scale,
                  const or_ray *rays, int64_t nrays, or_hit *hits,
                  uint32_t *stats, int threads);
void or_trace_tet(const or_tet_node *nodes, const uint32_t *indices,
                  const float *V, const uint32_t *F, const float *scalars,
                  const or_ray *rays, int64_t nrays, or_hit *hits,
                  or_tet_ext *ext, uint32_t *stats, int threads);
void or_trace_sol(const or_tet_node *nodes, const uint32_t *indices,
                  const float *V, const uint32_t *F, int nv, const float *scalars,
                  const or_ray *rays, int64_t nrays, or_hit *hits, or_sol_ext *ext,
                  uint32_t *stats, int threads);
/* One solid against one ray (double origin / direction / vertices):
 * 1 with both crossings' t, position, normal; 0 no crossing; -1 no axis
 * with |dir| > 0.1. */
int or_solid_isect(int nv, const double o[3], const double d[3], const double *verts,
                   float t[2], float pos[6], float nrm[6]);
void or_interpolate_solid(int nv, const float p[3], const double *verts, float d_out[8]);
void or_trace_lin(const or_tet_node *nodes, const uint32_t *indices, const float *V,
                  const uint32_t *F, const float *R, float rconst, int cap, const or_ray *rays,
                  int64_t nrays, or_hit *hits, or_lin_ext *ext, uint32_t *stats, int threads);
/* CylinderIsect for one line (radius rr) with the running distance tmax:
 * out = (t, u, v) as updated, returns the hit flag. */
int or_line_isect(const float p0[3], const float p1[3], float rr, const float o[3],
                  const float d[3], int cap, float tmax, float out[3], int *hit_cap);
void or_trace_sph(const or_sph_node *nodes, const uint32_t *indices,
                  const float *P, const float *R, float rconst,
                  const or_ray *rays, const double *prev_hit_t,
                  const float *prev_hit_n, int64_t nrays, or_hit *hits,
                  uint32_t *stats, int threads);

/* ---- 3D texture (FilterTexture3DFloat), single component. ---- */
void or_tex3d_sample(const float *vox, const int dims[3], const int clamp[3],
                     const float *uvw, int64_t n, float *out);

/* ---- Emission-absorption volume march (builder-defined; FP32 forward as
 * the GPU computes it, FP64 forward and adjoint as the gradient checker). */
typedef struct {
  int dims[3];
  int clamp[3];
  int tf_size;
  float step; /* in texture units */
} or_vol_params;
void or_volume_fwd_f32(const or_vol_params *p, const float *vox,
                       const float *tf, const or_ray *rays, int64_t nrays,
                       float *rgba, uint32_t *nsamples, int threads);
void or_volume_fwd_f64(const or_vol_params *p, const double *vox,
                       const double *tf, const or_ray *rays, int64_t nrays,
                       double *rgba, int threads);
void or_volume_bwd_f64(const or_vol_params *p, const double *vox,
                       const double *tf, const or_ray *rays, int64_t nrays,
                       const double *grad_rgba, double *grad_vox,
                       double *grad_tf);

/* ---- Two-level scene: the top-level BVHTree over element world boxes and
 * RenderGraph::Trace over triangle elements (gles/gles_render_graph.cc). */
typedef struct {
  double bmin[3];
  double bmax[3];
  int32_t child[2]; /* 0 = none */
  int32_t axis;
  int32_t is_leaf;
  uint32_t node_id; /* leaves: element index */
} or_tlas_node;     /* BVHNode, render/render_bvh_tree.h */

/* TransformBoundingBox: world box of a local box under T (row-vector
 * convention, p' = p T), gles/gles_render_graph.cc:19-79. out = bmin, bmax. */
void or_transform_box(const double bmin[3], const double bmax[3], const double T[16],
                      double out[6]);
/* BVHTree::BuildTree over n boxes (box i = bmin[3], bmax[3], node id i).
 * nodes must hold 2n entries; returns the node count. */
int64_t or_tlas_build(const double *boxes, int64_t n, or_tlas_node *nodes);

typedef struct {
  const or_tri_node *nodes;
  const uint32_t *indices;
  const float *V;
  const uint32_t *F;
  double eps_scale;
  double T[16];      /* local -> world */
  double invT[16];   /* world -> local */
  double invT33[16]; /* inverse of T with the translation cleared */
} or_tri_element;

/* RenderGraph::Trace per ray: hit t = world distance, u/v/prim_id of the
 * element's triangle, inst = element index (~0 on a miss). */
void or_scene_trace(const or_tlas_node *tlas, const or_tri_element *elems,
                    const or_ray *rays, int64_t nrays, or_hit *hits, uint32_t *inst,
                    int threads);

/* One element of a scene that mixes primitive kinds (RenderGraph::Trace's
 * switch, gles/gles_render_graph.cc:224-264): kind 0 triangles (nodes are
 * or_tri_node, V/F, eps_scale), 1 tetrahedra (or_tet_node, V/F), 2 spheres
 * (or_sph_node, positions V, radii R or rconst), 3 solids (or_tet_node, V/F,
 * nv vertices each), 4 lines (or_tet_node, positions V, segments F, radii R
 * or rconst, cap); transforms as in or_tri_element. */
typedef struct {
  int32_t kind;
  int32_t nv;
  int32_t cap;
  float rconst;
  const void *nodes;
  const uint32_t *indices;
  const float *V;
  const uint32_t *F;
  const float *R;
  double eps_scale;
  double T[16];
  double invT[16];
  double invT33[16];
} or_element;

/* RenderGraph::Trace per ray over elements of any kind: each element's hit
 * is its accelerator's own closest hit (its Traverse) of the ray moved into
 * its space; t = world distance, u/v/prim_id the element's. */
void or_scene_trace_mixed(const or_tlas_node *tlas, const or_element *elems,
                          const or_ray *rays, int64_t nrays, or_hit *hits, uint32_t *inst,
                          int threads);

/* Texture::Fetch of a dense 3D texture (byte / float / double, 1-4
 * interleaved components, optional per-axis remap tables): rgba[4 n]. */
void or_tex3d_fetch(const void *image, int format, int comps, const int dims[3],
                    const int clamp[3], const float *const remap[3], const int remap_n[3],
                    const float *uvw, int64_t n, float *rgba);

/* SparseVolumeAccel::Build + Sample / Texture::Fetch's sparse branch:
 * blocks = 9 ints per block (offset, extent, size), data[b] its voxels;
 * uvw (may be NULL) replaces pos by u * globalDim. values[4 n]. */
void or_sparse_sample(const int32_t *blocks, const void *const *data, int64_t nblocks,
                      int components, int format, const int global_dim[3], const double *pos,
                      const float *uvw, int64_t n, double *values);

/* BuildIntersection fields of a closest hit (80 bytes): position, geometric
 * and shading normal, tangent, binormal, vertex ids f0..f3; unset fields and
 * misses are 0. */
typedef struct {
  float position[3];
  uint32_t f0;
  float geometric[3];
  uint32_t f1;
  float normal[3];
  uint32_t f2;
  float tangent[3];
  uint32_t f3;
  float binormal[3];
  float pad;
} or_isect;

/* BuildIntersection for hits already traced: triangles
 * render/render_accel_triangle.cc:2439-2507, tetras (entry face from
 * ext->subface) render/render_accel_tetra.cc:2124-2196, spheres
 * render/render_accel_particle.cc:2968-3049. */
void or_isect_tri(const float *V, const uint32_t *F, const or_ray *rays, const or_hit *hits,
                  int64_t n, or_isect *out);
void or_isect_tet(const float *V, const uint32_t *F, const or_ray *rays, const or_hit *hits,
                  const or_tet_ext *ext, int64_t n, or_isect *out);
void or_isect_sph(const float *P, const or_ray *rays, const or_hit *hits, int64_t n,
                  or_isect *out);

#ifdef __cplusplus
}
#endif
#endif
