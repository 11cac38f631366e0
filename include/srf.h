/*
 * srf.h -- C ABI of the MI355X-native acceleration-structure build and ray
 * intersection library (libsrf.so).
 *
 * It replaces SURFACE's CPU accelerators behind gles_accel_builder and
 * gles_raytrace_engine. Every entry point below names the reference
 * interface it stands in for (paths relative to the SURFACE checkout).
 *
 * Conventions
 *  - Array arguments are DEVICE pointers (hipMalloc'd or any pointer a HIP
 *    kernel on the context's device may dereference) unless the parameter
 *    name ends in `_host`.
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream). All
 *    device work of a call is enqueued on it; calls return once the work is
 *    enqueued, except where a comment says the call synchronises.
 *  - A context and everything built from it belong to one device. Calls on
 *    one context must be serialised by the caller (the build scratch space
 *    is shared); accelerators may be traced from any number of streams.
 *  - Errors: every call returns srf_status; srf_last_error_string() gives a
 *    thread-local message for the last failing call. The reference reports
 *    failure as `bool` (render/render_accel_triangle.h:78) and asserts.
 *  - Floating point follows the reference bit for bit: no FMA contraction,
 *    IEEE division and square root, the reference's float/double mixing.
 */
#ifndef SRF_H_
#define SRF_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SRF_ABI_VERSION 1

typedef enum srf_status {
  SRF_OK = 0,
  SRF_ERR_INVALID_ARG = 1,
  SRF_ERR_OUT_OF_MEMORY = 2,
  SRF_ERR_HIP = 3,
  /* fewer than 1024 primitives: the reference's Build32 switches to its
   * SAH Build for those (render/render_accel_triangle.cc:1394-1397), which
   * this library does not provide */
  SRF_ERR_UNSUPPORTED = 4,
  SRF_ERR_INTERNAL = 5
} srf_status;

/* Primitive kinds, numbered as the GL modes that select them in
 * glDrawElements (gles/gles_primitive.cc:84-368; GL_TETRAHEDRONS_EXT is
 * include/GLES2/gl2.h:63). */
typedef enum srf_prim_kind {
  SRF_PRIM_POINTS = 0x0000,       /* GL_POINTS -> ParticleAccel  */
  SRF_PRIM_LINES = 0x0001,        /* GL_LINES -> LineAccel (gles/gles_accel_builder.cc:686-760) */
  SRF_PRIM_TRIANGLES = 0x0004,    /* GL_TRIANGLES -> TriangleAccel */
  SRF_PRIM_TETRAHEDRONS = 0x0010, /* GL_TETRAHEDRONS_EXT -> TetraAccel */
  /* GL_PYRAMIDS_EXT / GL_PRISMS_EXT / GL_HEXAHEDRONS_EXT -> SolidAccel
   * (include/GLES2/gl2.h:64-66, gles/gles_accel_builder.cc:880-968) */
  SRF_PRIM_PYRAMIDS = 0x0011,
  SRF_PRIM_PRISMS = 0x0012,
  SRF_PRIM_HEXAHEDRONS = 0x0013
} srf_prim_kind;

typedef struct srf_ctx srf_ctx;
typedef struct srf_accel srf_accel;
typedef struct srf_volume srf_volume;
typedef struct srf_scene srf_scene;
typedef struct srf_sparse_volume srf_sparse_volume;

/* Ray: the fields of render/render_ray.h:47-151 that traversal reads.
 * flags bit 0 = double_sided (Ray::double_sided; eye rays use 1,
 * gles/gles_raytrace_engine.cc:807). prev_prim_id = Ray::prev_prim_id
 * (0xFFFFFFFF = none). 32 bytes. */
typedef struct srf_ray {
  float org[3];
  uint32_t flags;
  float dir[3];
  uint32_t prev_prim_id;
} srf_ray;
#define SRF_RAY_DOUBLE_SIDED 1u
#define SRF_NO_PRIM 0xFFFFFFFFu

/* Closest hit: Intersection::{t,u,v,prim_id} (render/render_intersection.h:18-62).
 * A miss leaves t = FLT_MAX and prim_id = SRF_NO_PRIM, as Traverse does. */
typedef struct srf_hit {
  float t, u, v;
  uint32_t prim_id;
} srf_hit;

/* Tetra extras at the entry point: InterpolateTetra weights d0..d3
 * (render/render_accel_tetra.cc:148-241), the varying f0*d0+f1*d1+f2*d2+f3*d3
 * (gles/gles_common.cc:1074-1076) when vertex scalars are given,
 * Intersection::subface_id and the exit distance of the hit tetra. */
typedef struct srf_tet_hit_ext {
  float d[4];
  float scalar;
  uint32_t subface;
  float t_leave;
  float pad;
} srf_tet_hit_ext;

/* Solid extras at the reported crossing (SolidAccel::TestLeafNode,
 * render/render_accel_solid.cc:2206-2333): interpolate() weights d0..d7
 * of the solid's vertices (:247-435, mode 0; entries past the vertex count
 * are 0), Intersection::position and ::normal of the crossing (the face
 * point the Pluecker weights give, not org + t dir; an exit crossing's
 * normal is negated), the varying sum over k of scalar[nv * prim + k] * d_k
 * (gles/gles_common.cc:1078-1140) when scalars are given, and side = 0 for
 * the entry crossing, 1 for the exit (a ray starting inside). 64 bytes. */
typedef struct srf_solid_hit_ext {
  float d[8];
  float position[3];
  float scalar;
  float normal[3];
  uint32_t side;
} srf_solid_hit_ext;

/* Line extras (LineAccel's BuildIntersection, render/render_accel_line.cc:
 * 654-702): Intersection::position = org + t dir, ::normal (the cap's axis
 * direction, or from the axis point at v to the hit point), and whether an
 * end cap was hit. hit.u / hit.v are CylinderIsect's (cap: distance from
 * the axis and 0 / 1; side: 0 and the axial fraction). 32 bytes. */
typedef struct srf_line_hit_ext {
  float position[3];
  uint32_t hit_cap;
  float normal[3];
  float pad;
} srf_line_hit_ext;

/* The Intersection fields BuildIntersection fills for a closest hit
 * (render/render_intersection.h:18-62): position, geometric and shading
 * normal, tangent, binormal and the vertex ids f0..f3. Fields a kind leaves
 * unset (triangles and tetras: tangent, binormal; triangles: f3) and every
 * field of a miss are 0. 80 bytes. */
typedef struct srf_isect {
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
} srf_isect;

/* Sphere self-intersection inputs, Ray::prev_hit_t / prev_hit_normal
 * (render/render_ray.h:137-139), used when prev_prim_id names a sphere. */
typedef struct srf_sphere_prev {
  double prev_hit_t;
  float prev_hit_normal[3];
  float pad;
} srf_sphere_prev;

/* Geometry descriptors: the device-side counterparts of Mesh
 * (render/render_prim_triangle.h:22-56), Tetrahedron
 * (render/render_prim_tetra.h:30-37) and Particles
 * (render/render_prim_particle.h:19-27). FP32 positions only. */
typedef struct srf_tri_desc {
  const float *vertices; /* xyz per vertex */
  uint64_t num_vertices;
  const uint32_t *faces; /* 3 per triangle */
  uint64_t num_faces;
} srf_tri_desc;

typedef struct srf_tet_desc {
  const float *vertices;
  uint64_t num_vertices;
  const uint32_t *indices; /* 4 per tetrahedron */
  uint64_t num_tetras;
} srf_tet_desc;

/* Solid (render/render_prim_solid.h:33-40), FP32 positions: vertices_per_solid
 * = 5 pyramids, 6 prisms, 8 hexahedra, with the vertex order of
 * kPyramidFaces / kPrismFaces / kHexaFaces (render/render_accel_solid.cc:96-99). */
typedef struct srf_solid_desc {
  const float *vertices;
  uint64_t num_vertices;
  const uint32_t *indices; /* vertices_per_solid per solid */
  uint64_t num_solids;
  uint32_t vertices_per_solid;
} srf_solid_desc;

/* Lines (render/render_prim_line.h:30-38), FP32 positions: segments hold two
 * vertex ids per line, radius one value per vertex (NULL: constant_radius);
 * cap = LineBuildOptions::cap (render/render_accel_line.h:52-60, default on:
 * the end discs are part of the primitive). */
typedef struct srf_line_desc {
  const float *positions;
  uint64_t num_vertices;
  const uint32_t *segments;
  uint64_t num_lines;
  const float *radius;
  float constant_radius;
  uint32_t cap;
} srf_line_desc;

typedef struct srf_sphere_desc {
  const float *positions;
  const float *radius; /* per particle, or NULL to use constant_radius */
  float constant_radius;
  uint64_t num_particles;
} srf_sphere_desc;

/* Built-tree summary; srf_accel_get_info synchronises the build stream. */
typedef struct srf_accel_info {
  uint32_t kind; /* srf_prim_kind */
  uint64_t num_prims;
  uint64_t num_nodes;
  uint64_t num_leaves;
  uint32_t max_depth;   /* branch levels, root = 1 */
  double eps_scale;     /* triangles: scene extent (render/render_accel_triangle.cc:1451-1463) */
  uint64_t node_bytes;  /* bytes per node in srf_accel_export's reference layout */
} srf_accel_info;

/* NodeInfo32 (render/render_prefix_tree_util.h:51-61): left child =
 * child_index, right = child_index + 1; *_leaf = NODE_TYPE_LEAF. */
typedef struct srf_node_info {
  uint32_t child_index;
  uint32_t left_leaf;
  uint32_t right_leaf;
} srf_node_info;

/* One element of a two-level scene: RenderElement's accelerator and
 * matrices (gles/gles_render_graph.h:26-118). All three matrices are
 * row-major [4][4] in the reference's row-vector convention
 * (render/render_matrix.h:165-183: p' = p * M, row 3 = translation):
 * T local->world, invT = Inverse(T), invT33 = Inverse(T with row 3
 * cleared), as RenderElement::Update computes them
 * (gles/gles_render_graph.cc:81-101). */
typedef struct {
  const srf_accel *accel; /* any kind */
  double T[16];
  double invT[16];
  double invT33[16];
} srf_instance;

/* RenderElement::Update's matrices (gles/gles_render_graph.cc:81-101): from
 * inst->T, invT = Matrixd::Inverse(T) and invT33 = Matrixd::Inverse(T with
 * row 3 cleared), by the same Cramer's-rule inverse
 * (render/render_matrix.h:40-134), so a scene gets the reference's
 * matrices bit for bit. Host only. */
srf_status srf_instance_update(srf_instance *inst);

/* A top-level node as BVHTree keeps it (BVHNode, render/render_bvh_tree.h):
 * box, children (0 = none), split axis, leaf flag, element index. 72 B. */
typedef struct {
  double bmin[3];
  double bmax[3];
  int32_t child[2];
  int32_t axis;
  int32_t is_leaf;
  uint32_t node_id;
  uint32_t pad;
} srf_tlas_node;

/* ---------------------------------------------------------------- context */
const char *srf_last_error_string(void);
int srf_abi_version(void);
srf_status srf_ctx_create(int device, srf_ctx **out);
srf_status srf_ctx_destroy(srf_ctx *ctx);

/* ------------------------------------------------------------------ build */
/* TriangleAccel::Build32(const Mesh*, const TriangleBuildOptions&)
 * (render/render_accel_triangle.h:78, .cc:1389-1690), called from
 * AccelBuilder::AddTriangleData (gles/gles_accel_builder.cc:572). The
 * accelerator owns device copies of everything traversal needs; the
 * caller's buffers may be freed once the stream has passed the build. */
srf_status srf_accel_build_triangles(srf_ctx *ctx, const srf_tri_desc *desc,
                                     void *stream, srf_accel **out);
/* TetraAccel::Build32 (render/render_accel_tetra.h:91, .cc:1579-1898),
 * from AccelBuilder::AddTetraData (gles/gles_accel_builder.cc:867). */
srf_status srf_accel_build_tetras(srf_ctx *ctx, const srf_tet_desc *desc,
                                  void *stream, srf_accel **out);
/* SolidAccel::Build32 (render/render_accel_solid.h:91, .cc:1795-2124). The
 * reference's AccelBuilder::AddSolidData calls the SAH Build instead
 * (gles/gles_accel_builder.cc:967-968, Build32 left commented out); this is
 * the LBVH Build32 path, whose tree srf_accel_export reproduces. */
srf_status srf_accel_build_solids(srf_ctx *ctx, const srf_solid_desc *desc,
                                  void *stream, srf_accel **out);
/* Lines on the same LBVH as the other kinds (Morton code of the segment
 * midpoint, leaves of up to 5 lines). The reference has only a top-down
 * spatial-median LineAccel::Build (render/render_accel_line.h:79,
 * .cc:167-341), from AccelBuilder::AddLineData
 * (gles/gles_accel_builder.cc:686-760); its tree is not reproduced. Traversal
 * finds the closest crossing over all lines, which is what LineAccel::Traverse
 * returns whenever its boxes hold the cylinders it tests (always for a
 * constant radius). */
srf_status srf_accel_build_lines(srf_ctx *ctx, const srf_line_desc *desc, void *stream,
                                 srf_accel **out);
/* ParticleAccel::Build32 (render/render_accel_particle.h:77,
 * .cc:1560-1866), from AccelBuilder::AddParticleData
 * (gles/gles_accel_builder.cc:673). */
srf_status srf_accel_build_spheres(srf_ctx *ctx, const srf_sphere_desc *desc,
                                   void *stream, srf_accel **out);
/* Many meshes in one segmented build: each mesh gets the tree
 * TriangleAccel::Build32 would give it alone, but every stage is one launch
 * over all meshes. The reference builds draw calls one by one
 * (gles/gles_accel_builder.cc:34-117, one Build32 per AddTriangleData).
 * out[i] receives mesh i's accelerator; on failure none is returned. */
srf_status srf_accel_build_triangles_batch(srf_ctx *ctx, uint32_t num_meshes,
                                           const srf_tri_desc *descs_host,
                                           void *stream, srf_accel **out);
/* The batch rebuilt in place (accels from srf_accel_build_triangles_batch or
 * srf_accel_build_triangles, all distinct), as AccelBuilder's per-frame
 * rebuild of dynamic draws does one by one (gles/gles_accel_builder.cc:442-455). */
srf_status srf_accel_rebuild_triangles_batch(srf_ctx *ctx, uint32_t num_meshes,
                                             srf_accel *const *accels,
                                             const srf_tri_desc *descs_host, void *stream);

/* Build32 with HOST arrays, as AccelBuilder hands the reference its GL client
 * arrays (gles/gles_accel_builder.cc:545-572 -> TriangleAccel::Build32,
 * render/render_accel_triangle.h:78; TetraAccel / ParticleAccel / SolidAccel
 * ::Build32 and LineAccel::Build likewise). kind selects the descriptor
 * desc points to (srf_tri_desc for SRF_PRIM_TRIANGLES, srf_tet_desc,
 * srf_sphere_desc for SRF_PRIM_POINTS, srf_solid_desc for the three solid
 * kinds, srf_line_desc), whose arrays are in host memory. The arrays are
 * copied into device memory the accelerator owns (kept for
 * BuildIntersection); the call synchronises with the build, so the host
 * arrays may be reused once it returns. */
srf_status srf_accel_build_host(srf_ctx *ctx, srf_prim_kind kind, const void *desc,
                                void *stream, srf_accel **out);
/* srf_accel_rebuild with host arrays (same kind as the accelerator);
 * reuses the accelerator's device copies when the new arrays fit;
 * synchronises. */
srf_status srf_accel_rebuild_host(srf_ctx *ctx, srf_accel *accel, const void *desc,
                                  void *stream);

/* Build32 called again on an existing accelerator object (the reference
 * rebuilds in place: nodes_.clear(), render/render_accel_triangle.cc:1552;
 * AccelBuilder rebuilds dynamic draws every frame,
 * gles/gles_accel_builder.cc:442-455). desc must point to the descriptor
 * type of the accelerator's kind; device buffers are reused when the new
 * geometry fits, so no allocation or synchronisation happens then. Work
 * that still reads the old tree on another stream must be complete. */
srf_status srf_accel_rebuild(srf_ctx *ctx, srf_accel *accel, const void *desc,
                             void *stream);

/* TriangleAccel/TetraAccel/ParticleAccel::BoundingBox(double[3], double[3])
 * (render/render_accel_triangle.h:90), used for the top-level BVH
 * (gles/gles_primitive.cc:174,215,271). Synchronises. */
srf_status srf_accel_bounds(const srf_accel *accel, double bmin_host[3],
                            double bmax_host[3]);
srf_status srf_accel_get_info(const srf_accel *accel, srf_accel_info *info_host);
/* Tree in the reference's node layout and its permuted primitive indices:
 * the contents TriangleAccel::Dump writes (render/render_accel_triangle.cc:
 * 1692-1719), nodes in depth-first preorder. nodes_host must hold
 * info.num_nodes * info.node_bytes bytes, indices_host num_prims words.
 * Synchronises. */
srf_status srf_accel_export(const srf_accel *accel, void *nodes_host,
                            uint32_t *indices_host);
/* {Triangle,Tetra,Particle}Accel::Dump (render/render_accel_triangle.cc:
 * 1692-1719, render/render_accel_tetra.cc:1900-1928,
 * render/render_accel_particle.cc:1868-1896): unsigned long long node count,
 * the nodes in the reference layout (srf_accel_export), unsigned long long
 * index count, the 32-bit indices. A leaf's unused fields (axis, second box
 * slot), which the reference leaves uninitialised, are zero. Synchronises. */
srf_status srf_accel_dump(const srf_accel *accel, const char *path);
srf_status srf_accel_destroy(srf_accel *accel);

/* ------------------------------------------------------------------ trace */
/* {Triangle,Tetra,Particle}Accel::Traverse(Intersection&, Ray&) for a batch
 * of rays (render/render_accel_triangle.h:87), the loop the reference runs
 * per pixel under OpenMP (gles/gles_raytrace_engine.cc:662-880). */
srf_status srf_trace_closest(const srf_accel *accel, const srf_ray *rays,
                             uint64_t num_rays, srf_hit *hits, void *stream);
/* TetraAccel::Traverse plus BuildIntersection / InterpolateTetra
 * (render/render_accel_tetra.cc:2124-2289); vertex_scalars (one float per
 * vertex of the build's vertex array, may be NULL) gives ext->scalar. */
srf_status srf_trace_tetra(const srf_accel *accel, const srf_ray *rays,
                           uint64_t num_rays, const float *vertex_scalars,
                           srf_hit *hits, srf_tet_hit_ext *ext, void *stream);
/* SolidAccel::Traverse (render/render_accel_solid.cc:2345-2435) with its
 * BuildIntersection; index_scalars (one float per index position, i.e.
 * num_solids * vertices_per_solid, may be NULL) gives ext->scalar, ext may
 * be NULL. The reference never writes a solid hit's u and v (they come
 * from an uninitialised Intersection); they are 0 here. */
srf_status srf_trace_solids(const srf_accel *accel, const srf_ray *rays,
                            uint64_t num_rays, const float *index_scalars,
                            srf_hit *hits, srf_solid_hit_ext *ext, void *stream);
/* LineAccel::Traverse (render/render_accel_line.cc:704-774) with
 * CylinderIsect (:488-577) and BuildIntersection; ext may be NULL. */
srf_status srf_trace_lines(const srf_accel *accel, const srf_ray *rays, uint64_t num_rays,
                           srf_hit *hits, srf_line_hit_ext *ext, void *stream);
/* ParticleAccel::Traverse with the self-intersection inputs of each ray
 * (render/render_accel_particle.cc:2423-2512); prev may be NULL. */
srf_status srf_trace_spheres(const srf_accel *accel, const srf_ray *rays,
                             const srf_sphere_prev *prev, uint64_t num_rays,
                             srf_hit *hits, void *stream);
/* BuildIntersection for a batch of closest hits (hits from
 * srf_trace_closest / srf_trace_tetra / srf_trace_spheres on the same rays):
 * triangles render/render_accel_triangle.cc:2439-2507 (f0..f2, position =
 * org + t dir, normalised geometric normal), tetras
 * render/render_accel_tetra.cc:2124-2196 (f0..f3, position, normal of the
 * entry face ext->subface; the weights d0..d3 are in srf_tet_hit_ext),
 * spheres render/render_accel_particle.cc:2968-3049 (position, normal,
 * tangent, binormal; u and v are in the hit). The reference runs it inside
 * Traverse for every hit; here it is a separate pass that reads the caller's
 * geometry (device pointers in the descriptor, as Traverse reads the Mesh it
 * was built from), so rays that only need t and prim_id skip it. */
srf_status srf_build_intersection_triangles(const srf_tri_desc *mesh, const srf_ray *rays,
                                            const srf_hit *hits, uint64_t n, srf_isect *out,
                                            void *stream);
srf_status srf_build_intersection_tetras(const srf_tet_desc *tetras, const srf_ray *rays,
                                         const srf_hit *hits, const srf_tet_hit_ext *ext,
                                         uint64_t n, srf_isect *out, void *stream);
srf_status srf_build_intersection_spheres(const srf_sphere_desc *particles,
                                          const srf_ray *rays, const srf_hit *hits, uint64_t n,
                                          srf_isect *out, void *stream);

/* {Triangle,Tetra,Particle,Solid,Line}Accel::Traverse(Intersection&, Ray&)
 * (render/render_accel_triangle.h:87) for rays in HOST memory, as
 * RaytraceEngine calls it per pixel from its OpenMP loop
 * (gles/gles_raytrace_engine.cc:662-880); n = 1 is the per-ray form. The
 * rays go to the GPU in one batch and the results come back before the call
 * returns. hits_host gets t, u, v, prim_id; ext_host (may be NULL) the
 * kind's extras (srf_tet_hit_ext, srf_solid_hit_ext or srf_line_hit_ext,
 * without the tetra scalar varying); isect_host (may be NULL; triangles,
 * tetras, spheres) the BuildIntersection fields, computed from the geometry
 * of the last build (a device build's arrays must still be alive, as the
 * reference's Mesh must). Thread-safe: each host thread has its own stream
 * and device staging buffer, kept until the process ends. */
srf_status srf_trace_closest_host(const srf_accel *accel, const srf_ray *rays_host, uint64_t n,
                                  srf_hit *hits_host, void *ext_host, srf_isect *isect_host);

/* Occlusion: occluded[i] = 1 iff Traverse would report a hit with
 * t < t_max[i] (t_max NULL = FLT_MAX). The reference answers shadow() with
 * a full closest-hit Traverse (gles/gles_glsl_runtime.cc:179-217); this
 * stops at the first qualifying primitive instead. */
srf_status srf_trace_any(const srf_accel *accel, const srf_ray *rays,
                         const float *t_max, uint64_t num_rays,
                         uint32_t *occluded, void *stream);

/* Testing switch, no counterpart in the reference: with on != 0 the
 * triangle and tetra traversals of this accelerator use the kernels and
 * stack entry layouts meant for trees of 2^24 or more nodes (32-bit node
 * indices), so that they can be checked on trees small enough for the
 * oracle. Results are identical either way; only speed differs. */
srf_status srf_debug_force_wide_stack(srf_accel *accel, int on);

/* ------------------------------------------------------- LBVH stages */
/* Occlusion with ray sorting: the rays are traced in the order of a sort
 * key of key_bits bits (0: in the given order) -- the direction octant over
 * the 3D Morton code of the origin in the accelerator's scene box -- and
 * each flag goes back to its ray's own slot. mode: SRF_LAUNCH_PLAIN or
 * SRF_LAUNCH_PERSISTENT (a persistent grid taking 64 sorted rays per
 * wavefront from a work counter). Keys, order and counter live in ctx's
 * scratch. Triangles and spheres. Each flag equals srf_trace_any's. */
srf_status srf_trace_any_sorted(srf_ctx *ctx, const srf_accel *accel, const srf_ray *rays,
                                const float *t_max, uint64_t n, uint32_t key_bits, int mode,
                                uint32_t *occluded, void *stream);

/* srf_trace_any_sorted without allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It takes no
 * context: the keys, order, work counter and sort slices live in a static
 * device array of 26 MiB that each GPU the library runs on holds until the
 * process ends, enough for n <= 1048576 rays (SRF_NOALLOC_MAX_RAYS); larger n
 * is rejected with SRF_ERR_INVALID_ARG. Nothing is allocated or freed. Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialised by the caller. The
 * flags equal srf_trace_any_sorted's and srf_trace_any's. */
#define SRF_NOALLOC_MAX_RAYS 1048576u
srf_status srf_trace_any_sorted_noalloc(const srf_accel *accel, const srf_ray *rays,
                                        const float *t_max, uint64_t n, uint32_t key_bits,
                                        int mode, uint32_t *occluded, void *stream);

typedef struct {
  uint32_t key_bits;   /* 0 = no sort */
  uint32_t persistent; /* 0 or 1 */
  float ms;            /* measured time of the tuning batch */
} srf_any_config;

/* Autotuner for srf_trace_any_sorted: times key widths {0, 9, 15, 21, 27,
 * 30} with plain and persistent launches on the given rays and returns the
 * fastest. Synchronises. No counterpart in the reference (its rays are
 * traced one at a time). */
srf_status srf_autotune_any(srf_ctx *ctx, const srf_accel *accel, const srf_ray *rays,
                            const float *t_max, uint64_t n, uint32_t *occluded, void *stream,
                            srf_any_config *best);

/* Ambient-occlusion rays as a fragment shader issues them through shadow()
 * (gles/gles_glsl_runtime.cc:179-217), for a batch of sphere hits: for
 * every primary ray i with a hit on sphere c, per_hit cosine-weighted
 * directions about the outward normal n = normalize(p - c), p = o + t d,
 * drawn from a PCG32 stream seeded with (seed, i); origin p + eps r n,
 * t_max = dist, prev_prim_id = c (so the sphere itself is never reported).
 * Up to n * per_hit rays are written (out_pixel[j] = i); the count lands in
 * *out_count (device memory). Slots are handed out per wavefront, so the
 * order of the groups is not fixed. */
srf_status srf_ao_rays_spheres(const srf_ray *primary, const srf_hit *hits, uint64_t n,
                               const float *positions, const float *radius,
                               float constant_radius, uint32_t per_hit, float eps, float dist,
                               uint64_t seed, srf_ray *out_rays, float *out_tmax,
                               uint32_t *out_pixel, uint32_t *out_count, void *stream);

/* CalculateMortonCodes{TriangleFloat,TetraFloat}30 / CalculateMortonCodes30SIMD
 * (render/render_prefix_tree_util.cc:176-304) and CalculateMortonCodesSolidFloat30
 * (:347-382, the three solid kinds) for the scene box bmin/bmax (host
 * arrays). desc points to the srf_*_desc matching kind. */
srf_status srf_lbvh_morton30(srf_ctx *ctx, srf_prim_kind kind, const void *desc,
                             const float bmin_host[3], const float bmax_host[3],
                             uint32_t *codes, void *stream);
/* RadixSort30OMP (render/render_prefix_tree_util.h:296-402): stable sort of
 * {index = 0..n-1, code} by code. */
srf_status srf_lbvh_sort30(srf_ctx *ctx, const uint32_t *codes, uint64_t n,
                           uint32_t *sorted_codes, uint32_t *sorted_index,
                           void *stream);
/* ConstructBinaryRadixTree30 for i in [0, n-2]
 * (render/render_prefix_tree_util.cc:562-649). */
srf_status srf_lbvh_radix_tree30(srf_ctx *ctx, const uint32_t *sorted_codes,
                                 uint64_t n, srf_node_info *out, void *stream);

/* ----------------------------------------------------------------- volume */
/* ---- two-level scenes (RenderGraph) ---- */
/* RenderGraph::Build (gles/gles_render_graph.cc:103-143): world boxes of the
 * elements (TransformBoundingBox of each accelerator's BoundingBox(), :19-79)
 * and the top-level BVHTree over them (render/render_bvh_tree.cc:319-502),
 * built on the host; waits for the accelerators' builds. Elements may be
 * accelerators of any kind (triangles, tetrahedra, spheres, solids, lines),
 * as RenderGraph::Trace switches on the element's primitive type
 * (:224-264). The scene keeps the accelerators' device arrays: rebuilding
 * one of them requires a new scene, as the reference rebuilds its graph. */
srf_status srf_scene_create(srf_ctx *ctx, uint32_t num_instances,
                            const srf_instance *instances_host, srf_scene **out);
srf_status srf_scene_destroy(srf_scene *scene);
/* BVHTree::BoundingBox (render/render_bvh_tree.cc:656-669). */
srf_status srf_scene_bounds(const srf_scene *scene, double bmin_host[3], double bmax_host[3]);
/* The top-level BVH: *count nodes; nodes_host (NULL to query the count)
 * receives them in BVHTree order. */
srf_status srf_scene_export(const srf_scene *scene, srf_tlas_node *nodes_host, uint32_t *count);
/* Launch shape of the scene trace. PLAIN: one lane per ray, one block per
 * block_size rays. PERSISTENT: blocks_per_cu resident blocks per compute
 * unit, each wavefront taking 64-ray tiles from a work counter until none is
 * left (the counter is reset by a kernel, so the launch can be captured in a
 * hipGraph). TUNED: the configuration srf_scene_autotune picked (before any
 * tuning: persistent, 128 threads, 8 blocks per CU). */
typedef enum {
  SRF_LAUNCH_TUNED = 0,
  SRF_LAUNCH_PLAIN = 1,
  SRF_LAUNCH_PERSISTENT = 2
} srf_launch_mode;

typedef struct {
  uint32_t block_size;    /* 64, 128 or 256 */
  uint32_t persistent;    /* 0 or 1 */
  uint32_t blocks_per_cu; /* persistent grid: resident blocks per CU */
  float ms;               /* measured time of the tuning batch */
} srf_launch_config;

/* RenderGraph::Trace (gles/gles_render_graph.cc:145-314) for a batch of
 * rays: hits[i] = {world distance, u, v, prim_id} of the nearest element
 * hit, instance_ids[i] its element (~0 on a miss; may be NULL). The ray's
 * prev_prim_id is not used (no element is the ray's previous one).
 * mode: srf_launch_mode. A scene of triangle accelerators only is one
 * kernel (capturable in a hipGraph). A scene mixing kinds is traced in
 * element passes: each element's rays go to that accelerator's own
 * closest-hit launch (srf_trace_closest), then one kernel walks the top
 * level per ray over those hits; it allocates scratch (about 4 bytes per
 * element and ray plus 48 per ray entering an element's box), synchronises
 * the stream once per chunk of 2^26 / elements rays, cannot be captured in
 * a graph, and serialises calls on the same scene. A solid element's u and
 * v are 0 (the reference never writes them). */
srf_status srf_scene_trace(const srf_scene *scene, const srf_ray *rays, uint64_t n,
                           srf_hit *hits, uint32_t *instance_ids, int mode, void *stream);
/* Autotuner: traces the given rays with every launch configuration (block
 * size 64/128/256, plain or persistent with 2/4/8/16 blocks per CU), keeps
 * the fastest for SRF_LAUNCH_TUNED and reports it. Synchronises. No
 * counterpart in the reference (its CPU loop has no launch shape). */
srf_status srf_scene_autotune(srf_scene *scene, const srf_ray *rays, uint64_t n, srf_hit *hits,
                              uint32_t *instance_ids, void *stream, srf_launch_config *best);

/* glTexImage3D / lsglTexImage3DPointer for a single-component FP32 volume
 * (gles/gles_texture.cc:198-357); wrap_* = 0 REPEAT (the glTexImage3D
 * default, gles/gles_common.cc:1274-1276) or 1 CLAMP_TO_EDGE
 * (render/render_texture.h:23-24). voxels (x fastest) are referenced, not
 * copied, like lsglTexImage3DPointer's retained buffer. */
srf_status srf_volume_create(srf_ctx *ctx, const float *voxels, int width,
                             int height, int depth, int wrap_s, int wrap_t,
                             int wrap_r, srf_volume **out);
srf_status srf_volume_destroy(srf_volume *vol);

/* One block of a sparse volume, VolumeBlock (render/render_prim_volume.h:
 * 18-26): integer offset and extent in the global grid, cell count size[]
 * (the block's voxels, x fastest, `components` interleaved per voxel), and
 * its voxel data (DEVICE pointer, referenced). */
typedef struct srf_volume_block {
  int32_t offset[3];
  int32_t extent[3];
  int32_t size[3];
  int32_t id;
  const void *data;
} srf_volume_block;

/* SparseVolumeAccel::Build (render/render_accel_volume.cc:35-57): the
 * BVHTree over the blocks' boxes (built on the host, as the reference does).
 * global_dim = SparseVolume::globalDim (for srf_sparse_volume_fetch);
 * data_type SRF_TEX_BYTE / _FLOAT / _DOUBLE for all blocks. */
srf_status srf_sparse_volume_create(srf_ctx *ctx, const srf_volume_block *blocks_host,
                                    uint32_t num_blocks, const int global_dim[3],
                                    int components, int data_type, srf_sparse_volume **out);
srf_status srf_sparse_volume_destroy(srf_sparse_volume *vol);
/* SparseVolumeAccel::Sample (render/render_accel_volume.cc:59-139) at n
 * positions (3 doubles each): values[4 i..] = the nearest voxel of the first
 * block BVHTree::Locate finds (byte / 255, the rest 0, alpha 1 unless the
 * volume has 4 components), or (0, 0, 0, 0) outside every block. As in the
 * reference, the local offset is scaled by size[0] / extent[0] on all three
 * axes. */
srf_status srf_sparse_volume_sample(const srf_sparse_volume *vol, const double *positions,
                                    uint64_t n, double *values, void *stream);
/* Texture::Fetch's sparse branch (gles/gles_common.h:346-371): position =
 * (u, v, r) * globalDim in float, Sample, RGBA as float. */
srf_status srf_sparse_volume_fetch(const srf_sparse_volume *vol, const float *uvw, uint64_t n,
                                   float *rgba, void *stream);

/* glTexImage3D / lsglTexImage3DPointer for any dense 3D texture the
 * reference's Texture::Fetch samples (gles/gles_common.h:344-403): data_type
 * is LSGL_RENDER_TEXTURE3D_FORMAT_* (render/render_texture.h:16-21):
 * SRF_TEX_BYTE (FilterTexture3DByte, voxel / 255.5f,
 * render/render_texture.cc:832-944), SRF_TEX_FLOAT (FilterTexture3DFloat,
 * :946-1058) or SRF_TEX_DOUBLE (FilterTexture3DDouble, :1060-1172);
 * components = 1..4 interleaved per texel (x fastest); wrap_* as
 * srf_volume_create. image is referenced, not copied. The object is an
 * srf_volume: srf_volume_destroy frees it; the march and srf_volume_sample
 * accept it when it is FP32 with one component. */
#define SRF_TEX_BYTE 1
#define SRF_TEX_FLOAT 8
#define SRF_TEX_DOUBLE 16
srf_status srf_texture3d_create(srf_ctx *ctx, const void *image, int data_type, int components,
                                int width, int height, int depth, int wrap_s, int wrap_t,
                                int wrap_r, srf_volume **out);
/* The per-axis coordinate remap of Texture::Fetch (local::remap,
 * gles/gles_common.h:263-269, applied at :377-389 before filtering): the
 * coordinate x becomes table[clamp((int)(x * n), 0, n - 1)]. axis 0..2
 * (s, t, r); table is a device array of n floats, referenced; NULL clears. */
srf_status srf_texture3d_remap(srf_volume *tex, int axis, const float *table, int n);
/* Texture::Fetch at n coordinates uvw (3 floats each): rgba[4 i..4 i + 3];
 * a single component is splatted to all four, as the reference does. */
srf_status srf_texture3d_fetch(const srf_volume *tex, const float *uvw, uint64_t n, float *rgba,
                               void *stream);

/* texture3D -> FilterTexture3DFloat (render/render_texture.cc:946-1058) at
 * n texture coordinates uvw (3 floats each); out[i] = sampled value. */
srf_status srf_volume_sample(const srf_volume *vol, const float *uvw,
                             uint64_t n, float *out, void *stream);
/* Emission-absorption march through the unit cube, step in texture units,
 * transfer function tf (tf_size RGBA entries); rgba[4*i..] = (C, 1 - T).
 * The reference marches in user GLSL (examples/sparse_volume_render/
 * input.frag:99-141); the march itself is defined by this library.
 * checkpoints (optional, NULL = none): num_rays * slots floats, slots from
 * srf_volume_checkpoint_slots; receives each ray's transmittance at the
 * start of every 16-sample segment and after the last sample, for
 * srf_volume_march_bwd. */
srf_status srf_volume_march_fwd(const srf_volume *vol, const float *tf,
                                int tf_size, float step, const srf_ray *rays,
                                uint64_t num_rays, float *rgba,
                                float *checkpoints, void *stream);
/* Checkpoint floats per ray srf_volume_march_fwd writes for this step. */
srf_status srf_volume_checkpoint_slots(const srf_volume *vol, float step,
                                       uint32_t *slots_host);
/* Adjoint of srf_volume_march_fwd for L = sum(grad_rgba * rgba): ACCUMULATES
 * dL/dvoxel into grad_voxels (width*height*depth floats) and dL/dtf into
 * grad_tf (4*tf_size floats). Each ray is replayed in reverse from its
 * segment checkpoints (T_k = T_{k+1} / (1 - a_k), re-marched forward from
 * the segment start where 1 - a_k is small); no per-sample state is
 * stored. checkpoints: the forward pass's (same step, tf and rays), or NULL
 * to recompute them with one extra forward replay. */
srf_status srf_volume_march_bwd(const srf_volume *vol, const float *tf,
                                int tf_size, float step, const srf_ray *rays,
                                uint64_t num_rays, const float *checkpoints,
                                const float *grad_rgba, float *grad_voxels,
                                float *grad_tf, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SRF_H_ */
