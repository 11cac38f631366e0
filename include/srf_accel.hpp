// srf_accel.hpp -- header-only C++ host mirror of SURFACE's accelerator
// classes over the srf C ABI (srf.h). Names and meaning follow the
// reference so a caller such as gles_accel_builder / gles_raytrace_engine
// can swap types:
//
//   lsgl::render::TriangleAccel   render/render_accel_triangle.h:69-105
//   lsgl::render::TetraAccel      render/render_accel_tetra.h:82-125
//   lsgl::render::ParticleAccel   render/render_accel_particle.h:68-110
//   lsgl::render::SolidAccel      render/render_accel_solid.h:82-125
//   lsgl::render::LineAccel       render/render_accel_line.h:73-103
//
// Build32 returns false on failure (the reference's contract) and leaves the
// reason in LastError(). Build32 takes device arrays; Build32Host takes host
// arrays as the reference's Build32(const Mesh*) does. Traverse comes in
// three forms: a batch of device-resident rays (the fast path: generate,
// trace the batch, shade each hit on the host), a batch of host rays
// (TraverseHost), and the reference's per-ray signature
// Traverse(Intersection&, Ray&) -- one ray to the GPU and back, reentrant
// from many host threads, for callers that are not batched yet.
#ifndef SRF_ACCEL_HPP_
#define SRF_ACCEL_HPP_

#include <string>

#include "srf.h"

namespace srf {

/// The fields of lsgl::render::Ray (render/render_ray.h:47-151) that
/// Traverse reads.
struct Ray {
  float org[3];
  float dir[3];
  int double_sided = 1;
  unsigned int prev_prim_id = SRF_NO_PRIM;
};

/// lsgl::render::Intersection (render/render_intersection.h:18-62) without
/// the solid-only f4..f7 / d4..d7 and the render-element pointer.
struct Intersection {
  float t, u, v;
  unsigned int prim_id;
  unsigned int subface_id;
  float position[3], geometric[3], normal[3], tangent[3], binormal[3];
  unsigned int f0, f1, f2, f3;
  float d0, d1, d2, d3;
};

class Context {
 public:
  explicit Context(int device = 0) { status_ = srf_ctx_create(device, &ctx_); }
  ~Context() {
    if (ctx_) srf_ctx_destroy(ctx_);
  }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  bool ok() const { return status_ == SRF_OK; }
  srf_ctx *get() const { return ctx_; }

 private:
  srf_ctx *ctx_ = nullptr;
  srf_status status_;
};

namespace detail {

class AccelBase {
 public:
  explicit AccelBase(Context &ctx) : ctx_(ctx) {}
  ~AccelBase() {
    if (accel_) srf_accel_destroy(accel_);
  }
  AccelBase(const AccelBase &) = delete;
  AccelBase &operator=(const AccelBase &) = delete;

  /// BoundingBox(double bmin[3], double bmax[3]); synchronises with the build.
  void BoundingBox(double bmin[3], double bmax[3]) const {
    if (!accel_ || srf_accel_bounds(accel_, bmin, bmax) != SRF_OK) {
      for (int k = 0; k < 3; k++) {
        bmin[k] = 1.7976931348623157e308;
        bmax[k] = -1.7976931348623157e308;
      }
    }
  }

  /// Closest hit for rays[0..n) (device arrays) on `stream` (hipStream_t).
  bool Traverse(const srf_ray *rays, uint64_t n, srf_hit *hits, void *stream = nullptr) {
    return remember(srf_trace_closest(accel_, rays, n, hits, stream));
  }

  /// Closest hits of host rays (srf_trace_closest_host): hits, the kind's
  /// ext records (may be null) and the BuildIntersection fields (may be
  /// null) land in host memory before the call returns.
  bool TraverseHost(const srf_ray *rays_host, uint64_t n, srf_hit *hits_host,
                    void *ext_host = nullptr, srf_isect *isect_host = nullptr) {
    return remember(srf_trace_closest_host(accel_, rays_host, n, hits_host, ext_host,
                                           isect_host));
  }

  /// Traverse(Intersection&, Ray&) (render/render_accel_triangle.h:87): the
  /// closest hit of one host ray with its BuildIntersection fields; false on
  /// a miss (isect.t = FLT_MAX, prim_id = SRF_NO_PRIM, as the reference
  /// leaves them) or on failure (LastError()). Reentrant.
  bool Traverse(Intersection &isect, const Ray &ray) const {
    srf_ray r;
    for (int k = 0; k < 3; k++) {
      r.org[k] = ray.org[k];
      r.dir[k] = ray.dir[k];
    }
    r.flags = ray.double_sided ? SRF_RAY_DOUBLE_SIDED : 0u;
    r.prev_prim_id = ray.prev_prim_id;
    srf_hit h;
    srf_tet_hit_ext te{};
    srf_isect is{};
    const bool tet = kind_ == SRF_PRIM_TETRAHEDRONS;
    const bool fields = kind_ == SRF_PRIM_TRIANGLES || tet || kind_ == SRF_PRIM_POINTS;
    if (srf_trace_closest_host(accel_, &r, 1, &h, tet ? &te : nullptr, fields ? &is : nullptr) !=
        SRF_OK)
      return false;
    isect.t = h.t;
    isect.u = h.u;
    isect.v = h.v;
    isect.prim_id = h.prim_id;
    if (h.prim_id == SRF_NO_PRIM) return false;
    isect.subface_id = tet ? te.subface : 0u;
    for (int k = 0; k < 3; k++) {
      isect.position[k] = is.position[k];
      isect.geometric[k] = is.geometric[k];
      isect.normal[k] = is.normal[k];
      isect.tangent[k] = is.tangent[k];
      isect.binormal[k] = is.binormal[k];
    }
    isect.f0 = is.f0;
    isect.f1 = is.f1;
    isect.f2 = is.f2;
    isect.f3 = is.f3;
    isect.d0 = te.d[0];
    isect.d1 = te.d[1];
    isect.d2 = te.d[2];
    isect.d3 = te.d[3];
    return true;
  }

  /// Occlusion (any hit closer than t_max[i]) for rays[0..n).
  bool TraverseAny(const srf_ray *rays, const float *t_max, uint64_t n, uint32_t *occluded,
                   void *stream = nullptr) {
    return remember(srf_trace_any(accel_, rays, t_max, n, occluded, stream));
  }

  /// Tree in the reference layout and the permuted primitive indices, as
  /// TriangleAccel::Dump writes them; synchronises.
  bool Export(void *nodes_host, uint32_t *indices_host) const {
    return srf_accel_export(accel_, nodes_host, indices_host) == SRF_OK;
  }

  /// TriangleAccel::Dump / TetraAccel::Dump / ParticleAccel::Dump: the tree
  /// and indices in the reference's file format.
  bool Dump(const char *filename) { return remember(srf_accel_dump(accel_, filename)); }

  srf_accel *handle() const { return accel_; }
  const std::string &LastError() const { return error_; }

 protected:
  bool remember(srf_status s) {
    if (s != SRF_OK) error_ = srf_last_error_string();
    return s == SRF_OK;
  }
  template <class Desc, class Fn>
  bool build(const Desc &d, Fn first_build, void *stream) {
    if (accel_) return remember(srf_accel_rebuild(ctx_.get(), accel_, &d, stream));
    return remember(first_build(ctx_.get(), &d, stream, &accel_));
  }
  template <class Desc>
  bool build_host(srf_prim_kind kind, const Desc &d, void *stream) {
    if (accel_) return remember(srf_accel_rebuild_host(ctx_.get(), accel_, &d, stream));
    return remember(srf_accel_build_host(ctx_.get(), kind, &d, stream, &accel_));
  }

  Context &ctx_;
  srf_accel *accel_ = nullptr;
  srf_prim_kind kind_ = SRF_PRIM_TRIANGLES;
  std::string error_;
};

}  // namespace detail

/// TriangleAccel: Build32(const Mesh*) takes device arrays here.
class TriangleAccel : public detail::AccelBase {
 public:
  explicit TriangleAccel(Context &ctx) : AccelBase(ctx) { kind_ = SRF_PRIM_TRIANGLES; }
  bool Build32(const float *vertices, uint64_t num_vertices, const uint32_t *faces,
               uint64_t num_faces, void *stream = nullptr) {
    srf_tri_desc d{vertices, num_vertices, faces, num_faces};
    return build(d, srf_accel_build_triangles, stream);
  }
  /// Build32(const Mesh*) with the mesh's host arrays; synchronises.
  bool Build32Host(const float *vertices, uint64_t num_vertices, const uint32_t *faces,
                   uint64_t num_faces, void *stream = nullptr) {
    srf_tri_desc d{vertices, num_vertices, faces, num_faces};
    return build_host(SRF_PRIM_TRIANGLES, d, stream);
  }
};

/// TetraAccel: Build32(const Tetrahedron*) with device arrays.
class TetraAccel : public detail::AccelBase {
 public:
  explicit TetraAccel(Context &ctx) : AccelBase(ctx) { kind_ = SRF_PRIM_TETRAHEDRONS; }
  bool Build32(const float *vertices, uint64_t num_vertices, const uint32_t *indices,
               uint64_t num_tetras, void *stream = nullptr) {
    srf_tet_desc d{vertices, num_vertices, indices, num_tetras};
    return build(d, srf_accel_build_tetras, stream);
  }
  /// Build32(const Tetrahedron*) with host arrays; synchronises.
  bool Build32Host(const float *vertices, uint64_t num_vertices, const uint32_t *indices,
                   uint64_t num_tetras, void *stream = nullptr) {
    srf_tet_desc d{vertices, num_vertices, indices, num_tetras};
    return build_host(SRF_PRIM_TETRAHEDRONS, d, stream);
  }
  /// Traverse + BuildIntersection/InterpolateTetra (+ scalar varying).
  bool TraverseTetra(const srf_ray *rays, uint64_t n, const float *vertex_scalars, srf_hit *hits,
                     srf_tet_hit_ext *ext, void *stream = nullptr) {
    return remember(srf_trace_tetra(accel_, rays, n, vertex_scalars, hits, ext, stream));
  }
};

/// SolidAccel: Build32(const Solid*) with device arrays; vertices_per_solid
/// is Solid::numVertsPerSolid (5 pyramids, 6 prisms, 8 hexahedra).
class SolidAccel : public detail::AccelBase {
 public:
  explicit SolidAccel(Context &ctx) : AccelBase(ctx) { kind_ = SRF_PRIM_HEXAHEDRONS; }
  bool Build32(const float *vertices, uint64_t num_vertices, const uint32_t *indices,
               uint64_t num_solids, uint32_t vertices_per_solid, void *stream = nullptr) {
    kind_ = vertices_per_solid == 5 ? SRF_PRIM_PYRAMIDS
                                    : (vertices_per_solid == 6 ? SRF_PRIM_PRISMS
                                                               : SRF_PRIM_HEXAHEDRONS);
    srf_solid_desc d{vertices, num_vertices, indices, num_solids, vertices_per_solid};
    return build(d, srf_accel_build_solids, stream);
  }
  /// Traverse + the crossing's position, normal, interpolate() weights and
  /// the varying sum of index_scalars (one per index position).
  bool TraverseSolid(const srf_ray *rays, uint64_t n, const float *index_scalars, srf_hit *hits,
                     srf_solid_hit_ext *ext, void *stream = nullptr) {
    return remember(srf_trace_solids(accel_, rays, n, index_scalars, hits, ext, stream));
  }
};

/// LineAccel: Build(const Lines*, options) with device arrays; cap is
/// LineBuildOptions::cap. Built on the LBVH (the reference's own tree is a
/// top-down spatial-median one); hits are the closest crossing.
class LineAccel : public detail::AccelBase {
 public:
  explicit LineAccel(Context &ctx) : AccelBase(ctx) { kind_ = SRF_PRIM_LINES; }
  bool Build(const float *positions, uint64_t num_vertices, const uint32_t *segments,
             uint64_t num_lines, const float *radius, float constant_radius, bool cap = true,
             void *stream = nullptr) {
    srf_line_desc d{positions, num_vertices, segments, num_lines, radius, constant_radius,
                    cap ? 1u : 0u};
    return build(d, srf_accel_build_lines, stream);
  }
  /// Traverse + BuildIntersection's position and normal.
  bool TraverseLines(const srf_ray *rays, uint64_t n, srf_hit *hits, srf_line_hit_ext *ext,
                     void *stream = nullptr) {
    return remember(srf_trace_lines(accel_, rays, n, hits, ext, stream));
  }
};

/// ParticleAccel: Build32(const Particles*) with device arrays.
class ParticleAccel : public detail::AccelBase {
 public:
  explicit ParticleAccel(Context &ctx) : AccelBase(ctx) { kind_ = SRF_PRIM_POINTS; }
  bool Build32(const float *positions, const float *radius, float constant_radius,
               uint64_t num_particles, void *stream = nullptr) {
    srf_sphere_desc d{positions, radius, constant_radius, num_particles};
    return build(d, srf_accel_build_spheres, stream);
  }
  /// Build32(const Particles*) with host arrays; synchronises.
  bool Build32Host(const float *positions, const float *radius, float constant_radius,
                   uint64_t num_particles, void *stream = nullptr) {
    srf_sphere_desc d{positions, radius, constant_radius, num_particles};
    return build_host(SRF_PRIM_POINTS, d, stream);
  }
  /// Traverse with per-ray self-intersection inputs (prev_hit_t/normal).
  bool TraverseSelf(const srf_ray *rays, const srf_sphere_prev *prev, uint64_t n, srf_hit *hits,
                    void *stream = nullptr) {
    return remember(srf_trace_spheres(accel_, rays, prev, n, hits, stream));
  }
};

}  // namespace srf

#endif  // SRF_ACCEL_HPP_
