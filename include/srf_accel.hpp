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
// reason in LastError(). Traverse takes a batch of device-resident rays and
// writes device-resident hits; the per-ray CPU signature of the reference
// (Traverse(Intersection&, Ray&)) has no GPU counterpart -- callers batch
// their rays (generate, trace the batch, then shade each hit on the host).
#ifndef SRF_ACCEL_HPP_
#define SRF_ACCEL_HPP_

#include <string>

#include "srf.h"

namespace srf {

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

  Context &ctx_;
  srf_accel *accel_ = nullptr;
  std::string error_;
};

}  // namespace detail

/// TriangleAccel: Build32(const Mesh*) takes device arrays here.
class TriangleAccel : public detail::AccelBase {
 public:
  using AccelBase::AccelBase;
  bool Build32(const float *vertices, uint64_t num_vertices, const uint32_t *faces,
               uint64_t num_faces, void *stream = nullptr) {
    srf_tri_desc d{vertices, num_vertices, faces, num_faces};
    return build(d, srf_accel_build_triangles, stream);
  }
};

/// TetraAccel: Build32(const Tetrahedron*) with device arrays.
class TetraAccel : public detail::AccelBase {
 public:
  using AccelBase::AccelBase;
  bool Build32(const float *vertices, uint64_t num_vertices, const uint32_t *indices,
               uint64_t num_tetras, void *stream = nullptr) {
    srf_tet_desc d{vertices, num_vertices, indices, num_tetras};
    return build(d, srf_accel_build_tetras, stream);
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
  using AccelBase::AccelBase;
  bool Build32(const float *vertices, uint64_t num_vertices, const uint32_t *indices,
               uint64_t num_solids, uint32_t vertices_per_solid, void *stream = nullptr) {
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
  using AccelBase::AccelBase;
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
  using AccelBase::AccelBase;
  bool Build32(const float *positions, const float *radius, float constant_radius,
               uint64_t num_particles, void *stream = nullptr) {
    srf_sphere_desc d{positions, radius, constant_radius, num_particles};
    return build(d, srf_accel_build_spheres, stream);
  }
  /// Traverse with per-ray self-intersection inputs (prev_hit_t/normal).
  bool TraverseSelf(const srf_ray *rays, const srf_sphere_prev *prev, uint64_t n, srf_hit *hits,
                    void *stream = nullptr) {
    return remember(srf_trace_spheres(accel_, rays, prev, n, hits, stream));
  }
};

}  // namespace srf

#endif  // SRF_ACCEL_HPP_
