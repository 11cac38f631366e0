// srf_gl.hpp -- the GLES draw path of the ray-tracing front end over the srf C
// ABI (header-only C++, host side; needs the HIP runtime).
//
//   DrawContext::DrawElements  Context::glDrawElements
//                              (gles/gles_primitive.cc:84-368): the reference's
//                              checks and GL errors, the draw's accelerator from
//                              the draw cache (AccelBuilder::Build*Accel,
//                              gles/gles_accel_builder.cc:34-390: one build per
//                              element buffer, position buffer, count and offset
//                              while the buffers are unchanged), and one
//                              RenderElement with the lsgl_World matrix and the
//                              matrices RenderElement::Update derives from it
//   DrawContext::Finish        Context::glFinish (gles/gles_special.cc:22-80):
//                              the frame's RenderGraph (srf_scene over every
//                              element drawn since the last Finish, of any
//                              primitive kind), its eye rays traced as one batch,
//                              each hit shaded on the host (srf::FrameTracer),
//                              then the frame's elements cleared (the
//                              accelerators stay cached, as EndFrame keeps them)
//
// Geometry follows AccelBuilder: triangles, tetrahedra, solids and lines take
// their vertex ids from the element buffer (count / 3, / 4, / nv and / 2
// primitives); points are the first `count` vertices of the position array
// (AddParticleData uses the index count only), with radius lsgl_PointSize / 2
// or, when a per-vertex size array holds count + offset values, those values
// from `offset` on; lines likewise with glLineWidth / 2 and the lsgl_LineWidth
// array (indexed by vertex id). Positions are XYZ floats.
#ifndef SRF_GL_HPP_
#define SRF_GL_HPP_

#include <stdint.h>

#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "srf.h"
#include "srf_frame.hpp"

namespace srf {
namespace gl {

// GLenum values of GLES2 and the reference's extensions (include/GLES2/gl2.h)
constexpr uint32_t kPoints = 0x0000;
constexpr uint32_t kLines = 0x0001;
constexpr uint32_t kTriangles = 0x0004;
constexpr uint32_t kTetrahedronsExt = 0x0010;
constexpr uint32_t kPyramidsExt = 0x0011;
constexpr uint32_t kPrismsExt = 0x0012;
constexpr uint32_t kHexahedronsExt = 0x0013;
constexpr uint32_t kUnsignedInt = 0x1405;
constexpr uint32_t kNoError = 0;
constexpr uint32_t kInvalidEnum = 0x0500;
constexpr uint32_t kInvalidValue = 0x0501;
constexpr uint32_t kInvalidOperation = 0x0502;

// What glDrawElements reads from the bound state.
struct DrawBuffers {
  const uint32_t *elements = nullptr;  // GL_ELEMENT_ARRAY_BUFFER contents
  uint64_t num_elements = 0;
  const float *positions = nullptr;    // position attribute: XYZ floats
  uint64_t num_vertices = 0;
  float world[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};  // lsgl_World
  float point_size = 1.0f;              // lsgl_PointSize uniform
  const float *point_sizes = nullptr;   // lsgl_PointSize vertex attribute
  uint64_t num_point_sizes = 0;
  float line_width = 1.0f;              // glLineWidth
  const float *line_widths = nullptr;   // lsgl_LineWidth vertex attribute
  uint64_t num_line_widths = 0;
  bool line_cap = false;                // lsgl_LineCap uniform > 0
};

class DrawContext {
 public:
  // Shades pixel (x, y) hit by `ray`: hit is the world-space closest hit
  // (t = world distance), element the index of its draw in this frame.
  using Shader = std::function<bool(int x, int y, const srf_ray &ray, const srf_hit &hit,
                                    uint32_t element, float rgba[4])>;

  explicit DrawContext(Context &ctx) : ctx_(ctx) {}
  ~DrawContext() {
    for (auto &kv : cache_)
      if (kv.second) srf_accel_destroy(kv.second);
  }
  DrawContext(const DrawContext &) = delete;
  DrawContext &operator=(const DrawContext &) = delete;

  // glDrawElements(mode, count, type, (void *)offset) with the state in b.
  void DrawElements(uint32_t mode, int32_t count, uint32_t type, uint64_t offset,
                    const DrawBuffers &b) {
    if (count < 0) return fail(kInvalidValue);
    if (count == 0) return;
    if (nv_of(mode) == 0) return fail(kInvalidEnum);
    if (type != kUnsignedInt) return fail(kInvalidEnum);
    if (!b.elements || !b.positions) return fail(kInvalidOperation);
    // the reference reads past the buffers here; this front end refuses
    const uint64_t first = offset / 4;
    if (offset % 4 || first + (uint64_t)count > b.num_elements ||
        (mode == kPoints && (uint64_t)count > b.num_vertices))
      return fail(kInvalidValue);
    srf_accel *a = accel_for(mode, count, offset, b);
    if (!a) return fail(kInvalidOperation);
    srf_instance inst{};
    inst.accel = a;
    for (int k = 0; k < 16; k++) inst.T[k] = (double)b.world[k];  // float -> double upcast
    if (srf_instance_update(&inst) != SRF_OK) return fail(kInvalidOperation);
    frame_.push_back(inst);
  }

  // glFinish: renders the frame (rows Y-down into rgba and depth, as
  // FrameTracer::Render) and clears it. Returns false when the trace fails;
  // a frame with no draws renders nothing.
  bool Finish(const Camera &cam, int width, int height, const Shader &shade, float *rgba,
              float *depth, void *stream = nullptr) {
    if (frame_.empty()) return true;
    srf_scene *scene = nullptr;
    bool ok = srf_scene_create(ctx_.get(), (uint32_t)frame_.size(), frame_.data(), &scene) ==
              SRF_OK;
    if (ok) {
      ok = tracer_.RenderScene(scene, cam, width, height,
                               [&](int x, int y, const srf_ray &r, const srf_hit &h,
                                   uint32_t e, float c[4]) { return shade(x, y, r, h, e, c); },
                               rgba, depth, stream);
      srf_scene_destroy(scene);
    }
    frame_.clear();  // RenderGraph::Clear, currentDrawStackIndex = 0
    if (!ok) fail(kInvalidOperation);
    return ok;
  }

  // glGetError: the first error since the last call, then kNoError.
  uint32_t GetError() {
    const uint32_t e = error_;
    error_ = kNoError;
    return e;
  }

  // Drops the cached accelerators built from these buffers (their contents
  // changed: glBufferData / glBufferSubData).
  void Invalidate(const uint32_t *elements, const float *positions) {
    for (auto it = cache_.begin(); it != cache_.end();) {
      if (std::get<1>(it->first) == elements || std::get<2>(it->first) == positions) {
        if (it->second) srf_accel_destroy(it->second);
        it = cache_.erase(it);
      } else {
        ++it;
      }
    }
  }

  size_t FrameElements() const { return frame_.size(); }
  size_t CachedAccels() const { return cache_.size(); }
  const FrameTracer &Tracer() const { return tracer_; }

 private:
  // vertices per primitive of a drawing mode (0: not a mode glDrawElements
  // takes)
  static int nv_of(uint32_t mode) {
    switch (mode) {
      case kPoints: return 1;
      case kLines: return 2;
      case kTriangles: return 3;
      case kTetrahedronsExt: return 4;
      case kPyramidsExt: return 5;
      case kPrismsExt: return 6;
      case kHexahedronsExt: return 8;
      default: return 0;
    }
  }

  void fail(uint32_t e) {
    if (error_ == kNoError) error_ = e;  // glGetError keeps the first one
  }

  using Key = std::tuple<uint32_t, const uint32_t *, const float *, int32_t, uint64_t, uint64_t,
                         float, const float *, bool>;

  srf_accel *accel_for(uint32_t mode, int32_t count, uint64_t offset, const DrawBuffers &b) {
    const bool pts = mode == kPoints, lin = mode == kLines;
    const float size = pts ? b.point_size : (lin ? b.line_width : 0.0f);
    const float *sizes = pts ? b.point_sizes : (lin ? b.line_widths : nullptr);
    const uint64_t nsizes = pts ? b.num_point_sizes : (lin ? b.num_line_widths : 0);
    const bool use_sizes = sizes && nsizes >= (uint64_t)count + offset;
    const Key key{mode, b.elements, b.positions, count, offset, b.num_vertices, size,
                  use_sizes ? sizes : nullptr, lin && b.line_cap};
    auto it = cache_.find(key);
    if (it != cache_.end()) return it->second;
    const uint32_t *idx = b.elements + offset / 4;
    const int nv = nv_of(mode);
    srf_accel *a = nullptr;
    srf_status st = SRF_ERR_INVALID_ARG;
    if (mode == kTriangles) {
      srf_tri_desc d{b.positions, b.num_vertices, idx, (uint64_t)count / 3};
      st = srf_accel_build_host(ctx_.get(), SRF_PRIM_TRIANGLES, &d, nullptr, &a);
    } else if (mode == kTetrahedronsExt) {
      srf_tet_desc d{b.positions, b.num_vertices, idx, (uint64_t)count / 4};
      st = srf_accel_build_host(ctx_.get(), SRF_PRIM_TETRAHEDRONS, &d, nullptr, &a);
    } else if (mode == kPoints) {
      srf_sphere_desc d{b.positions, use_sizes ? sizes + offset : nullptr, size * 0.5f,
                        (uint64_t)count};
      st = srf_accel_build_host(ctx_.get(), SRF_PRIM_POINTS, &d, nullptr, &a);
    } else if (mode == kLines) {
      srf_line_desc d{b.positions, b.num_vertices, idx, (uint64_t)count / 2,
                      use_sizes ? sizes + offset : nullptr, size * 0.5f,
                      b.line_cap ? 1u : 0u};
      st = srf_accel_build_host(ctx_.get(), SRF_PRIM_LINES, &d, nullptr, &a);
    } else {
      srf_solid_desc d{b.positions, b.num_vertices, idx, (uint64_t)count / (uint64_t)nv,
                       (uint32_t)nv};
      const srf_prim_kind kind = nv == 5 ? SRF_PRIM_PYRAMIDS
                                         : (nv == 6 ? SRF_PRIM_PRISMS : SRF_PRIM_HEXAHEDRONS);
      st = srf_accel_build_host(ctx_.get(), kind, &d, nullptr, &a);
    }
    if (st != SRF_OK) return nullptr;
    cache_[key] = a;
    return a;
  }

  Context &ctx_;
  std::map<Key, srf_accel *> cache_;
  std::vector<srf_instance> frame_;
  FrameTracer tracer_;
  uint32_t error_ = kNoError;
};

}  // namespace gl
}  // namespace srf

#endif  // SRF_GL_HPP_
