// srf_frame.hpp -- the GLES ray-tracing front end's frame loop and shader-issued
// rays over the srf C ABI (header-only C++, host side; needs the HIP runtime
// for device buffers and streams).
//
//   Camera, BuildFrame   lsgl::render::Camera::buildFrame, perspective
//                        (render/render_camera.cc:17-63)
//   EyeRay, EyeRays      the eye ray of one pixel as RaytraceEngine::OnData
//                        forms it for one sample per pixel
//                        (gles/gles_raytrace_engine.cc:730-812)
//   FrameTracer::Render  the OnData loop (gles/gles_raytrace_engine.cc:547-954):
//                        the reference generates, traces and shades pixel by
//                        pixel in OpenMP rows; here a frame's eye rays are
//                        generated on the host, traced as one batch on the
//                        GPU, then shaded on the host (the fragment shader
//                        stays on the CPU), with the reference's Y-down colour
//                        and depth writes and its Mraycasts figure
//   AccelCache           gles_accel_builder's per-buffer accelerator cache
//                        (gles/gles_accel_builder.cc:34-117, 442-455): a mesh
//                        drawn again with the same buffers reuses its build
//   RayQueue             shader-issued rays (trace()/shadow(),
//                        gles/gles_glsl_runtime.cc:88-217): shading threads
//                        queue rays in their own queues, a queue goes to the
//                        GPU as one batch when it holds `batch` rays or on
//                        Flush(), and each ray's result is read back by ticket
//                        -- the reference traces each such ray synchronously.
//
// Ray arithmetic follows the reference's real3 operators (float, with
// normalize() scaling by the double 1/sqrt of the float squared length), so
// the rays are bit-identical to the reference's.
#ifndef SRF_FRAME_HPP_
#define SRF_FRAME_HPP_

#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdint.h>

#include <chrono>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "srf.h"
#include "srf_accel.hpp"

namespace srf {

struct Camera {
  float eye[3], lookat[3], up[3];
  float fov;  // vertical field of view, degrees
};

namespace frame_detail {

struct R3 {
  float v[3];
};
inline R3 r3(float x, float y, float z) { return R3{{x, y, z}}; }
inline R3 add(R3 a, R3 b) { return r3(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]); }
inline R3 sub(R3 a, R3 b) { return r3(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
inline R3 mul(float f, R3 a) { return r3(a.v[0] * f, a.v[1] * f, a.v[2] * f); }
inline R3 cross(R3 l, R3 r) {
  return r3(l.v[1] * r.v[2] - l.v[2] * r.v[1], l.v[2] * r.v[0] - l.v[0] * r.v[2],
            l.v[0] * r.v[1] - l.v[1] * r.v[0]);
}
// real3::normalize (render/render_common.h:69-83)
inline R3 normalize(R3 a) {
  const double len2 = a.v[0] * a.v[0] + a.v[1] * a.v[1] + a.v[2] * a.v[2];
  double len = 0.0;
  if (fabs(len2) > 1.0e-30) len = 1.0 / sqrt(len2);
  for (int k = 0; k < 3; k++) a.v[k] = (float)(a.v[k] * len);
  return a;
}

}  // namespace frame_detail

// Camera::buildFrame, perspective branch.
inline void BuildFrame(const Camera &c, int width, int height, float corner[3], float u[3],
                       float v[3]) {
  using namespace frame_detail;
  const float flen = (float)(0.5f * (float)height / tan(0.5 * (c.fov * M_PI / 180.0)));
  const R3 eye = r3(c.eye[0], c.eye[1], c.eye[2]);
  R3 look = sub(r3(c.lookat[0], c.lookat[1], c.lookat[2]), eye);
  R3 uu = normalize(cross(look, r3(c.up[0], c.up[1], c.up[2])));
  R3 vv = normalize(cross(look, uu));
  look = normalize(look);
  look = add(mul(flen, look), eye);
  const R3 half = mul(0.5f, add(mul((float)width, uu), mul((float)height, vv)));
  const R3 cr = sub(look, half);
  for (int k = 0; k < 3; k++) {
    corner[k] = cr.v[k];
    u[k] = uu.v[k];
    v[k] = vv.v[k];
  }
}

// The eye ray of framebuffer pixel (x, y) (one sample, the whole framebuffer
// as viewport). Returns false when the viewport conversion leaves the
// framebuffer (the reference skips that pixel).
inline bool EyeRay(const Camera &c, const float corner[3], const float u[3], const float v[3],
                   int width, int height, int x, int y, srf_ray &ray) {
  using namespace frame_detail;
  const float x01 = (x + 0.5f) * (1.0f / width), y01 = (y + 0.5f) * (1.0f / height);
  // ViewportConversion (gles/gles_raytrace_engine.cc:530-544), viewport (0, 0, w, h)
  const float xnd = 2.0f * (x01 - 0.5f), ynd = 2.0f * (y01 - 0.5f);
  const int px = (int)((xnd + 1.0f) * (0.5f * width) - 0);
  const int py = (int)((ynd + 1.0f) * (0.5f * height) + 0);
  if (px < 0 || py < 0 || px >= width || py >= height) return false;
  const float step = 1.0f, su = 0.0f, sv = 0.0f;
  const R3 cr = r3(corner[0], corner[1], corner[2]);
  const R3 ud = r3(u[0], u[1], u[2]), vd = r3(v[0], v[1], v[2]);
  const R3 org = r3(c.eye[0], c.eye[1], c.eye[2]);
  R3 dir = sub(add(add(cr, mul((float)px + su + step / 2.0f, ud)),
                   mul((float)py + sv - step / 2.0f + 1.0f, vd)),
               org);
  dir = normalize(dir);
  for (int k = 0; k < 3; k++) {
    ray.org[k] = org.v[k];
    ray.dir[k] = dir.v[k];
  }
  ray.flags = SRF_RAY_DOUBLE_SIDED;  // eye rays always test double-sided
  ray.prev_prim_id = SRF_NO_PRIM;
  return true;
}

// All eye rays of a frame in the reference's loop order (rows y = 0, 1, ...,
// pixels x within a row); `valid[i]` is 0 for a skipped pixel.
inline void EyeRays(const Camera &c, int width, int height, srf_ray *rays, uint8_t *valid) {
  float corner[3], u[3], v[3];
  BuildFrame(c, width, height, corner, u, v);
  for (int y = 0; y < height; y++)
    for (int x = 0; x < width; x++) {
      srf_ray &r = rays[(size_t)y * width + x];
      const bool ok = EyeRay(c, corner, u, v, width, height, x, y, r);
      if (!ok) r = srf_ray{};
      if (valid) valid[(size_t)y * width + x] = ok ? 1 : 0;
    }
}

namespace frame_detail {

// A device buffer that grows; freed with its owner.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  bool ensure(size_t n) {
    if (n <= bytes) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, n) != hipSuccess) return false;
    bytes = n;
    return true;
  }
};

}  // namespace frame_detail

// One frame: generate, trace as a batch, shade on the host.
class FrameTracer {
 public:
  // Shades the hit of pixel (x, y); returns false when the fragment is
  // discarded (nothing is written for that pixel, as in the reference).
  using Shader =
      std::function<bool(int x, int y, const srf_ray &ray, const srf_hit &hit, float rgba[4])>;

  // Renders into rgba (width * height * 4 floats) and depth (width * height,
  // may be null), rows written Y-down as the reference's framebuffer writes
  // do; pixels whose ray misses (or is discarded) are left untouched.
  bool Render(const srf_accel *accel, const Camera &cam, int width, int height,
              const Shader &shade, float *rgba, float *depth, void *stream = nullptr) {
    return render(
        cam, width, height,
        [&](const srf_ray *r, uint64_t n, srf_hit *h) {
          return srf_trace_closest(accel, r, n, h, stream) == SRF_OK;
        },
        [&](int x, int y, size_t i, float c[4]) { return shade(x, y, rays_[i], hits_[i], c); },
        rgba, depth, stream);
  }

  // The same frame over a two-level scene (RenderGraph::Trace per eye ray,
  // as OnData does for a frame of several draws): hits are world-space and
  // the shader also receives the element hit.
  using SceneShader = std::function<bool(int x, int y, const srf_ray &ray, const srf_hit &hit,
                                         uint32_t element, float rgba[4])>;
  bool RenderScene(const srf_scene *scene, const Camera &cam, int width, int height,
                   const SceneShader &shade, float *rgba, float *depth,
                   void *stream = nullptr) {
    const size_t n = (size_t)width * height;
    if (!delem_.ensure(n * sizeof(uint32_t))) return false;
    elems_.resize(n);
    return render(
        cam, width, height,
        [&](const srf_ray *r, uint64_t cnt, srf_hit *h) {
          hipStream_t s = reinterpret_cast<hipStream_t>(stream);
          return srf_scene_trace(scene, r, cnt, h, (uint32_t *)delem_.p, SRF_LAUNCH_TUNED,
                                 stream) == SRF_OK &&
                 hipMemcpyAsync(elems_.data(), delem_.p, cnt * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, s) == hipSuccess;
        },
        [&](int x, int y, size_t i, float c[4]) {
          return shade(x, y, rays_[i], hits_[i], elems_[i], c);
        },
        rgba, depth, stream);
  }

  double RenderMs() const { return ms_; }
  // million eye rays per second of the last frame, the figure
  // RaytraceEngine::OnData prints
  double Mraycasts() const { return ms_ > 0.0 ? (double)rays_traced_ / (ms_ * 1e3) : 0.0; }
  const std::vector<srf_hit> &Hits() const { return hits_; }

 private:
  // Generate, trace (trace(rays, n, hits) on device arrays, enqueued on
  // stream), read back, shade (shade_at(x, y, pixel, rgba)).
  template <class TraceFn, class ShadeFn>
  bool render(const Camera &cam, int width, int height, const TraceFn &trace,
              const ShadeFn &shade_at, float *rgba, float *depth, void *stream) {
    const auto t0 = std::chrono::steady_clock::now();
    const size_t n = (size_t)width * height;
    rays_.resize(n);
    valid_.resize(n);
    hits_.resize(n);
    EyeRays(cam, width, height, rays_.data(), valid_.data());
    if (!dray_.ensure(n * sizeof(srf_ray)) || !dhit_.ensure(n * sizeof(srf_hit))) return false;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (hipMemcpyAsync(dray_.p, rays_.data(), n * sizeof(srf_ray), hipMemcpyHostToDevice, s) !=
        hipSuccess)
      return false;
    if (!trace((const srf_ray *)dray_.p, n, (srf_hit *)dhit_.p)) return false;
    if (hipMemcpyAsync(hits_.data(), dhit_.p, n * sizeof(srf_hit), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return false;
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < height; y++) {
      for (int x = 0; x < width; x++) {
        const size_t i = (size_t)y * width + x;
        if (!valid_[i] || hits_[i].prim_id == SRF_NO_PRIM) continue;
        float col[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (!shade_at(x, y, i, col)) continue;
        const size_t o = (size_t)(height - y - 1) * width + x;
        for (int k = 0; k < 4; k++) rgba[4 * o + k] = col[k];
        if (depth) depth[o] = hits_[i].t;
      }
    }
    const auto t1 = std::chrono::steady_clock::now();
    ms_ = std::chrono::duration<double, std::milli>(t1 - t0).count();
    rays_traced_ = n;
    return true;
  }

  std::vector<srf_ray> rays_;
  std::vector<uint8_t> valid_;
  std::vector<srf_hit> hits_;
  std::vector<uint32_t> elems_;
  frame_detail::DevBuf dray_, dhit_, delem_;
  double ms_ = 0.0;
  size_t rays_traced_ = 0;
};

// Triangle accelerators built once per set of mesh buffers.
class AccelCache {
 public:
  explicit AccelCache(Context &ctx) : ctx_(ctx) {}

  // The accelerator for these host arrays, built on the first call (the
  // arrays are copied to the device; the accelerator keeps its own copies).
  // Returns null on failure. Invalidate() after changing the arrays' contents.
  TriangleAccel *Get(const float *vertices, uint64_t nv, const uint32_t *faces, uint64_t nf) {
    const Key k{vertices, nv, faces, nf};
    auto it = map_.find(k);
    if (it != map_.end()) return it->second.get();
    frame_detail::DevBuf dv, df;
    if (!dv.ensure(12 * nv) || !df.ensure(12 * nf)) return nullptr;
    if (hipMemcpy(dv.p, vertices, 12 * nv, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(df.p, faces, 12 * nf, hipMemcpyHostToDevice) != hipSuccess)
      return nullptr;
    std::unique_ptr<TriangleAccel> a(new TriangleAccel(ctx_));
    if (!a->Build32((const float *)dv.p, nv, (const uint32_t *)df.p, nf)) return nullptr;
    double bmin[3], bmax[3];
    a->BoundingBox(bmin, bmax);  // waits for the build before the copies go
    TriangleAccel *out = a.get();
    map_[k] = std::move(a);
    return out;
  }
  void Invalidate(const float *vertices, uint64_t nv, const uint32_t *faces, uint64_t nf) {
    map_.erase(Key{vertices, nv, faces, nf});
  }
  size_t Size() const { return map_.size(); }

 private:
  using Key = std::tuple<const float *, uint64_t, const uint32_t *, uint64_t>;
  Context &ctx_;
  std::map<Key, std::unique_ptr<TriangleAccel>> map_;
};

// Shader-issued rays, queued per shading thread and traced in batches.
class RayQueue {
 public:
  // kind: kClosest (trace(): t, u, v, prim) or kAny (occlusion against
  // t_max).
  enum Kind { kClosest = 0, kAny = 1 };

  RayQueue(const srf_accel *accel, int threads, size_t batch, Kind kind = kClosest)
      : accel_(accel), batch_(batch ? batch : 1), kind_(kind), q_(threads) {}

  // Queues a ray for thread `tid` (0 <= tid < threads; each thread uses its
  // own id, so no locking); returns its ticket. Traces the thread's pending
  // rays as one batch when `batch` of them are waiting.
  uint64_t Push(int tid, const srf_ray &ray, float t_max = 3.402823466e38f) {
    Queue &q = q_[tid];
    const uint64_t ticket = ((uint64_t)tid << 40) | q.rays.size();
    q.rays.push_back(ray);
    q.tmax.push_back(t_max);
    if (q.rays.size() - q.done >= batch_) q.ok = submit(q) && q.ok;
    return ticket;
  }

  // Traces every thread's pending rays (call from one thread, after the
  // shading threads have pushed). Returns false if any batch failed.
  bool Flush() {
    bool ok = true;
    for (Queue &q : q_) {
      if (q.rays.size() > q.done) ok = submit(q) && ok;
      ok = q.ok && ok;
    }
    return ok;
  }

  // Results by ticket (after the batch that holds the ray was traced).
  const srf_hit &Hit(uint64_t ticket) const { return q_[ticket >> 40].hits[ticket & kIdx]; }
  bool Occluded(uint64_t ticket) const { return q_[ticket >> 40].occ[ticket & kIdx] != 0; }
  // shadow() of gles_glsl_runtime.cc: the hit distance, or -FLT_MAX on a miss
  float ShadowT(uint64_t ticket) const {
    const srf_hit &h = Hit(ticket);
    return h.prim_id == SRF_NO_PRIM ? -3.402823466e38f : h.t;
  }
  size_t Batches() const {
    size_t b = 0;
    for (const Queue &q : q_) b += q.batches;
    return b;
  }

 private:
  static constexpr uint64_t kIdx = (1ull << 40) - 1;
  struct Queue {
    std::vector<srf_ray> rays;
    std::vector<float> tmax;
    std::vector<srf_hit> hits;
    std::vector<uint32_t> occ;
    size_t done = 0, batches = 0;
    bool ok = true;
    frame_detail::DevBuf dray, dtmax, dout;
    hipStream_t stream = nullptr;
    ~Queue() {
      if (stream) (void)hipStreamDestroy(stream);
    }
  };

  bool submit(Queue &q) {
    const size_t n = q.rays.size() - q.done;
    if (n == 0) return true;
    if (!q.stream && hipStreamCreateWithFlags(&q.stream, hipStreamNonBlocking) != hipSuccess)
      return false;
    const size_t out_bytes = kind_ == kClosest ? sizeof(srf_hit) : 4;
    if (!q.dray.ensure(n * sizeof(srf_ray)) || !q.dtmax.ensure(n * 4) ||
        !q.dout.ensure(n * out_bytes))
      return false;
    if (hipMemcpyAsync(q.dray.p, q.rays.data() + q.done, n * sizeof(srf_ray),
                       hipMemcpyHostToDevice, q.stream) != hipSuccess)
      return false;
    srf_status st;
    if (kind_ == kClosest) {
      st = srf_trace_closest(accel_, (const srf_ray *)q.dray.p, n, (srf_hit *)q.dout.p, q.stream);
      q.hits.resize(q.rays.size());
      if (st == SRF_OK &&
          hipMemcpyAsync(q.hits.data() + q.done, q.dout.p, n * sizeof(srf_hit),
                         hipMemcpyDeviceToHost, q.stream) != hipSuccess)
        return false;
    } else {
      if (hipMemcpyAsync(q.dtmax.p, q.tmax.data() + q.done, n * 4, hipMemcpyHostToDevice,
                         q.stream) != hipSuccess)
        return false;
      st = srf_trace_any(accel_, (const srf_ray *)q.dray.p, (const float *)q.dtmax.p, n,
                         (uint32_t *)q.dout.p, q.stream);
      q.occ.resize(q.rays.size());
      if (st == SRF_OK && hipMemcpyAsync(q.occ.data() + q.done, q.dout.p, n * 4,
                                         hipMemcpyDeviceToHost, q.stream) != hipSuccess)
        return false;
    }
    if (st != SRF_OK || hipStreamSynchronize(q.stream) != hipSuccess) return false;
    q.done = q.rays.size();
    q.batches++;
    return true;
  }

  const srf_accel *accel_;
  size_t batch_;
  Kind kind_;
  std::vector<Queue> q_;
};

}  // namespace srf

#endif  // SRF_FRAME_HPP_
