// C++ host test of the frame loop and shader-ray queues (include/srf_frame.hpp)
// on the GPU: a grid heightfield drawn through srf::AccelCache, one frame
// rendered by srf::FrameTracer with a depth shader, and shadow rays from the
// frame's hit points queued by 8 threads in srf::RayQueue. Checks: the cache
// builds once; every pixel's depth is the oracle's hit distance for the same
// eye ray (oracle/liboracle.so, the restatement of TriangleAccel::Traverse),
// written Y-down, and missed pixels stay untouched; queued closest and any-hit
// results equal the oracle's trace of the same shadow rays (and one batch
// trace of them). Prints "PASS".
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/srf_frame.hpp"
#include "../../oracle/srf_oracle.h"

#define CHECK(x)                                                   \
  do {                                                             \
    if (!(x)) {                                                    \
      std::fprintf(stderr, "FAILED: %s (line %d)\n", #x, __LINE__); \
      return 1;                                                    \
    }                                                              \
  } while (0)

int main() {
  const int G = 128;
  std::vector<float> V;
  for (int j = 0; j < G; j++)
    for (int i = 0; i < G; i++) {
      float x = i / float(G - 1), y = j / float(G - 1);
      V.push_back(x);
      V.push_back(y);
      V.push_back(0.06f * std::sin(11.0f * x) * std::cos(8.0f * y));
    }
  std::vector<uint32_t> F;
  for (int j = 0; j + 1 < G; j++)
    for (int i = 0; i + 1 < G; i++) {
      uint32_t a = j * G + i, b = a + 1, c = a + G, d = c + 1;
      F.insert(F.end(), {a, b, d, a, d, c});
    }
  const uint64_t nv = V.size() / 3, nf = F.size() / 3;

  srf::Context ctx(0);
  CHECK(ctx.ok());
  srf::AccelCache cache(ctx);
  srf::TriangleAccel *acc = cache.Get(V.data(), nv, F.data(), nf);
  CHECK(acc != nullptr);
  CHECK(cache.Get(V.data(), nv, F.data(), nf) == acc);  // drawn again: no rebuild
  CHECK(cache.Size() == 1);

  const int W = 160, H = 120;
  srf::Camera cam = {{0.5f, -0.6f, 0.7f}, {0.5f, 0.5f, 0.0f}, {0.0f, 0.0f, 1.0f}, 45.0f};
  std::vector<float> rgba(4 * W * H, -1.0f), depth(W * H, -1.0f);
  srf::FrameTracer ft;
  auto shade = [](int, int, const srf_ray &, const srf_hit &h, float c[4]) {
    c[0] = c[1] = c[2] = h.t;
    c[3] = 1.0f;
    return true;
  };
  CHECK(ft.Render(acc->handle(), cam, W, H, shade, rgba.data(), depth.data()));

  // the same eye rays through the oracle
  std::vector<srf_ray> rays(W * H);
  std::vector<uint8_t> valid(W * H);
  srf::EyeRays(cam, W, H, rays.data(), valid.data());
  std::vector<or_tri_node> nodes(2 * nf);
  std::vector<uint32_t> idx(nf);
  double bounds[6], eps;
  or_build_tri(V.data(), F.data(), nf, nodes.data(), idx.data(), bounds, &eps);
  std::vector<or_hit> want(rays.size());
  or_trace_tri(nodes.data(), idx.data(), V.data(), F.data(), eps,
               reinterpret_cast<const or_ray *>(rays.data()), rays.size(), want.data(), nullptr,
               1);
  int hits = 0;
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const size_t i = (size_t)y * W + x, o = (size_t)(H - y - 1) * W + x;
      CHECK(std::memcmp(&ft.Hits()[i], &want[i], sizeof(srf_hit)) == 0);
      if (valid[i] && want[i].prim_id != SRF_NO_PRIM) {
        CHECK(depth[o] == want[i].t && rgba[4 * o] == want[i].t && rgba[4 * o + 3] == 1.0f);
        hits++;
      } else {
        CHECK(depth[o] == -1.0f && rgba[4 * o] == -1.0f);
      }
    }
  CHECK(hits > W * H / 4);

  // shadow rays from every hit point toward a light, queued by 8 threads
  std::vector<srf_ray> sh;
  std::vector<float> tmax;
  // a low light, so that about half of the shadow rays are occluded
  const float light[3] = {2.5f, 0.5f, 0.2f};
  for (size_t i = 0; i < rays.size(); i++) {
    if (want[i].prim_id == SRF_NO_PRIM) continue;
    srf_ray r = rays[i];
    float d[3], len2 = 0.0f;
    for (int k = 0; k < 3; k++) {
      r.org[k] = rays[i].org[k] + rays[i].dir[k] * want[i].t;
      d[k] = light[k] - r.org[k];
      len2 += d[k] * d[k];
    }
    const float len = std::sqrt(len2);
    for (int k = 0; k < 3; k++) r.dir[k] = d[k] / len;
    r.prev_prim_id = want[i].prim_id;
    sh.push_back(r);
    tmax.push_back(len);
  }
  const int T = 8;
  srf::RayQueue qc(acc->handle(), T, 256, srf::RayQueue::kClosest);
  srf::RayQueue qa(acc->handle(), T, 256, srf::RayQueue::kAny);
  std::vector<uint64_t> tc(sh.size()), ta(sh.size());
  {
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
      th.emplace_back([&, t] {
        for (size_t i = t; i < sh.size(); i += T) {
          tc[i] = qc.Push(t, sh[i]);
          ta[i] = qa.Push(t, sh[i], tmax[i]);
        }
      });
    for (auto &x : th) x.join();
  }
  CHECK(qc.Flush());
  CHECK(qa.Flush());
  CHECK(qc.Batches() >= sh.size() / 256);

  srf_ray *dR;
  srf_hit *dH;
  float *dT;
  uint32_t *dO;
  CHECK(hipMalloc(&dR, sh.size() * sizeof(srf_ray)) == hipSuccess);
  CHECK(hipMalloc(&dH, sh.size() * sizeof(srf_hit)) == hipSuccess);
  CHECK(hipMalloc(&dT, sh.size() * 4) == hipSuccess);
  CHECK(hipMalloc(&dO, sh.size() * 4) == hipSuccess);
  CHECK(hipMemcpy(dR, sh.data(), sh.size() * sizeof(srf_ray), hipMemcpyHostToDevice) ==
        hipSuccess);
  CHECK(hipMemcpy(dT, tmax.data(), sh.size() * 4, hipMemcpyHostToDevice) == hipSuccess);
  CHECK(acc->Traverse(dR, sh.size(), dH));
  CHECK(acc->TraverseAny(dR, dT, sh.size(), dO));
  std::vector<srf_hit> bh(sh.size());
  std::vector<uint32_t> bo(sh.size());
  CHECK(hipMemcpy(bh.data(), dH, sh.size() * sizeof(srf_hit), hipMemcpyDeviceToHost) ==
        hipSuccess);
  CHECK(hipMemcpy(bo.data(), dO, sh.size() * 4, hipMemcpyDeviceToHost) == hipSuccess);
  // the queued results against the oracle's TriangleAccel::Traverse of the
  // same shadow rays (shadow() = closest-hit t, gles/gles_glsl_runtime.cc:
  // 179-217; occluded iff that t < tmax), and against the batch trace
  std::vector<or_hit> osh(sh.size());
  or_trace_tri(nodes.data(), idx.data(), V.data(), F.data(), eps,
               reinterpret_cast<const or_ray *>(sh.data()), sh.size(), osh.data(), nullptr, 8);
  int occluded = 0;
  for (size_t i = 0; i < sh.size(); i++) {
    CHECK(std::memcmp(&qc.Hit(tc[i]), &osh[i], sizeof(srf_hit)) == 0);
    CHECK(qc.ShadowT(tc[i]) == (osh[i].prim_id == SRF_NO_PRIM ? -3.402823466e38f : osh[i].t));
    const bool occ = osh[i].prim_id != SRF_NO_PRIM && osh[i].t < tmax[i];
    CHECK(qa.Occluded(ta[i]) == occ);
    CHECK(std::memcmp(&bh[i], &osh[i], sizeof(srf_hit)) == 0);
    CHECK((bo[i] != 0) == occ);
    occluded += occ;
  }
  CHECK(occluded > 0 && occluded < (int)sh.size());
  hipFree(dR);
  hipFree(dH);
  hipFree(dT);
  hipFree(dO);
  std::printf("PASS: %dx%d frame, %d hits, %.1f Mraycasts; %zu shadow rays (%d occluded) in "
              "%zu queued batches\n",
              W, H, hits, ft.Mraycasts(), sh.size(), occluded, qc.Batches());
  return 0;
}
