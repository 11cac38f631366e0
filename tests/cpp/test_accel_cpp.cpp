// C++ host test of the drop-in classes (include/srf_accel.hpp) over the C
// ABI: builds a grid heightfield with srf::TriangleAccel, traces a pinhole
// image of primary rays and checks tree and hits against the oracle's
// TriangleAccel::Build32 / Traverse restatement (oracle/liboracle.so).
// Prints "PASS" and returns 0 on success.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/srf_accel.hpp"
#include "../../oracle/srf_oracle.h"

#define CHECK(x)                                                   \
  do {                                                             \
    if (!(x)) {                                                    \
      std::fprintf(stderr, "FAILED: %s (line %d)\n", #x, __LINE__); \
      return 1;                                                    \
    }                                                              \
  } while (0)

int main() {
  const int G = 96;  // 96 x 96 grid -> 18,050 triangles
  std::vector<float> V;
  for (int j = 0; j < G; j++)
    for (int i = 0; i < G; i++) {
      float x = i / float(G - 1), y = j / float(G - 1);
      V.push_back(x);
      V.push_back(y);
      V.push_back(0.05f * std::sin(9.0f * x) * std::cos(7.0f * y));
    }
  std::vector<uint32_t> F;
  for (int j = 0; j + 1 < G; j++)
    for (int i = 0; i + 1 < G; i++) {
      uint32_t a = j * G + i, b = a + 1, c = a + G, d = c + 1;
      F.insert(F.end(), {a, b, d, a, d, c});
    }
  const uint64_t nv = V.size() / 3, nf = F.size() / 3;
  const int W = 128, H = 128;
  std::vector<srf_ray> rays(W * H);
  for (int py = 0; py < H; py++)
    for (int px = 0; px < W; px++) {
      srf_ray &r = rays[py * W + px];
      r.org[0] = 0.5f;
      r.org[1] = -0.3f;
      r.org[2] = 0.9f;
      float dx = (px + 0.5f) / W - 0.5f, dz = -(py + 0.5f) / H + 0.2f;
      float d[3] = {dx, 1.0f, dz - 0.8f};
      float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      for (int k = 0; k < 3; k++) r.dir[k] = d[k] / len;
      r.flags = SRF_RAY_DOUBLE_SIDED;
      r.prev_prim_id = SRF_NO_PRIM;
    }

  float *dV;
  uint32_t *dF;
  srf_ray *dR;
  srf_hit *dH;
  CHECK(hipMalloc(&dV, V.size() * 4) == hipSuccess);
  CHECK(hipMalloc(&dF, F.size() * 4) == hipSuccess);
  CHECK(hipMalloc(&dR, rays.size() * sizeof(srf_ray)) == hipSuccess);
  CHECK(hipMalloc(&dH, rays.size() * sizeof(srf_hit)) == hipSuccess);
  CHECK(hipMemcpy(dV, V.data(), V.size() * 4, hipMemcpyHostToDevice) == hipSuccess);
  CHECK(hipMemcpy(dF, F.data(), F.size() * 4, hipMemcpyHostToDevice) == hipSuccess);
  CHECK(hipMemcpy(dR, rays.data(), rays.size() * sizeof(srf_ray), hipMemcpyHostToDevice) ==
        hipSuccess);

  srf::Context ctx(0);
  CHECK(ctx.ok());
  srf::TriangleAccel accel(ctx);
  CHECK(accel.Build32(dV, nv, dF, nf));
  CHECK(accel.Traverse(dR, rays.size(), dH));
  std::vector<srf_hit> hits(rays.size());
  CHECK(hipMemcpy(hits.data(), dH, hits.size() * sizeof(srf_hit), hipMemcpyDeviceToHost) ==
        hipSuccess);
  srf_accel_info info;
  CHECK(srf_accel_get_info(accel.handle(), &info) == SRF_OK);
  std::vector<or_tri_node> nodes(info.num_nodes);
  std::vector<uint32_t> idx(nf);
  CHECK(accel.Export(nodes.data(), idx.data()));
  double bmin[3], bmax[3];
  accel.BoundingBox(bmin, bmax);

  std::vector<or_tri_node> want(2 * nf);
  std::vector<uint32_t> want_idx(nf);
  double bounds[6], eps;
  int64_t nn = or_build_tri(V.data(), F.data(), nf, want.data(), want_idx.data(), bounds, &eps);
  CHECK((uint64_t)nn == info.num_nodes);
  for (int64_t i = 0; i < nn; i++) {
    or_tri_node a = nodes[i], b = want[i];
    if (a.flag == 1) {  // the reference leaves a leaf's axis and second slot unset
      a.axis = b.axis = 0;
      std::memset(a.bmin[1], 0, 12);
      std::memset(a.bmax[1], 0, 12);
      std::memset(b.bmin[1], 0, 12);
      std::memset(b.bmax[1], 0, 12);
    }
    CHECK(std::memcmp(&a, &b, sizeof(a)) == 0);
  }
  CHECK(idx == want_idx);
  for (int k = 0; k < 3; k++) CHECK(bmin[k] == bounds[k] && bmax[k] == bounds[3 + k]);
  std::vector<or_hit> want_hits(rays.size());
  or_trace_tri(want.data(), want_idx.data(), V.data(), F.data(), eps,
               reinterpret_cast<const or_ray *>(rays.data()), rays.size(), want_hits.data(),
               nullptr, 1);
  int hit_count = 0;
  for (size_t i = 0; i < rays.size(); i++) {
    CHECK(std::memcmp(&hits[i], &want_hits[i], sizeof(srf_hit)) == 0);
    hit_count += hits[i].prim_id != SRF_NO_PRIM;
  }
  CHECK(hit_count > (int)rays.size() / 4);

  // Build32 from the host arrays gives the same tree; the per-ray
  // Traverse(Intersection&, Ray&), called from 4 host threads at once,
  // gives the oracle's hit and BuildIntersection fields
  // (render/render_accel_triangle.cc:2439-2507) for every ray.
  srf::TriangleAccel host_accel(ctx);
  CHECK(host_accel.Build32Host(V.data(), nv, F.data(), nf));
  std::vector<or_tri_node> hnodes(info.num_nodes);
  std::vector<uint32_t> hidx(nf);
  CHECK(host_accel.Export(hnodes.data(), hidx.data()));
  CHECK(std::memcmp(hnodes.data(), nodes.data(), nodes.size() * sizeof(or_tri_node)) == 0);
  CHECK(hidx == idx);
  std::vector<or_isect> want_is(rays.size());
  or_isect_tri(V.data(), F.data(), reinterpret_cast<const or_ray *>(rays.data()),
               want_hits.data(), rays.size(), want_is.data());
  std::vector<srf::Intersection> got(rays.size());
  std::vector<int> got_hit(rays.size(), -1);
  std::vector<std::thread> pool;
  for (int w = 0; w < 4; w++)
    pool.emplace_back([&, w]() {
      for (size_t i = w; i < rays.size(); i += 4 * 7) {
        srf::Ray r;
        for (int k = 0; k < 3; k++) {
          r.org[k] = rays[i].org[k];
          r.dir[k] = rays[i].dir[k];
        }
        got_hit[i] = host_accel.Traverse(got[i], r) ? 1 : 0;
      }
    });
  for (auto &t : pool) t.join();
  int checked = 0;
  for (size_t i = 0; i < rays.size(); i++) {
    if (got_hit[i] < 0) continue;
    checked++;
    const bool hit = want_hits[i].prim_id != SRF_NO_PRIM;
    CHECK(got_hit[i] == (hit ? 1 : 0));
    CHECK(got[i].prim_id == want_hits[i].prim_id);
    CHECK(std::memcmp(&got[i].t, &want_hits[i].t, 4) == 0);
    if (!hit) continue;
    CHECK(std::memcmp(&got[i].u, &want_hits[i].u, 4) == 0);
    CHECK(std::memcmp(&got[i].v, &want_hits[i].v, 4) == 0);
    CHECK(std::memcmp(got[i].position, want_is[i].position, 12) == 0);
    CHECK(std::memcmp(got[i].geometric, want_is[i].geometric, 12) == 0);
    CHECK(std::memcmp(got[i].normal, want_is[i].normal, 12) == 0);
    CHECK(got[i].f0 == want_is[i].f0 && got[i].f1 == want_is[i].f1 &&
          got[i].f2 == want_is[i].f2);
  }
  CHECK(checked > 500);
  // the batched host form gives the same hits
  std::vector<srf_hit> hh(rays.size());
  CHECK(host_accel.TraverseHost(rays.data(), rays.size(), hh.data()));
  CHECK(std::memcmp(hh.data(), want_hits.data(), hh.size() * sizeof(srf_hit)) == 0);

  // a mesh below Build32's 1024-primitive floor is refused, as documented
  srf::TriangleAccel small(ctx);
  CHECK(!small.Build32(dV, nv, dF, 1000));
  CHECK(!small.LastError().empty());
  hipFree(dV);
  hipFree(dF);
  hipFree(dR);
  hipFree(dH);
  std::printf("PASS: %llu triangles, %llu nodes, %d of %zu rays hit, bit-exact vs oracle\n",
              (unsigned long long)nf, (unsigned long long)info.num_nodes, hit_count,
              rays.size());
  return 0;
}
