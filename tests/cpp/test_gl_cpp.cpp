// C++ host test of the GLES draw path (include/srf_gl.hpp) on the GPU: four
// draws of different primitive kinds -- a triangle heightfield, a hexahedron
// grid, the same grid split into tetrahedra, and points -- each placed by its
// own lsgl_World matrix, then one Finish over all of them. Checks: the GL
// errors of invalid draws; the draw cache (a second frame of the same draws
// builds nothing); and every pixel's world hit distance, primitive and element
// equal the oracle's RenderGraph walk over the same elements and eye rays
// (oracle/liboracle.so: or_scene_trace_mixed, whose walk equals the
// reference's RenderGraph::Trace), with missed pixels untouched. Prints "PASS".
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/srf_gl.hpp"
#include "../../oracle/srf_oracle.h"

#define CHECK(x)                                                   \
  do {                                                             \
    if (!(x)) {                                                    \
      std::fprintf(stderr, "FAILED: %s (line %d)\n", #x, __LINE__); \
      return 1;                                                    \
    }                                                              \
  } while (0)

namespace {

struct Geo {
  std::vector<float> V;
  std::vector<uint32_t> I;
};

Geo heightfield(int g) {
  Geo m;
  for (int j = 0; j < g; j++)
    for (int i = 0; i < g; i++) {
      const float x = i / float(g - 1), y = j / float(g - 1);
      m.V.insert(m.V.end(), {x, y, 0.08f * std::sin(9.0f * x) * std::cos(7.0f * y)});
    }
  for (int j = 0; j + 1 < g; j++)
    for (int i = 0; i + 1 < g; i++) {
      const uint32_t a = j * g + i, b = a + 1, c = a + g, d = c + 1;
      m.I.insert(m.I.end(), {a, b, d, a, d, c});
    }
  return m;
}

// n^3 cells of size 1/n: hexahedra (corners 0-3 the lower face counter-
// clockwise seen from above, 4-7 above them) or 6 tetrahedra per cell (the
// Kuhn split along the main diagonal, each oriented with
// det(v1 - v0, v2 - v0, v3 - v0) < 0)
Geo cell_grid(int n, bool tets) {
  Geo m;
  const int n1 = n + 1;
  for (int z = 0; z < n1; z++)
    for (int y = 0; y < n1; y++)
      for (int x = 0; x < n1; x++) {
        // a small deterministic wobble keeps faces off the lattice planes
        const float w = 0.05f * std::sin(1.7f * x + 2.3f * y + 3.1f * z);
        m.V.insert(m.V.end(), {(x + w) / n, (y - w) / n, (z + 0.5f * w) / n});
      }
  auto at = [&](int x, int y, int z) { return (uint32_t)((z * n1 + y) * n1 + x); };
  static const int kuhn[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7},
                                 {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};
  for (int z = 0; z < n; z++)
    for (int y = 0; y < n; y++)
      for (int x = 0; x < n; x++) {
        uint32_t c[8];
        for (int k = 0; k < 8; k++) c[k] = at(x + (k & 1), y + ((k >> 1) & 1), z + (k >> 2));
        if (!tets) {
          m.I.insert(m.I.end(), {c[0], c[1], c[3], c[2], c[4], c[5], c[7], c[6]});
          continue;
        }
        for (auto &t : kuhn) {
          uint32_t q[4] = {c[t[0]], c[t[1]], c[t[2]], c[t[3]]};
          const float *a = &m.V[3 * q[0]], *b = &m.V[3 * q[1]], *cc = &m.V[3 * q[2]],
                      *d = &m.V[3 * q[3]];
          double u[3], v[3], w[3];
          for (int k = 0; k < 3; k++) u[k] = b[k] - a[k], v[k] = cc[k] - a[k], w[k] = d[k] - a[k];
          const double det = (u[1] * v[2] - u[2] * v[1]) * w[0] +
                             (u[2] * v[0] - u[0] * v[2]) * w[1] + (u[0] * v[1] - u[1] * v[0]) * w[2];
          if (det > 0) std::swap(q[1], q[2]);
          m.I.insert(m.I.end(), {q[0], q[1], q[2], q[3]});
        }
      }
  return m;
}

Geo points(int n) {
  Geo m;
  uint32_t s = 12345u;
  auto rnd = [&]() {
    s = s * 1664525u + 1013904223u;
    return (s >> 8) * (1.0f / 16777216.0f);
  };
  for (int i = 0; i < n; i++) {
    m.V.insert(m.V.end(), {rnd(), rnd(), 0.5f * rnd()});
    m.I.push_back((uint32_t)i);
  }
  return m;
}

void translate_rotate(float M[16], float tx, float ty, float tz, float ang) {
  // row-vector convention p' = p M: rotation about z, then row 3 translation
  const float c = std::cos(ang), s = std::sin(ang);
  const float r[16] = {c, s, 0, 0, -s, c, 0, 0, 0, 0, 1, 0, tx, ty, tz, 1};
  std::memcpy(M, r, sizeof(r));
}

}  // namespace

int main() {
  srf::Context ctx(0);
  CHECK(ctx.ok());
  srf::gl::DrawContext gl(ctx);

  const Geo hf = heightfield(48), hex = cell_grid(11, false), tet = cell_grid(8, true),
            pts = points(3000);
  srf::gl::DrawBuffers b[4];
  const Geo *geo[4] = {&hf, &hex, &tet, &pts};
  const uint32_t modes[4] = {srf::gl::kTriangles, srf::gl::kHexahedronsExt,
                             srf::gl::kTetrahedronsExt, srf::gl::kPoints};
  for (int e = 0; e < 4; e++) {
    b[e].elements = geo[e]->I.data();
    b[e].num_elements = geo[e]->I.size();
    b[e].positions = geo[e]->V.data();
    b[e].num_vertices = geo[e]->V.size() / 3;
  }
  translate_rotate(b[0].world, 0.0f, 0.0f, 0.0f, 0.0f);
  translate_rotate(b[1].world, 1.3f, 0.1f, 0.1f, 0.4f);
  translate_rotate(b[2].world, -1.2f, 0.2f, 0.15f, -0.3f);
  translate_rotate(b[3].world, 0.1f, 1.2f, 0.2f, 0.2f);
  b[3].point_size = 0.03f;  // radius 0.015

  // the reference's GL errors
  gl.DrawElements(0x0005 /* GL_TRIANGLE_STRIP */, 3, srf::gl::kUnsignedInt, 0, b[0]);
  CHECK(gl.GetError() == srf::gl::kInvalidEnum);
  gl.DrawElements(srf::gl::kTriangles, -3, srf::gl::kUnsignedInt, 0, b[0]);
  CHECK(gl.GetError() == srf::gl::kInvalidValue);
  gl.DrawElements(srf::gl::kTriangles, 3, 0x1403 /* GL_UNSIGNED_SHORT */, 0, b[0]);
  CHECK(gl.GetError() == srf::gl::kInvalidEnum);
  gl.DrawElements(srf::gl::kTriangles, 0, srf::gl::kUnsignedInt, 0, b[0]);
  CHECK(gl.GetError() == srf::gl::kNoError && gl.FrameElements() == 0);

  const int W = 200, H = 150;
  srf::Camera cam = {{0.0f, -3.2f, 2.4f}, {0.1f, 0.5f, 0.2f}, {0.0f, 0.0f, 1.0f}, 50.0f};
  auto shade = [](int, int, const srf_ray &, const srf_hit &h, uint32_t e, float c[4]) {
    c[0] = h.t;
    std::memcpy(&c[1], &h.prim_id, 4);
    std::memcpy(&c[2], &e, 4);
    c[3] = 1.0f;
    return true;
  };
  std::vector<float> rgba(4 * W * H), depth(W * H);
  for (int frame = 0; frame < 2; frame++) {
    for (int e = 0; e < 4; e++)
      gl.DrawElements(modes[e], (int32_t)geo[e]->I.size(), srf::gl::kUnsignedInt, 0, b[e]);
    CHECK(gl.GetError() == srf::gl::kNoError);
    CHECK(gl.FrameElements() == 4 && gl.CachedAccels() == 4);  // frame 2 rebuilds nothing
    std::fill(rgba.begin(), rgba.end(), -1.0f);
    std::fill(depth.begin(), depth.end(), -1.0f);
    CHECK(gl.Finish(cam, W, H, shade, rgba.data(), depth.data()));
    CHECK(gl.FrameElements() == 0);
  }

  // the oracle: the same elements (trees, element matrices, top level) and rays
  std::vector<or_element> els(4);
  std::vector<std::vector<char>> keep_nodes(4);
  std::vector<std::vector<uint32_t>> keep_idx(4);
  std::vector<double> boxes(6 * 4);
  for (int e = 0; e < 4; e++) {
    const Geo &g = *geo[e];
    const int64_t n = e == 0 ? (int64_t)g.I.size() / 3
                             : (e == 1 ? (int64_t)g.I.size() / 8
                                       : (e == 2 ? (int64_t)g.I.size() / 4 : (int64_t)g.I.size()));
    double bounds[6], eps = 0.0;
    keep_idx[e].resize(n);
    or_element &x = els[e];
    std::memset(&x, 0, sizeof(x));
    if (e == 0) {
      keep_nodes[e].resize(sizeof(or_tri_node) * 2 * n);
      or_build_tri(g.V.data(), g.I.data(), n, (or_tri_node *)keep_nodes[e].data(),
                   keep_idx[e].data(), bounds, &eps);
      x.kind = 0;
    } else if (e == 1 || e == 2) {
      keep_nodes[e].resize(sizeof(or_tet_node) * 2 * n);
      if (e == 1)
        or_build_sol(g.V.data(), g.I.data(), n, 8, (or_tet_node *)keep_nodes[e].data(),
                     keep_idx[e].data(), bounds);
      else
        or_build_tet(g.V.data(), g.I.data(), n, (or_tet_node *)keep_nodes[e].data(),
                     keep_idx[e].data(), bounds);
      x.kind = e == 1 ? 3 : 1;
      x.nv = 8;
    } else {
      keep_nodes[e].resize(sizeof(or_sph_node) * 2 * n);
      or_build_sph(g.V.data(), nullptr, 0.015f, n, (or_sph_node *)keep_nodes[e].data(),
                   keep_idx[e].data(), bounds);
      x.kind = 2;
      x.rconst = 0.015f;
    }
    x.nodes = keep_nodes[e].data();
    x.indices = keep_idx[e].data();
    x.V = g.V.data();
    x.F = e == 3 ? nullptr : g.I.data();
    x.eps_scale = eps;
    srf_instance inst{};
    for (int k = 0; k < 16; k++) inst.T[k] = (double)b[e].world[k];
    CHECK(srf_instance_update(&inst) == SRF_OK);
    std::memcpy(x.T, inst.T, sizeof(x.T));
    std::memcpy(x.invT, inst.invT, sizeof(x.invT));
    std::memcpy(x.invT33, inst.invT33, sizeof(x.invT33));
    or_transform_box(bounds, bounds + 3, x.T, &boxes[6 * e]);
  }
  std::vector<or_tlas_node> tl(8);
  or_tlas_build(boxes.data(), 4, tl.data());
  std::vector<srf_ray> rays(W * H);
  std::vector<uint8_t> valid(W * H);
  srf::EyeRays(cam, W, H, rays.data(), valid.data());
  std::vector<or_hit> want(rays.size());
  std::vector<uint32_t> wel(rays.size());
  or_scene_trace_mixed(tl.data(), els.data(), reinterpret_cast<const or_ray *>(rays.data()),
                       (int64_t)rays.size(), want.data(), wel.data(), 1);
  int hit = 0, per_el[4] = {0, 0, 0, 0};
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const size_t i = (size_t)y * W + x, o = (size_t)(H - y - 1) * W + x;
      if (!valid[i] || want[i].prim_id == 0xffffffffu) {
        CHECK(rgba[4 * o + 3] == -1.0f && depth[o] == -1.0f);
        continue;
      }
      uint32_t prim, el;
      std::memcpy(&prim, &rgba[4 * o + 1], 4);
      std::memcpy(&el, &rgba[4 * o + 2], 4);
      CHECK(std::memcmp(&rgba[4 * o], &want[i].t, 4) == 0);
      CHECK(prim == want[i].prim_id && el == wel[i] && el < 4);
      CHECK(std::memcmp(&depth[o], &want[i].t, 4) == 0);
      hit++;
      per_el[el]++;
    }
  std::printf("hit pixels %d (elements %d %d %d %d), %.1f Mraycasts\n", hit, per_el[0], per_el[1],
              per_el[2], per_el[3], gl.Tracer().Mraycasts());
  CHECK(hit > W * H / 10);
  CHECK(per_el[0] > 0 && per_el[1] > 0 && per_el[2] > 0 && per_el[3] > 0);
  std::printf("PASS\n");
  return 0;
}
