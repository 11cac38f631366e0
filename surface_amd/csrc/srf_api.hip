// srf_api.hip -- the extern "C" entry points of include/srf.h.
#include <float.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>

#include <vector>

#include "srf_internal.h"

namespace srf {

static thread_local char g_err[512] = "";

srf_status set_error(srf_status s, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return s;
}

srf_status hip_error(hipError_t e, const char *what) {
  if (e == hipErrorOutOfMemory)
    return set_error(SRF_ERR_OUT_OF_MEMORY, "%s: %s", what, hipGetErrorString(e));
  return set_error(SRF_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

srf_status ensure_scratch(Ctx *ctx, size_t bytes, hipStream_t stream) {
  if (ctx->scratch_bytes >= bytes) return SRF_OK;
  // the old scratch may still be read by enqueued work
  SRF_HIP(hipStreamSynchronize(stream));
  SRF_HIP(hipDeviceSynchronize());
  if (ctx->scratch) SRF_HIP(hipFree(ctx->scratch));
  ctx->scratch = nullptr;
  ctx->scratch_bytes = 0;
  size_t sz = bytes + bytes / 4 + (1u << 20);
  SRF_HIP(hipMalloc(&ctx->scratch, sz));
  ctx->scratch_bytes = sz;
  return SRF_OK;
}

srf_status ensure_host_stage(Ctx *ctx, size_t bytes) {
  SRF_HIP(hipEventSynchronize(ctx->stage_done));
  if (ctx->host_stage_bytes >= bytes) return SRF_OK;
  if (ctx->host_stage) SRF_HIP(hipHostFree(ctx->host_stage));
  ctx->host_stage = nullptr;
  ctx->host_stage_bytes = 0;
  size_t sz = bytes + 4096;
  SRF_HIP(hipHostMalloc(&ctx->host_stage, sz, hipHostMallocDefault));
  ctx->host_stage_bytes = sz;
  return SRF_OK;
}

void free_sub_buffers(Accel *a) {
  (void)hipFree(a->sub_nodes);
  (void)hipFree(a->sub_prims);
  (void)hipFree(a->sub_sidx);
  (void)hipFree(a->sub_params);
  a->sub_nodes = a->sub_prims = nullptr;
  a->sub_sidx = nullptr;
  a->sub_params = nullptr;
  a->sub_cap_prims = a->sub_cap_nodes = a->sub_cap_trees = 0;
  a->sub_nodes_used = a->sub_prims_used = 0;
  a->nbig = 0;
}

namespace {

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

size_t node_bytes(Kind) { return 64; }
// packed primitive record: triangle 48 B, tetra 64 B, sphere 16 B, solid
// 16 B per vertex
size_t prim_bytes(Kind k, uint32_t nvps) {
  if (k == kLin) return 32;
  return k == kTri ? 48 : (k == kTet ? 64 : (k == kSol ? 16 * (size_t)nvps : 16));
}

void free_buffers(Accel *a);

void free_accel(Accel *a) {
  if (!a) return;
  DeviceGuard g(a->device);
  if (a->built) (void)hipEventSynchronize(a->built);
  free_buffers(a);
  free_geometry(a);
  (void)hipFree(a->params);
  if (a->built) (void)hipEventDestroy(a->built);
  delete a;
}

hipError_t alloc_buffers(Accel *a, uint64_t n) {
  hipError_t e = hipSuccess;
  const Kind kind = a->kind;
  if (e == hipSuccess) e = hipMalloc(&a->nodes, node_bytes(kind) * (2 * n));
  if (e == hipSuccess) e = hipMalloc(&a->prims, prim_bytes(kind, a->nvps) * n);
  if (e == hipSuccess) e = hipMalloc((void **)&a->sorted_index, 4 * n);
  if (e == hipSuccess && kind == kTet) e = hipMalloc((void **)&a->sorted_faces, 16 * n);
  if (e == hipSuccess && kind == kSph) {
    // every big leaf holds more than kBigLeaf spheres
    a->big_cap = n / (kBigLeaf + 1) + 1;
    e = hipMalloc((void **)&a->big_info, 16 * a->big_cap);
    if (e == hipSuccess) e = hipMalloc((void **)&a->big_count, 4);
  }
  if (e == hipSuccess) a->capacity = n;
  return e;
}

void free_buffers(Accel *a) {
  (void)hipFree(a->nodes);
  (void)hipFree(a->prims);
  (void)hipFree(a->sorted_index);
  (void)hipFree(a->sorted_faces);
  (void)hipFree(a->big_info);
  (void)hipFree(a->big_count);
  a->nodes = a->prims = nullptr;
  a->sorted_index = a->sorted_faces = nullptr;
  a->big_info = nullptr;
  a->big_count = nullptr;
  a->big_cap = 0;
  a->capacity = 0;
  free_sub_buffers(a);
}

srf_status alloc_accel(Ctx *ctx, Kind kind, uint64_t n, uint64_t nv, Accel **out,
                       uint32_t nvps = 0) {
  Accel *a = new (std::nothrow) Accel();
  if (!a) return set_error(SRF_ERR_OUT_OF_MEMORY, "host allocation failed");
  memset(a, 0, sizeof(*a));
  a->kind = kind;
  a->device = ctx->device;
  a->n = n;
  a->nv = nv;
  a->nvps = nvps;
  hipError_t e = alloc_buffers(a, n);
  if (e == hipSuccess) e = hipMalloc((void **)&a->params, sizeof(BuildParams));
  if (e == hipSuccess) e = hipEventCreateWithFlags(&a->built, hipEventDisableTiming);
  if (e != hipSuccess) {
    free_accel(a);
    return hip_error(e, "accelerator allocation");
  }
  *out = a;
  return SRF_OK;
}

srf_status check_count(uint64_t n) {
  if (n < 1024)
    return set_error(SRF_ERR_UNSUPPORTED,
                     "%llu primitives: Build32 needs at least 1024 (the reference uses its "
                     "SAH Build below that)",
                     (unsigned long long)n);
  if (n >= (1ull << 28))
    return set_error(SRF_ERR_UNSUPPORTED, "%llu primitives exceed the 2^28 limit",
                     (unsigned long long)n);
  return SRF_OK;
}

srf_status build_common(srf_ctx *c, Kind kind, const void *desc, uint64_t n, uint64_t nv,
                        void *stream, srf_accel **out, uint32_t nvps = 0) {
  Ctx *ctx = reinterpret_cast<Ctx *>(c);
  if (!ctx || !desc || !out) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  srf_status cs = check_count(n);
  if (cs != SRF_OK) return cs;
  DeviceGuard g(ctx->device);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  Accel *a = nullptr;
  srf_status st = alloc_accel(ctx, kind, n, nv, &a, nvps);
  if (st != SRF_OK) return st;
  st = build_lbvh(ctx, kind, desc, s, a);
  if (st == SRF_OK) {
    record_geometry(a, desc);
    hipError_t e = hipEventRecord(a->built, s);
    if (e != hipSuccess) st = hip_error(e, "hipEventRecord");
  }
  if (st != SRF_OK) {
    free_accel(a);
    return st;
  }
  *out = reinterpret_cast<srf_accel *>(a);
  return SRF_OK;
}

srf_status fetch_params(const Accel *a, BuildParams *p) {
  DeviceGuard g(a->device);
  SRF_HIP(hipEventSynchronize(a->built));
  SRF_HIP(hipMemcpy(p, a->params, sizeof(BuildParams), hipMemcpyDeviceToHost));
  if (p->error)
    return set_error(SRF_ERR_INTERNAL, "a bounded device wait gave up during the build");
  return SRF_OK;
}

}  // namespace

srf_status fetch_params_for(const Accel *a, BuildParams *p) { return fetch_params(a, p); }

struct Scene;
srf_status scene_create(Ctx *ctx, uint32_t n, const srf_instance *inst, Scene **out);
void cramer_inverse(const double in[16], double m[16]);
srf_status scene_trace(const Scene *sc, const srf_ray *rays, uint64_t n, srf_hit *hits,
                       uint32_t *inst_out, int mode, hipStream_t stream);
srf_status scene_autotune(Scene *sc, const srf_ray *rays, uint64_t n, srf_hit *hits,
                          uint32_t *inst_out, hipStream_t stream, srf_launch_config *best);
void scene_free(Scene *sc);
srf_status scene_query(const Scene *sc, double bmin[3], double bmax[3], srf_tlas_node *nodes,
                       uint32_t *count);
int scene_device(const Scene *sc);
struct SparseVolume;
srf_status sparse_create(Ctx *ctx, const srf_volume_block *blocks, uint32_t count,
                         const int global_dim[3], int components, int format,
                         SparseVolume **out);
void sparse_free(SparseVolume *sv);
srf_status sparse_sample(const SparseVolume *sv, const double *pos, const float *uvw, uint64_t n,
                         double *out_d, float *out_f, hipStream_t s);

}  // namespace srf

using namespace srf;

extern "C" {

const char *srf_last_error_string(void) { return srf::g_err; }

int srf_abi_version(void) { return SRF_ABI_VERSION; }

srf_status srf_ctx_create(int device, srf_ctx **out) {
  if (!out) return set_error(SRF_ERR_INVALID_ARG, "null out");
  int count = 0;
  SRF_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count)
    return set_error(SRF_ERR_INVALID_ARG, "device %d out of range (%d devices)", device, count);
  Ctx *c = new (std::nothrow) Ctx();
  if (!c) return set_error(SRF_ERR_OUT_OF_MEMORY, "host allocation failed");
  c->device = device;
  c->scratch = nullptr;
  c->scratch_bytes = 0;
  c->host_stage = nullptr;
  c->host_stage_bytes = 0;
  {
    DeviceGuard g(device);
    hipError_t e = hipEventCreateWithFlags(&c->stage_done, hipEventDisableTiming);
    if (e != hipSuccess) {
      delete c;
      return hip_error(e, "hipEventCreate");
    }
  }
  *out = reinterpret_cast<srf_ctx *>(c);
  return SRF_OK;
}

srf_status srf_ctx_destroy(srf_ctx *ctx) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c) return SRF_OK;
  {
    DeviceGuard g(c->device);
    (void)hipDeviceSynchronize();
    if (c->scratch) (void)hipFree(c->scratch);
    if (c->host_stage) (void)hipHostFree(c->host_stage);
    (void)hipEventDestroy(c->stage_done);
  }
  delete c;
  return SRF_OK;
}

srf_status srf_accel_build_triangles(srf_ctx *ctx, const srf_tri_desc *d, void *stream,
                                     srf_accel **out) {
  if (!d || !d->vertices || !d->faces) return set_error(SRF_ERR_INVALID_ARG, "null mesh");
  return build_common(ctx, kTri, d, d->num_faces, d->num_vertices, stream, out);
}

srf_status srf_accel_build_tetras(srf_ctx *ctx, const srf_tet_desc *d, void *stream,
                                  srf_accel **out) {
  if (!d || !d->vertices || !d->indices) return set_error(SRF_ERR_INVALID_ARG, "null tetras");
  return build_common(ctx, kTet, d, d->num_tetras, d->num_vertices, stream, out);
}

srf_status srf_accel_build_solids(srf_ctx *ctx, const srf_solid_desc *d, void *stream,
                                  srf_accel **out) {
  if (!d || !d->vertices || !d->indices) return set_error(SRF_ERR_INVALID_ARG, "null solids");
  const uint32_t k = d->vertices_per_solid;
  if (k != 5 && k != 6 && k != 8)
    return set_error(SRF_ERR_INVALID_ARG,
                     "%u vertices per solid: 5 (pyramid), 6 (prism) or 8 (hexahedron)", k);
  return build_common(ctx, kSol, d, d->num_solids, d->num_vertices, stream, out, k);
}

srf_status srf_accel_build_lines(srf_ctx *ctx, const srf_line_desc *d, void *stream,
                                 srf_accel **out) {
  if (!d || !d->positions || !d->segments) return set_error(SRF_ERR_INVALID_ARG, "null lines");
  srf_status st = build_common(ctx, kLin, d, d->num_lines, d->num_vertices, stream, out);
  if (st == SRF_OK) reinterpret_cast<Accel *>(*out)->line_cap = d->cap ? 1u : 0u;
  return st;
}

srf_status srf_accel_build_spheres(srf_ctx *ctx, const srf_sphere_desc *d, void *stream,
                                   srf_accel **out) {
  if (!d || !d->positions) return set_error(SRF_ERR_INVALID_ARG, "null particles");
  return build_common(ctx, kSph, d, d->num_particles, d->num_particles, stream, out);
}

srf_status srf_accel_build_triangles_batch(srf_ctx *ctx, uint32_t num_meshes,
                                           const srf_tri_desc *descs, void *stream,
                                           srf_accel **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !descs || !out || num_meshes == 0)
    return set_error(SRF_ERR_INVALID_ARG, "null argument or empty batch");
  for (uint32_t m = 0; m < num_meshes; m++) out[m] = nullptr;
  for (uint32_t m = 0; m < num_meshes; m++) {
    if (!descs[m].vertices || !descs[m].faces)
      return set_error(SRF_ERR_INVALID_ARG, "mesh %u: null arrays", m);
    srf_status cs = check_count(descs[m].num_faces);
    if (cs != SRF_OK) return cs;
  }
  DeviceGuard g(c->device);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  std::vector<Accel *> accs(num_meshes, nullptr);
  std::vector<const void *> dp(num_meshes);
  srf_status st = SRF_OK;
  for (uint32_t m = 0; m < num_meshes && st == SRF_OK; m++) {
    st = alloc_accel(c, kTri, descs[m].num_faces, descs[m].num_vertices, &accs[m]);
    dp[m] = &descs[m];
  }
  if (st == SRF_OK) st = build_lbvh_segments(c, kTri, dp.data(), accs.data(), num_meshes, s);
  for (uint32_t m = 0; m < num_meshes && st == SRF_OK; m++) {
    record_geometry(accs[m], &descs[m]);
    hipError_t e = hipEventRecord(accs[m]->built, s);
    if (e != hipSuccess) st = hip_error(e, "hipEventRecord");
  }
  if (st != SRF_OK) {
    (void)hipStreamSynchronize(s);
    for (Accel *a : accs)
      if (a) free_accel(a);
    return st;
  }
  for (uint32_t m = 0; m < num_meshes; m++) out[m] = reinterpret_cast<srf_accel *>(accs[m]);
  return SRF_OK;
}

srf_status srf_accel_rebuild_triangles_batch(srf_ctx *ctx, uint32_t num_meshes,
                                             srf_accel *const *accels,
                                             const srf_tri_desc *descs, void *stream) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !descs || !accels || num_meshes == 0)
    return set_error(SRF_ERR_INVALID_ARG, "null argument or empty batch");
  std::vector<Accel *> accs(num_meshes);
  std::vector<const void *> dp(num_meshes);
  for (uint32_t m = 0; m < num_meshes; m++) {
    Accel *a = reinterpret_cast<Accel *>(accels[m]);
    if (!a || a->kind != kTri || a->device != c->device)
      return set_error(SRF_ERR_INVALID_ARG, "accelerator %u: null, not triangles, or on "
                                            "another device", m);
    if (!descs[m].vertices || !descs[m].faces)
      return set_error(SRF_ERR_INVALID_ARG, "mesh %u: null arrays", m);
    srf_status cs = check_count(descs[m].num_faces);
    if (cs != SRF_OK) return cs;
    for (uint32_t k = 0; k < m; k++)
      if (accs[k] == a) return set_error(SRF_ERR_INVALID_ARG, "accelerator %u repeated", m);
    accs[m] = a;
    dp[m] = &descs[m];
  }
  DeviceGuard g(c->device);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  for (uint32_t m = 0; m < num_meshes; m++) {
    Accel *a = accs[m];
    if (descs[m].num_faces > a->capacity) {
      SRF_HIP(hipEventSynchronize(a->built));
      free_buffers(a);
      SRF_HIP(alloc_buffers(a, descs[m].num_faces));
    }
    a->n = descs[m].num_faces;
    a->nv = descs[m].num_vertices;
  }
  srf_status st = build_lbvh_segments(c, kTri, dp.data(), accs.data(), num_meshes, s);
  if (st != SRF_OK) return st;
  for (uint32_t m = 0; m < num_meshes; m++) {
    record_geometry(accs[m], &descs[m]);
    SRF_HIP(hipEventRecord(accs[m]->built, s));
  }
  return SRF_OK;
}

srf_status srf_accel_rebuild(srf_ctx *ctx, srf_accel *accel, const void *desc, void *stream) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  Accel *a = reinterpret_cast<Accel *>(accel);
  if (!c || !a || !desc) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (c->device != a->device) return set_error(SRF_ERR_INVALID_ARG, "context/device mismatch");
  uint64_t n, nv;
  if (a->kind == kTri) {
    const srf_tri_desc *d = (const srf_tri_desc *)desc;
    if (!d->vertices || !d->faces) return set_error(SRF_ERR_INVALID_ARG, "null mesh");
    n = d->num_faces;
    nv = d->num_vertices;
  } else if (a->kind == kTet) {
    const srf_tet_desc *d = (const srf_tet_desc *)desc;
    if (!d->vertices || !d->indices) return set_error(SRF_ERR_INVALID_ARG, "null tetras");
    n = d->num_tetras;
    nv = d->num_vertices;
  } else if (a->kind == kLin) {
    const srf_line_desc *d = (const srf_line_desc *)desc;
    if (!d->positions || !d->segments) return set_error(SRF_ERR_INVALID_ARG, "null lines");
    n = d->num_lines;
    nv = d->num_vertices;
    a->line_cap = d->cap ? 1u : 0u;
  } else if (a->kind == kSol) {
    const srf_solid_desc *d = (const srf_solid_desc *)desc;
    if (!d->vertices || !d->indices) return set_error(SRF_ERR_INVALID_ARG, "null solids");
    if (d->vertices_per_solid != a->nvps)
      return set_error(SRF_ERR_INVALID_ARG, "%u vertices per solid; the accelerator has %u",
                       d->vertices_per_solid, a->nvps);
    n = d->num_solids;
    nv = d->num_vertices;
  } else {
    const srf_sphere_desc *d = (const srf_sphere_desc *)desc;
    if (!d->positions) return set_error(SRF_ERR_INVALID_ARG, "null particles");
    n = d->num_particles;
    nv = n;
  }
  srf_status st = check_count(n);
  if (st != SRF_OK) return st;
  DeviceGuard g(a->device);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (n > a->capacity) {
    SRF_HIP(hipEventSynchronize(a->built));
    free_buffers(a);
    SRF_HIP(alloc_buffers(a, n));
  }
  a->n = n;
  a->nv = nv;
  st = build_lbvh(c, a->kind, desc, s, a);
  if (st != SRF_OK) return st;
  record_geometry(a, desc);
  SRF_HIP(hipEventRecord(a->built, s));
  return SRF_OK;
}

srf_status srf_accel_bounds(const srf_accel *accel, double bmin[3], double bmax[3]) {
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!a || !bmin || !bmax) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  BuildParams p;
  srf_status st = fetch_params(a, &p);
  if (st != SRF_OK) return st;
  for (int k = 0; k < 3; k++) {
    bmin[k] = p.root_bmin[k];
    bmax[k] = p.root_bmax[k];
  }
  return SRF_OK;
}

srf_status srf_accel_get_info(const srf_accel *accel, srf_accel_info *info) {
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!a || !info) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  BuildParams p;
  srf_status st = fetch_params(a, &p);
  if (st != SRF_OK) return st;
  if (a->kind == kLin)
    info->kind = SRF_PRIM_LINES;
  else if (a->kind == kSol)
    info->kind = a->nvps == 5 ? SRF_PRIM_PYRAMIDS
                              : (a->nvps == 6 ? SRF_PRIM_PRISMS : SRF_PRIM_HEXAHEDRONS);
  else
    info->kind = a->kind == kTri ? SRF_PRIM_TRIANGLES
                                 : (a->kind == kTet ? SRF_PRIM_TETRAHEDRONS : SRF_PRIM_POINTS);
  info->num_prims = a->n;
  info->num_nodes = p.num_nodes;
  info->num_leaves = p.num_leaves;
  info->max_depth = p.max_steps + 1;
  info->eps_scale = p.eps_scale;
  info->node_bytes =
      a->kind == kTri ? 64 : ((a->kind == kTet || a->kind == kSol || a->kind == kLin) ? 40 : 60);
  return SRF_OK;
}

srf_status srf_accel_export(const srf_accel *accel, void *nodes_host, uint32_t *indices_host) {
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!a || !nodes_host || !indices_host) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  BuildParams p;
  srf_status st = fetch_params(a, &p);
  if (st != SRF_OK) return st;
  DeviceGuard g(a->device);
  const size_t nn = p.num_nodes;
  SRF_HIP(hipMemcpy(indices_host, a->sorted_index, 4 * a->n, hipMemcpyDeviceToHost));
  if (a->kind == kTri) {
    SRF_HIP(hipMemcpy(nodes_host, a->nodes, 64 * nn, hipMemcpyDeviceToHost));
    return SRF_OK;
  }
  uint32_t *tmp = new (std::nothrow) uint32_t[16 * nn];
  if (!tmp) return set_error(SRF_ERR_OUT_OF_MEMORY, "host allocation failed");
  hipError_t e = hipMemcpy(tmp, a->nodes, 64 * nn, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    delete[] tmp;
    return hip_error(e, "hipMemcpy");
  }
  uint8_t *dst = reinterpret_cast<uint8_t *>(nodes_host);
  if (a->kind == kTet || a->kind == kSol || a->kind == kLin) {
    // TetraNode / SolidNode / LineNode: bmin[3], bmax[3], flag, axis, data[2] (40 B); a node's own
    // box is the one its parent holds for it, the root's is BoundingBox()
    float rb[6];
    for (int k = 0; k < 3; k++) {
      rb[k] = p.root_bmin[k];
      rb[3 + k] = p.root_bmax[k];
    }
    memcpy(dst, rb, 24);
    for (size_t i = 0; i < nn; i++) {
      const uint32_t *w = tmp + 16 * i;
      uint32_t meta[4] = {w[12], w[13], w[14], w[15]};
      memcpy(dst + 40 * i + 24, meta, 16);
      if (w[12] == 0) {
        for (int side = 0; side < 2; side++) {
          uint32_t c = w[14 + side];
          if (c >= nn) {
            delete[] tmp;
            return set_error(SRF_ERR_INTERNAL, "child index out of range");
          }
          float box[6];
          memcpy(box, w + 3 * side, 12);
          memcpy(box + 3, w + 6 + 3 * side, 12);
          memcpy(dst + 40 * (size_t)c, box, 24);
        }
      }
    }
  } else {
    // ParticleNode: bmin[2][3], bmax[2][3], data[2], short flag, short axis (60 B)
    for (size_t i = 0; i < nn; i++) {
      const uint32_t *w = tmp + 16 * i;
      uint32_t o[15];
      memcpy(o, w, 48);
      // a leaf's second box slot is unused in the reference (left
      // uninitialised); here words 3 and 4 link a big leaf to its
      // secondary tree, which is not part of the exported layout
      if (w[12] != 0) o[3] = o[4] = 0u;
      o[12] = w[14];
      o[13] = w[15];
      o[14] = (w[12] & 0xffffu) | (w[13] << 16);
      memcpy(dst + 60 * i, o, 60);
    }
  }
  delete[] tmp;
  return SRF_OK;
}

srf_status srf_accel_dump(const srf_accel *accel, const char *path) {
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!a || !path) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  srf_accel_info info;
  srf_status st = srf_accel_get_info(accel, &info);
  if (st != SRF_OK) return st;
  std::vector<uint8_t> nodes(info.num_nodes * info.node_bytes);
  std::vector<uint32_t> idx(a->n);
  st = srf_accel_export(accel, nodes.data(), idx.data());
  if (st != SRF_OK) return st;
  FILE *fp = fopen(path, "wb");
  if (!fp) return set_error(SRF_ERR_INVALID_ARG, "cannot write %s", path);
  // unsigned long long count, node records, count, unsigned int indices
  const unsigned long long nn = info.num_nodes, ni = a->n;
  bool ok = fwrite(&nn, sizeof(nn), 1, fp) == 1;
  ok = ok && fwrite(nodes.data(), 1, nodes.size(), fp) == nodes.size();
  ok = ok && fwrite(&ni, sizeof(ni), 1, fp) == 1;
  ok = ok && fwrite(idx.data(), 4, idx.size(), fp) == idx.size();
  ok = (fclose(fp) == 0) && ok;
  if (!ok) return set_error(SRF_ERR_INVALID_ARG, "short write to %s", path);
  return SRF_OK;
}

srf_status srf_accel_destroy(srf_accel *accel) {
  free_accel(reinterpret_cast<Accel *>(accel));
  return SRF_OK;
}

srf_status srf_trace_closest(const srf_accel *accel, const srf_ray *rays, uint64_t n,
                             srf_hit *hits, void *stream) {
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!a || (n && (!rays || !hits))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays");
  DeviceGuard g(a->device);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (a->kind == kTri) return trace_tri(a, rays, n, hits, s);
  if (a->kind == kTet) return trace_tet(a, rays, n, nullptr, hits, nullptr, s);
  if (a->kind == kSol) return trace_sol(a, rays, n, nullptr, hits, nullptr, s);
  if (a->kind == kLin) return trace_lin(a, rays, n, hits, nullptr, s);
  return trace_sph(a, rays, nullptr, n, hits, s);
}

srf_status srf_trace_lines(const srf_accel *accel, const srf_ray *rays, uint64_t n,
                           srf_hit *hits, srf_line_hit_ext *ext, void *stream) {
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!a || (n && (!rays || !hits))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (a->kind != kLin) return set_error(SRF_ERR_INVALID_ARG, "not a line accelerator");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays");
  DeviceGuard g(a->device);
  return trace_lin(a, rays, n, hits, ext, reinterpret_cast<hipStream_t>(stream));
}

srf_status srf_trace_solids(const srf_accel *accel, const srf_ray *rays, uint64_t n,
                            const float *scalars, srf_hit *hits, srf_solid_hit_ext *ext,
                            void *stream) {
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!a || (n && (!rays || !hits))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (a->kind != kSol) return set_error(SRF_ERR_INVALID_ARG, "not a solid accelerator");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays");
  DeviceGuard g(a->device);
  return trace_sol(a, rays, n, scalars, hits, ext, reinterpret_cast<hipStream_t>(stream));
}

srf_status srf_trace_tetra(const srf_accel *accel, const srf_ray *rays, uint64_t n,
                           const float *scalars, srf_hit *hits, srf_tet_hit_ext *ext,
                           void *stream) {
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!a || (n && (!rays || !hits))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (a->kind != kTet) return set_error(SRF_ERR_INVALID_ARG, "not a tetrahedron accelerator");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays");
  DeviceGuard g(a->device);
  return trace_tet(a, rays, n, scalars, hits, ext, reinterpret_cast<hipStream_t>(stream));
}

srf_status srf_trace_spheres(const srf_accel *accel, const srf_ray *rays,
                             const srf_sphere_prev *prev, uint64_t n, srf_hit *hits,
                             void *stream) {
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!a || (n && (!rays || !hits))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (a->kind != kSph) return set_error(SRF_ERR_INVALID_ARG, "not a sphere accelerator");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays");
  DeviceGuard g(a->device);
  return trace_sph(a, rays, prev, n, hits, reinterpret_cast<hipStream_t>(stream));
}

srf_status srf_trace_any(const srf_accel *accel, const srf_ray *rays, const float *t_max,
                         uint64_t n, uint32_t *occluded, void *stream) {
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!a || (n && (!rays || !occluded))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays");
  DeviceGuard g(a->device);
  return trace_any(a, rays, t_max, n, occluded, reinterpret_cast<hipStream_t>(stream));
}

srf_status srf_debug_force_wide_stack(srf_accel *accel, int on) {
  Accel *a = reinterpret_cast<Accel *>(accel);
  if (!a) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  a->wide_stack = on ? 1u : 0u;
  return SRF_OK;
}

srf_status srf_lbvh_morton30(srf_ctx *ctx, srf_prim_kind kind, const void *desc,
                             const float bmin[3], const float bmax[3], uint32_t *codes,
                             void *stream) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !desc || !bmin || !bmax || !codes)
    return set_error(SRF_ERR_INVALID_ARG, "null argument");
  Kind k;
  if (kind == SRF_PRIM_TRIANGLES)
    k = kTri;
  else if (kind == SRF_PRIM_TETRAHEDRONS)
    k = kTet;
  else if (kind == SRF_PRIM_POINTS)
    k = kSph;
  else if (kind == SRF_PRIM_LINES)
    k = kLin;
  else if (kind == SRF_PRIM_PYRAMIDS || kind == SRF_PRIM_PRISMS || kind == SRF_PRIM_HEXAHEDRONS) {
    const uint32_t want = kind == SRF_PRIM_PYRAMIDS ? 5 : (kind == SRF_PRIM_PRISMS ? 6 : 8);
    if (((const srf_solid_desc *)desc)->vertices_per_solid != want)
      return set_error(SRF_ERR_INVALID_ARG, "descriptor does not hold %u vertices per solid",
                       want);
    k = kSol;
  } else
    return set_error(SRF_ERR_INVALID_ARG, "unknown primitive kind 0x%x", (unsigned)kind);
  DeviceGuard g(c->device);
  return launch_morton(k, desc, bmin, bmax, codes, reinterpret_cast<hipStream_t>(stream));
}

srf_status srf_lbvh_sort30(srf_ctx *ctx, const uint32_t *codes, uint64_t n,
                           uint32_t *sorted_codes, uint32_t *sorted_index, void *stream) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || (n && (!codes || !sorted_codes || !sorted_index)))
    return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n == 0) return SRF_OK;
  if (n >= (1ull << 30)) return set_error(SRF_ERR_INVALID_ARG, "too many keys");
  DeviceGuard g(c->device);
  return sort_pairs30(c, codes, (uint32_t)n, sorted_codes, sorted_index,
                      reinterpret_cast<hipStream_t>(stream));
}

srf_status srf_lbvh_radix_tree30(srf_ctx *ctx, const uint32_t *sorted_codes, uint64_t n,
                                 srf_node_info *out, void *stream) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || (n > 1 && (!sorted_codes || !out)))
    return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n < 2) return SRF_OK;
  if (n >= (1ull << 31)) return set_error(SRF_ERR_INVALID_ARG, "too many keys");
  DeviceGuard g(c->device);
  return launch_radix_tree(sorted_codes, (uint32_t)n, out, reinterpret_cast<hipStream_t>(stream));
}

srf_status srf_instance_update(srf_instance *inst) {
  if (!inst) return set_error(SRF_ERR_INVALID_ARG, "null instance");
  double t33[16];
  memcpy(t33, inst->T, sizeof(t33));
  t33[12] = t33[13] = t33[14] = 0.0;  // row 3: the translation
  cramer_inverse(inst->T, inst->invT);
  cramer_inverse(t33, inst->invT33);
  return SRF_OK;
}

srf_status srf_scene_create(srf_ctx *ctx, uint32_t n, const srf_instance *inst,
                            srf_scene **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !inst || !out || n == 0)
    return set_error(SRF_ERR_INVALID_ARG, "null argument or empty scene");
  *out = nullptr;
  DeviceGuard g(c->device);
  Scene *sc = nullptr;
  srf_status st = scene_create(c, n, inst, &sc);
  if (st != SRF_OK) return st;
  *out = reinterpret_cast<srf_scene *>(sc);
  return SRF_OK;
}

srf_status srf_scene_destroy(srf_scene *scene) {
  Scene *sc = reinterpret_cast<Scene *>(scene);
  if (!sc) return SRF_OK;
  DeviceGuard g(scene_device(sc));
  (void)hipDeviceSynchronize();
  scene_free(sc);
  return SRF_OK;
}

srf_status srf_scene_bounds(const srf_scene *scene, double bmin[3], double bmax[3]) {
  const Scene *sc = reinterpret_cast<const Scene *>(scene);
  if (!sc || !bmin || !bmax) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  return scene_query(sc, bmin, bmax, nullptr, nullptr);
}

srf_status srf_scene_export(const srf_scene *scene, srf_tlas_node *nodes, uint32_t *count) {
  const Scene *sc = reinterpret_cast<const Scene *>(scene);
  if (!sc || !count) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  return scene_query(sc, nullptr, nullptr, nodes, count);
}

srf_status srf_scene_trace(const srf_scene *scene, const srf_ray *rays, uint64_t n,
                           srf_hit *hits, uint32_t *inst_out, int mode, void *stream) {
  const Scene *sc = reinterpret_cast<const Scene *>(scene);
  if (!sc || (n && (!rays || !hits))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays in one call");
  DeviceGuard g(scene_device(sc));
  return scene_trace(sc, rays, n, hits, inst_out, mode, reinterpret_cast<hipStream_t>(stream));
}

srf_status srf_scene_autotune(srf_scene *scene, const srf_ray *rays, uint64_t n, srf_hit *hits,
                              uint32_t *inst_out, void *stream, srf_launch_config *best) {
  Scene *sc = reinterpret_cast<Scene *>(scene);
  if (!sc || !rays || !hits || n == 0) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays in one call");
  DeviceGuard g(scene_device(sc));
  return scene_autotune(sc, rays, n, hits, inst_out, reinterpret_cast<hipStream_t>(stream),
                        best);
}

srf_status srf_trace_any_sorted(srf_ctx *ctx, const srf_accel *accel, const srf_ray *rays,
                                const float *t_max, uint64_t n, uint32_t key_bits, int mode,
                                uint32_t *occluded, void *stream) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!c || !a || (n && (!rays || !occluded))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (c->device != a->device) return set_error(SRF_ERR_INVALID_ARG, "context/device mismatch");
  if (n >= (1ull << 31)) return set_error(SRF_ERR_INVALID_ARG, "too many rays in one call");
  if (mode != SRF_LAUNCH_PLAIN && mode != SRF_LAUNCH_PERSISTENT)
    return set_error(SRF_ERR_INVALID_ARG, "mode must be SRF_LAUNCH_PLAIN or _PERSISTENT");
  DeviceGuard g(a->device);
  return trace_any_sorted(c, a, rays, t_max, n, key_bits, mode, occluded,
                          reinterpret_cast<hipStream_t>(stream));
}

srf_status srf_autotune_any(srf_ctx *ctx, const srf_accel *accel, const srf_ray *rays,
                            const float *t_max, uint64_t n, uint32_t *occluded, void *stream,
                            srf_any_config *best) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!c || !a || !rays || !occluded || n == 0)
    return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n >= (1ull << 31)) return set_error(SRF_ERR_INVALID_ARG, "too many rays in one call");
  DeviceGuard g(a->device);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipEvent_t e0, e1;
  SRF_HIP(hipEventCreate(&e0));
  hipError_t he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hip_error(he, "hipEventCreate");
  }
  srf_any_config win{0u, 0u, FLT_MAX};
  srf_status st = SRF_OK;
  for (uint32_t bits : {0u, 9u, 15u, 21u, 27u, 30u}) {
    for (int mode : {(int)SRF_LAUNCH_PLAIN, (int)SRF_LAUNCH_PERSISTENT}) {
      st = trace_any_sorted(c, a, rays, t_max, n, bits, mode, occluded, s);  // warm-up
      float t = FLT_MAX;
      for (int rep = 0; rep < 2 && st == SRF_OK; rep++) {
        he = hipEventRecord(e0, s);
        if (he == hipSuccess) st = trace_any_sorted(c, a, rays, t_max, n, bits, mode, occluded, s);
        if (he == hipSuccess && st == SRF_OK) he = hipEventRecord(e1, s);
        if (he == hipSuccess && st == SRF_OK) he = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (he == hipSuccess && st == SRF_OK) he = hipEventElapsedTime(&ms, e0, e1);
        if (he != hipSuccess && st == SRF_OK) st = hip_error(he, "autotune timing");
        t = ms < t ? ms : t;
      }
      if (st != SRF_OK) break;
      if (t < win.ms) win = srf_any_config{bits, mode == SRF_LAUNCH_PERSISTENT ? 1u : 0u, t};
    }
    if (st != SRF_OK) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != SRF_OK) return st;
  if (best) *best = win;
  return SRF_OK;
}

srf_status srf_ao_rays_spheres(const srf_ray *primary, const srf_hit *hits, uint64_t n,
                               const float *positions, const float *radius, float rconst,
                               uint32_t per_hit, float eps, float dist, uint64_t seed,
                               srf_ray *out, float *out_tmax, uint32_t *out_pixel,
                               uint32_t *out_count, void *stream) {
  if (!out_count || (n && (!primary || !hits || !positions || !out || !out_tmax || !out_pixel)))
    return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (per_hit == 0 || per_hit > 64) return set_error(SRF_ERR_INVALID_ARG, "per_hit in [1, 64]");
  if (n * (uint64_t)per_hit >= (1ull << 32))
    return set_error(SRF_ERR_INVALID_ARG, "too many secondary rays in one call");
  return ao_rays_spheres(primary, hits, n, positions, radius, rconst, per_hit, eps, dist, seed,
                         out, out_tmax, out_pixel, out_count,
                         reinterpret_cast<hipStream_t>(stream));
}

srf_status srf_sparse_volume_create(srf_ctx *ctx, const srf_volume_block *blocks_host,
                                    uint32_t num_blocks, const int global_dim[3],
                                    int components, int data_type, srf_sparse_volume **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !out || (num_blocks && !blocks_host))
    return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (num_blocks == 0) return set_error(SRF_ERR_INVALID_ARG, "no volume blocks");
  if (components < 1) return set_error(SRF_ERR_INVALID_ARG, "components must be >= 1");
  if (data_type != SRF_TEX_BYTE && data_type != SRF_TEX_FLOAT && data_type != SRF_TEX_DOUBLE)
    return set_error(SRF_ERR_UNSUPPORTED, "voxel data type %d (byte 1, float 8, double 16)",
                     data_type);
  DeviceGuard g(c->device);
  SparseVolume *sv = nullptr;
  srf_status st = sparse_create(c, blocks_host, num_blocks, global_dim, components, data_type, &sv);
  if (st != SRF_OK) return st;
  *out = reinterpret_cast<srf_sparse_volume *>(sv);
  return SRF_OK;
}

srf_status srf_sparse_volume_destroy(srf_sparse_volume *vol) {
  sparse_free(reinterpret_cast<SparseVolume *>(vol));
  return SRF_OK;
}

srf_status srf_sparse_volume_sample(const srf_sparse_volume *vol, const double *positions,
                                    uint64_t n, double *values, void *stream) {
  const SparseVolume *sv = reinterpret_cast<const SparseVolume *>(vol);
  if (!sv || (n && (!positions || !values))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many samples");
  return sparse_sample(sv, positions, nullptr, n, values, nullptr,
                       reinterpret_cast<hipStream_t>(stream));
}

srf_status srf_sparse_volume_fetch(const srf_sparse_volume *vol, const float *uvw, uint64_t n,
                                   float *rgba, void *stream) {
  const SparseVolume *sv = reinterpret_cast<const SparseVolume *>(vol);
  if (!sv || (n && (!uvw || !rgba))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many samples");
  return sparse_sample(sv, nullptr, uvw, n, nullptr, rgba, reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
