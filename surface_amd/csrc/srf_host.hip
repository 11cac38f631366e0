// srf_host.hip -- the host-memory side of the boundary: builds from host
// arrays and closest-hit traces of host rays, for callers that hold their
// geometry and rays in host memory as the reference's callers do
// (AccelBuilder hands Build32 the GL client arrays,
// gles/gles_accel_builder.cc:545-572; RaytraceEngine calls Traverse per ray
// from its OpenMP loop, gles/gles_raytrace_engine.cc:662-880).
//
// A host build copies the arrays into device memory that the accelerator
// owns (it keeps them for BuildIntersection, as the reference's accelerator
// keeps its Mesh pointer), then runs the same device build; the call returns
// once the build is complete, so the host arrays may be reused at once.
// A host trace copies the rays to a per-thread device staging buffer, runs
// the batch traversal (and BuildIntersection when asked) on a per-thread
// stream and copies the results back before returning. Calls from many host
// threads run concurrently (each thread has its own stream and staging,
// kept until the process ends).
#include <string.h>

#include <vector>

#include "srf_internal.h"

namespace srf {

namespace {

// The arrays of a descriptor: pointer and byte count of each (vertices /
// positions, faces / indices / segments, radius), by kind.
struct DescArrays {
  Kind kind;
  const void *p[3];
  size_t bytes[3];
  uint32_t nvps;
};

srf_status desc_arrays(Kind kind, const void *desc, DescArrays *out) {
  memset(out, 0, sizeof(*out));
  out->kind = kind;
  if (kind == kTri) {
    const srf_tri_desc *d = (const srf_tri_desc *)desc;
    out->p[0] = d->vertices;
    out->bytes[0] = 12 * (size_t)d->num_vertices;
    out->p[1] = d->faces;
    out->bytes[1] = 12 * (size_t)d->num_faces;
  } else if (kind == kTet) {
    const srf_tet_desc *d = (const srf_tet_desc *)desc;
    out->p[0] = d->vertices;
    out->bytes[0] = 12 * (size_t)d->num_vertices;
    out->p[1] = d->indices;
    out->bytes[1] = 16 * (size_t)d->num_tetras;
  } else if (kind == kSol) {
    const srf_solid_desc *d = (const srf_solid_desc *)desc;
    out->p[0] = d->vertices;
    out->bytes[0] = 12 * (size_t)d->num_vertices;
    out->p[1] = d->indices;
    out->bytes[1] = 4 * (size_t)d->vertices_per_solid * d->num_solids;
    out->nvps = d->vertices_per_solid;
  } else if (kind == kLin) {
    const srf_line_desc *d = (const srf_line_desc *)desc;
    out->p[0] = d->positions;
    out->bytes[0] = 12 * (size_t)d->num_vertices;
    out->p[1] = d->segments;
    out->bytes[1] = 8 * (size_t)d->num_lines;
    out->p[2] = d->radius;
    out->bytes[2] = d->radius ? 4 * (size_t)d->num_vertices : 0;
  } else {
    const srf_sphere_desc *d = (const srf_sphere_desc *)desc;
    out->p[0] = d->positions;
    out->bytes[0] = 12 * (size_t)d->num_particles;
    out->p[2] = d->radius;
    out->bytes[2] = d->radius ? 4 * (size_t)d->num_particles : 0;
  }
  if (!out->p[0]) return set_error(SRF_ERR_INVALID_ARG, "null vertex array");
  if (kind != kSph && !out->p[1]) return set_error(SRF_ERR_INVALID_ARG, "null index array");
  return SRF_OK;
}

// The descriptor with its arrays replaced by p (same kind). buf must hold
// the largest descriptor.
void retarget(Kind kind, const void *desc, void *const p[3], void *buf) {
  if (kind == kTri) {
    srf_tri_desc d = *(const srf_tri_desc *)desc;
    d.vertices = (const float *)p[0];
    d.faces = (const uint32_t *)p[1];
    memcpy(buf, &d, sizeof(d));
  } else if (kind == kTet) {
    srf_tet_desc d = *(const srf_tet_desc *)desc;
    d.vertices = (const float *)p[0];
    d.indices = (const uint32_t *)p[1];
    memcpy(buf, &d, sizeof(d));
  } else if (kind == kSol) {
    srf_solid_desc d = *(const srf_solid_desc *)desc;
    d.vertices = (const float *)p[0];
    d.indices = (const uint32_t *)p[1];
    memcpy(buf, &d, sizeof(d));
  } else if (kind == kLin) {
    srf_line_desc d = *(const srf_line_desc *)desc;
    d.positions = (const float *)p[0];
    d.segments = (const uint32_t *)p[1];
    d.radius = (const float *)p[2];
    memcpy(buf, &d, sizeof(d));
  } else {
    srf_sphere_desc d = *(const srf_sphere_desc *)desc;
    d.positions = (const float *)p[0];
    d.radius = (const float *)p[2];
    memcpy(buf, &d, sizeof(d));
  }
}

srf_status kind_of(srf_prim_kind k, Kind *out) {
  switch (k) {
    case SRF_PRIM_TRIANGLES: *out = kTri; return SRF_OK;
    case SRF_PRIM_TETRAHEDRONS: *out = kTet; return SRF_OK;
    case SRF_PRIM_POINTS: *out = kSph; return SRF_OK;
    case SRF_PRIM_LINES: *out = kLin; return SRF_OK;
    case SRF_PRIM_PYRAMIDS:
    case SRF_PRIM_PRISMS:
    case SRF_PRIM_HEXAHEDRONS: *out = kSol; return SRF_OK;
  }
  return set_error(SRF_ERR_INVALID_ARG, "unknown primitive kind 0x%x", (unsigned)k);
}

// Device copies of the arrays in `own` (grown when too small), filled from
// host memory on stream s.
srf_status stage_arrays(const DescArrays &da, void *own[3], size_t own_bytes[3], hipStream_t s) {
  for (int k = 0; k < 3; k++) {
    if (!da.p[k] || da.bytes[k] == 0) continue;
    if (own_bytes[k] < da.bytes[k]) {
      if (own[k]) SRF_HIP(hipFree(own[k]));
      own[k] = nullptr;
      own_bytes[k] = 0;
      SRF_HIP(hipMalloc(&own[k], da.bytes[k]));
      own_bytes[k] = da.bytes[k];
    }
    SRF_HIP(hipMemcpyAsync(own[k], da.p[k], da.bytes[k], hipMemcpyHostToDevice, s));
  }
  return SRF_OK;
}

struct DevGuard {
  int prev = -1;
  explicit DevGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DevGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// Per host thread and device: a stream and a device staging buffer.
struct HostStage {
  int device;
  hipStream_t stream;
  char *buf;
  size_t cap;
};
thread_local std::vector<HostStage> tl_stage;

srf_status host_stage(int device, size_t bytes, HostStage **out) {
  HostStage *hs = nullptr;
  for (HostStage &h : tl_stage)
    if (h.device == device) hs = &h;
  if (!hs) {
    HostStage h{device, nullptr, nullptr, 0};
    SRF_HIP(hipStreamCreateWithFlags(&h.stream, hipStreamNonBlocking));
    tl_stage.push_back(h);
    hs = &tl_stage.back();
  }
  if (hs->cap < bytes) {
    if (hs->buf) SRF_HIP(hipFree(hs->buf));
    hs->buf = nullptr;
    hs->cap = 0;
    const size_t sz = bytes + bytes / 2 + 4096;
    SRF_HIP(hipMalloc((void **)&hs->buf, sz));
    hs->cap = sz;
  }
  *out = hs;
  return SRF_OK;
}

size_t ext_bytes(Kind k) {
  return k == kTet ? sizeof(srf_tet_hit_ext)
                   : (k == kSol ? sizeof(srf_solid_hit_ext)
                                : (k == kLin ? sizeof(srf_line_hit_ext) : 0));
}

inline size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

void record_geometry(Accel *a, const void *desc) {
  DescArrays da;
  if (desc_arrays(a->kind, desc, &da) != SRF_OK) return;
  for (int k = 0; k < 3; k++) a->geom[k] = da.p[k];
}

void free_geometry(Accel *a) {
  for (int k = 0; k < 3; k++) {
    if (a->geom_own[k]) (void)hipFree(a->geom_own[k]);
    a->geom_own[k] = nullptr;
    a->geom_bytes[k] = 0;
  }
}

}  // namespace srf

using namespace srf;

extern "C" srf_status srf_accel_build_host(srf_ctx *ctx, srf_prim_kind kind, const void *desc,
                                           void *stream, srf_accel **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !desc || !out) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  Kind k;
  srf_status st = kind_of(kind, &k);
  if (st != SRF_OK) return st;
  DescArrays da;
  st = desc_arrays(k, desc, &da);
  if (st != SRF_OK) return st;
  DevGuard g(c->device);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  void *own[3] = {nullptr, nullptr, nullptr};
  size_t own_bytes[3] = {0, 0, 0};
  st = stage_arrays(da, own, own_bytes, s);
  alignas(8) unsigned char dd[64];
  if (st == SRF_OK) {
    retarget(k, desc, own, dd);
    if (k == kTri)
      st = srf_accel_build_triangles(ctx, (const srf_tri_desc *)dd, stream, out);
    else if (k == kTet)
      st = srf_accel_build_tetras(ctx, (const srf_tet_desc *)dd, stream, out);
    else if (k == kSol)
      st = srf_accel_build_solids(ctx, (const srf_solid_desc *)dd, stream, out);
    else if (k == kLin)
      st = srf_accel_build_lines(ctx, (const srf_line_desc *)dd, stream, out);
    else
      st = srf_accel_build_spheres(ctx, (const srf_sphere_desc *)dd, stream, out);
  }
  const hipError_t e = hipStreamSynchronize(s);
  if (st == SRF_OK && e != hipSuccess) {
    srf_accel_destroy(*out);
    *out = nullptr;
    st = hip_error(e, "hipStreamSynchronize");
  }
  if (st != SRF_OK) {
    for (int q = 0; q < 3; q++)
      if (own[q]) (void)hipFree(own[q]);
    return st;
  }
  Accel *a = reinterpret_cast<Accel *>(*out);
  for (int q = 0; q < 3; q++) {
    a->geom_own[q] = own[q];
    a->geom_bytes[q] = own_bytes[q];
  }
  return SRF_OK;
}

extern "C" srf_status srf_accel_rebuild_host(srf_ctx *ctx, srf_accel *accel, const void *desc,
                                             void *stream) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  Accel *a = reinterpret_cast<Accel *>(accel);
  if (!c || !a || !desc) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  DescArrays da;
  srf_status st = desc_arrays(a->kind, desc, &da);
  if (st != SRF_OK) return st;
  DevGuard g(a->device);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // the old geometry may still be read by the last build or trace
  SRF_HIP(hipEventSynchronize(a->built));
  st = stage_arrays(da, a->geom_own, a->geom_bytes, s);
  if (st != SRF_OK) return st;
  void *p[3] = {a->geom_own[0], a->geom_own[1], da.p[2] ? a->geom_own[2] : nullptr};
  alignas(8) unsigned char dd[64];
  retarget(a->kind, desc, p, dd);
  st = srf_accel_rebuild(ctx, accel, dd, stream);
  const hipError_t e = hipStreamSynchronize(s);
  if (st == SRF_OK && e != hipSuccess) st = hip_error(e, "hipStreamSynchronize");
  return st;
}

extern "C" srf_status srf_trace_closest_host(const srf_accel *accel, const srf_ray *rays_host,
                                             uint64_t n, srf_hit *hits_host, void *ext_host,
                                             srf_isect *isect_host) {
  const Accel *a = reinterpret_cast<const Accel *>(accel);
  if (!a || (n && (!rays_host || !hits_host)))
    return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays");
  if (isect_host && (a->kind == kSol || a->kind == kLin))
    return set_error(SRF_ERR_UNSUPPORTED,
                     "solids and lines return position and normal in their ext records");
  if (ext_host && ext_bytes(a->kind) == 0)
    return set_error(SRF_ERR_INVALID_ARG, "this kind has no ext records");
  if (n == 0) return SRF_OK;
  DevGuard g(a->device);
  const size_t eb = ext_bytes(a->kind);
  const size_t o_hits = up256(32 * n), o_ext = o_hits + up256(16 * n),
               o_isect = o_ext + up256((eb ? eb : 16) * n), total = o_isect + up256(80 * n);
  HostStage *hs = nullptr;
  srf_status st = host_stage(a->device, total, &hs);
  if (st != SRF_OK) return st;
  hipStream_t s = hs->stream;
  srf_ray *rays = (srf_ray *)hs->buf;
  srf_hit *hits = (srf_hit *)(hs->buf + o_hits);
  void *ext = eb ? (void *)(hs->buf + o_ext) : nullptr;
  srf_isect *isect = (srf_isect *)(hs->buf + o_isect);
  SRF_HIP(hipStreamWaitEvent(s, a->built, 0));
  SRF_HIP(hipMemcpyAsync(rays, rays_host, 32 * n, hipMemcpyHostToDevice, s));
  if (a->kind == kTri)
    st = trace_tri(a, rays, n, hits, s);
  else if (a->kind == kTet)
    st = trace_tet(a, rays, n, nullptr, hits, (srf_tet_hit_ext *)ext, s);
  else if (a->kind == kSol)
    st = trace_sol(a, rays, n, nullptr, hits, (srf_solid_hit_ext *)ext, s);
  else if (a->kind == kLin)
    st = trace_lin(a, rays, n, hits, (srf_line_hit_ext *)ext, s);
  else
    st = trace_sph(a, rays, nullptr, n, hits, s);
  if (st == SRF_OK && isect_host)
    st = launch_isect(a, rays, hits, (const srf_tet_hit_ext *)ext, (uint32_t)n, isect, s);
  if (st != SRF_OK) {
    (void)hipStreamSynchronize(s);
    return st;
  }
  SRF_HIP(hipMemcpyAsync(hits_host, hits, 16 * n, hipMemcpyDeviceToHost, s));
  if (ext_host) SRF_HIP(hipMemcpyAsync(ext_host, ext, eb * n, hipMemcpyDeviceToHost, s));
  if (isect_host) SRF_HIP(hipMemcpyAsync(isect_host, isect, 80 * n, hipMemcpyDeviceToHost, s));
  SRF_HIP(hipStreamSynchronize(s));
  return SRF_OK;
}
