// srf_isect.hip -- BuildIntersection for a batch of closest hits on gfx950:
// the Intersection fields a shader reads beyond t, u, v and prim_id.
//
//   triangles  render/render_accel_triangle.cc:2439-2507: f0..f2, position
//              = org + t dir, geometric = shading normal = the normalised
//              cross product of (p1 - p0) and (p2 - p0)
//   tetras     render/render_accel_tetra.cc:2124-2196: f0..f3, position, the
//              normal of the entry face face_table[subface_id]
//   spheres    render/render_accel_particle.cc:2968-3049: f0 = f1 = f2 =
//              prim_id, position, normal = vnormalize(position - centre),
//              tangent and binormal of the (u, v) parameterisation
//
// Arithmetic follows the reference's real3 operations in float with its
// double steps (real3::normalize takes the float sum of squares to double
// and scales by 1.0 / sqrt in double; the sphere's tangent and binormal are
// formed in double from M_PI): no FMA contraction (-ffp-contract=off),
// correctly rounded division and square root. The fields the reference
// leaves unset for a kind (tangent and binormal of triangles and tetras,
// f3 of triangles) and every field of a miss are 0.
#include <float.h>
#include <math.h>

#include "srf_internal.h"

namespace srf {
namespace {

struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 f3sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 f3cross(F3 l, F3 r) {
  return {l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x};
}
// real3::normalize (render/render_common.h:69-82)
__device__ __forceinline__ F3 f3normalize(F3 v) {
  const double len2 = (double)((v.x * v.x + v.y * v.y) + v.z * v.z);
  double len = 0.0;
  if (fabs(len2) > 1.0e-30) len = 1.0 / sqrt(len2);
  return {(float)((double)v.x * len), (float)((double)v.y * len), (float)((double)v.z * len)};
}
__device__ __forceinline__ F3 load3(const float *p, uint32_t i) {
  return {p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]};
}
// ray.origin() + isect.t * ray.direction() in float
__device__ __forceinline__ F3 hit_point(const srf_ray &r, float t) {
  return {r.org[0] + t * r.dir[0], r.org[1] + t * r.dir[1], r.org[2] + t * r.dir[2]};
}

__device__ __forceinline__ void put(srf_isect *o, F3 pos, F3 geo, F3 nrm, F3 tan, F3 bin,
                                    uint32_t f0, uint32_t f1, uint32_t f2, uint32_t f3) {
  float4 *w = reinterpret_cast<float4 *>(o);
  w[0] = make_float4(pos.x, pos.y, pos.z, __uint_as_float(f0));
  w[1] = make_float4(geo.x, geo.y, geo.z, __uint_as_float(f1));
  w[2] = make_float4(nrm.x, nrm.y, nrm.z, __uint_as_float(f2));
  w[3] = make_float4(tan.x, tan.y, tan.z, __uint_as_float(f3));
  w[4] = make_float4(bin.x, bin.y, bin.z, 0.0f);
}

__global__ void __launch_bounds__(256)
k_isect_tri(const float *__restrict__ V, const uint32_t *__restrict__ F, uint64_t nfaces,
            const srf_ray *__restrict__ rays, const srf_hit *__restrict__ hits, uint32_t n,
            srf_isect *__restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const F3 z = {0.0f, 0.0f, 0.0f};
  const srf_hit h = hits[i];
  if (h.prim_id == SRF_NO_PRIM || h.prim_id >= nfaces) {
    put(out + i, z, z, z, z, z, 0u, 0u, 0u, 0u);
    return;
  }
  const uint32_t f0 = F[3 * (size_t)h.prim_id], f1 = F[3 * (size_t)h.prim_id + 1],
                 f2 = F[3 * (size_t)h.prim_id + 2];
  const F3 p0 = load3(V, f0), p1 = load3(V, f1), p2 = load3(V, f2);
  const F3 nrm = f3normalize(f3cross(f3sub(p1, p0), f3sub(p2, p0)));
  put(out + i, hit_point(rays[i], h.t), nrm, nrm, z, z, f0, f1, f2, 0u);
}

__global__ void __launch_bounds__(256)
k_isect_tet(const float *__restrict__ V, const uint32_t *__restrict__ F, uint64_t ntets,
            const srf_ray *__restrict__ rays, const srf_hit *__restrict__ hits,
            const srf_tet_hit_ext *__restrict__ ext, uint32_t n, srf_isect *__restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const F3 z = {0.0f, 0.0f, 0.0f};
  const srf_hit h = hits[i];
  if (h.prim_id == SRF_NO_PRIM || h.prim_id >= ntets) {
    put(out + i, z, z, z, z, z, 0u, 0u, 0u, 0u);
    return;
  }
  const uint32_t *fp = F + 4 * (size_t)h.prim_id;
  const uint4 f = make_uint4(fp[0], fp[1], fp[2], fp[3]);
  const F3 p0 = load3(V, f.x), p1 = load3(V, f.y), p2 = load3(V, f.z), p3 = load3(V, f.w);
  // face_table of render/render_accel_tetra.cc:2178: {0,1,3} {1,2,3} {2,0,3} {0,1,2}
  const uint32_t sf = ext[i].subface & 3u;
  const F3 a = sf == 1 ? p1 : (sf == 2 ? p2 : p0);
  const F3 b = sf == 0 ? p1 : (sf == 1 ? p2 : (sf == 2 ? p0 : p1));
  const F3 c = sf == 3 ? p2 : p3;
  const F3 nrm = f3normalize(f3cross(f3sub(b, a), f3sub(c, a)));
  put(out + i, hit_point(rays[i], h.t), nrm, nrm, z, z, f.x, f.y, f.z, f.w);
}

__global__ void __launch_bounds__(256)
k_isect_sph(const float *__restrict__ P, uint64_t np, const srf_ray *__restrict__ rays,
            const srf_hit *__restrict__ hits, uint32_t n, srf_isect *__restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const F3 z = {0.0f, 0.0f, 0.0f};
  const srf_hit h = hits[i];
  if (h.prim_id == SRF_NO_PRIM || h.prim_id >= np) {
    put(out + i, z, z, z, z, z, 0u, 0u, 0u, 0u);
    return;
  }
  const F3 pos = hit_point(rays[i], h.t);
  const F3 cp = f3sub(pos, load3(P, h.prim_id));
  // vnormalize (render/render_accel_particle.cc:2041-2056)
  F3 nrm = cp;
  const float len = sqrtf((cp.x * cp.x + cp.y * cp.y) + cp.z * cp.z);
  if (len > (float)(FLT_EPSILON * 16.0)) {
    const float il = (float)(1.0 / (double)len);
    nrm = {cp.x * il, cp.y * il, cp.z * il};
  }
  F3 tan, bin;
  const float xz2 = nrm.x * nrm.x + nrm.z * nrm.z;
  if ((double)xz2 > 0.0) {
    const float xz = sqrtf(xz2);
    const float inv = (float)(1.0 / (double)xz);
    tan = {(float)(-2.0 * M_PI * (double)nrm.x), (float)(2.0 * M_PI * (double)nrm.z), 0.0f};
    bin = {(float)(((-M_PI * (double)nrm.z) * (double)inv) * (double)nrm.y),
           (float)(((-M_PI * (double)nrm.x) * (double)inv) * (double)nrm.y),
           (float)(M_PI * (double)xz)};
  } else {
    tan = {0.0f, 0.0f, 1.0f};
    bin = {nrm.y > 0.0f ? 1.0f : -1.0f, 0.0f, 0.0f};
  }
  put(out + i, pos, nrm, nrm, tan, bin, h.prim_id, h.prim_id, h.prim_id, 0u);
}

inline uint32_t blocks256(uint64_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace

srf_status launch_isect(const Accel *a, const srf_ray *rays, const srf_hit *hits,
                        const srf_tet_hit_ext *ext, uint32_t n, srf_isect *out, hipStream_t s) {
  if (n == 0) return SRF_OK;
  const float *V = (const float *)a->geom[0];
  const uint32_t *F = (const uint32_t *)a->geom[1];
  if (!V || (a->kind != kSph && !F))
    return set_error(SRF_ERR_INVALID_ARG, "the accelerator's geometry is not recorded");
  if (a->kind == kTri)
    hipLaunchKernelGGL(k_isect_tri, dim3(blocks256(n)), dim3(256), 0, s, V, F, a->n, rays, hits,
                       n, out);
  else if (a->kind == kTet)
    hipLaunchKernelGGL(k_isect_tet, dim3(blocks256(n)), dim3(256), 0, s, V, F, a->n, rays, hits,
                       ext, n, out);
  else if (a->kind == kSph)
    hipLaunchKernelGGL(k_isect_sph, dim3(blocks256(n)), dim3(256), 0, s, V, a->n, rays, hits, n,
                       out);
  else
    return set_error(SRF_ERR_UNSUPPORTED, "BuildIntersection covers triangles, tetras, spheres");
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

}  // namespace srf

using namespace srf;

extern "C" srf_status srf_build_intersection_triangles(const srf_tri_desc *mesh,
                                                       const srf_ray *rays, const srf_hit *hits,
                                                       uint64_t n, srf_isect *out, void *stream) {
  if (!mesh || (n && (!rays || !hits || !out || !mesh->vertices || !mesh->faces)))
    return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays");
  if (n == 0) return SRF_OK;
  hipLaunchKernelGGL(k_isect_tri, dim3(blocks256(n)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), mesh->vertices, mesh->faces,
                     mesh->num_faces, rays, hits, (uint32_t)n, out);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

extern "C" srf_status srf_build_intersection_tetras(const srf_tet_desc *tetras,
                                                    const srf_ray *rays, const srf_hit *hits,
                                                    const srf_tet_hit_ext *ext, uint64_t n,
                                                    srf_isect *out, void *stream) {
  if (!tetras || (n && (!rays || !hits || !ext || !out || !tetras->vertices || !tetras->indices)))
    return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays");
  if (n == 0) return SRF_OK;
  hipLaunchKernelGGL(k_isect_tet, dim3(blocks256(n)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), tetras->vertices, tetras->indices,
                     tetras->num_tetras, rays, hits, ext, (uint32_t)n, out);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

extern "C" srf_status srf_build_intersection_spheres(const srf_sphere_desc *particles,
                                                     const srf_ray *rays, const srf_hit *hits,
                                                     uint64_t n, srf_isect *out, void *stream) {
  if (!particles || (n && (!rays || !hits || !out || !particles->positions)))
    return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays");
  if (n == 0) return SRF_OK;
  hipLaunchKernelGGL(k_isect_sph, dim3(blocks256(n)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), particles->positions,
                     particles->num_particles, rays, hits, (uint32_t)n, out);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}
