// srf_volume.hip -- 3D-texture sampling and the differentiable
// emission-absorption march on gfx950.
//
// Sampling reproduces FilterTexture3DFloat (render/render_texture.cc:946-1058)
// bit for bit: node-centred voxels (vox = u * (dim - 1)), REPEAT wraps with
// fasterfloorf, CLAMP_TO_EDGE leaves u unclamped (extrapolating), lerp in x,
// then y, then z with Lerp(t, a, b) = (1 - t) a + t b.
//
// March (no reference counterpart; the reference marches in user GLSL):
//   entry/exit by a float slab test against [0,1]^3 (tnear clamped to 0);
//   sample k at t_k = tnear + ((float)k + 0.5f) * step while t_k < tfar;
//   s_k = texture sample at o + t_k d; TF index x = clamp(s, 0, 1) * (N - 1),
//   i0 = min((int)x, N - 2), f = x - i0, rgba_k = Lerp(f, tf[i0], tf[i0 + 1]);
//   front to back: w = T a_k; C += w rgb_k; T *= (1 - a_k); out = (C, 1 - T).
// Adjoint (L = sum grad . out), replayed backwards without per-sample
// state. The forward pass can leave each ray's transmittance at the start of
// every 16-sample segment and after its last sample (without them the
// adjoint first re-marches the ray once to make them). Each segment is then
// walked from its end to its start recovering T_k = T_{k+1} / (1 - a_k). The
// replay's a_k is the forward's bit for bit (same samples, same arithmetic),
// so the forward product T_{k+1} = fl(T_k fl(1 - a_k)) is undone with one
// rounding of the product and one of the reciprocal per sample: at most
// about 3 ulp per sample, 45 ulp over a 16-sample segment, whatever a_k is.
// Only an opaque sample (1 - a_k = 0) or a T_{k+1} below the normal range
// loses T_k; such a sample gets it by re-marching forward from the segment's
// checkpoint instead.
// With T_k in hand the division-free recursions give the gradients:
//   R_{k-1} = a_k c_k + (1 - a_k) R_k,  P_{k-1} = (1 - a_k) P_k,
//   dL/dc_k = g_C T_k a_k,  dL/da_k = T_k (g_C . (c_k - R_k) + g_A P_k).
// TF gradients collect in registers while a ray stays in one TF bin, then
// in LDS per block (four copies, spread over the lanes); voxel gradients go
// out as float atomics, merged across consecutive samples that share a cell.
#include <float.h>
#include <math.h>

#include "srf_internal.h"

namespace srf {

struct Volume {
  int device;
  const float *vox;  // FP32 single-component voxels (the march); else null
  int dims[3];
  int clamp[3];
  // Texture3D (render/render_texture.h:34-44): image, data_type (1 byte,
  // 8 float, 16 double) and interleaved components; Texture::Fetch's
  // per-axis remap tables (gles/gles_common.h:263-269,377-389), device
  // arrays or null
  const void *image;
  int format;
  int comps;
  const float *remap[3];
  int remap_n[3];
};

namespace {

constexpr int kVolBlock = 128;
constexpr int kVolBwdBlock = 256;
constexpr int kMaxTf = 1024;  // entries; forward keeps the table in LDS
constexpr int kFwdLdsEntries = 1024;  // forward LDS reserved per block (float4 entries)
constexpr int kSegShift = 4;    // log2 of the samples per checkpointed segment
constexpr int kSegLen = 1 << kSegShift;
constexpr int kMaxSegNoCkpt = 8;  // LDS checkpoints per ray when recomputing
constexpr int kMaxSamples = 65536;
// The adjoint's TF-gradient table is kept in this many copies, lane l adding
// to copy l % kTfCopies: neighbouring rays sit in the same TF bin at the same
// step, and LDS atomics to one address from several lanes of an instruction
// serialise (measured on the 256^3 bench: 1, 2, 4, 8 copies -> adjoint 5.9,
// 5.9, 5.5, 6.3 ms; 8 no longer fit the LDS beside 12^3 voxel windows). Copy
// strides are 4N + 1 floats, so equal entries of two copies fall in
// different banks.
constexpr int kTfCopies = 4;

__device__ __forceinline__ int faster_floorf(float x) {
  // render/render_texture.cc:108-121
  if (x >= 0.0f) return (int)x;
  int y = (int)x;
  if (!(fabsf(x - (float)y) <= FLT_EPSILON)) y = y - 1;
  return y;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  int a = v < lo ? lo : v;
  return hi < a ? hi : a;
}

__device__ __forceinline__ float lerpf(float t, float a, float b) {
  return (1.f - t) * a + t * b;
}

struct Geom {
  uint32_t a[8];  // 000 100 010 110 001 101 011 111
  float dx, dy, dz;
  int cx, cy, cz, cx1, cy1, cz1;  // corner coordinates
};

__device__ __forceinline__ void tex_geom(const Volume &V, float u, float v, float r, Geom &g) {
  float uu = V.clamp[0] ? u : u - (float)faster_floorf(u);
  float vv = V.clamp[1] ? v : v - (float)faster_floorf(v);
  float rr = V.clamp[2] ? r : r - (float)faster_floorf(r);
  float vox0 = uu * ((float)V.dims[0] - 1.0f);
  float vox1 = vv * ((float)V.dims[1] - 1.0f);
  float vox2 = rr * ((float)V.dims[2] - 1.0f);
  int vx = clampi((int)vox0, 0, V.dims[0] - 1);
  int vy = clampi((int)vox1, 0, V.dims[1] - 1);
  int vz = clampi((int)vox2, 0, V.dims[2] - 1);
  int vx1 = (vx + 1 < V.dims[0] - 1) ? vx + 1 : V.dims[0] - 1;
  int vy1 = (vy + 1 < V.dims[1] - 1) ? vy + 1 : V.dims[1] - 1;
  int vz1 = (vz + 1 < V.dims[2] - 1) ? vz + 1 : V.dims[2] - 1;
  g.dx = vox0 - (float)vx;
  g.dy = vox1 - (float)vy;
  g.dz = vox2 - (float)vz;
  g.cx = vx;
  g.cy = vy;
  g.cz = vz;
  g.cx1 = vx1;
  g.cy1 = vy1;
  g.cz1 = vz1;
  const uint32_t sy = (uint32_t)V.dims[0], sz = (uint32_t)V.dims[0] * (uint32_t)V.dims[1];
  const uint32_t z0 = (uint32_t)vz * sz, z1 = (uint32_t)vz1 * sz;
  const uint32_t y0 = (uint32_t)vy * sy, y1 = (uint32_t)vy1 * sy;
  g.a[0] = z0 + y0 + vx;
  g.a[1] = z0 + y0 + vx1;
  g.a[2] = z0 + y1 + vx;
  g.a[3] = z0 + y1 + vx1;
  g.a[4] = z1 + y0 + vx;
  g.a[5] = z1 + y0 + vx1;
  g.a[6] = z1 + y1 + vx;
  g.a[7] = z1 + y1 + vx1;
}

__device__ __forceinline__ float tex_eval(const float *__restrict__ vox, const Geom &g) {
  float d000 = vox[g.a[0]], d100 = vox[g.a[1]], d010 = vox[g.a[2]], d110 = vox[g.a[3]];
  float d001 = vox[g.a[4]], d101 = vox[g.a[5]], d011 = vox[g.a[6]], d111 = vox[g.a[7]];
  float d00 = lerpf(g.dx, d000, d100);
  float d10 = lerpf(g.dx, d010, d110);
  float d01 = lerpf(g.dx, d001, d101);
  float d11 = lerpf(g.dx, d011, d111);
  float d0 = lerpf(g.dy, d00, d10);
  float d1 = lerpf(g.dy, d01, d11);
  return lerpf(g.dz, d0, d1);
}

// A cell's 8 voxels as four x-pairs: each pair is one 8-byte load (two
// adjacent voxels of a row), so a sample issues four gathers instead of
// eight. At the far x border (vx = w - 1, where the texture clamps vx1 to
// vx) the pair starting at w - 2 is loaded and its second voxel used twice,
// so no load leaves the row. The values equal tex_eval's loads.
typedef float f2a4 __attribute__((ext_vector_type(2), aligned(4)));

__device__ __forceinline__ void gather8(const float *__restrict__ vox, int w, const Geom &g,
                                        float v[8]) {
  const int bx = g.cx < w - 2 ? g.cx : w - 2;
  const bool lo_hi = g.cx != bx;   // vx is the pair's second voxel
  const bool hi_hi = g.cx1 != bx;  // vx1 is the pair's second voxel
  for (int j = 0; j < 4; j++) {
    const f2a4 p = *reinterpret_cast<const f2a4 *>(vox + (g.a[2 * j] - (uint32_t)g.cx + bx));
    v[2 * j] = lo_hi ? p.y : p.x;
    v[2 * j + 1] = hi_hi ? p.y : p.x;
  }
}

// The 8 voxels of the last cell a lane sampled: consecutive samples of a ray
// (half-voxel steps) share a cell about every other step, and only lanes
// whose cell changed gather (the gathers' cost goes with the cache lines
// the active lanes touch).
struct CellCache {
  uint32_t a0;  // the cell's first corner (the others follow from it)
  float d[8];
};

__device__ __forceinline__ void cell_fetch(const float *__restrict__ vox, int w, const Geom &g,
                                           CellCache &c) {
  if (g.a[0] != c.a0) {
    gather8(vox, w, g, c.d);
    c.a0 = g.a[0];
  }
}

__device__ __forceinline__ float trilerp(const Geom &g, const float v[8]) {
  float d00 = lerpf(g.dx, v[0], v[1]);
  float d10 = lerpf(g.dx, v[2], v[3]);
  float d01 = lerpf(g.dx, v[4], v[5]);
  float d11 = lerpf(g.dx, v[6], v[7]);
  float d0 = lerpf(g.dy, d00, d10);
  float d1 = lerpf(g.dy, d01, d11);
  return lerpf(g.dz, d0, d1);
}

__device__ __forceinline__ bool vol_span(const float o[3], const float d[3], float &tn,
                                         float &tf) {
  tn = 0.0f;
  tf = FLT_MAX;
  for (int k = 0; k < 3; k++) {
    if (d[k] == 0.0f) {
      if (o[k] < 0.0f || o[k] > 1.0f) return false;
      continue;
    }
    float inv = 1.0f / d[k];
    float t0 = (0.0f - o[k]) * inv, t1 = (1.0f - o[k]) * inv;
    float lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
    if (lo > tn) tn = lo;
    if (hi < tf) tf = hi;
  }
  return tn < tf;
}

__device__ __forceinline__ void tf_index(float s, int N, int &i0, float &f, bool &inside) {
  float x = s * (float)(N - 1);
  inside = (x > 0.0f) && (x < (float)(N - 1));
  if (!(x > 0.0f)) x = 0.0f;
  if (x > (float)(N - 1)) x = (float)(N - 1);
  int i = (int)x;
  if (i > N - 2) i = N - 2;
  i0 = i;
  f = x - (float)i;
}

__device__ __forceinline__ void load_ray(const srf_ray *rays, uint32_t i, float o[3],
                                         float d[3]) {
  const float4 *r4 = reinterpret_cast<const float4 *>(rays + i);
  float4 a = r4[0], b = r4[1];
  o[0] = a.x;
  o[1] = a.y;
  o[2] = a.z;
  d[0] = b.x;
  d[1] = b.y;
  d[2] = b.z;
}

// local::remap (gles/gles_common.h:263-269)
__device__ __forceinline__ float remap_coord(float x, const float *__restrict__ table, int n) {
  int idx = (int)(x * (float)n);
  idx = idx < n - 1 ? idx : n - 1;
  idx = idx > 0 ? idx : 0;
  return table[idx];
}

__device__ __forceinline__ double dlerp(double t, double a, double b) {
  // render/render_texture.cc:32-34
  return (1.0 - t) * a + t * b;
}

// Texture::Fetch of a dense texture (gles/gles_common.h:344-403): remap,
// then FilterTexture3DByte (voxel / 255.5f, render/render_texture.cc:832-944),
// FilterTexture3DFloat (:946-1058) or FilterTexture3DDouble (:1060-1172) over
// V.comps interleaved components; one component is splatted to RGBA.
__global__ void __launch_bounds__(256) k_tex_fetch(Volume V, const float *__restrict__ uvw,
                                                  uint32_t n, float4 *__restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float c[3] = {uvw[3 * (size_t)i], uvw[3 * (size_t)i + 1], uvw[3 * (size_t)i + 2]};
  for (int k = 0; k < 3; k++)
    if (V.remap[k]) c[k] = remap_coord(c[k], V.remap[k], V.remap_n[k]);
  Geom g;
  tex_geom(V, c[0], c[1], c[2], g);
  float vals[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  const size_t cs = (size_t)V.comps;
  for (int q = 0; q < 4; q++) {
    if (q >= V.comps) break;
    if (V.format == 16) {
      const double *d = (const double *)V.image;
      double D[8];
      for (int k = 0; k < 8; k++) D[k] = d[(size_t)g.a[k] * cs + q];
      const double dx = g.dx, dy = g.dy, dz = g.dz;
      const double d00 = dlerp(dx, D[0], D[1]), d10 = dlerp(dx, D[2], D[3]);
      const double d01 = dlerp(dx, D[4], D[5]), d11 = dlerp(dx, D[6], D[7]);
      const double d0 = dlerp(dy, d00, d10), d1 = dlerp(dy, d01, d11);
      vals[q] = (float)dlerp(dz, d0, d1);
    } else {
      float D[8];
      if (V.format == 1) {
        const unsigned char *b = (const unsigned char *)V.image;
        for (int k = 0; k < 8; k++) D[k] = (float)b[(size_t)g.a[k] * cs + q] / 255.5f;
      } else {
        const float *f = (const float *)V.image;
        for (int k = 0; k < 8; k++) D[k] = f[(size_t)g.a[k] * cs + q];
      }
      const float d00 = lerpf(g.dx, D[0], D[1]), d10 = lerpf(g.dx, D[2], D[3]);
      const float d01 = lerpf(g.dx, D[4], D[5]), d11 = lerpf(g.dx, D[6], D[7]);
      const float d0 = lerpf(g.dy, d00, d10), d1 = lerpf(g.dy, d01, d11);
      vals[q] = lerpf(g.dz, d0, d1);
    }
  }
  out[i] = V.comps == 1 ? make_float4(vals[0], vals[0], vals[0], vals[0])
                        : make_float4(vals[0], vals[1], vals[2], vals[3]);
}

__global__ void __launch_bounds__(256) k_vol_sample(Volume V, const float *__restrict__ uvw,
                                                   uint32_t n, float *__restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Geom g;
  tex_geom(V, uvw[3 * (size_t)i], uvw[3 * (size_t)i + 1], uvw[3 * (size_t)i + 2], g);
  out[i] = tex_eval(V.vox, g);
}

// log2 of the segment length for a ray of ns samples so that its
// checkpoints (segment starts plus the final transmittance) fit in `slots`:
// 16 samples, doubled as needed (only rays longer than the unit-direction
// bound ever need that). Segments are powers of two so that a sample's
// segment is a shift and its place in it a mask.
__device__ __forceinline__ int seg_shift_for(int ns, uint32_t slots) {
  int sh = kSegShift;
  while ((uint32_t)(((ns + (1 << sh) - 1) >> sh) + 1) > slots) sh++;
  return sh;
}

// The number of samples k with t_k = tn + (k + 0.5) step < tfar (at most
// kMaxSamples). t_k never decreases with k (k + 0.5 is exact, and a product
// and a sum with a fixed operand round monotonically), so the samples are a
// prefix 0..ns-1: start from the quotient and step to the first k whose t_k
// fails the test, as the march itself evaluates it.
__device__ __forceinline__ int count_samples(float tn, float tfar, float step) {
  const float e = (tfar - tn) / step;
  int ns = e > 0.0f ? (int)fminf(e, (float)kMaxSamples) : 0;
  while (ns > 0 && !((tn + ((float)(ns - 1) + 0.5f) * step) < tfar)) ns--;
  while (ns < kMaxSamples && (tn + ((float)ns + 0.5f) * step) < tfar) ns++;
  return ns;
}

// Geometry of sample k of the ray (o, d) entering at tn.
__device__ __forceinline__ void sample_geom(const Volume &V, const float o[3], const float d[3],
                                            float tn, float step, int k, Geom &g) {
  const float t = tn + ((float)k + 0.5f) * step;
  tex_geom(V, o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2], g);
}

__global__ void __launch_bounds__(kVolBlock) k_vol_fwd(Volume V, const float *__restrict__ tf,
                                                      int N, float step,
                                                      const srf_ray *__restrict__ rays,
                                                      uint32_t n, float *__restrict__ rgba,
                                                      float *__restrict__ ckpt, uint32_t slots) {
  extern __shared__ float4 s_tf[];  // N entries
  for (int k = threadIdx.x; k < N; k += kVolBlock)
    s_tf[k] = reinterpret_cast<const float4 *>(tf)[k];
  __syncthreads();
  uint32_t i = xcd_block() * kVolBlock + threadIdx.x;
  if (i >= n) return;
  float o[3], d[3];
  load_ray(rays, i, o, d);
  float C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, T = 1.0f;
  float *ck = ckpt ? ckpt + (size_t)i * slots : nullptr;
  float tn = 0.0f, tfar = 0.0f;
  int ns = 0;
  if (vol_span(o, d, tn, tfar)) ns = count_samples(tn, tfar, step);
  const int sh = ck ? seg_shift_for(ns, slots) : kSegShift;
  const int mask = (1 << sh) - 1;
  CellCache vc;
  vc.a0 = 0xffffffffu;
  for (int k = 0; k < ns; k++) {
    if (ck && (k & mask) == 0) ck[k >> sh] = T;
    Geom g;
    sample_geom(V, o, d, tn, step, k, g);
    cell_fetch(V.vox, V.dims[0], g, vc);
    const float s = trilerp(g, vc.d);
    int i0;
    float f;
    bool inside;
    tf_index(s, N, i0, f, inside);
    float4 a = s_tf[i0], b = s_tf[i0 + 1];
    float cr = lerpf(f, a.x, b.x), cg = lerpf(f, a.y, b.y), cb = lerpf(f, a.z, b.z);
    float ca = lerpf(f, a.w, b.w);
    float w = T * ca;
    C0 = C0 + w * cr;
    C1 = C1 + w * cg;
    C2 = C2 + w * cb;
    T = T * (1.0f - ca);
  }
  // transmittance after the last sample closes the last segment
  if (ck) ck[(ns + (1 << sh) - 1) >> sh] = T;
  reinterpret_cast<float4 *>(rgba)[i] = make_float4(C0, C1, C2, 1.0f - T);
}

__device__ __forceinline__ float sample_alpha(const Volume &V, const float4 *s_tf, int N,
                                              const float o[3], const float d[3], float t) {
  Geom g;
  tex_geom(V, o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2], g);
  float s = tex_eval(V.vox, g);
  int i0;
  float f;
  bool inside;
  tf_index(s, N, i0, f, inside);
  return lerpf(f, s_tf[i0].w, s_tf[i0 + 1].w);
}

// floats of the adjoint's TF-gradient copies, rounded up to a float4
__host__ __device__ inline int tf_grad_floats(int N) { return (kTfCopies * (4 * N + 1) + 3) & ~3; }

struct TfAcc {
  int bin;
  float g[8];
};

__device__ __forceinline__ void flush_tf(TfAcc &acc, float *s_gtf) {
  if (acc.bin < 0) return;
  for (int c = 0; c < 8; c++)
    if (acc.g[c] != 0.0f) atomicAdd(&s_gtf[4 * acc.bin + c], acc.g[c]);
  acc.bin = -1;
}

// Moves the accumulated pair (bin, bin + 1) to (i0, i0 + 1): an entry both
// pairs share stays in registers, the others go to the block's table.
__device__ __forceinline__ void tf_move(TfAcc &acc, int i0, float *s_gtf) {
  if (acc.bin == i0) return;
  if (acc.bin >= 0 && i0 == acc.bin + 1) {
    for (int c = 0; c < 4; c++) {
      if (acc.g[c] != 0.0f) atomicAdd(&s_gtf[4 * acc.bin + c], acc.g[c]);
      acc.g[c] = acc.g[4 + c];
      acc.g[4 + c] = 0.0f;
    }
  } else if (acc.bin >= 0 && i0 == acc.bin - 1) {
    for (int c = 0; c < 4; c++) {
      if (acc.g[4 + c] != 0.0f) atomicAdd(&s_gtf[4 * (acc.bin + 1) + c], acc.g[4 + c]);
      acc.g[4 + c] = acc.g[c];
      acc.g[c] = 0.0f;
    }
  } else {
    flush_tf(acc, s_gtf);
    for (int c = 0; c < 8; c++) acc.g[c] = 0.0f;
  }
  acc.bin = i0;
}

__device__ __forceinline__ int wave_max_i(int v) {
  for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
  for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return v;
}

// Voxel-gradient window of one wavefront: a W^3 block of voxels in LDS
// anchored at (bx, by, bz). Lanes whose 8 corners lie inside add there
// (LDS float atomics); the window goes out to the global gradient (one
// float atomic per touched voxel) when the wavefront's samples leave it.
template <int W>
struct VoxWindow {
  float *cell;  // this wavefront's W^3 floats
  int bx, by, bz;
  bool set;
};

template <int W>
__device__ __forceinline__ void window_flush(VoxWindow<W> &w, const Volume &V, float *gv,
                                             int lane) {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (w.set) {
    for (int j = lane; j < W * W * W; j += 64) {
      const float v = w.cell[j];
      if (v != 0.0f) {
        const int x = w.bx + j % W, y = w.by + (j / W) % W, z = w.bz + j / (W * W);
        // only voxels of the volume ever receive a contribution
        unsafeAtomicAdd(gv + ((size_t)z * V.dims[1] + y) * V.dims[0] + x, v);
        w.cell[j] = 0.0f;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  w.set = false;
}

// One corner's accumulated gradient to the wavefront's window (LDS) when it
// lies inside, else straight to the global gradient.
template <int W>
__device__ __forceinline__ void corner_out(const VoxWindow<W> &w, const Volume &V, float *gv,
                                           int x, int y, int z, float v) {
  if (v == 0.0f) return;
  const int rx = x - w.bx, ry = y - w.by, rz = z - w.bz;
  if (w.set && (unsigned)rx < (unsigned)W && (unsigned)ry < (unsigned)W &&
      (unsigned)rz < (unsigned)W)
    atomicAdd(w.cell + (rz * W + ry) * W + rx, v);
  else
    unsafeAtomicAdd(gv + ((size_t)z * V.dims[1] + y) * V.dims[0] + x, v);
}

// A lane's gradient for the 8 corners of its current cell, kept in
// registers while consecutive samples stay in (or next to) that cell.
// g[(z << 2) | (y << 1) | x]; corner coordinates x0/x1 etc. (equal at the
// volume's far border, where the texture clamps).
struct CellAcc {
  int x0, x1, y0, y1, z0, z1;
  float g[8];
  bool live;
};

template <int W>
__device__ __forceinline__ void cell_flush(CellAcc &a, const VoxWindow<W> &w, const Volume &V,
                                           float *gv) {
  if (!a.live) return;
  corner_out(w, V, gv, a.x0, a.y0, a.z0, a.g[0]);
  corner_out(w, V, gv, a.x1, a.y0, a.z0, a.g[1]);
  corner_out(w, V, gv, a.x0, a.y1, a.z0, a.g[2]);
  corner_out(w, V, gv, a.x1, a.y1, a.z0, a.g[3]);
  corner_out(w, V, gv, a.x0, a.y0, a.z1, a.g[4]);
  corner_out(w, V, gv, a.x1, a.y0, a.z1, a.g[5]);
  corner_out(w, V, gv, a.x0, a.y1, a.z1, a.g[6]);
  corner_out(w, V, gv, a.x1, a.y1, a.z1, a.g[7]);
  a.live = false;
}

// Relation of an axis' corner pair between two cells: 0 same, +1 the new
// low corner is the old high one, -1 the new high corner is the old low
// one, 2 unrelated.
__device__ __forceinline__ int axis_step(int o0, int o1, int n0, int n1) {
  if (n0 == o0 && n1 == o1) return 0;
  if (n0 == o1) return 1;
  if (n1 == o0) return -1;
  return 2;
}

// Moves the accumulator to the cell (nx0..nz1): corners both cells share
// stay in registers (one axis at a time), the others are written out.
template <int W>
__device__ __forceinline__ void cell_move(CellAcc &a, int nx0, int nx1, int ny0, int ny1,
                                          int nz0, int nz1, const VoxWindow<W> &w,
                                          const Volume &V, float *gv) {
  if (a.live) {
    const int sx = axis_step(a.x0, a.x1, nx0, nx1);
    const int sy = axis_step(a.y0, a.y1, ny0, ny1);
    const int sz = axis_step(a.z0, a.z1, nz0, nz1);
    if (sx == 2 || sy == 2 || sz == 2) {
      cell_flush(a, w, V, gv);
    } else {
      if (sx == 1) {  // keep the x1 face as x0
        corner_out(w, V, gv, a.x0, a.y0, a.z0, a.g[0]);
        corner_out(w, V, gv, a.x0, a.y1, a.z0, a.g[2]);
        corner_out(w, V, gv, a.x0, a.y0, a.z1, a.g[4]);
        corner_out(w, V, gv, a.x0, a.y1, a.z1, a.g[6]);
        a.g[0] = a.g[1]; a.g[2] = a.g[3]; a.g[4] = a.g[5]; a.g[6] = a.g[7];
        a.g[1] = a.g[3] = a.g[5] = a.g[7] = 0.0f;
      } else if (sx == -1) {  // keep the x0 face as x1
        corner_out(w, V, gv, a.x1, a.y0, a.z0, a.g[1]);
        corner_out(w, V, gv, a.x1, a.y1, a.z0, a.g[3]);
        corner_out(w, V, gv, a.x1, a.y0, a.z1, a.g[5]);
        corner_out(w, V, gv, a.x1, a.y1, a.z1, a.g[7]);
        a.g[1] = a.g[0]; a.g[3] = a.g[2]; a.g[5] = a.g[4]; a.g[7] = a.g[6];
        a.g[0] = a.g[2] = a.g[4] = a.g[6] = 0.0f;
      }
      a.x0 = nx0;
      a.x1 = nx1;
      if (sy == 1) {
        corner_out(w, V, gv, a.x0, a.y0, a.z0, a.g[0]);
        corner_out(w, V, gv, a.x1, a.y0, a.z0, a.g[1]);
        corner_out(w, V, gv, a.x0, a.y0, a.z1, a.g[4]);
        corner_out(w, V, gv, a.x1, a.y0, a.z1, a.g[5]);
        a.g[0] = a.g[2]; a.g[1] = a.g[3]; a.g[4] = a.g[6]; a.g[5] = a.g[7];
        a.g[2] = a.g[3] = a.g[6] = a.g[7] = 0.0f;
      } else if (sy == -1) {
        corner_out(w, V, gv, a.x0, a.y1, a.z0, a.g[2]);
        corner_out(w, V, gv, a.x1, a.y1, a.z0, a.g[3]);
        corner_out(w, V, gv, a.x0, a.y1, a.z1, a.g[6]);
        corner_out(w, V, gv, a.x1, a.y1, a.z1, a.g[7]);
        a.g[2] = a.g[0]; a.g[3] = a.g[1]; a.g[6] = a.g[4]; a.g[7] = a.g[5];
        a.g[0] = a.g[1] = a.g[4] = a.g[5] = 0.0f;
      }
      a.y0 = ny0;
      a.y1 = ny1;
      if (sz == 1) {
        corner_out(w, V, gv, a.x0, a.y0, a.z0, a.g[0]);
        corner_out(w, V, gv, a.x1, a.y0, a.z0, a.g[1]);
        corner_out(w, V, gv, a.x0, a.y1, a.z0, a.g[2]);
        corner_out(w, V, gv, a.x1, a.y1, a.z0, a.g[3]);
        a.g[0] = a.g[4]; a.g[1] = a.g[5]; a.g[2] = a.g[6]; a.g[3] = a.g[7];
        a.g[4] = a.g[5] = a.g[6] = a.g[7] = 0.0f;
      } else if (sz == -1) {
        corner_out(w, V, gv, a.x0, a.y0, a.z1, a.g[4]);
        corner_out(w, V, gv, a.x1, a.y0, a.z1, a.g[5]);
        corner_out(w, V, gv, a.x0, a.y1, a.z1, a.g[6]);
        corner_out(w, V, gv, a.x1, a.y1, a.z1, a.g[7]);
        a.g[4] = a.g[0]; a.g[5] = a.g[1]; a.g[6] = a.g[2]; a.g[7] = a.g[3];
        a.g[0] = a.g[1] = a.g[2] = a.g[3] = 0.0f;
      }
      a.z0 = nz0;
      a.z1 = nz1;
      return;
    }
  }
  a.x0 = nx0; a.x1 = nx1; a.y0 = ny0; a.y1 = ny1; a.z0 = nz0; a.z1 = nz1;
  for (int c = 0; c < 8; c++) a.g[c] = 0.0f;
  a.live = true;
}

// Per-ray constants of the adjoint's replay.
struct BwdRay {
  float o[3], d[3], tn;
  int ns, sh, mask, ck_stride;
  const float *ck;
  float4 gr;
};

// Per-lane running state of the replay: the transmittance checkpoints in
// use, the suffix sums R (colour behind the sample) and P (transmittance
// behind it), and the TF and voxel gradient accumulators.
struct BwdLane {
  float Tnext, Tstart, Tpre, R0, R1, R2, P;
  TfAcc tacc;
  CellAcc cacc;
};

// Sample k of the adjoint's replay (a lane idles while k >= ns).
// (d0, d1, d2) is the wavefront's direction.
template <int W>
__device__ __forceinline__ void bwd_sample(const Volume &V, const float4 *s_tf, int N,
                                           float step, int k, const BwdRay &R, BwdLane &L,
                                           CellCache &vc, float *my_gtf, VoxWindow<W> &win,
                                           float *grad_vox, int lane, float d0, float d1,
                                           float d2) {
  const bool act = k < R.ns;
  if (act && k < R.ns - 1 && (k & R.mask) == R.mask) {  // top sample of a lower segment
    L.Tstart = L.Tpre;
    const int sg = k >> R.sh;
    if (sg > 0) L.Tpre = R.ck[(sg - 1) * R.ck_stride];
  }
  float ds = 0.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f, fx = 0.0f, fy = 0.0f, fz = 0.0f;
  int cx = 0, cy = 0, cz = 0, cx1 = 0, cy1 = 0, cz1 = 0;
  if (act) {
    Geom g;
    sample_geom(V, R.o, R.d, R.tn, step, k, g);
    cell_fetch(V.vox, V.dims[0], g, vc);
    const float s = trilerp(g, vc.d);
    int i0;
    float f;
    bool inside;
    tf_index(s, N, i0, f, inside);
    const float4 ta = s_tf[i0], tb = s_tf[i0 + 1];
    const float c0r = lerpf(f, ta.x, tb.x), c1g = lerpf(f, ta.y, tb.y),
                c2b = lerpf(f, ta.z, tb.z);
    const float a = lerpf(f, ta.w, tb.w);
    // transmittance in front of sample k
    float Tk;
    if ((k & R.mask) == 0) {
      Tk = L.Tstart;
    } else if (1.0f - a > 0.0f && L.Tnext >= FLT_MIN) {
      Tk = L.Tnext * __builtin_amdgcn_rcpf(1.0f - a);
    } else {
      Tk = L.Tstart;
      for (int j = k & ~R.mask; j < k; j++)
        Tk = Tk * (1.0f - sample_alpha(V, s_tf, N, R.o, R.d, R.tn + ((float)j + 0.5f) * step));
    }
    L.Tnext = Tk;
    const float wa = Tk * a;
    const float g0 = R.gr.x * wa, g1 = R.gr.y * wa, g2 = R.gr.z * wa;
    const float ga =
        Tk * __builtin_fmaf(R.gr.x, c0r - L.R0,
                            __builtin_fmaf(R.gr.y, c1g - L.R1,
                                           __builtin_fmaf(R.gr.z, c2b - L.R2, R.gr.w * L.P)));
    const float fo = 1.0f - f;
    tf_move(L.tacc, i0, my_gtf);
    L.tacc.g[0] = __builtin_fmaf(fo, g0, L.tacc.g[0]);
    L.tacc.g[1] = __builtin_fmaf(fo, g1, L.tacc.g[1]);
    L.tacc.g[2] = __builtin_fmaf(fo, g2, L.tacc.g[2]);
    L.tacc.g[3] = __builtin_fmaf(fo, ga, L.tacc.g[3]);
    L.tacc.g[4] = __builtin_fmaf(f, g0, L.tacc.g[4]);
    L.tacc.g[5] = __builtin_fmaf(f, g1, L.tacc.g[5]);
    L.tacc.g[6] = __builtin_fmaf(f, g2, L.tacc.g[6]);
    L.tacc.g[7] = __builtin_fmaf(f, ga, L.tacc.g[7]);
    if (inside)
      ds = __builtin_fmaf(g0, tb.x - ta.x,
                          __builtin_fmaf(g1, tb.y - ta.y,
                                         __builtin_fmaf(g2, tb.z - ta.z, ga * (tb.w - ta.w)))) *
           (float)(N - 1);
    // the 8 corners: (cx|cx1, cy|cy1, cz|cz1), trilinear weights
    cx = g.cx;
    cy = g.cy;
    cz = g.cz;
    cx1 = g.cx1;
    cy1 = g.cy1;
    cz1 = g.cz1;
    fx = g.dx;
    fy = g.dy;
    fz = g.dz;
    ux = 1.0f - g.dx;
    uy = 1.0f - g.dy;
    uz = 1.0f - g.dz;
    L.R0 = __builtin_fmaf(a, c0r - L.R0, L.R0);
    L.R1 = __builtin_fmaf(a, c1g - L.R1, L.R1);
    L.R2 = __builtin_fmaf(a, c2b - L.R2, L.R2);
    L.P = (1.0f - a) * L.P;
  }
  const bool want = act && ds != 0.0f;
  const bool inwin = want && win.set && cx >= win.bx && cy >= win.by && cz >= win.bz &&
               cx1 < win.bx + W && cy1 < win.by + W && cz1 < win.bz + W;
  const uint64_t out = __ballot(want && !inwin);
  if (__popcll(out) > 8 || (!win.set && __ballot(want) != 0ull)) {
    // re-anchor: flush, then place the window over this step's corners,
    // leaving room in the direction the wavefront moves next
    window_flush<W>(win, V, grad_vox, lane);
    const int big = 1 << 30;
    const int lox = wave_min_i(want ? cx : big), hix = wave_max_i(want ? cx1 : -big);
    const int loy = wave_min_i(want ? cy : big), hiy = wave_max_i(want ? cy1 : -big);
    const int loz = wave_min_i(want ? cz : big), hiz = wave_max_i(want ? cz1 : -big);
    win.bx = (d0 > 0.0f && hix - lox < W) ? hix - (W - 1) : lox;
    win.by = (d1 > 0.0f && hiy - loy < W) ? hiy - (W - 1) : loy;
    win.bz = (d2 > 0.0f && hiz - loz < W) ? hiz - (W - 1) : loz;
    win.set = true;
  }
  if (want) {
    cell_move(L.cacc, cx, cx1, cy, cy1, cz, cz1, win, V, grad_vox);
    const float wz0 = ds * uz, wz1 = ds * fz;
    const float w00 = wz0 * uy, w10 = wz0 * fy, w01 = wz1 * uy, w11 = wz1 * fy;
    L.cacc.g[0] = __builtin_fmaf(w00, ux, L.cacc.g[0]);
    L.cacc.g[1] = __builtin_fmaf(w00, fx, L.cacc.g[1]);
    L.cacc.g[2] = __builtin_fmaf(w10, ux, L.cacc.g[2]);
    L.cacc.g[3] = __builtin_fmaf(w10, fx, L.cacc.g[3]);
    L.cacc.g[4] = __builtin_fmaf(w01, ux, L.cacc.g[4]);
    L.cacc.g[5] = __builtin_fmaf(w01, fx, L.cacc.g[5]);
    L.cacc.g[6] = __builtin_fmaf(w11, ux, L.cacc.g[6]);
    L.cacc.g[7] = __builtin_fmaf(w11, fx, L.cacc.g[7]);
  }
}

// Dynamic LDS: tf[N] (float4), tf gradient [4N], one voxel window per
// wavefront, and, when the forward pass left no checkpoints,
// [kMaxSegNoCkpt + 1][block] recomputed ones.
//
// The wavefront walks sample indices k from its longest ray's last sample
// down to 0 in lockstep (a lane idles while k is past its own ray), so the
// lanes of an 8x8 ray tile sample neighbouring voxels at the same time.
// The LDS of a block (about 48 KB with 12^3 windows and a 256-entry table)
// leaves room for three blocks per CU, three waves per SIMD, so the kernel
// may use the registers of three.
template <bool kHaveCkpt, int W>
__global__ void __launch_bounds__(kVolBwdBlock) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_vol_bwd(Volume V, const float *__restrict__ tf, int N, float step,
          const srf_ray *__restrict__ rays, uint32_t n, const float *__restrict__ ckpt,
          uint32_t slots, const float *__restrict__ grad_rgba, float *__restrict__ grad_vox,
          float *__restrict__ grad_tf) {
  extern __shared__ float4 smem4[];
  float4 *s_tf = smem4;
  float *s_gtf = reinterpret_cast<float *>(smem4 + N);
  const int gstride = 4 * N + 1;
  float *s_win = s_gtf + tf_grad_floats(N);
  float *s_ck = s_win + (kVolBwdBlock / 64) * W * W * W;
  for (int k = threadIdx.x; k < N; k += kVolBwdBlock)
    s_tf[k] = reinterpret_cast<const float4 *>(tf)[k];
  for (int k = threadIdx.x; k < kTfCopies * gstride; k += kVolBwdBlock) s_gtf[k] = 0.0f;
  for (int k = threadIdx.x; k < (kVolBwdBlock / 64) * W * W * W; k += kVolBwdBlock)
    s_win[k] = 0.0f;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  float *my_gtf = s_gtf + (lane % kTfCopies) * gstride;
  VoxWindow<W> win;
  win.cell = s_win + (threadIdx.x >> 6) * W * W * W;
  win.set = false;
  win.bx = win.by = win.bz = 0;

  const uint32_t i = xcd_block() * kVolBwdBlock + threadIdx.x;
  float o[3] = {0.0f, 0.0f, 0.0f}, d[3] = {0.0f, 0.0f, 1.0f}, tn = 0.0f, tfar = 0.0f;
  bool live = i < n;
  if (live) {
    load_ray(rays, i, o, d);
    live = vol_span(o, d, tn, tfar);
    if (!live) tn = 0.0f;
  }
  int ns = 0, sh = kSegShift;
  const float *ck = nullptr;
  int ck_stride = 1;
  float4 gr = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (live) {
    gr = reinterpret_cast<const float4 *>(grad_rgba)[i];
    ns = count_samples(tn, tfar, step);  // the forward march's sample count
    sh = seg_shift_for(ns, kHaveCkpt ? slots : (uint32_t)(kMaxSegNoCkpt + 1));
    if (kHaveCkpt) {
      ck = ckpt + (size_t)i * slots;
    } else {
      // one forward replay leaves the segment checkpoints in LDS
      float T = 1.0f;
      float *w = s_ck + threadIdx.x;
      for (int k = 0; k < ns; k++) {
        if ((k & ((1 << sh) - 1)) == 0) w[(k >> sh) * kVolBwdBlock] = T;
        T = T * (1.0f - sample_alpha(V, s_tf, N, o, d, tn + ((float)k + 0.5f) * step));
      }
      w[((ns + (1 << sh) - 1) >> sh) * kVolBwdBlock] = T;
      ck = w;
      ck_stride = kVolBwdBlock;
    }
  }
  const int mask = (1 << sh) - 1;
  // Transmittance after the top sample (Tnext) and at the start (Tstart) of
  // the ray's last segment, and the start of the segment below (Tpre, always
  // requested one segment ahead). Walking down, the T after a segment's top
  // sample is the start of the segment above, which the recursion has just
  // reached, so each segment costs one checkpoint load.
  float Tnext = 1.0f, Tstart = 1.0f, Tpre = 1.0f;
  if (ns > 0) {
    const int sg = (ns - 1) >> sh;
    Tnext = ck[(sg + 1) * ck_stride];
    Tstart = ck[sg * ck_stride];
    if (sg > 0) Tpre = ck[(sg - 1) * ck_stride];
  }
  // the wavefront's travel direction when walking backwards (for anchoring)
  const float d0 = __shfl(d[0], 0, 64), d1 = __shfl(d[1], 0, 64), d2 = __shfl(d[2], 0, 64);
  const int kmax = wave_max_i(ns);
  BwdRay R;
  for (int c = 0; c < 3; c++) {
    R.o[c] = o[c];
    R.d[c] = d[c];
  }
  R.tn = tn;
  R.ns = ns;
  R.sh = sh;
  R.mask = mask;
  R.ck = ck;
  R.ck_stride = ck_stride;
  R.gr = gr;
  BwdLane L;
  L.Tnext = Tnext;
  L.Tstart = Tstart;
  L.Tpre = Tpre;
  L.R0 = L.R1 = L.R2 = 0.0f;
  L.P = 1.0f;
  L.tacc.bin = -1;
  L.cacc.live = false;
  L.cacc.x0 = L.cacc.x1 = L.cacc.y0 = L.cacc.y1 = L.cacc.z0 = L.cacc.z1 = 0;
  for (int c = 0; c < 8; c++) L.cacc.g[c] = 0.0f;
  CellCache vc;
  vc.a0 = 0xffffffffu;
  for (int k = kmax - 1; k >= 0; k--)
    bwd_sample<W>(V, s_tf, N, step, k, R, L, vc, my_gtf, win, grad_vox, lane, d0, d1, d2);
  cell_flush(L.cacc, win, V, grad_vox);
  window_flush<W>(win, V, grad_vox, lane);
  flush_tf(L.tacc, my_gtf);
  __syncthreads();
  for (int k = threadIdx.x; k < 4 * N; k += kVolBwdBlock) {
    float v = 0.0f;
    for (int r = 0; r < kTfCopies; r++) v += s_gtf[r * gstride + k];
    if (v != 0.0f) unsafeAtomicAdd(grad_tf + k, v);
  }
}

// blocks for n rays, padded to a multiple of kXcds (see xcd_block)
inline uint32_t vblocks(uint64_t n, int b) {
  const uint64_t g = (n + b - 1) / b;
  return (uint32_t)((g + kXcdPad - 1) / kXcdPad * kXcdPad);
}

}  // namespace
}  // namespace srf

using namespace srf;

extern "C" {

srf_status srf_volume_create(srf_ctx *ctx, const float *voxels, int w, int h, int d,
                             int wrap_s, int wrap_t, int wrap_r, srf_volume **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !voxels || !out) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (w < 2 || h < 2 || d < 2)
    return set_error(SRF_ERR_INVALID_ARG, "volume dims must be >= 2 (got %d %d %d)", w, h, d);
  if ((uint64_t)w * h * d >= (1ull << 32))
    return set_error(SRF_ERR_UNSUPPORTED, "volume exceeds 2^32 voxels");
  for (int k : {wrap_s, wrap_t, wrap_r})
    if (k != 0 && k != 1) return set_error(SRF_ERR_INVALID_ARG, "wrap must be 0 or 1");
  Volume *v = new Volume();
  v->device = c->device;
  v->vox = voxels;
  v->image = voxels;
  v->format = 8;
  v->comps = 1;
  v->dims[0] = w;
  v->dims[1] = h;
  v->dims[2] = d;
  v->clamp[0] = wrap_s;
  v->clamp[1] = wrap_t;
  v->clamp[2] = wrap_r;
  *out = reinterpret_cast<srf_volume *>(v);
  return SRF_OK;
}

srf_status srf_texture3d_create(srf_ctx *ctx, const void *image, int data_type, int components,
                                int w, int h, int d, int wrap_s, int wrap_t, int wrap_r,
                                srf_volume **out) {
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  if (!c || !image || !out) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (data_type != SRF_TEX_BYTE && data_type != SRF_TEX_FLOAT && data_type != SRF_TEX_DOUBLE)
    return set_error(SRF_ERR_UNSUPPORTED, "texture data type %d (byte 1, float 8, double 16)",
                     data_type);
  if (components < 1 || components > 4)
    return set_error(SRF_ERR_INVALID_ARG, "components must be 1..4 (got %d)", components);
  if (w < 1 || h < 1 || d < 1)
    return set_error(SRF_ERR_INVALID_ARG, "texture dims must be >= 1 (got %d %d %d)", w, h, d);
  if ((uint64_t)w * h * d >= (1ull << 32))
    return set_error(SRF_ERR_UNSUPPORTED, "texture exceeds 2^32 texels");
  for (int k : {wrap_s, wrap_t, wrap_r})
    if (k != 0 && k != 1) return set_error(SRF_ERR_INVALID_ARG, "wrap must be 0 or 1");
  Volume *v = new Volume();
  v->device = c->device;
  v->image = image;
  v->format = data_type;
  v->comps = components;
  // the march reads FP32 single-component volumes of at least 2 voxels a side
  v->vox = (data_type == SRF_TEX_FLOAT && components == 1 && w >= 2 && h >= 2 && d >= 2)
               ? (const float *)image : nullptr;
  v->dims[0] = w;
  v->dims[1] = h;
  v->dims[2] = d;
  v->clamp[0] = wrap_s;
  v->clamp[1] = wrap_t;
  v->clamp[2] = wrap_r;
  *out = reinterpret_cast<srf_volume *>(v);
  return SRF_OK;
}

srf_status srf_texture3d_remap(srf_volume *tex, int axis, const float *table, int n) {
  Volume *v = reinterpret_cast<Volume *>(tex);
  if (!v) return set_error(SRF_ERR_INVALID_ARG, "null texture");
  if (axis < 0 || axis > 2) return set_error(SRF_ERR_INVALID_ARG, "axis must be 0, 1 or 2");
  if (table && n < 1) return set_error(SRF_ERR_INVALID_ARG, "remap table of %d entries", n);
  v->remap[axis] = table;
  v->remap_n[axis] = table ? n : 0;
  return SRF_OK;
}

srf_status srf_texture3d_fetch(const srf_volume *tex, const float *uvw, uint64_t n, float *rgba,
                               void *stream) {
  const Volume *v = reinterpret_cast<const Volume *>(tex);
  if (!v || (n && (!uvw || !rgba))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (n == 0) return SRF_OK;
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many samples");
  hipLaunchKernelGGL(k_tex_fetch, dim3(vblocks(n, 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), *v, uvw, (uint32_t)n,
                     reinterpret_cast<float4 *>(rgba));
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

srf_status srf_volume_destroy(srf_volume *vol) {
  delete reinterpret_cast<Volume *>(vol);
  return SRF_OK;
}

srf_status srf_volume_sample(const srf_volume *vol, const float *uvw, uint64_t n, float *out,
                             void *stream) {
  const Volume *v = reinterpret_cast<const Volume *>(vol);
  if (!v || (n && (!uvw || !out))) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (!v->vox)
    return set_error(SRF_ERR_UNSUPPORTED, "srf_volume_sample reads FP32 single-component "
                                          "volumes; use srf_texture3d_fetch");
  if (n == 0) return SRF_OK;
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many samples");
  hipLaunchKernelGGL(k_vol_sample, dim3(vblocks(n, 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), *v, uvw, (uint32_t)n, out);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

static srf_status check_march(const Volume *v, const float *tf, int tf_size, float step,
                              const srf_ray *rays, uint64_t n) {
  if (!v || !tf || (n && !rays)) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (!v->vox)
    return set_error(SRF_ERR_UNSUPPORTED,
                     "the march takes FP32 single-component volumes of >= 2 voxels a side");
  if (tf_size < 2 || tf_size > kMaxTf)
    return set_error(SRF_ERR_INVALID_ARG, "tf_size must be in [2, %d]", kMaxTf);
  if (!(step > 0.0f)) return set_error(SRF_ERR_INVALID_ARG, "step must be > 0");
  if (n >= (1ull << 32)) return set_error(SRF_ERR_INVALID_ARG, "too many rays");
  return SRF_OK;
}

static uint32_t checkpoint_slots(float step) {
  // longest unit-cube chord sqrt(3) for unit directions, plus rounding slack
  double max_samples = 1.7320508075688772 / (double)step + 2.0;
  double segs = ceil(max_samples / kSegLen);
  if (segs > 1e8) segs = 1e8;
  return (uint32_t)segs + 1;
}

srf_status srf_volume_checkpoint_slots(const srf_volume *vol, float step, uint32_t *slots) {
  if (!vol || !slots) return set_error(SRF_ERR_INVALID_ARG, "null argument");
  if (!(step > 0.0f)) return set_error(SRF_ERR_INVALID_ARG, "step must be > 0");
  *slots = checkpoint_slots(step);
  return SRF_OK;
}

srf_status srf_volume_march_fwd(const srf_volume *vol, const float *tf, int tf_size, float step,
                                const srf_ray *rays, uint64_t n, float *rgba, float *ckpt,
                                void *stream) {
  const Volume *v = reinterpret_cast<const Volume *>(vol);
  srf_status st = check_march(v, tf, tf_size, step, rays, n);
  if (st != SRF_OK) return st;
  if (n && !rgba) return set_error(SRF_ERR_INVALID_ARG, "null rgba");
  if (n == 0) return SRF_OK;
  // at least 16 KB of LDS per block: fewer resident waves keep more of the
  // gathered voxels in L1 (measured faster than full occupancy)
  const size_t lds = 16 * (size_t)(tf_size > kFwdLdsEntries ? tf_size : kFwdLdsEntries);
  hipLaunchKernelGGL(k_vol_fwd, dim3(vblocks(n, kVolBlock)), dim3(kVolBlock), lds,
                     reinterpret_cast<hipStream_t>(stream), *v, tf,
                     tf_size, step, rays,
                     (uint32_t)n, rgba, ckpt, checkpoint_slots(step));
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

srf_status srf_volume_march_bwd(const srf_volume *vol, const float *tf, int tf_size, float step,
                                const srf_ray *rays, uint64_t n, const float *ckpt,
                                const float *grad_rgba, float *grad_voxels, float *grad_tf,
                                void *stream) {
  const Volume *v = reinterpret_cast<const Volume *>(vol);
  srf_status st = check_march(v, tf, tf_size, step, rays, n);
  if (st != SRF_OK) return st;
  if (n && (!grad_rgba || !grad_voxels || !grad_tf))
    return set_error(SRF_ERR_INVALID_ARG, "null gradient buffer");
  if (n == 0) return SRF_OK;
  const uint32_t slots = checkpoint_slots(step);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const size_t lds_tf = 16 * (size_t)tf_size + 4 * (size_t)tf_grad_floats(tf_size);
  const size_t lds_ck = ckpt ? 0 : 4 * (size_t)(kMaxSegNoCkpt + 1) * kVolBwdBlock;
  const size_t waves = kVolBwdBlock / 64;
  const size_t lds12 = lds_tf + lds_ck + 4 * waves * 12 * 12 * 12;
  const size_t lds10 = lds_tf + lds_ck + 4 * waves * 10 * 10 * 10;
  const dim3 grid(vblocks(n, kVolBwdBlock)), block(kVolBwdBlock);
  if (lds12 <= 64 * 1024) {
    if (ckpt)
      hipLaunchKernelGGL((k_vol_bwd<true, 12>), grid, block, lds12, s, *v, tf, tf_size, step,
                         rays, (uint32_t)n, ckpt, slots, grad_rgba, grad_voxels, grad_tf);
    else
      hipLaunchKernelGGL((k_vol_bwd<false, 12>), grid, block, lds12, s, *v, tf, tf_size, step,
                         rays, (uint32_t)n, nullptr, 0u, grad_rgba, grad_voxels, grad_tf);
  } else {
    if (ckpt)
      hipLaunchKernelGGL((k_vol_bwd<true, 10>), grid, block, lds10, s, *v, tf, tf_size, step,
                         rays, (uint32_t)n, ckpt, slots, grad_rgba, grad_voxels, grad_tf);
    else
      hipLaunchKernelGGL((k_vol_bwd<false, 10>), grid, block, lds10, s, *v, tf, tf_size, step,
                         rays, (uint32_t)n, nullptr, 0u, grad_rgba, grad_voxels, grad_tf);
  }
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

}  // extern "C"
