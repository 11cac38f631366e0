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
// walked from its end to its start recovering T_k = T_{k+1} / (1 - a_k); a
// sample with 1 - a_k < 1/4 (where that division would amplify rounding)
// gets T_k by re-marching forward from the segment's checkpoint instead.
// With T_k in hand the division-free recursions give the gradients:
//   R_{k-1} = a_k c_k + (1 - a_k) R_k,  P_{k-1} = (1 - a_k) P_k,
//   dL/dc_k = g_C T_k a_k,  dL/da_k = T_k (g_C . (c_k - R_k) + g_A P_k).
// TF gradients collect in registers while a ray stays in one TF bin, then
// in LDS per block; voxel gradients go out as float atomics, merged across
// consecutive samples that share a cell.
#include <float.h>
#include <math.h>

#include "srf_internal.h"

namespace srf {

struct Volume {
  int device;
  const float *vox;
  int dims[3];
  int clamp[3];
};

namespace {

constexpr int kVolBlock = 128;
constexpr int kMaxTf = 1024;  // entries; forward keeps the table in LDS
constexpr int kSegLen = 16;     // samples per checkpointed segment
constexpr int kMaxSeg = 64;     // LDS checkpoints per ray when recomputing
constexpr int kMaxSamples = 65536;

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

__global__ void __launch_bounds__(256) k_vol_sample(Volume V, const float *__restrict__ uvw,
                                                   uint32_t n, float *__restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Geom g;
  tex_geom(V, uvw[3 * (size_t)i], uvw[3 * (size_t)i + 1], uvw[3 * (size_t)i + 2], g);
  out[i] = tex_eval(V.vox, g);
}

// Segment length for a ray of ns samples so that its checkpoints (segment
// starts plus the final transmittance) fit in `slots`: 16, doubled as needed
// (only rays longer than the unit-direction bound ever need that).
__device__ __forceinline__ int seg_len_for(int ns, uint32_t slots) {
  int s = kSegLen;
  while ((uint32_t)((ns + s - 1) / s + 1) > slots) s *= 2;
  return s;
}

__device__ __forceinline__ int count_samples(float tn, float tfar, float step) {
  int ns = 0;
  while (ns < kMaxSamples && (tn + ((float)ns + 0.5f) * step) < tfar) ns++;
  return ns;
}

__global__ void __launch_bounds__(kVolBlock) k_vol_fwd(Volume V, const float *__restrict__ tf,
                                                      int N, float step,
                                                      const srf_ray *__restrict__ rays,
                                                      uint32_t n, float *__restrict__ rgba,
                                                      float *__restrict__ ckpt, uint32_t slots) {
  __shared__ float4 s_tf[kMaxTf];
  for (int k = threadIdx.x; k < N; k += kVolBlock)
    s_tf[k] = reinterpret_cast<const float4 *>(tf)[k];
  __syncthreads();
  uint32_t i = blockIdx.x * kVolBlock + threadIdx.x;
  if (i >= n) return;
  float o[3], d[3];
  load_ray(rays, i, o, d);
  float C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, T = 1.0f;
  float *ck = ckpt ? ckpt + (size_t)i * slots : nullptr;
  float tn, tfar;
  int k = 0, seg = kSegLen;
  if (vol_span(o, d, tn, tfar)) {
    if (ck) seg = seg_len_for(count_samples(tn, tfar, step), slots);
    for (; k < kMaxSamples; k++) {
      float t = tn + ((float)k + 0.5f) * step;
      if (!(t < tfar)) break;
      if (ck && (k % seg) == 0) ck[k / seg] = T;
      Geom g;
      tex_geom(V, o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2], g);
      float s = tex_eval(V.vox, g);
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
  }
  // transmittance after the last sample closes the last segment
  if (ck) ck[(k + seg - 1) / seg] = T;
  reinterpret_cast<float4 *>(rgba)[i] = make_float4(C0, C1, C2, 1.0f - T);
}

struct VoxelAcc {
  uint32_t a[8];
  float g[8];
  bool live;
};

__device__ __forceinline__ void flush_vox(VoxelAcc &acc, float *gv) {
  if (!acc.live) return;
  for (int c = 0; c < 8; c++)
    if (acc.g[c] != 0.0f) unsafeAtomicAdd(gv + acc.a[c], acc.g[c]);
  acc.live = false;
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

// Dynamic LDS: tf[N] (float4), tf gradient [4N], and, when the forward pass
// left no checkpoints, [kMaxSeg][block] recomputed ones.
template <bool kHaveCkpt>
__global__ void __launch_bounds__(kVolBlock) k_vol_bwd(Volume V, const float *__restrict__ tf,
                                                      int N, float step,
                                                      const srf_ray *__restrict__ rays,
                                                      uint32_t n, const float *__restrict__ ckpt,
                                                      uint32_t slots,
                                                      const float *__restrict__ grad_rgba,
                                                      float *__restrict__ grad_vox,
                                                      float *__restrict__ grad_tf) {
  extern __shared__ float4 smem4[];
  float4 *s_tf = smem4;
  float *s_gtf = reinterpret_cast<float *>(smem4 + N);
  float *s_ck = s_gtf + 4 * N;
  for (int k = threadIdx.x; k < N; k += kVolBlock) {
    s_tf[k] = reinterpret_cast<const float4 *>(tf)[k];
    s_gtf[4 * k] = s_gtf[4 * k + 1] = s_gtf[4 * k + 2] = s_gtf[4 * k + 3] = 0.0f;
  }
  __syncthreads();
  const uint32_t i = blockIdx.x * kVolBlock + threadIdx.x;
  float o[3], d[3], tn = 0.0f, tfar = 0.0f;
  bool live = i < n;
  if (live) {
    load_ray(rays, i, o, d);
    live = vol_span(o, d, tn, tfar);
  }
  if (live) {
    const float4 gr = reinterpret_cast<const float4 *>(grad_rgba)[i];
    // sample count (same float arithmetic as the march)
    const int ns = count_samples(tn, tfar, step);
    const int seg_len = seg_len_for(ns, kHaveCkpt ? slots : (uint32_t)(kMaxSeg + 1));
    const float *ck;
    int ck_stride;
    if (kHaveCkpt) {
      ck = ckpt + (size_t)i * slots;
      ck_stride = 1;
    } else {
      // one forward replay leaves at most kMaxSeg + 1 checkpoints in LDS
      float T = 1.0f;
      float *w = s_ck + threadIdx.x;
      for (int k = 0; k < ns; k++) {
        if (k % seg_len == 0) w[(k / seg_len) * kVolBlock] = T;
        T = T * (1.0f - sample_alpha(V, s_tf, N, o, d, tn + ((float)k + 0.5f) * step));
      }
      w[((ns + seg_len - 1) / seg_len) * kVolBlock] = T;
      ck = w;
      ck_stride = kVolBlock;
    }
    float R0 = 0.0f, R1 = 0.0f, R2 = 0.0f, P = 1.0f;
    VoxelAcc vacc;
    vacc.live = false;
    TfAcc tacc;
    tacc.bin = -1;
    const int nseg = (ns + seg_len - 1) / seg_len;
    for (int sg = nseg - 1; sg >= 0; sg--) {
      const int k0 = sg * seg_len;
      const int k1 = (k0 + seg_len < ns) ? k0 + seg_len : ns;
      float Tnext = ck[(sg + 1) * ck_stride];  // T after sample k1 - 1
      const float Tstart = ck[sg * ck_stride];
      for (int k = k1 - 1; k >= k0; k--) {
        float t = tn + ((float)k + 0.5f) * step;
        Geom g;
        tex_geom(V, o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2], g);
        float s = tex_eval(V.vox, g);
        int i0;
        float f;
        bool inside;
        tf_index(s, N, i0, f, inside);
        const float4 ta = s_tf[i0], tb = s_tf[i0 + 1];
        float c0r = lerpf(f, ta.x, tb.x), c1g = lerpf(f, ta.y, tb.y), c2b = lerpf(f, ta.z, tb.z);
        float a = lerpf(f, ta.w, tb.w);
        // transmittance in front of sample k
        float Tk;
        if (k == k0) {
          Tk = Tstart;
        } else if (1.0f - a >= 0.25f) {
          Tk = Tnext / (1.0f - a);
        } else {
          Tk = Tstart;
          for (int j = k0; j < k; j++)
            Tk = Tk * (1.0f - sample_alpha(V, s_tf, N, o, d, tn + ((float)j + 0.5f) * step));
        }
        Tnext = Tk;
        float wa = Tk * a;
        float g0 = gr.x * wa, g1 = gr.y * wa, g2 = gr.z * wa;
        float ga = Tk * ((gr.x * (c0r - R0) + gr.y * (c1g - R1) + gr.z * (c2b - R2)) + gr.w * P);
        float fo = 1.0f - f;
        if (tacc.bin != i0) {
          flush_tf(tacc, s_gtf);
          tacc.bin = i0;
          for (int c = 0; c < 8; c++) tacc.g[c] = 0.0f;
        }
        tacc.g[0] += fo * g0;
        tacc.g[1] += fo * g1;
        tacc.g[2] += fo * g2;
        tacc.g[3] += fo * ga;
        tacc.g[4] += f * g0;
        tacc.g[5] += f * g1;
        tacc.g[6] += f * g2;
        tacc.g[7] += f * ga;
        if (inside) {
          float ds = ((g0 * (tb.x - ta.x) + g1 * (tb.y - ta.y)) +
                      (g2 * (tb.z - ta.z) + ga * (tb.w - ta.w))) *
                     (float)(N - 1);
          if (ds != 0.0f) {
            if (vacc.live && (vacc.a[0] != g.a[0] || vacc.a[7] != g.a[7])) flush_vox(vacc, grad_vox);
            if (!vacc.live) {
              for (int c = 0; c < 8; c++) {
                vacc.a[c] = g.a[c];
                vacc.g[c] = 0.0f;
              }
              vacc.live = true;
            }
            const float ux = 1.0f - g.dx, uy = 1.0f - g.dy, uz = 1.0f - g.dz;
            vacc.g[0] += ds * (ux * uy * uz);
            vacc.g[1] += ds * (g.dx * uy * uz);
            vacc.g[2] += ds * (ux * g.dy * uz);
            vacc.g[3] += ds * (g.dx * g.dy * uz);
            vacc.g[4] += ds * (ux * uy * g.dz);
            vacc.g[5] += ds * (g.dx * uy * g.dz);
            vacc.g[6] += ds * (ux * g.dy * g.dz);
            vacc.g[7] += ds * (g.dx * g.dy * g.dz);
          }
        }
        R0 = a * c0r + (1.0f - a) * R0;
        R1 = a * c1g + (1.0f - a) * R1;
        R2 = a * c2b + (1.0f - a) * R2;
        P = (1.0f - a) * P;
      }
    }
    flush_vox(vacc, grad_vox);
    flush_tf(tacc, s_gtf);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < 4 * N; k += kVolBlock) {
    float v = s_gtf[k];
    if (v != 0.0f) unsafeAtomicAdd(grad_tf + k, v);
  }
}

inline uint32_t vblocks(uint64_t n, int b) { return (uint32_t)((n + b - 1) / b); }

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
  v->dims[0] = w;
  v->dims[1] = h;
  v->dims[2] = d;
  v->clamp[0] = wrap_s;
  v->clamp[1] = wrap_t;
  v->clamp[2] = wrap_r;
  *out = reinterpret_cast<srf_volume *>(v);
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
  hipLaunchKernelGGL(k_vol_fwd, dim3(vblocks(n, kVolBlock)), dim3(kVolBlock), 0,
                     reinterpret_cast<hipStream_t>(stream), *v, tf, tf_size, step, rays,
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
  const size_t lds_tf = 32 * (size_t)tf_size;
  if (ckpt) {
    hipLaunchKernelGGL(k_vol_bwd<true>, dim3(vblocks(n, kVolBlock)), dim3(kVolBlock), lds_tf, s,
                       *v, tf, tf_size, step, rays, (uint32_t)n, ckpt, slots, grad_rgba,
                       grad_voxels, grad_tf);
  } else {
    const size_t lds = lds_tf + 4 * (size_t)(kMaxSeg + 1) * kVolBlock;
    hipLaunchKernelGGL(k_vol_bwd<false>, dim3(vblocks(n, kVolBlock)), dim3(kVolBlock), lds, s,
                       *v, tf, tf_size, step, rays, (uint32_t)n, nullptr, 0u, grad_rgba,
                       grad_voxels, grad_tf);
  }
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

}  // extern "C"
