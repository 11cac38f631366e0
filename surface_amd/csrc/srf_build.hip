// srf_build.hip -- LBVH construction on gfx950, stream-ordered, no host sync.
//
// Pipeline (one launch each unless noted), for n primitives:
//   bounds   fused min/max over every primitive's vertices, one partial per
//            block, the last block to finish (atomic ticket) reduces the
//            partials and derives inv/kEPS/epsScale; it also clears the
//            Morton histogram and the sort's tile counters.
//   morton   centroid -> 30-bit code + identity index, and all four 8-bit
//            digit histograms of the codes (LDS, then one global atomic per
//            bin per block).
//   sort     4 onesweep passes: LDS-staged tiles ranked with wave64 ballots,
//            tile digit totals chained by decoupled look-back, scattered
//            through LDS so each digit run leaves as one coalesced burst.
//   karras   one thread per internal node (Karras 2012): split, leaf flags,
//            range end, parent pointers.
//   classify one thread per internal node: branch or collapsed (range <= leaf
//            limit or depth > 22), by a bounded walk of the parent chain
//            that also yields the preorder terms; marks leaf starts.
//   scan     exclusive scan of the leaf starts (3 launches).
//   emit     one thread per branch: writes the node at its depth-first
//            preorder position pos = ancestors + 2*leaves_left - right_turns,
//            which is where the reference's serial builder puts it, and the
//            leaf records of its leaf children.
//   leaves   one thread per leaf: leaf box + packed primitives, then an
//            atomic bottom-up refit (second arrival merges and climbs).
#include <float.h>

#include "srf_internal.h"

namespace srf {
namespace {

constexpr int kBoundsBlocks = 1024;
constexpr int kSortBlock = 256;   // one thread per digit
constexpr int kSortItems = 15;    // keys per thread
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 26;
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint32_t part_by2_30(uint32_t n) {
  // PartBy2_30, render/render_prefix_tree_util.h:113-120
  n &= 0x000003ffu;
  n = (n ^ (n << 16)) & 0xff0000ffu;
  n = (n ^ (n << 8)) & 0x0300f00fu;
  n = (n ^ (n << 4)) & 0x030c30c3u;
  n = (n ^ (n << 2)) & 0x09249249u;
  return n;
}

// (uint32_t)(float): x86-64 truncates through int64; for |x| < 2^31 and NaN
// (both give 0) this equals the 32-bit truncation.
__device__ __forceinline__ uint32_t f2u_trunc(float x) { return (uint32_t)(int)x; }

__device__ __forceinline__ uint32_t morton30(float px, float py, float pz,
                                             const BuildParams &p) {
  // MortionCode30, render/render_prefix_tree_util.h:143-157
  uint32_t ix = f2u_trunc((px - p.bmin[0]) * p.inv[0]);
  uint32_t iy = f2u_trunc((py - p.bmin[1]) * p.inv[1]);
  uint32_t iz = f2u_trunc((pz - p.bmin[2]) * p.inv[2]);
  return (part_by2_30(ix) << 2) | (part_by2_30(iy) << 1) | part_by2_30(iz);
}

__device__ __forceinline__ int clz32(uint32_t x) { return x == 0 ? 32 : __clz(x); }

// std::min / std::max as the reference calls them: keep a unless b is
// strictly smaller / larger.
__device__ __forceinline__ float fmin_ref(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmax_ref(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ float3 load_v3(const float *V, uint32_t i) {
  const float *p = V + 3 * (size_t)i;
  return make_float3(p[0], p[1], p[2]);
}

struct Box {
  float mn[3], mx[3];
};

__device__ __forceinline__ void box_add(Box &b, float x, float y, float z, float e) {
  b.mn[0] = fmin_ref(b.mn[0], x - e);
  b.mn[1] = fmin_ref(b.mn[1], y - e);
  b.mn[2] = fmin_ref(b.mn[2], z - e);
  b.mx[0] = fmax_ref(b.mx[0], x + e);
  b.mx[1] = fmax_ref(b.mx[1], y + e);
  b.mx[2] = fmax_ref(b.mx[2], z + e);
}

__device__ __forceinline__ void box_merge(Box &a, const Box &b) {
  for (int k = 0; k < 3; k++) {
    a.mn[k] = fmin_ref(a.mn[k], b.mn[k]);
    a.mx[k] = fmax_ref(a.mx[k], b.mx[k]);
  }
}

__device__ __forceinline__ void box_init(Box &b) {
  for (int k = 0; k < 3; k++) {
    b.mn[k] = FLT_MAX;
    b.mx[k] = -FLT_MAX;
  }
}

struct PrimSrc {
  const float *V;
  const uint32_t *F;
  const float *R;
  float rconst;
  uint32_t n;
};

// Per-primitive contribution to the scene box:
//  tri  ComputeBoundingBoxOMP, epsScale 0  render/render_accel_triangle.cc:711-797,1430-1443
//  tet  kEPS = 16*FLT_EPSILON               render/render_accel_tetra.cc:971-1055
//  sph  pos -+ r -+ 16*FLT_EPSILON          render/render_accel_particle.cc:303-378
template <int K>
__device__ __forceinline__ void scene_add(Box &b, const PrimSrc &s, uint32_t i) {
  if (K == kTri) {
    for (int j = 0; j < 3; j++) {
      float3 v = load_v3(s.V, s.F[3 * (size_t)i + j]);
      box_add(b, v.x, v.y, v.z, FLT_EPSILON * 0.0f);
    }
  } else if (K == kTet) {
    for (int j = 0; j < 4; j++) {
      float3 v = load_v3(s.V, s.F[4 * (size_t)i + j]);
      box_add(b, v.x, v.y, v.z, FLT_EPSILON * 16);
    }
  } else {
    float3 p = load_v3(s.V, i);
    float r = s.R ? s.R[i] : s.rconst;
    const float e = (float)(FLT_EPSILON * 16.0);
    b.mn[0] = fmin_ref(b.mn[0], (p.x - r) - e);
    b.mn[1] = fmin_ref(b.mn[1], (p.y - r) - e);
    b.mn[2] = fmin_ref(b.mn[2], (p.z - r) - e);
    b.mx[0] = fmax_ref(b.mx[0], (p.x + r) + e);
    b.mx[1] = fmax_ref(b.mx[1], (p.y + r) + e);
    b.mx[2] = fmax_ref(b.mx[2], (p.z + r) + e);
  }
}

__device__ __forceinline__ void box_wave_reduce(Box &b) {
  for (int off = 32; off >= 1; off >>= 1) {
    Box o;
    for (int k = 0; k < 3; k++) {
      o.mn[k] = __shfl_xor(b.mn[k], off, 64);
      o.mx[k] = __shfl_xor(b.mx[k], off, 64);
    }
    box_merge(b, o);
  }
}

template <int K>
__global__ void __launch_bounds__(256) k_bounds(PrimSrc s, Box *partials,
                                               BuildParams *params,
                                               uint32_t *clear, uint32_t nclear) {
  Box b;
  box_init(b);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n;
       i += gridDim.x * blockDim.x)
    scene_add<K>(b, s, i);
  box_wave_reduce(b);
  __shared__ Box wbox[4];
  __shared__ bool last;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wbox[wave] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    Box r = wbox[0];
    for (int w = 1; w < 4; w++) box_merge(r, wbox[w]);
    partials[blockIdx.x] = r;
    __threadfence();
    uint32_t t = atomicAdd(&params->ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  // clear the histogram / tile counters the next stages accumulate into
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nclear;
       i += gridDim.x * blockDim.x)
    clear[i] = 0;
  __syncthreads();
  if (!last) return;
  __threadfence();
  box_init(b);
  for (uint32_t i = threadIdx.x; i < gridDim.x; i += blockDim.x) {
    Box o;
    for (int k = 0; k < 3; k++) {
      o.mn[k] = __hip_atomic_load(&partials[i].mn[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      o.mx[k] = __hip_atomic_load(&partials[i].mx[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    box_merge(b, o);
  }
  box_wave_reduce(b);
  if (lane == 0) wbox[wave] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    Box r = wbox[0];
    for (int w = 1; w < 4; w++) box_merge(r, wbox[w]);
    BuildParams p;
    for (int k = 0; k < 3; k++) {
      p.bmin[k] = r.mn[k];
      p.bmax[k] = r.mx[k];
      // kDIV / (bmax - bmin), render/render_prefix_tree_util.cc:242-245
      p.inv[k] = (float)1024 / (r.mx[k] - r.mn[k]);
    }
    // epsScale = max extent, render/render_accel_triangle.cc:1451-1463
    double bs0 = (double)(r.mx[0] - r.mn[0]);
    double bs1 = (double)(r.mx[1] - r.mn[1]);
    double bs2 = (double)(r.mx[2] - r.mn[2]);
    double es = bs0;
    if (es < bs1) es = bs1;
    if (es < bs2) es = bs2;
    p.eps_scale = es;
    if (K == kTri)
      p.kEPS = FLT_EPSILON * (float)es;  // ComputeBoundingBox30's real kEPS
    else if (K == kTet)
      p.kEPS = FLT_EPSILON * 16;
    else
      p.kEPS = (float)(FLT_EPSILON * 16.0);
    for (int k = 0; k < 3; k++) {
      p.root_bmin[k] = FLT_MAX;
      p.root_bmax[k] = -FLT_MAX;
    }
    p.num_leaves = 0;
    p.num_nodes = 0;
    p.max_steps = 0;
    p.error = 0;
    p.ticket = 0;
    p.pad[0] = p.pad[1] = p.pad[2] = 0;
    *params = p;
  }
}

// ---- Morton codes + digit histograms ----
// tri CalculateMortonCodesTriangleFloat30 render/render_prefix_tree_util.cc:230-265
// tet CalculateMortonCodesTetraFloat30    render/render_prefix_tree_util.cc:267-304
// sph CalculateMortonCodes30SIMD          render/render_prefix_tree_util.cc:176-228
template <int K>
__device__ __forceinline__ uint32_t prim_code(const PrimSrc &s, uint32_t i,
                                              const BuildParams &p) {
  if (K == kTri) {
    const float one_third = 1.0f / 3.0f;
    float3 a = load_v3(s.V, s.F[3 * (size_t)i]);
    float3 b = load_v3(s.V, s.F[3 * (size_t)i + 1]);
    float3 c = load_v3(s.V, s.F[3 * (size_t)i + 2]);
    return morton30(one_third * ((a.x + b.x) + c.x), one_third * ((a.y + b.y) + c.y),
                    one_third * ((a.z + b.z) + c.z), p);
  } else if (K == kTet) {
    const float q = 1.0f / 4.0f;
    float3 a = load_v3(s.V, s.F[4 * (size_t)i]);
    float3 b = load_v3(s.V, s.F[4 * (size_t)i + 1]);
    float3 c = load_v3(s.V, s.F[4 * (size_t)i + 2]);
    float3 d = load_v3(s.V, s.F[4 * (size_t)i + 3]);
    return morton30(q * (((a.x + b.x) + c.x) + d.x), q * (((a.y + b.y) + c.y) + d.y),
                    q * (((a.z + b.z) + c.z) + d.z), p);
  } else {
    float3 a = load_v3(s.V, i);
    return morton30(a.x, a.y, a.z, p);
  }
}

template <int K>
__global__ void __launch_bounds__(256) k_morton(PrimSrc s, const BuildParams *params,
                                               uint32_t *codes, uint32_t *hist) {
  __shared__ uint32_t h[4][256];
  for (int i = threadIdx.x; i < 4 * 256; i += blockDim.x) (&h[0][0])[i] = 0;
  __syncthreads();
  const BuildParams p = *params;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n;
       i += gridDim.x * blockDim.x) {
    uint32_t c = prim_code<K>(s, i, p);
    codes[i] = c;
    atomicAdd(&h[0][c & 255], 1u);
    atomicAdd(&h[1][(c >> 8) & 255], 1u);
    atomicAdd(&h[2][(c >> 16) & 255], 1u);
    atomicAdd(&h[3][c >> 24], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 4 * 256; i += blockDim.x) {
    uint32_t v = (&h[0][0])[i];
    if (v) atomicAdd(&hist[i], v);
  }
}

// Morton codes for an explicit scene box (the stage entry point).
template <int K>
__global__ void __launch_bounds__(256) k_morton_box(PrimSrc s, BuildParams p, uint32_t *codes) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n;
       i += gridDim.x * blockDim.x)
    codes[i] = prim_code<K>(s, i, p);
}

__global__ void __launch_bounds__(256) k_histogram(const uint32_t *codes, uint32_t n,
                                                  uint32_t *hist) {
  __shared__ uint32_t h[4][256];
  for (int i = threadIdx.x; i < 4 * 256; i += blockDim.x) (&h[0][0])[i] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    uint32_t c = codes[i];
    atomicAdd(&h[0][c & 255], 1u);
    atomicAdd(&h[1][(c >> 8) & 255], 1u);
    atomicAdd(&h[2][(c >> 16) & 255], 1u);
    atomicAdd(&h[3][c >> 24], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 4 * 256; i += blockDim.x) {
    uint32_t v = (&h[0][0])[i];
    if (v) atomicAdd(&hist[i], v);
  }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---- onesweep radix pass ----
// Block-wide exclusive scan of one value per thread (256 threads).
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t *tmp,
                                                        uint32_t *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) tmp[wave] = x;
  __syncthreads();
  uint32_t base = 0, all = 0;
  for (int w = 0; w < 4; w++) {
    uint32_t t = tmp[w];
    if (w < wave) base += t;
    all += t;
  }
  __syncthreads();
  if (total) *total = all;
  return base + x - v;
}

__global__ void __launch_bounds__(kSortBlock)
k_onesweep(const uint32_t *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
           uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out, uint32_t n,
           int shift, const uint32_t *__restrict__ hist /* this pass */,
           uint32_t *status /* [tiles][256] */, uint32_t *tile_counter,
           uint32_t *error) {
  __shared__ uint32_t s_keys[kSortTile];
  __shared__ uint32_t s_vals[kSortTile];
  __shared__ uint32_t s_whist[4][256];   // per-wave digit counts, then offsets
  __shared__ uint32_t s_gbase[256];      // global destination of digit run
  __shared__ uint32_t s_texcl[256];      // tile-exclusive digit prefix
  __shared__ uint32_t s_tmp[8];
  __shared__ uint32_t s_tile;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_tile = atomicAdd(tile_counter, 1u);
  for (int i = threadIdx.x; i < 4 * 256; i += kSortBlock) (&s_whist[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t tile_base = tile * kSortTile;
  const uint32_t seg = tile_base + wave * (kSortItems * 64);

  uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    uint32_t idx = seg + k * 64 + lane;
    bool ok = idx < n;
    key[k] = ok ? keys_in[idx] : 0xffffffffu;
    val[k] = ok ? (vals_in ? vals_in[idx] : idx) : 0u;
  }
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    uint32_t idx = seg + k * 64 + lane;
    bool ok = idx < n;
    uint32_t d = (key[k] >> shift) & 255u;
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    uint32_t before = __popcll(peers & lt_mask);
    uint32_t cnt = __popcll(peers);
    uint32_t prev = ok ? s_whist[wave][d] : 0u;
    rank[k] = prev + before;
    // the highest peer publishes the new running count
    bool top = ok && (before == cnt - 1);
    if (top) s_whist[wave][d] = prev + cnt;
  }
  __syncthreads();

  // per-digit wave offsets and tile totals (thread = digit)
  const uint32_t d = threadIdx.x;
  uint32_t run = 0;
  for (int w = 0; w < 4; w++) {
    uint32_t c = s_whist[w][d];
    s_whist[w][d] = run;
    run += c;
  }
  const uint32_t tile_count = run;
  // publish this tile's aggregate, then look back for the prefix
  uint32_t *my = status + (size_t)tile * 256 + d;
  if (tile == 0) {
    __hip_atomic_store(my, kFlagInc | tile_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_store(my, kFlagAgg | tile_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  uint32_t excl = 0;
  if (tile > 0) {
    int32_t j = (int32_t)tile - 1;
    uint32_t spins = 0;
    while (j >= 0) {
      uint32_t s = __hip_atomic_load(status + (size_t)j * 256 + d, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
      uint32_t flag = s & ~kValueMask;
      if (flag == 0) {
        if (++spins > kSpinLimit) {
          atomicOr(error, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      excl += s & kValueMask;
      if (flag == kFlagInc) break;
      j--;
    }
    __hip_atomic_store(my, kFlagInc | ((excl + tile_count) & kValueMask), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  // global exclusive prefix of the digit over the whole pass
  uint32_t gpre = block_excl_scan256(hist[d], s_tmp, nullptr);
  uint32_t texcl = block_excl_scan256(tile_count, s_tmp, nullptr);
  s_gbase[d] = gpre + excl;
  s_texcl[d] = texcl;
  __syncthreads();

  // stage the tile in digit order through LDS
#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    uint32_t idx = seg + k * 64 + lane;
    if (idx < n) {
      uint32_t dd = (key[k] >> shift) & 255u;
      uint32_t pos = s_texcl[dd] + s_whist[wave][dd] + rank[k];
      s_keys[pos] = key[k];
      s_vals[pos] = val[k];
    }
  }
  __syncthreads();
  const uint32_t valid = (n - tile_base) < (uint32_t)kSortTile ? (n - tile_base) : kSortTile;
  for (uint32_t i = threadIdx.x; i < valid; i += kSortBlock) {
    uint32_t kk = s_keys[i];
    uint32_t dd = (kk >> shift) & 255u;
    uint32_t dst = s_gbase[dd] + (i - s_texcl[dd]);
    if (dst < n) {
      keys_out[dst] = kk;
      vals_out[dst] = s_vals[i];
    }
  }
}

// ---- Karras radix tree ----
// Delta30, render/render_prefix_tree_util.cc:501-526
__device__ __forceinline__ int delta30(const uint32_t *c, int i, int j, int n) {
  if (j < 0 || j >= n) return -1;
  uint32_t ki = c[i], kj = c[j];
  int off = 0;
  if (ki == kj) {
    ki = (uint32_t)i;
    kj = (uint32_t)j;
    off = 32;
  }
  return clz32(ki ^ kj) + off;
}

// ConstructBinaryRadixTree30, render/render_prefix_tree_util.cc:562-649
__device__ __forceinline__ void radix_node30(const uint32_t *c, int i, int n, uint32_t &child,
                                             uint32_t &leafbits, int &jout) {
  int di = delta30(c, i, i + 1, n) - delta30(c, i, i - 1, n);
  int d = di >= 0 ? 1 : -1;
  int dmin = delta30(c, i, i - d, n);
  int lmax = 2;
  while (delta30(c, i, i + lmax * d, n) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta30(c, i, i + (l + t) * d, n) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta30(c, i, j, n);
  int s = 0;
  int k = 2;
  int t = (l + k - 1) / k;  // (int)ceil((double)l / k)
  if (t > 0) {
    for (;;) {
      if (delta30(c, i, i + (s + t) * d, n) > dnode) s += t;
      if (t == 1) break;
      k *= 2;
      t = (l + k - 1) / k;
    }
  }
  int r = i + s * d + (d < 0 ? d : 0);
  child = (uint32_t)r;
  int lo = i < j ? i : j, hi = i < j ? j : i;
  leafbits = (lo == r ? 1u : 0u) | (hi == r + 1 ? 2u : 0u);
  jout = j;
}

__global__ void __launch_bounds__(256) k_karras(const uint32_t *codes, uint32_t n,
                                               uint32_t *kchild, uint32_t *kleaf,
                                               uint32_t *kother, uint32_t *kparent) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n) return;
  uint32_t child, leafbits;
  int j;
  radix_node30(codes, (int)i, (int)n, child, leafbits, j);
  kchild[i] = child;
  kleaf[i] = leafbits;
  kother[i] = (uint32_t)j;
  // parent link of an internal child: parent << 1 | (1 for the right child)
  if (!(leafbits & 1u)) kparent[child] = i << 1;
  if (!(leafbits & 2u)) kparent[child + 1] = (i << 1) | 1u;
}

__global__ void __launch_bounds__(256) k_karras_nodeinfo(const uint32_t *codes, uint32_t n,
                                                        srf_node_info *out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n) return;
  uint32_t child, leafbits;
  int j;
  radix_node30(codes, (int)i, (int)n, child, leafbits, j);
  out[i].child_index = child;
  out[i].left_leaf = leafbits & 1u;
  out[i].right_leaf = (leafbits >> 1) & 1u;
}

// ---- collapse: which Karras nodes become branches ----
// BuildTreeRecursive32, render/render_accel_triangle.cc:895-992: a child range
// [l, r] becomes a leaf iff it is a Karras leaf, r - l <= MAX_LEAF, or its
// depth (root's children = 0) exceeds 22.
// cls[i] = 0 (not a branch) or 1 | steps << 1 | right_turns << 8.
__global__ void __launch_bounds__(256) k_classify(uint32_t n, int max_leaf,
                                                 const uint32_t *kchild, const uint32_t *kleaf,
                                                 const uint32_t *kother, const uint32_t *kparent,
                                                 uint32_t *cls, uint32_t *leaf_start,
                                                 BuildParams *params) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t steps = 0, turns = 0;
  bool branch = false;
  uint32_t l = 0, r = 0;
  if (i + 1 < n) {
    const uint32_t j = kother[i];
    l = i < j ? i : j;
    r = i < j ? j : i;
    if (i == 0) {
      branch = true;
    } else if (r - l > (uint32_t)max_leaf) {
      uint32_t c = i;
      while (c != 0 && steps <= (uint32_t)(kMaxDepth32 + 1)) {
        const uint32_t link = kparent[c];
        const uint32_t p = link >> 1;
        if (p + 1 >= n) {  // malformed parent link: never expected, stop safely
          atomicOr(&params->error, 8u);
          break;
        }
        turns += link & 1u;
        c = p;
        steps++;
      }
      branch = (c == 0) && steps <= (uint32_t)(kMaxDepth32 + 1);
    }
    cls[i] = branch ? (1u | (steps << 1) | (turns << 8)) : 0u;
  }
  // deepest branch (one atomic per wavefront)
  uint32_t m = branch ? steps : 0u;
  for (int off = 32; off >= 1; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
  if ((threadIdx.x & 63) == 0 && m) atomicMax(&params->max_steps, m);
  if (!branch) return;
  // children of this branch sit at depth `steps`
  const uint32_t g = kchild[i];
  const uint32_t lb = kleaf[i];
  const bool deep = steps > (uint32_t)kMaxDepth32;
  const bool left_leaf = (lb & 1u) || (g - l <= (uint32_t)max_leaf) || deep;
  const bool right_leaf = (lb & 2u) || (r - (g + 1) <= (uint32_t)max_leaf) || deep;
  if (left_leaf) leaf_start[l] = 1;
  if (right_leaf) leaf_start[g + 1] = 1;
}

// ---- exclusive scan (3 launches) ----
__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const uint32_t *in, uint32_t n,
                                                           uint32_t *partial) {
  uint32_t base = blockIdx.x * kScanTile;
  uint32_t sum = 0;
  for (int k = 0; k < kScanItems; k++) {
    uint32_t idx = base + k * kScanBlock + threadIdx.x;
    if (idx < n) sum += in[idx];
  }
  for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, 64);
  __shared__ uint32_t w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = w[0] + w[1] + w[2] + w[3];
}

__global__ void __launch_bounds__(1024) k_scan_partials(uint32_t *partial, uint32_t m) {
  // single block: exclusive scan of m partials in place
  __shared__ uint32_t s[1024];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < m; base += 1024) {
    uint32_t idx = base + threadIdx.x;
    uint32_t v = idx < m ? partial[idx] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      uint32_t t = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    uint32_t incl = s[threadIdx.x];
    if (idx < m) partial[idx] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += incl;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kScanBlock) k_scan_downsweep(const uint32_t *in, uint32_t n,
                                                              const uint32_t *partial,
                                                              uint32_t *out,
                                                              BuildParams *params) {
  // blocked arrangement: thread t owns items [t*kScanItems, (t+1)*kScanItems)
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
  for (int k = 0; k < kScanItems; k++) {
    uint32_t idx = base + k;
    v[k] = idx < n ? in[idx] : 0;
    sum += v[k];
  }
  __shared__ uint32_t tmp[8];
  uint32_t total;
  uint32_t pre = block_excl_scan256(sum, tmp, &total) + partial[blockIdx.x];
  for (int k = 0; k < kScanItems; k++) {
    uint32_t idx = base + k;
    if (idx < n) out[idx] = pre;
    pre += v[k];
    if (idx == n - 1) {
      params->num_leaves = pre;
      params->num_nodes = 2 * pre - 1;
    }
  }
}

// ---- node emission ----
struct NodeOut {
  void *nodes;
  uint32_t *parent_link;  // per node: parent pos << 1 | side
  uint32_t *branch_pos;   // per Karras internal node that is a branch: its node
  uint32_t *leaf_pos;     // per leaf: node position
  uint32_t *leaf_first;   // per leaf: first sorted primitive
  uint32_t *leaf_count;
};

__device__ __forceinline__ int split_axis(uint32_t code) {
  // GetSplitAxis, render/render_accel_triangle.cc:854-863 (asserts compiled out)
  return (clz32(code) - 2) % 3;
}

__device__ __forceinline__ void write_meta(void *nodes, uint32_t pos, bool leaf, int axis,
                                           uint32_t d0, uint32_t d1) {
  // words 12..15 of the 64-byte node: flag, axis, data[0], data[1]
  uint4 *w = reinterpret_cast<uint4 *>(nodes) + 4 * (size_t)pos + 3;
  *w = make_uint4(leaf ? 1u : 0u, (uint32_t)axis, d0, d1);
}

template <int K>
__global__ void __launch_bounds__(256) k_emit(uint32_t n, int max_leaf, const uint32_t *codes,
                                             const uint32_t *kchild, const uint32_t *kleaf,
                                             const uint32_t *kother, const uint32_t *cls,
                                             const uint32_t *leaf_rank, NodeOut o) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= n) return;
  uint32_t c = cls[i];
  if (!(c & 1u)) return;
  uint32_t steps = (c >> 1) & 127u, turns = c >> 8;
  uint32_t j = kother[i];
  uint32_t l = i < j ? i : j, r = i < j ? j : i;
  uint32_t g = kchild[i], lb = kleaf[i];
  uint32_t pos = steps + 2 * leaf_rank[l] - turns;
  uint32_t posL = pos + 1;
  uint32_t posR = steps + 2 * leaf_rank[g + 1] - turns;
  const uint32_t max_nodes = 2 * n - 1;
  if (pos >= max_nodes || posL >= max_nodes || posR >= max_nodes || g + 1 >= n) return;
  write_meta(o.nodes, pos, false, split_axis(codes[i]), posL, posR);
  o.branch_pos[i] = pos;
  o.parent_link[posL] = pos << 1;
  o.parent_link[posR] = (pos << 1) | 1u;
  bool deep = steps > (uint32_t)kMaxDepth32;
  if ((lb & 1u) || (g - l <= (uint32_t)max_leaf) || deep) {
    uint32_t li = leaf_rank[l];
    o.leaf_pos[li] = posL;
    o.leaf_first[li] = l;
    o.leaf_count[li] = g + 1 - l;
    write_meta(o.nodes, posL, true, 0, g + 1 - l, l);
  }
  if ((lb & 2u) || (r - (g + 1) <= (uint32_t)max_leaf) || deep) {
    uint32_t li = leaf_rank[g + 1];
    o.leaf_pos[li] = posR;
    o.leaf_first[li] = g + 1;
    o.leaf_count[li] = r - g;
    write_meta(o.nodes, posR, true, 0, r - g, g + 1);
  }
}

// ---- leaves: box, packing, atomic refit ----
template <int K>
__device__ __forceinline__ void leaf_box_pack(const PrimSrc &s, const uint32_t *sidx,
                                              uint32_t first, uint32_t count, float e,
                                              void *prims, uint32_t *sorted_faces, Box &b) {
  // tri ComputeBoundingBox30 render/render_accel_triangle.cc:660-709
  // tet ComputeBoundingBox30 render/render_accel_tetra.cc:921-969
  // sph ComputeBoundingBox30 render/render_accel_particle.cc:417-462
  for (uint32_t q = first; q < first + count; q++) {
    uint32_t id = sidx[q];
    if (K == kTri) {
      uint32_t f0 = s.F[3 * (size_t)id], f1 = s.F[3 * (size_t)id + 1],
               f2 = s.F[3 * (size_t)id + 2];
      float3 a = load_v3(s.V, f0), bb = load_v3(s.V, f1), c = load_v3(s.V, f2);
      if (q == first) {
        b.mn[0] = a.x - e; b.mn[1] = a.y - e; b.mn[2] = a.z - e;
        b.mx[0] = a.x + e; b.mx[1] = a.y + e; b.mx[2] = a.z + e;
      }
      box_add(b, a.x, a.y, a.z, e);
      box_add(b, bb.x, bb.y, bb.z, e);
      box_add(b, c.x, c.y, c.z, e);
      float4 *P = reinterpret_cast<float4 *>(prims) + 3 * (size_t)q;
      P[0] = make_float4(a.x, a.y, a.z, __uint_as_float(id));
      P[1] = make_float4(bb.x - a.x, bb.y - a.y, bb.z - a.z, 0.0f);
      P[2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
    } else if (K == kTet) {
      uint32_t f[4];
      float3 v[4];
      for (int k = 0; k < 4; k++) {
        f[k] = s.F[4 * (size_t)id + k];
        v[k] = load_v3(s.V, f[k]);
      }
      if (q == first) {
        b.mn[0] = v[0].x - e; b.mn[1] = v[0].y - e; b.mn[2] = v[0].z - e;
        b.mx[0] = v[0].x + e; b.mx[1] = v[0].y + e; b.mx[2] = v[0].z + e;
      }
      for (int k = 0; k < 4; k++) box_add(b, v[k].x, v[k].y, v[k].z, e);
      float4 *P = reinterpret_cast<float4 *>(prims) + 4 * (size_t)q;
      P[0] = make_float4(v[0].x, v[0].y, v[0].z, __uint_as_float(id));
      P[1] = make_float4(v[1].x, v[1].y, v[1].z, 0.0f);
      P[2] = make_float4(v[2].x, v[2].y, v[2].z, 0.0f);
      P[3] = make_float4(v[3].x, v[3].y, v[3].z, 0.0f);
      reinterpret_cast<uint4 *>(sorted_faces)[q] = make_uint4(f[0], f[1], f[2], f[3]);
    } else {
      float3 p = load_v3(s.V, id);
      float r = s.R ? s.R[id] : s.rconst;
      float lo0 = (p.x - r) - e, lo1 = (p.y - r) - e, lo2 = (p.z - r) - e;
      float hi0 = (p.x + r) + e, hi1 = (p.y + r) + e, hi2 = (p.z + r) + e;
      if (q == first) {
        b.mn[0] = lo0; b.mn[1] = lo1; b.mn[2] = lo2;
        b.mx[0] = hi0; b.mx[1] = hi1; b.mx[2] = hi2;
      }
      b.mn[0] = fmin_ref(b.mn[0], lo0); b.mn[1] = fmin_ref(b.mn[1], lo1);
      b.mn[2] = fmin_ref(b.mn[2], lo2);
      b.mx[0] = fmax_ref(b.mx[0], hi0); b.mx[1] = fmax_ref(b.mx[1], hi1);
      b.mx[2] = fmax_ref(b.mx[2], hi2);
      reinterpret_cast<float4 *>(prims)[q] = make_float4(p.x, p.y, p.z, r);
    }
  }
}

__device__ __forceinline__ void store_box_slot(float *node_words, int slot, const Box &b) {
  // TriangleNode: bmin[slot][k] at word 3*slot+k, bmax[slot][k] at 6+3*slot+k
  for (int k = 0; k < 3; k++) {
    node_words[3 * slot + k] = b.mn[k];
    node_words[6 + 3 * slot + k] = b.mx[k];
  }
}

// MergeBoundingBox(left, right), render/render_accel_triangle.cc:804-814
__device__ __forceinline__ Box merge_lr(const Box &l, const Box &r) {
  Box m;
  for (int k = 0; k < 3; k++) {
    m.mn[k] = fmin_ref(l.mn[k], r.mn[k]);
    m.mx[k] = fmax_ref(l.mx[k], r.mx[k]);
  }
  return m;
}

template <int K>
__global__ void __launch_bounds__(256) k_leaves(PrimSrc s, const uint32_t *sidx, NodeOut o,
                                               void *prims, uint32_t *sorted_faces,
                                               BuildParams *params) {
  uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nleaves = params->num_leaves;
  if (li >= nleaves || li >= s.n) return;
  const float e = params->kEPS;
  uint32_t first = o.leaf_first[li], count = o.leaf_count[li];
  if (first >= s.n || count > s.n - first) {
    atomicOr(&params->error, 2u);
    return;
  }
  Box b;
  box_init(b);
  leaf_box_pack<K>(s, sidx, first, count, e, prims, sorted_faces, b);
  uint32_t node = o.leaf_pos[li];
  float *words = reinterpret_cast<float *>(o.nodes);
  {
    // a leaf keeps its own box in slot 0 and leaves slot 1 unused
    // (MakeLeaf32, render/render_accel_triangle.cc:865-890)
    float4 *w4 = reinterpret_cast<float4 *>(o.nodes) + 4 * (size_t)node;
    w4[0] = make_float4(b.mn[0], b.mn[1], b.mn[2], 0.0f);
    w4[1] = make_float4(0.0f, 0.0f, b.mx[0], b.mx[1]);
    w4[2] = make_float4(b.mx[2], 0.0f, 0.0f, 0.0f);
  }
  // every branch keeps both child boxes (TriangleNode / ParticleNode layout;
  // a tetra node's own box is the one its parent keeps for it)
  uint32_t link = o.parent_link[node];
  uint32_t p = link >> 1;
  if (node == 0 || node >= 2 * s.n - 1 || p >= 2 * s.n - 1) {
    atomicOr(&params->error, 4u);
    return;
  }
  store_box_slot(words + 16 * (size_t)p, (int)(link & 1u), b);
}

// One level of the bottom-up refit: every branch whose children already hold
// their boxes (deeper levels ran in earlier launches) merges them,
// MergeBoundingBox(left, right) (render/render_accel_triangle.cc:804-814),
// into its parent's slot; the root's merge is BoundingBox().
__global__ void __launch_bounds__(256) k_refit_level(uint32_t n, uint32_t level,
                                                    const uint32_t *cls, NodeOut o,
                                                    BuildParams *params) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i + 1 < n;
       i += gridDim.x * blockDim.x) {
    const uint32_t c = cls[i];
    if (!(c & 1u) || ((c >> 1) & 127u) != level) continue;
    const uint32_t pos = o.branch_pos[i];
    if (pos >= 2 * n - 1) {
      atomicOr(&params->error, 16u);
      continue;
    }
    const float4 *w4 = reinterpret_cast<const float4 *>(o.nodes) + 4 * (size_t)pos;
    const float4 a = w4[0], b = w4[1], cc = w4[2];
    Box l, r;
    l.mn[0] = a.x; l.mn[1] = a.y; l.mn[2] = a.z;
    r.mn[0] = a.w; r.mn[1] = b.x; r.mn[2] = b.y;
    l.mx[0] = b.z; l.mx[1] = b.w; l.mx[2] = cc.x;
    r.mx[0] = cc.y; r.mx[1] = cc.z; r.mx[2] = cc.w;
    const Box m = merge_lr(l, r);
    if (pos == 0) {
      for (int k = 0; k < 3; k++) {
        params->root_bmin[k] = m.mn[k];
        params->root_bmax[k] = m.mx[k];
      }
      continue;
    }
    const uint32_t link = o.parent_link[pos];
    const uint32_t p = link >> 1;
    if (p >= 2 * n - 1) {
      atomicOr(&params->error, 16u);
      continue;
    }
    store_box_slot(reinterpret_cast<float *>(o.nodes) + 16 * (size_t)p, (int)(link & 1u), m);
  }
}

inline uint32_t blocks_for(uint64_t n, int block, uint32_t cap = 0x7fffffffu) {
  uint64_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  return (uint32_t)(b < cap ? b : cap);
}

PrimSrc make_src(Kind kind, const void *desc) {
  PrimSrc s{};
  if (kind == kTri) {
    const srf_tri_desc *d = (const srf_tri_desc *)desc;
    s.V = d->vertices;
    s.F = d->faces;
    s.n = (uint32_t)d->num_faces;
  } else if (kind == kTet) {
    const srf_tet_desc *d = (const srf_tet_desc *)desc;
    s.V = d->vertices;
    s.F = d->indices;
    s.n = (uint32_t)d->num_tetras;
  } else {
    const srf_sphere_desc *d = (const srf_sphere_desc *)desc;
    s.V = d->positions;
    s.R = d->radius;
    s.rconst = d->constant_radius;
    s.n = (uint32_t)d->num_particles;
  }
  return s;
}

uint32_t sort_tiles(uint32_t n) { return (n + kSortTile - 1) / kSortTile; }

struct SortScratch {
  uint32_t *keys_b, *vals_b, *vals_a, *status, *counters, *hist;
};

// Runs the 4 onesweep passes; keys start in keys_a (vals implicit identity),
// the sorted pairs end in (keys_a, vals_a). hist/counters must be zero
// before the histogram was accumulated.
srf_status run_sort_passes(uint32_t *keys_a, uint32_t n, const SortScratch &ss,
                           hipStream_t stream) {
  const uint32_t tiles = sort_tiles(n);
  uint32_t *src_k = keys_a, *src_v = nullptr, *dst_k = ss.keys_b, *dst_v = ss.vals_b;
  for (int pass = 0; pass < 4; pass++) {
    hipLaunchKernelGGL(k_onesweep, dim3(tiles), dim3(kSortBlock), 0, stream, src_k, src_v,
                       dst_k, dst_v, n, pass * 8, ss.hist + 256 * pass,
                       ss.status + (size_t)pass * tiles * 256, ss.counters + pass,
                       ss.counters + 4);
    SRF_HIP(hipGetLastError());
    uint32_t *tk = src_k, *tv = src_v;
    src_k = dst_k;
    src_v = dst_v;
    dst_k = tk;
    dst_v = (pass == 0) ? ss.vals_a : tv;
  }
  return SRF_OK;
}

size_t sort_scratch_words(uint32_t n) {
  return 3 * (size_t)n + 4 * (size_t)sort_tiles(n) * 256 + 8 + 1024;
}

}  // namespace

// ------------------------------------------------------------------ host

srf_status scratch_bytes_for_build(uint64_t n, size_t *bytes) {
  // codes + sort scratch + Karras (child, leaf, other, parent) + cls +
  // leaf_start + leaf_rank + scan partials + per-node link/counter +
  // per-leaf pos/first/count + bounds partials, each slice 256-B aligned.
  size_t words = n + sort_scratch_words((uint32_t)n) + 4 * n + n + 2 * n +
                 (n / kScanTile + 2) + 2 * (2 * n) + 3 * n;
  *bytes = 4 * words + sizeof(Box) * kBoundsBlocks + 64 * 256;
  return SRF_OK;
}

srf_status build_lbvh(Ctx *ctx, Kind kind, const void *desc, hipStream_t stream,
                      Accel *acc) {
  const PrimSrc src = make_src(kind, desc);
  const uint32_t n = src.n;
  size_t need = 0;
  scratch_bytes_for_build(n, &need);
  srf_status st = ensure_scratch(ctx, need, stream);
  if (st != SRF_OK) return st;
  Arena ar{(char *)ctx->scratch, ctx->scratch_bytes, 0};
  uint32_t *codes = ar.take<uint32_t>(n);
  SortScratch ss;
  ss.keys_b = ar.take<uint32_t>(n);
  ss.vals_b = ar.take<uint32_t>(n);
  ss.vals_a = acc->sorted_index;  // final sorted index lands in the accel
  const uint32_t tiles = sort_tiles(n);
  ss.status = ar.take<uint32_t>((size_t)4 * tiles * 256);
  ss.counters = ar.take<uint32_t>(8);
  ss.hist = ar.take<uint32_t>(1024);
  uint32_t *kchild = ar.take<uint32_t>(n);
  uint32_t *kleaf = ar.take<uint32_t>(n);
  uint32_t *kother = ar.take<uint32_t>(n);
  uint32_t *kparent = ar.take<uint32_t>(n);
  uint32_t *cls = ar.take<uint32_t>(n);
  uint32_t *leaf_start = ar.take<uint32_t>(n);
  uint32_t *leaf_rank = ar.take<uint32_t>(n);
  const uint32_t scan_blocks = blocks_for(n, kScanTile);
  uint32_t *partial = ar.take<uint32_t>(scan_blocks + 1);
  NodeOut o;
  o.nodes = acc->nodes;
  o.parent_link = ar.take<uint32_t>(2 * (size_t)n);
  o.branch_pos = ar.take<uint32_t>(n);
  o.leaf_pos = ar.take<uint32_t>(n);
  o.leaf_first = ar.take<uint32_t>(n);
  o.leaf_count = ar.take<uint32_t>(n);
  Box *partials = ar.take<Box>(kBoundsBlocks);
  if (ar.off > ar.size) return set_error(SRF_ERR_INTERNAL, "scratch layout overflow");

  // zero the ticket (params is fully rewritten by the last bounds block)
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, stream,
                     reinterpret_cast<uint32_t *>(acc->params), 0u,
                     sizeof(BuildParams) / 4);
  // the region [status, hist] is contiguous: clear it inside k_bounds
  uint32_t *clear_begin = ss.status;
  uint32_t nclear = (uint32_t)((ss.hist + 1024) - clear_begin);
  const uint32_t bblocks = blocks_for(n, 256, kBoundsBlocks);
  const uint32_t pblocks = blocks_for(n, 256, 4096);
  const uint32_t nblocks = blocks_for(n, 256);
  int max_leaf = kind == kTri ? kTriMaxLeaf : (kind == kTet ? kTetMaxLeaf : kSphMaxLeaf);

#define SRF_DISPATCH(KERNEL, ...)                                                      \
  do {                                                                                 \
    if (kind == kTri)                                                                  \
      hipLaunchKernelGGL(KERNEL<kTri>, __VA_ARGS__);                                   \
    else if (kind == kTet)                                                             \
      hipLaunchKernelGGL(KERNEL<kTet>, __VA_ARGS__);                                   \
    else                                                                               \
      hipLaunchKernelGGL(KERNEL<kSph>, __VA_ARGS__);                                   \
    SRF_HIP(hipGetLastError());                                                        \
  } while (0)

  SRF_DISPATCH(k_bounds, dim3(bblocks), dim3(256), 0, stream, src, partials, acc->params,
               clear_begin, nclear);
  SRF_DISPATCH(k_morton, dim3(pblocks), dim3(256), 0, stream, src, acc->params, codes,
               ss.hist);
  st = run_sort_passes(codes, n, ss, stream);
  if (st != SRF_OK) return st;
  // sorted codes in `codes`, sorted index in acc->sorted_index
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256, 2048)), dim3(256), 0, stream,
                     leaf_start, 0u, (size_t)n);
  hipLaunchKernelGGL(k_karras, dim3(nblocks), dim3(256), 0, stream, codes, n, kchild, kleaf,
                     kother, kparent);
  hipLaunchKernelGGL(k_classify, dim3(nblocks), dim3(256), 0, stream, n, max_leaf, kchild,
                     kleaf, kother, kparent, cls, leaf_start, acc->params);
  hipLaunchKernelGGL(k_scan_reduce, dim3(scan_blocks), dim3(kScanBlock), 0, stream,
                     leaf_start, n, partial);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, stream, partial, scan_blocks);
  hipLaunchKernelGGL(k_scan_downsweep, dim3(scan_blocks), dim3(kScanBlock), 0, stream,
                     leaf_start, n, partial, leaf_rank, acc->params);
  SRF_HIP(hipGetLastError());
  SRF_DISPATCH(k_emit, dim3(nblocks), dim3(256), 0, stream, n, max_leaf, codes, kchild, kleaf,
               kother, cls, leaf_rank, o);
  SRF_DISPATCH(k_leaves, dim3(nblocks), dim3(256), 0, stream, src, acc->sorted_index, o,
               acc->prims, acc->sorted_faces, acc->params);
  // bottom-up refit, one level per launch (levels are branch steps 23..0)
  const uint32_t rblocks = blocks_for(n, 256, 2048);
  for (int level = kMaxBranchLevels - 1; level >= 0; level--) {
    hipLaunchKernelGGL(k_refit_level, dim3(rblocks), dim3(256), 0, stream, n, (uint32_t)level,
                       cls, o, acc->params);
  }
  SRF_HIP(hipGetLastError());
#undef SRF_DISPATCH
  return SRF_OK;
}

srf_status launch_morton(Kind kind, const void *desc, const float bmin[3],
                         const float bmax[3], uint32_t *codes, hipStream_t stream) {
  PrimSrc s = make_src(kind, desc);
  BuildParams p{};
  for (int k = 0; k < 3; k++) {
    p.bmin[k] = bmin[k];
    p.bmax[k] = bmax[k];
    p.inv[k] = (float)1024 / (bmax[k] - bmin[k]);
  }
  uint32_t blocks = blocks_for(s.n, 256, 4096);
  if (kind == kTri)
    hipLaunchKernelGGL(k_morton_box<kTri>, dim3(blocks), dim3(256), 0, stream, s, p, codes);
  else if (kind == kTet)
    hipLaunchKernelGGL(k_morton_box<kTet>, dim3(blocks), dim3(256), 0, stream, s, p, codes);
  else
    hipLaunchKernelGGL(k_morton_box<kSph>, dim3(blocks), dim3(256), 0, stream, s, p, codes);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

srf_status sort_pairs30(Ctx *ctx, const uint32_t *codes_in, uint32_t n, uint32_t *codes_out,
                        uint32_t *index_out, hipStream_t stream) {
  size_t need = 4 * (2 * (size_t)n + sort_scratch_words(n)) + 8 * 256;
  srf_status st = ensure_scratch(ctx, need, stream);
  if (st != SRF_OK) return st;
  Arena ar{(char *)ctx->scratch, ctx->scratch_bytes, 0};
  SortScratch ss;
  ss.keys_b = ar.take<uint32_t>(n);
  ss.vals_b = ar.take<uint32_t>(n);
  ss.vals_a = index_out;
  const uint32_t tiles = sort_tiles(n);
  ss.status = ar.take<uint32_t>((size_t)4 * tiles * 256);
  ss.counters = ar.take<uint32_t>(8);
  ss.hist = ar.take<uint32_t>(1024);
  if (ar.off > ar.size) return set_error(SRF_ERR_INTERNAL, "scratch layout overflow");
  size_t nclear = (size_t)((ss.hist + 1024) - ss.status);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(nclear, 256, 2048)), dim3(256), 0, stream,
                     ss.status, 0u, nclear);
  SRF_HIP(hipMemcpyAsync(codes_out, codes_in, 4 * (size_t)n, hipMemcpyDeviceToDevice, stream));
  hipLaunchKernelGGL(k_histogram, dim3(blocks_for(n, 256, 4096)), dim3(256), 0, stream,
                     codes_out, n, ss.hist);
  SRF_HIP(hipGetLastError());
  return run_sort_passes(codes_out, n, ss, stream);
}

srf_status launch_radix_tree(const uint32_t *sorted_codes, uint32_t n, srf_node_info *out,
                             hipStream_t stream) {
  hipLaunchKernelGGL(k_karras_nodeinfo, dim3(blocks_for(n, 256)), dim3(256), 0, stream,
                     sorted_codes, n, out);
  SRF_HIP(hipGetLastError());
  return SRF_OK;
}

}  // namespace srf
