/*
 * srf_oracle.c -- CPU restatement of SURFACE's acceleration-structure build,
 * closest-hit traversal and 3D-texture sampling.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path. The product
 * never links or calls it; tests/, the smoke check and the bench's
 * cpu_baseline leg load it through oracle/oracle.py.
 *
 * Each function cites the reference file:line whose arithmetic it follows
 * (paths relative to the SURFACE checkout). The arithmetic is restated
 * operation by operation, with the same float/double mixing and evaluation
 * order, and is compiled with -ffp-contract=off and without -ffast-math, as
 * the reference's x86-64 Release build (no -march, hence no FMA) is.
 *
 * Pins (see tests/test_oracle_*.py): the known-answer Morton vector of
 * test/cctest/test-prefix-tree.cc:27-41, the Karras invariants of
 * test-prefix-tree.cc:418-510,622-625, byte comparison with trees and hits
 * produced by the reference itself (oracle/_ref, built from the reference
 * sources by oracle/Makefile) and committed golden fixtures, brute-force
 * closest hit over all primitives, analytic sphere/tetra/trilinear cases.
 * The emission-absorption volume march has no reference counterpart (the
 * reference leaves marching to user GLSL); its f64 adjoint is pinned by
 * central finite differences of the f64 forward.
 */
#include "srf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TRI_MAX_LEAF 16 /* render/render_accel_triangle.cc:34 */
#define TET_MAX_LEAF 4  /* render/render_accel_tetra.cc:34 */
#define SPH_MAX_LEAF 16 /* render/render_accel_particle.cc:51 */
#define MAX_DEPTH32 22  /* render/render_accel_triangle.cc:35 */
#define STACK_MAX 512   /* render/render_accel_triangle.cc:1723 */

/* ------------------------------------------------------------------------ */
/* Morton codes                                                             */
/* ------------------------------------------------------------------------ */

/* PartBy2_30, render/render_prefix_tree_util.h:113-120 */
static uint32_t part_by2_30(uint32_t n) {
  n &= 0x000003ffu;
  n = (n ^ (n << 16)) & 0xff0000ffu;
  n = (n ^ (n << 8)) & 0x0300f00fu;
  n = (n ^ (n << 4)) & 0x030c30c3u;
  n = (n ^ (n << 2)) & 0x09249249u;
  return n;
}

/* float -> uint32_t as x86-64 g++ emits it (cvttss2si to 64 bit, keep the
 * low word); identical to a plain truncation for |x| < 2^31. */
static uint32_t f2u_trunc(float x) { return (uint32_t)(int64_t)x; }

/* MortionCode30, render/render_prefix_tree_util.h:143-157 */
uint32_t or_morton30_point(const float p[3], const float bmin[3], float invx,
                           float invy, float invz) {
  uint32_t ix = f2u_trunc((p[0] - bmin[0]) * invx);
  uint32_t iy = f2u_trunc((p[1] - bmin[1]) * invy);
  uint32_t iz = f2u_trunc((p[2] - bmin[2]) * invz);
  return (part_by2_30(ix) << 2) | (part_by2_30(iy) << 1) | part_by2_30(iz);
}

/* kDIV / (bmax - bmin): int over float, i.e. a float division.
 * render/render_prefix_tree_util.cc:242-245 */
static void morton_inv(const float bmin[3], const float bmax[3], float inv[3]) {
  for (int k = 0; k < 3; k++) inv[k] = (float)1024 / (bmax[k] - bmin[k]);
}

/* CalculateMortonCodesTriangleFloat30, render/render_prefix_tree_util.cc:230-265 */
void or_morton_tri(const float *V, const uint32_t *F, int64_t n,
                   const float bmin[3], const float bmax[3], uint32_t *codes) {
  float inv[3];
  morton_inv(bmin, bmax, inv);
  const float one_third = 1.0f / 3.0f;
  for (int64_t i = 0; i < n; i++) {
    const float *p0 = V + 3 * (size_t)F[3 * i + 0];
    const float *p1 = V + 3 * (size_t)F[3 * i + 1];
    const float *p2 = V + 3 * (size_t)F[3 * i + 2];
    float c[3];
    for (int k = 0; k < 3; k++) c[k] = one_third * ((p0[k] + p1[k]) + p2[k]);
    codes[i] = or_morton30_point(c, bmin, inv[0], inv[1], inv[2]);
  }
}

/* CalculateMortonCodesTetraFloat30, render/render_prefix_tree_util.cc:267-304 */
void or_morton_tet(const float *V, const uint32_t *F, int64_t n,
                   const float bmin[3], const float bmax[3], uint32_t *codes) {
  float inv[3];
  morton_inv(bmin, bmax, inv);
  const float one_fourth = 1.0f / 4.0f;
  for (int64_t i = 0; i < n; i++) {
    const float *p0 = V + 3 * (size_t)F[4 * i + 0];
    const float *p1 = V + 3 * (size_t)F[4 * i + 1];
    const float *p2 = V + 3 * (size_t)F[4 * i + 2];
    const float *p3 = V + 3 * (size_t)F[4 * i + 3];
    float c[3];
    for (int k = 0; k < 3; k++)
      c[k] = one_fourth * (((p0[k] + p1[k]) + p2[k]) + p3[k]);
    codes[i] = or_morton30_point(c, bmin, inv[0], inv[1], inv[2]);
  }
}

/* CalculateMortonCodesSolidFloat30, render/render_prefix_tree_util.cc:347-382:
 * the centroid is the float sum of the solid's vertices from (0, 0, 0), in
 * index order, times the float reciprocal of the vertex count. */
void or_morton_sol(const float *V, const uint32_t *F, int64_t n, int nv,
                   const float bmin[3], const float bmax[3], uint32_t *codes) {
  float inv[3];
  morton_inv(bmin, bmax, inv);
  const float inv_div = 1.0f / (float)nv;
  for (int64_t i = 0; i < n; i++) {
    float c[3] = {0.0f, 0.0f, 0.0f};
    for (int j = 0; j < nv; j++) {
      const float *p = V + 3 * (size_t)F[(size_t)nv * i + j];
      for (int k = 0; k < 3; k++) c[k] = c[k] + p[k];
    }
    for (int k = 0; k < 3; k++) c[k] *= inv_div;
    codes[i] = or_morton30_point(c, bmin, inv[0], inv[1], inv[2]);
  }
}

/* CalculateMortonCodes30SIMD, render/render_prefix_tree_util.cc:176-228: the
 * SSE2 lanes do the same sub/mul/truncate as the scalar MortionCode30, and
 * the missing 10-bit mask of vpartby2 only matters for ix > 1024, which the
 * scene box (radius + eps padded) rules out. */
void or_morton_pts(const float *P, int64_t n, const float bmin[3],
                   const float bmax[3], uint32_t *codes) {
  float inv[3];
  morton_inv(bmin, bmax, inv);
  for (int64_t i = 0; i < n; i++)
    codes[i] = or_morton30_point(P + 3 * i, bmin, inv[0], inv[1], inv[2]);
}

/* ------------------------------------------------------------------------ */
/* Scene bounds                                                             */
/* ------------------------------------------------------------------------ */

/* ComputeBoundingBoxOMP<float> with epsScale = 0 (kEPS = 0):
 * render/render_accel_triangle.cc:711-797 called at :1438 with :1430. */
void or_scene_bounds_tri(const float *V, const uint32_t *F, int64_t n,
                         float bmin[3], float bmax[3]) {
  const float kEPS = FLT_EPSILON * 0.0f;
  const float *v0 = V + 3 * (size_t)F[0];
  for (int k = 0; k < 3; k++) {
    bmin[k] = v0[k] - kEPS;
    bmax[k] = v0[k] + kEPS;
  }
  for (int64_t i = 0; i < n; i++)
    for (int j = 0; j < 3; j++) {
      const float *v = V + 3 * (size_t)F[3 * i + j];
      for (int k = 0; k < 3; k++) {
        float mn = v[k] - kEPS, mx = v[k] + kEPS;
        if (mn < bmin[k]) bmin[k] = mn; /* std::min(a, b) keeps a unless b < a */
        if (mx > bmax[k]) bmax[k] = mx;
      }
    }
}

/* ComputeBoundingBoxOMP (tetra), render/render_accel_tetra.cc:971-1055:
 * kEPS = FLT_EPSILON * 16. */
void or_scene_bounds_tet(const float *V, const uint32_t *F, int64_t n,
                         float bmin[3], float bmax[3]) {
  const float kEPS = FLT_EPSILON * 16;
  const float *v0 = V + 3 * (size_t)F[0];
  for (int k = 0; k < 3; k++) {
    bmin[k] = v0[k] - kEPS;
    bmax[k] = v0[k] + kEPS;
  }
  for (int64_t i = 0; i < n; i++)
    for (int j = 0; j < 4; j++) {
      const float *v = V + 3 * (size_t)F[4 * i + j];
      for (int k = 0; k < 3; k++) {
        float mn = v[k] - kEPS, mx = v[k] + kEPS;
        if (mn < bmin[k]) bmin[k] = mn;
        if (mx > bmax[k]) bmax[k] = mx;
      }
    }
}

/* ComputeBoundingBoxOMP (solid), render/render_accel_solid.cc:1172-1262:
 * every vertex of every solid -+ 16 * FLT_EPSILON. */
void or_scene_bounds_sol(const float *V, const uint32_t *F, int64_t n, int nv,
                         float bmin[3], float bmax[3]) {
  const float kEPS = FLT_EPSILON * 16;
  const float *v0 = V + 3 * (size_t)F[0];
  for (int k = 0; k < 3; k++) {
    bmin[k] = v0[k] - kEPS;
    bmax[k] = v0[k] + kEPS;
  }
  for (int64_t i = 0; i < n; i++)
    for (int j = 0; j < nv; j++) {
      const float *v = V + 3 * (size_t)F[(size_t)nv * i + j];
      for (int k = 0; k < 3; k++) {
        float mn = v[k] - kEPS, mx = v[k] + kEPS;
        if (mn < bmin[k]) bmin[k] = mn;
        if (mx > bmax[k]) bmax[k] = mx;
      }
    }
}

/* GetRadius, render/render_accel_particle.cc:272-278 */
static float sph_radius(const float *R, float rconst, int64_t i) {
  return R ? R[i] : rconst;
}

/* ComputeBoundingBoxOMP (particles), render/render_accel_particle.cc:303-378:
 * pos - radius - kEPS with kEPS = (float)(FLT_EPSILON * 16.0). */
void or_scene_bounds_sph(const float *P, const float *R, float rconst,
                         int64_t n, float bmin[3], float bmax[3]) {
  const float kEPS = (float)(FLT_EPSILON * 16.0);
  float r = sph_radius(R, rconst, 0);
  for (int k = 0; k < 3; k++) {
    bmin[k] = (P[k] - r) - kEPS;
    bmax[k] = (P[k] + r) + kEPS;
  }
  for (int64_t i = 0; i < n; i++) {
    r = sph_radius(R, rconst, i);
    for (int k = 0; k < 3; k++) {
      float mn = (P[3 * i + k] - r) - kEPS, mx = (P[3 * i + k] + r) + kEPS;
      if (mn < bmin[k]) bmin[k] = mn;
      if (mx > bmax[k]) bmax[k] = mx;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Stable key sort and Karras radix tree                                    */
/* ------------------------------------------------------------------------ */

/* RadixSort30 / RadixSort30OMP, render/render_prefix_tree_util.h:296-457:
 * four stable 8-bit LSD passes over {index, code}; the permutation equals a
 * stable sort by code of the identity index sequence. */
void or_sort_keys(const uint32_t *codes, int64_t n, uint32_t *sorted_codes,
                  uint32_t *sorted_index) {
  uint32_t *kc = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
  uint32_t *ki = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
  uint32_t *ac = sorted_codes, *ai = sorted_index;
  for (int64_t i = 0; i < n; i++) {
    ac[i] = codes[i];
    ai[i] = (uint32_t)i;
  }
  uint32_t *src_c = ac, *src_i = ai, *dst_c = kc, *dst_i = ki;
  for (unsigned shift = 0; shift < 32; shift += 8) {
    size_t count[257];
    memset(count, 0, sizeof(count));
    for (int64_t i = 0; i < n; i++) count[((src_c[i] >> shift) & 0xff) + 1]++;
    for (int b = 0; b < 256; b++) count[b + 1] += count[b];
    for (int64_t i = 0; i < n; i++) {
      size_t d = count[(src_c[i] >> shift) & 0xff]++;
      dst_c[d] = src_c[i];
      dst_i[d] = src_i[i];
    }
    uint32_t *t;
    t = src_c; src_c = dst_c; dst_c = t;
    t = src_i; src_i = dst_i; dst_i = t;
  }
  /* four passes: the result is back in the caller's arrays */
  free(kc);
  free(ki);
}

/* CountLeadingZeros32, render/render_prefix_tree_util.cc:447-483 */
static int clz32(uint32_t x) {
  if (x == 0) return 32;
  int n = 0;
  while (!(x & 0x80000000u)) {
    n++;
    x <<= 1;
  }
  return n;
}

/* Delta30, render/render_prefix_tree_util.cc:501-526: equal codes fall back to
 * comparing the sorted positions, offset by 32. */
static int delta30(const uint32_t *c, int32_t i, int32_t j, int32_t n) {
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

/* ConstructBinaryRadixTree30, render/render_prefix_tree_util.cc:562-649
 * (Karras 2012, Fig. 4). */
static or_nodeinfo radix_node30(const uint32_t *c, int i, int n, int *jout) {
  int dp = delta30(c, i, i + 1, n) - delta30(c, i, i - 1, n);
  int d = dp >= 0 ? 1 : -1;
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
  int t = (l + k - 1) / k; /* (int)ceil((double)l / k) for l >= 0 */
  if (t > 0) {
    for (;;) {
      if (delta30(c, i, i + (s + t) * d, n) > dnode) s += t;
      if (t == 1) break;
      k *= 2;
      t = (l + k - 1) / k;
    }
  }
  int r = i + s * d + (d < 0 ? d : 0);
  or_nodeinfo ni;
  ni.child = (uint32_t)r;
  ni.left_type = ((i < j ? i : j) == r) ? 1 : 0;
  ni.right_type = ((i > j ? i : j) == r + 1) ? 1 : 0;
  if (jout) *jout = j;
  return ni;
}

void or_radix_tree30(const uint32_t *sorted_codes, int64_t n,
                     or_nodeinfo *out) {
  for (int64_t i = 0; i + 1 < n; i++)
    out[i] = radix_node30(sorted_codes, (int)i, (int)n, NULL);
}

/* GetSplitAxis, render/render_accel_triangle.cc:854-863 (asserts compiled out:
 * code 0 gives (32 - 2) % 3 = 0). */
static int split_axis(uint32_t code) { return (clz32(code) - 2) % 3; }

/* ------------------------------------------------------------------------ */
/* Collapse + boxes: BuildTreeRecursive32 (serial), which emits DFS preorder */
/* ------------------------------------------------------------------------ */

typedef struct {
  int kind; /* 0 tri, 1 tet, 2 sphere, 3 solid (pyramid / prism / hexahedron), 4 line */
  int nv;   /* solids: vertices per solid (5, 6, 8) */
  const float *V;
  const uint32_t *F;
  const float *R;
  float rconst;
  float kEPS;
  const uint32_t *sidx;  /* sorted primitive index */
  const uint32_t *scode; /* sorted codes */
  const or_nodeinfo *info;
  int max_leaf;
  void *nodes;
  int64_t count;
} build_ctx;

/* Leaf box over prims [l, r) in sorted order.
 * tri: ComputeBoundingBox30 render/render_accel_triangle.cc:660-709
 * tet: ComputeBoundingBox30 render/render_accel_tetra.cc:921-969
 * sph: ComputeBoundingBox30 render/render_accel_particle.cc:424-462 */
static void lin_box(const float *V, const uint32_t *F, const float *R, float rconst, uint32_t i,
                    float mn[3], float mx[3]);

static void leaf_box(const build_ctx *b, uint32_t l, uint32_t r, float mn[3],
                     float mx[3]) {
  const float e = b->kEPS;
  if (b->kind == 4) {
    lin_box(b->V, b->F, b->R, b->rconst, b->sidx[l], mn, mx);
    for (uint32_t i = l; i < r; i++) {
      float lo[3], hi[3];
      lin_box(b->V, b->F, b->R, b->rconst, b->sidx[i], lo, hi);
      for (int k = 0; k < 3; k++) {
        if (mn[k] > lo[k]) mn[k] = lo[k];
        if (mx[k] < hi[k]) mx[k] = hi[k];
      }
    }
    return;
  }
  if (b->kind == 2) {
    uint32_t idx = b->sidx[l];
    float rad = sph_radius(b->R, b->rconst, idx);
    for (int k = 0; k < 3; k++) {
      mn[k] = (b->V[3 * idx + k] - rad) - e;
      mx[k] = (b->V[3 * idx + k] + rad) + e;
    }
    for (uint32_t i = l; i < r; i++) {
      idx = b->sidx[i];
      rad = sph_radius(b->R, b->rconst, idx);
      for (int k = 0; k < 3; k++) {
        float lo = (b->V[3 * (size_t)idx + k] - rad) - e;
        float hi = (b->V[3 * (size_t)idx + k] + rad) + e;
        if (mn[k] > lo) mn[k] = lo;
        if (mx[k] < hi) mx[k] = hi;
      }
    }
    return;
  }
  const int nv = b->kind == 0 ? 3 : (b->kind == 1 ? 4 : b->nv);
  uint32_t idx = b->sidx[l];
  const float *v0 = b->V + 3 * (size_t)b->F[(size_t)nv * idx];
  for (int k = 0; k < 3; k++) {
    mn[k] = v0[k] - e;
    mx[k] = v0[k] + e;
  }
  for (uint32_t i = l; i < r; i++) {
    idx = b->sidx[i];
    for (int j = 0; j < nv; j++) {
      const float *v = b->V + 3 * (size_t)b->F[(size_t)nv * idx + j];
      for (int k = 0; k < 3; k++) {
        float lo = v[k] - e, hi = v[k] + e;
        if (mn[k] > lo) mn[k] = lo;
        if (mx[k] < hi) mx[k] = hi;
      }
    }
  }
}

/* MergeBoundingBox, render/render_accel_triangle.cc:804-814 */
static void merge_box(float mn[3], float mx[3], const float lmn[3],
                      const float lmx[3], const float rmn[3],
                      const float rmx[3]) {
  for (int k = 0; k < 3; k++) {
    mn[k] = (rmn[k] < lmn[k]) ? rmn[k] : lmn[k];
    mx[k] = (lmx[k] < rmx[k]) ? rmx[k] : lmx[k];
  }
}

static int64_t push_node(build_ctx *b) { return b->count++; }

static void set_leaf(build_ctx *b, int64_t at, uint32_t l, uint32_t end,
                     const float mn[3], const float mx[3]) {
  if (b->kind == 1 || b->kind == 3 || b->kind == 4) {
    or_tet_node *nd = (or_tet_node *)b->nodes + at;
    memset(nd, 0, sizeof(*nd));
    memcpy(nd->bmin, mn, 12);
    memcpy(nd->bmax, mx, 12);
    nd->flag = 1;
    nd->data[0] = end - l;
    nd->data[1] = l;
  } else if (b->kind == 0) {
    or_tri_node *nd = (or_tri_node *)b->nodes + at;
    memset(nd, 0, sizeof(*nd));
    memcpy(nd->bmin[0], mn, 12);
    memcpy(nd->bmax[0], mx, 12);
    nd->flag = 1;
    nd->data[0] = end - l;
    nd->data[1] = l;
  } else {
    or_sph_node *nd = (or_sph_node *)b->nodes + at;
    memset(nd, 0, sizeof(*nd));
    memcpy(nd->bmin[0], mn, 12);
    memcpy(nd->bmax[0], mx, 12);
    nd->flag = 1;
    nd->data[0] = end - l;
    nd->data[1] = l;
  }
}

static void set_branch(build_ctx *b, int64_t at, int axis, int64_t lc,
                       int64_t rc, const float lmn[3], const float lmx[3],
                       const float rmn[3], const float rmx[3],
                       const float mn[3], const float mx[3]) {
  if (b->kind == 1 || b->kind == 3 || b->kind == 4) {
    /* TetraNode keeps its own (merged) box: render/render_accel_tetra.cc:1225-1235;
     * SolidNode has the same 40-byte layout and rule,
     * render/render_accel_solid.h:24-43, render/render_accel_solid.cc:1436-1450 */
    or_tet_node *nd = (or_tet_node *)b->nodes + at;
    memset(nd, 0, sizeof(*nd));
    memcpy(nd->bmin, mn, 12);
    memcpy(nd->bmax, mx, 12);
    nd->flag = 0;
    nd->axis = axis;
    nd->data[0] = (uint32_t)lc;
    nd->data[1] = (uint32_t)rc;
  } else if (b->kind == 0) {
    or_tri_node *nd = (or_tri_node *)b->nodes + at;
    memset(nd, 0, sizeof(*nd));
    memcpy(nd->bmin[0], lmn, 12);
    memcpy(nd->bmax[0], lmx, 12);
    memcpy(nd->bmin[1], rmn, 12);
    memcpy(nd->bmax[1], rmx, 12);
    nd->flag = 0;
    nd->axis = axis;
    nd->data[0] = (uint32_t)lc;
    nd->data[1] = (uint32_t)rc;
  } else {
    or_sph_node *nd = (or_sph_node *)b->nodes + at;
    memset(nd, 0, sizeof(*nd));
    memcpy(nd->bmin[0], lmn, 12);
    memcpy(nd->bmax[0], lmx, 12);
    memcpy(nd->bmin[1], rmn, 12);
    memcpy(nd->bmax[1], rmx, 12);
    nd->flag = 0;
    nd->axis = (int16_t)axis;
    nd->data[0] = (uint32_t)lc;
    nd->data[1] = (uint32_t)rc;
  }
}

/* BuildTreeRecursive32, render/render_accel_triangle.cc:895-992 (tetra
 * render/render_accel_tetra.cc:1154-1240, particle
 * render/render_accel_particle.cc:559-661). Range [l, r] is inclusive. */
static int64_t build_rec(build_ctx *b, uint32_t root, uint32_t l, uint32_t r,
                         int is_leaf, int depth, float mn[3], float mx[3]) {
  uint32_t n = r - l;
  if (is_leaf || n <= (uint32_t)b->max_leaf || depth > MAX_DEPTH32) {
    leaf_box(b, l, r + 1, mn, mx);
    int64_t at = push_node(b);
    set_leaf(b, at, l, r + 1, mn, mx);
    return at;
  }
  uint32_t mid = b->info[root].child;
  int ll = b->info[root].left_type == 1, rl = b->info[root].right_type == 1;
  int axis = split_axis(b->scode[root]);
  int64_t at = push_node(b);
  float lmn[3], lmx[3], rmn[3], rmx[3];
  int64_t lc = build_rec(b, mid, l, mid, ll, depth + 1, lmn, lmx);
  int64_t rc = build_rec(b, mid + 1, mid + 1, r, rl, depth + 1, rmn, rmx);
  merge_box(mn, mx, lmn, lmx, rmn, rmx);
  set_branch(b, at, axis, lc, rc, lmn, lmx, rmn, rmx, mn, mx);
  return at;
}

/* Root handling of Build32: the root is always a branch over [0, n-1] whose
 * children start at depth 0. render/render_accel_triangle.cc:1547-1642 */
static int64_t build_common(build_ctx *b, int64_t n, double bounds[6]) {
  uint32_t *codes = (uint32_t *)malloc(4 * (size_t)n);
  uint32_t *scode = (uint32_t *)malloc(4 * (size_t)n);
  or_nodeinfo *info = (or_nodeinfo *)malloc(sizeof(or_nodeinfo) * (size_t)n);
  float bmin[3], bmax[3];
  if (b->kind == 0) {
    or_scene_bounds_tri(b->V, b->F, n, bmin, bmax);
    or_morton_tri(b->V, b->F, n, bmin, bmax, codes);
  } else if (b->kind == 1) {
    or_scene_bounds_tet(b->V, b->F, n, bmin, bmax);
    or_morton_tet(b->V, b->F, n, bmin, bmax, codes);
  } else if (b->kind == 3) {
    or_scene_bounds_sol(b->V, b->F, n, b->nv, bmin, bmax);
    or_morton_sol(b->V, b->F, n, b->nv, bmin, bmax, codes);
  } else if (b->kind == 4) {
    or_scene_bounds_lin(b->V, b->F, b->R, b->rconst, n, bmin, bmax);
    or_morton_lin(b->V, b->F, n, bmin, bmax, codes);
  } else {
    or_scene_bounds_sph(b->V, b->R, b->rconst, n, bmin, bmax);
    or_morton_pts(b->V, n, bmin, bmax, codes);
  }
  or_sort_keys(codes, n, scode, (uint32_t *)b->sidx);
  or_radix_tree30(scode, n, info);
  b->scode = scode;
  b->info = info;
  b->count = 0;
  int64_t root = push_node(b);
  uint32_t mid = info[0].child;
  float lmn[3], lmx[3], rmn[3], rmx[3], mn[3], mx[3];
  int64_t lc = build_rec(b, mid, 0, mid, info[0].left_type == 1, 0, lmn, lmx);
  int64_t rc = build_rec(b, mid + 1, mid + 1, (uint32_t)(n - 1),
                         info[0].right_type == 1, 0, rmn, rmx);
  merge_box(mn, mx, lmn, lmx, rmn, rmx);
  set_branch(b, root, split_axis(scode[0]), lc, rc, lmn, lmx, rmn, rmx, mn, mx);
  for (int k = 0; k < 3; k++) {
    bounds[k] = mn[k];
    bounds[3 + k] = mx[k];
  }
  free(codes);
  free(scode);
  free(info);
  return b->count;
}

/* TriangleAccel::Build32, render/render_accel_triangle.cc:1389-1690 */
int64_t or_build_tri(const float *V, const uint32_t *F, int64_t n,
                     or_tri_node *nodes, uint32_t *indices, double bounds[6],
                     double *eps_scale) {
  float bmin[3], bmax[3];
  or_scene_bounds_tri(V, F, n, bmin, bmax);
  /* epsScale = max extent, each extent a float subtraction held in double:
   * render/render_accel_triangle.cc:1451-1463 */
  double bs[3];
  for (int k = 0; k < 3; k++) bs[k] = (double)(bmax[k] - bmin[k]);
  double es = bs[0];
  if (es < bs[1]) es = bs[1];
  if (es < bs[2]) es = bs[2];
  if (eps_scale) *eps_scale = es;
  build_ctx b;
  memset(&b, 0, sizeof(b));
  b.kind = 0;
  b.V = V;
  b.F = F;
  b.kEPS = FLT_EPSILON * (float)es; /* ComputeBoundingBox30 real kEPS */
  b.sidx = indices;
  b.max_leaf = TRI_MAX_LEAF;
  b.nodes = nodes;
  return build_common(&b, n, bounds);
}

/* TetraAccel::Build32, render/render_accel_tetra.cc:1579-1898 */
int64_t or_build_tet(const float *V, const uint32_t *F, int64_t n,
                     or_tet_node *nodes, uint32_t *indices, double bounds[6]) {
  build_ctx b;
  memset(&b, 0, sizeof(b));
  b.kind = 1;
  b.V = V;
  b.F = F;
  b.kEPS = FLT_EPSILON * 16;
  b.sidx = indices;
  b.max_leaf = TET_MAX_LEAF;
  b.nodes = nodes;
  return build_common(&b, n, bounds);
}

/* ParticleAccel::Build32, render/render_accel_particle.cc:1560-1866 */
int64_t or_build_sph(const float *P, const float *R, float rconst, int64_t n,
                     or_sph_node *nodes, uint32_t *indices, double bounds[6]) {
  build_ctx b;
  memset(&b, 0, sizeof(b));
  b.kind = 2;
  b.V = P;
  b.R = R;
  b.rconst = rconst;
  b.kEPS = (float)(FLT_EPSILON * 16.0);
  b.sidx = indices;
  b.max_leaf = SPH_MAX_LEAF;
  b.nodes = nodes;
  return build_common(&b, n, bounds);
}

/* SolidAccel::Build32, render/render_accel_solid.cc:1795-2124: the tetra
 * sequence (MAX_LEAF_ELEMENTS 4 and depth cap 22, :34-36; leaf boxes of
 * vertices -+ 16 * FLT_EPSILON, ComputeBoundingBox30 :1120-1169) over
 * solids of nv = 5 (pyramid), 6 (prism) or 8 (hexahedron) vertices. */
int64_t or_build_sol(const float *V, const uint32_t *F, int64_t n, int nv,
                     or_tet_node *nodes, uint32_t *indices, double bounds[6]) {
  build_ctx b;
  memset(&b, 0, sizeof(b));
  b.kind = 3;
  b.nv = nv;
  b.V = V;
  b.F = F;
  b.kEPS = FLT_EPSILON * 16;
  b.sidx = indices;
  b.max_leaf = TET_MAX_LEAF;
  b.nodes = nodes;
  return build_common(&b, n, bounds);
}

/* ------------------------------------------------------------------------ */
/* Traversal                                                                */
/* ------------------------------------------------------------------------ */

typedef struct {
  float o[3], d[3], inv[3];
  int sign[3];
} ray_setup;

/* Traverse prologue, render/render_accel_triangle.cc:2523-2538:
 * rayInvDir = (float)(1.0 / dir) (no safe invert). */
static void setup_ray(const or_ray *r, ray_setup *s) {
  for (int k = 0; k < 3; k++) {
    s->o[k] = r->org[k];
    s->d[k] = r->dir[k];
    s->sign[k] = r->dir[k] < 0.0 ? 1 : 0;
    s->inv[k] = (float)(1.0 / (double)r->dir[k]);
  }
}

/* IntersectRayAABB, render/render_accel_triangle.cc:1815-1861 (two-box nodes:
 * tmax clamped by maxT). clamp_tmax = 0 selects the tetra variant
 * render/render_accel_tetra.cc:1934-1978 (tmin <= maxT instead). */
static int ray_aabb(const ray_setup *s, float maxT, const float bmin[3],
                    const float bmax[3], int clamp_tmax) {
  float mnx = s->sign[0] ? bmax[0] : bmin[0];
  float mny = s->sign[1] ? bmax[1] : bmin[1];
  float mnz = s->sign[2] ? bmax[2] : bmin[2];
  float mxx = s->sign[0] ? bmin[0] : bmax[0];
  float mxy = s->sign[1] ? bmin[1] : bmax[1];
  float mxz = s->sign[2] ? bmin[2] : bmax[2];
  float tminx = (mnx - s->o[0]) * s->inv[0];
  float tmaxx = (mxx - s->o[0]) * s->inv[0];
  float tminy = (mny - s->o[1]) * s->inv[1];
  float tmaxy = (mxy - s->o[1]) * s->inv[1];
  float tmin = (tminx > tminy) ? tminx : tminy;
  float tmax = (tmaxx < tmaxy) ? tmaxx : tmaxy;
  float tminz = (mnz - s->o[2]) * s->inv[2];
  float tmaxz = (mxz - s->o[2]) * s->inv[2];
  tmin = (tmin > tminz) ? tmin : tminz;
  tmax = (tmax < tmaxz) ? tmax : tmaxz;
  if (clamp_tmax) {
    tmax = (maxT < tmax) ? maxT : tmax; /* std::min(tmax, maxT) */
    return (tmax > 0.0f) && (tmin <= tmax);
  }
  return (tmax > 0.0f) && (tmin <= tmax) && (tmin <= maxT);
}

/* TriangleIsect (Moller-Trumbore), render/render_accel_triangle.cc:1887-1938 */
static int tri_isect(float *tio, float *uo, float *vo, const float *p0,
                     const float *p1, const float *p2, const ray_setup *s,
                     int double_sided, float eps_scale) {
  const float kEPS = FLT_EPSILON * eps_scale;
  float e1[3], e2[3], p[3], q[3], sv[3];
  for (int k = 0; k < 3; k++) {
    e1[k] = p1[k] - p0[k];
    e2[k] = p2[k] - p0[k];
  }
  const float *d = s->d;
  p[0] = d[1] * e2[2] - d[2] * e2[1];
  p[1] = d[2] * e2[0] - d[0] * e2[2];
  p[2] = d[0] * e2[1] - d[1] * e2[0];
  float det = (e1[0] * p[0] + e1[1] * p[1]) + e1[2] * p[2];
  if (double_sided) {
    if (fabsf(det) < kEPS) return 0;
  } else {
    if (det < kEPS) return 0;
  }
  float invDet = (float)(1.0 / (double)det);
  for (int k = 0; k < 3; k++) sv[k] = s->o[k] - p0[k];
  q[0] = sv[1] * e1[2] - sv[2] * e1[1];
  q[1] = sv[2] * e1[0] - sv[0] * e1[2];
  q[2] = sv[0] * e1[1] - sv[1] * e1[0];
  float u = ((sv[0] * p[0] + sv[1] * p[1]) + sv[2] * p[2]) * invDet;
  float v = ((q[0] * d[0] + q[1] * d[1]) + q[2] * d[2]) * invDet;
  float t = ((e2[0] * q[0] + e2[1] * q[1]) + e2[2] * q[2]) * invDet;
  if (u < 0.0 || u > 1.0) return 0;
  if (v < 0.0 || u + v > 1.0f) return 0;
  if (t < 0.0 || t > *tio) return 0;
  *tio = t;
  *uo = u;
  *vo = v;
  return 1;
}

/* TriangleAccel::Traverse, render/render_accel_triangle.cc:2511-2669, with
 * TestLeafNode (scalar path) :2386-2433. */
static void trace_tri_one(const or_tri_node *nodes, const uint32_t *indices,
                          const float *V, const uint32_t *F, float eps_scale,
                          const or_ray *r, or_hit *h, uint32_t *st) {
  ray_setup s;
  setup_ray(r, &s);
  float hitT = FLT_MAX;
  h->t = hitT;
  h->u = 0.0f;
  h->v = 0.0f;
  h->prim_id = 0xffffffffu;
  int stack[STACK_MAX];
  int sp = 0;
  stack[0] = 0;
  uint32_t nvisit = 0, ntest = 0;
  const int ds = (int)(r->flags & 1u);
  while (sp >= 0) {
    const or_tri_node *nd = nodes + stack[sp--];
    nvisit++;
    if (nd->flag == 0) {
      int hit0 = ray_aabb(&s, hitT, nd->bmin[0], nd->bmax[0], 1);
      int hit1 = ray_aabb(&s, hitT, nd->bmin[1], nd->bmax[1], 1);
      int hits[2] = {hit0, hit1};
      if (hit0 || hit1) {
        int nearc = s.sign[nd->axis], farc = 1 - nearc;
        if (hits[farc]) stack[++sp] = (int)nd->data[farc];
        if (hits[nearc]) stack[++sp] = (int)nd->data[nearc];
      }
    } else {
      float t = h->t;
      int any = 0;
      for (uint32_t i = 0; i < nd->data[0]; i++) {
        uint32_t fi = indices[i + nd->data[1]];
        if (fi == r->prev_prim_id) continue;
        ntest++;
        const float *p0 = V + 3 * (size_t)F[3 * (size_t)fi + 0];
        const float *p1 = V + 3 * (size_t)F[3 * (size_t)fi + 1];
        const float *p2 = V + 3 * (size_t)F[3 * (size_t)fi + 2];
        float u, v;
        if (tri_isect(&t, &u, &v, p0, p1, p2, &s, ds, eps_scale)) {
          h->t = t;
          h->u = u;
          h->v = v;
          h->prim_id = fi;
          any = 1;
        }
      }
      if (any) hitT = h->t;
    }
  }
  if (st) {
    st[0] = nvisit;
    st[1] = ntest;
  }
}

void or_trace_tri(const or_tri_node *nodes, const uint32_t *indices,
                  const float *V, const uint32_t *F, double eps_scale,
                  const or_ray *rays, int64_t nrays, or_hit *hits,
                  uint32_t *stats, int threads) {
  const float es = (float)eps_scale; /* TriangleIsect takes real epsScale */
  (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
#endif
  for (int64_t i = 0; i < nrays; i++)
    trace_tri_one(nodes, indices, V, F, es, rays + i, hits + i,
                  stats ? stats + 2 * i : NULL);
}

/* ---- tetra ---- */

typedef struct {
  double x, y, z;
} d3;

static d3 d3sub(d3 a, d3 b) { d3 c = {a.x - b.x, a.y - b.y, a.z - b.z}; return c; }
static d3 d3add(d3 a, d3 b) { d3 c = {a.x + b.x, a.y + b.y, a.z + b.z}; return c; }
/* operator*(double f, const double3 &v) = (v.x * f, ...), render/render_common.h:172-174 */
static d3 d3scale(double f, d3 v) { d3 c = {v.x * f, v.y * f, v.z * f}; return c; }
/* vcrossd / vdotd, render/render_accel_tetra.cc:75-86 */
static d3 d3cross(d3 a, d3 b) {
  d3 c = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  return c;
}
static double d3dot(d3 a, d3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

/* fsign, render/render_accel_tetra.cc:89-93 */
static int fsign(double x) {
  const double eps = DBL_EPSILON;
  return x > eps ? 1 : (x < -eps ? -1 : 0);
}

/* Pluecker(v0, v1) = {d = v1 - v0, c = v1 x v0} and its inner product,
 * render/render_accel_tetra.cc:98-115 */
typedef struct {
  d3 d, c;
} pluecker;
static pluecker mk_pl(d3 v0, d3 v1) {
  pluecker p;
  p.d = d3sub(v1, v0);
  p.c = d3cross(v1, v0);
  return p;
}
static double pl_dot(pluecker a, pluecker b) {
  return d3dot(a.d, b.c) + d3dot(b.d, a.c);
}

/* ComputeParametricDist, render/render_accel_tetra.cc:123-139 */
static void param_dist(d3 o, d3 d, d3 ep, d3 lp, double *te, double *tl) {
  if (d.x != 0.0) {
    double inv = 1.0 / d.x;
    *te = (ep.x - o.x) * inv;
    *tl = (lp.x - o.x) * inv;
  } else if (d.y != 0.0) {
    double inv = 1.0 / d.y;
    *te = (ep.y - o.y) * inv;
    *tl = (lp.y - o.y) * inv;
  } else {
    double inv = 1.0 / d.z;
    *te = (ep.z - o.z) * inv;
    *tl = (lp.z - o.z) * inv;
  }
}

/* (1 - a - b) * A + a * B + b * C as the double3 operators evaluate it. */
static d3 bary3(double a, double b, d3 A, d3 B, d3 C) {
  return d3add(d3add(d3scale((1 - a) - b, A), d3scale(a, B)), d3scale(b, C));
}

/* RayTetraPluecker (Platis & Theoharis 2003), render/render_accel_tetra.cc:243-567 */
static int ray_tet_pluecker(d3 orig, d3 dir, const d3 vert[4], int *enterFace,
                            int *leaveFace, double *uE1, double *uE2,
                            double *uL1, double *uL2, double *tE,
                            double *tL) {
  *enterFace = -1;
  *leaveFace = -1;
  double uAB = 0, uAC = 0, uDB = 0, uDC = 0, uBC = 0, uAD = 0;
  int signAB = -2, signAC = -2, signDB = -2, signDC = -2, signBC = -2,
      signAD = -2;
  d3 dest = d3add(orig, dir);
  pluecker plRay = mk_pl(orig, dest);
  int nextSign = 0;
  d3 ep = {0, 0, 0}, lp = {0, 0, 0};

  uAB = pl_dot(plRay, mk_pl(vert[0], vert[1]));
  signAB = fsign(uAB);
  uAC = pl_dot(plRay, mk_pl(vert[0], vert[2]));
  signAC = fsign(uAC);

  if ((signAC == -signAB) || (signAC == 0) || (signAB == 0)) {
    uBC = pl_dot(plRay, mk_pl(vert[1], vert[2]));
    signBC = fsign(uBC);
    int signABC = signAB;
    if (signABC == 0) {
      signABC = -signAC;
      if (signABC == 0) signABC = signBC;
    }
    if ((signABC != 0) && ((signBC == signABC) || (signBC == 0))) {
      double invVolABC = 1.0 / ((uAB + uBC) - uAC);
      if (signABC == 1) {
        *enterFace = 3;
        *uE1 = -uAC * invVolABC;
        *uE2 = uAB * invVolABC;
        ep = bary3(*uE1, *uE2, vert[0], vert[1], vert[2]);
        nextSign = -1;
      } else {
        *leaveFace = 3;
        *uL1 = -uAC * invVolABC;
        *uL2 = uAB * invVolABC;
        lp = bary3(*uL1, *uL2, vert[0], vert[1], vert[2]);
        nextSign = 1;
      }
      uAD = pl_dot(plRay, mk_pl(vert[0], vert[3]));
      signAD = fsign(uAD);
      if ((signAD == nextSign) || (signAD == 0)) {
        uDB = pl_dot(plRay, mk_pl(vert[3], vert[1]));
        signDB = fsign(uDB);
        if ((signDB == nextSign) ||
            ((signDB == 0) && ((signAD != 0) || (signAB != 0)))) {
          double invVolBAD = 1.0 / ((uAD + uDB) - uAB);
          if (nextSign == 1) {
            *enterFace = 2;
            *uE1 = uDB * invVolBAD;
            *uE2 = -uAB * invVolBAD;
            ep = bary3(*uE1, *uE2, vert[1], vert[0], vert[3]);
          } else {
            *leaveFace = 2;
            *uL1 = uDB * invVolBAD;
            *uL2 = -uAB * invVolBAD;
            lp = bary3(*uL1, *uL2, vert[1], vert[0], vert[3]);
          }
          param_dist(orig, dir, ep, lp, tE, tL);
          return 1;
        }
      }
      uDC = pl_dot(plRay, mk_pl(vert[3], vert[2]));
      signDC = fsign(uDC);
      if ((signDC == -nextSign) ||
          ((signDC == 0) && ((signAD != 0) || (signAC != 0)))) {
        double invVolCDA = 1.0 / ((uAC - uDC) - uAD);
        if (nextSign == 1) {
          *enterFace = 1;
          *uE1 = uAC * invVolCDA;
          *uE2 = -uDC * invVolCDA;
          ep = bary3(*uE1, *uE2, vert[2], vert[3], vert[0]);
        } else {
          *leaveFace = 1;
          *uL1 = uAC * invVolCDA;
          *uL2 = -uDC * invVolCDA;
          lp = bary3(*uL1, *uL2, vert[2], vert[3], vert[0]);
        }
        param_dist(orig, dir, ep, lp, tE, tL);
        return 1;
      } else {
        if (signDB == -2) uDB = pl_dot(plRay, mk_pl(vert[3], vert[1]));
        double invVolDCB = 1.0 / ((uDC - uBC) - uDB);
        if (nextSign == 1) {
          *enterFace = 0;
          *uE1 = -uDB * invVolDCB;
          *uE2 = uDC * invVolDCB;
          ep = bary3(*uE1, *uE2, vert[3], vert[2], vert[1]);
        } else {
          *leaveFace = 0;
          *uL1 = -uDB * invVolDCB;
          *uL2 = uDC * invVolDCB;
          lp = bary3(*uL1, *uL2, vert[3], vert[2], vert[1]);
        }
        param_dist(orig, dir, ep, lp, tE, tL);
        return 1;
      }
    }
  }

  uAD = pl_dot(plRay, mk_pl(vert[0], vert[3]));
  signAD = fsign(uAD);
  if ((signAD == -signAB) || (signAB == 0) || (signAD == 0)) {
    uDB = pl_dot(plRay, mk_pl(vert[3], vert[1]));
    signDB = fsign(uDB);
    int signBAD = -signAB;
    if (signBAD == 0) {
      signBAD = signAD;
      if (signBAD == 0) signBAD = signDB;
    }
    if ((signBAD != 0) && ((signDB == signBAD) || (signDB == 0))) {
      double invVolBAD = 1.0 / ((uAD + uDB) - uAB);
      if (signBAD == 1) {
        *enterFace = 2;
        *uE1 = uDB * invVolBAD;
        *uE2 = -uAB * invVolBAD;
        ep = bary3(*uE1, *uE2, vert[1], vert[0], vert[3]);
        nextSign = -1;
      } else {
        *leaveFace = 2;
        *uL1 = uDB * invVolBAD;
        *uL2 = -uAB * invVolBAD;
        lp = bary3(*uL1, *uL2, vert[1], vert[0], vert[3]);
        nextSign = 1;
      }
      uDC = pl_dot(plRay, mk_pl(vert[3], vert[2]));
      signDC = fsign(uDC);
      if ((signDC == -nextSign) ||
          ((signDC == 0) && ((signAD != 0) || (signAC != 0)))) {
        double invVolCDA = 1.0 / ((uAC - uDC) - uAD);
        if (nextSign == 1) {
          *enterFace = 1;
          *uE1 = uAC * invVolCDA;
          *uE2 = -uDC * invVolCDA;
          ep = bary3(*uE1, *uE2, vert[2], vert[3], vert[0]);
        } else {
          *leaveFace = 1;
          *uL1 = uAC * invVolCDA;
          *uL2 = -uDC * invVolCDA;
          lp = bary3(*uL1, *uL2, vert[2], vert[3], vert[0]);
        }
        param_dist(orig, dir, ep, lp, tE, tL);
        return 1;
      } else {
        if (signBC == -2) uBC = pl_dot(plRay, mk_pl(vert[1], vert[2]));
        double invVolDCB = 1.0 / ((uDC - uBC) - uDB);
        if (nextSign == 1) {
          *enterFace = 0;
          *uE1 = -uDB * invVolDCB;
          *uE2 = uDC * invVolDCB;
          ep = bary3(*uE1, *uE2, vert[3], vert[2], vert[1]);
        } else {
          *leaveFace = 0;
          *uL1 = -uDB * invVolDCB;
          *uL2 = uDC * invVolDCB;
          lp = bary3(*uL1, *uL2, vert[3], vert[2], vert[1]);
        }
        param_dist(orig, dir, ep, lp, tE, tL);
        return 1;
      }
    }
  }

  if ((-signAD == signAC) || (signAC == 0) || (signAD == 0)) {
    uDC = pl_dot(plRay, mk_pl(vert[3], vert[2]));
    signDC = fsign(uDC);
    int signCDA = signAC;
    if (signCDA == 0) {
      signCDA = -signAD;
      if (signCDA == 0) signCDA = -signDC;
    }
    if ((signCDA != 0) && ((-signDC == signCDA) || (signDC == 0))) {
      double invVolCDA = 1.0 / ((uAC - uDC) - uAD);
      if (signBC == -2) uBC = pl_dot(plRay, mk_pl(vert[1], vert[2]));
      if (signDB == -2) uDB = pl_dot(plRay, mk_pl(vert[3], vert[1]));
      double invVolDCB = 1.0 / ((uDC - uBC) - uDB);
      if (signCDA == 1) {
        *enterFace = 1;
        *uE1 = uAC * invVolCDA;
        *uE2 = -uDC * invVolCDA;
        ep = bary3(*uE1, *uE2, vert[2], vert[3], vert[0]);
        *leaveFace = 0;
        *uL1 = -uDB * invVolDCB;
        *uL2 = uDC * invVolDCB;
        lp = bary3(*uL1, *uL2, vert[3], vert[2], vert[1]);
      } else {
        *leaveFace = 1;
        *uL1 = uAC * invVolCDA;
        *uL2 = -uDC * invVolCDA;
        lp = bary3(*uL1, *uL2, vert[2], vert[3], vert[0]);
        *enterFace = 0;
        *uE1 = -uDB * invVolDCB;
        *uE2 = uDC * invVolDCB;
        ep = bary3(*uE1, *uE2, vert[3], vert[2], vert[1]);
      }
      param_dist(orig, dir, ep, lp, tE, tL);
      return 1;
    }
  }
  return 0;
}

/* float real3 helpers for InterpolateTetra (render/render_common.h:30-160) */
typedef struct {
  float v[3];
} f3;
static f3 f3sub(f3 a, f3 b) {
  f3 c = {{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]}};
  return c;
}
static f3 f3add(f3 a, f3 b) {
  f3 c = {{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]}};
  return c;
}
static f3 f3mul(f3 a, float f) {
  f3 c = {{a.v[0] * f, a.v[1] * f, a.v[2] * f}};
  return c;
}
static f3 f3div(f3 a, float f) {
  f3 c = {{a.v[0] / f, a.v[1] / f, a.v[2] / f}};
  return c;
}
static f3 f3cross(f3 l, f3 r) {
  f3 c = {{l.v[1] * r.v[2] - l.v[2] * r.v[1], l.v[2] * r.v[0] - l.v[0] * r.v[2],
           l.v[0] * r.v[1] - l.v[1] * r.v[0]}};
  return c;
}
static float f3dot(f3 l, f3 r) {
  return (l.v[0] * r.v[0] + l.v[1] * r.v[1]) + l.v[2] * r.v[2];
}
/* real3::length: sqrt of the float sum of squares (correctly rounded). */
static float f3len(f3 a) {
  float s = (a.v[0] * a.v[0] + a.v[1] * a.v[1]) + a.v[2] * a.v[2];
  return (float)sqrt((double)s);
}

static int contains3(int v, const int l[3]) {
  return l[0] == v || l[1] == v || l[2] == v;
}

/* InterpolateTetra (interpolate_mode 0), render/render_accel_tetra.cc:148-241 */
static void interpolate_tetra(float d_out[4], f3 pt, const f3 verts[4]) {
  static const int faces[4][3] = {{0, 2, 1}, {1, 2, 3}, {0, 3, 2}, {0, 1, 3}};
  const float kEPS = (float)0.001;
  f3 edges[6];
  edges[0] = f3sub(verts[1], verts[0]);
  edges[1] = f3sub(verts[2], verts[1]);
  edges[2] = f3sub(verts[0], verts[2]);
  edges[3] = f3sub(verts[3], verts[0]);
  edges[4] = f3sub(verts[3], verts[1]);
  edges[5] = f3sub(verts[3], verts[2]);
  for (int v = 0; v < 4; v++) {
    f3 rayorg = verts[v];
    f3 raydir = f3sub(pt, verts[v]);
    f3 raypc = f3cross(raydir, rayorg);
    float d = f3len(f3sub(pt, rayorg));
    float ws[6];
    int cw0[6], cw1[6];
    for (int i = 0; i < 6; i++) {
      ws[i] = f3dot(raydir, f3cross(edges[i], verts[i % 3])) +
              f3dot(edges[i], raypc);
      cw0[i] = ws[i] >= -kEPS;
      cw1[i] = ws[i] <= kEPS;
    }
    f3 cp = {{(float)1e16, (float)1e16, (float)1e16}};
    if (cw0[0] && cw0[1] && cw0[2] && !contains3(v, faces[0]))
      cp = f3div(f3add(f3add(f3mul(verts[2], ws[0]), f3mul(verts[0], ws[1])),
                       f3mul(verts[1], ws[2])),
                 (ws[0] + ws[1]) + ws[2]);
    else if (cw1[1] && cw0[4] && cw1[5] && !contains3(v, faces[1]))
      cp = f3div(f3sub(f3add(f3mul(verts[3], ws[1]), f3mul(verts[1], ws[5])),
                       f3mul(verts[2], ws[4])),
                 (ws[5] + ws[1]) - ws[4]);
    else if (cw1[2] && cw1[3] && cw0[5] && !contains3(v, faces[2]))
      cp = f3div(f3sub(f3add(f3mul(verts[3], ws[2]), f3mul(verts[2], ws[3])),
                       f3mul(verts[0], ws[5])),
                 (ws[2] + ws[3]) - ws[5]);
    else if (cw1[0] && cw0[3] && cw1[4] && !contains3(v, faces[3]))
      cp = f3div(f3sub(f3add(f3mul(verts[3], ws[0]), f3mul(verts[0], ws[4])),
                       f3mul(verts[1], ws[3])),
                 (ws[0] + ws[4]) - ws[3]);
    float dv = d / f3len(f3sub(cp, rayorg));
    if (dv < 0)
      dv = 0;
    else if (!(dv < 1) && !(dv > 0))
      dv = 1;
    d_out[v] = 1 - dv;
  }
  float ds = 0, w = 0, dd[4];
  for (int i = 0; i < 4; i++) ds += f3len(f3sub(pt, verts[i]));
  for (int i = 0; i < 4; i++) dd[i] = ds / f3len(f3sub(pt, verts[i]));
  for (int i = 0; i < 4; i++) {
    d_out[i] *= dd[i];
    w += d_out[i];
  }
  for (int i = 0; i < 4; i++) d_out[i] = d_out[i] / w;
}

/* TetraAccel::Traverse, render/render_accel_tetra.cc:2200-2289; TestLeafNode
 * :1980-2122 (float vertices widened to double); BuildIntersection
 * :2124-2196; shader-side varying f0*d0 + f1*d1 + f2*d2 + f3*d3,
 * gles/gles_common.cc:1074-1076. */
static void trace_tet_one(const or_tet_node *nodes, const uint32_t *indices,
                          const float *V, const uint32_t *F,
                          const float *scalars, const or_ray *r, or_hit *h,
                          or_tet_ext *ext, uint32_t *st) {
  ray_setup s;
  setup_ray(r, &s);
  float hitT = FLT_MAX;
  h->t = hitT;
  h->u = 0.0f;
  h->v = 0.0f;
  h->prim_id = 0xffffffffu;
  uint32_t subface = 0;
  double tleave_best = 0.0;
  int stack[STACK_MAX];
  int sp = 0;
  stack[0] = 0;
  uint32_t nvisit = 0, ntest = 0;
  d3 ro = {s.o[0], s.o[1], s.o[2]}, rd = {s.d[0], s.d[1], s.d[2]};
  while (sp >= 0) {
    const or_tet_node *nd = nodes + stack[sp--];
    nvisit++;
    int hit = ray_aabb(&s, hitT, nd->bmin, nd->bmax, 0);
    if (nd->flag == 0) {
      if (hit) {
        int nearc = s.sign[nd->axis], farc = 1 - nearc;
        stack[++sp] = (int)nd->data[farc];
        stack[++sp] = (int)nd->data[nearc];
      }
    } else if (hit) {
      float t = h->t;
      int any = 0;
      for (uint32_t i = 0; i < nd->data[0]; i++) {
        uint32_t fi = indices[i + nd->data[1]];
        if (fi == r->prev_prim_id) continue;
        ntest++;
        d3 vtx[4];
        for (int j = 0; j < 4; j++) {
          const float *p = V + 3 * (size_t)F[4 * (size_t)fi + j];
          vtx[j].x = p[0];
          vtx[j].y = p[1];
          vtx[j].z = p[2];
        }
        int ef, lf;
        double eu = 0, lu = 0, ev = 0, lv = 0, et = 0, lt = 0;
        /* argument order of the call site, render/render_accel_tetra.cc:2103:
         * (enterU, leaveU, enterV, leaveV) land in (uEnter1, uEnter2,
         * uLeave1, uLeave2), so isect.v = uLeave1. */
        if (ray_tet_pluecker(ro, rd, vtx, &ef, &lf, &eu, &lu, &ev, &lv, &et,
                             &lt)) {
          if ((et >= 0.0) && (et < (double)t)) {
            h->t = (float)et;
            h->u = (float)eu;
            h->v = (float)ev;
            h->prim_id = fi;
            subface = (uint32_t)ef;
            tleave_best = lt;
            t = (float)et;
            any = 1;
          }
        }
      }
      if (any) hitT = h->t;
    }
  }
  if (st) {
    st[0] = nvisit;
    st[1] = ntest;
  }
  if (ext) {
    memset(ext, 0, sizeof(*ext));
    if (h->t < FLT_MAX) {
      f3 p[4];
      uint32_t fidx[4];
      for (int j = 0; j < 4; j++) {
        fidx[j] = F[4 * (size_t)h->prim_id + j];
        for (int k = 0; k < 3; k++) p[j].v[k] = V[3 * (size_t)fidx[j] + k];
      }
      f3 pos;
      for (int k = 0; k < 3; k++) pos.v[k] = s.o[k] + s.d[k] * h->t;
      interpolate_tetra(ext->d, pos, p);
      ext->subface = subface;
      ext->t_leave = (float)tleave_best;
      if (scalars) {
        float f0 = scalars[fidx[0]], f1 = scalars[fidx[1]];
        float f2 = scalars[fidx[2]], f3v = scalars[fidx[3]];
        ext->scalar = ((f0 * ext->d[0] + f1 * ext->d[1]) + f2 * ext->d[2]) +
                      f3v * ext->d[3];
      }
    }
  }
}

void or_trace_tet(const or_tet_node *nodes, const uint32_t *indices,
                  const float *V, const uint32_t *F, const float *scalars,
                  const or_ray *rays, int64_t nrays, or_hit *hits,
                  or_tet_ext *ext, uint32_t *stats, int threads) {
  (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
#endif
  for (int64_t i = 0; i < nrays; i++)
    trace_tet_one(nodes, indices, V, F, scalars, rays + i, hits + i,
                  ext ? ext + i : NULL, stats ? stats + 2 * i : NULL);
}

/* ---- solids (pyramid, prism, hexahedron) ---- */

/* GetEdges, render/render_accel_solid.cc:119-154 (double) and :156-191
 * (float): edge i runs from vertex i % vmod to vertex i % vmod + edge i. */
static int sol_edge_count(int nv) { return nv == 5 ? 8 : (nv == 6 ? 9 : 12); }
static int sol_vmod(int nv) { return nv == 5 ? 4 : nv; }
static const int kSolEdgeEnds[3][12][2] = {
    /* pyramid */ {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 4}, {2, 4}, {3, 4}},
    /* prism */ {{0, 1}, {1, 2}, {2, 0}, {3, 4}, {4, 5}, {5, 3}, {0, 3}, {1, 4}, {2, 5}},
    /* hexahedron */
    {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6},
     {3, 7}}};
static int sol_type(int nv) { return nv == 5 ? 0 : (nv == 6 ? 1 : 2); }

/* Result of Intersect{Pyramid,Prism,Hexa}D for one of the two crossings:
 * Intersection::{t, position, normal}. */
typedef struct {
  float t;
  float pos[3];
  float nrm[3];
} sol_cross;

/* toreal3(num) / den: the double numerator rounded to float, divided by the
 * float-rounded denominator (real3::operator/(real), render/render_common.h:57-59) */
static void sol_pos_div(float out[3], d3 num, double den) {
  const float f = (float)den;
  out[0] = (float)num.x / f;
  out[1] = (float)num.y / f;
  out[2] = (float)num.z / f;
}

/* SetCrossPoint_Sq (double), render/render_accel_solid.cc:438-444 */
static void sol_pos_sq(float out[3], d3 v0, d3 v1, d3 v2, double ws0, double ws1, double ws2,
                       double ws3, d3 e2, d3 e3, d3 rd) {
  const double w = (ws2 + ws3) + d3dot(d3cross(e2, e3), rd);
  const d3 num = d3add(d3add(d3scale(ws0, v2), d3scale(ws1, v0)), d3scale(w, v1));
  const double den = (ws0 + ws1) + w;
  out[0] = (float)(num.x / den);
  out[1] = (float)(num.y / den);
  out[2] = (float)(num.z / den);
}

/* toreal3(normalize(vcrossd(a, b))), render/render_accel_solid.cc:69-85 */
static void sol_nrm(float out[3], d3 a, d3 b) {
  const d3 c = d3cross(a, b);
  const double len = sqrt((c.x * c.x + c.y * c.y) + c.z * c.z);
  out[0] = (float)(c.x / len);
  out[1] = (float)(c.y / len);
  out[2] = (float)(c.z / len);
}

/* (v4 * -wa + vb * wb + vc * -wc) / (-wa + wb - wc), the pyramid's side
 * faces, render/render_accel_solid.cc:484-495 */
static void sol_pos_pyr(float out[3], d3 v4, d3 vb, d3 vc, double wa, double wb, double wc) {
  const d3 num = d3add(d3add(d3scale(-wa, v4), d3scale(wb, vb)), d3scale(-wc, vc));
  sol_pos_div(out, num, ((-wa) + wb) - wc);
}

/* (va * w0 + vb * w1 + vc * w2) / (w0 + w1 + w2), the prism's triangles,
 * render/render_accel_solid.cc:585-590 */
static void sol_pos_tri(float out[3], d3 va, d3 vb, d3 vc, double w0, double w1, double w2) {
  const d3 num = d3add(d3add(d3scale(w0, va), d3scale(w1, vb)), d3scale(w2, vc));
  sol_pos_div(out, num, (w0 + w1) + w2);
}

/* Intersect{Pyramid,Prism,Hexa}D, render/render_accel_solid.cc:455-505,
 * 559-611, 664-716: Pluecker side products of the ray with every edge,
 * then for each of the two crossings the first face whose edges all lie on
 * one side (the "<= 0" side for c[0] of prism / hexahedron and for the
 * pyramid's base, see the loops' index pairs), its cross point and outward
 * normal; t from the cross point along the last axis with |dir| > 0.1.
 * Returns 0 when a crossing finds no face, 1 with both crossings, -1 when
 * no axis qualifies (the reference leaves t unset; treated as no hit). */
int or_solid_isect(int nv, const double o_in[3], const double d_in[3], const double *vin,
                   float out_t[2], float out_pos[6], float out_nrm[6]) {
  d3 v[8], e[12], ro = {o_in[0], o_in[1], o_in[2]}, rd = {d_in[0], d_in[1], d_in[2]};
  double ws[12];
  int cw[2][12];
  const int ty = sol_type(nv), ne = sol_edge_count(nv), vm = sol_vmod(nv);
  for (int k = 0; k < nv; k++) {
    v[k].x = vin[3 * k];
    v[k].y = vin[3 * k + 1];
    v[k].z = vin[3 * k + 2];
  }
  for (int i = 0; i < ne; i++) e[i] = d3sub(v[kSolEdgeEnds[ty][i][1]], v[kSolEdgeEnds[ty][i][0]]);
  const d3 raypc = d3cross(rd, ro);
  for (int i = 0; i < ne; i++) {
    const d3 epc = d3cross(e[i], v[i % vm]);
    ws[i] = d3dot(rd, epc) + d3dot(e[i], raypc);
    cw[0][i] = ws[i] >= 0;
    cw[1][i] = ws[i] <= 0;
  }
  sol_cross is[2];
  for (int pass = 0; pass < 2; pass++) {
    sol_cross *c;
    if (nv == 5) {
      /* for (int i = 1, n = 0; n < 2; i--, n++) fills isects[n] */
      const int n = pass, i = 1 - pass;
      c = &is[n];
      if (cw[i][0] && cw[i][1] && cw[i][2] && cw[i][3]) {
        sol_pos_sq(c->pos, v[0], v[1], v[2], ws[0], ws[1], ws[2], ws[3], e[2], e[3], rd);
        sol_nrm(c->nrm, e[1], e[0]);
      } else if (cw[n][0] && cw[i][4] && cw[n][5]) {
        sol_pos_pyr(c->pos, v[4], v[1], v[0], ws[0], ws[4], ws[5]);
        sol_nrm(c->nrm, e[4], e[0]);
      } else if (cw[n][1] && cw[i][5] && cw[n][6]) {
        sol_pos_pyr(c->pos, v[4], v[2], v[1], ws[1], ws[5], ws[6]);
        sol_nrm(c->nrm, e[5], e[1]);
      } else if (cw[n][2] && cw[i][6] && cw[n][7]) {
        sol_pos_pyr(c->pos, v[4], v[3], v[2], ws[2], ws[6], ws[7]);
        sol_nrm(c->nrm, e[6], e[2]);
      } else if (cw[n][3] && cw[i][7] && cw[n][4]) {
        sol_pos_pyr(c->pos, v[4], v[0], v[3], ws[3], ws[7], ws[4]);
        sol_nrm(c->nrm, e[7], e[3]);
      } else {
        return 0;
      }
    } else if (nv == 6) {
      /* for (int i = 0, n = 1; i < 2; i++, n--) fills isects[i] */
      const int i = pass, n = 1 - pass;
      c = &is[i];
      if (cw[n][0] && cw[n][1] && cw[n][2]) {
        sol_pos_tri(c->pos, v[2], v[0], v[1], ws[0], ws[1], ws[2]);
        sol_nrm(c->nrm, e[0], e[1]);
      } else if (cw[i][3] && cw[i][4] && cw[i][5]) {
        sol_pos_tri(c->pos, v[5], v[3], v[4], ws[3], ws[4], ws[5]);
        sol_nrm(c->nrm, e[3], e[4]);
      } else if (cw[i][2] && cw[i][6] && cw[n][5] && cw[n][8]) {
        sol_pos_sq(c->pos, v[2], v[0], v[3], ws[2], ws[6], -ws[5], -ws[8], e[5], e[8], rd);
        sol_nrm(c->nrm, e[2], e[6]);
      } else if (cw[i][0] && cw[i][7] && cw[n][3] && cw[n][6]) {
        sol_pos_sq(c->pos, v[0], v[1], v[4], ws[0], ws[7], -ws[3], -ws[6], e[3], e[6], rd);
        sol_nrm(c->nrm, e[0], e[7]);
      } else if (cw[i][1] && cw[i][8] && cw[n][4] && cw[n][7]) {
        sol_pos_sq(c->pos, v[1], v[2], v[5], ws[1], ws[8], -ws[4], -ws[7], e[4], e[7], rd);
        sol_nrm(c->nrm, e[1], e[8]);
      } else {
        return 0;
      }
    } else {
      const int i = pass, n = 1 - pass;
      c = &is[i];
      if (cw[n][0] && cw[n][1] && cw[n][2] && cw[n][3]) {
        sol_pos_sq(c->pos, v[0], v[1], v[2], ws[0], ws[1], ws[2], ws[3], e[2], e[3], rd);
        sol_nrm(c->nrm, e[0], e[1]);
      } else if (cw[i][4] && cw[i][5] && cw[i][6] && cw[i][7]) {
        sol_pos_sq(c->pos, v[4], v[5], v[6], ws[4], ws[5], ws[6], ws[7], e[6], e[7], rd);
        sol_nrm(c->nrm, e[4], e[5]);
      } else if (cw[i][0] && cw[i][9] && cw[n][4] && cw[n][8]) {
        sol_pos_sq(c->pos, v[0], v[1], v[5], ws[0], ws[9], -ws[4], -ws[8], e[4], e[8], rd);
        sol_nrm(c->nrm, e[0], e[9]);
      } else if (cw[i][1] && cw[i][10] && cw[n][5] && cw[n][9]) {
        sol_pos_sq(c->pos, v[1], v[2], v[6], ws[1], ws[10], -ws[5], -ws[9], e[5], e[9], rd);
        sol_nrm(c->nrm, e[1], e[10]);
      } else if (cw[i][2] && cw[i][11] && cw[n][6] && cw[n][10]) {
        sol_pos_sq(c->pos, v[2], v[3], v[7], ws[2], ws[11], -ws[6], -ws[10], e[6], e[10], rd);
        sol_nrm(c->nrm, e[2], e[11]);
      } else if (cw[i][3] && cw[i][8] && cw[n][7] && cw[n][11]) {
        sol_pos_sq(c->pos, v[3], v[0], v[4], ws[3], ws[8], -ws[7], -ws[11], e[7], e[11], rd);
        sol_nrm(c->nrm, e[3], e[8]);
      } else {
        return 0;
      }
    }
  }
  int axis = -1;
  for (int j = 0; j < 3; j++) {
    const double dj = j == 0 ? rd.x : (j == 1 ? rd.y : rd.z);
    if (dj > 0.1 || dj < -0.1) axis = j;
  }
  if (axis < 0) return -1;
  {
    const double oj = axis == 0 ? ro.x : (axis == 1 ? ro.y : ro.z);
    const double dj = axis == 0 ? rd.x : (axis == 1 ? rd.y : rd.z);
    for (int k = 0; k < 2; k++) out_t[k] = (float)(((double)is[k].pos[axis] - oj) / dj);
  }
  for (int k = 0; k < 3; k++) is[1].nrm[k] = -is[1].nrm[k];
  for (int k = 0; k < 2; k++)
    for (int c = 0; c < 3; c++) {
      out_pos[3 * k + c] = is[k].pos[c];
      out_nrm[3 * k + c] = is[k].nrm[c];
    }
  return 1;
}

/* getCP_sq (float vertices, double weights), render/render_accel_solid.cc:
 * 238-244: the double w rounds to float where it scales a vertex and the
 * denominator rounds to float before the division. */
static f3 sol_cp_sq_f(f3 v0, f3 v1, f3 v2, double ws0, double ws1, double ws2, double ws3,
                      f3 e2, f3 e3, f3 rd) {
  const double w = (ws2 + ws3) + (double)f3dot(f3cross(e2, e3), rd);
  const f3 num = f3add(f3add(f3mul(v2, (float)ws0), f3mul(v0, (float)ws1)), f3mul(v1, (float)w));
  return f3div(num, (float)((ws0 + ws1) + w));
}

/* (vx * -wa + vy * wb + vz * -wc) / (-wa + wb - wc), all float */
static f3 sol_cp_pyr_f(f3 vx, f3 vy, f3 vz, float wa, float wb, float wc) {
  const f3 num = f3add(f3add(f3mul(vx, -wa), f3mul(vy, wb)), f3mul(vz, -wc));
  return f3div(num, ((-wa) + wb) - wc);
}

static f3 sol_cp_tri_f(f3 va, f3 vb, f3 vc, float w0, float w1, float w2) {
  const f3 num = f3add(f3add(f3mul(va, w0), f3mul(vb, w1)), f3mul(vc, w2));
  return f3div(num, (w0 + w1) + w2);
}

static int contains4(int v, const int l[4]) {
  return l[0] == v || l[1] == v || l[2] == v || l[3] == v;
}

/* Initial value of interpolate()'s kEPS accumulator, which the reference
 * leaves uninitialised (render/render_accel_solid.cc:271-275); 0 here. */
#ifndef OR_SOLID_KEPS_INIT
#define OR_SOLID_KEPS_INIT 0.0f
#endif

/* interpolate() with interpolate_mode 0, render/render_accel_solid.cc:247-435:
 * per-vertex weights at p from the ray (vertex -> p) crossing the opposite
 * faces, everything in float after moving vertex 0 to the origin. */
void or_interpolate_solid(int nv, const float p[3], const double *vin, float d_out[8]) {
  static const int faces[3][6][4] = {
      {{0, 3, 2, 1}, {0, 1, 4, -1}, {1, 2, 4, -1}, {2, 3, 4, -1}, {0, 4, 3, -1}},
      {{0, 2, 1, -1}, {3, 4, 5, -1}, {0, 3, 5, 2}, {0, 1, 4, 3}, {1, 2, 5, 4}},
      {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {0, 4, 7, 3}}};
  const int ty = sol_type(nv), ne = sol_edge_count(nv), vm = sol_vmod(nv);
  f3 pt = {{p[0], p[1], p[2]}}, verts[8], edges[12], edgepc[12];
  for (int i = 0; i < nv; i++)
    for (int k = 0; k < 3; k++) verts[i].v[k] = (float)vin[3 * i + k];
  pt = f3sub(pt, verts[0]);
  for (int i = nv - 1; i >= 0; i--) verts[i] = f3sub(verts[i], verts[0]);
  for (int i = 0; i < ne; i++)
    edges[i] = f3sub(verts[kSolEdgeEnds[ty][i][1]], verts[kSolEdgeEnds[ty][i][0]]);
  float kEPS = OR_SOLID_KEPS_INIT;
  for (int i = 0; i < ne; i++) kEPS += f3len(edges[i]);
  kEPS = (float)(((double)kEPS * 1e-6) / ne);
  for (int i = 0; i < ne; i++) edgepc[i] = f3cross(edges[i], verts[i % vm]);
  for (int v = 0; v < nv; v++) {
    const f3 rayorg = verts[v];
    const f3 raydir = f3sub(pt, rayorg);
    const f3 raypc = f3cross(raydir, rayorg);
    const float d = f3len(f3sub(pt, rayorg));
    float ws[12];
    int c0[12], c1[12];
    for (int i = 0; i < ne; i++) {
      ws[i] = f3dot(raydir, edgepc[i]) + f3dot(edges[i], raypc);
      c0[i] = ws[i] >= -kEPS;
      c1[i] = ws[i] <= kEPS;
    }
    const int(*fc)[4] = faces[ty];
    const f3 *V = verts;
    f3 cp;
    int found = 1;
    if (nv == 5) {
      if (c1[0] && c1[1] && c1[2] && c1[3] && !contains4(v, fc[0]))
        cp = sol_cp_sq_f(V[0], V[1], V[2], ws[0], ws[1], ws[2], ws[3], edges[2], edges[3], raydir);
      else if (c0[0] && c1[4] && c0[5] && !contains4(v, fc[1]))
        cp = sol_cp_pyr_f(V[4], V[1], V[0], ws[0], ws[4], ws[5]);
      else if (c0[1] && c1[5] && c0[6] && !contains4(v, fc[2]))
        cp = sol_cp_pyr_f(V[4], V[2], V[1], ws[1], ws[5], ws[6]);
      else if (c0[2] && c1[6] && c0[7] && !contains4(v, fc[3]))
        cp = sol_cp_pyr_f(V[4], V[3], V[2], ws[2], ws[6], ws[7]);
      else if (c0[3] && c1[7] && c0[4] && !contains4(v, fc[4]))
        cp = sol_cp_pyr_f(V[4], V[0], V[3], ws[3], ws[7], ws[4]);
      else
        found = 0;
    } else if (nv == 6) {
      if (c0[0] && c0[1] && c0[2] && !contains4(v, fc[0]))
        cp = sol_cp_tri_f(V[2], V[0], V[1], ws[0], ws[1], ws[2]);
      else if (c1[3] && c1[4] && c1[5] && !contains4(v, fc[1]))
        cp = sol_cp_tri_f(V[5], V[3], V[4], ws[3], ws[4], ws[5]);
      else if (c1[2] && c1[6] && c0[5] && c0[8] && !contains4(v, fc[2]))
        cp = sol_cp_sq_f(V[2], V[0], V[3], ws[2], ws[6], -ws[5], -ws[8], edges[5], edges[8],
                         raydir);
      else if (c1[0] && c1[7] && c0[3] && c0[6] && !contains4(v, fc[3]))
        cp = sol_cp_sq_f(V[0], V[1], V[4], ws[0], ws[7], -ws[3], -ws[6], edges[3], edges[6],
                         raydir);
      else if (c1[1] && c1[8] && c0[4] && c0[7] && !contains4(v, fc[4]))
        cp = sol_cp_sq_f(V[1], V[2], V[5], ws[1], ws[8], -ws[4], -ws[7], edges[4], edges[7],
                         raydir);
      else
        found = 0;
    } else {
      if (c0[0] && c0[1] && c0[2] && c0[3] && !contains4(v, fc[0]))
        cp = sol_cp_sq_f(V[0], V[1], V[2], ws[0], ws[1], ws[2], ws[3], edges[2], edges[3], raydir);
      else if (c1[4] && c1[5] && c1[6] && c1[7] && !contains4(v, fc[1]))
        cp = sol_cp_sq_f(V[4], V[5], V[6], ws[4], ws[5], ws[6], ws[7], edges[6], edges[7], raydir);
      else if (c1[0] && c1[9] && c0[4] && c0[8] && !contains4(v, fc[2]))
        cp = sol_cp_sq_f(V[0], V[1], V[5], ws[0], ws[9], -ws[4], -ws[8], edges[4], edges[8],
                         raydir);
      else if (c1[1] && c1[10] && c0[5] && c0[9] && !contains4(v, fc[3]))
        cp = sol_cp_sq_f(V[1], V[2], V[6], ws[1], ws[10], -ws[5], -ws[9], edges[5], edges[9],
                         raydir);
      else if (c1[2] && c1[11] && c0[6] && c0[10] && !contains4(v, fc[4]))
        cp = sol_cp_sq_f(V[2], V[3], V[7], ws[2], ws[11], -ws[6], -ws[10], edges[6], edges[10],
                         raydir);
      else if (c1[3] && c1[8] && c0[7] && c0[11] && !contains4(v, fc[5]))
        cp = sol_cp_sq_f(V[3], V[0], V[4], ws[3], ws[8], -ws[7], -ws[11], edges[7], edges[11],
                         raydir);
      else
        found = 0;
    }
    if (!found) {
      d_out[v] = 0;
      continue;
    }
    float dv = d / f3len(f3sub(cp, rayorg));
    if (dv < 0)
      dv = 0;
    else if (!(dv < 1) && !(dv > 0))
      dv = 1;
    d_out[v] = 1 - dv;
  }
  float ds = 0, w = 0, dd[8];
  for (int i = 0; i < nv; i++) ds += f3len(f3sub(pt, verts[i]));
  for (int i = 0; i < nv; i++) dd[i] = ds / f3len(f3sub(pt, verts[i]));
  for (int i = 0; i < nv; i++) {
    d_out[i] *= dd[i];
    w += d_out[i];
  }
  for (int i = 0; i < nv; i++) d_out[i] = d_out[i] / w;
  for (int i = nv; i < 8; i++) d_out[i] = 0;
}

/* SolidAccel::Traverse, render/render_accel_solid.cc:2345-2435, with
 * TestLeafNode :2206-2333. TestLeafNode always returns false (its `hit` is
 * shadowed inside the loop), so the search distance never narrows: every
 * node whose box the ray enters is visited, and the closest crossing wins
 * (entry before exit of one solid, strict <, first in visit order). u and v
 * come from never-written Intersection fields in the reference and are 0
 * here; subface_id is 0. */
static void trace_sol_one(const or_tet_node *nodes, const uint32_t *indices, const float *V,
                          const uint32_t *F, int nv, const float *scalars, const or_ray *r,
                          or_hit *h, or_sol_ext *ext, uint32_t *st) {
  ray_setup s;
  setup_ray(r, &s);
  h->t = FLT_MAX;
  h->u = 0.0f;
  h->v = 0.0f;
  h->prim_id = 0xffffffffu;
  float best_pos[3] = {0, 0, 0}, best_nrm[3] = {0, 0, 0};
  uint32_t best_side = 0;
  double bestv[24];
  int stack[STACK_MAX];
  int sp = 0;
  stack[0] = 0;
  uint32_t nvisit = 0, ntest = 0;
  const double ro[3] = {s.o[0], s.o[1], s.o[2]}, rd[3] = {s.d[0], s.d[1], s.d[2]};
  while (sp >= 0) {
    const or_tet_node *nd = nodes + stack[sp--];
    nvisit++;
    const int hit = ray_aabb(&s, FLT_MAX, nd->bmin, nd->bmax, 0);
    if (nd->flag == 0) {
      if (hit) {
        int nearc = s.sign[nd->axis], farc = 1 - nearc;
        stack[++sp] = (int)nd->data[farc];
        stack[++sp] = (int)nd->data[nearc];
      }
    } else if (hit) {
      float t = h->t;
      for (uint32_t i = 0; i < nd->data[0]; i++) {
        const uint32_t pi = indices[i + nd->data[1]];
        if (pi == r->prev_prim_id) continue;
        ntest++;
        double vtx[24];
        for (int j = 0; j < nv; j++) {
          const float *p = V + 3 * (size_t)F[(size_t)nv * pi + j];
          for (int k = 0; k < 3; k++) vtx[3 * j + k] = p[k];
        }
        float ts[2], pos[6], nrm[6];
        if (or_solid_isect(nv, ro, rd, vtx, ts, pos, nrm) != 1) continue;
        for (int k = 0; k < 2; k++) {
          if ((ts[k] >= 0.0) && (ts[k] < t)) {
            h->t = ts[k];
            h->prim_id = pi;
            memcpy(best_pos, pos + 3 * k, 12);
            memcpy(best_nrm, nrm + 3 * k, 12);
            best_side = (uint32_t)k;
            memcpy(bestv, vtx, sizeof(double) * 3 * nv);
            t = h->t;
          }
        }
      }
    }
  }
  if (st) {
    st[0] = nvisit;
    st[1] = ntest;
  }
  if (ext) {
    memset(ext, 0, sizeof(*ext));
    if (h->t < FLT_MAX) {
      or_interpolate_solid(nv, best_pos, bestv, ext->d);
      memcpy(ext->pos, best_pos, 12);
      memcpy(ext->nrm, best_nrm, 12);
      ext->side = best_side;
      if (scalars) {
        /* f_k = nv * prim + k (TestLeafNode :2284-2291), the varying sum of
         * gles/gles_common.cc:1078-1140 from left to right */
        float acc = scalars[(size_t)nv * h->prim_id] * ext->d[0];
        for (int k = 1; k < nv; k++) acc = acc + scalars[(size_t)nv * h->prim_id + k] * ext->d[k];
        ext->scalar = acc;
      }
    }
  }
}

void or_trace_sol(const or_tet_node *nodes, const uint32_t *indices, const float *V,
                  const uint32_t *F, int nv, const float *scalars, const or_ray *rays,
                  int64_t nrays, or_hit *hits, or_sol_ext *ext, uint32_t *stats, int threads) {
  (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
#endif
  for (int64_t i = 0; i < nrays; i++)
    trace_sol_one(nodes, indices, V, F, nv, scalars, rays + i, hits + i, ext ? ext + i : NULL,
                  stats ? stats + 2 * i : NULL);
}

/* ---- lines (capped cylinders) ---- */

/* The reference builds lines only with LineAccel::Build, a top-down
 * spatial-median tree (render/render_accel_line.cc:167-341); here they go
 * through the same Karras LBVH as the other kinds: Morton code of the
 * segment midpoint (p0 + p1) * 0.5f, leaves of r - l <= 4, depth cap 22, and
 * per-line boxes that hold the whole cylinder CylinderIsect intersects --
 * radius max(r0, r1) around both endpoints (the reference's
 * GetBoundingBoxOfLine, :83-121, uses r0 and r1 separately) -- widened by
 * 16 FLT_EPSILON of each bound's magnitude. */
static float lin_radius(const float *R, float rconst, uint32_t vtx) {
  return R ? R[vtx] : rconst;
}

static void lin_box(const float *V, const uint32_t *F, const float *R, float rconst, uint32_t i,
                    float mn[3], float mx[3]) {
  const float kE = FLT_EPSILON * 16.0f;
  const uint32_t a = F[2 * (size_t)i], b = F[2 * (size_t)i + 1];
  const float r0 = lin_radius(R, rconst, a), r1 = lin_radius(R, rconst, b);
  const float rr = r0 > r1 ? r0 : r1;
  for (int k = 0; k < 3; k++) {
    const float p = V[3 * (size_t)a + k], q = V[3 * (size_t)b + k];
    float lo = p - rr, hi = p + rr;
    const float lo2 = q - rr, hi2 = q + rr;
    if (lo2 < lo) lo = lo2;
    if (hi2 > hi) hi = hi2;
    mn[k] = lo - fabsf(lo) * kE;
    mx[k] = hi + fabsf(hi) * kE;
  }
}

void or_scene_bounds_lin(const float *V, const uint32_t *F, const float *R, float rconst,
                         int64_t n, float bmin[3], float bmax[3]) {
  lin_box(V, F, R, rconst, 0, bmin, bmax);
  for (int64_t i = 0; i < n; i++) {
    float mn[3], mx[3];
    lin_box(V, F, R, rconst, (uint32_t)i, mn, mx);
    for (int k = 0; k < 3; k++) {
      if (mn[k] < bmin[k]) bmin[k] = mn[k];
      if (mx[k] > bmax[k]) bmax[k] = mx[k];
    }
  }
}

void or_morton_lin(const float *V, const uint32_t *F, int64_t n, const float bmin[3],
                   const float bmax[3], uint32_t *codes) {
  float inv[3];
  morton_inv(bmin, bmax, inv);
  for (int64_t i = 0; i < n; i++) {
    const float *p = V + 3 * (size_t)F[2 * i], *q = V + 3 * (size_t)F[2 * i + 1];
    float c[3];
    for (int k = 0; k < 3; k++) c[k] = (p[k] + q[k]) * 0.5f;
    codes[i] = or_morton30_point(c, bmin, inv[0], inv[1], inv[2]);
  }
}

int64_t or_build_lin(const float *V, const uint32_t *F, const float *R, float rconst, int64_t n,
                     or_tet_node *nodes, uint32_t *indices, double bounds[6]) {
  build_ctx b;
  memset(&b, 0, sizeof(b));
  b.kind = 4;
  b.V = V;
  b.F = F;
  b.R = R;
  b.rconst = rconst;
  b.sidx = indices;
  b.max_leaf = TET_MAX_LEAF;
  b.nodes = nodes;
  return build_common(&b, n, bounds);
}

/* real3 helpers of render/render_accel_line.cc:422-454 (float) */
static float l3dot(const float a[3], const float b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static void l3normalize(float v[3], const float a[3]) {
  const float kE = FLT_EPSILON * 16.0f;
  const float len = (float)sqrt((double)l3dot(a, a));
  v[0] = a[0];
  v[1] = a[1];
  v[2] = a[2];
  if (len > kE) {
    const float inv = (float)(1.0 / len);
    v[0] *= inv;
    v[1] *= inv;
    v[2] *= inv;
  }
}

/* solve2e, render/render_accel_line.cc:456-486: A t^2 + 2 B t + C = 0 */
static int lin_solve2e(float root[2], float A, float B, float C) {
  if (fabsf(A) <= FLT_EPSILON) {
    root[0] = -C / B;
    return 1;
  }
  const float D = B * B - A * C;
  if (D < 0) return 0;
  if (D == 0) {
    root[0] = -B / A;
    return 1;
  }
  /* fabs and sqrt are the C library's double functions here, so the sum
   * and the division run in double and round once (matches the compiled
   * reference on every tested ray; the float reading differs in 1 ulp) */
  float x1 = (float)((fabs((double)B) + sqrt((double)D)) / (double)A);
  if (B >= 0.0) x1 = -x1;
  float x2 = C / (A * x1);
  if (x1 > x2) {
    const float tmp = x1;
    x1 = x2;
    x2 = tmp;
  }
  root[0] = x1;
  root[1] = x2;
  return 2;
}

/* CylinderIsect, render/render_accel_line.cc:488-577: the end caps (discs of
 * radius max(r0, r1) on the planes through p0 and p1 normal to the axis,
 * measured along the normalised direction, strict <), then the side (the
 * smaller root only, t <= tmax and <= the cap's t, axial fraction in [0, 1]). */
static int lin_isect(float *pt, float *pu, float *pv, int *hit_cap, const float p0[3],
                     const float p1[3], float rr, const float o[3], const float dir[3], int cap) {
  const float kEPSl = 1.0e-6f;
  const float tmax = *pt;
  float d[3], m[3];
  for (int k = 0; k < 3; k++) {
    d[k] = p1[k] - p0[k];
    m[k] = o[k] - p0[k];
  }
  const float md = l3dot(m, d), nd = l3dot(dir, d), dd = l3dot(d, d);
  *hit_cap = 0;
  float capT = FLT_MAX;
  if (cap) {
    float dn0[3], dn1[3], rd[3], tmp[3];
    for (int k = 0; k < 3; k++) tmp[k] = p0[k] - p1[k];
    l3normalize(dn0, tmp);
    for (int k = 0; k < 3; k++) dn1[k] = -dn0[k];
    l3normalize(rd, dir);
    if (fabsf(l3dot(dir, dn0)) > kEPSl) {
      const float p0D = -l3dot(p0, dn0), p1D = -l3dot(p1, dn1);
      const float p0T = -(l3dot(o, dn0) + p0D) / l3dot(rd, dn0);
      const float p1T = -(l3dot(o, dn1) + p1D) / l3dot(rd, dn1);
      float q0[3], q1[3], e0[3], e1[3];
      for (int k = 0; k < 3; k++) {
        q0[k] = o[k] + p0T * rd[k];
        q1[k] = o[k] + p1T * rd[k];
        e0[k] = q0[k] - p0[k];
        e1[k] = q1[k] - p1[k];
      }
      const float qp0 = l3dot(e0, e0), qp1 = l3dot(e1, e1);
      if (p0T > 0.0 && p0T < tmax && (qp0 < rr * rr)) {
        *hit_cap = 1;
        capT = p0T;
        *pt = capT;
        *pu = (float)sqrt((double)qp0);
        *pv = 0;
      }
      if (p1T > 0.0 && p1T < tmax && p1T < capT && (qp1 < rr * rr)) {
        *hit_cap = 1;
        capT = p1T;
        *pt = capT;
        *pu = (float)sqrt((double)qp1);
        *pv = 1.0f;
      }
    }
  }
  if (md <= 0.0 && nd <= 0.0) return *hit_cap;
  if (md >= dd && nd >= 0.0) return *hit_cap;
  const float nn = l3dot(dir, dir), mn = l3dot(m, dir);
  const float A = dd * nn - nd * nd;
  const float kk = l3dot(m, m) - rr * rr;
  const float C = dd * kk - md * md;
  const float B = dd * mn - nd * md;
  float root[2] = {0, 0};
  if (lin_solve2e(root, A, B, C)) {
    const float t = root[0];
    if (0 <= t && t <= tmax && t <= capT) {
      float s = md + t * nd;
      s /= dd;
      if (0 <= s && s <= 1) {
        *hit_cap = 0;
        *pt = t;
        *pu = 0;
        *pv = s;
        return 1;
      }
    }
  }
  return *hit_cap;
}

int or_line_isect(const float p0[3], const float p1[3], float rr, const float o[3],
                  const float d[3], int cap, float tmax, float out[3], int *hit_cap) {
  float t = tmax, u = 0, v = 0;
  const int h = lin_isect(&t, &u, &v, hit_cap, p0, p1, rr, o, d, cap);
  out[0] = t;
  out[1] = u;
  out[2] = v;
  return h;
}

/* LineAccel::Traverse's TestLeafNode (render/render_accel_line.cc:579-633)
 * and BuildIntersection (:654-702) over the LBVH: every node whose box the
 * ray enters (IntersectRayAABB, :377-420) is visited, nearer child first,
 * with no narrowing by the hit distance (a computed crossing of a grazing
 * ray may lie outside its padded box), so the closest crossing over all
 * lines is found; ties follow this visit order (caps: first, side: last). */
static void trace_lin_one(const or_tet_node *nodes, const uint32_t *indices, const float *V,
                          const uint32_t *F, const float *R, float rconst, int cap,
                          const or_ray *r, or_hit *h, or_lin_ext *ext, uint32_t *st) {
  ray_setup s;
  setup_ray(r, &s);
  float t = FLT_MAX, bu = 0, bv = 0;
  uint32_t best = 0xffffffffu;
  int best_cap = 0;
  int stack[STACK_MAX];
  int sp = 0;
  stack[0] = 0;
  uint32_t nvisit = 0, ntest = 0;
  while (sp >= 0) {
    const or_tet_node *nd = nodes + stack[sp--];
    nvisit++;
    if (!ray_aabb(&s, FLT_MAX, nd->bmin, nd->bmax, 0)) continue;
    if (nd->flag == 0) {
      int nearc = s.sign[nd->axis], farc = 1 - nearc;
      stack[++sp] = (int)nd->data[farc];
      stack[++sp] = (int)nd->data[nearc];
      continue;
    }
    for (uint32_t i = 0; i < nd->data[0]; i++) {
      const uint32_t li = indices[i + nd->data[1]];
      ntest++;
      const uint32_t a = F[2 * (size_t)li], b = F[2 * (size_t)li + 1];
      const float r0 = lin_radius(R, rconst, a), r1 = lin_radius(R, rconst, b);
      float u, v;
      int hc;
      if (lin_isect(&t, &u, &v, &hc, V + 3 * (size_t)a, V + 3 * (size_t)b, r0 > r1 ? r0 : r1,
                    s.o, s.d, cap)) {
        bu = u;
        bv = v;
        best = li;
        best_cap = hc;
      }
    }
  }
  h->t = t;
  h->u = best != 0xffffffffu ? bu : 0.0f;
  h->v = best != 0xffffffffu ? bv : 0.0f;
  h->prim_id = best;
  if (st) {
    st[0] = nvisit;
    st[1] = ntest;
  }
  if (ext) {
    memset(ext, 0, sizeof(*ext));
    if (best != 0xffffffffu) {
      const float *p0 = V + 3 * (size_t)F[2 * (size_t)best];
      const float *p1 = V + 3 * (size_t)F[2 * (size_t)best + 1];
      float pos[3], n[3], dl[3];
      for (int k = 0; k < 3; k++) {
        pos[k] = r->org[k] + t * r->dir[k];
        dl[k] = p1[k] - p0[k];
      }
      if (best_cap) {
        float c[3], pc[3];
        for (int k = 0; k < 3; k++) {
          c[k] = 0.5f * dl[k] + p0[k];
          pc[k] = pos[k] - c[k];
        }
        l3normalize(n, dl);
        if (!(l3dot(pc, n) > 0.0))
          for (int k = 0; k < 3; k++) n[k] = -n[k];
      } else {
        float pc[3];
        for (int k = 0; k < 3; k++) pc[k] = pos[k] - (p0[k] + bv * dl[k]);
        l3normalize(n, pc);
      }
      memcpy(ext->pos, pos, 12);
      memcpy(ext->nrm, n, 12);
      ext->hit_cap = (uint32_t)best_cap;
    }
  }
}

void or_trace_lin(const or_tet_node *nodes, const uint32_t *indices, const float *V,
                  const uint32_t *F, const float *R, float rconst, int cap, const or_ray *rays,
                  int64_t nrays, or_hit *hits, or_lin_ext *ext, uint32_t *stats, int threads) {
  (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
#endif
  for (int64_t i = 0; i < nrays; i++)
    trace_lin_one(nodes, indices, V, F, R, rconst, cap, rays + i, hits + i, ext ? ext + i : NULL,
                  stats ? stats + 2 * i : NULL);
}

/* ---- spheres ---- */

/* SphereIsect, render/render_accel_particle.cc:2423-2512 */
static int sphere_isect(float *tio, const float *c, float radius,
                        const ray_setup *s, double prevT, const float *prevN,
                        int self) {
  float oc[3];
  for (int k = 0; k < 3; k++) oc[k] = s->o[k] - c[k];
  const float *d = s->d;
  float a = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
  float b = (float)(2.0 * (double)((d[0] * oc[0] + d[1] * oc[1]) + d[2] * oc[2]));
  float cc = ((oc[0] * oc[0] + oc[1] * oc[1]) + oc[2] * oc[2]) - radius * radius;
  float disc = (float)((double)(b * b) - (4.0 * (double)a) * (double)cc);
  float t0, t1;
  if (disc < 0.0) return 0;
  if (disc == 0.0) {
    t0 = t1 = (float)(-0.5 * (double)(b / a));
  } else {
    float dsq = (float)sqrt((double)disc);
    float q;
    if (b < 0)
      q = (float)((double)(-b - dsq) / 2.0);
    else
      q = (float)((double)(-b + dsq) / 2.0);
    t0 = q / a;
    t1 = cc / q;
  }
  if (t0 > t1) {
    float tmp = t0;
    t0 = t1;
    t1 = tmp;
  }
  if (t1 < 0) return 0;
  if (self) {
    float t = t1;
    float nrm[3];
    for (int k = 0; k < 3; k++) nrm[k] = (s->o[k] + t * d[k]) - c[k];
    float nd = (nrm[0] * prevN[0] + nrm[1] * prevN[1]) + nrm[2] * prevN[2];
    if ((t < *tio) && ((double)t < prevT) && (nd <= 0.0)) {
      *tio = t;
      return 1;
    }
  } else {
    float t = (t0 < 0) ? t1 : t0;
    if (t < *tio) {
      *tio = t;
      return 1;
    }
  }
  return 0;
}

/* ParticleAccel::Traverse, render/render_accel_particle.cc:3053-3218 with the
 * scalar TestLeafNode :2809-2829; u, v from BuildIntersection :2968-3049. */
static void trace_sph_one(const or_sph_node *nodes, const uint32_t *indices,
                          const float *P, const float *R, float rconst,
                          const or_ray *r, double prevT, const float *prevN,
                          or_hit *h, uint32_t *st) {
  static const float zero3[3] = {0, 0, 0};
  ray_setup s;
  setup_ray(r, &s);
  float hitT = FLT_MAX;
  h->t = hitT;
  h->u = 0.0f;
  h->v = 0.0f;
  h->prim_id = 0xffffffffu;
  int stack[STACK_MAX];
  int sp = 0;
  stack[0] = 0;
  uint32_t nvisit = 0, ntest = 0;
  if (!prevN) prevN = zero3;
  while (sp >= 0) {
    const or_sph_node *nd = nodes + stack[sp--];
    nvisit++;
    if (nd->flag == 0) {
      int hit0 = ray_aabb(&s, hitT, nd->bmin[0], nd->bmax[0], 1);
      int hit1 = ray_aabb(&s, hitT, nd->bmin[1], nd->bmax[1], 1);
      int hits[2] = {hit0, hit1};
      if (hit0 || hit1) {
        int nearc = s.sign[nd->axis], farc = 1 - nearc;
        if (hits[farc]) stack[++sp] = (int)nd->data[farc];
        if (hits[nearc]) stack[++sp] = (int)nd->data[nearc];
      }
    } else {
      float t = h->t;
      int any = 0;
      for (uint32_t i = 0; i < nd->data[0]; i++) {
        uint32_t pi = indices[i + nd->data[1]];
        ntest++;
        if (sphere_isect(&t, P + 3 * (size_t)pi, sph_radius(R, rconst, pi), &s,
                         prevT, prevN, r->prev_prim_id == pi)) {
          h->t = t;
          h->prim_id = pi;
          any = 1;
        }
      }
      if (any) hitT = h->t;
    }
  }
  if (h->t < FLT_MAX) {
    /* vnormalize, render/render_accel_particle.cc:2041-2050, then the
     * spherical u, v of BuildIntersection (libm; tolerance-only parity). */
    float cp[3];
    for (int k = 0; k < 3; k++)
      cp[k] = (s.o[k] + h->t * s.d[k]) - P[3 * (size_t)h->prim_id + k];
    const float kEPS = (float)(FLT_EPSILON * 16.0);
    float len = (float)sqrt((double)((cp[0] * cp[0] + cp[1] * cp[1]) + cp[2] * cp[2]));
    float n[3] = {cp[0], cp[1], cp[2]};
    if (len > kEPS) {
      float invlen = (float)(1.0 / (double)len);
      for (int k = 0; k < 3; k++) n[k] = cp[k] * invlen;
    }
    h->u = (float)((atan2((double)n[0], (double)n[2]) + M_PI) * 0.5 * (1.0 / M_PI));
    h->v = (float)(acos((double)n[1]) / M_PI);
  }
  if (st) {
    st[0] = nvisit;
    st[1] = ntest;
  }
}

void or_trace_sph(const or_sph_node *nodes, const uint32_t *indices,
                  const float *P, const float *R, float rconst,
                  const or_ray *rays, const double *prev_hit_t,
                  const float *prev_hit_n, int64_t nrays, or_hit *hits,
                  uint32_t *stats, int threads) {
  (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
#endif
  for (int64_t i = 0; i < nrays; i++)
    trace_sph_one(nodes, indices, P, R, rconst, rays + i,
                  prev_hit_t ? prev_hit_t[i] : 0.0,
                  prev_hit_n ? prev_hit_n + 3 * i : NULL, hits + i,
                  stats ? stats + 2 * i : NULL);
}

/* ------------------------------------------------------------------------ */
/* 3D texture and volume march                                              */
/* ------------------------------------------------------------------------ */

/* fasterfloorf, render/render_texture.cc:108-121 */
static int faster_floorf(float x) {
  if (x >= 0) return (int)x;
  int y = (int)x;
  if (fabsf(x - (float)y) <= FLT_EPSILON) {
  } else {
    y = y - 1;
  }
  return y;
}

static int clampi(int v, int lo, int hi) {
  int a = v < lo ? lo : v; /* std::max(v, low) */
  return hi < a ? hi : a;  /* std::min(., high) */
}

/* Geometry of one trilinear fetch: the 8 corner addresses and fractions.
 * FilterTexture3DFloat, render/render_texture.cc:946-1011 */
typedef struct {
  int64_t a[8]; /* 000 100 010 110 001 101 011 111 */
  float dx, dy, dz;
} tri_geom;

static void tex3d_geom(const int dims[3], const int clamp[3], float u, float v,
                       float r, tri_geom *g) {
  float uu = clamp[0] ? u : u - (float)faster_floorf(u);
  float vv = clamp[1] ? v : v - (float)faster_floorf(v);
  float rr = clamp[2] ? r : r - (float)faster_floorf(r);
  float vox0 = uu * ((float)dims[0] - 1.0f);
  float vox1 = vv * ((float)dims[1] - 1.0f);
  float vox2 = rr * ((float)dims[2] - 1.0f);
  int vx = (int)vox0, vy = (int)vox1, vz = (int)vox2;
  vx = clampi(vx, 0, dims[0] - 1);
  vy = clampi(vy, 0, dims[1] - 1);
  vz = clampi(vz, 0, dims[2] - 1);
  int vx1 = (vx + 1 < dims[0] - 1) ? vx + 1 : dims[0] - 1;
  int vy1 = (vy + 1 < dims[1] - 1) ? vy + 1 : dims[1] - 1;
  int vz1 = (vz + 1 < dims[2] - 1) ? vz + 1 : dims[2] - 1;
  g->dx = vox0 - (float)vx;
  g->dy = vox1 - (float)vy;
  g->dz = vox2 - (float)vz;
  int64_t sx = 1, sy = dims[0], sz = (int64_t)dims[0] * dims[1];
  int xs[2] = {vx, vx1}, ys[2] = {vy, vy1}, zs[2] = {vz, vz1};
  for (int c = 0; c < 8; c++)
    g->a[c] = zs[(c >> 2) & 1] * sz + ys[(c >> 1) & 1] * sy + xs[c & 1] * sx;
}

/* Lerp, render/render_texture.cc:97 */
static float lerpf(float t, float a, float b) { return (1.f - t) * a + t * b; }

static float tex3d_eval(const float *vox, const tri_geom *g) {
  float d00 = lerpf(g->dx, vox[g->a[0]], vox[g->a[1]]);
  float d10 = lerpf(g->dx, vox[g->a[2]], vox[g->a[3]]);
  float d01 = lerpf(g->dx, vox[g->a[4]], vox[g->a[5]]);
  float d11 = lerpf(g->dx, vox[g->a[6]], vox[g->a[7]]);
  float d0 = lerpf(g->dy, d00, d10);
  float d1 = lerpf(g->dy, d01, d11);
  return lerpf(g->dz, d0, d1);
}

void or_tex3d_sample(const float *vox, const int dims[3], const int clamp[3],
                     const float *uvw, int64_t n, float *out) {
  for (int64_t i = 0; i < n; i++) {
    tri_geom g;
    tex3d_geom(dims, clamp, uvw[3 * i], uvw[3 * i + 1], uvw[3 * i + 2], &g);
    out[i] = tex3d_eval(vox, &g);
  }
}

/* Texture::Fetch (gles/gles_common.h:344-403) for a dense 3D texture: the
 * optional per-axis remap (local::remap, :263-269: idx = (int)(x * n)
 * clamped to [0, n-1], table[idx]), then FilterTexture3DByte (voxel / 255.5f,
 * render/render_texture.cc:410-413,832-944), FilterTexture3DFloat
 * (:946-1058) or FilterTexture3DDouble (double lerps from the float
 * fractions, :1060-1172) with `comps` interleaved components; one component
 * is splatted to RGBA, more fill rgba[0..comps) (the rest 0). format: 1 byte,
 * 8 float, 16 double (LSGL_RENDER_TEXTURE3D_FORMAT_*). */
static float remap_coord(float x, const float *table, int n) {
  int idx = (int)(x * (float)n);
  idx = idx < n - 1 ? idx : n - 1; /* std::min(n - 1, idx) */
  idx = idx > 0 ? idx : 0;         /* std::max(., 0) */
  return table[idx];
}

static double dlerp_(double t, double a, double b) { return (1.0 - t) * a + t * b; }

void or_tex3d_fetch(const void *image, int format, int comps, const int dims[3],
                    const int clamp[3], const float *const remap[3], const int remap_n[3],
                    const float *uvw, int64_t n, float *rgba) {
  for (int64_t i = 0; i < n; i++) {
    float c3[3] = {uvw[3 * i], uvw[3 * i + 1], uvw[3 * i + 2]};
    for (int k = 0; k < 3; k++)
      if (remap && remap[k] && remap_n[k] > 0) c3[k] = remap_coord(c3[k], remap[k], remap_n[k]);
    tri_geom g;
    tex3d_geom(dims, clamp, c3[0], c3[1], c3[2], &g);
    float vals[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int c = 0; c < comps; c++) {
      if (format == 16) {
        const double *d = (const double *)image;
        double D[8];
        for (int q = 0; q < 8; q++) D[q] = d[g.a[q] * comps + c];
        double dx = g.dx, dy = g.dy, dz = g.dz;
        double d00 = dlerp_(dx, D[0], D[1]), d10 = dlerp_(dx, D[2], D[3]);
        double d01 = dlerp_(dx, D[4], D[5]), d11 = dlerp_(dx, D[6], D[7]);
        double d0 = dlerp_(dy, d00, d10), d1 = dlerp_(dy, d01, d11);
        vals[c] = (float)dlerp_(dz, d0, d1);
      } else {
        float D[8];
        for (int q = 0; q < 8; q++) {
          const int64_t e = g.a[q] * comps + c;
          D[q] = format == 1 ? ((const unsigned char *)image)[e] / 255.5f
                             : ((const float *)image)[e];
        }
        float d00 = lerpf(g.dx, D[0], D[1]), d10 = lerpf(g.dx, D[2], D[3]);
        float d01 = lerpf(g.dx, D[4], D[5]), d11 = lerpf(g.dx, D[6], D[7]);
        float d0 = lerpf(g.dy, d00, d10), d1 = lerpf(g.dy, d01, d11);
        vals[c] = lerpf(g.dz, d0, d1);
      }
    }
    for (int k = 0; k < 4; k++) rgba[4 * i + k] = comps == 1 ? vals[0] : vals[k];
  }
}

/* Volume march spec (builder-defined; the reference leaves marching to user
 * GLSL, e.g. examples/sparse_volume_render/input.frag:99-141):
 *  - the volume fills texture space [0,1]^3, rays are given in that space;
 *  - entry/exit by a float slab test against [0,1]^3, tnear clamped to 0;
 *  - sample k at t_k = tnear + ((float)k + 0.5f) * step while t_k < tfar;
 *  - s_k = FilterTexture3DFloat at o + t_k * d (single component);
 *  - TF lookup x = clamp(s, 0, 1) * (N - 1), i0 = min((int)x, N - 2),
 *    f = x - i0, rgba = Lerp(f, tf[i0], tf[i0 + 1]) per channel;
 *  - front to back: w = T * a; C += w * rgb; T *= (1 - a); out = (C, 1 - T).
 */
#define VOL_MAX_SAMPLES 65536

static int vol_span(const or_ray *r, float *tnear, float *tfar) {
  float tn = 0.0f, tf = FLT_MAX;
  for (int k = 0; k < 3; k++) {
    float o = r->org[k], d = r->dir[k];
    if (d == 0.0f) {
      if (o < 0.0f || o > 1.0f) return 0;
      continue;
    }
    float inv = 1.0f / d;
    float t0 = (0.0f - o) * inv, t1 = (1.0f - o) * inv;
    float lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
    if (lo > tn) tn = lo;
    if (hi < tf) tf = hi;
  }
  *tnear = tn;
  *tfar = tf;
  return tn < tf;
}

static float vol_sample_t(float tnear, int k, float step) {
  return tnear + ((float)k + 0.5f) * step;
}

static void tf_index(float s, int N, int *i0, float *f, int *inside) {
  float x = s * (float)(N - 1);
  *inside = (x > 0.0f) && (x < (float)(N - 1));
  if (!(x > 0.0f)) x = 0.0f;
  if (x > (float)(N - 1)) x = (float)(N - 1);
  int i = (int)x;
  if (i > N - 2) i = N - 2;
  *i0 = i;
  *f = x - (float)i;
}

static void vol_geom_at(const or_vol_params *p, const or_ray *r, float t,
                        tri_geom *g) {
  float pos[3];
  for (int k = 0; k < 3; k++) pos[k] = r->org[k] + t * r->dir[k];
  tex3d_geom(p->dims, p->clamp, pos[0], pos[1], pos[2], g);
}

void or_volume_fwd_f32(const or_vol_params *p, const float *vox,
                       const float *tf, const or_ray *rays, int64_t nrays,
                       float *rgba, uint32_t *nsamples, int threads) {
  (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
#endif
  for (int64_t i = 0; i < nrays; i++) {
    const or_ray *r = rays + i;
    float C[3] = {0, 0, 0}, T = 1.0f;
    float tn, tfar;
    uint32_t ns = 0;
    if (vol_span(r, &tn, &tfar)) {
      for (int k = 0; k < VOL_MAX_SAMPLES; k++) {
        float t = vol_sample_t(tn, k, p->step);
        if (!(t < tfar)) break;
        tri_geom g;
        vol_geom_at(p, r, t, &g);
        float s = tex3d_eval(vox, &g);
        int i0, inside;
        float f;
        tf_index(s, p->tf_size, &i0, &f, &inside);
        float c[4];
        for (int ch = 0; ch < 4; ch++)
          c[ch] = lerpf(f, tf[4 * i0 + ch], tf[4 * (i0 + 1) + ch]);
        float w = T * c[3];
        for (int ch = 0; ch < 3; ch++) C[ch] = C[ch] + w * c[ch];
        T = T * (1.0f - c[3]);
        ns++;
      }
    }
    rgba[4 * i + 0] = C[0];
    rgba[4 * i + 1] = C[1];
    rgba[4 * i + 2] = C[2];
    rgba[4 * i + 3] = 1.0f - T;
    if (nsamples) nsamples[i] = ns;
  }
}

/* The f64 checker: identical sample geometry (computed in float as above),
 * values, TF and compositing in double. */
static double lerpd(double t, double a, double b) { return (1.0 - t) * a + t * b; }

static double tex3d_eval_d(const double *vox, const tri_geom *g, double w[8]) {
  double dx = g->dx, dy = g->dy, dz = g->dz;
  w[0] = (1 - dx) * (1 - dy) * (1 - dz);
  w[1] = dx * (1 - dy) * (1 - dz);
  w[2] = (1 - dx) * dy * (1 - dz);
  w[3] = dx * dy * (1 - dz);
  w[4] = (1 - dx) * (1 - dy) * dz;
  w[5] = dx * (1 - dy) * dz;
  w[6] = (1 - dx) * dy * dz;
  w[7] = dx * dy * dz;
  double d00 = lerpd(dx, vox[g->a[0]], vox[g->a[1]]);
  double d10 = lerpd(dx, vox[g->a[2]], vox[g->a[3]]);
  double d01 = lerpd(dx, vox[g->a[4]], vox[g->a[5]]);
  double d11 = lerpd(dx, vox[g->a[6]], vox[g->a[7]]);
  return lerpd(dz, lerpd(dy, d00, d10), lerpd(dy, d01, d11));
}

static void tf_index_d(double s, int N, int *i0, double *f, int *inside) {
  double x = s * (double)(N - 1);
  *inside = (x > 0.0) && (x < (double)(N - 1));
  if (!(x > 0.0)) x = 0.0;
  if (x > (double)(N - 1)) x = (double)(N - 1);
  int i = (int)x;
  if (i > N - 2) i = N - 2;
  *i0 = i;
  *f = x - (double)i;
}

void or_volume_fwd_f64(const or_vol_params *p, const double *vox,
                       const double *tf, const or_ray *rays, int64_t nrays,
                       double *rgba, int threads) {
  (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
#endif
  for (int64_t i = 0; i < nrays; i++) {
    const or_ray *r = rays + i;
    double C[3] = {0, 0, 0}, T = 1.0;
    float tn, tfar;
    if (vol_span(r, &tn, &tfar)) {
      for (int k = 0; k < VOL_MAX_SAMPLES; k++) {
        float t = vol_sample_t(tn, k, p->step);
        if (!(t < tfar)) break;
        tri_geom g;
        vol_geom_at(p, r, t, &g);
        double w8[8];
        double s = tex3d_eval_d(vox, &g, w8);
        int i0, inside;
        double f;
        tf_index_d(s, p->tf_size, &i0, &f, &inside);
        double c[4];
        for (int ch = 0; ch < 4; ch++)
          c[ch] = lerpd(f, tf[4 * i0 + ch], tf[4 * (i0 + 1) + ch]);
        double w = T * c[3];
        for (int ch = 0; ch < 3; ch++) C[ch] += w * c[ch];
        T *= (1.0 - c[3]);
      }
    }
    rgba[4 * i + 0] = C[0];
    rgba[4 * i + 1] = C[1];
    rgba[4 * i + 2] = C[2];
    rgba[4 * i + 3] = 1.0 - T;
  }
}

/* Adjoint of or_volume_fwd_f64 for L = sum grad_rgba . rgba, by the reverse
 * recursion R_{k-1} = a_k c_k + (1 - a_k) R_k, P_{k-1} = (1 - a_k) P_k:
 *   dL/dc_k = gC T_k a_k,  dL/da_k = T_k (gC . (c_k - R_k) + gA P_k).
 * Stored forward T_k here (it is the checker; the GPU recomputes them). */
void or_volume_bwd_f64(const or_vol_params *p, const double *vox,
                       const double *tf, const or_ray *rays, int64_t nrays,
                       const double *grad_rgba, double *grad_vox,
                       double *grad_tf) {
  const int N = p->tf_size;
  double *Tk = (double *)malloc(sizeof(double) * VOL_MAX_SAMPLES);
  for (int64_t i = 0; i < nrays; i++) {
    const or_ray *r = rays + i;
    float tn, tfar;
    if (!vol_span(r, &tn, &tfar)) continue;
    int ns = 0;
    double T = 1.0;
    for (int k = 0; k < VOL_MAX_SAMPLES; k++) {
      float t = vol_sample_t(tn, k, p->step);
      if (!(t < tfar)) break;
      tri_geom g;
      vol_geom_at(p, r, t, &g);
      double w8[8];
      double s = tex3d_eval_d(vox, &g, w8);
      int i0, inside;
      double f;
      tf_index_d(s, N, &i0, &f, &inside);
      double a = lerpd(f, tf[4 * i0 + 3], tf[4 * (i0 + 1) + 3]);
      Tk[k] = T;
      T *= (1.0 - a);
      ns++;
    }
    const double *gr = grad_rgba + 4 * i;
    double R[3] = {0, 0, 0}, P = 1.0;
    for (int k = ns - 1; k >= 0; k--) {
      float t = vol_sample_t(tn, k, p->step);
      tri_geom g;
      vol_geom_at(p, r, t, &g);
      double w8[8];
      double s = tex3d_eval_d(vox, &g, w8);
      int i0, inside;
      double f;
      tf_index_d(s, N, &i0, &f, &inside);
      double c[4];
      for (int ch = 0; ch < 4; ch++)
        c[ch] = lerpd(f, tf[4 * i0 + ch], tf[4 * (i0 + 1) + ch]);
      double a = c[3];
      double gc[4];
      double dot = 0.0;
      for (int ch = 0; ch < 3; ch++) {
        gc[ch] = gr[ch] * Tk[k] * a;
        dot += gr[ch] * (c[ch] - R[ch]);
      }
      gc[3] = Tk[k] * (dot + gr[3] * P);
      for (int ch = 0; ch < 4; ch++) {
        grad_tf[4 * i0 + ch] += (1.0 - f) * gc[ch];
        grad_tf[4 * (i0 + 1) + ch] += f * gc[ch];
      }
      if (inside) {
        double ds = 0.0;
        for (int ch = 0; ch < 4; ch++)
          ds += gc[ch] * (tf[4 * (i0 + 1) + ch] - tf[4 * i0 + ch]) * (double)(N - 1);
        for (int c8 = 0; c8 < 8; c8++) grad_vox[g.a[c8]] += ds * w8[c8];
      }
      for (int ch = 0; ch < 3; ch++) R[ch] = a * c[ch] + (1.0 - a) * R[ch];
      P = (1.0 - a) * P;
    }
  }
  free(Tk);
}


/* ===================================================================
 * Two-level scene
 * =================================================================== */

/* Matrixd::MultV, render/render_matrix.h:165-183: row vector times the
 * matrix, sums left to right. */
static void mult_v(const double m[16], const double v[3], double out[3]) {
  for (int k = 0; k < 3; k++)
    out[k] = ((m[0 * 4 + k] * v[0] + m[1 * 4 + k] * v[1]) + m[2 * 4 + k] * v[2]) + m[3 * 4 + k];
}

/* std::min / std::max: keep the first unless the second is strictly
 * smaller / larger. */
static double dmin_ref(double a, double b) { return (b < a) ? b : a; }
static double dmax_ref(double a, double b) { return (a < b) ? b : a; }

/* TransformBoundingBox, gles/gles_render_graph.cc:19-79: the 8 corners
 * (x fastest, then y, then z) transformed, min/max from corner 0 on. */
void or_transform_box(const double bmin[3], const double bmax[3], const double T[16],
                      double out[6]) {
  double xb[8][3];
  for (int c = 0; c < 8; c++) {
    double p[3] = {(c & 1) ? bmax[0] : bmin[0], (c & 2) ? bmax[1] : bmin[1],
                   (c & 4) ? bmax[2] : bmin[2]};
    mult_v(T, p, xb[c]);
  }
  for (int k = 0; k < 3; k++) {
    out[k] = xb[0][k];
    out[3 + k] = xb[0][k];
  }
  for (int c = 1; c < 8; c++)
    for (int k = 0; k < 3; k++) {
      out[k] = dmin_ref(xb[c][k], out[k]);
      out[3 + k] = dmax_ref(xb[c][k], out[3 + k]);
    }
}

typedef struct {
  double bmin[3], bmax[3];
  uint32_t id;
} tlas_item; /* BVHData */

/* CalculateBoundingBox over items [lo, hi), render/render_bvh_tree.cc:129-147 */
static void items_box(const tlas_item *it, int64_t lo, int64_t hi, double mn[3], double mx[3]) {
  for (int k = 0; k < 3; k++) {
    mn[k] = it[lo].bmin[k];
    mx[k] = it[lo].bmax[k];
  }
  for (int64_t i = lo + 1; i < hi; i++)
    for (int k = 0; k < 3; k++) {
      mn[k] = dmin_ref(mn[k], it[i].bmin[k]);
      mx[k] = dmax_ref(mx[k], it[i].bmax[k]);
    }
}

/* PartitionNodes, render/render_bvh_tree.cc:83-127: items after the first
 * whose bmax on the axis lies below the split move forward (swaps), the
 * first item is exchanged with the last moved one, and the split index
 * counts every item of the range below the split. */
static int64_t partition_items(tlas_item *it, int64_t lo, int64_t hi, double pos, int axis) {
  tlas_item first = it[lo];
  int64_t pivot = lo;
  for (int64_t i = lo + 1; i < hi; i++) {
    if (it[i].bmax[axis] < pos) {
      pivot++;
      tlas_item t = it[pivot];
      it[pivot] = it[i];
      it[i] = t;
    }
  }
  it[lo] = it[pivot];
  it[pivot] = first;
  int64_t split = lo;
  for (int64_t i = lo; i < hi; i++)
    if (it[i].bmax[axis] < pos) split++;
  return split;
}

typedef struct {
  or_tlas_node *nodes;
  int64_t count;
  tlas_item *items;
} tlas_builder;

static void pad_box(double mn[3], double mx[3]) {
  const double eps = 1.0e-6; /* render/render_bvh_tree.cc:404-410 */
  for (int k = 0; k < 3; k++) {
    double ext = mx[k] - mn[k];
    mn[k] = mn[k] - ext * eps;
    mx[k] = mx[k] + ext * eps;
  }
}

/* ConstructTree, render/render_bvh_tree.cc:319-472: spatial median of the
 * node box along the longest axis, then the next two axes in turn; object
 * median when no axis separates; children appended depth first. */
static void tlas_construct(tlas_builder *b, int64_t root, int64_t lo, int64_t hi) {
  static const int order[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
  const int64_t n = hi - lo;
  or_tlas_node *nd = b->nodes + root;
  if (n == 1) {
    nd->is_leaf = 1;
    nd->child[0] = nd->child[1] = 0;
    nd->node_id = b->items[lo].id;
    return;
  }
  double size[3], mn[3];
  for (int k = 0; k < 3; k++) {
    size[k] = nd->bmax[k] - nd->bmin[k];
    mn[k] = nd->bmin[k];
  }
  double longest = size[0];
  int first_axis = 0;
  if (longest < size[1]) {
    longest = size[1];
    first_axis = 1;
  }
  if (longest < size[2]) first_axis = 2;
  int64_t split = lo, nleft = 0, nright = n;
  int axis = 0;
  for (int i = 0; i < 3; i++) {
    axis = order[first_axis][i];
    double pos = mn[axis] + size[axis] * 0.5;
    split = partition_items(b->items, lo, hi, pos, axis);
    if (split != lo && split != hi) {
      nleft = split - lo;
      nright = n - nleft;
      break;
    }
  }
  if (split == lo || split == hi) {
    nleft = n / 2;
    nright = n - nleft;
    split = lo + (hi - lo) / 2;
  }
  b->nodes[root].axis = axis;
  b->nodes[root].is_leaf = 0;
  if (nleft > 0) {
    int64_t c = b->count++;
    or_tlas_node *cn = b->nodes + c;
    items_box(b->items, lo, split, cn->bmin, cn->bmax);
    pad_box(cn->bmin, cn->bmax);
    b->nodes[root].child[0] = (int32_t)c;
    tlas_construct(b, c, lo, split);
  } else {
    b->nodes[root].child[0] = 0;
  }
  if (nright > 0) {
    int64_t c = b->count++;
    or_tlas_node *cn = b->nodes + c;
    items_box(b->items, split, hi, cn->bmin, cn->bmax);
    pad_box(cn->bmin, cn->bmax);
    b->nodes[root].child[1] = (int32_t)c;
    tlas_construct(b, c, split, hi);
  } else {
    b->nodes[root].child[1] = 0;
  }
}

/* BVHTree::BuildTree, render/render_bvh_tree.cc:478-502: root = unpadded
 * scene box (CalcSceneBBox). */
int64_t or_tlas_build(const double *boxes, int64_t n, or_tlas_node *nodes) {
  if (n <= 0) return 0;
  tlas_item *items = (tlas_item *)malloc(sizeof(tlas_item) * (size_t)n);
  for (int64_t i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) {
      items[i].bmin[k] = boxes[6 * i + k];
      items[i].bmax[k] = boxes[6 * i + 3 + k];
    }
    items[i].id = (uint32_t)i;
  }
  memset(nodes, 0, sizeof(or_tlas_node) * (size_t)(2 * n));
  tlas_builder b = {nodes, 1, items};
  items_box(items, 0, n, nodes[0].bmin, nodes[0].bmax);
  tlas_construct(&b, 0, 0, n);
  free(items);
  return b.count;
}

typedef struct {
  int sign[3];
  double org[3], inv[3];
} tlas_ray;

/* SetupRay, render/render_bvh_tree.cc:40-71 (kEPS = 2e-15) */
static void tlas_setup(const double org[3], const double dir[3], tlas_ray *r) {
  for (int k = 0; k < 3; k++) {
    r->org[k] = org[k];
    r->sign[k] = dir[k] < 0.0 ? 1 : 0;
    if (fabs(dir[k]) > 2.0e-15)
      r->inv[k] = 1.0 / dir[k];
    else
      r->inv[k] = dir[k] < 0.0 ? -FLT_MAX : FLT_MAX;
  }
}

/* TestRayAABB, render/render_bvh_tree.cc:198-255 */
static int tlas_box(const tlas_ray *r, const double mn[3], const double mx[3], double *tmin_out,
                    double *tmax_out) {
  double lo[3], hi[3];
  for (int k = 0; k < 3; k++) {
    lo[k] = r->sign[k] ? mx[k] : mn[k];
    hi[k] = r->sign[k] ? mn[k] : mx[k];
  }
  double tminx = (lo[0] - r->org[0]) * r->inv[0], tmaxx = (hi[0] - r->org[0]) * r->inv[0];
  double tminy = (lo[1] - r->org[1]) * r->inv[1], tmaxy = (hi[1] - r->org[1]) * r->inv[1];
  double tmin = (tminx > tminy) ? tminx : tminy;
  double tmax = (tmaxx < tmaxy) ? tmaxx : tmaxy;
  double tminz = (lo[2] - r->org[2]) * r->inv[2], tmaxz = (hi[2] - r->org[2]) * r->inv[2];
  tmin = (tmin > tminz) ? tmin : tminz;
  tmax = (tmax < tmaxz) ? tmax : tmaxz;
  if (tmax > 0.0 && tmin <= tmax) {
    *tmin_out = tmin;
    *tmax_out = tmax;
    return 1;
  }
  return 0;
}

typedef struct {
  double tmin;
  uint32_t id;
} tlas_isect;

/* BVHTree::Trace, render/render_bvh_tree.cc:504-586: every leaf whose box the
 * ray enters (maxdist = DBL_MAX), in traversal order, then ordered by tMin.
 * The reference's std::sort is unstable in general; it is an insertion sort
 * (stable) for up to 16 entries, and a stable order is used here. Returns
 * the count (capacity: number of leaves). */
static int64_t tlas_trace(const or_tlas_node *tl, const tlas_ray *r, tlas_isect *out) {
  int64_t cnt = 0;
  int64_t stack[256];
  int sp = 0;
  const or_tlas_node *nd = tl;
  for (;;) {
    if (nd->is_leaf) {
      double t0, t1;
      if (tlas_box(r, nd->bmin, nd->bmax, &t0, &t1) && t0 <= DBL_MAX) {
        out[cnt].tmin = t0;
        out[cnt].id = nd->node_id;
        cnt++;
      }
      if (sp < 1) break;
      nd = tl + stack[--sp];
    } else {
      double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
      int hl = nd->child[0] > 0 && tlas_box(r, tl[nd->child[0]].bmin, tl[nd->child[0]].bmax, &a0, &a1);
      int hr = nd->child[1] > 0 && tlas_box(r, tl[nd->child[1]].bmin, tl[nd->child[1]].bmax, &b0, &b1);
      if (!hl && !hr) {
        if (sp < 1) break;
        nd = tl + stack[--sp];
      } else if (hl && !hr) {
        nd = tl + nd->child[0];
      } else if (!hl && hr) {
        nd = tl + nd->child[1];
      } else {
        int o = r->sign[nd->axis];
        stack[sp++] = nd->child[1 - o];
        nd = tl + nd->child[o];
      }
    }
  }
  for (int64_t i = 1; i < cnt; i++) { /* stable insertion sort by tmin */
    tlas_isect v = out[i];
    int64_t j = i;
    while (j > 0 && v.tmin < out[j - 1].tmin) {
      out[j] = out[j - 1];
      j--;
    }
    out[j] = v;
  }
  return cnt;
}

static int64_t tlas_leaves(const or_tlas_node *tl) {
  /* leaves reachable from the root */
  int64_t stack[256], cnt = 0;
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const or_tlas_node *nd = tl + stack[--sp];
    if (nd->is_leaf) {
      cnt++;
      continue;
    }
    if (nd->child[0] > 0) stack[sp++] = nd->child[0];
    if (nd->child[1] > 0) stack[sp++] = nd->child[1];
  }
  return cnt;
}

/* RenderGraph::Trace, gles/gles_render_graph.cc:145-314, triangle elements */
static void scene_trace_one(const or_tlas_node *tl, const or_tri_element *el, tlas_isect *list,
                            const or_ray *r, or_hit *h, uint32_t *inst) {
  double ro[3] = {r->org[0], r->org[1], r->org[2]};
  double rd[3] = {r->dir[0], r->dir[1], r->dir[2]};
  tlas_ray tr;
  tlas_setup(ro, rd, &tr);
  h->t = FLT_MAX;
  h->u = h->v = 0.0f;
  h->prim_id = 0xffffffffu;
  *inst = 0xffffffffu;
  int64_t cnt = tlas_trace(tl, &tr, list);
  double nearest = DBL_MAX;
  for (int64_t i = 0; i < cnt; i++) {
    if (nearest < list[i].tmin) continue;
    const or_tri_element *e = el + list[i].id;
    double lo[3], ld[3];
    mult_v(e->invT, ro, lo);
    mult_v(e->invT33, rd, ld);
    or_ray lr;
    for (int k = 0; k < 3; k++) {
      lr.org[k] = (float)lo[k];
      lr.dir[k] = (float)ld[k];
    }
    lr.flags = r->flags & 1u;
    lr.prev_prim_id = 0xffffffffu;
    or_hit bh;
    trace_tri_one(e->nodes, e->indices, e->V, e->F, (float)e->eps_scale, &lr, &bh, NULL);
    if (bh.prim_id == 0xffffffffu) continue;
    /* world distance of the local hit point */
    float lp[3], wp[3];
    for (int k = 0; k < 3; k++) lp[k] = (float)(lo[k] + (double)bh.t * ld[k]);
    double lpd[3] = {lp[0], lp[1], lp[2]}, wpd[3];
    mult_v(e->T, lpd, wpd);
    for (int k = 0; k < 3; k++) wp[k] = (float)wpd[k];
    float px = wp[0] - r->org[0], py = wp[1] - r->org[1], pz = wp[2] - r->org[2];
    float tw = sqrtf((px * px + py * py) + pz * pz);
    if (tw < nearest) {
      nearest = tw;
      *h = bh;
      h->t = tw;
      *inst = list[i].id;
    }
  }
}

void or_scene_trace(const or_tlas_node *tlas, const or_tri_element *elems,
                    const or_ray *rays, int64_t nrays, or_hit *hits, uint32_t *inst,
                    int threads) {
  const int64_t nleaves = tlas_leaves(tlas);
  (void)threads;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
#endif
  {
    tlas_isect *list = (tlas_isect *)malloc(sizeof(tlas_isect) * (size_t)(nleaves + 1));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (int64_t i = 0; i < nrays; i++)
      scene_trace_one(tlas, elems, list, rays + i, hits + i, inst + i);
    free(list);
  }
}


/* RenderGraph::Trace with the primitive-kind switch of
 * gles/gles_render_graph.cc:224-264; the loop is scene_trace_one's. */
static void element_trace_one(const or_element *e, const or_ray *lr, or_hit *bh) {
  switch (e->kind) {
    case 0:
      trace_tri_one((const or_tri_node *)e->nodes, e->indices, e->V, e->F, (float)e->eps_scale,
                    lr, bh, NULL);
      break;
    case 1:
      trace_tet_one((const or_tet_node *)e->nodes, e->indices, e->V, e->F, NULL, lr, bh, NULL,
                    NULL);
      break;
    case 2:
      trace_sph_one((const or_sph_node *)e->nodes, e->indices, e->V, e->R, e->rconst, lr, 0.0,
                    NULL, bh, NULL);
      break;
    case 3:
      trace_sol_one((const or_tet_node *)e->nodes, e->indices, e->V, e->F, e->nv, NULL, lr, bh,
                    NULL, NULL);
      break;
    default:
      trace_lin_one((const or_tet_node *)e->nodes, e->indices, e->V, e->F, e->R, e->rconst,
                    e->cap, lr, bh, NULL, NULL);
      break;
  }
}

static void scene_trace_one_mixed(const or_tlas_node *tl, const or_element *el,
                                  tlas_isect *list, const or_ray *r, or_hit *h,
                                  uint32_t *inst) {
  double ro[3] = {r->org[0], r->org[1], r->org[2]};
  double rd[3] = {r->dir[0], r->dir[1], r->dir[2]};
  tlas_ray tr;
  tlas_setup(ro, rd, &tr);
  h->t = FLT_MAX;
  h->u = h->v = 0.0f;
  h->prim_id = 0xffffffffu;
  *inst = 0xffffffffu;
  int64_t cnt = tlas_trace(tl, &tr, list);
  double nearest = DBL_MAX;
  for (int64_t i = 0; i < cnt; i++) {
    if (nearest < list[i].tmin) continue;
    const or_element *e = el + list[i].id;
    double lo[3], ld[3];
    mult_v(e->invT, ro, lo);
    mult_v(e->invT33, rd, ld);
    or_ray lr;
    for (int k = 0; k < 3; k++) {
      lr.org[k] = (float)lo[k];
      lr.dir[k] = (float)ld[k];
    }
    lr.flags = r->flags & 1u;
    lr.prev_prim_id = 0xffffffffu;
    or_hit bh;
    element_trace_one(e, &lr, &bh);
    if (bh.prim_id == 0xffffffffu) continue;
    float lp[3], wp[3];
    for (int k = 0; k < 3; k++) lp[k] = (float)(lo[k] + (double)bh.t * ld[k]);
    double lpd[3] = {lp[0], lp[1], lp[2]}, wpd[3];
    mult_v(e->T, lpd, wpd);
    for (int k = 0; k < 3; k++) wp[k] = (float)wpd[k];
    float px = wp[0] - r->org[0], py = wp[1] - r->org[1], pz = wp[2] - r->org[2];
    float tw = sqrtf((px * px + py * py) + pz * pz);
    if (tw < nearest) {
      nearest = tw;
      *h = bh;
      h->t = tw;
      *inst = list[i].id;
    }
  }
}

void or_scene_trace_mixed(const or_tlas_node *tlas, const or_element *elems,
                          const or_ray *rays, int64_t nrays, or_hit *hits, uint32_t *inst,
                          int threads) {
  const int64_t nleaves = tlas_leaves(tlas);
  (void)threads;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
#endif
  {
    tlas_isect *list = (tlas_isect *)malloc(sizeof(tlas_isect) * (size_t)(nleaves + 1));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (int64_t i = 0; i < nrays; i++)
      scene_trace_one_mixed(tlas, elems, list, rays + i, hits + i, inst + i);
    free(list);
  }
}

/* ------------------------------------------------------ BuildIntersection */

typedef struct {
  float x, y, z;
} v3f;

static v3f v3f_at(const float *p, uint32_t i) {
  v3f r = {p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]};
  return r;
}
static v3f v3f_sub(v3f a, v3f b) {
  v3f r = {a.x - b.x, a.y - b.y, a.z - b.z};
  return r;
}
/* cross (render/render_common.h:161-166) */
static v3f v3f_cross(v3f l, v3f r) {
  v3f o = {l.y * r.z - l.z * r.y, l.z * r.x - l.x * r.z, l.x * r.y - l.y * r.x};
  return o;
}
/* real3::normalize (render/render_common.h:69-82): float sum of squares
 * widened to double, 1.0 / sqrt in double, each component scaled in double */
static v3f v3f_normalize(v3f v) {
  double length2 = (double)(v.x * v.x + v.y * v.y + v.z * v.z);
  double length = 0.0;
  if (fabs(length2) > 1.0e-30) length = 1.0 / sqrt(length2);
  v3f r = {(float)(v.x * length), (float)(v.y * length), (float)(v.z * length)};
  return r;
}
/* ray.origin() + isect.t * ray.direction() */
static v3f hit_pos(const or_ray *r, float t) {
  float tx = r->dir[0] * t, ty = r->dir[1] * t, tz = r->dir[2] * t;
  v3f p = {r->org[0] + tx, r->org[1] + ty, r->org[2] + tz};
  return p;
}
static void isect_put(or_isect *o, v3f pos, v3f geo, v3f nrm, v3f tan, v3f bin, uint32_t f0,
                      uint32_t f1, uint32_t f2, uint32_t f3) {
  memset(o, 0, sizeof(*o));
  o->position[0] = pos.x, o->position[1] = pos.y, o->position[2] = pos.z;
  o->geometric[0] = geo.x, o->geometric[1] = geo.y, o->geometric[2] = geo.z;
  o->normal[0] = nrm.x, o->normal[1] = nrm.y, o->normal[2] = nrm.z;
  o->tangent[0] = tan.x, o->tangent[1] = tan.y, o->tangent[2] = tan.z;
  o->binormal[0] = bin.x, o->binormal[1] = bin.y, o->binormal[2] = bin.z;
  o->f0 = f0, o->f1 = f1, o->f2 = f2, o->f3 = f3;
}

/* render/render_accel_triangle.cc:2439-2507 (float vertices) */
void or_isect_tri(const float *V, const uint32_t *F, const or_ray *rays, const or_hit *hits,
                  int64_t n, or_isect *out) {
  const v3f z = {0, 0, 0};
  for (int64_t i = 0; i < n; i++) {
    if (hits[i].prim_id == 0xFFFFFFFFu) {
      isect_put(&out[i], z, z, z, z, z, 0, 0, 0, 0);
      continue;
    }
    const uint32_t *f = F + 3 * (size_t)hits[i].prim_id;
    v3f p0 = v3f_at(V, f[0]), p1 = v3f_at(V, f[1]), p2 = v3f_at(V, f[2]);
    v3f n3 = v3f_normalize(v3f_cross(v3f_sub(p1, p0), v3f_sub(p2, p0)));
    isect_put(&out[i], hit_pos(&rays[i], hits[i].t), n3, n3, z, z, f[0], f[1], f[2], 0);
  }
}

/* render/render_accel_tetra.cc:2124-2196; face_table at :2178 */
void or_isect_tet(const float *V, const uint32_t *F, const or_ray *rays, const or_hit *hits,
                  const or_tet_ext *ext, int64_t n, or_isect *out) {
  static const int face_table[4][3] = {{0, 1, 3}, {1, 2, 3}, {2, 0, 3}, {0, 1, 2}};
  const v3f z = {0, 0, 0};
  for (int64_t i = 0; i < n; i++) {
    if (hits[i].prim_id == 0xFFFFFFFFu) {
      isect_put(&out[i], z, z, z, z, z, 0, 0, 0, 0);
      continue;
    }
    const uint32_t *f = F + 4 * (size_t)hits[i].prim_id;
    v3f p[4];
    for (int k = 0; k < 4; k++) p[k] = v3f_at(V, f[k]);
    const int *ft = face_table[ext[i].subface & 3u];
    v3f n3 = v3f_normalize(v3f_cross(v3f_sub(p[ft[1]], p[ft[0]]), v3f_sub(p[ft[2]], p[ft[0]])));
    isect_put(&out[i], hit_pos(&rays[i], hits[i].t), n3, n3, z, z, f[0], f[1], f[2], f[3]);
  }
}

/* render/render_accel_particle.cc:2968-3049 with vnormalize (:2041-2056) */
void or_isect_sph(const float *P, const or_ray *rays, const or_hit *hits, int64_t n,
                  or_isect *out) {
  const v3f z = {0, 0, 0};
  for (int64_t i = 0; i < n; i++) {
    const uint32_t id = hits[i].prim_id;
    if (id == 0xFFFFFFFFu) {
      isect_put(&out[i], z, z, z, z, z, 0, 0, 0, 0);
      continue;
    }
    v3f pos = hit_pos(&rays[i], hits[i].t);
    v3f cp = v3f_sub(pos, v3f_at(P, id));
    const float kEPS = (float)(FLT_EPSILON * 16.0);
    v3f nn = cp;
    float len = sqrtf(cp.x * cp.x + cp.y * cp.y + cp.z * cp.z);
    if (len > kEPS) {
      float invlen = (float)(1.0 / len);
      nn.x = cp.x * invlen, nn.y = cp.y * invlen, nn.z = cp.z * invlen;
    }
    v3f tan, bin;
    float xz2 = nn.x * nn.x + nn.z * nn.z;
    if (xz2 > 0.0) {
      float xz = sqrtf(xz2);
      float inv = (float)(1.0 / xz);
      tan.x = (float)(-2.0 * M_PI * nn.x);
      tan.y = (float)(2.0 * M_PI * nn.z);
      tan.z = 0.0f;
      bin.x = (float)(-M_PI * nn.z * inv * nn.y);
      bin.y = (float)(-M_PI * nn.x * inv * nn.y);
      bin.z = (float)(M_PI * xz);
    } else {
      tan.x = 0.0f, tan.y = 0.0f, tan.z = 1.0f;
      bin.x = nn.y > 0.0 ? 1.0f : -1.0f, bin.y = 0.0f, bin.z = 0.0f;
    }
    isect_put(&out[i], pos, nn, nn, tan, bin, id, id, id, 0);
  }
}


/* -------------------------------------------------- SparseVolumeAccel */

/* BVHTree::Locate's first hit (render/render_bvh_tree.cc:588-654 with
 * TestPointAABB / TestPointBVHNode, :173-209): -1 when no leaf box holds p. */
static int sparse_locate(const or_tlas_node *nodes, int64_t nn, const double p[3]) {
  int64_t stack[256];
  int sp = 0;
  int64_t node = 0;
  if (nn <= 0) return -1;
#define INSIDE(nd)                                                                        \
  (p[0] >= (nd)->bmin[0] && p[0] <= (nd)->bmax[0] && p[1] >= (nd)->bmin[1] &&             \
   p[1] <= (nd)->bmax[1] && p[2] >= (nd)->bmin[2] && p[2] <= (nd)->bmax[2])
  for (;;) {
    const or_tlas_node *nd = &nodes[node];
    if (nd->is_leaf) {
      if (INSIDE(nd)) return (int)nd->node_id;
      if (sp < 1) return -1;
      node = stack[--sp];
      continue;
    }
    int l = nd->child[0] > 0 && INSIDE(&nodes[nd->child[0]]);
    int r = nd->child[1] > 0 && INSIDE(&nodes[nd->child[1]]);
    if (l && r) {
      stack[sp++] = nd->child[1];
      node = nd->child[0];
    } else if (l) {
      node = nd->child[0];
    } else if (r) {
      node = nd->child[1];
    } else {
      if (sp < 1) return -1;
      node = stack[--sp];
    }
  }
#undef INSIDE
}

/* SparseVolumeAccel::Build + Sample (render/render_accel_volume.cc:35-139):
 * blocks (offset[3], extent[3], size[3] per block, 9 ints) with voxel data
 * data[b]; uvw != NULL takes Texture::Fetch's position u * globalDim in
 * float (gles/gles_common.h:353-363) instead of pos. values[4 n]. */
void or_sparse_sample(const int32_t *blocks, const void *const *data, int64_t nblocks,
                      int components, int format, const int global_dim[3], const double *pos,
                      const float *uvw, int64_t n, double *values) {
  double *boxes = (double *)malloc(sizeof(double) * 6 * (size_t)nblocks);
  for (int64_t b = 0; b < nblocks; b++)
    for (int k = 0; k < 3; k++) {
      boxes[6 * b + k] = (double)blocks[9 * b + k];
      boxes[6 * b + 3 + k] = (double)(blocks[9 * b + k] + blocks[9 * b + 3 + k]);
    }
  or_tlas_node *nodes = (or_tlas_node *)malloc(sizeof(or_tlas_node) * 2 * (size_t)nblocks);
  const int64_t nn = or_tlas_build(boxes, nblocks, nodes);
  const int comps = components >= 4 ? 4 : components;
  for (int64_t i = 0; i < n; i++) {
    double p[3];
    for (int k = 0; k < 3; k++)
      p[k] = uvw ? (double)(uvw[3 * i + k] * (float)global_dim[k]) : pos[3 * i + k];
    double *v = values + 4 * i;
    v[0] = v[1] = v[2] = 0.0;
    v[3] = 1.0;
    const int b = sparse_locate(nodes, nn, p);
    if (b < 0) {
      v[3] = 0.0;
      continue;
    }
    const int32_t *off = blocks + 9 * b, *ext = off + 3, *size = off + 6;
    double x = p[0] - off[0], y = p[1] - off[1], z = p[2] - off[2];
    double scaleX = size[0] / (double)ext[0];
    x = x * scaleX;
    y = y * scaleX; /* the reference scales every axis by scaleX */
    z = z * scaleX;
    int ix = (int)x, iy = (int)y, iz = (int)z;
    ix = ix < size[0] - 1 ? ix : size[0] - 1;
    ix = ix > 0 ? ix : 0;
    iy = iy < size[1] - 1 ? iy : size[1] - 1;
    iy = iy > 0 ? iy : 0;
    iz = iz < size[2] - 1 ? iz : size[2] - 1;
    iz = iz > 0 ? iz : 0;
    size_t idx = (size_t)(iz * size[0] * size[1] + iy * size[0] + ix);
    for (int c = 0; c < comps; c++) {
      size_t e = (size_t)comps * idx + c;
      if (format == 1)
        v[c] = ((const unsigned char *)data[b])[e] / 255.0f;
      else if (format == 8)
        v[c] = ((const float *)data[b])[e];
      else
        v[c] = (float)((const double *)data[b])[e];
    }
  }
  free(nodes);
  free(boxes);
}
