// ref_driver.cc -- a thin extern "C" shim over the SURFACE reference's own
// render/ sources, compiled together with them (from where they lie in the
// reference checkout) by oracle/Makefile into oracle/_ref/. It exists only
// to pin the oracle: tests and the bench's cpu_baseline leg call it through
// ctypes. It contains no algorithm of its own -- every result comes from the
// reference's functions named below.
//
//   Morton        CalculateMortonCodes{TriangleFloat,TetraFloat}30,
//                 CalculateMortonCodes30{,SIMD}   render/render_prefix_tree_util.cc
//   sort          RadixSort30 / RadixSort30OMP   render/render_prefix_tree_util.h:296-457
//   Karras        ConstructBinaryRadixTree30     render/render_prefix_tree_util.cc:562-649
//   build         {Triangle,Tetra,Particle,Solid}Accel::Build32 + Dump,
//                 LineAccel::Build (the reference's only line builder)
//   trace         {Triangle,Tetra,Particle,Solid,Line}Accel::Traverse, OpenMP
//                 schedule(dynamic,1) over image rows as in
//                 gles/gles_raytrace_engine.cc:662-665
//   texture       FilterTexture3DFloat           render/render_texture.cc:946-1058
//   scene         RenderElement (Update: world box, inverse matrices),
//                 RenderGraph::Build / Trace    gles/gles_render_graph.cc:19-314,
//                 BVHTree::BuildTree            render/render_bvh_tree.cc:319-502
//   eye rays      Camera::buildFrame            render/render_camera.cc:17-63, then
//                 the per-pixel ray expression of RaytraceEngine::OnData and its
//                 ViewportConversion (gles/gles_raytrace_engine.cc:530-544,
//                 730-812), written out here in the reference's real3 types
//                 because that file needs the GL context to compile
#include <cstdint>  // render_prim_line.h uses uint32_t without including it
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <unistd.h>
#include <fcntl.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "render_accel_line.h"
#include "render_accel_particle.h"
#include "render_accel_solid.h"
#include "render_camera.h"
#include "render_accel_tetra.h"
#include "render_accel_triangle.h"
#include "render_accel_volume.h"
#include "render_prefix_tree_util.h"
#include "render_texture.h"
#include "render_bvh_tree.h"
#include "gles_render_graph.h"

using namespace lsgl::render;

namespace {

// Build32 prints its stage timings; keep the caller's stdout clean.
struct QuietStdout {
  int saved;
  QuietStdout() {
    fflush(stdout);
    saved = dup(1);
    int devnull = open("/dev/null", O_WRONLY);
    dup2(devnull, 1);
    close(devnull);
  }
  ~QuietStdout() {
    fflush(stdout);
    dup2(saved, 1);
    close(saved);
  }
};

struct RayIn {
  float org[3];
  unsigned int flags;
  float dir[3];
  unsigned int prev_prim_id;
};
struct HitOut {
  float t, u, v;
  unsigned int prim_id;
};
struct TetExt {
  float d[4];
  float scalar;
  unsigned int subface;
  float t_leave;
  float pad;
};

struct SolExt {
  float d[8];
  float pos[3];
  float scalar;
  float nrm[3];
  unsigned int side;
};

struct LinExt {
  float pos[3];
  unsigned int hit_cap;
  float nrm[3];
  float pad;
};

struct RefAccel {
  int kind;  // 0 tri, 1 tet, 2 sphere, 3 solid, 4 line
  Mesh mesh;
  Tetrahedron tet;
  Particles part;
  Solid sol;
  Lines lin;
  LineAccel lin_accel;
  TriangleAccel tri_accel;
  TetraAccel tet_accel;
  ParticleAccel sph_accel;
  SolidAccel sol_accel;
  ~RefAccel() {
    // Mesh::~Mesh deletes its arrays; they belong to the caller.
    mesh.vertices = 0;
    mesh.faces = 0;
    mesh.dvertices = 0;
    mesh.normals = 0;
  }
};

Ray make_ray(const RayIn &r) {
  real3 o(r.org[0], r.org[1], r.org[2]);
  real3 d(r.dir[0], r.dir[1], r.dir[2]);
  Ray ray(o, d);
  ray.double_sided = (int)(r.flags & 1u);
  ray.prev_prim_id = r.prev_prim_id;
  ray.prev_hit_t = 0.0;
  ray.prev_hit_normal = real3(0.0f, 0.0f, 0.0f);
  return ray;
}

}  // namespace

extern "C" {

int ref_has_openmp(void) {
#ifdef _OPENMP
  return 1;
#else
  return 0;
#endif
}

void ref_morton_tri(const float *V, const unsigned *F, long long n,
                    const float bmin[3], const float bmax[3], unsigned *codes) {
  real3 mn(bmin[0], bmin[1], bmin[2]), mx(bmax[0], bmax[1], bmax[2]);
  CalculateMortonCodesTriangleFloat30(codes, V, F, mn, mx, 0, n);
}

void ref_morton_tet(const float *V, const unsigned *F, long long n,
                    const float bmin[3], const float bmax[3], unsigned *codes) {
  real3 mn(bmin[0], bmin[1], bmin[2]), mx(bmax[0], bmax[1], bmax[2]);
  CalculateMortonCodesTetraFloat30(codes, V, F, mn, mx, 0, n);
}

void ref_morton_pts(const float *P, long long n, const float bmin[3],
                    const float bmax[3], unsigned *codes, int simd) {
  real3 mn(bmin[0], bmin[1], bmin[2]), mx(bmax[0], bmax[1], bmax[2]);
  if (simd)
    CalculateMortonCodes30SIMD(codes, P, mn, mx, 0, n);
  else
    CalculateMortonCodes30(codes, P, mn, mx, 0, n);
}

void ref_sort_keys(const unsigned *codes, long long n, unsigned *out_codes,
                   unsigned *out_index) {
  std::vector<IndexKey30> keys(n);
  for (long long i = 0; i < n; i++) {
    keys[i].index = (uint32_t)i;
    keys[i].code = codes[i];
  }
#ifdef _OPENMP
  RadixSort30OMP(&keys[0], &keys[0] + n);
#else
  RadixSort30(&keys[0], &keys[0] + n);
#endif
  for (long long i = 0; i < n; i++) {
    out_codes[i] = keys[i].code;
    out_index[i] = keys[i].index;
  }
}

void ref_radix_tree30(const unsigned *sorted_codes, long long n,
                      unsigned *out /* 3 words per node */) {
  std::vector<IndexKey30> keys(n);
  for (long long i = 0; i < n; i++) {
    keys[i].index = (uint32_t)i;
    keys[i].code = sorted_codes[i];
  }
  for (long long i = 0; i + 1 < n; i++) {
    NodeInfo32 ni = ConstructBinaryRadixTree30(&keys[0], (int)i, (uint32_t)n);
    out[3 * i + 0] = ni.childIndex;
    out[3 * i + 1] = (unsigned)ni.leftType;
    out[3 * i + 2] = (unsigned)ni.rightType;
  }
}

void *ref_build_tri(const float *V, long long nv, unsigned *F, long long nf) {
  RefAccel *a = new RefAccel();
  a->kind = 0;
  a->mesh.nvertices = nv;
  a->mesh.vertices = V;
  a->mesh.nfaces = nf;
  a->mesh.faces = F;
  QuietStdout q;
  a->tri_accel.Build32(&a->mesh, TriangleBuildOptions());
  return a;
}

void *ref_build_tet(const float *V, unsigned *F, long long n) {
  RefAccel *a = new RefAccel();
  a->kind = 1;
  memset(&a->tet, 0, sizeof(a->tet));
  a->tet.numTetrahedrons = n;
  a->tet.vertices = V;
  a->tet.faces = F;
  QuietStdout q;
  a->tet_accel.Build32(&a->tet, TetraBuildOptions());
  return a;
}

void *ref_build_sph(const float *P, float *R, float rconst, long long n) {
  RefAccel *a = new RefAccel();
  a->kind = 2;
  memset(&a->part, 0, sizeof(a->part));
  a->part.numParticles = n;
  a->part.positions = P;
  a->part.radius = R;
  a->part.constantRadius = rconst;
  QuietStdout q;
  a->sph_accel.Build32(&a->part, ParticleBuildOptions());
  return a;
}

// SolidAccel::Build32 (render/render_accel_solid.cc:1795-2124); nv = 5
// pyramids, 6 prisms, 8 hexahedra.
void *ref_build_sol(const float *V, unsigned *F, long long n, int nv) {
  RefAccel *a = new RefAccel();
  a->kind = 3;
  memset(&a->sol, 0, sizeof(a->sol));
  a->sol.numVertsPerSolid = nv;
  a->sol.numSolids = n;
  a->sol.vertices = V;
  a->sol.isDoublePrecisionPos = false;
  a->sol.indices = F;
  QuietStdout q;
  a->sol_accel.Build32(&a->sol, SolidBuildOptions());
  return a;
}

// LineAccel::Build (render/render_accel_line.cc:298-341) with
// LineBuildOptions (minLeafPrimitives 4, maxTreeDepth 512, cap).
void *ref_build_lin(const float *P, unsigned *segs, const float *R, float rconst, long long n,
                    int cap) {
  RefAccel *a = new RefAccel();
  a->kind = 4;
  memset(&a->lin, 0, sizeof(a->lin));
  a->lin.numLines = n;
  a->lin.positions = P;
  a->lin.isDoublePrecisionPos = false;
  a->lin.radius = R;
  a->lin.constantRadius = rconst;
  a->lin.segments = segs;
  LineBuildOptions opt;
  opt.cap = cap != 0;
  QuietStdout q;
  a->lin_accel.Build(&a->lin, opt);
  return a;
}

int ref_dump(void *h, const char *path) {
  RefAccel *a = (RefAccel *)h;
  if (a->kind == 0) return a->tri_accel.Dump(path) ? 0 : 1;
  if (a->kind == 1) return a->tet_accel.Dump(path) ? 0 : 1;
  if (a->kind == 3) return a->sol_accel.Dump(path) ? 0 : 1;
  if (a->kind == 4) return a->lin_accel.Dump(path) ? 0 : 1;
  return a->sph_accel.Dump(path) ? 0 : 1;
}

void ref_bounds(void *h, double bmin[3], double bmax[3]) {
  RefAccel *a = (RefAccel *)h;
  if (a->kind == 0)
    a->tri_accel.BoundingBox(bmin, bmax);
  else if (a->kind == 1)
    a->tet_accel.BoundingBox(bmin, bmax);
  else if (a->kind == 3)
    a->sol_accel.BoundingBox(bmin, bmax);
  else if (a->kind == 4)
    a->lin_accel.BoundingBox(bmin, bmax);
  else
    a->sph_accel.BoundingBox(bmin, bmax);
}

void ref_free(void *h) { delete (RefAccel *)h; }

// Trace n rays as rows of `width` rays, OpenMP dynamic over rows. For tetra,
// ext receives InterpolateTetra weights and, given scalars, the varying
// f0*d0 + f1*d1 + f2*d2 + f3*d3 (gles/gles_common.cc:1074-1076); for solids
// (SolExt) the interpolate() weights of the solid's vertices (the rest 0),
// the position and normal, and the varying sum over f0..f(nv-1)
// (gles/gles_common.cc:1078-1140). A solid's Intersection::u/v are never
// written by the reference (uninitialised); hits[i].u/v carry them as they
// are.
void ref_trace(void *h, const void *rays_in, long long n, long long width,
               void *hits_out, void *ext_out, const float *scalars,
               int threads) {
  RefAccel *a = (RefAccel *)h;
  const RayIn *rays = (const RayIn *)rays_in;
  HitOut *hits = (HitOut *)hits_out;
  TetExt *ext = (TetExt *)ext_out;
  if (width <= 0) width = n;
  long long rows = (n + width - 1) / width;
  (void)threads;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (long long y = 0; y < rows; y++) {
    long long end = (y + 1) * width < n ? (y + 1) * width : n;
    for (long long i = y * width; i < end; i++) {
      Ray ray = make_ray(rays[i]);
      Intersection isect;
      isect.subface_id = 0;
      isect.d0 = isect.d1 = isect.d2 = isect.d3 = 0.0f;
      bool hit = false;
      if (a->kind == 0)
        hit = a->tri_accel.Traverse(isect, ray);
      else if (a->kind == 1)
        hit = a->tet_accel.Traverse(isect, ray);
      else if (a->kind == 3)
        hit = a->sol_accel.Traverse(isect, ray);
      else if (a->kind == 4)
        hit = a->lin_accel.Traverse(isect, ray);
      else
        hit = a->sph_accel.Traverse(isect, ray);
      if (!hit && a->kind == 4) {
        // LineAccel::Traverse leaves the caller's Intersection untouched on
        // a miss (render/render_accel_line.cc:766-773): report the miss
        isect.t = 3.402823466e+38f;
        isect.u = isect.v = 0.0f;
        isect.prim_id = 0xffffffffu;
      }
      hits[i].t = isect.t;
      hits[i].u = isect.u;
      hits[i].v = isect.v;
      hits[i].prim_id = isect.prim_id;
      if (ext && a->kind == 4) {
        // position and normal of BuildIntersection (:654-702); whether a
        // cap was hit is not part of the reference's Intersection
        LinExt *le = (LinExt *)ext_out + i;
        memset(le, 0, sizeof(LinExt));
        if (hit)
          for (int k = 0; k < 3; k++) {
            le->pos[k] = isect.position[k];
            le->nrm[k] = isect.normal[k];
          }
      } else if (ext && a->kind == 3) {
        SolExt *se = (SolExt *)ext_out + i;
        memset(se, 0, sizeof(SolExt));
        if (hit) {
          const int nv = a->sol.numVertsPerSolid;
          const float dd[8] = {isect.d0, isect.d1, isect.d2, isect.d3,
                               isect.d4, isect.d5, isect.d6, isect.d7};
          const unsigned ff[8] = {isect.f0, isect.f1, isect.f2, isect.f3,
                                  isect.f4, isect.f5, isect.f6, isect.f7};
          for (int k = 0; k < nv; k++) se->d[k] = dd[k];
          for (int k = 0; k < 3; k++) {
            se->pos[k] = isect.position[k];
            se->nrm[k] = isect.normal[k];
          }
          if (scalars) {
            float acc = scalars[ff[0]] * dd[0];
            for (int k = 1; k < nv; k++) acc = acc + scalars[ff[k]] * dd[k];
            se->scalar = acc;
          }
        }
      } else if (ext) {
        memset(&ext[i], 0, sizeof(TetExt));
        if (hit && a->kind == 1) {
          ext[i].d[0] = isect.d0;
          ext[i].d[1] = isect.d1;
          ext[i].d[2] = isect.d2;
          ext[i].d[3] = isect.d3;
          ext[i].subface = isect.subface_id;
          if (scalars)
            ext[i].scalar = scalars[isect.f0] * isect.d0 +
                            scalars[isect.f1] * isect.d1 +
                            scalars[isect.f2] * isect.d2 +
                            scalars[isect.f3] * isect.d3;
        }
      }
    }
  }
}

// The whole Intersection that {Triangle,Tetra,Particle}Accel::Traverse fills
// through its BuildIntersection (render/render_accel_triangle.cc:2439-2507,
// render/render_accel_tetra.cc:2124-2196,
// render/render_accel_particle.cc:2968-3049): position, geometric, normal,
// tangent, binormal, f0..f3, 80 bytes per ray, from a zeroed Intersection
// (the fields a kind leaves unset stay 0; a miss is all 0).
struct IsectOut {
  float position[3];
  unsigned int f0;
  float geometric[3];
  unsigned int f1;
  float normal[3];
  unsigned int f2;
  float tangent[3];
  unsigned int f3;
  float binormal[3];
  float pad;
};

void ref_trace_isect(void *h, const void *rays_in, long long n, void *out_v) {
  RefAccel *a = (RefAccel *)h;
  const RayIn *rays = (const RayIn *)rays_in;
  IsectOut *out = (IsectOut *)out_v;
  for (long long i = 0; i < n; i++) {
    Ray ray = make_ray(rays[i]);
    Intersection isect;
    memset(&isect, 0, sizeof(isect));
    bool hit = false;
    if (a->kind == 0)
      hit = a->tri_accel.Traverse(isect, ray);
    else if (a->kind == 1)
      hit = a->tet_accel.Traverse(isect, ray);
    else if (a->kind == 2)
      hit = a->sph_accel.Traverse(isect, ray);
    IsectOut &o = out[i];
    memset(&o, 0, sizeof(o));
    if (!hit) continue;
    for (int k = 0; k < 3; k++) {
      o.position[k] = isect.position[k];
      o.geometric[k] = isect.geometric[k];
      o.normal[k] = isect.normal[k];
      o.tangent[k] = isect.tangent[k];
      o.binormal[k] = isect.binormal[k];
    }
    o.f0 = isect.f0;
    o.f1 = isect.f1;
    o.f2 = isect.f2;
    o.f3 = isect.f3;
  }
}

void ref_tex3d_sample(float *vox, const int dims[3], const int clamp[3],
                      const float *uvw, long long n, float *out) {
  Texture3D tex;
  memset(&tex, 0, sizeof(tex));
  tex.image = (unsigned char *)vox;
  tex.width = dims[0];
  tex.height = dims[1];
  tex.depth = dims[2];
  tex.components = 1;
  tex.data_type = LSGL_RENDER_TEXTURE3D_FORMAT_FLOAT;
  for (long long i = 0; i < n; i++) {
    float rgba[4];
    FilterTexture3DFloat(rgba, &tex, uvw[3 * i], uvw[3 * i + 1], uvw[3 * i + 2],
                         clamp[0] != 0, clamp[1] != 0, clamp[2] != 0);
    out[i] = rgba[0];
  }
}

// FilterTexture3DByte / Float / Double (render/render_texture.cc:832-1172)
// on a texture of `comps` interleaved components; data_type is the
// reference's LSGL_RENDER_TEXTURE3D_FORMAT_* (1 byte, 8 float, 16 double).
void ref_tex3d_fetch(void *image, int data_type, int comps, const int dims[3],
                     const int clamp[3], const float *uvw, long long n, float *rgba) {
  Texture3D tex;
  memset(&tex, 0, sizeof(tex));
  tex.image = (unsigned char *)image;
  tex.width = dims[0];
  tex.height = dims[1];
  tex.depth = dims[2];
  tex.components = comps;
  tex.data_type = data_type;
  for (long long i = 0; i < n; i++) {
    float *o = rgba + 4 * i;
    const bool cs = clamp[0] != 0, ct = clamp[1] != 0, cr = clamp[2] != 0;
    if (data_type == LSGL_RENDER_TEXTURE3D_FORMAT_BYTE)
      FilterTexture3DByte(o, &tex, uvw[3 * i], uvw[3 * i + 1], uvw[3 * i + 2], cs, ct, cr);
    else if (data_type == LSGL_RENDER_TEXTURE3D_FORMAT_DOUBLE)
      FilterTexture3DDouble(o, &tex, uvw[3 * i], uvw[3 * i + 1], uvw[3 * i + 2], cs, ct, cr);
    else
      FilterTexture3DFloat(o, &tex, uvw[3 * i], uvw[3 * i + 1], uvw[3 * i + 2], cs, ct, cr);
  }
}

// SparseVolumeAccel::Build + Sample (render/render_accel_volume.cc:35-139)
// over blocks given as 9 ints each (offset, extent, size) with their data.
void ref_sparse_sample(const int *blocks, const void *const *data, long long nblocks,
                       int components, int format, const double *pos, long long n,
                       double *values) {
  SparseVolume sv;
  sv.components = components;
  sv.format = format;
  sv.maxLevel = 0;
  for (long long b = 0; b < nblocks; b++) {
    VolumeBlock vb;
    memset(&vb, 0, sizeof(vb));
    for (int k = 0; k < 3; k++) {
      vb.offset[k] = blocks[9 * b + k];
      vb.extent[k] = blocks[9 * b + 3 + k];
      vb.size[k] = blocks[9 * b + 6 + k];
    }
    vb.id = (int)b;
    vb.data = (const unsigned char *)data[b];
    sv.blocks.push_back(vb);
  }
  SparseVolumeAccel accel;
  accel.Build(&sv);
  for (long long i = 0; i < n; i++) accel.Sample(values + 4 * i, pos + 3 * i);
}

// ---- two-level scenes (RenderGraph over accelerators of any kind) ----
struct RefScene {
  lsgl::RenderGraph graph;
};

// The RenderElement primitive type and accelerator of a built RefAccel, as
// glDrawElements records them (gles/gles_primitive.cc:137-293).
static void element_of(RefAccel *a, lsgl::AccelBuilder::PrimitiveType *ty, GLenum *mode,
                       unsigned char **accel, double bmin[3], double bmax[3]) {
  typedef lsgl::AccelBuilder AB;
  if (a->kind == 1) {
    *ty = AB::PRIMITIVE_TETRAHEDRONS, *mode = GL_TETRAHEDRONS_EXT;
    *accel = reinterpret_cast<unsigned char *>(&a->tet_accel);
  } else if (a->kind == 2) {
    *ty = AB::PRIMITIVE_POINTS, *mode = GL_POINTS;
    *accel = reinterpret_cast<unsigned char *>(&a->sph_accel);
  } else if (a->kind == 3) {
    const int nv = a->sol.numVertsPerSolid;
    *ty = nv == 5 ? AB::PRIMITIVE_PYRAMIDS : (nv == 6 ? AB::PRIMITIVE_PRISMS
                                                       : AB::PRIMITIVE_HEXAHEDRONS);
    *mode = nv == 5 ? GL_PYRAMIDS_EXT : (nv == 6 ? GL_PRISMS_EXT : GL_HEXAHEDRONS_EXT);
    *accel = reinterpret_cast<unsigned char *>(&a->sol_accel);
  } else if (a->kind == 4) {
    *ty = AB::PRIMITIVE_LINES, *mode = GL_LINES;
    *accel = reinterpret_cast<unsigned char *>(&a->lin_accel);
  } else {
    *ty = AB::PRIMITIVE_TRIANGLES, *mode = GL_TRIANGLES;
    *accel = reinterpret_cast<unsigned char *>(&a->tri_accel);
  }
  ref_bounds(a, bmin, bmax);
}

// Builds a RenderGraph of n elements (accels from ref_build_*, any kind;
// T[i] = 16 doubles, row-major T_[4][4]); writes each element's inverse
// matrices and world box as RenderElement::Update computes them, and the
// top-level BVHTree over those boxes (nodes: 6 doubles box + child[2] +
// axis + isLeaf + nodeID each, see BVHNode).
void *ref_scene_build(int n, void **accels, const double *T, double *invT_out,
                      double *invT33_out, double *xbox_out, double *tlas_box,
                      int *tlas_meta, int *tlas_count, int tlas_cap) {
  RefScene *s = new RefScene;
  BVHTree tree;
  for (int i = 0; i < n; i++) {
    RefAccel *a = (RefAccel *)accels[i];
    double bmin[3], bmax[3], M[4][4];
    lsgl::AccelBuilder::PrimitiveType ty;
    GLenum mode;
    unsigned char *acc;
    element_of(a, &ty, &mode, &acc, bmin, bmax);
    memcpy(M, T + 16 * i, sizeof(M));
    lsgl::RenderElement e(ty, mode, acc, bmin, bmax, M, 0, NULL, FragmentState(),
                          lsgl::ShadingState());
    s->graph.AddRenderElement(e);
    memcpy(invT_out + 16 * i, e.invT_, sizeof(M));
    memcpy(invT33_out + 16 * i, e.invT33_, sizeof(M));
    memcpy(xbox_out + 6 * i, e.xbmin_, 3 * sizeof(double));
    memcpy(xbox_out + 6 * i + 3, e.xbmax_, 3 * sizeof(double));
    BVHData d;
    memcpy(d.bmin, e.xbmin_, sizeof(d.bmin));
    memcpy(d.bmax, e.xbmax_, sizeof(d.bmax));
    d.nodeID = (unsigned)i;
    tree.AddNode(d);
  }
  {
    QuietStdout q;
    s->graph.Build();
    tree.BuildTree();
  }
  const std::vector<BVHNode> &tn = tree.GetNodesTree();
  *tlas_count = (int)tn.size();
  for (size_t k = 0; k < tn.size() && (int)k < tlas_cap; k++) {
    memcpy(tlas_box + 6 * k, tn[k].bmin, 3 * sizeof(double));
    memcpy(tlas_box + 6 * k + 3, tn[k].bmax, 3 * sizeof(double));
    tlas_meta[5 * k + 0] = tn[k].child[0];
    tlas_meta[5 * k + 1] = tn[k].child[1];
    tlas_meta[5 * k + 2] = tn[k].axis;
    tlas_meta[5 * k + 3] = tn[k].isLeaf ? 1 : 0;
    tlas_meta[5 * k + 4] = (int)tn[k].nodeID;
  }
  return s;
}

void ref_scene_free(void *h) { delete (RefScene *)h; }

// RenderGraph::Trace per ray (OpenMP dynamic over rows of `width`): hit t is
// the world distance, u, v, prim_id the element's; inst = element index.
void ref_scene_trace(void *h, const void *rays_in, long long n, long long width,
                     void *hits_out, unsigned *inst_out, int threads) {
  RefScene *s = (RefScene *)h;
  const RayIn *rays = (const RayIn *)rays_in;
  HitOut *hits = (HitOut *)hits_out;
  if (width <= 0) width = n;
  long long rows = (n + width - 1) / width;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (long long y = 0; y < rows; y++) {
    for (long long x = 0; x < width; x++) {
      long long i = y * width + x;
      if (i >= n) break;
      Ray ray = make_ray(rays[i]);
      ray.prev_prim_id = (unsigned int)(-1);
      ray.prev_node = NULL;
      Intersection isect;
      isect.t = 0.0f;
      isect.u = isect.v = 0.0f;
      isect.prim_id = (unsigned int)(-1);
      isect.renderElement = NULL;
      bool hit = s->graph.Trace(isect, ray);
      if (hit) {
        hits[i].t = isect.t;
        hits[i].u = isect.u;
        hits[i].v = isect.v;
        hits[i].prim_id = isect.prim_id;
        inst_out[i] = (unsigned)((const lsgl::RenderElement *)isect.renderElement)->nodeID_;
      } else {
        hits[i].t = 3.402823466e+38f;
        hits[i].u = hits[i].v = 0.0f;
        hits[i].prim_id = (unsigned int)(-1);
        inst_out[i] = (unsigned int)(-1);
      }
    }
  }
}

void ref_eye_rays(const float *eye, const float *lookat, const float *up, float fov, int width,
                  int height, float *org_out, float *dir_out, unsigned char *valid) {
  using namespace lsgl::render;
  Camera cam(real3(eye[0], eye[1], eye[2]), real3(lookat[0], lookat[1], lookat[2]),
             real3(up[0], up[1], up[2]), fov);
  real3 corner, udir, vdir;
  cam.buildFrame(corner, udir, vdir, width, height);
  const float invWidth = 1.0f / width;
  const float invHeight = 1.0f / height;
  const int step = 1;
  for (int y = 0; y < height; y++)
    for (int x = 0; x < width; x++) {
      const size_t i = (size_t)y * width + x;
      float x01 = (x + 0.5f) * invWidth;
      float y01 = (y + 0.5f) * invHeight;
      float xnd = 2.0f * (x01 - 0.5f);
      float ynd = 2.0f * (y01 - 0.5f);
      int px = (xnd + 1.0f) * (0.5f * width) - 0;
      int py = (ynd + 1.0f) * (0.5f * height) + 0;
      valid[i] = !((px < 0) || (py < 0) || (px >= width) || (py >= height));
      float u = (0 + 0.0f) / (float)1, v = (0 + 0.0f) / (float)1;
      real3 org = cam.getEye();
      real3 dir = (corner + ((float)px + u + step / 2.0f) * udir +
                   ((float)py + v - step / 2.0f + 1.0f) * vdir) - org;
      dir.normalize();
      for (int k = 0; k < 3; k++) {
        org_out[3 * i + k] = valid[i] ? org[k] : 0.0f;
        dir_out[3 * i + k] = valid[i] ? dir[k] : 0.0f;
      }
    }
}

}  // extern "C"
