// Plain data shared by the host scene compiler and the HIP kernels.
//
// Everything here is POD so the same bytes are uploaded to HBM unchanged.
// Geometry is kept in FP64 (the reference's RREAL, common/fvect.h:16) for the
// primitive tests; the BVH bounds are FP32 and widened outward so the FP32 box
// test never rejects a primitive the FP64 test would accept.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define RG_HD __host__ __device__ __forceinline__
#else
#define RG_HD inline
#endif

namespace rg {

// Scene tables are reached through pointers held in structs (one DevScene per
// scene of a batch, read from device memory).  Device code types those
// pointers as global (address space 1) so every load through them is a
// global load rather than a flat one (flat loads are slower and also count
// against the LDS counter).  Host code sees plain pointers; the layout is the
// same 64-bit address either way.
#if defined(__HIP_DEVICE_COMPILE__)
#define RG_GLOBAL __attribute__((address_space(1)))
#else
#define RG_GLOBAL
#endif
template <class T>
inline const RG_GLOBAL T* to_global(const T* p) {
  return (const RG_GLOBAL T*)p;
}

constexpr double FTINY = 1e-6;   // common/fvect.h:17
constexpr double FHUGE = 1e10;   // common/fvect.h:20
constexpr double PI_D = 3.14159265358979323846;

// ray type flags (rt/ray.h:27-44)
enum : uint16_t {
  PRIMARY = 01, RSHADOW = 02, REFLECTED = 04, REFRACTED = 010, TRANS = 020,
  RAMBIENT = 040, RSPECULAR = 0100, TSHADOW = 0200, TAMBIENT = 0400,
  TSPECULAR = 01000,
  SHADOW = RSHADOW | TSHADOW,
  AMBIENT = RAMBIENT | TAMBIENT,
  SPECULAR = RSPECULAR | TSPECULAR,
  RAYREFL = RSHADOW | REFLECTED | RAMBIENT | RSPECULAR,
};

// primitive kinds (PK_TRIANGLE: one triangle of a compiled mesh, rt/o_mesh.c)
enum PrimKind : int32_t {
  PK_POLYGON = 0, PK_SPHERE = 1, PK_BUBBLE = 2, PK_RING = 3,
  PK_CONE = 4, PK_CUP = 5, PK_CYLINDER = 6, PK_TUBE = 7, PK_TRIANGLE = 8,
};

// Mesh triangle geometry block (world coordinates, FP64):
//   v0[3] v1[3] v2[3]  vertices in stored order
//   n[3]               unit normal as the reference reports it (the mesh-space
//                      (v0-v2)x(v1-v0) carried through the transform)
//   flags              TRI_REV0..2: edge (v0,v1) / (v1,v2) / (v2,v0) runs
//                      against vertex-id order; TRI_NORMALS: vn follow;
//                      TRI_PHONG: the modifier is named "Phong"
//   vn0[3] vn1[3] vn2[3]  transformed vertex normals (unnormalised)
enum : int32_t { TRI_REV0 = 1, TRI_REV1 = 2, TRI_REV2 = 4, TRI_NORMALS = 8, TRI_PHONG = 16 };

// material kinds (only the set the GPU path shades)
enum MatKind : int32_t {
  MK_PLASTIC = 0, MK_METAL = 1, MK_TRANS = 2, MK_GLASS = 3,
  MK_LIGHT = 4, MK_GLOW = 5, MK_ILLUM = 6, MK_NONE = 7,
};

struct Material {
  int32_t kind;      // MatKind
  int32_t nfargs;
  double f[8];       // real arguments (plastic 5, trans 7, glass 3/4, light 3, glow 4)
  int32_t illum_alt; // illum: alternate material index (-1 = void / transparent)
  int32_t transp;    // istransp() of rt/otypes.h (trans/glass/illum-void)
};

// Source flags (rt/source.h)
enum : int32_t {
  SDISTANT = 01, SSKIP = 02, SPROX = 04, SSPOT = 010, SVIRTUAL = 020,
  SFLAT = 040, SCYL = 0100, SFOLLOW = 0200, SCIR = 01000,
};

struct Source {
  double sloc[3];    // direction (distant) or centre
  double snorm[3];
  double ss[3][3];   // sampling vectors SU, SV, SW
  double ss2;        // solid angle (distant) or area
  double srad;
  double prox;       // glow proximity radius
  int32_t flags;
  int32_t prim;      // primitive index for local sources, -1 for distant
  int32_t mat;       // material index of the emitter
  int32_t modslot;   // tracked-modifier slot of the source object (-1 none)
  int32_t objmod;    // modifier object number (tie rule / identity)
  int32_t pad;
};

// Shadow cache of one light source (rt/srcobstr.c): for each cell, the
// opaque object last found to block the source (SHADCACHE of rt/source.h:19-20).
// Distant sources: 2S x 2S cells over the cube face the source shines on,
// indexed by where a ray origin projects along the source direction; flat
// sources: 3 S^2 cells over directions of approach (front and four half
// sides); others: 6 S^2 cells of a direction cube.
constexpr int kShadCache = 20;
enum : int32_t { SC_NONE = 0, SC_DISTANT = 1, SC_FLAT = 2, SC_SPHERE = 3 };
struct SrcCache {
  int32_t kind;      // SC_*
  int32_t ax;        // distant: dominant axis of the source direction
  int32_t cell0;     // first cell of this source in the blocker array
  int32_t ncell;
  double o[3];       // distant: origin of the cell grid on the cube face
  double e1, e2;     // distant: inverse extents of the grid
  double u[3], v[3]; // flat: in-plane axes
};

// Prim.flat bits: PF_FLAT isflat() of the object reported as hit;
// PF_COMPLEX part of a mesh or an instance, whose hits the reference never
// records as shadow blockers (rt/srcobstr.c:298-305).
enum : int32_t { PF_FLAT = 1, PF_COMPLEX = 2 };

// Per-primitive record (all kinds); geometry lives in a shared FP64 pool.
struct Prim {
  int32_t kind;      // PrimKind
  int32_t mat;       // resolved material index (-1: none)
  int32_t omod;      // modifier object number, the "later modifier wins" key
  int32_t modslot;   // tracked-modifier slot (rcontrib), -1 if untracked
  int32_t goff;      // offset into geometry pool (doubles)
  int32_t nv;        // polygon vertex count
  int32_t flat;      // PF_FLAT | PF_COMPLEX bits
  int32_t srcidx;    // source index if this primitive is a local source, else -1
};

// Bin functions evaluated per binned ray: util/reinhartb.cal, util/klems_*.cal,
// util/disk2square.cal (Shirley-Chiu scbin, SCdim held in `mf`) and
// util/cieskyscan.cal (cbin).  BK_HALFSPACE is rfluxmtx's uniform receiver
// surface, bin 0 for rays that meet it from the front:
// if(-Dx*nx-Dy*ny-Dz*nz,0,-1), util/rfluxmtx.c:465-470.
enum BinKind : int32_t { BK_CONST = 0, BK_REINHART = 1, BK_KLEMS_FULL = 2,
                         BK_KLEMS_HALF = 3, BK_KLEMS_QUARTER = 4, BK_HALFSPACE = 5,
                         BK_SHIRLEY_CHIU = 6, BK_CIE = 7 };

struct BinSpec {
  int32_t kind;      // BinKind
  int32_t nbins;
  int32_t col0;      // first output column (in units of bins)
  int32_t cbin;      // constant bin for BK_CONST
  double mf;         // Reinhart subdivision MF, or Shirley-Chiu SCdim
  double rhs;        // +1 right-handed, -1 left-handed
  double nrm[3];     // rNx,rNy,rNz (Reinhart) or Nx,Ny,Nz (Klems)
  double up[3];      // Ux,Uy,Uz
};

// Ray-tracing parameters (subset of RAYPARAMS, rt/ray.h:157-190)
struct Params {
  int32_t ambounce;     // -ab
  int32_t ambdiv;       // -ad
  int32_t maxdepth;     // -lr
  int32_t backvis;      // -bv
  double minweight;     // -lw
  double dstrsrc;       // -dj
  double srcsizerat;    // -ds
  double specthresh;    // -st
  double specjitter;    // -ss
  double ambval[3];     // -av
  int32_t directvis;    // -dv
  int32_t contrib;      // rcontrib -V
  uint64_t seed;
};

}  // namespace rg
