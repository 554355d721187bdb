// Shading, source sampling and bin functions for the wavefront renderer.
//
// The renderer runs Radiance's ray tree breadth-first.  Every ray carries its
// cumulative coefficient (the product of rcoef up to the primary ray that
// raycontrib() of rt/raytrace.c:408-442 would form), so emission found at a
// leaf is added straight into the sensor's output: rtrace's returned value
// and rcontrib's bins are both sums over leaves of coefficient x emission.
#pragma once
#include "rg_geom.hpp"

namespace rg {

// A spawn record: one parent event that produces `count` child rays.
struct alignas(16) Spawn {
  double org[3];     // child origin (parent hit point)
  float d0[3];       // explicit direction, or hemisphere normal
  float d1[3];       // hemisphere tangent ux
  float coef[3];     // cumulative coefficient of each child (before roulette)
  float cw;          // child weight before roulette
  float crc;         // largest channel of the child's own rcoef
  uint32_t sensor;
  uint32_t count;    // number of children
  uint16_t ns;       // hemisphere divisions per axis (0: explicit ray)
  uint16_t crtype;   // child's cumulative ray type
  uint16_t rtype;    // child's ray type
  int16_t rsrc;      // child's aimed source
  uint8_t rlvl;      // child's reflection level
  uint8_t adepth;    // child's ambient recursion depth
  uint16_t flags : 3;    // SPF_*
  uint16_t scene : 13;   // scene of a batched launch (0 otherwise; at most kMaxBatchScenes)
  uint64_t key;      // explicit: child key; hemisphere: parent key
};
static_assert(sizeof(Spawn) == 96, "Spawn layout");
constexpr int kMaxBatchScenes = 4096;   // scenes of one batched launch

enum : uint8_t { SPF_RR = 1, SPF_FORCE = 2, SPF_BACK = 4 };

// Adjoint bookkeeping (path replay): the differentiable vertices a ray's
// coefficient was formed at, as a list of words.  An entry is a header word
// followed by 0, 1 or 3 payload words (float bits), by kind:
//   header bits 0-11  material index      bits 12-15 kind (PV_*)
//          bit 16     Russian roulette after this vertex rescaled by 1/weight
//          bits 17-31 Fresnel estimate fest of a pure-specular vertex, x 32767
//   PV_GSPEC, PV_GTRANS: 1 payload, the score d log p / d roughness summed
//     over the lobe trials up to the accepted one;
//   PV_DIR_R, PV_DIR_T: 3 payloads, the direct-light factors of dirnorm()
//     with the material's colour terms taken out: diffuse a, lobe b, and
//     d log(lobe) / d roughness.
// hdr: bits 0-4 word count, 5-9 first word since the last roulette reset,
// bit 10 overflow (more than PATH_WORDS words).
constexpr int PATH_WORDS = 15;
constexpr int ADJ_COLS = 7;      // gradient columns per material
enum : uint32_t {
  PV_AMB = 1,            // front ambient: rho * rdiff
  PV_AMB_RBLT = 2,       // front ambient with the specular part folded in: rho * rdiff + scolor
  PV_SPEC = 3,           // mirror reflection: scolor
  PV_GSPEC = 4,          // Gaussian lobe reflection sample: scolor
  PV_TSPEC = 5,          // straight transmission: rho * tspec
  PV_GTRANS = 6,         // Gaussian lobe transmission sample: rho * tspec
  PV_AMB_BACK = 7,       // back ambient: rho * tdiff
  PV_AMB_BACK_TBLT = 8,  // back ambient with the transmitted lobe folded in: rho * trans
  PV_DIR_R = 9,          // shadow ray on the reflection side: rho * rdiff * a + scolor * b
  PV_DIR_T = 10          // shadow ray on the transmission side: rho * tdiff * a + rho * tspec * b
};
constexpr uint32_t PV_RR = 1u << 16;
RG_HD int path_payload(uint32_t kind) {
  return (kind == PV_GSPEC || kind == PV_GTRANS) ? 1 : ((kind == PV_DIR_R || kind == PV_DIR_T) ? 3 : 0);
}
struct alignas(16) PathRec {
  uint32_t e[PATH_WORDS];
  uint32_t hdr;
};
static_assert(sizeof(PathRec) == 64, "PathRec layout");

// An entry to append: header word and its payload.
struct PathAdd {
  uint32_t w[4];
  int n;               // 0: nothing to append
};
RG_HD PathAdd path_entry(int mat, uint32_t kind, double fest) {
  PathAdd a;
  uint32_t q = (uint32_t)(fest <= 0 ? 0.0 : (fest >= 1 ? 32767.0 : fest * 32767.0 + 0.5));
  a.w[0] = (uint32_t)(mat & 0xfff) | (kind << 12) | (q << 17);
  a.w[1] = a.w[2] = a.w[3] = 0u;
  a.n = 1 + path_payload(kind);
  return a;
}
RG_HD PathAdd path_none() {
  PathAdd a;
  a.w[0] = a.w[1] = a.w[2] = a.w[3] = 0u;
  a.n = 0;
  return a;
}
RG_HD PathRec path_push(PathRec p, const PathAdd& a) {
  if (!a.n) return p;
  uint32_t n = p.hdr & 31u;
  if (n + (uint32_t)a.n > (uint32_t)PATH_WORDS) { p.hdr |= 1024u; return p; }
  for (int i = 0; i < a.n; ++i) p.e[n + i] = a.w[i];
  p.hdr = (p.hdr & ~31u) | (n + (uint32_t)a.n);
  return p;
}
// A roulette survivor's coefficient was divided by its weight, the product
// of the own-coefficient maxima since the last reset: flag those vertices.
RG_HD void path_roulette(PathRec& p) {
  uint32_t n = p.hdr & 31u, i = (p.hdr >> 5) & 31u;
  while (i < n) {
    p.e[i] |= PV_RR;
    i += 1u + (uint32_t)path_payload((p.e[i] >> 12) & 15u);
  }
  p.hdr = (p.hdr & ~(31u << 5)) | (n << 5);
}

// State of one traced ray, reconstructed from its spawn record.
struct RayS {
  V3 o, d;
  C3 coef;         // cumulative coefficient
  float w;         // rweight
  float crc;       // sintens of own rcoef
  uint32_t sensor;
  uint16_t crtype, rtype;
  int rsrc;
  int rlvl;
  int adepth;
  int scene;
  uint64_t key;
  PathRec path;    // adjoint replay only
};

struct RunCfg {
  Params p;
  int32_t mode;        // 0 rtrace value, 1 rcontrib bins
  int32_t irrad;       // -i (Lambertian replacement at first hit)
  int32_t ncols;       // output doubles per sensor
  int32_t nbinspec;
  int32_t accum;       // rays per output record (rcontrib -c)
  int32_t child_order; // hemisphere children handed out 0: by stratum row, 1: in 8 x 8 blocks of strata
  int32_t nscenes;     // >0: batched launch over `scenes`
  const DevScene* scenes;      // per-scene tables of a batched launch
  const uint32_t* scene_start; // first input ray of each scene (nscenes+1)
  const BinSpec* bins; // per tracked-modifier slot
  double* out;         // [nsensor][ncols]
  double fix;          // > 0: deposits add up as int64 multiples of 1 / fix in `out` (order-independent,
                       // so results repeat bit for bit); 0: FP64 atomics
  unsigned long long* nrays;
  unsigned long long* visits;  // optional: [0] BVH nodes visited, [1] primitives tested
  uint32_t* err;       // error flags
  // adjoint replay (adj != 0): instead of writing `out`, each binned
  // deposit adds d(loss)/d(material parameter) into grad
  int32_t adj;
  const double* adj_w;         // [nsensor][3] d(loss)/d(E), already / accum
  const double* adj_L;         // [ncols] radiance of each output column
  const int32_t* grad_row;     // material index -> grad row (-1: not differentiated)
  double* grad;                // [rows][ADJ_COLS]: rho_r, rho_g, rho_b, specularity, roughness, transmissivity,
                               // transmitted specularity (the last two for trans only)
};

enum : uint32_t { ERR_STACK = 1, ERR_CAPACITY = 2, ERR_NOMAT = 4, ERR_GEN = 8, ERR_PATH = 16, ERR_FIXED = 32 };

// ---- bins (util/reinhartb.cal:19-47, util/klems_full.cal:10-43,
//      util/disk2square.cal:34-90, util/cieskyscan.cal:21-53) ------------
RG_HD double cal_asin_deg(double x) {   // Asin() of reinhartb.cal
  const double DEGREE = PI_D / 180;
  double v = (x - 1 > 0) ? PI_D / 2 : ((-1 - x > 0) ? -PI_D / 2 : asin(x));
  return v / DEGREE;
}
RG_HD double cal_acos_deg(double x) {   // Acos() of klems_full.cal
  const double DEGREE = PI_D / 180;
  double v = (x - 1 > 0) ? 0 : ((-1 - x > 0) ? PI_D : acos(x));
  return v / DEGREE;
}
RG_HD double cal_atan2_deg(double y, double x) {   // Atan2() / posangle()
  const double DEGREE = PI_D / 180;
  double a = atan2(y, x);
  return ((-a > 0) ? a + 2 * PI_D : a) / DEGREE;
}

// tnaz() of reinhartb.cal:35 and the Klems row tables (element 0 = row count)
__device__ __constant__ static const int kTnaz[7] = {30, 30, 24, 24, 18, 12, 6};
__device__ __constant__ static const double kKlems[6][10] = {
    {9, 5, 15, 25, 35, 45, 55, 65, 75, 90},         // kpola, klems_full.cal:13
    {9, 1, 8, 16, 20, 24, 24, 24, 16, 12},          // knaz
    {7, 6.5, 19.5, 32.5, 45.5, 58.5, 71.5, 90, 0, 0},   // khpola, klems_half.cal:15
    {7, 1, 8, 12, 16, 20, 12, 8, 0, 0},             // khnaz
    {5, 9, 27, 45, 63, 90, 0, 0, 0, 0},             // kqpola, klems_quarter.cal:15
    {5, 1, 8, 12, 12, 8, 0, 0, 0, 0}};              // kqnaz

__device__ inline int reinhart_rnaz(int r, double mf) {
  if (r - (7 * mf - .5) > 0) return 1;
  int k = (int)floor((r + .5) / mf);   // select() is 1-based: tnaz(k+1) = tnaz[k]
  if (k < 0 || k > 6) return 0;
  return (int)(mf * kTnaz[k]);
}

__device__ inline double reinhart_bin(V3 D, const BinSpec& b) {
  double rNx = b.nrm[0], rNy = b.nrm[1], rNz = b.nrm[2];
  double Ux = b.up[0], Uy = b.up[1], Uz = b.up[2];
  double inc_dz = -D.x * rNx - D.y * rNy - D.z * rNz;
  double inc_rx = -b.rhs * (D.x * (Uy * rNz - Uz * rNy) + D.y * (Uz * rNx - Ux * rNz) + D.z * (Ux * rNy - Uy * rNx));
  double inc_ry = D.x * Ux + D.y * Uy + D.z * Uz + inc_dz * (rNx * Ux + rNy * Uy + rNz * Uz);
  double r_alt = cal_asin_deg(inc_dz);
  double r_azi = cal_atan2_deg(inc_rx, inc_ry);
  if (!(r_alt > 0)) return -1;
  double alpha = 90 / (b.mf * 7 + .5);
  int r_row = (int)floor(r_alt / alpha);
  int naz = reinhart_rnaz(r_row, b.mf);
  if (naz <= 0) return -1;
  double r_inc = 360.0 / naz;
  double r_azn = (359.9999 - .5 * r_inc - r_azi > 0) ? floor((r_azi + .5 * r_inc) / r_inc) : 0;
  double acc = 0;
  for (int r = 0; r < r_row; ++r) acc += reinhart_rnaz(r, b.mf);
  return acc + r_azn;
}

__device__ inline double klems_bin(V3 D, const BinSpec& b, int which) {
  const double* kpola = kKlems[2 * which];
  const double* knaz = kKlems[2 * which + 1];
  double Nx = b.nrm[0], Ny = b.nrm[1], Nz = b.nrm[2];
  double Ux = b.up[0], Uy = b.up[1], Uz = b.up[2];
  double pol = cal_acos_deg(-D.x * Nx - D.y * Ny - D.z * Nz);
  double azi = cal_atan2_deg(-D.x * Ux - D.y * Uy - D.z * Uz + (Nx * D.x + Ny * D.y + Nz * D.z) * (Nx * Ux + Ny * Uy + Nz * Uz),
                             -b.rhs * (D.x * (Uy * Nz - Uz * Ny) + D.y * (Uz * Nx - Ux * Nz) + D.z * (Ux * Ny - Uy * Nx)));
  if (pol - 90 > 0) return -1;
  int nrows = (int)kpola[0];
  int r = 1;   // kfindrow(1, pol)
  while (!(r - kpola[0] + .5 > 0) && (pol - kpola[r] > 0)) r++;
  if (r > nrows) r = nrows;
  double acc = 0;
  for (int k = 1; k < r; ++k) acc += knaz[k];
  double inc = 360 / knaz[r];
  double azn = ((360 - .5 * inc) - azi > 0) ? floor((azi + .5 * inc) / inc) : 0;
  return acc + azn;
}

// Oriented axis values shared by rbin, scbin and cbin (reinhartb.cal:27-30)
struct IncAxes { double dz, rx, ry; };
__device__ inline IncAxes inc_axes(V3 D, const BinSpec& b) {
  double rNx = b.nrm[0], rNy = b.nrm[1], rNz = b.nrm[2];
  double Ux = b.up[0], Uy = b.up[1], Uz = b.up[2];
  IncAxes a;
  a.dz = -D.x * rNx - D.y * rNy - D.z * rNz;
  a.rx = -b.rhs * (D.x * (Uy * rNz - Uz * rNy) + D.y * (Uz * rNx - Ux * rNz) + D.z * (Ux * rNy - Uy * rNx));
  a.ry = D.x * Ux + D.y * Uy + D.z * Uz + a.dz * (rNx * Ux + rNy * Uy + rNz * Uz);
  return a;
}

// scbin of util/disk2square.cal:67-90: the incident direction projected to
// the unit disk, mapped to the square (:34-64), binned SCdim x SCdim
__device__ inline double shirley_chiu_bin(V3 D, const BinSpec& b) {
  IncAxes a = inc_axes(D, b);
  if (!(a.dz > 0)) return -1;
  double den2 = a.rx * a.rx + a.ry * a.ry;
  double radf = (den2 - 1e-7 > 0) ? sqrt((1 - a.dz * a.dz) / den2) : 0;
  double x = a.rx * radf, y = -a.ry * radf;
  double r = sqrt(x * x + y * y), phi = atan2(y, x);
  if (-phi - PI_D / 4 > 0) phi += 2 * PI_D;
  int rgn = (int)floor((phi + PI_D / 4) / (PI_D / 2)) + 1;
  double q = r / (PI_D / 4), sa, sb;
  if (rgn == 1) { sa = r; sb = phi * q; }
  else if (rgn == 2) { sa = (PI_D / 2 - phi) * q; sb = r; }
  else if (rgn == 3) { sa = -r; sb = (PI_D - phi) * q; }
  else if (rgn == 4) { sa = (phi - 3 * PI_D / 2) * q; sb = -r; }
  else { sa = r; sb = -r; }
  return floor((sa + 1) / 2 * b.mf) * b.mf + floor((sb + 1) / 2 * b.mf);
}

// cbin of util/cieskyscan.cal:21-53: 145 CIE sky-scanner patches, 12-degree
// rows, azimuth from -U winding alternately by row
__device__ inline double cie_bin(V3 D, const BinSpec& b) {
  IncAxes a = inc_axes(D, b);
  double r_alt = cal_asin_deg(a.dz);
  if (!(r_alt > 0)) return -1;
  double c_azi = cal_atan2_deg(-a.rx, -a.ry);
  double row = floor(r_alt / (90 / 7.5));
  int ri = (int)row;
  if (ri < 0 || ri > 7) return -1;
  double naz = ri < 7 ? (double)kTnaz[ri] : 1.0;   // cnaz() = tnaz() plus the zenith patch
  double acc = 0;
  for (int k = 0; k < ri; ++k) acc += kTnaz[k];
  double fwd = floor(row / 2 + 1.25) - row / 2 - .75;
  double inc = 360 / naz;
  double azn = (359.9999 - c_azi > 0) ? (fwd > 0 ? floor(c_azi / inc) : naz - 1 - floor(c_azi / inc)) : 0;
  return acc + azn;
}

__device__ inline int eval_bin(V3 D, const BinSpec& b) {
  double bval;
  switch (b.kind) {
    case BK_CONST: bval = b.cbin; break;
    case BK_REINHART: bval = reinhart_bin(D, b); break;
    case BK_KLEMS_FULL: bval = klems_bin(D, b, 0); break;
    case BK_KLEMS_HALF: bval = klems_bin(D, b, 1); break;
    case BK_KLEMS_QUARTER: bval = klems_bin(D, b, 2); break;
    case BK_HALFSPACE: bval = (-D.x * b.nrm[0] - D.y * b.nrm[1] - D.z * b.nrm[2] > 0) ? 0.0 : -1.0; break;
    case BK_SHIRLEY_CHIU: bval = shirley_chiu_bin(D, b); break;
    default: bval = cie_bin(D, b); break;
  }
  if (bval <= -.5) return -1;
  int bn = (int)(bval + .5);
  if (bn >= b.nbins) return -1;
  return bn;
}

}  // namespace rg
