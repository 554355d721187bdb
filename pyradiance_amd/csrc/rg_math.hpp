// Small vector helpers and the counter-based random number generator shared by
// every kernel of the renderer.
//
// Random numbers: each ray carries a 64-bit key derived from its path address
// (sensor, then the child slot taken at every spawn).  A sample is a pure
// function of (key, index), so the same path draws the same numbers no matter
// how the wavefront schedules it.
#pragma once
#include <math.h>
#include <stdint.h>
#include "rg_types.hpp"

namespace rg {

struct V3 {
  double x, y, z;
};
RG_HD V3 v3(double x, double y, double z) { return V3{x, y, z}; }
RG_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RG_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RG_HD V3 operator*(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
RG_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
RG_HD double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RG_HD V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RG_HD double comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// normalize() of common/fvect.c:132-155: returns the length, leaves zero
// vectors alone; within FTINY of unit length it takes the first-order
// reciprocal 2 - (1 + |v|^2) / 2.  The arithmetic is in the order the
// reference's -O3 -ffast-math build evaluates it, |v|^2 = (x^2 + z^2) + y^2,
// without fused multiply-adds: source directions feed the shadow cache's
// cell indices, where the last bit can decide a cell.
RG_HD double normalize(V3& v) {
#pragma clang fp contract(off)
  double l2 = (v.x * v.x + v.z * v.z) + v.y * v.y;
  if (l2 == 0.0) return 0.0;
  double l, inv;
  if ((l2 <= 1.0 + 1e-6) & (l2 >= 1.0 - 1e-6)) {
    l = (l2 + 1.0) * 0.5;
    inv = 2.0 - l;
  } else {
    l = sqrt(l2);
    inv = 1.0 / l;
  }
  v = V3{v.x * inv, v.y * inv, v.z * inv};
  return l;
}
RG_HD V3 load3(const double* p) { return V3{p[0], p[1], p[2]}; }

struct C3 {            // RGB coefficient / colour
  float r, g, b;
};
RG_HD C3 c3(float r, float g, float b) { return C3{r, g, b}; }
RG_HD C3 operator*(C3 a, C3 b) { return C3{a.r * b.r, a.g * b.g, a.b * b.b}; }
RG_HD C3 operator*(C3 a, double s) { return C3{(float)(a.r * s), (float)(a.g * s), (float)(a.b * s)}; }
RG_HD C3 operator+(C3 a, C3 b) { return C3{a.r + b.r, a.g + b.g, a.b + b.b}; }
RG_HD double cmax(C3 a) {   // sintens(): largest channel
  return a.r > a.g ? (a.r > a.b ? a.r : a.b) : (a.g > a.b ? a.g : a.b);
}
RG_HD double cmean(C3 a) { return (1.0 / 3.0) * ((double)a.r + a.g + a.b); }
// pbright() with the default RGB primaries (common/color.h:257): CIE Y weights
RG_HD double cbright(C3 a) { return 0.265074126 * a.r + 0.670114631 * a.g + 0.064811243 * a.b; }

// ---- random numbers --------------------------------------------------------
RG_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
RG_HD uint64_t sensor_key(uint64_t seed, uint64_t sensor) {
  return mix64(seed * 0x9e3779b97f4a7c15ULL ^ mix64(sensor + 1));
}
RG_HD uint64_t child_key(uint64_t k, uint64_t slot) {
  return mix64(k ^ mix64(slot + 0x9e3779b97f4a7c15ULL));
}
// uniform double in [0,1) with 53 random bits
RG_HD double rnd(uint64_t k, uint32_t i) {
  return (double)(mix64(k + 0x632be59bd9b4e019ULL * (uint64_t)(i + 1)) >> 11) * (1.0 / 9007199254740992.0);
}

// child-slot numbering of spawned rays (part of the path address)
enum : uint64_t {
  SLOT_TRANS = 1, SLOT_REFL = 2, SLOT_RSPEC = 3, SLOT_TSPEC = 4, SLOT_PASS = 5,
  SLOT_SHADOW = 64,                  // + source index
  SLOT_HEMI_FRONT = 1ULL << 20,      // + sample index
  SLOT_HEMI_BACK = 2ULL << 20,       // + sample index
};
// per-ray random-number indices
enum : uint32_t {
  RN_RR = 0, RN_AMB0 = 1, RN_AMB1 = 2, RN_SRC0 = 3, RN_SRC1 = 4, RN_SRC2 = 5,
  RN_FRAME_FRONT = 10, RN_FRAME_BACK = 11, RN_GFRAME = 12, RN_GAUSS = 16,
};

// getperpendicular() of common/fvect.c:159-192 with randomize = 0
RG_HD V3 perpendicular(V3 v) {
  int i;
  for (i = 3; i--;) {
    double c = comp(v, i);
    if ((-0.6 < c) & (c < 0.6)) break;
  }
  V3 v1 = V3{0, 0, 0};
  if (i < 0) i = 0;
  if (i == 0) v1.x = 1.0; else if (i == 1) v1.y = 1.0; else v1.z = 1.0;
  V3 r = cross(v1, v);
  normalize(r);
  return r;
}

// A random rotation of the tangent frame about n (stands in for the random
// coordinate order of getperpendicular(randomize=1)).
RG_HD void tangent_frame(V3 n, double u, V3& ux, V3& uy) {
  V3 a = perpendicular(n);
  V3 b = cross(n, a);
  double ang = 2.0 * PI_D * u;
  double c = cos(ang), s = sin(ang);
  ux = a * c + b * s;
  normalize(ux);
  uy = cross(n, ux);
}

// square2disk() of common/disk2square.c:44-86 (Shirley-Chiu concentric map)
RG_HD void square2disk(double ds[2], double seedx, double seedy) {
  double phi, r;
  double a = 2. * seedx - 1;
  double b = 2. * seedy - 1;
  if (a > -b) {
    if (a > b) { r = a; phi = (PI_D / 4.) * (b / a); }
    else { r = b; phi = (PI_D / 4.) * (2. - (a / b)); }
  } else {
    if (a < b) { r = -a; phi = (PI_D / 4.) * (4. + (b / a)); }
    else {
      r = -b;
      if (b != 0.) phi = (PI_D / 4.) * (6. - (a / b));
      else phi = 0.;
    }
  }
  r *= 0.9999999999999;
  ds[0] = r * cos(phi);
  ds[1] = r * sin(phi);
}

}  // namespace rg
