// Ray-primitive tests and the closest-hit BVH walk.
//
// The primitive tests restate the reference's FP64 intersection code:
//   polygon  rt/o_face.c:16-62 + common/face.c:121-162 (crossing-number test)
//   sphere   rt/sphere.c:16-83, quadratic() common/zeroes.c:19-55
//   cone/cup/cylinder/tube/ring  rt/o_cone.c:17-146, getcone() common/cone.c
// and the tie rule for coincident hits of rt/raytrace.c:536-575 (rayreject).
// BVH boxes are FP32, widened at build time, and only prune work; every
// accepted hit is decided by the FP64 test.
#pragma once
#include "rg_math.hpp"

namespace rg {

// Two child boxes per internal node: one 64-byte fetch tests both children.
struct alignas(16) BVHNode {
  float lmin[3]; int32_t left;    // child ref: >=0 internal node, <0 leaf ~pos
  float lmax[3]; int32_t right;
  float rmin[3]; int32_t pad0;
  float rmax[3]; int32_t pad1;
};

// Four child boxes per node, coordinates grouped by axis: one 128-byte fetch
// tests four children.  Built from the binary tree by pulling each child's
// children up (a leaf child stays as it is); child refs are indices into the
// four-wide array (root 0) or leaf refs; unused slots hold INT32_MIN.
struct alignas(16) BVH4Node {
  float lox[4], loy[4], loz[4];
  float hix[4], hiy[4], hiz[4];
  int32_t child[4];
  int32_t pad[4];
};

// Planes of the first kBigPlanes room-sized primitives, in one contiguous
// table: the walk estimates their hit distances from these before any
// primitive test, with one load per entry instead of the chain big index ->
// primitive record -> geometry, and tests a polygon among them from the
// entry's plane, geometry offset and vertex count.  Polygons hold their
// plane (normal, offset); other shapes, and entries past the primitive
// count, hold zeros (the estimate then defers to the full test).
constexpr int kBigPlanes = 8;
struct alignas(16) BigPlane {
  double n[3];
  double off;
  int32_t prim;      // primitive index (-1: no entry)
  int32_t kind;      // PrimKind
  int32_t goff;      // geometry offset (doubles)
  int32_t nv;        // polygon vertex count
};

struct DevScene {
  const RG_GLOBAL Prim* prims;
  const RG_GLOBAL double* geom;
  const RG_GLOBAL Material* mats;
  const RG_GLOBAL Source* srcs;
  const RG_GLOBAL BVHNode* nodes;
  const RG_GLOBAL BVH4Node* nodes4;    // four-wide form of `nodes` (root 0); null: none
  const RG_GLOBAL int32_t* leaf_prim;   // sorted leaf position -> primitive index
  const RG_GLOBAL int32_t* big;         // primitives kept out of the BVH, tested first
  const RG_GLOBAL BigPlane* bigpl;      // kBigPlanes entries, see BigPlane
  int32_t nbig;
  int32_t nbvh;                 // primitives inside the BVH
  int32_t nprims;
  int32_t nmats;
  int32_t nsrcs;
  int32_t root;                 // root node (<0: single leaf)
  int32_t lamb;                 // index of the built-in Lambertian material
  int32_t shadow_on;            // shadow cache in use (caches filled)
  double cusize;                // scene cube size (ambient distance clamp)
  const RG_GLOBAL SrcCache* scache;             // per source
  const RG_GLOBAL unsigned long long* shadow_obs;  // blocker cells: (sequence+1) << 32 | primitive; 0 empty
  RG_GLOBAL uint32_t* src_tested;               // per source: shadow-tested during the current call
};

struct HitRec {
  double t;                     // rot
  double rod;
  int32_t prim;                 // -1: none
};

RG_HD bool ray_reject(const DevScene& S, int prim, double t, double rod, const HitRec& h) {
  if ((t <= FTINY) | (t > h.t + FTINY)) return true;
  if (t < h.t - FTINY) return false;
  if (h.prim == prim) return true;
  if (h.prim < 0) return false;
  int mn = S.prims[prim].mat, mr = S.prims[h.prim].mat;
  if (mn < 0) {
    if (mr >= 0) return true;
  } else if (mr < 0) {
    return false;
  } else if (S.mats[mn].transp) {
    if (!S.mats[mr].transp) return true;
  } else if (S.mats[mr].transp) {
    return false;
  }
  if (rod <= 0) {
    if (h.rod > 0) return true;
  } else if (h.rod <= 0) {
    return false;
  }
  return S.prims[h.prim].omod >= S.prims[prim].omod;
}

// polygon geometry block: n[3], offset, ax, then nv vertices
RG_HD bool in_face(V3 p, const double* g, int nv) {
  int ax = (int)g[4];
  int xi = ax + 1; if (xi >= 3) xi -= 3;
  int yi = xi + 1; if (yi >= 3) yi -= 3;
  double x = comp(p, xi), y = comp(p, yi);
  const double* va = g + 5;
  const double* p0 = va + 3 * (nv - 1);
  const double* p1 = va;
  int ncross = 0;
  for (int n = nv; n--;) {
    if (fabs(p0[yi] - y) <= FTINY && fabs(p1[yi] - y) <= FTINY && ((p0[xi] > x) ^ (p1[xi] > x)))
      return true;
    if ((p0[yi] > y) ^ (p1[yi] > y)) {
      int tst = (p0[xi] > x) + (p1[xi] > x);
      if (tst == 2) {
        ncross++;
      } else if (tst) {
        double prodA = (p0[yi] - y) * (p1[xi] - x);
        double prodB = (p0[xi] - x) * (p1[yi] - y);
        if (fabs(prodA - prodB) <= FTINY) return true;
        ncross += (p1[yi] > p0[yi]) ^ (prodA > prodB);
      } else if (fabs(p0[xi] - x) <= FTINY && fabs(p1[xi] - x) <= FTINY) {
        return true;
      }
    }
    p0 = p1;
    p1 += 3;
  }
  return ncross & 1;
}

RG_HD int quadratic(double r[2], double a, double b, double c) {
  int first;
  if (a < -FTINY) first = 1;
  else if (a > FTINY) first = 0;
  else if (fabs(b) > FTINY) { r[0] = -c / b; return 1; }
  else return 0;
  b *= 0.5;
  double disc = b * b - a * c;
  if (disc < -FTINY * FTINY) return 0;
  if (disc <= FTINY * FTINY) { r[0] = -b / a; return 1; }
  disc = sqrt(disc);
  r[first] = (-b - disc) / a;
  r[1 - first] = (-b + disc) / a;
  return 2;
}

// Signed volume of the ray against one triangle edge (volume_sign() of
// rt/o_mesh.c:60-95): computed with the edge's lower-id vertex first, so the
// two triangles sharing an edge see opposite signs and a ray through the
// edge is claimed by exactly one of them.  `rev`: the stored order runs
// against the id order.  Never 0.
RG_HD int tri_vsign(V3 o, V3 d, V3 p, V3 q, bool rev) {
  if (rev) { V3 t = p; p = q; q = t; }
  V3 a = p - o, b = q - o;
  double vol = a.x * (b.y * d.z - b.z * d.y) + a.y * (b.z * d.x - b.x * d.z) + a.z * (b.x * d.y - b.y * d.x);
  int sg = vol > .0 ? 1 : -1;
  return rev ? -sg : sg;
}

// Barycentric weights of p in (v1, v2, v3): comp_baryc() / eval_baryc() of
// common/tmesh.c:44-110 (projection onto the dominant plane).
RG_HD bool tri_baryc(double wt[3], V3 p, V3 v1, V3 v2, V3 v3) {
  V3 vab = v1 - v2, vcb = v3 - v2;
  V3 va = cross(vab, vcb);
  int ax = va.y * va.y > va.x * va.x;
  if (va.z * va.z > comp(va, ax) * comp(va, ax)) ax = 2;
  int ax0 = (ax + 1) % 3, ax1 = (ax + 2) % 3;
  double tm[2][3];
  for (int i = 0; i < 2; ++i) {
    double ab0 = comp(v1, ax0) - comp(v2, ax0), cb0 = comp(v3, ax0) - comp(v2, ax0);
    double ab1 = comp(v1, ax1) - comp(v2, ax1), cb1 = comp(v3, ax1) - comp(v2, ax1);
    double dd = cb0 * cb0 + cb1 * cb1;
    if (dd <= FTINY * FTINY) return false;
    dd = (cb0 * ab0 + cb1 * ab1) / dd;
    double a0 = ab0 - cb0 * dd, a1 = ab1 - cb1 * dd;
    dd = a0 * a0 + a1 * a1;
    if (dd <= FTINY * FTINY) return false;
    dd = 1.0 / dd;
    tm[i][0] = a0 * dd;
    tm[i][1] = a1 * dd;
    tm[i][2] = -(comp(v2, ax0) * tm[i][0] + comp(v2, ax1) * tm[i][1]);
    V3 t = v1; v1 = v2; v2 = v3; v3 = t;
  }
  double u = comp(p, ax0), v = comp(p, ax1);
  wt[0] = u * tm[0][0] + v * tm[0][1] + tm[0][2];
  wt[1] = u * tm[1][0] + v * tm[1][1] + tm[1][2];
  wt[2] = 1. - wt[1] - wt[0];
  return true;
}

// Test one primitive; updates h when it wins.
// Polygon pi (plane normal n and offset off, geometry g, nv vertices):
// o_face.c's plane distance, hit rejection and inside test.
RG_HD void hit_polygon(const DevScene& S, int pi, V3 n, double off, const double* g, int nv, V3 o, V3 d,
                       HitRec& h) {
  double rdot = -dot(d, n);
  if ((rdot <= FTINY) & (rdot >= -FTINY)) return;
  double t = (dot(o, n) - off) / rdot;
  if (ray_reject(S, pi, t, rdot, h)) return;
  V3 p = o + d * t;
  if (!in_face(p, g, nv)) return;
  h.t = t; h.rod = rdot; h.prim = pi;
}

RG_HD void hit_prim(const DevScene& S, int pi, V3 o, V3 d, HitRec& h) {
  const Prim& P = S.prims[pi];
  const double* g = S.geom + P.goff;
  switch (P.kind) {
    case PK_TRIANGLE: {
      // mesh_hit() of rt/o_mesh.c:98-141: equal edge-volume signs, then the
      // plane distance along (v0-v2)x(v1-v0)
      V3 a = load3(g), b = load3(g + 3), c = load3(g + 6);
      int fl = (int)g[12];
      int s1 = tri_vsign(o, d, a, b, fl & TRI_REV0);
      int s2 = tri_vsign(o, d, b, c, fl & TRI_REV1);
      if (s1 != s2) return;
      if (s2 != tri_vsign(o, d, c, a, fl & TRI_REV2)) return;
      V3 nrm = cross(a - c, b - a);
      double dn = dot(d, nrm);
      if (dn == 0.0) return;
      double t = dot(a - o, nrm) / dn;
      double rod = -dot(d, load3(g + 9));
      if (ray_reject(S, pi, t, rod, h)) return;
      h.t = t; h.rod = rod; h.prim = pi;
      return;
    }
    case PK_POLYGON:
      hit_polygon(S, pi, load3(g), g[3], g, P.nv, o, d, h);
      return;
    case PK_SPHERE:
    case PK_BUBBLE: {
      double a = dot(d, d);
      V3 oc = o - load3(g);
      double b = 2.0 * dot(d, oc);
      double c = dot(oc, oc) - g[3] * g[3];
      double root[2];
      int nr = quadratic(root, a, b, c);
      int i;
      double t = 0;
      for (i = 0; i < nr; i++)
        if ((t = root[i]) > FTINY) break;
      if (i >= nr) return;
      double rod = 1 - 2 * ((i > 0) ^ (P.kind == PK_BUBBLE));
      if (ray_reject(S, pi, t, rod, h)) return;
      // the reference's rod for the hit is recomputed from the normal
      double rad = g[3] * (1 - 2 * (P.kind == PK_BUBBLE));
      V3 p = o + d * t;
      V3 nrm = (p - load3(g)) * (1.0 / rad);
      h.t = t; h.rod = -dot(d, nrm); h.prim = pi;
      return;
    }
    case PK_RING: {
      // g: centre[3], axis[3] (unit), r0, r1
      V3 ad = load3(g + 3);
      double dz = dot(d, ad);
      if ((dz <= FTINY) & (dz >= -FTINY)) return;
      double t = -dot(o - load3(g), ad) / dz;
      if (ray_reject(S, pi, t, -dz, h)) return;
      V3 p = o + d * t - load3(g);
      double a = dot(p, p);   // in-plane radius^2 (p lies in the ring plane)
      if (a > g[7] * g[7] || a < g[6] * g[6]) return;
      h.t = t; h.rod = -dz; h.prim = pi;
      return;
    }
    default: {
      // cone, cup, cylinder, tube. g: p0[3], ad[3] (unit axis), al, r0, r1, sl
      V3 p0 = load3(g), ad = load3(g + 3);
      double al = g[6], r0 = g[7], r1 = g[8];
      bool cyl = (P.kind == PK_CYLINDER) | (P.kind == PK_TUBE);
      V3 w = o - p0;
      double dd = dot(d, ad), wd = dot(w, ad);
      V3 dp = d - ad * dd, wp = w - ad * wd;        // components normal to the axis
      double a, b, c;
      if (cyl) {
        a = dot(dp, dp);
        b = 2.0 * dot(dp, wp);
        c = dot(wp, wp) - r0 * r0;
      } else {
        // radius along the axis: rr(h) = r0 + k*h, k = (r1-r0)/al
        double k = (r1 - r0) / al;
        double rw = r0 + k * wd, rd = k * dd;
        a = dot(dp, dp) - rd * rd;
        b = 2.0 * (dot(dp, wp) - rw * rd);
        c = dot(wp, wp) - rw * rw;
      }
      double root[2];
      int nr = quadratic(root, a, b, c);
      bool inward = (P.kind == PK_CUP) | (P.kind == PK_TUBE);
      for (int rn = 0; rn < nr; rn++) {
        double t = root[rn];
        if (t <= FTINY) continue;
        if (t > h.t + FTINY) break;
        V3 p = o + d * t;
        double hb = dot(p - p0, ad);
        if (hb < 0.0 || hb > al) continue;
        if (!cyl) {      // reject the mirror nappe of the double cone
          double rr = r0 + (r1 - r0) * hb / al;
          if (rr < 0.0) continue;
        }
        double rod_sign = 1 - 2 * ((rn > 0) ^ inward);
        if (ray_reject(S, pi, t, rod_sign, h)) break;
        V3 rdx = p - p0;
        V3 nrm;
        if (cyl) {
          double aa = (P.kind == PK_CYLINDER) ? r0 : -r0;
          nrm = (rdx - ad * hb) * (1.0 / aa);
        } else {
          double cc = r1 - r0;
          double aa = r0 + hb * cc / al;
          if (P.kind == PK_CUP) { cc = -cc; aa = -aa; }
          nrm = (rdx - ad * hb) * (1.0 / aa);
          nrm = (nrm * al - ad * cc) * (1.0 / g[9]);
        }
        double l2 = dot(nrm, nrm);
        if ((l2 > 1. + FTINY) | (l2 < 1. - FTINY)) nrm = nrm * (1. / (.5 + .5 * l2));
        h.t = t; h.rod = -dot(d, nrm); h.prim = pi;
        return;
      }
      return;
    }
  }
}

// Surface normal at a hit (the reference's ron), FP64.
RG_HD V3 prim_normal_k(const DevScene& S, int kind, int goff, V3 p) {
  const double* g = S.geom + goff;
  switch (kind) {
    case PK_POLYGON: return load3(g);
    case PK_TRIANGLE: return load3(g + 9);
    case PK_SPHERE:
    case PK_BUBBLE: {
      double rad = g[3] * (1 - 2 * (kind == PK_BUBBLE));
      return (p - load3(g)) * (1.0 / rad);
    }
    case PK_RING: return load3(g + 3);
    default: {
      V3 p0 = load3(g), ad = load3(g + 3);
      double al = g[6], r0 = g[7], r1 = g[8];
      V3 rdx = p - p0;
      double hb = dot(rdx, ad);
      V3 nrm;
      if ((kind == PK_CYLINDER) | (kind == PK_TUBE)) {
        double aa = (kind == PK_CYLINDER) ? r0 : -r0;
        nrm = (rdx - ad * hb) * (1.0 / aa);
      } else {
        double cc = r1 - r0;
        double aa = r0 + hb * cc / al;
        if (kind == PK_CUP) { cc = -cc; aa = -aa; }
        nrm = (rdx - ad * hb) * (1.0 / aa);
        nrm = (nrm * al - ad * cc) * (1.0 / g[9]);
      }
      double l2 = dot(nrm, nrm);
      if ((l2 > 1. + FTINY) | (l2 < 1. - FTINY)) nrm = nrm * (1. / (.5 + .5 * l2));
      return nrm;
    }
  }
}
RG_HD V3 prim_normal(const DevScene& S, int pi, V3 p) {
  const Prim& P = S.prims[pi];
  return prim_normal_k(S, P.kind, P.goff, p);
}

// Normal perturbation at a hit (the reference's pert): mesh triangles with
// vertex normals interpolate them barycentrically (rt/o_mesh.c:192-204);
// everything else is unperturbed.
RG_HD V3 prim_pert_k(const DevScene& S, int kind, int goff, V3 p, V3 ron) {
  V3 z = {0.0, 0.0, 0.0};
  if (kind != PK_TRIANGLE) return z;
  const double* g = S.geom + goff;
  if (!((int)g[12] & TRI_NORMALS)) return z;
  double wt[3];
  if (!tri_baryc(wt, p, load3(g), load3(g + 3), load3(g + 6))) return z;
  V3 pn = load3(g + 13) * wt[0] + load3(g + 16) * wt[1] + load3(g + 19) * wt[2];
  double l2 = dot(pn, pn);
  if (l2 == 0.0) return z;
  pn = pn * (1.0 / sqrt(l2));
  return pn - ron;
}
RG_HD V3 prim_pert(const DevScene& S, int pi, V3 p, V3 ron) {
  const Prim& P = S.prims[pi];
  return prim_pert_k(S, P.kind, P.goff, p, ron);
}

RG_HD bool prim_phong_k(const DevScene& S, int kind, int goff) {
  return kind == PK_TRIANGLE && ((int)S.geom[goff + 12] & TRI_PHONG);
}
RG_HD bool prim_phong(const DevScene& S, int pi) {
  const Prim& P = S.prims[pi];
  return prim_phong_k(S, P.kind, P.goff);
}

}  // namespace rg
