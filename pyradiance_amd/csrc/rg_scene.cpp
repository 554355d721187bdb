// Scene compilation: the host half of subsystem (a).
//
//   findmaterial()          rt/initotypes.c:113-145
//   getface()               common/face.c:35-117
//   getcone() radii rules   common/cone.c:20-120
//   o_sphere() radius rule  rt/sphere.c:31-39
//   marksources()           rt/source.c:47-142
//   ssetsrc/fsetsrc/sphsetsrc/rsetsrc + setflatss  rt/srcsupp.c:60-260
//   xf()                    common/xf.c:40-150
//   mesh triangles          rt/o_mesh.c:146-217, common/mesh.c:174-278
#include "rg_scene.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <unordered_map>

namespace rg {

namespace {

struct D3 {
  double v[3];
};
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross3(double* r, const double* a, const double* b) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
// normalize() of common/fvect.c:132-155, first-order reciprocal near unit
// length included, in the operation order of the reference's -O3
// -ffast-math build ((x^2 + z^2) + y^2; source directions feed the shadow
// cache's cell indices)
inline double norm3(double* v) {
  double l2 = (v[0] * v[0] + v[2] * v[2]) + v[1] * v[1];
  if (l2 == 0.0) return 0;
  double l, inv;
  if ((l2 <= 1.0 + FTINY) & (l2 >= 1.0 - FTINY)) {
    l = (l2 + 1.0) * 0.5;
    inv = 2.0 - l;
  } else {
    l = sqrt(l2);
    inv = 1.0 / l;
  }
  v[0] *= inv; v[1] *= inv; v[2] *= inv;
  return l;
}
// getperpendicular(randomize=0), common/fvect.c:159-192
bool perpendicular3(double* vp, const double* v) {
  int i;
  for (i = 3; i--;)
    if ((-0.6 < v[i]) & (v[i] < 0.6)) break;
  if (i < 0) return false;
  double v1[3] = {0, 0, 0};
  v1[i] = 1.0;
  cross3(vp, v1, v);
  return norm3(vp) > 0.0;
}

[[noreturn]] void objerr(const RObj& o, const std::string& m) {
  throw RadError(o.type + " \"" + o.name + "\": " + m);
}

// xf() of common/xf.c:40-150: transform options -> 4x4 matrix for row
// vectors (p' = p M) and the scale factor (negative after an odd number of
// mirrors).  Returns the number of arguments used.
struct Xf {
  double m[4][4];
  double sca;
};
void ident4(double m[4][4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) m[i][j] = i == j;
}
void mult4(double r[4][4], const double a[4][4], const double b[4][4]) {
  double t[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) t[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j] + a[i][3] * b[3][j];
  memcpy(r, t, sizeof t);
}
bool is_number(const std::string& w, bool integer) {
  if (w.empty()) return false;
  char* e = nullptr;
  if (integer) strtol(w.c_str(), &e, 10);
  else strtod(w.c_str(), &e);
  return e && *e == '\0';
}
int parse_xf(Xf& ret, const std::vector<std::string>& av, size_t start) {
  const int ac = (int)(av.size() - start);
  auto A = [&](int i) -> const std::string& { return av[start + i]; };
  auto ok = [&](int i, size_t optlen, const char* types) {   // checkarg()
    if (A(i).size() != optlen) return false;
    int n = (int)strlen(types);
    if (i + n >= ac) return false;
    for (int k = 0; k < n; ++k)
      if (!is_number(A(i + 1 + k), types[k] == 'i')) return false;
    return true;
  };
  double xfmat[4][4], m4[4][4];
  double xfsca = 1.0;
  int icnt = 1;
  ident4(ret.m);
  ret.sca = 1.0;
  ident4(xfmat);
  int i;
  for (i = 0; i < ac && A(i)[0] == '-' && A(i).size() > 1; ++i) {
    ident4(m4);
    const std::string& o = A(i);
    bool done = false;
    switch (o[1]) {
      case 't':
        if (!ok(i, 2, "fff")) { done = true; break; }
        for (int k = 0; k < 3; ++k) m4[3][k] = atof(A(++i).c_str());
        break;
      case 'r': {
        if (o.size() != 3 || (o[2] != 'x' && o[2] != 'y' && o[2] != 'z') || !ok(i, 3, "f")) { done = true; break; }
        double a = (PI_D / 180.) * atof(A(++i).c_str()), cs = cos(a), sn = sin(a);
        if (o[2] == 'x') { m4[1][1] = m4[2][2] = cs; m4[1][2] = sn; m4[2][1] = -sn; }
        else if (o[2] == 'y') { m4[0][0] = m4[2][2] = cs; m4[2][0] = sn; m4[0][2] = -sn; }
        else { m4[0][0] = m4[1][1] = cs; m4[0][1] = sn; m4[1][0] = -sn; }
        break;
      }
      case 's': {
        if (!ok(i, 2, "f")) { done = true; break; }
        double v = atof(A(i + 1).c_str());
        if (v == 0.0) { done = true; break; }
        ++i;
        xfsca *= v;
        m4[0][0] = m4[1][1] = m4[2][2] = v;
        break;
      }
      case 'm':
        if (o.size() != 3 || (o[2] != 'x' && o[2] != 'y' && o[2] != 'z')) { done = true; break; }
        xfsca *= -1.0;
        m4[o[2] - 'x'][o[2] - 'x'] = -1.0;
        break;
      case 'i':
        if (!ok(i, 2, "i")) { done = true; break; }
        while (icnt-- > 0) {
          mult4(ret.m, ret.m, xfmat);
          ret.sca *= xfsca;
        }
        icnt = atoi(A(++i).c_str());
        ident4(xfmat);
        xfsca = 1.0;
        continue;
      default:
        done = true;
    }
    if (done) break;
    mult4(xfmat, xfmat, m4);
  }
  while (icnt-- > 0) {
    mult4(ret.m, ret.m, xfmat);
    ret.sca *= xfsca;
  }
  return i;
}
void xf_point(double r[3], const double p[3], const Xf& x) {   // multp3()
  for (int k = 0; k < 3; ++k) r[k] = p[0] * x.m[0][k] + p[1] * x.m[1][k] + p[2] * x.m[2][k] + x.m[3][k];
}
void xf_vector(double r[3], const double v[3], const Xf& x) {  // multv3()
  for (int k = 0; k < 3; ++k) r[k] = v[0] * x.m[0][k] + v[1] * x.m[1][k] + v[2] * x.m[2][k];
}

// X followed by the parent transform P (row vectors: M = Mx * Mp)
void compose_xf(Xf& x, const Xf& p) {
  mult4(x.m, x.m, p.m);
  x.sca *= p.sca;
}

// A surface object's real arguments carried through a transform: points
// transformed, axis vectors rotated, radii scaled; a mirrored polygon's
// vertex order is reversed so that its normal stays the transformed normal.
RObj xf_object(const RObj& o, const Xf& X) {
  RObj t = o;
  auto& f = t.fargs;
  const double sca = fabs(X.sca);
  auto pt = [&](size_t k) {
    double r[3];
    xf_point(r, &o.fargs[k], X);
    f[k] = r[0]; f[k + 1] = r[1]; f[k + 2] = r[2];
  };
  if (o.type == "polygon") {
    for (size_t k = 0; k + 2 < f.size(); k += 3) pt(k);
    if (X.sca < 0) {
      size_t nv = f.size() / 3;
      for (size_t a = 1, b = nv - 1; a < b; ++a, --b)
        for (int c = 0; c < 3; ++c) std::swap(f[3 * a + c], f[3 * b + c]);
    }
  } else if (o.type == "sphere" || o.type == "bubble") {
    if (f.size() >= 4) { pt(0); f[3] *= sca; }
  } else if (o.type == "ring") {
    if (f.size() >= 8) {
      pt(0);
      double v[3];
      xf_vector(v, &o.fargs[3], X);
      f[3] = v[0]; f[4] = v[1]; f[5] = v[2];
      f[6] *= sca; f[7] *= sca;
    }
  } else if (f.size() >= 7) {          // cone, cup, cylinder, tube
    pt(0);
    pt(3);
    for (size_t k = 6; k < f.size(); ++k) f[k] *= sca;
  }
  return t;
}

int material_kind(const std::string& t) {
  if (t == "plastic") return MK_PLASTIC;
  if (t == "metal") return MK_METAL;
  if (t == "trans") return MK_TRANS;
  if (t == "glass") return MK_GLASS;
  if (t == "light") return MK_LIGHT;
  if (t == "glow") return MK_GLOW;
  if (t == "illum") return MK_ILLUM;
  return -1;
}

}  // namespace

// Shadow-cache layout and cast rays of every source, as initobscache() sets
// them up (rt/srcobstr.c:63-183): proximity (glow) and skipped sources get
// no cache; sources from the 200th on (MAX2SHADE, :20-21) get cells but no
// cast rays.  Cast rays are listed in the reference's loop order, which
// decides which of two rays aimed at one cell records its blocker.
static void shadow_cache_setup(CompiledScene& cs) {
  constexpr int S = kShadCache;
  cs.scache.assign(cs.srcs.size(), SrcCache{});
  cs.ncells = 0;
  cs.cast_rays.clear();
  cs.cast_src.clear();
  auto cast = [&](int sn, const double* o, const double* d) {
    cs.cast_rays.insert(cs.cast_rays.end(), {o[0], o[1], o[2], d[0], d[1], d[2]});
    cs.cast_src.push_back(sn);
  };
  for (size_t sn = 0; sn < cs.srcs.size(); ++sn) {
    const Source& src = cs.srcs[sn];
    SrcCache& C = cs.scache[sn];
    memset(&C, 0, sizeof C);
    if (src.flags & (SSKIP | SPROX | SSPOT | SVIRTUAL)) continue;
    if (src.flags & SDISTANT) {
      C.kind = SC_DISTANT;
      C.ncell = 4 * S * S;
    } else if (src.flags & SFLAT) {
      C.kind = SC_FLAT;
      C.ncell = S * S * 3 + (S & 1) * S * 4;
    } else {
      C.kind = SC_SPHERE;
      C.ncell = S * S * 6;
    }
    C.cell0 = cs.ncells;
    cs.ncells += C.ncell;
    if (C.kind == SC_DISTANT) {
      double amax = 0;
      int ax = 0;
      for (int a = 3; a--;)
        if (fabs(src.sloc[a]) > amax) { amax = fabs(src.sloc[a]); ax = a; }
      const int ax1 = (ax + 1) % 3, ax2 = (ax + 2) % 3;
      C.ax = ax;
      for (int a = 0; a < 3; ++a) C.o[a] = cs.cuorg[a];
      if (src.sloc[ax] > 0) C.o[ax] += cs.cube_size;
      if (src.sloc[ax1] < 0) C.o[ax1] += cs.cube_size * src.sloc[ax1] / amax;
      if (src.sloc[ax2] < 0) C.o[ax2] += cs.cube_size * src.sloc[ax2] / amax;
      C.e1 = 1. / (cs.cube_size * (1. + fabs(src.sloc[ax1]) / amax));
      C.e2 = 1. / (cs.cube_size * (1. + fabs(src.sloc[ax2]) / amax));
    } else if (C.kind == SC_FLAT) {
      for (int a = 0; a < 3; ++a) C.u[a] = src.ss[0][a];
      norm3(C.u);
      cross3(C.v, src.snorm, C.u);
    }
    if (sn >= 200) continue;
    double org[3], dir[3], d;
    if (C.kind == SC_DISTANT) {
      const int ax1 = (C.ax + 1) % 3, ax2 = (C.ax + 2) % 3;
      for (int a = 0; a < 3; ++a) dir[a] = -src.sloc[a];
      for (int i = 2 * S; i--;)
        for (int j = 2 * S; j--;) {
          // (i + .5) / (2 S e1) as the reference's -O3 -ffast-math build
          // evaluates it (rt/srcobstr.c:127-131): (i + .5) * 0.025 / e1
          for (int a = 0; a < 3; ++a) org[a] = C.o[a];
          org[ax1] += (i + .5) * (1.0 / (2 * S)) / C.e1;
          org[ax2] += (j + .5) * (1.0 / (2 * S)) / C.e2;
          cast((int)sn, org, dir);
        }
    } else if (C.kind == SC_FLAT) {
      d = 0.01 * src.srad;
      for (int a = 0; a < 3; ++a) org[a] = src.sloc[a] + src.snorm[a] * d;
      for (int i = S; i--;)
        for (int j = S; j--;) {
          d = 2. / S * (i + .5) - 1.;
          for (int a = 0; a < 3; ++a) dir[a] = src.snorm[a] + C.u[a] * d;
          d = 2. / S * (j + .5) - 1.;
          for (int a = 0; a < 3; ++a) dir[a] += C.v[a] * d;
          norm3(dir);
          cast((int)sn, org, dir);
        }
      for (int k = 2; k--;)
        for (int i = S; i--;)
          for (int j = S >> 1; j--;) {
            d = 2. / S * (i + .5) - 1.;
            for (int a = 0; a < 3; ++a) dir[a] = k ? C.u[a] + C.v[a] * d : C.v[a] + C.u[a] * d;
            d = 1. - 2. / S * (j + .5);
            for (int a = 0; a < 3; ++a) dir[a] += src.snorm[a] * d;
            norm3(dir);
            cast((int)sn, org, dir);
            d = 2. * dot3(dir, src.snorm);
            for (int a = 0; a < 3; ++a) dir[a] = d * src.snorm[a] - dir[a];
            cast((int)sn, org, dir);
          }
    } else {
      for (int k = 6; k--;) {
        const int ax = k % 3, ax1 = (k + 1) % 3, ax2 = (k + 2) % 3;
        for (int i = S; i--;)
          for (int j = S; j--;) {
            dir[0] = dir[1] = dir[2] = 0.;
            dir[ax] = k < 3 ? 1. : -1.;
            dir[ax1] = 2. / S * (i + .5) - 1.;
            dir[ax2] = 2. / S * (j + .5) - 1.;
            norm3(dir);
            d = 1.05 * src.srad;
            for (int a = 0; a < 3; ++a) org[a] = src.sloc[a] + dir[a] * d;
            cast((int)sn, org, dir);
          }
      }
    }
  }
}

CompiledScene compile_scene(const SceneDB& db) {
  CompiledScene cs;
  const auto& objs = db.objs;
  const int nobj = (int)objs.size();
  std::unordered_map<int, int> matidx;   // material object -> table index

  // findmaterial(): follow the modifier chain from modifier m to the first
  // material (for object oi, m is its own modifier)
  auto findmaterial_from = [&](int oi, int m) -> int {
    int guard = 0;
    while (m >= 0) {
      if (++guard > 10000) objerr(objs[oi], "modifier loop");
      const RObj& mo = objs[m];
      if (SceneDB::is_material(mo.type)) return m;
      if (mo.type == "alias") {
        if (!mo.sargs.empty()) {
          int r = db.modifier(mo.sargs[0]);
          if (r < 0) objerr(mo, "bad reference");
          m = r;
          continue;
        }
        m = mo.omod;
        continue;
      }
      // patterns, textures and mixtures are not part of the GPU material set
      objerr(mo, "modifier type \"" + mo.type + "\" is not supported on the GPU path");
    }
    return -1;
  };
  // the chain depends only on the modifier: remembered per modifier object
  // (a large scene has millions of surfaces and a few hundred modifiers)
  std::vector<int> mat_of_mod(nobj, INT32_MIN);
  auto findmaterial = [&](int oi) -> int {
    const int m = objs[oi].omod;
    if (m < 0) return findmaterial_from(oi, m);
    int& r = mat_of_mod[m];
    if (r == INT32_MIN) r = findmaterial_from(oi, m);
    return r;
  };

  std::function<int(int)> material_index = [&](int mo) -> int {
    if (mo < 0) return -1;
    auto it = matidx.find(mo);
    if (it != matidx.end()) return it->second;
    const RObj& m = objs[mo];
    int kind = material_kind(m.type);
    if (kind < 0) objerr(m, "material type \"" + m.type + "\" is not supported on the GPU path");
    // a material's own modifiers (patterns, textures: raytexture() and the
    // pattern functions of rt/p_func.c, rt/t_func.c) change its colour or
    // normal; the GPU material set has none, so such a material is refused
    // instead of being shaded as if unmodified
    for (int q = m.omod, guard = 0; q >= 0;) {
      if (++guard > 10000) objerr(m, "modifier loop");
      const RObj& pm = objs[q];
      if (pm.type == "alias" && !pm.sargs.empty()) {
        q = db.modifier(pm.sargs[0]);
        if (q < 0) objerr(pm, "bad reference");
        continue;
      }
      if (pm.type == "alias") { q = pm.omod; continue; }
      objerr(m, "modifier \"" + pm.name + "\" of type \"" + pm.type + "\" (a pattern or texture) is not supported on the GPU path");
    }
    Material M;
    memset(&M, 0, sizeof M);
    M.kind = kind;
    M.nfargs = (int)m.fargs.size();
    M.illum_alt = -1;
    int need = kind == MK_TRANS ? 7 : (kind == MK_PLASTIC || kind == MK_METAL) ? 5
             : kind == MK_GLOW ? 4 : kind == MK_GLASS ? -1 : 3;
    if (need > 0 && M.nfargs != need) objerr(m, "bad number of arguments");
    if (kind == MK_GLASS && M.nfargs != 3 && M.nfargs != 4) objerr(m, "bad arguments");
    for (int i = 0; i < M.nfargs && i < 8; ++i) M.f[i] = m.fargs[i];
    M.transp = (kind == MK_TRANS || kind == MK_GLASS) ? 1 : 0;
    int idx = (int)cs.mats.size();
    cs.mats.push_back(M);
    cs.mat_names.push_back(m.name);
    matidx[mo] = idx;
    if (kind == MK_ILLUM && !m.sargs.empty() && m.sargs[0] != "void") {
      int alt = db.modifier(m.sargs[0]);
      if (alt < 0) objerr(m, "unknown alternate material");
      int ai = material_index(alt);
      cs.mats[idx].illum_alt = ai;
    }
    return idx;
  };

  for (int i = 0; i < 3; ++i) { cs.bmin[i] = 1e300; cs.bmax[i] = -1e300; }
  auto grow = [&](const double* lo, const double* hi) {
    for (int i = 0; i < 3; ++i) {
      cs.bmin[i] = std::min(cs.bmin[i], lo[i]);
      cs.bmax[i] = std::max(cs.bmax[i], hi[i]);
    }
  };

  // the octree cube's box, grown as oconv's add2bbox does (ot/bbox.c:23-140):
  // polygon vertices, a sphere's six axis extremes, the end circles of the
  // cone family (only the outer circle of a ring)
  double cb_lo[3] = {1e300, 1e300, 1e300}, cb_hi[3] = {-1e300, -1e300, -1e300};
  auto cube_point = [&](const double* p) {
    for (int a = 0; a < 3; ++a) {
      cb_lo[a] = std::min(cb_lo[a], p[a]);
      cb_hi[a] = std::max(cb_hi[a], p[a]);
    }
  };
  auto cube_circle = [&](const double* c, const double* n, double rad) {
    for (int a = 0; a < 3; ++a) {
      double r = sqrt(1. - n[a] * n[a]);
      cb_hi[a] = std::max(cb_hi[a], c[a] + r * rad);
      cb_lo[a] = std::min(cb_lo[a], c[a] - r * rad);
    }
  };

  // pass 1: primitives (tables reserved from the argument counts: a
  // polygon's geometry is its 3 nv coordinates plus 5 plane values)
  {
    size_t nf = 0;
    for (const RObj& o : objs) nf += o.fargs.size() + 5;
    cs.prims.reserve((size_t)nobj);
    cs.geom.reserve(nf);
  }
  std::vector<int> obj_prim(nobj, -1);
  std::vector<std::array<double, 6>> dbounds;
  // One surface object -> one primitive (-1 if degenerate).  mo: material
  // object, omod: modifier that ranks coincident hits and names the
  // primitive for rcontrib; flat: isflat() of the object reported as hit.
  auto add_surface = [&](const RObj& o, int mo, int omod, bool flat) -> int {
    Prim P;
    memset(&P, 0, sizeof P);
    P.mat = material_index(mo);
    P.omod = omod;
    P.modslot = -1;
    P.srcidx = -1;
    P.goff = (int)cs.geom.size();
    P.nv = 0;
    P.flat = 0;
    double lo[3], hi[3];
    const auto& f = o.fargs;
    if (o.type == "polygon") {
      if (f.size() < 9 || f.size() % 3) objerr(o, "bad # arguments");
      int nv = (int)f.size() / 3;
      auto V = [&](int k) { return &f[3 * k]; };
      if (nv > 3) {
        double d2 = 0;
        for (int a = 0; a < 3; ++a) d2 += (V(0)[a] - V(nv - 1)[a]) * (V(0)[a] - V(nv - 1)[a]);
        if (d2 <= FTINY * FTINY) nv--;
      }
      double n[3] = {0, 0, 0}, v1[3], v2[3], v3[3];
      for (int a = 0; a < 3; ++a) v1[a] = V(1)[a] - V(0)[a];
      for (int k = 2; k < nv; ++k) {
        for (int a = 0; a < 3; ++a) v2[a] = V(k)[a] - V(0)[a];
        cross3(v3, v1, v2);
        for (int a = 0; a < 3; ++a) n[a] += v3[a];
        memcpy(v1, v2, sizeof v1);
      }
      double area = norm3(n);
      if (area == 0.0) {
        cs.warnings.push_back("polygon \"" + o.name + "\": zero area");
        return -1;
      }
      double offset = dot3(n, V(0));
      for (int k = 1; k < nv; ++k) offset += dot3(n, V(k));
      offset /= (double)nv;
      int ax = fabs(n[1]) > fabs(n[0]);
      if (fabs(n[2]) > fabs(n[ax])) ax = 2;
      P.kind = PK_POLYGON;
      P.nv = nv;
      P.flat = PF_FLAT;
      cs.geom.insert(cs.geom.end(), {n[0], n[1], n[2], offset, (double)ax});
      for (int a = 0; a < 3; ++a) { lo[a] = 1e300; hi[a] = -1e300; }
      for (int k = 0; k < nv; ++k) {
        for (int a = 0; a < 3; ++a) {
          cs.geom.push_back(V(k)[a]);
          lo[a] = std::min(lo[a], V(k)[a]);
          hi[a] = std::max(hi[a], V(k)[a]);
        }
        cube_point(V(k));
      }
    } else if (o.type == "sphere" || o.type == "bubble") {
      if (f.size() != 4) objerr(o, "bad # arguments");
      bool bubble = o.type == "bubble";
      double r = f[3];
      if (r < -FTINY) { bubble = !bubble; r = -r; cs.warnings.push_back(o.name + ": negative radius"); }
      else if (r <= FTINY) objerr(o, "zero radius");
      P.kind = bubble ? PK_BUBBLE : PK_SPHERE;
      cs.geom.insert(cs.geom.end(), {f[0], f[1], f[2], r});
      for (int a = 0; a < 3; ++a) { lo[a] = f[a] - r; hi[a] = f[a] + r; }
      for (int a = 0; a < 3; ++a) {
        double v[3] = {f[0], f[1], f[2]};
        v[a] -= r;
        cube_point(v);
        v[a] += 2.0 * r;
        cube_point(v);
      }
    } else {
      // cone family (getcone)
      bool ring = o.type == "ring";
      bool cylfam = o.type == "cylinder" || o.type == "tube";
      std::string t = o.type;
      double P0[3], P1[3], R0, R1, ad[3];
      if (cylfam) {
        if (f.size() != 7) objerr(o, "bad # arguments");
        double r = f[6];
        if (r < -FTINY) { t = (t == "cylinder") ? "tube" : "cylinder"; r = -r; }
        else if (r <= FTINY) { cs.warnings.push_back(o.name + ": illegal radii"); return -1; }
        for (int a = 0; a < 3; ++a) { P0[a] = f[a]; P1[a] = f[3 + a]; }
        R0 = R1 = r;
      } else {
        if (f.size() != 8) objerr(o, "bad # arguments");
        double c6 = f[6], c7 = f[7];
        int s0 = c6 < -FTINY ? -1 : (c6 > FTINY ? 1 : 0);
        int s1 = c7 < -FTINY ? -1 : (c7 > FTINY ? 1 : 0);
        if (s0 + s1 == 0) { cs.warnings.push_back(o.name + ": illegal radii"); return -1; }
        if ((s0 < 0) | (s1 < 0)) {
          if (ring) { cs.warnings.push_back(o.name + ": illegal radii"); return -1; }
          t = (t == "cone") ? "cup" : "cone";
        }
        c6 *= s0;
        c7 *= s1;
        if (c7 - c6 > FTINY) {
          for (int a = 0; a < 3; ++a) { P0[a] = f[a]; P1[a] = f[3 + a]; }
          R0 = c6; R1 = c7;
        } else if (c6 - c7 > FTINY) {
          for (int a = 0; a < 3; ++a) { P0[a] = ring ? f[a] : f[3 + a]; P1[a] = ring ? f[3 + a] : f[a]; }
          R0 = c7; R1 = c6;
        } else {
          if (ring) { cs.warnings.push_back(o.name + ": illegal radii"); return -1; }
          t = (t == "cone") ? "cylinder" : "tube";
          for (int a = 0; a < 3; ++a) { P0[a] = f[a]; P1[a] = f[3 + a]; }
          R0 = R1 = c6;
          cylfam = true;
        }
      }
      double al;
      if (ring) {
        for (int a = 0; a < 3; ++a) ad[a] = f[3 + a];
        al = norm3(ad);
      } else {
        for (int a = 0; a < 3; ++a) ad[a] = P1[a] - P0[a];
        al = norm3(ad);
      }
      if (al == 0.0) { cs.warnings.push_back(o.name + ": unknown orientation"); return -1; }
      if (ring) {
        P.kind = PK_RING;
        P.flat = PF_FLAT;
        cs.geom.insert(cs.geom.end(), {f[0], f[1], f[2], ad[0], ad[1], ad[2], R0, R1});
        for (int a = 0; a < 3; ++a) { lo[a] = f[a] - R1; hi[a] = f[a] + R1; }
        cube_circle(f.data(), ad, R1);
      } else {
        cube_circle(P0, ad, R0);
        cube_circle(P1, ad, R1);
        P.kind = t == "cone" ? PK_CONE : t == "cup" ? PK_CUP : t == "cylinder" ? PK_CYLINDER : PK_TUBE;
        double sl = cylfam ? al : sqrt((R1 - R0) * (R1 - R0) + al * al);
        cs.geom.insert(cs.geom.end(), {P0[0], P0[1], P0[2], ad[0], ad[1], ad[2], al, R0, R1, sl});
        double rr = std::max(R0, R1);
        for (int a = 0; a < 3; ++a) {
          lo[a] = std::min(P0[a], P1[a]) - rr;
          hi[a] = std::max(P0[a], P1[a]) + rr;
        }
      }
    }
    if (!flat) P.flat = 0;
    int idx = (int)cs.prims.size();
    cs.prims.push_back(P);
    cs.prim_modname.push_back(objs[omod].name);
    dbounds.push_back({lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]});
    grow(lo, hi);
    return idx;
  };
  // mesh objects: one triangle primitive per mesh triangle, in world space
  // (the reference transforms the ray into the mesh instead; o_mesh.c:146-217)
  // parent: transform of the enclosing instances (nullptr at top level);
  // ovr: the instance object whose modifier replaces every material (-1 none)
  auto add_mesh = [&](int oi, const Xf* parent, int ovr) {
    const RObj& o = objs[oi];
    auto itm = db.obj_mesh.find(oi);
    if (itm == db.obj_mesh.end()) objerr(o, "mesh file not loaded");
    const MeshData& M = db.meshes[itm->second];
    Xf X;
    if (parse_xf(X, o.sargs, 1) != (int)o.sargs.size() - 1) objerr(o, "bad transform");
    if (parent) compose_xf(X, *parent);
    const bool mirror = X.sca < 0;
    const size_t nv = M.has_n.size();
    std::vector<double> W(3 * nv), WN(3 * nv);
    for (size_t v = 0; v < nv; ++v) {
      xf_point(&W[3 * v], &M.xyz[3 * v], X);
      xf_vector(&WN[3 * v], &M.nrm[3 * v], X);
    }
    int objmat = -2;
    for (const auto& t : M.tris) {
      // the triangle's own material when the mesh's modifier is void
      const bool pseudo = ovr < 0 && o.omod < 0 && t.mat >= 0;
      const int mod = ovr >= 0 ? objs[ovr].omod : pseudo ? t.mat : o.omod;
      if (mod < 0) continue;                         // void: never shaded
      int mo;
      if (pseudo) {
        mo = findmaterial_from(oi, mod);
      } else {
        if (objmat == -2) objmat = findmaterial(ovr >= 0 ? ovr : oi);
        mo = objmat;
      }
      if (mo < 0) continue;
      const double *a = &W[3 * t.v[0]], *b = &W[3 * t.v[1]], *c = &W[3 * t.v[2]];
      double e1[3], e2[3], n[3];
      for (int k = 0; k < 3; ++k) { e1[k] = a[k] - c[k]; e2[k] = b[k] - a[k]; }
      cross3(n, e1, e2);
      if (norm3(n) == 0.0) continue;                 // degenerate: always tangent
      if (mirror) for (int k = 0; k < 3; ++k) n[k] = -n[k];
      Prim P;
      memset(&P, 0, sizeof P);
      P.kind = PK_TRIANGLE;
      P.mat = material_index(mo);
      P.omod = mod;
      P.modslot = -1;
      P.srcidx = -1;
      P.goff = (int)cs.geom.size();
      P.nv = 3;
      P.flat = (pseudo ? PF_FLAT : 0) | PF_COMPLEX;   // the pseudo object is a face
      const bool normals = M.has_n[t.v[0]] && M.has_n[t.v[1]] && M.has_n[t.v[2]];
      int fl = (t.v[0] > t.v[1] ? TRI_REV0 : 0) | (t.v[1] > t.v[2] ? TRI_REV1 : 0) |
               (t.v[2] > t.v[0] ? TRI_REV2 : 0) | (normals ? TRI_NORMALS : 0) |
               (objs[mod].name == "Phong" ? TRI_PHONG : 0);
      cs.geom.insert(cs.geom.end(), {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], n[0], n[1], n[2],
                                     (double)fl});
      if (normals)
        for (int k = 0; k < 3; ++k) {
          const double* vn = &WN[3 * t.v[k]];
          cs.geom.insert(cs.geom.end(), {vn[0], vn[1], vn[2]});
        }
      double lo[3], hi[3];
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min({a[k], b[k], c[k]});
        hi[k] = std::max({a[k], b[k], c[k]});
      }
      cube_point(a);
      cube_point(b);
      cube_point(c);
      cs.prims.push_back(P);
      cs.prim_modname.push_back(objs[mod].name);
      dbounds.push_back({lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]});
      grow(lo, hi);
    }
  };
  // instance objects: the instanced octree's surfaces, transformed into
  // place (the reference transforms the ray instead, rt/o_instance.c:16-71);
  // a non-void instance modifier replaces every material inside
  std::function<void(int, const Xf*, int, int)> add_instance = [&](int oi, const Xf* parent, int ovr, int depth) {
    const RObj& io = objs[oi];
    if (depth > 32) objerr(io, "instances nested too deeply");
    auto itb = db.obj_inst.find(oi);
    if (itb == db.obj_inst.end()) objerr(io, "octree not loaded");
    const SceneDB::Block blk = db.inst_blocks[itb->second];
    Xf X;
    if (parse_xf(X, io.sargs, 1) != (int)io.sargs.size() - 1) objerr(io, "bad transform");
    if (parent) compose_xf(X, *parent);
    const int over = ovr >= 0 ? ovr : (io.omod >= 0 ? oi : -1);
    const int overmat = over >= 0 ? findmaterial(over) : -1;
    for (int j = blk.first; j < blk.first + blk.count; ++j) {
      const RObj& o = objs[j];
      if (o.type == "mesh") {
        add_mesh(j, &X, over);
      } else if (o.type == "instance") {
        add_instance(j, &X, over, depth + 1);
      } else if (SceneDB::is_surface(o.type) && o.type != "source") {
        int pi = -1;
        if (over >= 0) {
          if (overmat >= 0) pi = add_surface(xf_object(o, X), overmat, objs[over].omod, false);
        } else if (o.omod >= 0) {
          int mo = findmaterial(j);
          if (mo >= 0) pi = add_surface(xf_object(o, X), mo, o.omod, true);
        }
        if (pi >= 0) cs.prims[pi].flat |= PF_COMPLEX;
      }
    }
  };
  for (int oi = 0; oi < nobj; ++oi) {
    const RObj& o = objs[oi];
    if (!db.hidden.empty() && db.hidden[oi]) continue;   // inside an instanced octree
    if (o.type == "mesh") {
      if (o.omod >= 0) {
        int mo = findmaterial(oi);
        if (mo >= 0) {
          const std::string& mt = objs[mo].type;
          if (mt == "light" || mt == "glow" || mt == "illum" || mt == "spotlight") objerr(o, "illegal material");
        }
      }
      add_mesh(oi, nullptr, -1);
      continue;
    }
    if (o.type == "instance") {
      add_instance(oi, nullptr, -1, 0);
      continue;
    }
    if (!SceneDB::is_surface(o.type) || o.type == "source") continue;
    if (o.omod < 0) continue;   // no modifier: never shaded
    int mo = findmaterial(oi);
    obj_prim[oi] = add_surface(o, mo, o.omod, true);
  }
  if (cs.prims.empty()) {
    for (int i = 0; i < 3; ++i) { cs.bmin[i] = 0; cs.bmax[i] = 1; }
  }
  double ext = std::max({cs.bmax[0] - cs.bmin[0], cs.bmax[1] - cs.bmin[1], cs.bmax[2] - cs.bmin[2], 1e-3});
  cs.cusize = ext;
  // the octree cube: the file's for octree input, else as oconv sets it
  // (ot/oconv.c:126-139: box widened by 10 FTINY, largest side, centred)
  if (db.cusize > 0) {
    for (int a = 0; a < 3; ++a) cs.cuorg[a] = db.cuorg[a];
    cs.cube_size = db.cusize;
  } else if (cb_lo[0] <= cb_hi[0]) {
    cs.cube_size = 0;
    for (int a = 0; a < 3; ++a) {
      cb_lo[a] -= 10 * FTINY;
      cb_hi[a] += 10 * FTINY;
      cs.cube_size = std::max(cs.cube_size, cb_hi[a] - cb_lo[a]);
    }
    for (int a = 0; a < 3; ++a) cs.cuorg[a] = (cb_hi[a] + cb_lo[a] - cs.cube_size) * .5;
    // the reference reads the cube back from the octree oconv wrote, where
    // it is text of 12 significant digits (ot/writeoct.c:44-47)
    char buf[40];
    for (int a = 0; a < 3; ++a) {
      snprintf(buf, sizeof buf, "%.12g", cs.cuorg[a]);
      cs.cuorg[a] = strtod(buf, nullptr);
    }
    snprintf(buf, sizeof buf, "%.12g", cs.cube_size);
    cs.cube_size = strtod(buf, nullptr);
  }
  // FP32 bounds, widened so that the FP32 slab test stays conservative
  double eps = 1e-5 + 4e-6 * ext;
  cs.lo.resize(3 * cs.prims.size());
  cs.hi.resize(3 * cs.prims.size());
  for (size_t k = 0; k < cs.prims.size(); ++k)
    for (int a = 0; a < 3; ++a) {
      float l = (float)(dbounds[k][a] - eps), h = (float)(dbounds[k][3 + a] + eps);
      cs.lo[3 * k + a] = nextafterf(l, -INFINITY);
      cs.hi[3 * k + a] = nextafterf(h, INFINITY);
    }

  // pass 2: light sources (marksources)
  for (int oi = 0; oi < nobj; ++oi) {
    const RObj& o = objs[oi];
    if (!SceneDB::is_surface(o.type) || o.omod < 0) continue;
    if (!db.hidden.empty() && db.hidden[oi]) continue;   // sources inside instances are not marked
    int mo = findmaterial(oi);
    if (mo < 0) continue;
    const RObj& m = objs[mo];
    if (m.type != "light" && m.type != "glow" && m.type != "illum" && m.type != "spotlight") continue;
    if (m.type == "spotlight") objerr(m, "spotlight sources are not supported on the GPU path");
    size_t need = m.type == "glow" ? 4 : 3;
    if (m.fargs.size() != need) objerr(m, "bad # arguments");
    if (m.fargs[0] <= FTINY && (m.fargs[1] <= FTINY) & (m.fargs[2] <= FTINY)) continue;
    if (m.type == "glow" && o.type != "source" && m.fargs[3] <= FTINY) continue;
    Source S;
    memset(&S, 0, sizeof S);
    S.prim = -1;
    S.mat = material_index(mo);
    S.objmod = o.omod;
    S.modslot = -1;
    const auto& f = o.fargs;
    auto setflatss = [&]() {
      double u[3];
      if (!perpendicular3(u, S.snorm)) objerr(o, "bad source orientation");
      double mult = .5 * sqrt(S.ss2);
      for (int a = 0; a < 3; ++a) S.ss[0][a] = u[a] * mult;
      cross3(S.ss[1], S.snorm, S.ss[0]);
    };
    if (o.type == "source") {
      if (f.size() != 4) objerr(o, "bad arguments");
      S.flags |= SDISTANT | SCIR;
      for (int a = 0; a < 3; ++a) S.sloc[a] = f[a];
      if (norm3(S.sloc) == 0.0) objerr(o, "zero direction");
      double theta = PI_D / 180.0 / 2.0 * f[3];
      if (theta <= FTINY) objerr(o, "zero size");
      S.ss2 = 2.0 * PI_D * (1.0 - cos(theta));
      S.srad = sqrt(S.ss2 / PI_D);
      memcpy(S.snorm, S.sloc, sizeof S.snorm);
      setflatss();
      S.ss[2][0] = S.ss[2][1] = S.ss[2][2] = 0.0;
    } else {
      int pi = obj_prim[oi];
      if (pi < 0) continue;
      const Prim& P = cs.prims[pi];
      const double* g = cs.geom.data() + P.goff;
      S.prim = pi;
      if (P.kind == PK_POLYGON) {
        int nv = P.nv;
        const double* va = g + 5;
        for (int a = 0; a < 3; ++a) {
          S.sloc[a] = 0;
          for (int k = 0; k < nv; ++k) S.sloc[a] += va[3 * k + a];
          S.sloc[a] /= (double)nv;
        }
        S.flags |= SFLAT;
        for (int a = 0; a < 3; ++a) S.snorm[a] = g[a];
        // area = half the fan cross-product length
        double n[3] = {0, 0, 0}, v1[3], v2[3], v3[3];
        for (int a = 0; a < 3; ++a) v1[a] = va[3 + a] - va[a];
        for (int k = 2; k < nv; ++k) {
          for (int a = 0; a < 3; ++a) v2[a] = va[3 * k + a] - va[a];
          cross3(v3, v1, v2);
          for (int a = 0; a < 3; ++a) n[a] += v3[a];
          memcpy(v1, v2, sizeof v1);
        }
        S.ss2 = 0.5 * sqrt(dot3(n, n));
        S.srad = 0;
        for (int k = 0; k < nv; ++k) {
          double d2 = 0;
          for (int a = 0; a < 3; ++a) d2 += (va[3 * k + a] - S.sloc[a]) * (va[3 * k + a] - S.sloc[a]);
          S.srad = std::max(S.srad, d2);
        }
        S.srad = sqrt(S.srad);
        if (nv == 4) {
          for (int a = 0; a < 3; ++a) {
            S.ss[0][a] = .5 * (va[3 + a] - va[a]);
            S.ss[1][a] = .5 * (va[9 + a] - va[a]);
          }
        } else if (nv == 3) {
          auto dist2line = [&](const double* p, const double* p1, const double* p2) {
            double d[3], c[3], e[3];
            for (int a = 0; a < 3; ++a) { d[a] = p2[a] - p1[a]; e[a] = p[a] - p1[a]; }
            cross3(c, d, e);
            return dot3(c, c) / dot3(d, d);
          };
          int near0 = 2;
          double dmin = dist2line(S.sloc, va + 6, va);
          for (int k = 0; k < 2; ++k) {
            double d2 = dist2line(S.sloc, va + 3 * k, va + 3 * (k + 1));
            if (d2 >= dmin) continue;
            near0 = k;
            dmin = d2;
          }
          int k1 = (near0 + 1) % 3;
          for (int a = 0; a < 3; ++a) S.ss[0][a] = va[3 * k1 + a] - va[3 * near0 + a];
          norm3(S.ss[0]);
          dmin = sqrt(dmin);
          for (int a = 0; a < 3; ++a) S.ss[0][a] *= dmin;
          cross3(S.ss[1], S.snorm, S.ss[0]);
        } else {
          setflatss();
        }
      } else if (P.kind == PK_SPHERE || P.kind == PK_BUBBLE) {
        S.flags |= SCIR;
        for (int a = 0; a < 3; ++a) S.sloc[a] = g[a];
        S.srad = g[3];
        S.ss2 = PI_D * S.srad * S.srad;
        for (int a = 0; a < 3; ++a) S.ss[a][a] = 0.7236 * g[3];
      } else if (P.kind == PK_RING) {
        if (g[6] > 0.0) objerr(o, "cannot hit source center");
        for (int a = 0; a < 3; ++a) { S.sloc[a] = g[a]; S.snorm[a] = g[3 + a]; }
        S.flags |= SFLAT | SCIR;
        S.srad = g[7];
        S.ss2 = PI_D * S.srad * S.srad;
        setflatss();
      } else {
        objerr(o, "illegal material (this surface type cannot be a light source on the GPU path)");
      }
      cs.prims[pi].srcidx = (int)cs.srcs.size();
    }
    if (m.type == "glow") {
      S.flags |= SPROX;
      S.prox = m.fargs[3];
      if (S.flags & SDISTANT) S.flags |= SSKIP;
    }
    cs.srcs.push_back(S);
    cs.src_modname.push_back(objs[o.omod].name);
  }
  if (cs.srcs.size() > 32000) throw RadError("too many light sources for the GPU path");
  shadow_cache_setup(cs);

  // built-in Lambertian used by -I / -i (ρ = π, rt/raytrace.c:24-28)
  Material L;
  memset(&L, 0, sizeof L);
  L.kind = MK_PLASTIC;
  L.nfargs = 5;
  L.f[0] = L.f[1] = L.f[2] = PI_D;
  L.illum_alt = -1;
  cs.lamb = (int)cs.mats.size();
  cs.mats.push_back(L);
  cs.mat_names.push_back("");          // not addressable by name
  return cs;
}

}  // namespace rg
