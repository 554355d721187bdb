// Host-side check of scene input and compilation (scene_io.cpp, rg_scene.cpp)
// together with the primitive tests of rg_geom.hpp, which are host/device
// code: scenes are loaded and compiled as the library does, then rays are
// intersected by brute force over every primitive, so meshes and transforms
// can be checked against the reference's rtrace without a GPU.
#include <cstdio>
#include <cstring>
#include <string>

#include "rg_geom.hpp"
#include "rg_scene.hpp"
#include "scene_io.hpp"

using namespace rg;

namespace {
CompiledScene g_cs;

int fail(char* err, int errlen, const std::string& m) {
  snprintf(err, (size_t)errlen, "%s", m.c_str());
  return -1;
}
}  // namespace

extern "C" {

// Reads scene text as the library does and copies the real arguments of
// object `obj` (at most cap); returns their count, or -1 with the message.
int scene_text_fargs(const char* text, int obj, double* out, int cap, char* err, int errlen) {
  try {
    SceneDB db;
    db.load_rad_text(text, "scene text");
    if (obj < 0 || obj >= (int)db.objs.size()) return fail(err, errlen, "no such object");
    const auto& f = db.objs[(size_t)obj].fargs;
    for (size_t i = 0; i < f.size() && (int)i < cap; ++i) out[i] = f[i];
    return (int)f.size();
  } catch (const std::exception& e) {
    return fail(err, errlen, e.what());
  }
}


// Loads the files in order and compiles them; returns the primitive count.
int scene_compile(const char** files, int nfiles, char* err, int errlen) {
  try {
    SceneDB db;
    for (int i = 0; i < nfiles; ++i) db.load_any(files[i]);
    db.resolve_instances();
    db.resolve_meshes();
    g_cs = compile_scene(db);
    return (int)g_cs.prims.size();
  } catch (const std::exception& e) {
    return fail(err, errlen, e.what());
  }
}

// Per primitive: kind, material index, flat flag, geometry offset.
void scene_prims(int* out) {
  for (size_t i = 0; i < g_cs.prims.size(); ++i) {
    const Prim& P = g_cs.prims[i];
    out[4 * i] = P.kind;
    out[4 * i + 1] = P.mat;
    out[4 * i + 2] = P.flat & PF_FLAT;
    out[4 * i + 3] = P.goff;
  }
}

int scene_geom(double* out, int cap) {
  int n = (int)g_cs.geom.size();
  if (n <= cap) memcpy(out, g_cs.geom.data(), sizeof(double) * (size_t)n);
  return n;
}

// Name of material index m.
const char* scene_mat_name(int m) { return g_cs.mat_names[(size_t)m].c_str(); }

// Closest hit by brute force.  Per ray: t (0 on a miss), unperturbed normal,
// perturbed normal (raynormal(), rt/raytrace.c:446-480), primitive.
void scene_trace(const double* rays, int n, double* out) {
  DevScene S;
  memset(&S, 0, sizeof S);
  S.prims = g_cs.prims.data();
  S.geom = g_cs.geom.data();
  S.mats = g_cs.mats.data();
  S.nprims = (int)g_cs.prims.size();
  for (int i = 0; i < n; ++i) {
    V3 o = v3(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]);
    V3 d = v3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
    normalize(d);
    HitRec h;
    h.t = FHUGE;
    h.rod = 1.0;
    h.prim = -1;
    for (int p = 0; p < S.nprims; ++p) hit_prim(S, p, o, d, h);
    double* r = out + 8 * i;
    memset(r, 0, sizeof(double) * 8);
    r[7] = h.prim;
    if (h.prim < 0) continue;
    V3 p = o + d * h.t;
    V3 ron = prim_normal(S, h.prim, p);
    double rod = -dot(d, ron);
    V3 pert = prim_pert(S, h.prim, p, ron);
    V3 pn = ron + pert;
    double l2 = dot(pn, pn);
    if (l2 == 0.0) {
      pn = ron;
    } else {
      pn = pn * (1.0 / sqrt(l2));
      double nd = -dot(pn, d);
      if ((nd > 0.0) != (rod > 0.0)) pn = pn + d * (2.0 * nd);
    }
    r[0] = h.t;
    r[1] = ron.x; r[2] = ron.y; r[3] = ron.z;
    r[4] = pn.x; r[5] = pn.y; r[6] = pn.z;
  }
}
}
