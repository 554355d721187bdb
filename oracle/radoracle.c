/*
 * radoracle.c - CPU restatement of the reference rtrace / rcontrib ray path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU product.  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
 *
 * It restates, function by function (citations are to the reference tree):
 *   rayorigin        rt/raytrace.c:40-135      Russian roulette, weights
 *   raytrace         rt/raytrace.c:163-180     localhit / sourcehit / trace
 *   raytrans         rt/raytrace.c:194-207
 *   rayreject        rt/raytrace.c:536-575     coincident-hit tie rules
 *   raycontrib       rt/raytrace.c:408-442     product over the parent chain
 *   o_face + inface  rt/o_face.c:16-62, common/face.c:121-162
 *   getface          common/face.c:35-117
 *   o_sphere         rt/sphere.c:16-83, quadratic common/zeroes.c:19-55
 *   o_cone, getcone  rt/o_cone.c:17-146, common/cone.c:20-120
 *   m_normal         rt/normal.c:177-360, dirnorm :72-173, gaussamp :364-495
 *   m_glass          rt/glass.c:46-165
 *   m_light + macros rt/source.c:683-793
 *   direct           rt/source.c:398-557 (every source tested: -dt 0)
 *   srcray/nextssamp rt/source.c:219-257, rt/srcsamp.c:36-144 (one partition)
 *   shadow cache     rt/srcobstr.c:28-339 (castshadow, initobscache,
 *                    srcobstructp, srcblocker, srcblocked; SHADCACHE 20 of
 *                    rt/source.h:19-20), with the cube of ot/oconv.c:120-139
 *                    and ot/bbox.c:23-140; it decides which shadow rays are
 *                    traced (and so the ray count), never a value
 *   sourcehit        rt/source.c:317-377
 *   marksources      rt/source.c:47-142, ssetsrc etc. rt/srcsupp.c:60-260
 *   multambient      rt/ambient.c:229-330 (-aa 0 branch)
 *   samp_hemi        rt/ambcomp.c:351-424, ambsample :177-250
 *   square2disk      common/disk2square.c:44-86
 *   trace_contrib    rt/rcontrib.c:272-316, eval_irrad :321-338
 *   rayirrad         rt/rtrace.c:416-431
 *   rbin             util/reinhartb.cal:19-47; kbin util/klems_full.cal:10-43
 *   scbin, cbin      util/disk2square.cal:34-90, util/cieskyscan.cal:21-53
 *
 * Deliberate differences from the reference, shared with the GPU path:
 *   - random numbers come from a counter-based generator keyed by the ray's
 *     path address (sensor, then child slot per spawn) instead of drand48 /
 *     urand streams; the ambient collision trade-in (ambcomp.c:81-175) is
 *     not reproduced, so each stratum gets one uniform jittered sample;
 *   - tangent frames are a fixed perpendicular rotated by a random angle;
 *   - all sources with non-zero potential are shadow-tested (-dt 0);
 *   - local sources are sampled as one partition (-ds 0).
 */
#include "radoracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define FTINY 1e-6
#define FHUGE 1e10
#define PI 3.14159265358979323846

/* ray types, rt/ray.h:27-44 */
#define PRIMARY 01
#define RSHADOW 02
#define REFLECTED 04
#define TRANS 020
#define RAMBIENT 040
#define RSPECULAR 0100
#define TSHADOW 0200
#define TAMBIENT 0400
#define TSPECULAR 01000
#define SHADOW (RSHADOW | TSHADOW)
#define AMBIENT (RAMBIENT | TAMBIENT)
#define SPECULAR (RSPECULAR | TSPECULAR)
#define RAYREFL (RSHADOW | REFLECTED | RAMBIENT | RSPECULAR)

/* source flags, rt/source.h */
#define SDISTANT 01
#define SSKIP 02
#define SPROX 04
#define SFLAT 040
#define SCYL 0100
#define SFOLLOW 0200
#define SCIR 01000

enum { T_POLY, T_SPHERE, T_BUBBLE, T_RING, T_CONE, T_CUP, T_CYL, T_TUBE, T_SOURCE,
       M_PLASTIC, M_METAL, M_TRANS, M_GLASS, M_LIGHT, M_GLOW, M_ILLUM, T_OTHER };

typedef struct {
  int type, mod, mat;     /* type, modifier object, material object (-1 none) */
  char name[128];
  int nf;
  double f[16];
  double *va;             /* polygon vertices (nv*3) */
  int nv, ax, alt;        /* polygon; illum alternate material object */
  double norm[3], offset; /* polygon plane */
  double p0[3], ad[3], al, r0, r1, sl;  /* cone family, ring: p0 = centre */
  int ok;                 /* usable surface */
  int src;                /* source index, -1 */
} OObj;

typedef struct {
  double sloc[3], snorm[3], ss[3][3], ss2, srad, prox;
  int flags, obj, mat;
} OSrc;

struct ro_scene {
  OObj *o;
  int n, cap;
  OSrc *s;
  int ns;
  double *vpool;
  double cuorg[3], cusize;   /* the octree's cube as oconv sets it */
};

/* ---- counter-based random numbers (same generator as the GPU path) ---- */
static unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
unsigned long long ro_sensor_key(unsigned long long seed, unsigned long long sensor) {
  return mix64(seed * 0x9e3779b97f4a7c15ULL ^ mix64(sensor + 1));
}
unsigned long long ro_child_key(unsigned long long k, unsigned long long slot) {
  return mix64(k ^ mix64(slot + 0x9e3779b97f4a7c15ULL));
}
double ro_rnd(unsigned long long k, unsigned int i) {
  return (double)(mix64(k + 0x632be59bd9b4e019ULL * (unsigned long long)(i + 1)) >> 11) *
         (1.0 / 9007199254740992.0);
}
#define SLOT_TRANS 1ULL
#define SLOT_REFL 2ULL
#define SLOT_RSPEC 3ULL
#define SLOT_TSPEC 4ULL
#define SLOT_PASS 5ULL
#define SLOT_SHADOW 64ULL
#define SLOT_HEMI_FRONT (1ULL << 20)
#define SLOT_HEMI_BACK (2ULL << 20)
#define RN_RR 0
#define RN_AMB0 1
#define RN_AMB1 2
#define RN_SRC0 3
#define RN_FRAME_FRONT 10
#define RN_FRAME_BACK 11
#define RN_GFRAME 12
#define RN_GAUSS 16

/* ---- vectors ---- */
#define DOT(a, b) ((a)[0] * (b)[0] + (a)[1] * (b)[1] + (a)[2] * (b)[2])
static void vcross(double *r, const double *a, const double *b) {
  double t[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  r[0] = t[0]; r[1] = t[1]; r[2] = t[2];
}
/* normalize() of common/fvect.c:132-155, with its first-order reciprocal
 * within FTINY of unit length, in the operation order of the reference's
 * -O3 -ffast-math build: |v|^2 = (x^2 + z^2) + y^2, 2 - (|v|^2 + 1) / 2 */
static double vnorm(double *v) {
  double l2 = (v[0] * v[0] + v[2] * v[2]) + v[1] * v[1], l, s;
  if (l2 == 0.0) return 0;
  if ((l2 <= 1.0 + FTINY) & (l2 >= 1.0 - FTINY)) {
    l = (l2 + 1.0) * 0.5;
    s = 2.0 - l;
  } else {
    l = sqrt(l2);
    s = 1.0 / l;
  }
  v[0] *= s; v[1] *= s; v[2] *= s;
  return l;
}
static double cmaxd(const double *c) { return c[0] > c[1] ? (c[0] > c[2] ? c[0] : c[2]) : (c[1] > c[2] ? c[1] : c[2]); }
static double cmeand(const double *c) { return (c[0] + c[1] + c[2]) / 3.0; }
static double cbright(const double *c) { return .265074126 * c[0] + .670114631 * c[1] + .064811243 * c[2]; }

/* getperpendicular(randomize=0): common/fvect.c:159-192 */
static int perp(double *vp, const double *v) {
  int i;
  double v1[3] = {0, 0, 0};
  for (i = 3; i--;)
    if ((-0.6 < v[i]) & (v[i] < 0.6)) break;
  if (i < 0) i = 0;
  v1[i] = 1.0;
  vcross(vp, v1, v);
  return vnorm(vp) > 0.0;
}
static void frame(const double *n, double u, double *ux, double *uy) {
  double a[3], b[3], ang = 2.0 * PI * u, c = cos(ang), s = sin(ang);
  int i;
  perp(a, n);
  vcross(b, n, a);
  for (i = 0; i < 3; i++) ux[i] = a[i] * c + b[i] * s;
  vnorm(ux);
  vcross(uy, n, ux);
}
static void square2disk(double ds[2], double seedx, double seedy) {
  double phi, r, a = 2. * seedx - 1, b = 2. * seedy - 1;
  if (a > -b) {
    if (a > b) { r = a; phi = (PI / 4.) * (b / a); }
    else { r = b; phi = (PI / 4.) * (2. - (a / b)); }
  } else {
    if (a < b) { r = -a; phi = (PI / 4.) * (4. + (b / a)); }
    else { r = -b; phi = b != 0. ? (PI / 4.) * (6. - (a / b)) : 0.; }
  }
  r *= 0.9999999999999;
  ds[0] = r * cos(phi);
  ds[1] = r * sin(phi);
}

/* ---- scene parsing ---- */
static int otype_of(const char *t) {
  static const char *names[] = {"polygon", "sphere", "bubble", "ring", "cone", "cup", "cylinder", "tube",
                                "source", "plastic", "metal", "trans", "glass", "light", "glow", "illum"};
  int i;
  for (i = 0; i < 16; i++)
    if (!strcmp(t, names[i])) return i;
  return T_OTHER;
}
static int is_surface(int t) { return t <= T_SOURCE; }
static int is_material(int t) { return t >= M_PLASTIC && t <= M_ILLUM; }

static const char *next_word(const char **pp, char *w, int wl) {
  const char *p = *pp;
  int n = 0;
  for (;;) {
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
    if (*p == '#') { while (*p && *p != '\n') p++; continue; }
    break;
  }
  if (!*p) return NULL;
  while (*p && !(*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) {
    if (n < wl - 1) w[n++] = *p;
    p++;
  }
  w[n] = 0;
  *pp = p;
  return w;
}

static int findmod(const ro_scene *s, const char *name, int before) {
  int i;
  for (i = before - 1; i >= 0; i--)
    if (!is_surface(s->o[i].type) && !strcmp(s->o[i].name, name)) return i;
  return -1;
}

#define PERR(...) do { snprintf(err, errlen, __VA_ARGS__); goto fail; } while (0)

static int setup_face(OObj *o) {   /* getface() */
  int i, k;
  double v1[3], v2[3], v3[3];
  o->nv = o->nf / 3;
  if (o->nv > 3) {
    double d2 = 0;
    for (k = 0; k < 3; k++) d2 += (o->va[k] - o->va[3 * (o->nv - 1) + k]) * (o->va[k] - o->va[3 * (o->nv - 1) + k]);
    if (d2 <= FTINY * FTINY) o->nv--;
  }
  o->norm[0] = o->norm[1] = o->norm[2] = 0;
  for (k = 0; k < 3; k++) v1[k] = o->va[3 + k] - o->va[k];
  for (i = 2; i < o->nv; i++) {
    for (k = 0; k < 3; k++) v2[k] = o->va[3 * i + k] - o->va[k];
    vcross(v3, v1, v2);
    for (k = 0; k < 3; k++) o->norm[k] += v3[k];
    memcpy(v1, v2, sizeof v1);
  }
  if (vnorm(o->norm) == 0.0) return 0;
  o->offset = DOT(o->norm, o->va);
  for (i = 1; i < o->nv; i++) o->offset += DOT(o->norm, o->va + 3 * i);
  o->offset /= (double)o->nv;
  o->ax = fabs(o->norm[1]) > fabs(o->norm[0]);
  if (fabs(o->norm[2]) > fabs(o->norm[o->ax])) o->ax = 2;
  return 1;
}

static int setup_cone(OObj *o) {   /* getcone() */
  int k;
  double *ca = o->f;
  if (o->type == T_CYL || o->type == T_TUBE) {
    if (o->nf != 7) return 0;
    if (ca[6] < -FTINY) { o->type = o->type == T_CYL ? T_TUBE : T_CYL; ca[6] = -ca[6]; }
    else if (ca[6] <= FTINY) return 0;
    for (k = 0; k < 3; k++) { o->p0[k] = ca[k]; o->ad[k] = ca[3 + k] - ca[k]; }
    o->r0 = o->r1 = ca[6];
  } else {
    int s0, s1;
    if (o->nf != 8) return 0;
    s0 = ca[6] < -FTINY ? -1 : ca[6] > FTINY;
    s1 = ca[7] < -FTINY ? -1 : ca[7] > FTINY;
    if (s0 + s1 == 0) return 0;
    if ((s0 < 0) | (s1 < 0)) {
      if (o->type == T_RING) return 0;
      o->type = o->type == T_CONE ? T_CUP : T_CONE;
    }
    ca[6] *= s0;
    ca[7] *= s1;
    if (ca[7] - ca[6] > FTINY) {
      for (k = 0; k < 3; k++) { o->p0[k] = ca[k]; o->ad[k] = ca[3 + k] - ca[k]; }
      o->r0 = ca[6]; o->r1 = ca[7];
    } else if (ca[6] - ca[7] > FTINY) {
      if (o->type == T_RING) {
        for (k = 0; k < 3; k++) o->p0[k] = ca[k];
      } else {
        for (k = 0; k < 3; k++) { o->p0[k] = ca[3 + k]; o->ad[k] = ca[k] - ca[3 + k]; }
      }
      o->r0 = ca[7]; o->r1 = ca[6];
    } else {
      if (o->type == T_RING) return 0;
      o->type = o->type == T_CONE ? T_CYL : T_TUBE;
      for (k = 0; k < 3; k++) { o->p0[k] = ca[k]; o->ad[k] = ca[3 + k] - ca[k]; }
      o->r0 = o->r1 = ca[6];
    }
    if (o->type == T_RING) {
      for (k = 0; k < 3; k++) { o->p0[k] = ca[k]; o->ad[k] = ca[3 + k]; }
    }
  }
  o->al = vnorm(o->ad);
  if (o->al == 0.0) return 0;
  if (o->type == T_RING) o->sl = o->r1 - o->r0;
  else if (o->type == T_CONE || o->type == T_CUP) o->sl = sqrt((o->r1 - o->r0) * (o->r1 - o->r0) + o->al * o->al);
  else o->sl = o->al;
  return 1;
}

static int findmaterial(const ro_scene *s, int oi) {
  int m = s->o[oi].mod, guard = 0;
  while (m >= 0 && guard++ < 10000) {
    if (is_material(s->o[m].type)) return m;
    m = s->o[m].mod;
  }
  return -1;
}

static void setflatss(OSrc *S) {
  double u[3], mult = .5 * sqrt(S->ss2);
  int k;
  perp(u, S->snorm);
  for (k = 0; k < 3; k++) S->ss[0][k] = u[k] * mult;
  vcross(S->ss[1], S->snorm, S->ss[0]);
}

/* bounding cube of the surfaces as oconv computes it (ot/oconv.c:120-139,
 * add2bbox ot/bbox.c:23-95, circle2bbox :117-140), margin 10*FTINY */
static void box_point(double *lo, double *hi, const double *p) {
  int k;
  for (k = 0; k < 3; k++) {
    if (p[k] < lo[k]) lo[k] = p[k];
    if (p[k] > hi[k]) hi[k] = p[k];
  }
}
static void box_circle(double *lo, double *hi, const double *c, const double *n, double rad) {
  int k;
  for (k = 0; k < 3; k++) {
    double r = sqrt(1. - n[k] * n[k]), d = c[k] + r * rad;
    if (d > hi[k]) hi[k] = d;
    d = c[k] - r * rad;
    if (d < lo[k]) lo[k] = d;
  }
}
static void scene_cube(ro_scene *s) {
  double lo[3] = {FHUGE, FHUGE, FHUGE}, hi[3] = {-FHUGE, -FHUGE, -FHUGE};
  int i, k;
  for (i = 0; i < s->n; i++) {
    const OObj *o = &s->o[i];
    if (!is_surface(o->type) || o->type == T_SOURCE || !o->ok) continue;
    if (o->type == T_POLY) {
      for (k = 0; k < o->nv; k++) box_point(lo, hi, o->va + 3 * k);
    } else if (o->type == T_SPHERE || o->type == T_BUBBLE) {
      for (k = 0; k < 3; k++) {
        double v[3] = {o->f[0], o->f[1], o->f[2]};
        v[k] -= o->f[3];
        box_point(lo, hi, v);
        v[k] += 2.0 * o->f[3];
        box_point(lo, hi, v);
      }
    } else {
      double p1[3];
      for (k = 0; k < 3; k++) p1[k] = o->type == T_RING ? o->p0[k] : o->p0[k] + o->ad[k] * o->al;
      if (o->type != T_RING) box_circle(lo, hi, o->p0, o->ad, o->r0);
      box_circle(lo, hi, p1, o->ad, o->r1);
    }
  }
  s->cusize = 0;
  if (lo[0] > hi[0]) return;
  for (k = 0; k < 3; k++) { lo[k] -= 10 * FTINY; hi[k] += 10 * FTINY; }
  for (k = 0; k < 3; k++)
    if (hi[k] - lo[k] > s->cusize) s->cusize = hi[k] - lo[k];
  for (k = 0; k < 3; k++) s->cuorg[k] = (hi[k] + lo[k] - s->cusize) * .5;
  /* as read back from the octree: 12 significant digits (ot/writeoct.c:44-47) */
  {
    char buf[40];
    for (k = 0; k < 3; k++) {
      snprintf(buf, sizeof buf, "%.12g", s->cuorg[k]);
      s->cuorg[k] = strtod(buf, NULL);
    }
    snprintf(buf, sizeof buf, "%.12g", s->cusize);
    s->cusize = strtod(buf, NULL);
  }
}

ro_scene *ro_scene_parse(const char *text, char *err, int errlen) {
  ro_scene *s = (ro_scene *)calloc(1, sizeof(ro_scene));
  const char *p = text;
  char w[256];
  int i, k;
  while (next_word(&p, w, sizeof w)) {
    OObj o;
    int ns, ni, nf;
    memset(&o, 0, sizeof o);
    o.src = -1;
    o.alt = -1;
    if (w[0] == '!') PERR("inline commands are not supported by the oracle");
    if (!strcmp(w, "void")) o.mod = -1;
    else if ((o.mod = findmod(s, w, s->n)) < 0) PERR("undefined modifier \"%s\"", w);
    if (!next_word(&p, w, sizeof w)) PERR("unexpected EOF");
    o.type = otype_of(w);
    if (o.type == T_OTHER) PERR("type \"%s\" not supported by the oracle", w);
    if (!next_word(&p, w, sizeof w)) PERR("unexpected EOF");
    snprintf(o.name, sizeof o.name, "%s", w);
    if (!next_word(&p, w, sizeof w)) PERR("bad arguments");
    ns = atoi(w);
    for (i = 0; i < ns; i++) {
      if (!next_word(&p, w, sizeof w)) PERR("bad arguments");
      if (i == 0 && o.type == M_ILLUM && strcmp(w, "void")) o.alt = findmod(s, w, s->n);
    }
    if (!next_word(&p, w, sizeof w)) PERR("bad arguments");
    ni = atoi(w);
    if (ni) PERR("integer arguments not supported");
    if (!next_word(&p, w, sizeof w)) PERR("bad arguments");
    nf = atoi(w);
    o.nf = nf;
    if (o.type == T_POLY) {
      o.va = (double *)malloc(sizeof(double) * (nf > 0 ? nf : 1));
      for (i = 0; i < nf; i++) {
        if (!next_word(&p, w, sizeof w)) PERR("bad arguments");
        o.va[i] = atof(w);
      }
    } else {
      if (nf > 16) PERR("too many real arguments");
      for (i = 0; i < nf; i++) {
        if (!next_word(&p, w, sizeof w)) PERR("bad arguments");
        o.f[i] = atof(w);
      }
    }
    if (s->n == s->cap) {
      s->cap = s->cap ? 2 * s->cap : 256;
      s->o = (OObj *)realloc(s->o, sizeof(OObj) * s->cap);
    }
    s->o[s->n++] = o;
  }
  /* surfaces */
  for (i = 0; i < s->n; i++) {
    OObj *o = &s->o[i];
    if (!is_surface(o->type) || o->type == T_SOURCE) continue;
    o->mat = findmaterial(s, i);
    if (o->mod < 0) continue;
    if (o->type == T_POLY) {
      if (o->nf < 9 || o->nf % 3) PERR("polygon %s: bad # arguments", o->name);
      o->ok = setup_face(o);
    } else if (o->type == T_SPHERE || o->type == T_BUBBLE) {
      if (o->nf != 4) PERR("sphere %s: bad # arguments", o->name);
      if (o->f[3] < -FTINY) { o->type = o->type == T_SPHERE ? T_BUBBLE : T_SPHERE; o->f[3] = -o->f[3]; }
      else if (o->f[3] <= FTINY) PERR("sphere %s: zero radius", o->name);
      o->ok = 1;
    } else {
      o->ok = setup_cone(o);
    }
  }
  /* sources: marksources() */
  for (i = 0; i < s->n; i++) {
    OObj *o = &s->o[i];
    OSrc S;
    int mi;
    OObj *m;
    if (!is_surface(o->type) || o->mod < 0) continue;
    if (o->type != T_SOURCE && !o->ok) continue;
    mi = findmaterial(s, i);
    if (mi < 0) continue;
    m = &s->o[mi];
    if (m->type != M_LIGHT && m->type != M_GLOW && m->type != M_ILLUM) continue;
    if (m->f[0] <= FTINY && (m->f[1] <= FTINY) & (m->f[2] <= FTINY)) continue;
    if (m->type == M_GLOW && o->type != T_SOURCE && m->f[3] <= FTINY) continue;
    memset(&S, 0, sizeof S);
    S.obj = i;
    S.mat = mi;
    if (o->type == T_SOURCE) {   /* ssetsrc */
      double theta;
      S.flags |= SDISTANT | SCIR;
      for (k = 0; k < 3; k++) S.sloc[k] = o->f[k];
      if (vnorm(S.sloc) == 0.0) PERR("source %s: zero direction", o->name);
      theta = PI / 180.0 / 2.0 * o->f[3];
      if (theta <= FTINY) PERR("source %s: zero size", o->name);
      S.ss2 = 2.0 * PI * (1.0 - cos(theta));
      S.srad = sqrt(S.ss2 / PI);
      memcpy(S.snorm, S.sloc, sizeof S.snorm);
      setflatss(&S);
    } else if (o->type == T_POLY) {   /* fsetsrc */
      int j;
      double n[3] = {0, 0, 0}, v1[3], v2[3], v3[3];
      for (k = 0; k < 3; k++) {
        S.sloc[k] = 0;
        for (j = 0; j < o->nv; j++) S.sloc[k] += o->va[3 * j + k];
        S.sloc[k] /= o->nv;
      }
      S.flags |= SFLAT;
      memcpy(S.snorm, o->norm, sizeof S.snorm);
      for (k = 0; k < 3; k++) v1[k] = o->va[3 + k] - o->va[k];
      for (j = 2; j < o->nv; j++) {
        for (k = 0; k < 3; k++) v2[k] = o->va[3 * j + k] - o->va[k];
        vcross(v3, v1, v2);
        for (k = 0; k < 3; k++) n[k] += v3[k];
        memcpy(v1, v2, sizeof v1);
      }
      S.ss2 = 0.5 * sqrt(DOT(n, n));
      for (j = 0; j < o->nv; j++) {
        double d2 = 0;
        for (k = 0; k < 3; k++) d2 += (o->va[3 * j + k] - S.sloc[k]) * (o->va[3 * j + k] - S.sloc[k]);
        if (d2 > S.srad) S.srad = d2;
      }
      S.srad = sqrt(S.srad);
      if (o->nv == 4) {
        for (k = 0; k < 3; k++) {
          S.ss[0][k] = .5 * (o->va[3 + k] - o->va[k]);
          S.ss[1][k] = .5 * (o->va[9 + k] - o->va[k]);
        }
      } else {
        setflatss(&S);
      }
    } else if (o->type == T_SPHERE || o->type == T_BUBBLE) {   /* sphsetsrc */
      S.flags |= SCIR;
      for (k = 0; k < 3; k++) { S.sloc[k] = o->f[k]; S.ss[k][k] = 0.7236 * o->f[3]; }
      S.srad = o->f[3];
      S.ss2 = PI * S.srad * S.srad;
    } else if (o->type == T_RING) {   /* rsetsrc */
      S.flags |= SFLAT | SCIR;
      memcpy(S.sloc, o->p0, sizeof S.sloc);
      memcpy(S.snorm, o->ad, sizeof S.snorm);
      S.srad = o->r1;
      S.ss2 = PI * S.srad * S.srad;
      setflatss(&S);
    } else {
      PERR("%s: surface type cannot be a source in the oracle", o->name);
    }
    if (m->type == M_GLOW) {
      S.flags |= SPROX;
      S.prox = m->f[3];
      if (S.flags & SDISTANT) S.flags |= SSKIP;
    }
    o->src = s->ns;
    s->s = (OSrc *)realloc(s->s, sizeof(OSrc) * (s->ns + 1));
    s->s[s->ns++] = S;
  }
  scene_cube(s);
  return s;
fail:
  ro_scene_free(s);
  return NULL;
}

void ro_scene_free(ro_scene *s) {
  int i;
  if (!s) return;
  for (i = 0; i < s->n; i++) free(s->o[i].va);
  free(s->o);
  free(s->s);
  free(s);
}

void ro_scene_counts(const ro_scene *s, long long *out) {
  int i, np = 0, nm = 0;
  for (i = 0; i < s->n; i++) {
    if (is_surface(s->o[i].type) && s->o[i].type != T_SOURCE && s->o[i].ok) np++;
    if (is_material(s->o[i].type)) nm++;
  }
  out[0] = s->n; out[1] = np; out[2] = nm; out[3] = s->ns;
}

/* ---- rays ---- */
typedef struct ORay {
  double rorg[3], rdir[3], rop[3], ron[3], rod, rot;
  const struct ORay *parent;
  int ro;              /* object index hit, or -1 */
  int rsrc, rlvl, rtype, crtype, adepth;
  float rweight;
  double rcoef[3], rcol[3];
  unsigned long long key;
  long sensor;
} ORay;

typedef struct {
  const ro_scene *s;
  const ro_params *p;
  int mode;            /* 0 rtrace, 1 rcontrib */
  int irrad;           /* -i */
  const ro_mod *mods;
  int nmods;
  int *modslot;        /* per object: tracked modifier slot */
  int *col0;
  int ncols;
  double *out;         /* current record row */
  unsigned long long nrays;
  int lamb;            /* -1: use built-in Lambertian */
  struct OCache *cache;  /* shadow cache per source (one run = one process) */
} Ctx;

static const double LAMBF[5] = {PI, PI, PI, 0, 0};

/* rayorigin(): returns 0 to trace, -1 to skip */
static int rayorigin(const Ctx *c, ORay *r, int rt, const ORay *ro, const double *rc, unsigned long long slot) {
  double rw = 1.0;
  int k;
  if (rc == NULL) {
    r->rcoef[0] = r->rcoef[1] = r->rcoef[2] = 1.0;
  } else {
    rw = cmaxd(rc);
    if (rw > 1.0) rw = 1.0;
    for (k = 0; k < 3; k++) r->rcoef[k] = rc[k];
  }
  r->parent = ro;
  r->rlvl = ro->rlvl;
  r->rsrc = ro->rsrc;
  if (rt & RAYREFL) { r->rlvl++; r->rsrc = -1; }
  r->crtype = ro->crtype | (r->rtype = rt);
  memcpy(r->rorg, ro->rop, sizeof r->rorg);
  r->rweight = (float)(ro->rweight * rw);
  r->adepth = ro->adepth;
  r->sensor = ro->sensor;
  r->key = ro_child_key(ro->key, slot);
  r->ro = -1;
  r->rot = FHUGE;
  r->rcol[0] = r->rcol[1] = r->rcol[2] = 0;
  if (r->rweight <= 0.0) return -1;
  if (r->crtype & SHADOW) return 0;
  if ((c->p->maxdepth <= 0) & (rc != NULL)) {
    if ((c->p->maxdepth < 0) & (r->rlvl > -c->p->maxdepth)) return -1;
    if (r->rweight >= c->p->minweight) return 0;
    if (ro_rnd(r->key, RN_RR) > r->rweight / c->p->minweight) return -1;
    rw = c->p->minweight / r->rweight;
    for (k = 0; k < 3; k++) r->rcoef[k] *= rw;
    r->rweight = (float)c->p->minweight;
    return 0;
  }
  return (r->rweight >= c->p->minweight) & (r->rlvl <= abs(c->p->maxdepth)) ? 0 : -1;
}

/* rayreject() */
static int rayreject(const ro_scene *s, int oi, const ORay *r, double t, double rod) {
  int mn, mr;
  if ((t <= FTINY) | (t > r->rot + FTINY)) return 1;
  if (t < r->rot - FTINY) return 0;
  if (oi == r->ro) return 1;
  if (r->ro < 0) return 0;
  mn = s->o[oi].mat;
  mr = s->o[r->ro].mat;
  if (mn < 0) { if (mr >= 0) return 1; }
  else if (mr < 0) return 0;
  else if (s->o[mn].type == M_TRANS || s->o[mn].type == M_GLASS) {
    if (!(s->o[mr].type == M_TRANS || s->o[mr].type == M_GLASS)) return 1;
  } else if (s->o[mr].type == M_TRANS || s->o[mr].type == M_GLASS) return 0;
  if (rod <= 0) { if (r->rod > 0) return 1; }
  else if (r->rod <= 0) return 0;
  return s->o[r->ro].mod >= s->o[oi].mod;
}

static int inface(const double *p, const OObj *f) {
  int ncross = 0, n = f->nv, xi, yi, tst;
  double x, y;
  const double *p0, *p1;
  if ((xi = f->ax + 1) >= 3) xi -= 3;
  if ((yi = xi + 1) >= 3) yi -= 3;
  x = p[xi];
  y = p[yi];
  p0 = f->va + 3 * (n - 1);
  p1 = f->va;
  while (n--) {
    if (fabs(p0[yi] - y) <= FTINY && fabs(p1[yi] - y) <= FTINY && ((p0[xi] > x) ^ (p1[xi] > x))) return 1;
    if ((p0[yi] > y) ^ (p1[yi] > y)) {
      tst = (p0[xi] > x) + (p1[xi] > x);
      if (tst == 2) ncross++;
      else if (tst) {
        double prodA = (p0[yi] - y) * (p1[xi] - x), prodB = (p0[xi] - x) * (p1[yi] - y);
        if (fabs(prodA - prodB) <= FTINY) return 1;
        ncross += (p1[yi] > p0[yi]) ^ (prodA > prodB);
      } else if (fabs(p0[xi] - x) <= FTINY && fabs(p1[xi] - x) <= FTINY) return 1;
    }
    p0 = p1;
    p1 += 3;
  }
  return ncross & 1;
}

static int quadratic(double *r, double a, double b, double c) {
  double disc;
  int first;
  if (a < -FTINY) first = 1;
  else if (a > FTINY) first = 0;
  else if (fabs(b) > FTINY) { r[0] = -c / b; return 1; }
  else return 0;
  b *= 0.5;
  disc = b * b - a * c;
  if (disc < -FTINY * FTINY) return 0;
  if (disc <= FTINY * FTINY) { r[0] = -b / a; return 1; }
  disc = sqrt(disc);
  r[first] = (-b - disc) / a;
  r[1 - first] = (-b + disc) / a;
  return 2;
}

static void set_hit(ORay *r, int oi, double t, const double *n) {
  int k;
  r->ro = oi;
  r->rot = t;
  for (k = 0; k < 3; k++) { r->rop[k] = r->rorg[k] + r->rdir[k] * t; r->ron[k] = n[k]; }
  r->rod = -DOT(r->rdir, r->ron);
}

/* one object: o_face / o_sphere / o_cone */
static int hit_obj(const ro_scene *s, int oi, ORay *r) {
  const OObj *o = &s->o[oi];
  int k;
  if (o->type == T_POLY) {
    double rdot = -DOT(r->rdir, o->norm), t, pi[3];
    if ((rdot <= FTINY) & (rdot >= -FTINY)) return 0;
    t = (DOT(r->rorg, o->norm) - o->offset) / rdot;
    if (rayreject(s, oi, r, t, rdot)) return 0;
    for (k = 0; k < 3; k++) pi[k] = r->rorg[k] + r->rdir[k] * t;
    if (!inface(pi, o)) return 0;
    set_hit(r, oi, t, o->norm);
    return 1;
  }
  if (o->type == T_SPHERE || o->type == T_BUBBLE) {
    double a = 0, b = 0, c = 0, root[2], t = 0, n[3], rad;
    int nr, i;
    for (k = 0; k < 3; k++) {
      double d = r->rorg[k] - o->f[k];
      a += r->rdir[k] * r->rdir[k];
      b += 2.0 * r->rdir[k] * d;
      c += d * d;
    }
    c -= o->f[3] * o->f[3];
    nr = quadratic(root, a, b, c);
    for (i = 0; i < nr; i++)
      if ((t = root[i]) > FTINY) break;
    if (i >= nr) return 0;
    if (rayreject(s, oi, r, t, 1 - 2 * ((i > 0) ^ (o->type == T_BUBBLE)))) return 0;
    rad = o->f[3] * (1 - 2 * (o->type == T_BUBBLE));
    for (k = 0; k < 3; k++) n[k] = (r->rorg[k] + r->rdir[k] * t - o->f[k]) / rad;
    set_hit(r, oi, t, n);
    return 1;
  }
  if (o->type == T_RING) {
    double dz = DOT(r->rdir, o->ad), t, q[3], a;
    double oc[3] = {r->rorg[0] - o->p0[0], r->rorg[1] - o->p0[1], r->rorg[2] - o->p0[2]};
    if ((dz <= FTINY) & (dz >= -FTINY)) return 0;
    t = -DOT(oc, o->ad) / dz;
    if (rayreject(s, oi, r, t, -dz)) return 0;
    for (k = 0; k < 3; k++) q[k] = oc[k] + r->rdir[k] * t;
    a = DOT(q, q);
    if (a > o->r1 * o->r1 || a < o->r0 * o->r0) return 0;
    set_hit(r, oi, t, o->ad);
    return 1;
  }
  {
    /* cone family: solve |w_perp + t d_perp|^2 = (rad(h))^2 along the axis */
    int cyl = (o->type == T_CYL) | (o->type == T_TUBE), rn, nr;
    int inward = (o->type == T_CUP) | (o->type == T_TUBE);
    double w[3], dp[3], wp[3], dd, wd, a, b, c, root[2];
    for (k = 0; k < 3; k++) w[k] = r->rorg[k] - o->p0[k];
    dd = DOT(r->rdir, o->ad);
    wd = DOT(w, o->ad);
    for (k = 0; k < 3; k++) { dp[k] = r->rdir[k] - o->ad[k] * dd; wp[k] = w[k] - o->ad[k] * wd; }
    if (cyl) {
      a = DOT(dp, dp);
      b = 2.0 * DOT(dp, wp);
      c = DOT(wp, wp) - o->r0 * o->r0;
    } else {
      double kk = (o->r1 - o->r0) / o->al, rw = o->r0 + kk * wd, rd = kk * dd;
      a = DOT(dp, dp) - rd * rd;
      b = 2.0 * (DOT(dp, wp) - rw * rd);
      c = DOT(wp, wp) - rw * rw;
    }
    nr = quadratic(root, a, b, c);
    for (rn = 0; rn < nr; rn++) {
      double t = root[rn], p[3], hb, n[3], aa, cc, l2;
      if (t <= FTINY) continue;
      if (t > r->rot + FTINY) break;
      for (k = 0; k < 3; k++) p[k] = r->rorg[k] + r->rdir[k] * t - o->p0[k];
      hb = DOT(p, o->ad);
      if (hb < 0.0 || hb > o->al) continue;
      if (!cyl && o->r0 + (o->r1 - o->r0) * hb / o->al < 0.0) continue;
      if (rayreject(s, oi, r, t, 1 - 2 * ((rn > 0) ^ inward))) break;
      if (cyl) {
        aa = o->type == T_CYL ? o->r0 : -o->r0;
        for (k = 0; k < 3; k++) n[k] = (p[k] - hb * o->ad[k]) / aa;
      } else {
        cc = o->r1 - o->r0;
        aa = o->r0 + hb * cc / o->al;
        if (o->type == T_CUP) { cc = -cc; aa = -aa; }
        for (k = 0; k < 3; k++) n[k] = (p[k] - hb * o->ad[k]) / aa;
        for (k = 0; k < 3; k++) n[k] = (o->al * n[k] - cc * o->ad[k]) / o->sl;
      }
      l2 = DOT(n, n);
      if ((l2 > 1. + FTINY) | (l2 < 1. - FTINY)) { double sc = 1. / (.5 + .5 * l2); n[0] *= sc; n[1] *= sc; n[2] *= sc; }
      set_hit(r, oi, t, n);
      return 1;
    }
    return 0;
  }
}

/* localhit() by brute force over every surface */
static int localhit(Ctx *c, ORay *r) {
  int i;
  c->nrays++;
  r->ro = -1;
  r->rot = FHUGE;
  r->rod = 1.0;
  for (i = 0; i < c->s->n; i++) {
    const OObj *o = &c->s->o[i];
    if (!is_surface(o->type) || o->type == T_SOURCE || !o->ok) continue;
    hit_obj(c->s, i, r);
  }
  return r->ro >= 0;
}


/* ---- shadow cache (rt/srcobstr.c) ----
 * The reference keeps, per light source, the last opaque object found to
 * block it in each of SHADCACHE^2-ish cells (distant sources: cells over
 * the cube face the source shines on, indexed by where the ray origin
 * projects along the source direction; local sources: cells over the
 * directions of approach).  At a source's first shadow test it fills the
 * cells by casting a grid of rays (castshadow), which the ray counter
 * counts; afterwards a shadow ray whose cell's blocker intersects it is not
 * traced, and a traced shadow ray found blocked records its blocker. */
#define SHADCACHE 20
typedef struct OCache {
  int init, len, ax;
  double o[3], e1, e2, u[3], v[3];
  int *obs;            /* object index per cell, -1: none */
} OCache;

static int src_cacheable(const OSrc *S) { return !(S->flags & (SSKIP | SPROX)); }
static void cache_init(Ctx *c, int sn);

/* srcobstructp(): cell of ray r aimed at source sn, -1 outside */
static int cache_cell(const Ctx *c, int sn, const ORay *r) {
  const OSrc *S = &c->s->s[sn];
  const OCache *C = &c->cache[sn];
  int ondx;
  if (S->flags & SDISTANT) {
    int ax = C->ax, ax1 = (ax + 1) % 3, ax2 = (ax + 2) % 3;
    double t = (C->o[ax] - r->rorg[ax]) / S->sloc[ax];
    if (t <= FTINY) return -1;
    /* in the operation order of the reference's -O3 -ffast-math build of
     * rt/srcobstr.c:216-221 (reassociated to ((t sl - o) + ro) (e 40)):
     * shadow-ray origins on the cube's centre lines fall exactly on a cell
     * boundary, where only that order gives the reference's cell */
    ondx = 2 * SHADCACHE * (int)(((t * S->sloc[ax1] - C->o[ax1]) + r->rorg[ax1]) * (C->e1 * (2 * SHADCACHE)));
    ondx += (int)(((t * S->sloc[ax2] - C->o[ax2]) + r->rorg[ax2]) * (C->e2 * (2 * SHADCACHE)));
    if ((ondx < 0) | (ondx >= 4 * SHADCACHE * SHADCACHE)) return -1;
  } else if (S->flags & SFLAT) {
    double sd[3], sd0m, sd1m;
    sd[0] = -DOT(r->rdir, C->u);
    sd[1] = -DOT(r->rdir, C->v);
    sd[2] = -DOT(r->rdir, S->snorm);
    if (sd[2] < 0) return -1;
    sd0m = fabs(sd[0]);
    sd1m = fabs(sd[1]);
    if (sd[2] >= sd0m && sd[2] >= sd1m) {
      ondx = SHADCACHE * (int)(SHADCACHE * (.5 - FTINY) * (1. + sd[0] / sd[2]));
      ondx += (int)(SHADCACHE * (.5 - FTINY) * (1. + sd[1] / sd[2]));
    } else if (sd0m >= sd1m) {
      ondx = SHADCACHE * SHADCACHE;
      if (sd[0] < 0) ondx += ((SHADCACHE + 1) >> 1) * SHADCACHE;
      ondx += SHADCACHE * (int)(SHADCACHE * (.5 - FTINY) * (1. - sd[2] / sd0m));
      ondx += (int)(SHADCACHE * (.5 - FTINY) * (1. + sd[1] / sd0m));
    } else {
      ondx = SHADCACHE * SHADCACHE + ((SHADCACHE + 1) >> 1) * SHADCACHE * 2;
      if (sd[1] < 0) ondx += ((SHADCACHE + 1) >> 1) * SHADCACHE;
      ondx += SHADCACHE * (int)(SHADCACHE * (.5 - FTINY) * (1. - sd[2] / sd1m));
      ondx += (int)(SHADCACHE * (.5 - FTINY) * (1. + sd[0] / sd1m));
    }
  } else {
    int ax = 0, ax1, ax2, k;
    double amax = 0;
    for (k = 3; k--;)
      if (fabs(r->rdir[k]) > amax) { amax = fabs(r->rdir[k]); ax = k; }
    ax1 = (ax + 1) % 3;
    ax2 = (ax + 2) % 3;
    ondx = 2 * SHADCACHE * SHADCACHE * ax;
    if (r->rdir[ax] < 0) ondx += SHADCACHE * SHADCACHE;
    ondx += SHADCACHE * (int)(SHADCACHE * (.5 - FTINY) * (1. + r->rdir[ax1] / amax));
    ondx += (int)(SHADCACHE * (.5 - FTINY) * (1. + r->rdir[ax2] / amax));
  }
  return ondx;
}

static int mat_opaque(const ro_scene *s, int oi) {
  int m = s->o[oi].mat, t;
  if (m < 0) return 0;
  t = s->o[m].type;
  return t == M_PLASTIC || t == M_METAL || t == M_LIGHT || t == M_GLOW;
}

/* srcblocker(): record what blocked ray r (aimed at r->rsrc); 1 if recorded */
static int src_blocker(Ctx *c, const ORay *r) {
  int cell;
  if (r->ro < 0 || r->rsrc < 0 || c->s->s[r->rsrc].obj == r->ro) return 0;
  if (!mat_opaque(c->s, r->ro)) return 0;
  if (!src_cacheable(&c->s->s[r->rsrc])) return 1;   /* written to a dummy cell */
  if (!c->cache[r->rsrc].init) cache_init(c, r->rsrc);
  cell = cache_cell(c, r->rsrc, r);
  if (cell >= 0) c->cache[r->rsrc].obs[cell] = r->ro;
  return 1;
}

/* castshadow() */
static void cast_shadow(Ctx *c, int sn, const double *org, const double *dir) {
  ORay rt;
  int k;
  memset(&rt, 0, sizeof rt);
  memcpy(rt.rorg, org, sizeof rt.rorg);
  memcpy(rt.rdir, dir, sizeof rt.rdir);
  rt.crtype = rt.rtype = PRIMARY;
  rt.rsrc = -1;
  while (localhit(c, &rt)) {
    ORay rt1 = rt;
    rt1.crtype |= rt1.rtype = RSHADOW;
    for (k = 0; k < 3; k++) { rt1.rdir[k] = -rt.rdir[k]; rt1.rorg[k] = rt.rop[k] - rt.rdir[k]; }
    rt1.rod = -rt.rod;
    rt1.rot = 1.;
    rt1.rsrc = sn;
    if (src_blocker(c, &rt1)) return;
    for (k = 0; k < 3; k++) rt.rorg[k] = rt.rop[k] + rt.rdir[k] * 10. * FTINY;
  }
}

/* initobscache() */
static void cache_init(Ctx *c, int sn) {
  const OSrc *S = &c->s->s[sn];
  OCache *C = &c->cache[sn];
  double rorg[3], rdir[3], d;
  int i, j, k;
  C->init = 1;
  if (S->flags & SDISTANT) C->len = 4 * SHADCACHE * SHADCACHE;
  else if (S->flags & SFLAT) C->len = SHADCACHE * SHADCACHE * 3 + (SHADCACHE & 1) * SHADCACHE * 4;
  else C->len = SHADCACHE * SHADCACHE * 6;
  C->obs = (int *)malloc(sizeof(int) * C->len);
  for (i = 0; i < C->len; i++) C->obs[i] = -1;
  if (S->flags & SDISTANT) {
    double amax = 0;
    int ax = 0, ax1, ax2;
    for (k = 3; k--;)
      if (fabs(S->sloc[k]) > amax) { amax = fabs(S->sloc[k]); ax = k; }
    C->ax = ax;
    ax1 = (ax + 1) % 3;
    ax2 = (ax + 2) % 3;
    memcpy(C->o, c->s->cuorg, sizeof C->o);
    if (S->sloc[ax] > 0) C->o[ax] += c->s->cusize;
    if (S->sloc[ax1] < 0) C->o[ax1] += c->s->cusize * S->sloc[ax1] / amax;
    if (S->sloc[ax2] < 0) C->o[ax2] += c->s->cusize * S->sloc[ax2] / amax;
    C->e1 = 1. / (c->s->cusize * (1. + fabs(S->sloc[ax1]) / amax));
    C->e2 = 1. / (c->s->cusize * (1. + fabs(S->sloc[ax2]) / amax));
  } else if (S->flags & SFLAT) {
    memcpy(C->u, S->ss[0], sizeof C->u);
    vnorm(C->u);
    vcross(C->v, S->snorm, C->u);
  }
  if (sn >= 200) return;   /* MAX2SHADE */
  if (S->flags & SDISTANT) {
    int ax1 = (C->ax + 1) % 3, ax2 = (C->ax + 2) % 3;
    for (k = 0; k < 3; k++) rdir[k] = -S->sloc[k];
    for (i = 2 * SHADCACHE; i--;)
      for (j = 2 * SHADCACHE; j--;) {
        /* (i + .5) / (2 SHADCACHE e1) as the reference build evaluates it:
         * (i + .5) * 0.025 / e1 (rt/srcobstr.c:127-131) */
        memcpy(rorg, C->o, sizeof rorg);
        rorg[ax1] += (i + .5) * (1.0 / (2 * SHADCACHE)) / C->e1;
        rorg[ax2] += (j + .5) * (1.0 / (2 * SHADCACHE)) / C->e2;
        cast_shadow(c, sn, rorg, rdir);
      }
  } else if (S->flags & SFLAT) {
    d = 0.01 * S->srad;
    for (k = 0; k < 3; k++) rorg[k] = S->sloc[k] + S->snorm[k] * d;
    for (i = SHADCACHE; i--;)
      for (j = SHADCACHE; j--;) {
        d = 2. / SHADCACHE * (i + .5) - 1.;
        for (k = 0; k < 3; k++) rdir[k] = S->snorm[k] + C->u[k] * d;
        d = 2. / SHADCACHE * (j + .5) - 1.;
        for (k = 0; k < 3; k++) rdir[k] += C->v[k] * d;
        vnorm(rdir);
        cast_shadow(c, sn, rorg, rdir);
      }
    for (int kk = 2; kk--;)
      for (i = SHADCACHE; i--;)
        for (j = SHADCACHE >> 1; j--;) {
          d = 2. / SHADCACHE * (i + .5) - 1.;
          for (k = 0; k < 3; k++) rdir[k] = kk ? C->u[k] + C->v[k] * d : C->v[k] + C->u[k] * d;
          d = 1. - 2. / SHADCACHE * (j + .5);
          for (k = 0; k < 3; k++) rdir[k] += S->snorm[k] * d;
          vnorm(rdir);
          cast_shadow(c, sn, rorg, rdir);
          d = 2. * DOT(rdir, S->snorm);
          for (k = 0; k < 3; k++) rdir[k] = d * S->snorm[k] - rdir[k];
          cast_shadow(c, sn, rorg, rdir);
        }
  } else {
    for (int kk = 6; kk--;) {
      int ax = kk % 3, ax1 = (kk + 1) % 3, ax2 = (kk + 2) % 3;
      for (i = SHADCACHE; i--;)
        for (j = SHADCACHE; j--;) {
          rdir[0] = rdir[1] = rdir[2] = 0.;
          rdir[ax] = kk < 3 ? 1. : -1.;
          rdir[ax1] = 2. / SHADCACHE * (i + .5) - 1.;
          rdir[ax2] = 2. / SHADCACHE * (j + .5) - 1.;
          vnorm(rdir);
          d = 1.05 * S->srad;
          for (k = 0; k < 3; k++) rorg[k] = S->sloc[k] + rdir[k] * d;
          cast_shadow(c, sn, rorg, rdir);
        }
    }
  }
}

/* srcblocked(): 1 when the cached blocker of the ray's cell intersects it
 * (in front of a local source) */
static int src_blocked(Ctx *c, const ORay *sr) {
  const OSrc *S = &c->s->s[sr->rsrc];
  ORay t = *sr;
  int cell, ob;
  if (!src_cacheable(S)) return 0;
  if (!c->cache[sr->rsrc].init) cache_init(c, sr->rsrc);
  cell = cache_cell(c, sr->rsrc, sr);
  if (cell < 0 || (ob = c->cache[sr->rsrc].obs[cell]) < 0) return 0;
  t.ro = -1;
  t.rot = FHUGE;
  t.rod = 1.0;
  if (!hit_obj(c->s, ob, &t)) return 0;
  if (S->flags & SDISTANT) return 1;
  return !hit_obj(c->s, S->obj, &t);
}

static void cache_free(Ctx *c) {
  int i;
  if (!c->cache) return;
  for (i = 0; i < c->s->ns; i++) free(c->cache[i].obs);
  free(c->cache);
}

void ro_closest(const ro_scene *s, const double *rays, long n, int *prim, double *t) {
  Ctx c;
  long i;
  int j, k, np = 0;
  int *pmap = (int *)malloc(sizeof(int) * (s->n + 1));
  memset(&c, 0, sizeof c);
  c.s = s;
  for (j = 0; j < s->n; j++) {
    const OObj *o = &s->o[j];
    pmap[j] = (is_surface(o->type) && o->type != T_SOURCE && o->ok) ? np++ : -1;
  }
  for (i = 0; i < n; i++) {
    ORay r;
    memset(&r, 0, sizeof r);
    for (k = 0; k < 3; k++) { r.rorg[k] = rays[6 * i + k]; r.rdir[k] = rays[6 * i + 3 + k]; }
    vnorm(r.rdir);
    localhit(&c, &r);
    prim[i] = r.ro >= 0 ? pmap[r.ro] : -1;
    t[i] = r.rot;
  }
  free(pmap);
}

/* ---- bins ---- */
static double asin_deg(double x) { double v = (x - 1 > 0) ? PI / 2 : ((-1 - x > 0) ? -PI / 2 : asin(x)); return v / (PI / 180); }
static double acos_deg(double x) { double v = (x - 1 > 0) ? 0 : ((-1 - x > 0) ? PI : acos(x)); return v / (PI / 180); }
static double atan2_deg(double y, double x) { double a = atan2(y, x); return ((-a > 0) ? a + 2 * PI : a) / (PI / 180); }
static double rnaz(double r, double mf) {
  static const double tnaz[7] = {30, 30, 24, 24, 18, 12, 6};
  int k;
  if (r - (7 * mf - .5) > 0) return 1;
  k = (int)floor((r + .5) / mf);
  return (k < 0 || k > 6) ? 0 : mf * tnaz[k];
}
static double raccum(double r, double mf) { return (r - .5 > 0) ? rnaz(r - 1, mf) + raccum(r - 1, mf) : 0; }
static int bin_of(const ro_mod *m, const double *D) {
  double bval;
  if (m->bin_kind == 0) {
    bval = m->cbin;
  } else if (m->bin_kind == 5) {   /* rfluxmtx h=u surface: if(-D.n,0,-1), util/rfluxmtx.c:465-470 */
    bval = (-D[0] * m->nrm[0] - D[1] * m->nrm[1] - D[2] * m->nrm[2] > 0) ? 0 : -1;
  } else if (m->bin_kind == 6) {   /* scbin, util/disk2square.cal:67-90 and :34-64 */
    const double *N = m->nrm, *U = m->up;
    double inc_dz = -D[0] * N[0] - D[1] * N[1] - D[2] * N[2];
    double inc_rx = -m->rhs * (D[0] * (U[1] * N[2] - U[2] * N[1]) + D[1] * (U[2] * N[0] - U[0] * N[2]) + D[2] * (U[0] * N[1] - U[1] * N[0]));
    double inc_ry = D[0] * U[0] + D[1] * U[1] + D[2] * U[2] + inc_dz * (N[0] * U[0] + N[1] * U[1] + N[2] * U[2]);
    double den2 = inc_rx * inc_rx + inc_ry * inc_ry;
    double radf = (den2 - 1e-7 > 0) ? sqrt((1 - inc_dz * inc_dz) / den2) : 0;
    double x = inc_rx * radf, y = -inc_ry * radf;
    double r = sqrt(x * x + y * y), phi = atan2(y, x), a, b;
    if (-phi - PI / 4 > 0) phi += 2 * PI;                    /* norm_radians */
    switch ((int)floor((phi + PI / 4) / (PI / 2)) + 1) {    /* in_disk_rgn */
      case 1: a = r; b = phi * r / (PI / 4); break;
      case 2: a = (PI / 2 - phi) * r / (PI / 4); b = r; break;
      case 3: a = -r; b = (PI - phi) * r / (PI / 4); break;
      case 4: a = (phi - 3 * PI / 2) * r / (PI / 4); b = -r; break;
      default: a = r; b = -r; break;
    }
    bval = (inc_dz > 0) ? floor((a + 1) / 2 * m->mf) * m->mf + floor((b + 1) / 2 * m->mf) : -1;
  } else if (m->bin_kind == 7) {   /* cbin, util/cieskyscan.cal:21-53 */
    static const double cnaz[8] = {30, 30, 24, 24, 18, 12, 6, 1};
    const double *N = m->nrm, *U = m->up;
    double inc_dz = -D[0] * N[0] - D[1] * N[1] - D[2] * N[2];
    double inc_rx = -m->rhs * (D[0] * (U[1] * N[2] - U[2] * N[1]) + D[1] * (U[2] * N[0] - U[0] * N[2]) + D[2] * (U[0] * N[1] - U[1] * N[0]));
    double inc_ry = D[0] * U[0] + D[1] * U[1] + D[2] * U[2] + inc_dz * (N[0] * U[0] + N[1] * U[1] + N[2] * U[2]);
    double r_alt = asin_deg(inc_dz), c_azi = atan2_deg(-inc_rx, -inc_ry);
    double row = floor(r_alt / (90 / 7.5)), inc, fwd, acc = 0;
    int k, ri = (int)row;
    if (!(r_alt > 0) || ri < 0 || ri > 7) {
      bval = -1;
    } else {
      inc = 360 / cnaz[ri];
      fwd = floor(row / 2 + 1.25) - row / 2 - .75;
      for (k = 0; k < ri; k++) acc += cnaz[k];   /* caccum */
      bval = acc + ((359.9999 - c_azi > 0) ? (fwd > 0 ? floor(c_azi / inc) : cnaz[ri] - 1 - floor(c_azi / inc)) : 0);
    }
  } else if (m->bin_kind == 1) {
    const double *N = m->nrm, *U = m->up;
    double inc_dz = -D[0] * N[0] - D[1] * N[1] - D[2] * N[2];
    double inc_rx = -m->rhs * (D[0] * (U[1] * N[2] - U[2] * N[1]) + D[1] * (U[2] * N[0] - U[0] * N[2]) + D[2] * (U[0] * N[1] - U[1] * N[0]));
    double inc_ry = D[0] * U[0] + D[1] * U[1] + D[2] * U[2] + inc_dz * (N[0] * U[0] + N[1] * U[1] + N[2] * U[2]);
    double r_alt = asin_deg(inc_dz), r_azi = atan2_deg(inc_rx, inc_ry);
    double alpha = 90 / (m->mf * 7 + .5), r_row = floor(r_alt / alpha), r_inc = 360 / rnaz(r_row, m->mf);
    double r_azn = (359.9999 - .5 * r_inc - r_azi > 0) ? floor((r_azi + .5 * r_inc) / r_inc) : 0;
    bval = (r_alt > 0) ? raccum(r_row, m->mf) + r_azn : -1;
  } else {
    static const double pf[] = {9, 5, 15, 25, 35, 45, 55, 65, 75, 90}, nf[] = {9, 1, 8, 16, 20, 24, 24, 24, 16, 12};
    static const double ph[] = {7, 6.5, 19.5, 32.5, 45.5, 58.5, 71.5, 90}, nh[] = {7, 1, 8, 12, 16, 20, 12, 8};
    static const double pq[] = {5, 9, 27, 45, 63, 90}, nq[] = {5, 1, 8, 12, 12, 8};
    const double *kp = m->bin_kind == 2 ? pf : m->bin_kind == 3 ? ph : pq;
    const double *kn = m->bin_kind == 2 ? nf : m->bin_kind == 3 ? nh : nq;
    const double *N = m->nrm, *U = m->up;
    double pol = acos_deg(-D[0] * N[0] - D[1] * N[1] - D[2] * N[2]);
    double azi = atan2_deg(-D[0] * U[0] - D[1] * U[1] - D[2] * U[2] + (N[0] * D[0] + N[1] * D[1] + N[2] * D[2]) * (N[0] * U[0] + N[1] * U[1] + N[2] * U[2]),
                           -m->rhs * (D[0] * (U[1] * N[2] - U[2] * N[1]) + D[1] * (U[2] * N[0] - U[0] * N[2]) + D[2] * (U[0] * N[1] - U[1] * N[0])));
    int row = 1, k;
    double acc = 0, inc;
    if (pol - 90 > 0) {
      bval = -1;
    } else {
      while (!(row - kp[0] + .5 > 0) && pol - kp[row] > 0) row++;   /* kfindrow */
      for (k = 1; k < row; k++) acc += kn[k];                        /* kaccum */
      inc = 360 / kn[row];
      bval = acc + (((360 - .5 * inc) - azi > 0) ? floor((azi + .5 * inc) / inc) : 0);
    }
  }
  if (bval <= -.5) return -1;
  if ((int)(bval + .5) >= m->nbins) return -1;
  return (int)(bval + .5);
}
void ro_bins(const ro_mod *m, const double *dirs, long n, int *bins) {
  long i;
  for (i = 0; i < n; i++) bins[i] = bin_of(m, dirs + 3 * i);
}

/* trace_contrib(): bin this ray's cumulative coefficient */
static void trace_contrib(Ctx *c, const ORay *r) {
  int slot, bn, k;
  double contr[3] = {1, 1, 1};
  const ORay *q;
  if (c->mode != 1) return;
  if (r->ro == -1) return;
  if (r->rsrc >= 0 && c->s->s[r->rsrc].obj != r->ro) return;
  slot = c->modslot[r->ro];
  if (slot < 0) return;
  if (c->p->contrib) {
    if (!(r->rcoef[0] * r->rcol[0] > FTINY || r->rcoef[1] * r->rcol[1] > FTINY || r->rcoef[2] * r->rcol[2] > FTINY)) return;
  } else if (cmaxd(r->rcoef) <= FTINY) {
    return;
  }
  bn = bin_of(&c->mods[slot], r->rdir);
  if (bn < 0) return;
  for (q = r; q != NULL; q = q->parent)   /* raycontrib() */
    for (k = 0; k < 3; k++) contr[k] *= q->rcoef[k];
  if (c->p->contrib)
    for (k = 0; k < 3; k++) contr[k] *= r->rcol[k];
  for (k = 0; k < 3; k++) c->out[3 * (c->col0[slot] + bn) + k] += contr[k];
}

static void raytrace(Ctx *c, ORay *r);
static int rayshade(Ctx *c, ORay *r, int mat);

static void raytrans(Ctx *c, ORay *r) {
  ORay tr;
  rayorigin(c, &tr, TRANS, r, NULL, SLOT_PASS);   /* always continue */
  memcpy(tr.rdir, r->rdir, sizeof tr.rdir);
  raytrace(c, &tr);
  memcpy(r->rcol, tr.rcol, sizeof r->rcol);
}

/* sourcehit(): returns source index or -1 */
static int sourcehit(Ctx *c, ORay *r) {
  int glowsrc = -1, transrc = -1, first = 0, last = c->s->ns - 1, i;
  if (r->rsrc >= 0) first = last = r->rsrc;
  for (i = first; i <= last; i++) {
    const OSrc *S = &c->s->s[i];
    if (!(S->flags & SDISTANT)) continue;
    if (2. * PI * (1. - DOT(S->sloc, r->rdir)) > S->ss2) continue;
    if (i == r->rsrc) return i;
    if (S->flags & SSKIP) { if (glowsrc < 0) glowsrc = i; continue; }
    if (c->s->o[S->mat].type == M_ILLUM && c->s->o[S->mat].alt < 0) { if (transrc < 0) transrc = i; continue; }
    return i;
  }
  if (transrc >= 0 && r->crtype & (AMBIENT | SPECULAR)) return -1;
  return glowsrc;
}

/* m_light() */
static int m_light(Ctx *c, const OObj *m, ORay *r) {
  int distglow = m->type == M_GLOW && m->f[3] >= -FTINY && r->rot > m->f[3];
  if (r->crtype & (AMBIENT | SPECULAR) && !(r->crtype & SHADOW || r->rod < 0.0 || distglow)) {
    r->rcoef[0] = r->rcoef[1] = r->rcoef[2] = 0;
    return 1;
  }
  if (r->rsrc >= 0 && c->s->s[r->rsrc].obj != r->ro) {
    const OSrc *S = &c->s->s[r->rsrc];
    int weak = c->s->o[S->mat].type == M_ILLUM || c->s->o[S->mat].type == M_GLOW;
    if (m->type != M_ILLUM || (r->rod > 0.0 && weak)) {
      r->rcoef[0] = r->rcoef[1] = r->rcoef[2] = 0;
      return 1;
    }
  }
  if (m->type == M_ILLUM && (r->rsrc < 0 || c->s->s[r->rsrc].obj != r->ro)) {
    if (m->alt >= 0 && r->ro >= 0) return rayshade(c, r, m->alt);
    raytrans(c, r);
    return 1;
  }
  if (r->rod < 0.0) {
    if (!c->p->backvis) raytrans(c, r);
    return 1;
  }
  r->rcol[0] = m->f[0];
  r->rcol[1] = m->f[1];
  r->rcol[2] = m->f[2];
  return 1;
}

/* ---- m_normal ---- */
#define FRESNE(ci) (exp(-5.85 * (ci)) - 0.00202943064)
#define FRESTHRESH 0.017999
#define SP_REFL 01
#define SP_TRAN 02
#define SP_PURE 04
#define SP_FLAT 010
#define SP_RBLT 020
#define SP_TBLT 040
typedef struct {
  int specfl;
  double mcolor[3], scolor[3], prdir[3], pnorm[3];
  double alpha2, rdiff, rspec, trans, tdiff, tspec, pdot;
  ORay *rp;
} ND;

static void dirnorm(const Ctx *c, double *cval, const ND *np, const double *ldir, double omega) {
  double ldot = DOT(np->pnorm, ldir), lrdiff = np->rdiff, ltdiff = np->tdiff, dtmp, d2, d3, d4, vtmp[3];
  int k;
  cval[0] = cval[1] = cval[2] = 0;
  if (ldot < 0.0 ? np->trans <= FTINY : np->trans >= 1.0 - FTINY) return;
  if (np->specfl & SP_PURE && np->rspec >= FRESTHRESH && (lrdiff > FTINY) | (ltdiff > FTINY)) {
    dtmp = 1. - FRESNE(fabs(ldot));
    lrdiff *= dtmp;
    ltdiff *= dtmp;
  }
  if ((ldot > FTINY) & (lrdiff > FTINY))
    for (k = 0; k < 3; k++) cval[k] += np->mcolor[k] * ldot * omega * lrdiff * (1.0 / PI);
  if ((ldot < -FTINY) & (ltdiff > FTINY))
    for (k = 0; k < 3; k++) cval[k] += np->mcolor[k] * -ldot * omega * ltdiff * (1.0 / PI);
  if ((ldot > FTINY) & ((np->specfl & (SP_REFL | SP_PURE)) == SP_REFL)) {
    dtmp = np->alpha2;
    if (np->specfl & SP_FLAT) dtmp += (1. - c->p->dstrsrc) * omega * (0.25 / PI);
    for (k = 0; k < 3; k++) vtmp[k] = ldir[k] - np->rp->rdir[k];
    d2 = DOT(vtmp, np->pnorm);
    d2 *= d2;
    d3 = DOT(vtmp, vtmp);
    d4 = (d3 - d2) / d2;
    dtmp = exp(-d4 / dtmp) * d3 / (PI * d2 * d2 * dtmp);
    if (dtmp > FTINY)
      for (k = 0; k < 3; k++) cval[k] += np->scolor[k] * dtmp * ldot * omega;
  }
  if ((ldot < -FTINY) & ((np->specfl & (SP_TRAN | SP_PURE)) == SP_TRAN)) {
    dtmp = np->alpha2 + omega * (1.0 / PI);
    dtmp = exp((2. * DOT(np->prdir, ldir) - 2.) / dtmp) / (PI * dtmp);
    if (dtmp > FTINY)
      for (k = 0; k < 3; k++) cval[k] += np->mcolor[k] * dtmp * np->tspec * omega * sqrt(-ldot / np->pdot);
  }
}

/* direct(): every source with a positive potential is shadow-tested */
static void direct(Ctx *c, ORay *r, const ND *nd) {
  int sn, k;
  for (sn = 0; sn < c->s->ns; sn++) {
    const OSrc *S = &c->s->s[sn];
    double vpos[3] = {0, 0, 0}, dir[3], dist, dom, coef[3], val[3];
    ORay sr;
    unsigned long long ks;
    if (S->flags & SSKIP) continue;
    if ((S->flags & (SPROX | SDISTANT)) == SPROX) {
      double dv[3] = {r->rop[0] - S->sloc[0], r->rop[1] - S->sloc[1], r->rop[2] - S->sloc[2]};
      if (DOT(dv, dv) > (S->prox + S->srad) * (S->prox + S->srad)) continue;
    }
    ks = ro_child_key(r->key, SLOT_SHADOW + sn);
    if (c->p->dstrsrc > FTINY) {
      vpos[0] = ro_rnd(ks, RN_SRC0);
      vpos[1] = ro_rnd(ks, RN_SRC0 + 1);
      vpos[2] = (S->flags & SFLAT) ? 0.5 : ro_rnd(ks, RN_SRC0 + 2);
      for (k = 0; k < 3; k++) vpos[k] = c->p->dstrsrc * (1. - 2. * vpos[k]);
    }
    if ((S->flags & SCIR) && c->p->dstrsrc > 0.7) {
      double trim[3], d;
      if (S->flags & (SFLAT | SDISTANT)) {
        d = 1.12837917;
        trim[0] = d * sqrt(1.0 - 0.5 * vpos[1] * vpos[1]);
        trim[1] = d * sqrt(1.0 - 0.5 * vpos[0] * vpos[0]);
        trim[2] = 0.0;
      } else {
        trim[2] = trim[0] = vpos[0] * vpos[0];
        d = vpos[1] * vpos[1];
        if (d > trim[2]) trim[2] = d;
        trim[0] += d;
        d = vpos[2] * vpos[2];
        if (d > trim[2]) trim[2] = d;
        trim[0] += d;
        if (trim[0] > FTINY * FTINY) {
          d = 1.0 / 0.7236;
          trim[2] = trim[1] = trim[0] = d * sqrt(trim[2] / trim[0]);
        } else {
          trim[2] = trim[1] = trim[0] = 0.0;
        }
      }
      for (k = 0; k < 3; k++) vpos[k] *= trim[k];
    }
    for (k = 0; k < 3; k++)
      dir[k] = S->sloc[k] + vpos[0] * S->ss[0][k] + vpos[1] * S->ss[1][k] + vpos[2] * S->ss[2][k];
    if (!(S->flags & SDISTANT))
      for (k = 0; k < 3; k++) dir[k] -= r->rop[k];
    if ((dist = vnorm(dir)) == 0.0) continue;
    if (S->flags & SFLAT) dom = -DOT(S->snorm, dir);
    else dom = 1.0;
    if (S->flags & SDISTANT) {
      dom *= S->ss2;
    } else {
      if (dom <= 1e-4) continue;
      dom *= S->ss2 / (dist * dist);
      if ((S->flags & SPROX) && dist > S->prox) continue;
    }
    /* shadow cache check of the source ray (rt/source.c:437-446) */
    rayorigin(c, &sr, RSHADOW, r, NULL, SLOT_SHADOW + sn);
    memcpy(sr.rdir, dir, sizeof dir);
    sr.rsrc = sn;
    if (src_blocked(c, &sr)) continue;
    dirnorm(c, coef, nd, dir, dom);
    if (!(cmaxd(coef) > 0.0)) continue;
    /* potential: srcvalue() */
    rayorigin(c, &sr, RSHADOW, r, NULL, SLOT_SHADOW + sn);
    memcpy(sr.rdir, dir, sizeof dir);
    sr.rsrc = sn;
    if (S->flags & SDISTANT) {
      if (sourcehit(c, &sr) != sn) continue;
      sr.ro = S->obj;
      sr.rot = FHUGE;
      sr.rod = 1.0;
    } else {
      sr.rot = FHUGE;
      sr.ro = -1;
      if (!hit_obj(c->s, S->obj, &sr)) continue;
    }
    m_light(c, &c->s->o[S->mat], &sr);
    for (k = 0; k < 3; k++) val[k] = sr.rcol[k] * coef[k];
    if (!(cbright(val) > 0.0)) continue;
    /* shadow test */
    rayorigin(c, &sr, ((r->rod > 0) ^ (DOT(r->ron, dir) > 0)) ? TSHADOW : RSHADOW, r, NULL, SLOT_SHADOW + sn);
    memcpy(sr.rcoef, coef, sizeof coef);
    memcpy(sr.rdir, dir, sizeof dir);
    sr.rsrc = sn;
    if (localhit(c, &sr) && (sr.ro != S->obj || c->mode == 1)) {
      const int hit = sr.ro;
      if (!rayshade(c, &sr, c->s->o[sr.ro].mat)) raytrans(c, &sr);
      trace_contrib(c, &sr);
      if (cmeand(sr.rcol) <= FTINY) {
        ORay b = sr;
        b.ro = hit;
        src_blocker(c, &b);   /* rt/source.c:512-518 */
        continue;
      }
      for (k = 0; k < 3; k++) val[k] = sr.rcol[k] * sr.rcoef[k];
    } else if (c->mode == 1 && (S->flags & SDISTANT)) {
      int si = sourcehit(c, &sr);
      if (si >= 0) {
        sr.ro = c->s->s[si].obj;
        sr.rod = 1.0;
        sr.rot = FHUGE;
        rayshade(c, &sr, c->s->s[si].mat);
        trace_contrib(c, &sr);
      }
    }
    for (k = 0; k < 3; k++) r->rcol[k] += val[k];
  }
}

/* multambient(), -aa 0 branch: stratified hemisphere */
static void multambient(Ctx *c, double *aval, ORay *r, const double *nrm) {
  const ro_params *p = c->p;
  double rdot = DOT(nrm, r->ron), d, wt, acoef[3], onrm[3], ux[3], uy[3], acol[3] = {0, 0, 0};
  int sgn = 1 - 2 * (rdot < 0), n, i, j, k, back, ok = 0;
  if (p->ambdiv <= 0 || r->adepth >= p->ambounce) goto dumbamb;
  d = cmaxd(aval);
  if (d <= FTINY) goto dumbamb;
  wt = r->rweight;
  d *= 0.8 * r->rweight / (p->ambdiv * p->minweight + 1e-20);
  if (wt > d) wt = d;
  n = (int)(sqrt(p->ambdiv * wt) + 0.5);
  if (n < 1) n = 1;
  back = sgn < 0;
  for (k = 0; k < 3; k++) { onrm[k] = back ? -r->ron[k] : r->ron[k]; acoef[k] = aval[k] / ((double)n * n); }
  frame(onrm, ro_rnd(r->key, back ? RN_FRAME_BACK : RN_FRAME_FRONT), ux, uy);
  for (i = n; i--;)
    for (j = n; j--;) {
      ORay ar;
      double spt[2], zd;
      unsigned long long slot = (back ? SLOT_HEMI_BACK : SLOT_HEMI_FRONT) + (unsigned long long)(i * n + j);
      if (rayorigin(c, &ar, back ? TAMBIENT : RAMBIENT, r, acoef, slot) < 0) continue;
      ar.adepth = r->adepth + 1;
      square2disk(spt, (j + ro_rnd(ar.key, RN_AMB1)) / n, (i + ro_rnd(ar.key, RN_AMB0)) / n);
      zd = 1. - spt[0] * spt[0] - spt[1] * spt[1];
      zd = zd > 0 ? sqrt(zd) : 0;
      for (k = 0; k < 3; k++) ar.rdir[k] = spt[0] * ux[k] + spt[1] * uy[k] + zd * onrm[k];
      vnorm(ar.rdir);
      raytrace(c, &ar);
      for (k = 0; k < 3; k++) acol[k] += ar.rcol[k] * ar.rcoef[k];
      ok++;
    }
  if (!ok) goto dumbamb;
  for (k = 0; k < 3; k++) aval[k] = acol[k];
  return;
dumbamb:
  for (k = 0; k < 3; k++) aval[k] *= p->ambval[k];
}

static void gaussamp(Ctx *c, ND *np) {
  double u[3], v[3];
  int k, nt;
  ORay sr;
  if ((np->specfl & (SP_REFL | SP_RBLT)) != SP_REFL && (np->specfl & (SP_TRAN | SP_TBLT)) != SP_TRAN) return;
  frame(np->pnorm, ro_rnd(np->rp->key, RN_GFRAME), u, v);
  if ((np->specfl & (SP_REFL | SP_RBLT)) == SP_REFL &&
      rayorigin(c, &sr, RSPECULAR, np->rp, np->scolor, SLOT_RSPEC) == 0) {
    for (nt = 0; nt < 10; nt++) {
      double rv0 = ro_rnd(sr.key, RN_GAUSS + 2 * nt), rv1 = ro_rnd(sr.key, RN_GAUSS + 2 * nt + 1);
      double d = 2.0 * PI * rv0, cosp = cos(d), sinp = sin(d), h[3];
      if ((0. <= c->p->specjitter) & (c->p->specjitter < 1.)) rv1 = 1.0 - c->p->specjitter * rv1;
      d = rv1 <= FTINY ? 1.0 : sqrt(np->alpha2 * -log(rv1));
      for (k = 0; k < 3; k++) h[k] = np->pnorm[k] + d * (cosp * u[k] + sinp * v[k]);
      d = -2.0 * DOT(h, np->rp->rdir) / (1.0 + d * d);
      for (k = 0; k < 3; k++) sr.rdir[k] = np->rp->rdir[k] + h[k] * d;
      if (DOT(sr.rdir, np->rp->ron) <= FTINY) continue;
      vnorm(sr.rdir);
      raytrace(c, &sr);
      for (k = 0; k < 3; k++) np->rp->rcol[k] += sr.rcol[k] * sr.rcoef[k];
      break;
    }
  }
  {
    double tc[3];
    for (k = 0; k < 3; k++) tc[k] = np->mcolor[k] * np->tspec;
    if ((np->specfl & (SP_TRAN | SP_TBLT)) == SP_TRAN && rayorigin(c, &sr, TSPECULAR, np->rp, tc, SLOT_TSPEC) == 0) {
      for (nt = 0; nt < 10; nt++) {
        double rv0 = ro_rnd(sr.key, RN_GAUSS + 2 * nt), rv1 = ro_rnd(sr.key, RN_GAUSS + 2 * nt + 1);
        double d = 2.0 * PI * rv0, cosp = cos(d), sinp = sin(d);
        if ((0. <= c->p->specjitter) & (c->p->specjitter < 1.)) rv1 = 1.0 - c->p->specjitter * rv1;
        d = rv1 <= FTINY ? 1.0 : sqrt(np->alpha2 * -log(rv1));
        for (k = 0; k < 3; k++) sr.rdir[k] = np->prdir[k] + d * (cosp * u[k] + sinp * v[k]);
        if (DOT(sr.rdir, np->rp->ron) >= -FTINY) continue;
        vnorm(sr.rdir);
        raytrace(c, &sr);
        for (k = 0; k < 3; k++) np->rp->rcol[k] += sr.rcol[k] * sr.rcoef[k];
        break;
      }
    }
  }
}

static int m_normal(Ctx *c, const OObj *m, const double *fa, int otype, ORay *r) {
  ND nd;
  double fest = 0, sctmp[3];
  int k;
  if (r->crtype & SHADOW && otype != M_TRANS) return 1;
  if (r->rod < 0.0) {
    if (!c->p->backvis) { raytrans(c, r); return 1; }
    r->rod = -r->rod;
    for (k = 0; k < 3; k++) r->ron[k] = -r->ron[k];
  }
  (void)m;
  memset(&nd, 0, sizeof nd);
  nd.rp = r;
  for (k = 0; k < 3; k++) nd.mcolor[k] = fa[k];
  nd.alpha2 = fa[4] * fa[4];
  if (nd.alpha2 <= FTINY) nd.specfl |= SP_PURE;
  memcpy(nd.pnorm, r->ron, sizeof nd.pnorm);
  nd.pdot = r->rod;
  if (r->ro >= 0 && (c->s->o[r->ro].type == T_POLY || c->s->o[r->ro].type == T_RING)) nd.specfl |= SP_FLAT;
  if (nd.pdot < .001) nd.pdot = .001;
  nd.rspec = fa[3];
  if (nd.specfl & SP_PURE && nd.rspec >= FRESTHRESH) {
    fest = FRESNE(nd.pdot);
    nd.rspec += fest * (1. - nd.rspec);
  }
  memcpy(nd.prdir, r->rdir, sizeof nd.prdir);
  if (otype == M_TRANS) {
    nd.trans = fa[5] * (1.0 - nd.rspec);
    nd.tspec = nd.trans * fa[6];
    nd.tdiff = nd.trans - nd.tspec;
    if (nd.tspec > FTINY) {
      nd.specfl |= SP_TRAN;
      if (!(nd.specfl & SP_PURE) && c->p->specthresh >= nd.tspec - FTINY) nd.specfl |= SP_TBLT;
    }
  }
  nd.rdiff = 1.0 - nd.trans - nd.rspec;
  if ((nd.specfl & (SP_TRAN | SP_PURE | SP_TBLT)) == (SP_TRAN | SP_PURE)) {
    ORay lr;
    double tc[3];
    for (k = 0; k < 3; k++) tc[k] = nd.mcolor[k] * nd.tspec;
    if (rayorigin(c, &lr, TRANS, r, tc, SLOT_TRANS) == 0) {
      memcpy(lr.rdir, nd.prdir, sizeof lr.rdir);
      raytrace(c, &lr);
      for (k = 0; k < 3; k++) r->rcol[k] += lr.rcol[k] * lr.rcoef[k];
    }
  }
  if (r->crtype & SHADOW) return 1;
  if (nd.rspec > FTINY) {
    nd.specfl |= SP_REFL;
    if (otype != M_METAL) {
      nd.scolor[0] = nd.scolor[1] = nd.scolor[2] = nd.rspec;
    } else if (fest > FTINY) {
      double d = fa[3] * (1. - fest);
      for (k = 0; k < 3; k++) nd.scolor[k] = fest + nd.mcolor[k] * d;
    } else {
      for (k = 0; k < 3; k++) nd.scolor[k] = nd.mcolor[k] * nd.rspec;
    }
    if (!(nd.specfl & SP_PURE) && c->p->specthresh >= nd.rspec - FTINY) nd.specfl |= SP_RBLT;
  }
  if ((nd.specfl & (SP_REFL | SP_PURE | SP_RBLT)) == (SP_REFL | SP_PURE)) {
    ORay lr;
    if (rayorigin(c, &lr, REFLECTED, r, nd.scolor, SLOT_REFL) == 0) {
      for (k = 0; k < 3; k++) lr.rdir[k] = r->rdir[k] + nd.pnorm[k] * 2. * nd.pdot;
      vnorm(lr.rdir);
      raytrace(c, &lr);
      for (k = 0; k < 3; k++) r->rcol[k] += lr.rcol[k] * lr.rcoef[k];
    }
  }
  if (nd.specfl & SP_PURE && nd.rdiff <= FTINY && nd.tdiff <= FTINY) return 1;
  if (!(nd.specfl & SP_PURE)) gaussamp(c, &nd);
  if (nd.rdiff > FTINY) {
    for (k = 0; k < 3; k++) sctmp[k] = nd.mcolor[k] * nd.rdiff + ((nd.specfl & SP_RBLT) ? nd.scolor[k] : 0);
    multambient(c, sctmp, r, nd.pnorm);
    for (k = 0; k < 3; k++) r->rcol[k] += sctmp[k];
  }
  if (nd.tdiff > FTINY) {
    double bnorm[3] = {-nd.pnorm[0], -nd.pnorm[1], -nd.pnorm[2]};
    for (k = 0; k < 3; k++) sctmp[k] = nd.mcolor[k] * ((nd.specfl & SP_TBLT) ? nd.trans : nd.tdiff);
    multambient(c, sctmp, r, bnorm);
    for (k = 0; k < 3; k++) r->rcol[k] += sctmp[k];
  }
  direct(c, r, &nd);
  return 1;
}

static int m_glass(Ctx *c, const OObj *m, ORay *r) {
  double mc[3], rindex = m->nf == 4 ? m->f[3] : 1.52, pdot, pnorm[3], cos2, r1e, r1m, sc[3];
  int hastrans, k;
  ORay p;
  if (!c->p->backvis && r->rod <= 0.0) { raytrans(c, r); return 1; }
  for (k = 0; k < 3; k++) mc[k] = m->f[k];
  if ((hastrans = cmaxd(mc) > 1e-15)) {
    for (k = 0; k < 3; k++) if (mc[k] < 1e-15) mc[k] = 1e-15;
  } else if (r->crtype & SHADOW) return 1;
  if (r->rod < 0.0) { r->rod = -r->rod; for (k = 0; k < 3; k++) r->ron[k] = -r->ron[k]; }
  memcpy(pnorm, r->ron, sizeof pnorm);
  pdot = r->rod;
  cos2 = sqrt((1.0 - 1.0 / (rindex * rindex)) + pdot * pdot / (rindex * rindex));
  if (hastrans) for (k = 0; k < 3; k++) mc[k] = pow(mc[k], 1.0 / cos2);
  r1e = (pdot - rindex * cos2) / (pdot + rindex * cos2);
  r1e *= r1e;
  r1m = (1.0 / pdot - rindex / cos2) / (1.0 / pdot + rindex / cos2);
  r1m *= r1m;
  if (hastrans) {
    for (k = 0; k < 3; k++) {
      double d = mc[k];
      sc[k] = .5 * (1.0 - r1e) * (1.0 - r1e) * d / (1.0 - r1e * r1e * d * d) + .5 * (1.0 - r1m) * (1.0 - r1m) * d / (1.0 - r1m * r1m * d * d);
    }
    if (rayorigin(c, &p, TRANS, r, sc, SLOT_TRANS) == 0) {
      memcpy(p.rdir, r->rdir, sizeof p.rdir);
      raytrace(c, &p);
      for (k = 0; k < 3; k++) r->rcol[k] += p.rcol[k] * p.rcoef[k];
    }
  }
  if (r->crtype & SHADOW) return 1;
  for (k = 0; k < 3; k++) {
    double d = mc[k] * mc[k];
    sc[k] = .5 * r1e * (1.0 + (1.0 - 2.0 * r1e) * d) / (1.0 - r1e * r1e * d) + .5 * r1m * (1.0 + (1.0 - 2.0 * r1m) * d) / (1.0 - r1m * r1m * d);
  }
  if (rayorigin(c, &p, REFLECTED, r, sc, SLOT_REFL) == 0) {
    for (k = 0; k < 3; k++) p.rdir[k] = r->rdir[k] + pnorm[k] * 2. * pdot;
    vnorm(p.rdir);
    raytrace(c, &p);
    for (k = 0; k < 3; k++) r->rcol[k] += p.rcol[k] * p.rcoef[k];
  }
  return 1;
}

/* rayshade() with material object `mat`; returns 0 when there is none */
static int rayshade(Ctx *c, ORay *r, int mat) {
  const OObj *m;
  int tst_irrad = c->irrad && !(r->crtype & ~(PRIMARY | TRANS));
  if (mat < 0) return 0;
  m = &c->s->o[mat];
  if (tst_irrad && m->type != M_LIGHT && m->type != M_GLOW && m->type != M_ILLUM) {
    if (m->type == M_TRANS || m->type == M_GLASS) { raytrans(c, r); return 1; }
    return m_normal(c, NULL, LAMBF, M_PLASTIC, r);
  }
  switch (m->type) {
    case M_PLASTIC: case M_METAL: case M_TRANS: return m_normal(c, m, m->f, m->type, r);
    case M_GLASS: return m_glass(c, m, r);
    default: return m_light(c, m, r);
  }
}

static void raytrace(Ctx *c, ORay *r) {
  if (localhit(c, r)) {
    if (!rayshade(c, r, c->s->o[r->ro].mat)) raytrans(c, r);
  } else {
    int si = sourcehit(c, r);
    if (si >= 0) {
      r->ro = c->s->s[si].obj;
      r->rod = 1.0;
      r->rot = FHUGE;
      rayshade(c, r, c->s->s[si].mat);
    }
  }
  trace_contrib(c, r);
}

/* one sensor record value into out (ncols doubles) */
static void eval_ray(Ctx *c, const double *ry, long ridx, long sensor, int imm_irrad) {
  ORay r;
  double org[3], dir[3];
  int k;
  memset(&r, 0, sizeof r);
  for (k = 0; k < 3; k++) { org[k] = ry[k]; dir[k] = ry[3 + k]; }
  vnorm(dir);
  r.crtype = r.rtype = PRIMARY;
  r.rsrc = -1;
  r.rweight = 1.0f;
  r.rcoef[0] = r.rcoef[1] = r.rcoef[2] = 1.0;
  r.key = ro_sensor_key(c->p->seed, (unsigned long long)ridx);
  r.sensor = sensor;
  r.ro = -1;
  if (imm_irrad) {
    for (k = 0; k < 3; k++) {
      /* rtcompute() and rayirrad() (rt/rtrace.c:416-449): origin 1.1e-4
       * out along the normal, the pretended hit 1e-5 back from there */
      r.rorg[k] = org[k] + dir[k] * 1.1e-4;
      r.rdir[k] = -dir[k];
      r.rop[k] = r.rorg[k] + r.rdir[k] * 1e-5;
      r.ron[k] = dir[k];
    }
    r.rot = 1e-5;
    r.rod = 1.0;
    m_normal(c, NULL, LAMBF, M_PLASTIC, &r);
  } else {
    memcpy(r.rorg, org, sizeof org);
    memcpy(r.rdir, dir, sizeof dir);
    raytrace(c, &r);
  }
  if (c->mode == 0)
    for (k = 0; k < 3; k++) c->out[k] += r.rcol[k];
}

int ro_trace(const ro_scene *s, const ro_params *p, const double *rays, long n, int flags, double *out,
             unsigned long long *nrays, char *err, int errlen) {
  Ctx c;
  long i;
  memset(&c, 0, sizeof c);
  c.s = s;
  c.p = p;
  c.mode = 0;
  c.irrad = (flags & 16) != 0;
  c.modslot = (int *)calloc(s->n + 1, sizeof(int));
  c.cache = (OCache *)calloc(s->ns + 1, sizeof(OCache));
  for (i = 0; i < n; i++) {
    c.out = out + 3 * i;
    out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = 0;
    eval_ray(&c, rays + 6 * i, i, i, (flags & 8) != 0);
  }
  free(c.modslot);
  cache_free(&c);
  if (nrays) *nrays = c.nrays;
  (void)err; (void)errlen;
  return 0;
}

int ro_contrib(const ro_scene *s, const ro_params *p, const double *rays, long nrec, int accum, int flags,
               const ro_mod *mods, int nmods, double *out, unsigned long long *nrays, char *err, int errlen) {
  Ctx c;
  long i, j;
  int k, col = 0;
  memset(&c, 0, sizeof c);
  c.s = s;
  c.p = p;
  c.mode = 1;
  c.irrad = (flags & 16) != 0;
  c.mods = mods;
  c.nmods = nmods;
  c.col0 = (int *)calloc(nmods + 1, sizeof(int));
  for (k = 0; k < nmods; k++) { c.col0[k] = col; col += mods[k].nbins; }
  c.ncols = 3 * col;
  c.cache = (OCache *)calloc(s->ns + 1, sizeof(OCache));
  c.modslot = (int *)malloc(sizeof(int) * (s->n + 1));
  for (i = 0; i < s->n; i++) {
    c.modslot[i] = -1;
    if (s->o[i].mod >= 0)
      for (k = 0; k < nmods; k++)
        if (!strcmp(s->o[s->o[i].mod].name, mods[k].name)) c.modslot[i] = k;
  }
  for (i = 0; i < nrec; i++) {
    c.out = out + (size_t)i * c.ncols;
    memset(c.out, 0, sizeof(double) * c.ncols);
    for (j = 0; j < accum; j++) eval_ray(&c, rays + 6 * (i * accum + j), i * accum + j, i, (flags & 8) != 0);
    if (accum > 1)
      for (k = 0; k < c.ncols; k++) c.out[k] /= accum;
  }
  free(c.modslot);
  free(c.col0);
  cache_free(&c);
  if (nrays) *nrays = c.nrays;
  (void)err; (void)errlen;
  return 0;
}
