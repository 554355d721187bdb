// Wavefront renderer for the rtrace / rcontrib ray tree on CDNA4 (gfx950).
//
// One fused kernel per generation of the tree: each thread takes one child of
// a spawn record (binary search over the exclusive scan of child counts),
// applies the spawn rules of rayorigin() (rt/raytrace.c:40-135), walks the BVH
// for the closest hit, shades the hit and appends the spawn records of the
// next generation.  Spawn records are allocated with one atomic per wavefront
// (ballot + mbcnt), emission is added into the sensor's output with FP64
// atomics.  The BVH stack lives in LDS, column-major per lane.
#include <hip/hip_runtime.h>
#include "rg_shade.hpp"
#include "rg_render.hpp"

namespace rg {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int BLOCK = 128;       // shading and seed kernels
constexpr int TBLOCK = 256;      // traversal kernel
#ifndef RG_LSTACK
#define RG_LSTACK 16
#endif
constexpr int LSTACK = RG_LSTACK;                    // per-lane traversal stack entries held in LDS
constexpr int PSTACK = kTraceStackEntries - LSTACK;  // further entries in private (scratch) memory

// ---- wave-level helpers -----------------------------------------------------
__device__ __forceinline__ uint32_t lane_rank(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}
// Allocate one slot per active lane with a single atomic per wavefront.
__device__ __forceinline__ uint32_t wave_alloc(uint32_t* ctr) {
  unsigned long long m = __ballot(1);
  uint32_t leader = __ffsll((long long)m) - 1;
  uint32_t base = 0;
  if (__lane_id() == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  return base + lane_rank(m);
}
__device__ __forceinline__ void wave_count(unsigned long long* ctr) {
  unsigned long long m = __ballot(1);
  uint32_t leader = __ffsll((long long)m) - 1;
  if (__lane_id() == leader) atomicAdd(ctr, (unsigned long long)__popcll(m));
}

struct Sink {
  Spawn* buf;
  uint32_t* count;
  uint32_t cap;
  uint32_t* err;
  PathRec* pbuf;       // adjoint replay: path record of each spawn (else null)
  uint32_t* counts;    // child count of each spawn, compact (the scan's input)
};

// Append a spawn; in adjoint replay also its path: the parent's plus `entry`
// (none: the parent's vertex is not a differentiated one).
__device__ __forceinline__ void emit(const Sink& sk, const Spawn& s, const RayS& r, const PathAdd& entry) {
  uint32_t i = wave_alloc(sk.count);
  if (i < sk.cap) {
    sk.buf[i] = s;
    sk.counts[i] = s.count;
    if (sk.pbuf) sk.pbuf[i] = path_push(r.path, entry);
  } else {
    atomicOr(sk.err, ERR_CAPACITY);
  }
}

__device__ __forceinline__ void emit(const Sink& sk, const Spawn& s, const RayS& r) { emit(sk, s, r, path_none()); }

// Bin accumulation.  With c.fix > 0 each channel is rounded to a multiple
// of 1 / c.fix and added as a 64-bit integer: integer sums do not depend on
// the order the deposits arrive in, so a record's value is the same bit for
// bit whatever the launch shape or scheduling.  A deposit beyond 2^60 / fix
// flags ERR_FIXED (the range the host chose from the record's bound).
__device__ __forceinline__ void deposit(const RunCfg& c, uint32_t sensor, int col, C3 v) {
  double* o = c.out + (size_t)sensor * c.ncols + col;
  if (c.fix > 0.0) {
    unsigned long long* q = (unsigned long long*)o;
    const double x[3] = {(double)v.r * c.fix, (double)v.g * c.fix, (double)v.b * c.fix};
    for (int k = 0; k < 3; ++k) {
      if (!(fabs(x[k]) < 0x1p60)) { atomicOr(c.err, ERR_FIXED); continue; }
      const long long n = llrint(x[k]);
      if (n) atomicAdd(q + k, (unsigned long long)n);
    }
    return;
  }
  unsafeAtomicAdd(o + 0, (double)v.r);
  unsafeAtomicAdd(o + 1, (double)v.g);
  unsafeAtomicAdd(o + 2, (double)v.b);
}

// ---- BVH closest hit (FP32 boxes, FP64 primitive tests) --------------------
// (lo - o) * inv as one FMA per plane: o * inv is precomputed per ray (the
// boxes are widened by far more than the extra rounding this allows).
__device__ __forceinline__ bool slab(const float* lo, const float* hi, float ox, float oy, float oz,
                                     float ix, float iy, float iz, float tmax, float& tn) {
  float t0x = fmaf(lo[0], ix, -ox), t1x = fmaf(hi[0], ix, -ox);
  float t0y = fmaf(lo[1], iy, -oy), t1y = fmaf(hi[1], iy, -oy);
  float t0z = fmaf(lo[2], iz, -oz), t1z = fmaf(hi[2], iz, -oz);
  float tnear = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
  float tfar = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
  tn = tnear;
  return (tnear <= tfar) & (tfar >= 0.0f) & (tnear <= tmax);
}

__device__ __forceinline__ uint64_t pack_entry(int32_t ref, float t) {
  return ((uint64_t)__float_as_uint(t) << 32) | (uint32_t)ref;
}

// Room-sized primitives, tested before the walk.  Polygons among the first
// eight are taken nearest plane first (plane distance from FP64 dot products,
// divided in FP32 only to order them); once the next estimate lies beyond
// the current hit by a margin far above the estimate's error, the FP64 test
// would reject it and every later one (t > hit + FTINY), so they are skipped.
// Planes the ray nearly parallels, and other shapes, get the full test.
// te: this lane's NF estimates, in LDS column-major (te[k * BS] for entry k).
constexpr int NBIG_ORDER = kBigPlanes;
template <int BS>
__device__ __forceinline__ void big_hits(const DevScene& S, V3 o, V3 d, HitRec& h, uint32_t& nprim, float* te) {
  constexpr int NF = NBIG_ORDER;
  const int nb = S.nbig;
  uint32_t full = 0;        // big primitives that get the full test unordered
  // all NF plane records requested at once (the table always holds NF)
  BigPlane bp[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) bp[k] = S.bigpl[k];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    te[k * BS] = INFINITY;
    if (k < nb) {
      // a zero normal (not a polygon) fails the test below: full test
      const double* g = bp[k].n;
      double rdot = -(d.x * g[0] + d.y * g[1] + d.z * g[2]);
      double num = o.x * g[0] + o.y * g[1] + o.z * g[2] - bp[k].off;
      bool f = true;
      if (fabs(rdot) > 4 * FTINY) {
        float t = (float)num / (float)rdot;
        if (t > -1e-3f) te[k * BS] = t;   // behind the origin: the FP64 test rejects t <= FTINY
        f = false;
      }
      if (f) full |= 1u << k;
    }
  }
  // one call site for the primitive test: unordered ones first, then the
  // ordered polygons nearest first
  int kk = 0;
  for (;;) {
    int pk = -1;
    if (full) {
      pk = __ffs(full) - 1;
      full &= full - 1;
    } else if (kk < nb - NF) {
      pk = NF + kk++;
    } else {
      float best = INFINITY;
      int bk = -1;
#pragma unroll 1
      for (int k = 0; k < NF; ++k) {
        const float tk = te[k * BS];
        if (tk < best) { best = tk; bk = k; }
      }
      if (bk < 0 || best > (float)(h.t + FTINY) * 1.0001f + 1e-4f) break;
      te[bk * BS] = INFINITY;
      pk = bk;
    }
    if (pk < NF) {
      // plane, geometry offset and vertex count from the table entry (one
      // load, where the record's own chain is three)
      const BigPlane B = S.bigpl[pk];
      if (B.kind == PK_POLYGON)
        hit_polygon(S, B.prim, v3(B.n[0], B.n[1], B.n[2]), B.off, S.geom + B.goff, B.nv, o, d, h);
      else
        hit_prim(S, B.prim, o, d, h);
    } else {
      hit_prim(S, S.big[pk], o, d, h);
    }
    ++nprim;
  }
}

// Stack: the top LSTACK entries of each lane live in LDS (column-major, one
// bank per lane); deeper entries spill to a private array that only deep
// subtrees touch.
template <int BS>
__device__ void closest_hit(const DevScene& S, V3 o, V3 d, HitRec& h, int32_t* stk, float* stn, float* tes,
                            uint32_t* err, unsigned long long* visits) {
  uint32_t nvis = 0, nprim = 0;
  h.t = FHUGE; h.rod = 1.0; h.prim = -1;
  // room-sized primitives first: their hit bounds the walk below
  big_hits<BS>(S, o, d, h, nprim, tes + threadIdx.x);
  if (S.nbvh == 0) goto done;
  if (S.root < 0) { hit_prim(S, S.leaf_prim[~S.root], o, d, h); ++nprim; goto done; }
  {
  const float ox = (float)o.x, oy = (float)o.y, oz = (float)o.z;
  auto inv = [](double v) -> float {
    float f = (float)v;
    if (fabsf(f) < 1e-20f) f = f < 0 ? -1e-20f : 1e-20f;
    return 1.0f / f;
  };
  const float ix = inv(d.x), iy = inv(d.y), iz = inv(d.z);
  const float oix = ox * ix, oiy = oy * iy, oiz = oz * iz;
  const int lane = threadIdx.x;
  // Deeper entries pack (ref, tnear) in 64 bits: one private store / load
  // each, and a different access type from the LDS pair above, so the two
  // paths are not merged into one flat access.
  uint64_t deep[PSTACK];
  int sp = 0;
  int node = S.root;
  const BVHNode* __restrict__ nodes = S.nodes;
  const int nbvh = S.nbvh;
  // Walk internal nodes until every lane of the wavefront holds a postponed
  // leaf (or is done), then run the FP64 primitive tests together: testing a
  // leaf as soon as one lane meets it would stall the other 63 lanes on it.
  // Child refs: >= 0 internal node, < 0 primitive ~ref, DONE: stack empty.
  constexpr int32_t DONE = INT32_MIN;
  auto pop = [&](float lim) -> int32_t {
    // skip subtrees whose box starts beyond the current closest hit
    while (sp > 0) {
      --sp;
      int32_t cand;
      float ct;
      if (sp < LSTACK) {
        cand = stk[sp * BS + lane];
        ct = stn[sp * BS + lane];
      } else {
        uint64_t e = deep[sp - LSTACK];
        cand = (int32_t)(uint32_t)e;
        ct = __uint_as_float((uint32_t)(e >> 32));
      }
      if (ct <= lim) return cand;
    }
    return DONE;
  };
  int32_t leaf = -1;             // postponed primitive, -1: none
  int iter = 0;                  // every internal node is entered at most once
  while (node != DONE) {
    while (node >= 0) {
      if (++iter > nbvh + 8) { atomicOr(err, ERR_STACK); node = DONE; break; }
      const BVHNode N = nodes[node];
      ++nvis;
      float tmax = (float)(h.t + 2 * FTINY) * 1.000001f;
      float tl, tr;
      bool hl = slab(N.lmin, N.lmax, oix, oiy, oiz, ix, iy, iz, tmax, tl);
      bool hr = slab(N.rmin, N.rmax, oix, oiy, oiz, ix, iy, iz, tmax, tr);
      if (hl & hr) {
        bool lfirst = tl <= tr;
        int32_t farc = lfirst ? N.right : N.left;
        float fart = lfirst ? tr : tl;
        if (sp < LSTACK) { stk[sp * BS + lane] = farc; stn[sp * BS + lane] = fart; }
        else if (sp < LSTACK + PSTACK) deep[sp - LSTACK] = pack_entry(farc, fart);
        else { atomicOr(err, ERR_STACK); node = DONE; break; }
        ++sp;
        node = lfirst ? N.left : N.right;
      } else if (hl) {
        node = N.left;
      } else if (hr) {
        node = N.right;
      } else {
        node = pop(tmax);
      }
      if ((node < 0) & (node != DONE) & (leaf < 0)) {
        leaf = ~node;
        node = pop((float)(h.t + 2 * FTINY) * 1.000001f);
      }
      if (__ballot(leaf < 0) == 0ull) break;
    }
    while (leaf >= 0) {
      hit_prim(S, leaf, o, d, h);
      ++nprim;
      leaf = -1;
      if ((node < 0) & (node != DONE)) {
        leaf = ~node;
        node = pop((float)(h.t + 2 * FTINY) * 1.000001f);
      }
    }
  }
  }
done:
  if (visits) {
    atomicAdd(visits, (unsigned long long)nvis);
    atomicAdd(visits + 1, (unsigned long long)nprim);
  }
}

// ---- spawn helpers ----------------------------------------------------------
// Child ray of type rt with own coefficient rc (nullptr: rayorigin(...,NULL)).
__device__ __forceinline__ Spawn make_child(const RayS& r, V3 org, uint16_t rt, const C3* rc, uint64_t slot, uint8_t flags) {
  Spawn s;
  s.org[0] = org.x; s.org[1] = org.y; s.org[2] = org.z;
  s.d0[0] = s.d0[1] = s.d0[2] = 0.f;
  s.d1[0] = s.d1[1] = s.d1[2] = 0.f;
  double rw = 1.0;
  C3 own = c3(1, 1, 1);
  if (rc) {
    own = *rc;
    rw = cmax(own);
    if (rw > 1.0) rw = 1.0;
    flags |= SPF_RR;
  }
  C3 cum = r.coef * own;
  s.coef[0] = cum.r; s.coef[1] = cum.g; s.coef[2] = cum.b;
  s.cw = (float)(r.w * rw);
  s.crc = (float)cmax(own);
  s.sensor = r.sensor;
  s.count = 1;
  s.ns = 0;
  s.rtype = rt;
  s.crtype = r.crtype | rt;
  if (rt & RAYREFL) { s.rlvl = (uint8_t)(r.rlvl + 1); s.rsrc = -1; }
  else { s.rlvl = (uint8_t)r.rlvl; s.rsrc = (int16_t)r.rsrc; }
  s.adepth = (uint8_t)r.adepth;
  s.flags = flags;
  s.scene = (uint16_t)r.scene;
  s.key = child_key(r.key, slot);
  return s;
}

__device__ __forceinline__ void set_dir(Spawn& s, V3 d) {
  s.d0[0] = (float)d.x; s.d0[1] = (float)d.y; s.d0[2] = (float)d.z;
}

// rayorigin() admission for a child built from a spawn record.
// Returns false when the child is not traced.
__device__ __forceinline__ bool admit(const RunCfg& c, RayS& r, uint8_t flags) {
  if (!(r.w > 0.0f)) return (flags & SPF_FORCE) != 0;
  if (r.crtype & SHADOW) return true;
  const Params& p = c.p;
  if ((p.maxdepth <= 0) & ((flags & SPF_RR) != 0)) {
    if ((p.maxdepth < 0) & (r.rlvl > -p.maxdepth)) return false;
    if (r.w >= p.minweight) return true;
    if (rnd(r.key, RN_RR) > r.w / p.minweight) return false;
    double sc = p.minweight / r.w;
    r.coef = r.coef * sc;
    r.crc = (float)(r.crc * sc);
    r.w = (float)p.minweight;
    if (c.adj) path_roulette(r.path);
    return true;
  }
  if (flags & SPF_FORCE) return true;
  int md = p.maxdepth < 0 ? -p.maxdepth : p.maxdepth;
  return (r.w >= p.minweight) & (r.rlvl <= md);
}

// ---- light sources -----------------------------------------------------------
// Object identity of what a ray hit: prim index, or -(2+source) for distant.
__device__ __forceinline__ int src_obj(const Source& s, int si) { return s.prim >= 0 ? s.prim : -(2 + si); }

// m_light() rules (rt/source.c:683-793).  Returns emission in *em, or sets
// *pass for a transparent pass-through; false when the ray is discarded.
__device__ __forceinline__ bool light_rules(const DevScene& S, const RunCfg& c, const RayS& r, const Material& m, int hitobj,
                            double rot, double rod, C3* em, bool* pass) {
  *pass = false;
  auto distglow = [&](double dist) {
    return m.kind == MK_GLOW && m.f[3] >= -FTINY && dist > m.f[3];
  };
  // badcomponent
  if ((r.crtype & (AMBIENT | SPECULAR)) && !((r.crtype & SHADOW) || rod < 0.0 || distglow(rot))) return false;
  // wrongsource
  if (r.rsrc >= 0) {
    const Source& sr = S.srcs[r.rsrc];
    if (src_obj(sr, r.rsrc) != hitobj) {
      bool illumblock = !(sr.flags & SVIRTUAL) && rod > 0.0 &&
                        (S.mats[sr.mat].kind == MK_ILLUM || S.mats[sr.mat].kind == MK_GLOW);
      if (m.kind != MK_ILLUM || illumblock) return false;
    }
  }
  // passillum
  if (m.kind == MK_ILLUM && (r.rsrc < 0 || src_obj(S.srcs[r.rsrc], r.rsrc) != hitobj)) {
    *pass = true;
    return true;
  }
  if (rod < 0.0) {           // behind the emitter: black, or pass through with -bv-
    *em = c3(0, 0, 0);
    if (!c.p.backvis) *pass = true;
    return true;
  }
  *em = c3((float)m.f[0], (float)m.f[1], (float)m.f[2]);
  return true;
}

// sourcehit() of rt/source.c:317-377 for a ray that left the scene.
__device__ __forceinline__ int source_hit(const DevScene& S, const RayS& r) {
  int glowsrc = -1, transrc = -1;
  int first = 0, last = S.nsrcs - 1;
  if (r.rsrc >= 0) first = last = r.rsrc;
  for (int i = first; i <= last; ++i) {
    const Source& s = S.srcs[i];
    if ((s.flags & (SDISTANT | SVIRTUAL)) != SDISTANT) continue;
    if (2. * PI_D * (1. - dot(load3(s.sloc), r.d)) > s.ss2) continue;
    if (i == r.rsrc) return i;
    if (s.flags & SSKIP) { if (glowsrc < 0) glowsrc = i; continue; }
    const Material& m = S.mats[s.mat];
    if (m.kind == MK_ILLUM && m.illum_alt < 0) { if (transrc < 0) transrc = i; continue; }
    return i;
  }
  if (transrc >= 0 && (r.crtype & (AMBIENT | SPECULAR))) return -1;
  return glowsrc;
}

#define FRESNE(ci) (exp(-5.85 * (ci)) - 0.00202943064)
constexpr double FRESTHRESH = 0.017999;

// Own coefficient of one differentiated vertex, own[c], and its derivative
// dOwn[c][q] with respect to the material's real arguments q (red, green,
// blue, specularity, roughness and, for trans, transmissivity and
// transmitted specularity), as m_normal() forms them (rt/normal.c:177-360):
//   rspec = spec + fest (1 - spec)           (fest: Fresnel estimate, pure specular only)
//   trans = f5 (1 - rspec), tspec = trans f6, tdiff = trans - tspec   (trans only)
//   rdiff = 1 - trans - rspec
//   scolor = rspec (plastic, trans); metal: rho rspec, or fest + rho spec (1 - fest) when fest > 0.
// Shadow-ray entries carry dirnorm()'s factors with these colour terms taken
// out (rt/normal.c:72-173): the reflection side is rho rdiff a + scolor b,
// the transmission side rho tdiff a + rho tspec b, and b's roughness
// derivative is b g.
__device__ __forceinline__ void vertex_own(const Material& M, uint32_t kind, double fest, const uint32_t* pay,
                                           double own[3], double dOwn[3][ADJ_COLS]) {
  const double sp = M.f[3];
  const double rspec = sp + fest * (1.0 - sp);
  double d_rspec[ADJ_COLS] = {0, 0, 0, 1.0 - fest, 0, 0, 0};
  double trans = 0, tspec = 0, tdiff = 0;
  double d_trans[ADJ_COLS] = {}, d_tspec[ADJ_COLS] = {}, d_tdiff[ADJ_COLS] = {}, d_rdiff[ADJ_COLS];
  if (M.kind == MK_TRANS) {
    trans = M.f[5] * (1.0 - rspec);
    d_trans[3] = -M.f[5] * (1.0 - fest);
    d_trans[5] = 1.0 - rspec;
    tspec = trans * M.f[6];
    for (int q = 0; q < ADJ_COLS; ++q) d_tspec[q] = M.f[6] * d_trans[q];
    d_tspec[6] += trans;
    tdiff = trans - tspec;
    for (int q = 0; q < ADJ_COLS; ++q) d_tdiff[q] = d_trans[q] - d_tspec[q];
  }
  const double rdiff = 1.0 - trans - rspec;
  for (int q = 0; q < ADJ_COLS; ++q) d_rdiff[q] = -d_trans[q] - d_rspec[q];
  const double a = kind >= PV_DIR_R ? (double)__uint_as_float(pay[0]) : 0.0;
  const double b = kind >= PV_DIR_R ? (double)__uint_as_float(pay[1]) : 0.0;
  const double g = kind >= PV_DIR_R ? (double)__uint_as_float(pay[2]) : 0.0;
  for (int ch = 0; ch < 3; ++ch) {
    const double rho = M.f[ch];
    // scolor and its derivative
    double sc, dsc[ADJ_COLS] = {};
    if (M.kind != MK_METAL) {
      sc = rspec;
      dsc[3] = d_rspec[3];
    } else if (fest > FTINY) {
      sc = fest + rho * sp * (1.0 - fest);
      dsc[ch] = sp * (1.0 - fest);
      dsc[3] = rho * (1.0 - fest);
    } else {
      sc = rho * rspec;
      dsc[ch] = rspec;
      dsc[3] = rho * d_rspec[3];
    }
    // rho * X for the X the kind uses (and a second term on some kinds)
    double x = 0, dx[ADJ_COLS] = {}, x2 = 0, dx2[ADJ_COLS] = {};
    double wx = 1.0, wsc = 0.0, wx2 = 0.0;       // own = wx rho x + wsc scolor + wx2 rho x2
    switch (kind) {
      case PV_AMB: x = rdiff; for (int q = 0; q < ADJ_COLS; ++q) dx[q] = d_rdiff[q]; break;
      case PV_AMB_RBLT: x = rdiff; for (int q = 0; q < ADJ_COLS; ++q) dx[q] = d_rdiff[q]; wsc = 1.0; break;
      case PV_SPEC:
      case PV_GSPEC: wx = 0.0; wsc = 1.0; break;
      case PV_TSPEC:
      case PV_GTRANS: x = tspec; for (int q = 0; q < ADJ_COLS; ++q) dx[q] = d_tspec[q]; break;
      case PV_AMB_BACK: x = tdiff; for (int q = 0; q < ADJ_COLS; ++q) dx[q] = d_tdiff[q]; break;
      case PV_AMB_BACK_TBLT: x = trans; for (int q = 0; q < ADJ_COLS; ++q) dx[q] = d_trans[q]; break;
      case PV_DIR_R: x = rdiff; for (int q = 0; q < ADJ_COLS; ++q) dx[q] = d_rdiff[q]; wx = a; wsc = b; break;
      default:       // PV_DIR_T
        x = tdiff; for (int q = 0; q < ADJ_COLS; ++q) dx[q] = d_tdiff[q]; wx = a;
        x2 = tspec; for (int q = 0; q < ADJ_COLS; ++q) dx2[q] = d_tspec[q]; wx2 = b;
        break;
    }
    own[ch] = wx * rho * x + wsc * sc + wx2 * rho * x2;
    for (int q = 0; q < ADJ_COLS; ++q) dOwn[ch][q] = wx * rho * dx[q] + wsc * dsc[q] + wx2 * rho * dx2[q];
    dOwn[ch][ch] += wx * x + wx2 * x2;
    if (kind == PV_DIR_R) dOwn[ch][4] += b * sc * g;
    else if (kind == PV_DIR_T) dOwn[ch][4] += b * rho * tspec * g;
  }
}

// Path-replay adjoint of one binned deposit.  The deposit is
//   contr_c = L-independent path coefficient; loss weight w_c = dLoss/dE_c * L_c * contr_c,
// and contr_c is a product of the own coefficients of the path's vertices,
// times minweight / weight at every Russian roulette survival.  For each
// differentiated vertex the derivative of log(own_c) follows from
// vertex_own(); a roulette survival divides by the largest channel of each
// own coefficient since the previous reset, which subtracts that channel's
// derivative.  Sampling decisions (strata counts, roulette) are held fixed.
// Gaussian lobe samples (gaussamp(), rt/normal.c:364-495) draw the lobe's
// tan^2 as alpha^2 u with u exponential, so a sample's density moves with
// the roughness alpha: their entries carry the score d log p / d alpha =
// (2 / alpha)(u - 1) of every trial up to the accepted one, which makes the
// roughness derivative of the estimator's expectation exact in the mean
// (likelihood-ratio form); the lobe directions themselves are held fixed.
__device__ __forceinline__ void adjoint_deposit(const DevScene& S, const RunCfg& c, const RayS& r, int col, C3 contr) {
  if (r.path.hdr & 1024u) { atomicOr(c.err, ERR_PATH); return; }
  const double* aw = c.adj_w + 3 * (size_t)r.sensor;
  const double* L = c.adj_L + col;
  const double w[3] = {aw[0] * L[0] * contr.r, aw[1] * L[1] * contr.g, aw[2] * L[2] * contr.b};
  if (w[0] == 0.0 && w[1] == 0.0 && w[2] == 0.0) return;
  const uint32_t n = r.path.hdr & 31u;
  uint32_t i = 0;
  while (i < n) {
    const uint32_t e = r.path.e[i];
    const int mi = (int)(e & 0xfffu);
    const uint32_t kind = (e >> 12) & 15u;
    const int np = path_payload(kind);
    const uint32_t* pay = &r.path.e[i + 1 < (uint32_t)PATH_WORDS ? i + 1 : i];
    i += 1u + (uint32_t)np;
    const int row = c.grad_row[mi];
    if (row < 0) continue;
    const Material& M = S.mats[mi];
    const double fest = (double)(e >> 17) * (1.0 / 32767.0);
    double own[3], dOwn[3][ADJ_COLS], D[3][ADJ_COLS];
    vertex_own(M, kind, fest, pay, own, dOwn);
    for (int ch = 0; ch < 3; ++ch)
      for (int q = 0; q < ADJ_COLS; ++q) D[ch][q] = own[ch] > 0 ? dOwn[ch][q] / own[ch] : 0.0;
    if (e & PV_RR) {                          // roulette divided by the largest channel
      int cm = 0;
      if ((float)own[1] > (float)own[cm]) cm = 1;
      if ((float)own[2] > (float)own[cm]) cm = 2;
      double Dm[ADJ_COLS];
      for (int q = 0; q < ADJ_COLS; ++q) Dm[q] = D[cm][q];
      for (int ch = 0; ch < 3; ++ch)
        for (int q = 0; q < ADJ_COLS; ++q) D[ch][q] -= Dm[q];
    }
    if (kind == PV_GSPEC || kind == PV_GTRANS) {
      const double sc = (double)__uint_as_float(pay[0]);
      for (int ch = 0; ch < 3; ++ch) D[ch][4] += sc;
    }
    double* g = c.grad + ADJ_COLS * (size_t)row;
    for (int q = 0; q < ADJ_COLS; ++q) {
      double v = w[0] * D[0][q] + w[1] * D[1][q] + w[2] * D[2][q];
      if (v != 0.0) unsafeAtomicAdd(g + q, v);
    }
  }
}

// rcontrib's trace callback (rt/rcontrib.c:272-316) in coefficient mode.
__device__ __forceinline__ void trace_contrib(const DevScene& S, const RunCfg& c, const RayS& r, int modslot, int hitobj,
                              bool zeroed, C3 rcol) {
  if (c.mode != 1 || modslot < 0 || zeroed) return;
  if (r.rsrc >= 0 && src_obj(S.srcs[r.rsrc], r.rsrc) != hitobj) return;
  C3 contr = r.coef;
  if (c.p.contrib) {
    if (!((r.crc * rcol.r > FTINY) || (r.crc * rcol.g > FTINY) || (r.crc * rcol.b > FTINY))) return;
    contr = contr * rcol;
  } else if (r.crc <= FTINY) {
    return;
  }
  const BinSpec& b = c.bins[modslot];
  int bn = eval_bin(r.d, b);
  if (bn < 0) return;
  if (c.adj) adjoint_deposit(S, c, r, 3 * (b.col0 + bn), contr);
  else deposit(c, r.sensor, 3 * (b.col0 + bn), contr);
}

// ---- materials ------------------------------------------------------------
struct HitS {
  V3 rop, ron;     // hit point, normal (flipped by the material when needed)
  V3 pert;         // normal perturbation (mesh vertex normals), zero if none
  double rod;      // -DOT(rdir, ron)
  double rot;
  int prim;        // -1 for the irradiance pseudo-hit
  int flat;
  int phong;       // modifier named "Phong": perturbation not applied to transmission
};

// flipsurface() of rt/raytrace.c:519-531
__device__ __forceinline__ void flip_surface(HitS& hs) {
  hs.rod = -hs.rod;
  hs.ron = -hs.ron;
  hs.pert = -hs.pert;
}

// raynormal() of rt/raytrace.c:446-480: perturbed normal, kept on the ray's
// side of the surface
__device__ __forceinline__ double ray_normal(V3& norm, const HitS& hs, V3 rdir) {
  norm = hs.ron + hs.pert;
  double l2 = dot(norm, norm);
  if (l2 == 0.0) {
    norm = hs.ron;
    return hs.rod;
  }
  norm = norm * (1.0 / sqrt(l2));
  double newdot = -dot(norm, rdir);
  if ((newdot > 0.0) != (hs.rod > 0.0)) {
    norm = norm + rdir * (2.0 * newdot);
    newdot = -newdot;
  }
  return newdot;
}

enum { SP_REFL = 01, SP_TRAN = 02, SP_PURE = 04, SP_FLAT = 010, SP_RBLT = 020, SP_TBLT = 040 };

struct NormDat {
  int specfl;
  C3 mcolor, scolor;
  V3 prdir, pnorm;
  double alpha2, rdiff, rspec, trans, tdiff, tspec, pdot;
  double rough;     // the material's roughness argument (alpha2 = rough^2)
};

// dirnorm()'s factors with the material's colour terms taken out, for the
// adjoint: side 0 (reflection) or 1 (transmission), diffuse factor a, lobe
// factor b and d log(lobe) / d roughness g (see vertex_own()).
struct DirParts {
  int side;
  double a, b, g;
};

// dirnorm() of rt/normal.c:72-173
__device__ __forceinline__ C3 dirnorm(const NormDat& np, const RunCfg& c, V3 rdir, V3 ldir, double omega,
                                      DirParts* pt = nullptr) {
  C3 scval = c3(0, 0, 0);
  double ldot = dot(np.pnorm, ldir);
  if (pt) { pt->side = ldot < 0.0; pt->a = pt->b = pt->g = 0.0; }
  if (ldot < 0.0 ? np.trans <= FTINY : np.trans >= 1.0 - FTINY) return scval;
  double lrdiff = np.rdiff, ltdiff = np.tdiff, kf = 1.0;
  if ((np.specfl & SP_PURE) && np.rspec >= FRESTHRESH && ((lrdiff > FTINY) | (ltdiff > FTINY))) {
    double dtmp = 1. - FRESNE(fabs(ldot));
    lrdiff *= dtmp;
    ltdiff *= dtmp;
    kf = dtmp;
  }
  if ((ldot > FTINY) & (lrdiff > FTINY)) {
    scval = scval + np.mcolor * (ldot * omega * lrdiff * (1.0 / PI_D));
    if (pt) pt->a = ldot * omega * kf * (1.0 / PI_D);
  }
  if ((ldot < -FTINY) & (ltdiff > FTINY)) {
    scval = scval + np.mcolor * (-ldot * omega * ltdiff * (1.0 / PI_D));
    if (pt) pt->a = -ldot * omega * kf * (1.0 / PI_D);
  }
  if ((ldot > FTINY) & ((np.specfl & (SP_REFL | SP_PURE)) == SP_REFL)) {
    double dtmp = np.alpha2;
    if (np.specfl & SP_FLAT) dtmp += (1. - c.p.dstrsrc) * omega * (0.25 / PI_D);
    V3 vtmp = ldir - rdir;
    double d2 = dot(vtmp, np.pnorm);
    d2 *= d2;
    double d3 = dot(vtmp, vtmp);
    double d4 = (d3 - d2) / d2;
    const double aw = dtmp;
    dtmp = exp(-d4 / dtmp) * d3 / (PI_D * d2 * d2 * dtmp);
    if (dtmp > FTINY) {
      scval = scval + np.scolor * (dtmp * ldot * omega);
      // d log(lobe) / d alpha2 = d4 / aw^2 - 1 / aw; d alpha2 / d roughness = 2 roughness
      if (pt) { pt->b = dtmp * ldot * omega; pt->g = 2.0 * np.rough * (d4 / (aw * aw) - 1.0 / aw); }
    }
  }
  if ((ldot < -FTINY) & ((np.specfl & (SP_TRAN | SP_PURE)) == SP_TRAN)) {
    double dtmp = np.alpha2 + omega * (1.0 / PI_D);
    const double bw = dtmp, cw = 2. * dot(np.prdir, ldir) - 2.;
    dtmp = exp(cw / dtmp) / (PI_D * dtmp);
    if (dtmp > FTINY) {
      scval = scval + np.mcolor * (dtmp * np.tspec * omega * sqrt(-ldot / np.pdot));
      if (pt) { pt->b = dtmp * omega * sqrt(-ldot / np.pdot); pt->g = 2.0 * np.rough * (-cw / (bw * bw) - 1.0 / bw); }
    }
  }
  return scval;
}

// ---- shadow cache (rt/srcobstr.c) ---------------------------------------------
// srcobstructp() (rt/srcobstr.c:196-278): the cell of a ray from o along d
// aimed at source src; -1 when the ray falls outside the grid.
__device__ __forceinline__ int shadow_cell(const Source& src, const SrcCache& C, V3 o, V3 d) {
  // no fused multiply-adds: shadow-ray origins on the cube's centre lines
  // fall exactly on a cell boundary, where the rounding decides the cell
#pragma clang fp contract(off)
  constexpr int S = kShadCache;
  int ondx;
  if (C.kind == SC_DISTANT) {
    const int ax = C.ax, ax1 = ax == 2 ? 0 : ax + 1, ax2 = ax == 0 ? 2 : ax - 1;
    const double sl[3] = {src.sloc[0], src.sloc[1], src.sloc[2]};
    const double ro[3] = {o.x, o.y, o.z};
    const double t = (C.o[ax] - ro[ax]) / sl[ax];
    if (t <= FTINY) return -1;
    // in the operation order of the reference's -O3 -ffast-math build
    // (reassociated to ((t sl - o) + ro) (e 2S))
    ondx = 2 * S * (int)(((t * sl[ax1] - C.o[ax1]) + ro[ax1]) * (C.e1 * (2 * S)));
    ondx += (int)(((t * sl[ax2] - C.o[ax2]) + ro[ax2]) * (C.e2 * (2 * S)));
    if ((ondx < 0) | (ondx >= 4 * S * S)) return -1;
  } else if (C.kind == SC_FLAT) {
    const double sd0 = -dot(d, load3(C.u)), sd1 = -dot(d, load3(C.v)), sd2 = -dot(d, load3(src.snorm));
    if (sd2 < 0) return -1;
    const double sd0m = fabs(sd0), sd1m = fabs(sd1);
    if (sd2 >= sd0m && sd2 >= sd1m) {
      ondx = S * (int)(S * (.5 - FTINY) * (1. + sd0 / sd2));
      ondx += (int)(S * (.5 - FTINY) * (1. + sd1 / sd2));
    } else if (sd0m >= sd1m) {
      ondx = S * S;
      if (sd0 < 0) ondx += ((S + 1) >> 1) * S;
      ondx += S * (int)(S * (.5 - FTINY) * (1. - sd2 / sd0m));
      ondx += (int)(S * (.5 - FTINY) * (1. + sd1 / sd0m));
    } else {
      ondx = S * S + ((S + 1) >> 1) * S * 2;
      if (sd1 < 0) ondx += ((S + 1) >> 1) * S;
      ondx += S * (int)(S * (.5 - FTINY) * (1. - sd2 / sd1m));
      ondx += (int)(S * (.5 - FTINY) * (1. + sd0 / sd1m));
    }
    if ((ondx < 0) | (ondx >= C.ncell)) return -1;
  } else {
    const double dd[3] = {d.x, d.y, d.z};
    int ax = 0;
    double amax = 0;
    for (int k = 3; k--;)
      if (fabs(dd[k]) > amax) { amax = fabs(dd[k]); ax = k; }
    const int ax1 = ax == 2 ? 0 : ax + 1, ax2 = ax == 0 ? 2 : ax - 1;
    ondx = 2 * S * S * ax;
    if (dd[ax] < 0) ondx += S * S;
    ondx += S * (int)(S * (.5 - FTINY) * (1. + dd[ax1] / amax));
    ondx += (int)(S * (.5 - FTINY) * (1. + dd[ax2] / amax));
    if ((ondx < 0) | (ondx >= C.ncell)) return -1;
  }
  return ondx;
}

// srcblocked() (rt/srcobstr.c:323-339): the cached blocker of the shadow
// ray's cell intersects it (and, for a local source, before the source).
// Marks the source as shadow-tested in this call.
__device__ __forceinline__ bool shadow_blocked(const DevScene& S, int sn, V3 o, V3 d) {
  if (!S.shadow_on) return false;
  const SrcCache& C = S.scache[sn];
  if (C.kind == SC_NONE) return false;
  if (!S.src_tested[sn]) S.src_tested[sn] = 1u;
  const Source& src = S.srcs[sn];
  const int cell = shadow_cell(src, C, o, d);
  if (cell < 0) return false;
  const unsigned long long e = S.shadow_obs[C.cell0 + cell];
  if (!(e >> 32)) return false;
  HitRec h;
  h.t = FHUGE; h.rod = 1.0; h.prim = -1;
  hit_prim(S, (int)(uint32_t)e, o, d, h);
  if (h.prim < 0) return false;
  if (src.flags & SDISTANT) return true;
  hit_prim(S, src.prim, o, d, h);
  return h.prim != src.prim;
}

// direct() of rt/source.c:398-557 with every source tested (no shadow
// threshold / untested-source statistics): one shadow spawn per source with
// non-zero potential.
// mi: material of the vertex when it is differentiated (adjoint replay), else -1.
__device__ __forceinline__ void direct_light(const DevScene& S, const RunCfg& c, const RayS& r, const HitS& hs, const NormDat& nd,
                             const Sink& sk, int mi = -1, double fest = 0.0) {
  for (int sn = 0; sn < S.nsrcs; ++sn) {
    const Source& src = S.srcs[sn];
    if (src.flags & SSKIP) continue;
    if ((src.flags & (SPROX | SDISTANT)) == SPROX) {
      V3 dv = hs.rop - load3(src.sloc);
      double pr = src.prox + src.srad;
      if (dot(dv, dv) > pr * pr) continue;
    }
    uint64_t kshadow = child_key(r.key, SLOT_SHADOW + sn);
    // nextssamp() of rt/srcsamp.c:36-144 with a single partition
    double vpos[3] = {0, 0, 0};
    if (c.p.dstrsrc > FTINY) {
      vpos[0] = rnd(kshadow, RN_SRC0);
      vpos[1] = rnd(kshadow, RN_SRC1);
      vpos[2] = (src.flags & SFLAT) ? 0.5 : rnd(kshadow, RN_SRC2);
      for (int i = 0; i < 3; ++i) vpos[i] = c.p.dstrsrc * (1. - 2. * vpos[i]);
    }
    if ((src.flags & SCIR) && (c.p.dstrsrc > 0.7)) {
      double trim[3];
      if (src.flags & (SFLAT | SDISTANT)) {
        double d = 1.12837917;
        trim[0] = d * sqrt(1.0 - 0.5 * vpos[1] * vpos[1]);
        trim[1] = d * sqrt(1.0 - 0.5 * vpos[0] * vpos[0]);
        trim[2] = 0.0;
      } else {
        trim[2] = trim[0] = vpos[0] * vpos[0];
        double d = vpos[1] * vpos[1];
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
      for (int i = 0; i < 3; ++i) vpos[i] *= trim[i];
    }
    V3 dir = load3(src.sloc) + load3(src.ss[0]) * vpos[0] + load3(src.ss[1]) * vpos[1] + load3(src.ss[2]) * vpos[2];
    if (!(src.flags & SDISTANT)) dir = dir - hs.rop;
    double dist = normalize(dir);
    if (dist == 0.0) continue;
    double dom;
    if (src.flags & SFLAT) dom = -dot(load3(src.snorm), dir);
    else if (src.flags & SCYL) {
      V3 dv = load3(src.ss[0]);
      double dd = dot(dir, dv);
      dd *= dd / dot(dv, dv);
      dom = sqrt(1. - dd);
    } else dom = 1.0;
    if (src.flags & SDISTANT) {
      dom *= src.ss2;
    } else {
      if (dom <= 1e-4) continue;
      dom *= src.ss2 / (dist * dist);
      if ((src.flags & SPROX) && dist > src.prox) continue;
    }
    // the shadow cache decides before anything else (rt/source.c:437-446)
    if (shadow_blocked(S, sn, hs.rop, dir)) continue;
    DirParts parts;
    C3 coef = dirnorm(nd, c, r.d, dir, dom, mi >= 0 ? &parts : nullptr);
    if (!(cmax(coef) > 0.0)) continue;
    // potential: the ray must reach the source (srcvalue, rt/source.c:261-299)
    RayS sr = r;
    sr.d = dir;
    sr.rsrc = sn;
    sr.crtype = r.crtype | RSHADOW;
    const Material& sm = S.mats[src.mat];
    C3 em = c3(0, 0, 0);
    bool pass;
    if (src.flags & SDISTANT) {
      if (2. * PI_D * (1. - dot(load3(src.sloc), dir)) > src.ss2) continue;
      if (!light_rules(S, c, sr, sm, src_obj(src, sn), FHUGE, 1.0, &em, &pass) || pass) continue;
    } else {
      HitRec h1; h1.t = FHUGE; h1.rod = 1.0; h1.prim = -1;
      hit_prim(S, src.prim, hs.rop, dir, h1);
      if (h1.prim < 0) continue;
      if (!light_rules(S, c, sr, sm, src.prim, h1.t, h1.rod, &em, &pass) || pass) continue;
    }
    if (!(cbright(coef * em) > 0.0)) continue;
    bool thru = (hs.rod > 0) ^ (dot(hs.ron, dir) > 0);
    Spawn s = make_child(r, hs.rop, thru ? TSHADOW : RSHADOW, nullptr, SLOT_SHADOW + sn, 0);
    // own rcoef of the shadow ray is the source coefficient (rt/source.c:511-513)
    C3 cum = r.coef * coef;
    s.coef[0] = cum.r; s.coef[1] = cum.g; s.coef[2] = cum.b;
    s.crc = (float)cmax(coef);
    s.rsrc = (int16_t)sn;
    set_dir(s, dir);
    if (mi >= 0) {
      PathAdd e = path_entry(mi, parts.side ? PV_DIR_T : PV_DIR_R, fest);
      e.w[1] = __float_as_uint((float)parts.a);
      e.w[2] = __float_as_uint((float)parts.b);
      e.w[3] = __float_as_uint((float)parts.g);
      emit(sk, s, r, e);
    } else {
      emit(sk, s, r);
    }
  }
}

// multambient() + samp_hemi() set-up (rt/ambient.c:229-330, rt/ambcomp.c:351-424)
// for -aa 0: one hemisphere spawn of n*n stratified children.
__device__ __forceinline__ void ambient(const DevScene& S, const RunCfg& c, const RayS& r, const HitS& hs, C3 aval, V3 nrm,
                        const Sink& sk, PathAdd entry = path_none()) {
  const Params& p = c.p;
  bool dumb = (p.ambdiv <= 0) || (r.adepth >= p.ambounce);
  double rdot = dot(nrm, hs.ron);
  int sgn = 1 - 2 * (rdot < 0);
  if (!dumb) {
    double d = cmax(aval);
    if (d <= FTINY) dumb = true;
    else {
      double wt = r.w;
      d *= 0.8 * r.w / (p.ambdiv * p.minweight + 1e-20);
      if (wt > d) wt = d;
      int n = (int)(sqrt(p.ambdiv * wt) + 0.5);
      if (n < 1) n = 1;
      if (n > 65535) n = 65535;
      bool back = sgn < 0;
      V3 onrm = back ? -hs.ron : hs.ron;
      C3 acoef = aval * (1.0 / ((double)n * n));
      Spawn s = make_child(r, hs.rop, back ? TAMBIENT : RAMBIENT, &acoef, 0, back ? SPF_BACK : 0);
      s.key = r.key;   // children derive their keys from the parent key
      V3 ux, uy;
      tangent_frame(onrm, rnd(r.key, back ? RN_FRAME_BACK : RN_FRAME_FRONT), ux, uy);
      s.d0[0] = (float)onrm.x; s.d0[1] = (float)onrm.y; s.d0[2] = (float)onrm.z;
      s.d1[0] = (float)ux.x; s.d1[1] = (float)ux.y; s.d1[2] = (float)ux.z;
      s.ns = (uint16_t)n;
      s.count = (uint32_t)n * (uint32_t)n;
      s.adepth = (uint8_t)(r.adepth + 1);
      emit(sk, s, r, entry);
      return;
    }
  }
  // dumbamb: constant ambient value -av
  if (c.mode == 0) {
    C3 av = aval * c3((float)p.ambval[0], (float)p.ambval[1], (float)p.ambval[2]);
    if (cmax(av) > 0) deposit(c, r.sensor, 0, r.coef * av);
  }
}

// gaussamp() of rt/normal.c:364-495 with one sample per component (-ss <= 1.5).
// mi >= 0 (adjoint replay of a differentiated material): each child's path
// entry carries the roughness score of its trials (see adjoint_deposit()):
// trial k draws tan^2 = alpha^2 u_k with u_k = -ln rv1 exponential, whose
// log density -ln alpha^2 - u_k has derivative (2 / alpha)(u_k - 1).
__device__ __forceinline__ void gauss_sample(const DevScene& S, const RunCfg& c, const RayS& r, const HitS& hs, const NormDat& np,
                             const Sink& sk, int mi) {
  if ((np.specfl & (SP_REFL | SP_RBLT)) != SP_REFL && (np.specfl & (SP_TRAN | SP_TBLT)) != SP_TRAN) return;
  V3 u, v;
  tangent_frame(np.pnorm, rnd(r.key, RN_GFRAME), u, v);
  const double sfac = mi >= 0 ? 2.0 / np.rough : 0.0;
  if ((np.specfl & (SP_REFL | SP_RBLT)) == SP_REFL) {
    C3 sc = np.scolor;
    Spawn s = make_child(r, hs.rop, RSPECULAR, &sc, SLOT_RSPEC, 0);
    bool ok = false;
    V3 dir;
    double score = 0.0;
    for (int ntrials = 0; ntrials < 10; ++ntrials) {
      double rv0 = rnd(s.key, RN_GAUSS + 2 * ntrials), rv1 = rnd(s.key, RN_GAUSS + 2 * ntrials + 1);
      double d = 2.0 * PI_D * rv0;
      double cosp = cos(d), sinp = sin(d);
      if ((0. <= c.p.specjitter) & (c.p.specjitter < 1.)) rv1 = 1.0 - c.p.specjitter * rv1;
      d = (rv1 <= FTINY) ? 1.0 : sqrt(np.alpha2 * -log(rv1));
      if (mi >= 0) score += sfac * (d * d / np.alpha2 - 1.0);
      V3 h = np.pnorm + (u * cosp + v * sinp) * d;
      d = -2.0 * dot(h, r.d) / (1.0 + d * d);
      dir = r.d + h * d;
      if (dot(dir, hs.ron) <= FTINY) continue;
      normalize(dir);
      ok = true;
      break;
    }
    if (ok) {
      set_dir(s, dir);
      if (mi >= 0) {
        PathAdd e = path_entry(mi, PV_GSPEC, 0.0);
        e.w[1] = __float_as_uint((float)score);
        emit(sk, s, r, e);
      } else {
        emit(sk, s, r);
      }
    }
  }
  if ((np.specfl & (SP_TRAN | SP_TBLT)) == SP_TRAN) {
    C3 tc = np.mcolor * np.tspec;
    Spawn s = make_child(r, hs.rop, TSPECULAR, &tc, SLOT_TSPEC, 0);
    bool ok = false;
    V3 dir;
    double score = 0.0;
    for (int ntrials = 0; ntrials < 10; ++ntrials) {
      double rv0 = rnd(s.key, RN_GAUSS + 2 * ntrials), rv1 = rnd(s.key, RN_GAUSS + 2 * ntrials + 1);
      double d = 2.0 * PI_D * rv0;
      double cosp = cos(d), sinp = sin(d);
      if ((0. <= c.p.specjitter) & (c.p.specjitter < 1.)) rv1 = 1.0 - c.p.specjitter * rv1;
      d = (rv1 <= FTINY) ? 1.0 : sqrt(np.alpha2 * -log(rv1));
      if (mi >= 0) score += sfac * (d * d / np.alpha2 - 1.0);
      dir = np.prdir + (u * cosp + v * sinp) * d;
      if (dot(dir, hs.ron) >= -FTINY) continue;
      normalize(dir);
      ok = true;
      break;
    }
    if (ok) {
      set_dir(s, dir);
      if (mi >= 0) {
        PathAdd e = path_entry(mi, PV_GTRANS, 0.0);
        e.w[1] = __float_as_uint((float)score);
        emit(sk, s, r, e);
      } else {
        emit(sk, s, r);
      }
    }
  }
}

// m_normal() of rt/normal.c:177-360 (plastic, metal, trans)
// mi: material index of m for the adjoint path (-1: not differentiated).
__device__ __forceinline__ void shade_normal(const DevScene& S, const RunCfg& c, const RayS& r, HitS hs, const Material& m, int mi,
                             const Sink& sk) {
  if ((r.crtype & SHADOW) && m.kind != MK_TRANS) return;
  if (hs.rod < 0.0) {
    if (!c.p.backvis) {
      Spawn s = make_child(r, hs.rop, TRANS, nullptr, SLOT_PASS, SPF_FORCE);
      set_dir(s, r.d);
      emit(sk, s, r);
      return;
    }
    flip_surface(hs);
  }
  NormDat nd;
  nd.mcolor = c3((float)m.f[0], (float)m.f[1], (float)m.f[2]);
  nd.specfl = 0;
  nd.rough = m.f[4];
  nd.alpha2 = m.f[4] * m.f[4];
  if (nd.alpha2 <= FTINY) nd.specfl |= SP_PURE;
  const bool hastexture = dot(hs.pert, hs.pert) > FTINY * FTINY;
  if (hastexture) {
    nd.pdot = ray_normal(nd.pnorm, hs, r.d);
  } else {
    nd.pnorm = hs.ron;
    nd.pdot = hs.rod;
  }
  if (!hastexture && hs.prim >= 0 && hs.flat) nd.specfl |= SP_FLAT;
  if (nd.pdot < .001) nd.pdot = .001;
  nd.rspec = m.f[3];
  double fest = 0.;
  if ((nd.specfl & SP_PURE) && nd.rspec >= FRESTHRESH) {
    fest = FRESNE(nd.pdot);
    nd.rspec += fest * (1. - nd.rspec);
  }
  nd.prdir = r.d;
  if (m.kind == MK_TRANS) {
    nd.trans = m.f[5] * (1.0 - nd.rspec);
    nd.tspec = nd.trans * m.f[6];
    nd.tdiff = nd.trans - nd.tspec;
    if (nd.tspec > FTINY) {
      nd.specfl |= SP_TRAN;
      if (!(nd.specfl & SP_PURE) && c.p.specthresh >= nd.tspec - FTINY) nd.specfl |= SP_TBLT;
      if (hastexture && !(r.crtype & (SHADOW | AMBIENT)) && !hs.phong) {
        V3 pd = r.d - hs.pert;
        if (dot(pd, hs.ron) < -FTINY) {
          normalize(pd);
          nd.prdir = pd;
        }
      }
    }
  } else {
    nd.tdiff = nd.tspec = nd.trans = 0.0;
  }
  nd.rdiff = 1.0 - nd.trans - nd.rspec;
  // adjoint path entries of this vertex (differentiated materials only)
  const bool dv = c.adj && mi >= 0 && c.grad_row[mi] >= 0;
  if ((nd.specfl & (SP_TRAN | SP_PURE | SP_TBLT)) == (SP_TRAN | SP_PURE)) {
    C3 tc = nd.mcolor * nd.tspec;
    Spawn s = make_child(r, hs.rop, TRANS, &tc, SLOT_TRANS, 0);
    set_dir(s, nd.prdir);
    emit(sk, s, r, dv ? path_entry(mi, PV_TSPEC, fest) : path_none());
  }
  if (r.crtype & SHADOW) return;
  nd.scolor = c3(0, 0, 0);
  if (nd.rspec > FTINY) {
    nd.specfl |= SP_REFL;
    if (m.kind != MK_METAL) {
      nd.scolor = c3((float)nd.rspec, (float)nd.rspec, (float)nd.rspec);
    } else if (fest > FTINY) {
      double d = m.f[3] * (1. - fest);
      nd.scolor = c3((float)(fest + nd.mcolor.r * d), (float)(fest + nd.mcolor.g * d), (float)(fest + nd.mcolor.b * d));
    } else {
      nd.scolor = nd.mcolor * nd.rspec;
    }
    if (!(nd.specfl & SP_PURE) && c.p.specthresh >= nd.rspec - FTINY) nd.specfl |= SP_RBLT;
  }
  if ((nd.specfl & (SP_REFL | SP_PURE | SP_RBLT)) == (SP_REFL | SP_PURE)) {
    C3 sc = nd.scolor;
    Spawn s = make_child(r, hs.rop, REFLECTED, &sc, SLOT_REFL, 0);
    V3 dir = r.d + nd.pnorm * (2. * nd.pdot);
    if (hastexture && dot(dir, hs.ron) <= FTINY) dir = r.d + hs.ron * (2. * hs.rod);   // penetration
    normalize(dir);
    set_dir(s, dir);
    emit(sk, s, r, dv ? path_entry(mi, PV_SPEC, fest) : path_none());
  }
  if ((nd.specfl & SP_PURE) && nd.rdiff <= FTINY && nd.tdiff <= FTINY) return;
  if (!(nd.specfl & SP_PURE)) gauss_sample(S, c, r, hs, nd, sk, dv ? mi : -1);
  if (nd.rdiff > FTINY) {
    C3 a = nd.mcolor * nd.rdiff;
    if (nd.specfl & SP_RBLT) a = a + nd.scolor;
    ambient(S, c, r, hs, a, nd.pnorm, sk,
            dv ? path_entry(mi, (nd.specfl & SP_RBLT) ? PV_AMB_RBLT : PV_AMB, fest) : path_none());
  }
  if (nd.tdiff > FTINY) {
    C3 a = nd.mcolor * ((nd.specfl & SP_TBLT) ? nd.trans : nd.tdiff);
    ambient(S, c, r, hs, a, -nd.pnorm, sk,
            dv ? path_entry(mi, (nd.specfl & SP_TBLT) ? PV_AMB_BACK_TBLT : PV_AMB_BACK, fest) : path_none());
  }
  direct_light(S, c, r, hs, nd, sk, dv ? mi : -1, fest);
}

// m_glass() of rt/glass.c:46-165
__device__ __forceinline__ void shade_glass(const RunCfg& c, const RayS& r, HitS hs, const Material& m, const Sink& sk) {
  double rindex = m.nfargs == 4 ? m.f[3] : 1.52;
  if (!c.p.backvis && hs.rod <= 0.0) {
    Spawn s = make_child(r, hs.rop, TRANS, nullptr, SLOT_PASS, SPF_FORCE);
    set_dir(s, r.d);
    emit(sk, s, r);
    return;
  }
  double mc[3] = {m.f[0], m.f[1], m.f[2]};
  bool hastrans = fmax(fmax(mc[0], mc[1]), mc[2]) > 1e-15;
  if (hastrans) {
    for (int i = 0; i < 3; ++i) if (mc[i] < 1e-15) mc[i] = 1e-15;
  } else if (r.crtype & SHADOW) {
    return;
  }
  if (hs.rod < 0.0) flip_surface(hs);
  const bool hastexture = dot(hs.pert, hs.pert) > FTINY * FTINY;
  V3 pnorm = hs.ron;
  double pdot = hs.rod;
  if (hastexture) pdot = ray_normal(pnorm, hs, r.d);
  double cos2 = sqrt((1.0 - 1.0 / (rindex * rindex)) + pdot * pdot / (rindex * rindex));
  if (hastrans) for (int i = 0; i < 3; ++i) mc[i] = pow(mc[i], 1.0 / cos2);
  double r1e = (pdot - rindex * cos2) / (pdot + rindex * cos2);
  r1e *= r1e;
  double r1m = (1.0 / pdot - rindex / cos2) / (1.0 / pdot + rindex / cos2);
  r1m *= r1m;
  if (hastrans) {
    double ct[3];
    for (int i = 0; i < 3; ++i) {
      double d = mc[i];
      ct[i] = .5 * (1.0 - r1e) * (1.0 - r1e) * d / (1.0 - r1e * r1e * d * d) +
              .5 * (1.0 - r1m) * (1.0 - r1m) * d / (1.0 - r1m * r1m * d * d);
    }
    C3 tc = c3((float)ct[0], (float)ct[1], (float)ct[2]);
    Spawn s = make_child(r, hs.rop, TRANS, &tc, SLOT_TRANS, 0);
    V3 td = r.d;
    if (hastexture && !(r.crtype & (SHADOW | AMBIENT)) && !hs.phong) {
      V3 pd = r.d + hs.pert * (2. * (1. - rindex));
      if (dot(pd, pd) > 0.0) { normalize(pd); td = pd; }
    }
    set_dir(s, td);
    emit(sk, s, r);
  }
  if (r.crtype & SHADOW) return;
  double cr[3];
  for (int i = 0; i < 3; ++i) {
    double d = mc[i];
    d *= d;
    cr[i] = .5 * r1e * (1.0 + (1.0 - 2.0 * r1e) * d) / (1.0 - r1e * r1e * d) +
            .5 * r1m * (1.0 + (1.0 - 2.0 * r1m) * d) / (1.0 - r1m * r1m * d);
  }
  C3 rc = c3((float)cr[0], (float)cr[1], (float)cr[2]);
  Spawn s = make_child(r, hs.rop, REFLECTED, &rc, SLOT_REFL, 0);
  V3 dir = r.d + pnorm * (2. * pdot);
  normalize(dir);
  set_dir(s, dir);
  emit(sk, s, r);
}

// rayshade() for a surface hit, followed by rcontrib's trace callback.
__device__ __forceinline__ void shade_hit(const DevScene& S, const RunCfg& c, const RayS& r, const HitRec& h, const Sink& sk,
                                          const Prim& P) {
  HitS hs;
  hs.rot = h.t;
  hs.rop = r.o + r.d * h.t;
  hs.ron = prim_normal_k(S, P.kind, P.goff, hs.rop);
  hs.rod = -dot(r.d, hs.ron);
  hs.pert = prim_pert_k(S, P.kind, P.goff, hs.rop, hs.ron);
  hs.prim = h.prim;
  hs.flat = P.flat & PF_FLAT;
  hs.phong = prim_phong_k(S, P.kind, P.goff);
  if (P.mat < 0) { atomicOr(c.err, ERR_NOMAT); return; }
  const Material& m = S.mats[P.mat];
  bool zeroed = false;
  C3 rcol = c3(0, 0, 0);
  // -i: Lambertian replacement at the first qualifying surface (raytirrad)
  bool tst_irrad = c.irrad && !(r.crtype & ~(PRIMARY | TRANS));
  if (tst_irrad && m.kind != MK_LIGHT && m.kind != MK_GLOW && m.kind != MK_ILLUM) {
    if (m.transp) {
      Spawn s = make_child(r, hs.rop, TRANS, nullptr, SLOT_PASS, SPF_FORCE);
      set_dir(s, r.d);
      emit(sk, s, r);
    } else {
      shade_normal(S, c, r, hs, S.mats[S.lamb], -1, sk);
    }
  } else {
    switch (m.kind) {
      case MK_PLASTIC:
      case MK_METAL:
      case MK_TRANS:
        shade_normal(S, c, r, hs, m, P.mat, sk);
        break;
      case MK_GLASS:
        shade_glass(c, r, hs, m, sk);
        break;
      default: {   // light, glow, illum
        C3 em = c3(0, 0, 0);
        bool pass;
        if (!light_rules(S, c, r, m, h.prim, hs.rot, hs.rod, &em, &pass)) {
          zeroed = true;
        } else if (pass) {
          if (m.kind == MK_ILLUM && m.illum_alt >= 0) {
            const Material& alt = S.mats[m.illum_alt];
            if (alt.kind == MK_GLASS) shade_glass(c, r, hs, alt, sk);
            else if (alt.kind <= MK_TRANS) shade_normal(S, c, r, hs, alt, -1, sk);
          } else {
            Spawn s = make_child(r, hs.rop, TRANS, nullptr, SLOT_PASS, SPF_FORCE);
            set_dir(s, r.d);
            emit(sk, s, r);
          }
        } else {
          rcol = em;
          if (c.mode == 0) deposit(c, r.sensor, 0, r.coef * em);
        }
      }
    }
  }
  trace_contrib(S, c, r, P.modslot, h.prim, zeroed, rcol);
}

// A ray that hit nothing: distant sources (sourcehit + m_light).
__device__ __forceinline__ void shade_miss(const DevScene& S, const RunCfg& c, const RayS& r) {
  int si = source_hit(S, r);
  if (si < 0) return;
  const Source& src = S.srcs[si];
  const Material& m = S.mats[src.mat];
  C3 em = c3(0, 0, 0);
  bool pass;
  bool ok = light_rules(S, c, r, m, -(2 + si), FHUGE, 1.0, &em, &pass);
  if (!ok || pass) {
    trace_contrib(S, c, r, src.modslot, -(2 + si), true, em);
    return;
  }
  if (c.mode == 0) deposit(c, r.sensor, 0, r.coef * em);
  trace_contrib(S, c, r, src.modslot, -(2 + si), false, em);
}

// ---- kernels -----------------------------------------------------------------
// Scene tables of a ray: the kernel argument, or its entry of a batched launch.
__device__ __forceinline__ DevScene scene_of(const DevScene& Sarg, const RunCfg& c, int scene) {
  return c.nscenes ? c.scenes[scene] : Sarg;
}
__device__ __forceinline__ int scene_of_ray(const RunCfg& c, uint32_t ray) {
  if (!c.nscenes) return 0;
  int lo = 0, hi = c.nscenes;       // last k with scene_start[k] <= ray
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (c.scene_start[mid] <= ray) lo = mid; else hi = mid;
  }
  return lo;
}

// Generation 0 for -I: the pseudo-hit on a Lambertian surface at each sensor
// (rt/rtrace.c:416-431, rt/rcontrib.c:321-338).
// ADJ: adjoint replay (path records on, deposits into the gradient); the
// forward instantiation compiles all of that away.
template <bool ADJ>
__global__ void __launch_bounds__(BLOCK) k_irrad_seed(DevScene Sarg, RunCfg c, const double* rays, uint32_t nsens,
                                                      uint32_t s0, Sink sk) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nsens) return;
  c.adj = ADJ;
  if (!ADJ) sk.pbuf = nullptr;
  const int scene = scene_of_ray(c, s0 + i);
  const DevScene& S = scene_of(Sarg, c, scene);
  const double* ry = rays + 6 * (size_t)i;
  V3 org = load3(ry), dir = load3(ry + 3);
  normalize(dir);
  RayS r;
  {
#pragma clang fp contract(off)
    r.o = V3{org.x + dir.x * 1.1e-4, org.y + dir.y * 1.1e-4, org.z + dir.z * 1.1e-4};
  }
  r.d = -dir;
  r.coef = c3(1, 1, 1);
  r.w = 1.0f;
  r.crc = 1.0f;
  r.sensor = (s0 + i) / (uint32_t)c.accum;
  r.crtype = r.rtype = PRIMARY;
  r.rsrc = -1;
  r.rlvl = 0;
  r.adepth = 0;
  r.scene = scene;
  r.key = sensor_key(c.p.seed + scene, s0 + i - (c.nscenes ? c.scene_start[scene] : 0));
  if (ADJ) r.path = PathRec{};
  HitS hs;
  // rayirrad() (rt/rtrace.c:416-431): the pretended hit 1e-5 back along
  // the ray, computed as the reference does (shadow-cache cells read it)
  hs.rot = 1e-5;
  {
#pragma clang fp contract(off)
    hs.rop = V3{r.o.x + r.d.x * 1e-5, r.o.y + r.d.y * 1e-5, r.o.z + r.d.z * 1e-5};
  }
  hs.ron = dir;
  hs.rod = 1.0;
  hs.prim = -1;
  hs.flat = 0;
  hs.pert = V3{0.0, 0.0, 0.0};
  hs.phong = 0;
  shade_normal(S, c, r, hs, S.mats[S.lamb], -1, sk);
}

// Generation 0 for radiance rays: primary rays as explicit spawns.
__global__ void __launch_bounds__(BLOCK) k_primary_seed(RunCfg c, const double* rays, uint32_t nsens, uint32_t s0,
                                                        Spawn* out, uint32_t* counts) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nsens) return;
  const double* ry = rays + 6 * (size_t)i;
  V3 dir = load3(ry + 3);
  normalize(dir);
  Spawn s;
  s.org[0] = ry[0]; s.org[1] = ry[1]; s.org[2] = ry[2];
  set_dir(s, dir);
  s.d1[0] = s.d1[1] = s.d1[2] = 0.f;
  s.coef[0] = s.coef[1] = s.coef[2] = 1.f;
  s.cw = 1.f;
  s.crc = 1.f;
  s.sensor = (s0 + i) / (uint32_t)c.accum;
  s.count = 1;
  s.ns = 0;
  s.crtype = s.rtype = PRIMARY;
  s.rsrc = -1;
  s.rlvl = 0;
  s.adepth = 0;
  s.flags = SPF_FORCE;
  const int scene = scene_of_ray(c, s0 + i);
  s.scene = (uint16_t)scene;
  s.key = sensor_key(c.p.seed + scene, s0 + i - (c.nscenes ? c.scene_start[scene] : 0));
  out[i] = s;
  counts[i] = 1u;
}

__global__ void __launch_bounds__(BLOCK) k_counts(const Spawn* sp, uint32_t n, uint32_t* counts) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) counts[i] = sp[i].count;
}

// Child ci of spawn record sp: origin, direction and rayorigin() admission.
// pp: the spawn's path record in adjoint replay (null otherwise).
__device__ __forceinline__ bool build_ray(const RunCfg& c, const Spawn& sp, uint32_t ci, RayS& r,
                                          const PathRec* pp = nullptr) {
  if (c.adj) r.path = pp ? *pp : PathRec{};
  r.o = v3(sp.org[0], sp.org[1], sp.org[2]);
  r.coef = c3(sp.coef[0], sp.coef[1], sp.coef[2]);
  r.w = sp.cw;
  r.crc = sp.crc;
  r.sensor = sp.sensor;
  r.crtype = sp.crtype;
  r.rtype = sp.rtype;
  r.rsrc = sp.rsrc;
  r.rlvl = sp.rlvl;
  r.adepth = sp.adepth;
  r.scene = sp.scene;
  if (sp.ns) {
    // stratified cosine-weighted hemisphere sample (rt/ambcomp.c:177-216)
    const uint32_t ns = sp.ns;
    uint32_t i, j;
    if (c.child_order) {
      // children in strips of 8 stratum rows, column by column within a
      // strip: 64 consecutive children (a wavefront) cover an 8 x 8 block of
      // strata, a compact patch of the hemisphere, instead of a row of
      // strata sweeping across it
      const uint32_t strip = min(ci / (8u * ns), (ns - 1u) / 8u);
      const uint32_t r0 = ci - 8u * ns * strip;
      const uint32_t h = min(8u, ns - 8u * strip);
      j = r0 / h;
      i = 8u * strip + r0 % h;
    } else {
      i = ci / ns;
      j = ci % ns;
    }
    // the sample of stratum (i, j) is drawn from its own key, whatever order
    // the children are handed out in
    r.key = child_key(sp.key, ((sp.flags & SPF_BACK) ? SLOT_HEMI_BACK : SLOT_HEMI_FRONT) + i * ns + j);
    double ss0 = rnd(r.key, RN_AMB0), ss1 = rnd(r.key, RN_AMB1);
    double spt[2];
    square2disk(spt, (j + ss1) / ns, (i + ss0) / ns);
    double zd = sqrt(fmax(0.0, 1. - spt[0] * spt[0] - spt[1] * spt[1]));
    V3 onrm = v3(sp.d0[0], sp.d0[1], sp.d0[2]);
    V3 ux = v3(sp.d1[0], sp.d1[1], sp.d1[2]);
    V3 uy = cross(onrm, ux);
    r.d = ux * spt[0] + uy * spt[1] + onrm * zd;
  } else {
    r.key = sp.key;
    r.d = v3(sp.d0[0], sp.d0[1], sp.d0[2]);
  }
  normalize(r.d);
  return admit(c, r, sp.flags);
}

__device__ __forceinline__ uint32_t find_spawn(const uint64_t* offs, uint32_t nspawn, uint64_t gid) {
  uint32_t lo = 0, hi = nspawn;      // last s with offs[s] <= gid
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (offs[mid] <= gid) lo = mid; else hi = mid;
  }
  return lo;
}

// Spawn record si and its first child index, requested together and waited
// for together (the empty asm statement takes the loaded vectors as inputs,
// which keeps the compiler from splitting the record's fields over several
// round trips at their first uses).
__device__ __forceinline__ void load_spawn(const Spawn* in, const uint64_t* offs, uint32_t si, Spawn& sp,
                                           uint64_t& off) {
  static_assert(sizeof(Spawn) == 6 * sizeof(u32x4), "Spawn layout");
  const u32x4* spp = reinterpret_cast<const u32x4*>(in + si);
  const u32x4 s0 = spp[0], s1 = spp[1], s2 = spp[2], s3 = spp[3], s4 = spp[4], s5 = spp[5];
  off = offs[si];
  asm volatile("" :: "v"(s0), "v"(s1), "v"(s2), "v"(s3), "v"(s4), "v"(s5), "v"(off));
  const u32x4 sv[6] = {s0, s1, s2, s3, s4, s5};
  __builtin_memcpy(&sp, sv, sizeof sp);
}

// Spawn of the first child of every 64-child chunk of a window [c0, c0 +
// nth), and at index nchunk the spawn of the window's last child: one search
// per chunk, all in parallel, ahead of the walk.  The walk's refill then
// finds a child's spawn between its chunk's two entries (usually the same
// record) instead of searching the whole scan from a single lane, a chain of
// about log2(nspawn) dependent loads that every lane of the wavefront waited
// for.
constexpr uint32_t CHUNK_LOG = 6;
__global__ void __launch_bounds__(BLOCK) k_chunk_spawn(const uint64_t* offs, uint32_t nspawn, uint64_t c0, uint32_t nth,
                                                       uint32_t* first) {
  const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
  const uint32_t nch = (nth + (1u << CHUNK_LOG) - 1) >> CHUNK_LOG;
  if (k > nch) return;
  const uint64_t gid = c0 + (k < nch ? (uint64_t)k << CHUNK_LOG : (uint64_t)nth - 1);
  first[k] = find_spawn(offs, nspawn, gid);
}

// Traversal result of one child ray (prim -2: not traced).
struct alignas(16) TraceOut {
  double t;
  int32_t prim;
  float dx, dy, dz;
  uint32_t spawn;
  uint32_t pad;
};

// One generation, part 1: children [c0, c0+nthreads) of `in` -> closest hits.
#ifndef RG_TRACE_WAVES
#define RG_TRACE_WAVES 4     // waves per SIMD the register budget of k_trace targets
#endif
__global__ void __launch_bounds__(TBLOCK) __attribute__((amdgpu_waves_per_eu(RG_TRACE_WAVES)))
k_trace(DevScene Sarg, RunCfg c, const Spawn* in, const uint64_t* offs, uint32_t nspawn, uint64_t c0,
        uint32_t nthreads, TraceOut* outp) {
  __shared__ int32_t stk[LSTACK * TBLOCK];
  __shared__ float stn[LSTACK * TBLOCK];
  __shared__ float tes[NBIG_ORDER * TBLOCK];
  uint32_t t = blockIdx.x * TBLOCK + threadIdx.x;
  if (t >= nthreads) return;
  c.adj = 0;      // admission does not depend on the path record
  uint64_t gid = c0 + t;
  // The wavefront's children are consecutive: one lane searches the whole
  // table for the first child, the others only the few records after it
  // (every record has at least one child).
  unsigned long long act = __ballot(1);
  uint32_t leader = __ffsll((long long)act) - 1;
  uint64_t gid0 = __shfl(gid, leader);
  uint32_t base = 0;
  if (__lane_id() == leader) base = find_spawn(offs, nspawn, gid0);
  base = __shfl(base, leader);
  uint32_t lo = base, hi = (uint32_t)min((uint64_t)nspawn, (uint64_t)base + (gid - gid0) + 1);
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (offs[mid] <= gid) lo = mid; else hi = mid;
  }
  uint32_t si = lo;
  RayS r;
  TraceOut o;
  o.spawn = si;
  o.pad = 0;
  if (!build_ray(c, in[si], (uint32_t)(gid - offs[si]), r)) {
    o.prim = -2;
    o.t = 0;
    o.dx = o.dy = o.dz = 0.f;
    outp[t] = o;
    return;
  }
  wave_count(c.nrays);
  HitRec h;
  closest_hit<TBLOCK>(scene_of(Sarg, c, r.scene), r.o, r.d, h, stk, stn, tes, c.err, c.visits);
  o.t = h.t;
  o.prim = h.prim;
  o.dx = (float)r.d.x; o.dy = (float)r.d.y; o.dz = (float)r.d.z;
  outp[t] = o;
}

// Persistent form of k_trace with dynamic ray fetch.  A fixed grid of
// wavefronts loops; when at least `refill` lanes of a wavefront have finished
// their ray (or all have), the wavefront takes the next batch of child
// indices from `work` with one atomic and the idle lanes start new rays,
// while the others carry on.  Lanes therefore do not idle until the longest
// ray of their wavefront ends, which is what bounds the one-ray-per-thread
// kernel (its VALU lane utilisation is about a third).  A child's spawn lies
// between the entries of its 64-child chunk in `first` (k_chunk_spawn).  Same
// admission, closest-hit walk and output record as k_trace.  BATCH: rays of several
// scenes (each lane keeps its ray's scene tables); otherwise the kernel
// argument's scene, held in scalar registers.  WIDE: the walk fetches the
// four-wide tree (four child boxes per node; the children hit are taken
// nearest first, the others pushed farthest first).
template <int BS, bool BATCH, bool WIDE>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(RG_TRACE_WAVES)))
k_trace_pf(DevScene Sarg, RunCfg c, const Spawn* in, const uint64_t* offs, const uint32_t* first, uint64_t c0,
           uint32_t nthreads, TraceOut* outp, uint32_t* work, uint32_t refill) {
  __shared__ int32_t stk[LSTACK * BS];
  __shared__ float stn[LSTACK * BS];
  __shared__ float tes[NBIG_ORDER * BS];
  constexpr int32_t DONE = INT32_MIN;
  c.adj = 0;
  const int lane = threadIdx.x;
  const uint32_t wl = __lane_id();
  uint64_t deep[PSTACK];
  int32_t tcur = -1;              // window index of this lane's ray, -1: idle
  bool exhausted = false;         // no child left to take (wave-uniform)
  DevScene S = Sarg;
  V3 o = v3(0, 0, 0), d = v3(0, 0, 1);
  HitRec h;
  h.t = FHUGE; h.rod = 1.0; h.prim = -1;
  int32_t node = DONE, leaf = -1;
  int sp = 0, iter = 0;
  uint32_t si = 0;
  float ox = 0, oy = 0, oz = 0, ix = 0, iy = 0, iz = 0;
  uint32_t nvis = 0, nprim = 0;
  for (;;) {
    // ---- refill idle lanes ----
    unsigned long long idle = __ballot(tcur < 0);
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (!exhausted && nidle && (nidle >= refill || nidle == (uint32_t)__popcll(__ballot(1)))) {
      const uint32_t leader = __ffsll((long long)idle) - 1;
      uint32_t base = 0;
      if (wl == leader) base = atomicAdd(work, nidle);
      base = __shfl(base, leader);
      if ((uint64_t)base + nidle >= nthreads) exhausted = true;
      if (tcur < 0) {
        const uint32_t my = base + lane_rank(idle);
        if (my < nthreads) {
          const uint64_t gid = c0 + my;
          // last spawn with offs <= gid, between the spawn of this chunk's
          // first child and that of the next chunk's (or the window's last child)
          const uint32_t ck = my >> CHUNK_LOG;
          uint32_t lo = first[ck], hi = first[ck + 1] + 1;
          if (hi - lo > 2) {
            // a run of one-child spawns (the later generations) puts child
            // gid at spawn lo + (gid - offs[lo]): try that spawn first, two
            // dependent loads instead of a search of six or seven
            const uint64_t o0 = offs[lo];
            const uint32_t g = (uint32_t)min((uint64_t)(hi - 2), (uint64_t)lo + (gid - o0));
            const uint64_t og = offs[g], og1 = offs[g + 1];
            if (og <= gid && gid < og1) { lo = g; hi = g + 1; }
          }
          while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (offs[mid] <= gid) lo = mid; else hi = mid;
          }
          si = lo;
          RayS r;
          Spawn srec;
          uint64_t soff;
          load_spawn(in, offs, si, srec, soff);
          if (!build_ray(c, srec, (uint32_t)(gid - soff), r)) {
            TraceOut to;
            to.spawn = si; to.pad = 0; to.prim = -2; to.t = 0;
            to.dx = to.dy = to.dz = 0.f;
            outp[my] = to;
          } else {
            wave_count(c.nrays);
            tcur = (int32_t)my;
            if (BATCH) S = scene_of(Sarg, c, r.scene);
            o = r.o;
            d = r.d;
            h.t = FHUGE; h.rod = 1.0; h.prim = -1;
            big_hits<BS>(S, o, d, h, nprim, tes + lane);
            node = DONE;
            leaf = -1;
            sp = 0;
            iter = 0;
            if (S.nbvh > 0) {
              if (S.root < 0) { hit_prim(S, S.leaf_prim[~S.root], o, d, h); ++nprim; }
              else node = S.root;
            }
            auto inv = [](double v) -> float {
              float f = (float)v;
              if (fabsf(f) < 1e-20f) f = f < 0 ? -1e-20f : 1e-20f;
              return 1.0f / f;
            };
            ix = inv(d.x); iy = inv(d.y); iz = inv(d.z);
            ox = (float)o.x * ix; oy = (float)o.y * iy; oz = (float)o.z * iz;   // origin x inverse direction
          }
        }
      }
    }
    if (__ballot(tcur >= 0) == 0ull) {
      if (exhausted) break;
      continue;
    }
    // ---- walk: internal nodes until every lane holds a postponed leaf ----
    auto pop = [&](float lim) -> int32_t {
      while (sp > 0) {
        --sp;
        int32_t cand;
        float ct;
        if (sp < LSTACK) {
          cand = stk[sp * BS + lane];
          ct = stn[sp * BS + lane];
        } else {
          uint64_t e = deep[sp - LSTACK];
          cand = (int32_t)(uint32_t)e;
          ct = __uint_as_float((uint32_t)(e >> 32));
        }
        if (ct <= lim) return cand;
      }
      return DONE;
    };
    auto push = [&](int32_t ref, float t) -> bool {
      if (sp < LSTACK) { stk[sp * BS + lane] = ref; stn[sp * BS + lane] = t; }
      else if (sp < LSTACK + PSTACK) deep[sp - LSTACK] = pack_entry(ref, t);
      else return false;
      ++sp;
      return true;
    };
    if (WIDE) {
      const BVH4Node* __restrict__ n4 = S.nodes4;
      while (node >= 0) {
        if (++iter > S.nbvh + 8) { atomicOr(c.err, ERR_STACK); node = DONE; leaf = -1; break; }
        const BVH4Node& N = n4[node];
        ++nvis;
        const float tmax = (float)(h.t + 2 * FTINY) * 1.000001f;
        const float kInf = __int_as_float(0x7f800000);
        float t0, t1, t2, t3;
        int32_t c0r = N.child[0], c1r = N.child[1], c2r = N.child[2], c3r = N.child[3];
        {
          float tk, lo[3], hi[3];
          lo[0] = N.lox[0]; lo[1] = N.loy[0]; lo[2] = N.loz[0]; hi[0] = N.hix[0]; hi[1] = N.hiy[0]; hi[2] = N.hiz[0];
          t0 = (slab(lo, hi, ox, oy, oz, ix, iy, iz, tmax, tk) & (c0r != DONE)) ? tk : kInf;
          lo[0] = N.lox[1]; lo[1] = N.loy[1]; lo[2] = N.loz[1]; hi[0] = N.hix[1]; hi[1] = N.hiy[1]; hi[2] = N.hiz[1];
          t1 = (slab(lo, hi, ox, oy, oz, ix, iy, iz, tmax, tk) & (c1r != DONE)) ? tk : kInf;
          lo[0] = N.lox[2]; lo[1] = N.loy[2]; lo[2] = N.loz[2]; hi[0] = N.hix[2]; hi[1] = N.hiy[2]; hi[2] = N.hiz[2];
          t2 = (slab(lo, hi, ox, oy, oz, ix, iy, iz, tmax, tk) & (c2r != DONE)) ? tk : kInf;
          lo[0] = N.lox[3]; lo[1] = N.loy[3]; lo[2] = N.loz[3]; hi[0] = N.hix[3]; hi[1] = N.hiy[3]; hi[2] = N.hiz[3];
          t3 = (slab(lo, hi, ox, oy, oz, ix, iy, iz, tmax, tk) & (c3r != DONE)) ? tk : kInf;
        }
        // ascending entry distance, equal distances in slot order: five
        // compare-exchanges on (distance, slot)
        int s0 = 0, s1 = 1, s2 = 2, s3 = 3;
        auto cx = [](float& ta, int32_t& ca, int& sa, float& tb, int32_t& cb, int& sb) {
          const bool sw = (ta > tb) | ((ta == tb) & (sa > sb));
          if (sw) {
            float tt = ta; ta = tb; tb = tt;
            int32_t cc = ca; ca = cb; cb = cc;
            int ss = sa; sa = sb; sb = ss;
          }
        };
        cx(t0, c0r, s0, t1, c1r, s1);
        cx(t2, c2r, s2, t3, c3r, s3);
        cx(t0, c0r, s0, t2, c2r, s2);
        cx(t1, c1r, s1, t3, c3r, s3);
        cx(t1, c1r, s1, t2, c2r, s2);
        if (t0 == kInf) {
          node = pop(tmax);
        } else {
          bool ok = true;
          if (t3 != kInf) ok &= push(c3r, t3);
          if (t2 != kInf) ok &= push(c2r, t2);
          if (t1 != kInf) ok &= push(c1r, t1);
          if (!ok) { atomicOr(c.err, ERR_STACK); node = DONE; leaf = -1; break; }
          node = c0r;
        }
        if ((node < 0) & (node != DONE) & (leaf < 0)) {
          leaf = ~node;
          node = pop((float)(h.t + 2 * FTINY) * 1.000001f);
        }
        if (__ballot(leaf < 0) == 0ull) break;
      }
    }
    const BVHNode* __restrict__ nodes = S.nodes;
    while (!WIDE && node >= 0) {
      if (++iter > S.nbvh + 8) { atomicOr(c.err, ERR_STACK); node = DONE; leaf = -1; break; }
      const BVHNode N = nodes[node];
      ++nvis;
      float tmax = (float)(h.t + 2 * FTINY) * 1.000001f;
      float tl, tr;
      bool hl = slab(N.lmin, N.lmax, ox, oy, oz, ix, iy, iz, tmax, tl);
      bool hr = slab(N.rmin, N.rmax, ox, oy, oz, ix, iy, iz, tmax, tr);
      if (hl & hr) {
        bool lfirst = tl <= tr;
        int32_t farc = lfirst ? N.right : N.left;
        float fart = lfirst ? tr : tl;
        if (sp < LSTACK) { stk[sp * BS + lane] = farc; stn[sp * BS + lane] = fart; }
        else if (sp < LSTACK + PSTACK) deep[sp - LSTACK] = pack_entry(farc, fart);
        else { atomicOr(c.err, ERR_STACK); node = DONE; leaf = -1; break; }
        ++sp;
        node = lfirst ? N.left : N.right;
      } else if (hl) {
        node = N.left;
      } else if (hr) {
        node = N.right;
      } else {
        node = pop(tmax);
      }
      if ((node < 0) & (node != DONE) & (leaf < 0)) {
        leaf = ~node;
        node = pop((float)(h.t + 2 * FTINY) * 1.000001f);
      }
      if (__ballot(leaf < 0) == 0ull) break;
    }
    while (leaf >= 0) {
      hit_prim(S, leaf, o, d, h);
      ++nprim;
      leaf = -1;
      if ((node < 0) & (node != DONE)) {
        leaf = ~node;
        node = pop((float)(h.t + 2 * FTINY) * 1.000001f);
      }
    }
    // ---- finished rays: write the record, free the lane ----
    if ((tcur >= 0) & (node == DONE) & (leaf < 0)) {
      TraceOut to;
      to.spawn = si; to.pad = 0;
      to.t = h.t;
      to.prim = h.prim;
      to.dx = (float)d.x; to.dy = (float)d.y; to.dz = (float)d.z;
      outp[tcur] = to;
      tcur = -1;
    }
  }
  if (c.visits) {
    atomicAdd(c.visits, (unsigned long long)nvis);
    atomicAdd(c.visits + 1, (unsigned long long)nprim);
  }
}

static void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// First intersection only (what rtrace -oLN prints: rt/rtrace.c:758-804):
// distance (FHUGE on a miss), the primitive's unflipped normal, its index.
__global__ void __launch_bounds__(TBLOCK) k_first_hit(DevScene S, const double* rays, uint32_t n, double* t_out,
                                                      double* n_out, int32_t* p_out, uint32_t* err) {
  __shared__ int32_t stk[LSTACK * TBLOCK];
  __shared__ float stn[LSTACK * TBLOCK];
  __shared__ float tes[NBIG_ORDER * TBLOCK];
  uint32_t i = blockIdx.x * TBLOCK + threadIdx.x;
  if (i >= n) return;
  V3 o = load3(rays + 6 * (size_t)i), d = load3(rays + 6 * (size_t)i + 3);
  normalize(d);
  HitRec h;
  closest_hit<TBLOCK>(S, o, d, h, stk, stn, tes, err, nullptr);
  V3 nrm = v3(0, 0, 0);
  if (h.prim >= 0) nrm = prim_normal(S, h.prim, o + d * h.t);
  t_out[i] = h.prim >= 0 ? h.t : FHUGE;
  n_out[3 * (size_t)i + 0] = nrm.x;
  n_out[3 * (size_t)i + 1] = nrm.y;
  n_out[3 * (size_t)i + 2] = nrm.z;
  p_out[i] = h.prim;
}

void first_hit(const DevScene& S, const double* d_rays, uint32_t n, double* d_t, double* d_n, int32_t* d_p,
               uint32_t* d_err, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_first_hit, dim3((n + TBLOCK - 1) / TBLOCK), dim3(TBLOCK), 0, st, S, d_rays, n, d_t, d_n, d_p,
                     d_err);
  check(hipGetLastError(), "k_first_hit launch");
}

// Shadow-cache fill (castshadow() and initobscache() of rt/srcobstr.c:28-183):
// each cast ray walks from its start point until it meets an object that
// srcblocker() (:298-320) would record - opaque material, not the source
// itself, not part of a mesh or an instance - and files that object under
// the cell its reversed ray falls in.  Passing a transparent or unrecorded
// object costs another closest-hit query, as in the reference.  Cells of
// two rays: the later ray of the reference's order wins (the sequence
// number sits in the high word).  counts[source] += closest-hit queries.
__global__ void __launch_bounds__(TBLOCK) k_shadow_cast(DevScene S, const double* rays, const int32_t* rsrc, uint32_t n,
                                                        unsigned long long* obs, unsigned long long* counts,
                                                        uint32_t* err) {
  __shared__ int32_t stk[LSTACK * TBLOCK];
  __shared__ float stn[LSTACK * TBLOCK];
  __shared__ float tes[NBIG_ORDER * TBLOCK];
  const uint32_t i = blockIdx.x * TBLOCK + threadIdx.x;
  if (i >= n) return;
  const int sn = rsrc[i];
  const Source& src = S.srcs[sn];
  const SrcCache& C = S.scache[sn];
  const uint32_t seq = i;
  V3 o = load3(rays + 6 * (size_t)i), d = load3(rays + 6 * (size_t)i + 3);
  unsigned long long nq = 0;
  for (int pass = 0; pass < 4096; ++pass) {
    HitRec h;
    closest_hit<TBLOCK>(S, o, d, h, stk, stn, tes, err, nullptr);
    ++nq;
    if (h.prim < 0) break;
    const V3 rop = o + d * h.t;
    const Prim& P = S.prims[h.prim];
    int mk = P.mat >= 0 ? S.mats[P.mat].kind : MK_NONE;
    const bool opaque = mk == MK_PLASTIC || mk == MK_METAL || mk == MK_LIGHT || mk == MK_GLOW;
    if (opaque && !(P.flat & PF_COMPLEX) && h.prim != src.prim) {
      const int cell = shadow_cell(src, C, rop - d, -d);
      if (cell >= 0)
        atomicMax(obs + C.cell0 + cell, ((unsigned long long)(seq + 1) << 32) | (uint32_t)h.prim);
      break;
    }
    o = rop + d * (10. * FTINY);
  }
  atomicAdd(counts + sn, nq);
}

void shadow_cast(const DevScene& S, const double* d_rays, const int32_t* d_src, uint32_t n, unsigned long long* d_obs,
                 unsigned long long* d_counts, uint32_t* d_err, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_shadow_cast, dim3((n + TBLOCK - 1) / TBLOCK), dim3(TBLOCK), 0, st, S, d_rays, d_src, n, d_obs,
                     d_counts, d_err);
  check(hipGetLastError(), "k_shadow_cast launch");
}

// One generation, part 2: shade the hits and append the next spawn records.
// pin: path records of `in` (adjoint replay only).
#ifndef RG_SHADE_WAVES
#define RG_SHADE_WAVES 2     // waves per SIMD the register budget of k_shade targets
#endif
template <bool ADJ>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(RG_SHADE_WAVES))) k_shade(DevScene Sarg, RunCfg c, const Spawn* in, const PathRec* pin,
                                                 const uint64_t* offs, uint64_t c0, uint32_t nthreads,
                                                 const TraceOut* tr, Sink sk) {
  uint32_t t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= nthreads) return;
  c.adj = ADJ;
  if (!ADJ) sk.pbuf = nullptr;
  // The loads a shaded ray starts with are issued together and waited for
  // together: the 32-byte trace record (two 16-byte loads), then its spawn
  // record, the spawn's first child index and, with one scene, the hit
  // primitive's record.  The empty asm statements take the loaded vectors as
  // inputs, which keeps the compiler from sinking a load below the early exit
  // or to its first use, where each would cost a round trip of its own.
  TraceOut o;
  {
    static_assert(sizeof(TraceOut) == 2 * sizeof(u32x4), "TraceOut layout");
    const u32x4* tp = reinterpret_cast<const u32x4*>(tr + t);
    const u32x4 a = tp[0], b = tp[1];
    asm volatile("" :: "v"(a), "v"(b));
    __builtin_memcpy(&o, &a, sizeof a);
    __builtin_memcpy(reinterpret_cast<char*>(&o) + sizeof a, &b, sizeof b);
  }
  if (o.prim == -2) return;
  const uint32_t si = o.spawn;
  Spawn sp;
  uint64_t off;
  Prim P{};
  const bool early = !c.nscenes && o.prim >= 0;
  {
    static_assert(sizeof(Spawn) == 6 * sizeof(u32x4), "Spawn layout");
    const u32x4* spp = reinterpret_cast<const u32x4*>(in + si);
    const u32x4 s0 = spp[0], s1 = spp[1], s2 = spp[2], s3 = spp[3], s4 = spp[4], s5 = spp[5];
    off = offs[si];
    u32x4 p0 = {0u, 0u, 0u, 0u}, p1 = {0u, 0u, 0u, 0u};
    if (early) {
      const u32x4* pp = reinterpret_cast<const u32x4*>(Sarg.prims + o.prim);
      p0 = pp[0];
      p1 = pp[1];
    }
    asm volatile("" :: "v"(s0), "v"(s1), "v"(s2), "v"(s3), "v"(s4), "v"(s5), "v"(off), "v"(p0), "v"(p1));
    const u32x4 sv[6] = {s0, s1, s2, s3, s4, s5};
    __builtin_memcpy(&sp, sv, sizeof sp);
    static_assert(sizeof(Prim) == 2 * sizeof(u32x4), "Prim layout");
    const u32x4 pv[2] = {p0, p1};
    __builtin_memcpy(&P, pv, sizeof P);
  }
  RayS r;
  build_ray(c, sp, (uint32_t)(c0 + t - off), r, ADJ ? pin + si : nullptr);
  HitRec h;
  h.t = o.t;
  h.prim = o.prim;
  h.rod = 1.0;
  const DevScene& S = scene_of(Sarg, c, r.scene);
  if (h.prim >= 0) {
    if (!early) P = S.prims[h.prim];
    shade_hit(S, c, r, h, sk, P);
  } else {
    shade_miss(S, c, r);
  }
}

// ---- grouping a batched generation by scene ----------------------------------
// Between generations of a batched launch the spawn records are regrouped by
// scene (counting sort on Spawn.scene), so that the next windows trace one
// scene at a time instead of interleaving every scene's tables in the caches.
constexpr int SORT_BLOCK = 256;
constexpr int SORT_ITEMS = 4;

constexpr int NSC = kMaxBatchScenes;        // histogram bins (scenes)
constexpr int SC_PER_T = NSC / SORT_BLOCK;   // bins per thread

__global__ void __launch_bounds__(SORT_BLOCK) k_scene_hist(const Spawn* in, uint32_t n, uint32_t* hist) {
  __shared__ uint32_t h[NSC];
  for (int q = 0; q < SC_PER_T; ++q) h[threadIdx.x + q * SORT_BLOCK] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * SORT_BLOCK * SORT_ITEMS + threadIdx.x;
  for (int k = 0; k < SORT_ITEMS; ++k) {
    size_t i = base + (size_t)k * SORT_BLOCK;
    if (i < n) atomicAdd(&h[in[i].scene], 1u);
  }
  __syncthreads();
  for (int q = 0; q < SC_PER_T; ++q) {
    const int b = threadIdx.x + q * SORT_BLOCK;
    if (h[b]) atomicAdd(&hist[b], h[b]);
  }
}

// exclusive scan of the NSC scene counts (one block of 1024, 4 each)
__global__ void __launch_bounds__(1024) k_scene_start(const uint32_t* hist, uint32_t* start) {
  static_assert(NSC == 4 * 1024, "scan shape");
  __shared__ uint32_t sh[1024];
  const int t = threadIdx.x;
  const uint32_t v0 = hist[4 * t], v1 = hist[4 * t + 1], v2 = hist[4 * t + 2], v3 = hist[4 * t + 3];
  const uint32_t tot = v0 + v1 + v2 + v3;
  sh[t] = tot;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    uint32_t a = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  const uint32_t ex = sh[t] - tot;
  start[4 * t] = ex;
  start[4 * t + 1] = ex + v0;
  start[4 * t + 2] = ex + v0 + v1;
  start[4 * t + 3] = ex + v0 + v1 + v2;
}

// Each block reserves one range per scene it holds, then places its records
// (and their adjoint path records) at LDS-ranked positions in that range.
__global__ void __launch_bounds__(SORT_BLOCK) k_scene_scatter(const Spawn* in, const PathRec* pin, uint32_t n,
                                                              const uint32_t* start, uint32_t* fill, Spawn* out,
                                                              PathRec* pout) {
  __shared__ uint32_t h[NSC];
  __shared__ uint32_t b[NSC];
  for (int q = 0; q < SC_PER_T; ++q) h[threadIdx.x + q * SORT_BLOCK] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * SORT_BLOCK * SORT_ITEMS + threadIdx.x;
  uint32_t sc[SORT_ITEMS], rank[SORT_ITEMS];
  for (int k = 0; k < SORT_ITEMS; ++k) {
    size_t i = base + (size_t)k * SORT_BLOCK;
    sc[k] = i < n ? in[i].scene : 0u;
    rank[k] = i < n ? atomicAdd(&h[sc[k]], 1u) : 0u;
  }
  __syncthreads();
  for (int q = 0; q < SC_PER_T; ++q) {
    const int c = threadIdx.x + q * SORT_BLOCK;
    if (h[c]) b[c] = start[c] + atomicAdd(&fill[c], h[c]);
  }
  __syncthreads();
  for (int k = 0; k < SORT_ITEMS; ++k) {
    size_t i = base + (size_t)k * SORT_BLOCK;
    if (i < n) {
      uint32_t pos = b[sc[k]] + rank[k];
      out[pos] = in[i];
      if (pin) pout[pos] = pin[i];
    }
  }
}

// ---- exclusive scan of uint32 counts (three-phase, LDS per block) ----------
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
    uint32_t a = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  total = sh[SCAN_BLOCK - 1];
  uint32_t incl = sh[t];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_tiles(const uint32_t* in, uint32_t* out, uint32_t n,
                                                           uint32_t* tile_sums) {
  __shared__ uint32_t sh[SCAN_BLOCK];
  size_t base = (size_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS], s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) { v[k] = base + k < n ? in[base + k] : 0; s += v[k]; }
  uint32_t total;
  uint32_t ex = block_excl_scan(s, sh, total);
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_add(uint32_t* out, uint32_t n, const uint32_t* tile_offs) {
  size_t base = (size_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint32_t a = tile_offs[blockIdx.x];
  for (int k = 0; k < SCAN_ITEMS; ++k)
    if (base + k < n) out[base + k] += a;
}

// The same in 64 bits, for child counts whose totals can pass 2^32 (a
// million -I sensors at -ad 4096 already spawn 2^32 children): an inclusive
// scan within each wavefront by lane shuffles, the wavefront totals through
// LDS, two barriers per tile.  sh holds at least SCAN_BLOCK / 64 entries.
__device__ __forceinline__ uint64_t block_excl_scan64(uint64_t v, uint64_t* sh, uint64_t& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t a = __shfl_up(incl, off, 64);
    if (lane >= off) incl += a;
  }
  if (lane == 63) sh[w] = incl;
  __syncthreads();
  uint64_t pre = 0;
  total = 0;
#pragma unroll
  for (int k = 0; k < SCAN_BLOCK / 64; ++k) {
    const uint64_t x = sh[k];
    pre += k < w ? x : 0;
    total += x;
  }
  __syncthreads();      // sh is rewritten by the next call
  return pre + incl - v;
}

template <class TI>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_tiles64(const TI* in, uint64_t* out, uint32_t n,
                                                             uint64_t* tile_sums) {
  __shared__ uint64_t sh[SCAN_BLOCK];
  size_t base = (size_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint64_t v[SCAN_ITEMS], s = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) { v[k] = base + k < n ? (uint64_t)in[base + k] : 0; s += v[k]; }
  uint64_t total;
  uint64_t ex = block_excl_scan64(s, sh, total);
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_add64(uint64_t* out, uint32_t n, const uint64_t* tile_offs) {
  size_t base = (size_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint64_t a = tile_offs[blockIdx.x];
  for (int k = 0; k < SCAN_ITEMS; ++k)
    if (base + k < n) out[base + k] += a;
}

}  // namespace rg

// ---- host side -----------------------------------------------------------------
namespace rg {


// out[i] = sum(in[0..i)), out[n] = total.  `out` holds n+1 entries.
void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, DevArena& arena, hipStream_t st, int level) {
  if (n == 0) {
    check(hipMemsetAsync(out, 0, sizeof(uint32_t), st), "scan memset");
    return;
  }
  uint32_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  uint32_t* sums = (uint32_t*)arena.get("scan_sums" + std::to_string(level), sizeof(uint32_t) * (ntiles + 1), st, true);
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(SCAN_BLOCK), 0, st, in, out, n, sums);
  if (ntiles > 1) {
    uint32_t* soffs = (uint32_t*)arena.get("scan_offs" + std::to_string(level), sizeof(uint32_t) * (ntiles + 1), st, true);
    exclusive_scan(sums, soffs, ntiles, arena, st, level + 1);
    hipLaunchKernelGGL(k_scan_add, dim3(ntiles), dim3(SCAN_BLOCK), 0, st, out, n, soffs);
    check(hipMemcpyAsync(out + n, soffs + ntiles, sizeof(uint32_t), hipMemcpyDeviceToDevice, st), "scan total");
  } else {
    check(hipMemcpyAsync(out + n, sums, sizeof(uint32_t), hipMemcpyDeviceToDevice, st), "scan total");
  }
}

// The same scan with its length n = min(*d_n, cap) read on the device, so
// that it follows the kernel that produced the counts without a host round
// trip: persistent blocks walk the tiles, one block scans the tile sums (and
// writes the total to *d_total and out[n]), persistent blocks add the tile
// offsets.
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_tiles_dev(const uint32_t* in, uint64_t* out, const uint32_t* d_n,
                                                               uint32_t cap, uint64_t* tile_sums) {
  __shared__ uint64_t sh[SCAN_BLOCK];
  const uint32_t n = min(*d_n, cap);
  const uint32_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const size_t base = (size_t)tile * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint64_t v[SCAN_ITEMS], s = 0;
    for (int k = 0; k < SCAN_ITEMS; ++k) { v[k] = base + k < n ? (uint64_t)in[base + k] : 0; s += v[k]; }
    uint64_t total;
    uint64_t ex = block_excl_scan64(s, sh, total);
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      if (base + k < n) out[base + k] = ex;
      ex += v[k];
    }
    if (threadIdx.x == 0) tile_sums[tile] = total;
  }
}

constexpr int SUMS_BLOCK = 1024;
__global__ void __launch_bounds__(SUMS_BLOCK) k_scan_sums_dev(const uint32_t* d_n, uint32_t cap, uint64_t* tile_sums,
                                                              uint64_t* out, uint64_t* d_total) {
  __shared__ uint64_t sh[SUMS_BLOCK / 64];
  const uint32_t n = min(*d_n, cap);
  const uint32_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  const uint32_t per = (ntiles + SUMS_BLOCK - 1) / SUMS_BLOCK;
  const uint32_t t = threadIdx.x, b0 = min(t * per, ntiles), b1 = min(b0 + per, ntiles);
  uint64_t s = 0;
  for (uint32_t k = b0; k < b1; ++k) s += tile_sums[k];
  // exclusive scan of the per-thread sums: lane shuffles, then the 16
  // wavefront totals through LDS
  const int lane = t & 63, w = t >> 6;
  uint64_t incl = s;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t a = __shfl_up(incl, off, 64);
    if (lane >= off) incl += a;
  }
  if (lane == 63) sh[w] = incl;
  __syncthreads();
  uint64_t pre = 0, total = 0;
#pragma unroll
  for (int k = 0; k < SUMS_BLOCK / 64; ++k) {
    const uint64_t x = sh[k];
    pre += k < w ? x : 0;
    total += x;
  }
  uint64_t ex = pre + incl - s;
  for (uint32_t k = b0; k < b1; ++k) {
    const uint64_t v = tile_sums[k];
    tile_sums[k] = ex;
    ex += v;
  }
  if (t == SUMS_BLOCK - 1) {
    out[n] = total;
    *d_total = total;
  }
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_add_dev(uint64_t* out, const uint32_t* d_n, uint32_t cap,
                                                             const uint64_t* tile_offs) {
  const uint32_t n = min(*d_n, cap);
  const uint32_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const size_t base = (size_t)tile * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    const uint64_t a = tile_offs[tile];
    for (int k = 0; k < SCAN_ITEMS; ++k)
      if (base + k < n) out[base + k] += a;
  }
}

// out[i] = sum(in[0..i)) in 64 bits, out[n] = total.  `out` holds n+1 entries.
template <class TI>
static void exclusive_scan64(const TI* in, uint64_t* out, uint32_t n, DevArena& arena, hipStream_t st, int level) {
  if (n == 0) {
    check(hipMemsetAsync(out, 0, sizeof(uint64_t), st), "scan memset");
    return;
  }
  uint32_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  uint64_t* sums = (uint64_t*)arena.get("scan64_sums" + std::to_string(level), sizeof(uint64_t) * (ntiles + 1), st, true);
  hipLaunchKernelGGL(k_scan_tiles64<TI>, dim3(ntiles), dim3(SCAN_BLOCK), 0, st, in, out, n, sums);
  if (ntiles > 1) {
    uint64_t* soffs = (uint64_t*)arena.get("scan64_offs" + std::to_string(level), sizeof(uint64_t) * (ntiles + 1), st, true);
    exclusive_scan64<uint64_t>(sums, soffs, ntiles, arena, st, level + 1);
    hipLaunchKernelGGL(k_scan_add64, dim3(ntiles), dim3(SCAN_BLOCK), 0, st, out, n, soffs);
    check(hipMemcpyAsync(out + n, soffs + ntiles, sizeof(uint64_t), hipMemcpyDeviceToDevice, st), "scan total");
  } else {
    check(hipMemcpyAsync(out + n, sums, sizeof(uint64_t), hipMemcpyDeviceToDevice, st), "scan total");
  }
}

void render(const DevScene& S, RunCfg cfg, const double* d_rays, uint32_t nsens, bool imm_irrad, int max_srcs,
            const Tuning& tn, DevArena& arena, hipStream_t st, RenderStats* stats) {
  // Work proceeds in waves of at most `window` children; each child appends
  // at most `per_child` spawn records, so a wave never overflows its buffer.
  const uint32_t per_child = 6u + (uint32_t)std::max(max_srcs, S.nsrcs);
  const uint32_t window = std::max<uint32_t>(tn.window, 1024u);
  int dev = 0, ncu = 256;
  check(hipGetDevice(&dev), "device");
  check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "device attribute");
  const uint32_t pf_blocks = (uint32_t)std::max(1, ncu) * std::max<uint32_t>(tn.blocks_per_cu, 1u);
  const long max_batches = 1L << 22;
  uint32_t* d_cnt = (uint32_t*)arena.get("ctr", 64, st, true);
  uint32_t* h_cnt = arena.host_u32(16);   // [0, 8): frame totals and seeds; [8, 16): windows in flight

  const bool adj = cfg.adj != 0;
  struct Batch { Spawn* buf; PathRec* pbuf; uint32_t n; int slot; };
  int nslots = 0;
  std::vector<int> free_slots;
  auto new_slot = [&]() -> int {
    if (!free_slots.empty()) { int s = free_slots.back(); free_slots.pop_back(); return s; }
    return nslots++;
  };
  auto slot_buf = [&](int s, size_t cap) -> Spawn* {
    return (Spawn*)arena.get("spawn" + std::to_string(s), sizeof(Spawn) * cap, st, false);
  };
  // adjoint replay: one path record beside every spawn record
  auto slot_path = [&](int s, size_t cap) -> PathRec* {
    return adj ? (PathRec*)arena.get("path" + std::to_string(s), sizeof(PathRec) * cap, st, false) : nullptr;
  };
  // Spawn buffers: at most slot_bytes each (a window's children each append
  // at most per_child records), so capacities fit 32 bits and scenes with
  // many light sources shrink their windows instead of failing to allocate.
  // A slot is held to 1/24 of the device's memory (about six slots are live
  // at the deepest point of a ray tree, each with its scan and trace
  // records), and to 16 GiB.
  size_t dev_mem = 0;
  check(hipDeviceTotalMem(&dev_mem, dev), "device memory");
  const size_t slot_bytes = std::max<size_t>(size_t(64) << 20, std::min<size_t>(size_t(16) << 30, dev_mem / 24));
  const uint32_t slot_children = (uint32_t)std::max<size_t>(64, slot_bytes / (sizeof(Spawn) * per_child));
  const uint32_t win = std::min(window, slot_children);
  long gens = 0;
  // Events: two per window in flight (its read-back, and the trace / shade
  // timing when asked for), reused in turn.
  struct Events {
    std::vector<hipEvent_t> ev;
    ~Events() { for (hipEvent_t e : ev) (void)hipEventDestroy(e); }
  } evs;
  const bool timing = stats && stats->timing;
  constexpr int kInFlight = 2;
  for (int k = 0; k < kInFlight; ++k) {
    hipEvent_t e;
    check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
    evs.ev.push_back(e);
    if (timing)
      for (int q = 0; q < 3; ++q) {
        check(hipEventCreate(&e), "event");
        evs.ev.push_back(e);
      }
  }
  const int ev_per = timing ? 4 : 1;
  static_assert(8 + 4 * kInFlight <= 16, "host counters");
  uint32_t* h_win = h_cnt + 8;   // per in-flight window: spawn count, work counter, child total (64 bits)
  // Depth first, one window at a time: after each window its children are
  // opened and walked before the next window of the same batch, so live
  // spawn buffers are bounded by the depth of the ray tree, not by the size
  // of a generation.
  struct Frame { Batch bt; uint64_t* offs; uint64_t total; uint64_t c0; };
  std::vector<Frame> frames;
  // compact child counts and their exclusive scan, per slot
  auto slot_counts = [&](int s, size_t cap) -> uint32_t* {
    return (uint32_t*)arena.get("counts" + std::to_string(s), sizeof(uint32_t) * (cap + 1), st, false);
  };
  auto slot_offs = [&](int s, size_t cap) -> uint64_t* {
    return (uint64_t*)arena.get("offs" + std::to_string(s), sizeof(uint64_t) * (cap + 1), st, false);
  };
  // A batch whose size the host knows: scan its child counts (recounted
  // from the records after a scene regroup), read the total.
  auto open_frame = [&](const Batch& b, bool recount) {
    if (b.n == 0) { free_slots.push_back(b.slot); return; }
    if (++gens > max_batches) throw std::runtime_error("ray tree did not terminate");
    uint32_t* counts = slot_counts(b.slot, b.n);
    uint64_t* offs = slot_offs(b.slot, b.n);
    if (recount) hipLaunchKernelGGL(k_counts, dim3((b.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, b.buf, b.n, counts);
    exclusive_scan64<uint32_t>(counts, offs, b.n, arena, st, 0);
    check(hipMemcpyAsync(h_cnt, offs + b.n, sizeof(uint64_t), hipMemcpyDeviceToHost, st), "total");
    check(hipStreamSynchronize(st), "scan");
    const uint64_t total = (uint64_t)h_cnt[0] | ((uint64_t)h_cnt[1] << 32);   // little-endian 64-bit total
    frames.push_back({b, offs, total, 0});
  };
  const uint32_t scan_grid = 2u * (uint32_t)std::max(1, ncu);
  // Generation 0 per chunk of sensors (buffers are sized to the work at hand
  // and kept by the arena for later calls, so a context that only ever
  // traces a few sensors stays small); each chunk's tree is walked to the
  // end before the next chunk is seeded.
  const uint32_t chunk = std::max<uint32_t>(1u, std::min<uint32_t>(nsens, slot_children));
  for (uint32_t s0 = 0; s0 < nsens; s0 += chunk) {
    const uint32_t ns = std::min(chunk, nsens - s0);
    const size_t batch_cap = (size_t)ns * per_child;
    const double* rays0 = d_rays + 6 * (size_t)s0;
    {
      int s = new_slot();
      Spawn* b = slot_buf(s, batch_cap);
      PathRec* pb = slot_path(s, batch_cap);
      uint32_t* bc = slot_counts(s, batch_cap);
      check(hipMemsetAsync(d_cnt, 0, sizeof(uint32_t), st), "memset");
      uint32_t grid = (ns + BLOCK - 1) / BLOCK;
      if (imm_irrad) {
        Sink sk{b, d_cnt, (uint32_t)batch_cap, cfg.err, pb, bc};
        if (adj) hipLaunchKernelGGL(k_irrad_seed<true>, dim3(grid), dim3(BLOCK), 0, st, S, cfg, rays0, ns, s0, sk);
        else hipLaunchKernelGGL(k_irrad_seed<false>, dim3(grid), dim3(BLOCK), 0, st, S, cfg, rays0, ns, s0, sk);
        if (cfg.nscenes <= 1) {
          // child counts scanned on the device: one read-back of the spawn
          // count and the child total
          uint64_t* offs = slot_offs(s, batch_cap);
          uint64_t* sums =
              (uint64_t*)arena.get("scan_dev_sums", sizeof(uint64_t) * (batch_cap / SCAN_TILE + 2), st, false);
          hipLaunchKernelGGL(k_scan_tiles_dev, dim3(scan_grid), dim3(SCAN_BLOCK), 0, st, bc, offs, d_cnt,
                             (uint32_t)batch_cap, sums);
          hipLaunchKernelGGL(k_scan_sums_dev, dim3(1), dim3(SUMS_BLOCK), 0, st, d_cnt, (uint32_t)batch_cap, sums, offs,
                             (uint64_t*)(d_cnt + 2));
          hipLaunchKernelGGL(k_scan_add_dev, dim3(scan_grid), dim3(SCAN_BLOCK), 0, st, offs, d_cnt,
                             (uint32_t)batch_cap, sums);
          check(hipGetLastError(), "scan launch");
          check(hipMemcpyAsync(h_cnt, d_cnt, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "count");
          check(hipStreamSynchronize(st), "seed");
          const uint32_t nseed = std::min<uint32_t>(h_cnt[0], (uint32_t)batch_cap);
          if (nseed) {
            if (++gens > max_batches) throw std::runtime_error("ray tree did not terminate");
            const uint64_t total = (uint64_t)h_cnt[2] | ((uint64_t)h_cnt[3] << 32);
            frames.push_back({Batch{b, pb, nseed, s}, offs, total, 0});
          } else {
            free_slots.push_back(s);
          }
        } else {
          check(hipMemcpyAsync(h_cnt, d_cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, st), "count");
          check(hipStreamSynchronize(st), "seed");
          open_frame({b, pb, std::min<uint32_t>(h_cnt[0], (uint32_t)batch_cap), s}, false);
        }
      } else {
        hipLaunchKernelGGL(k_primary_seed, dim3(grid), dim3(BLOCK), 0, st, cfg, rays0, ns, s0, b, bc);
        if (pb) check(hipMemsetAsync(pb, 0, sizeof(PathRec) * ns, st), "memset");   // empty paths
        open_frame({b, pb, ns, s}, false);
      }
    }
    // With tn.inflight = 2 windows are pipelined: while the GPU runs one, the
    // host waits for the counts of the one before and pushes its children's
    // frame.  All work is on one stream, so windows run in the order they
    // were queued and a slot freed on the host is only overwritten after the
    // kernels reading it.  Batched launches regroup each window's spawns by
    // scene before their frame opens, and run one window at a time.
    struct Pending { Batch out; size_t cap; uint32_t nth; int pair; };
    std::vector<Pending> inflight;
    int next_pair = 0;
    const size_t depth = cfg.nscenes <= 1 ? (size_t)std::max(1, std::min(tn.inflight, kInFlight)) : 1;
    auto launch_window = [&]() {
      const Batch bt = frames.back().bt;
      const uint64_t* offs = frames.back().offs;
      const uint64_t c0 = frames.back().c0;
      const uint64_t total = frames.back().total;
      uint32_t nth = (uint32_t)std::min<uint64_t>(win, total - c0);
      frames.back().c0 += nth;
      size_t cap = (size_t)nth * per_child;
      int s = new_slot();
      Spawn* ob = slot_buf(s, cap);
      PathRec* opb = slot_path(s, cap);
      uint32_t* ocnt = slot_counts(s, cap);
      uint64_t* ooffs = slot_offs(s, cap);
      TraceOut* tout = (TraceOut*)arena.get("trace_out", sizeof(TraceOut) * nth, st, false);
      const int pair = next_pair;
      next_pair = (next_pair + 1) % kInFlight;
      hipEvent_t* e = &evs.ev[pair * ev_per];
      check(hipMemsetAsync(d_cnt, 0, 2 * sizeof(uint32_t), st), "memset");   // spawn count, work counter
      Sink sk{ob, d_cnt, (uint32_t)cap, cfg.err, opb, ocnt};
      if (timing) check(hipEventRecord(e[1], st), "event");
      const uint32_t tblocks = (nth + TBLOCK - 1) / TBLOCK;
      if (tn.trace_kernel == 1) {
        const uint32_t nch = (nth + (1u << CHUNK_LOG) - 1) >> CHUNK_LOG;
        uint32_t* first = (uint32_t*)arena.get("chunk_spawn", sizeof(uint32_t) * (nch + 1), st, false);
        hipLaunchKernelGGL(k_chunk_spawn, dim3((nch + BLOCK) / BLOCK), dim3(BLOCK), 0, st, offs, bt.n, c0, nth, first);
        const uint32_t nb = std::min(tblocks, pf_blocks);
        if (cfg.nscenes && tn.wide)
          hipLaunchKernelGGL((k_trace_pf<TBLOCK, true, true>), dim3(nb), dim3(TBLOCK), 0, st, S, cfg, bt.buf, offs,
                             first, c0, nth, tout, d_cnt + 1, tn.refill);
        else if (cfg.nscenes)
          hipLaunchKernelGGL((k_trace_pf<TBLOCK, true, false>), dim3(nb), dim3(TBLOCK), 0, st, S, cfg, bt.buf, offs,
                             first, c0, nth, tout, d_cnt + 1, tn.refill);
        else if (tn.wide)
          hipLaunchKernelGGL((k_trace_pf<TBLOCK, false, true>), dim3(nb), dim3(TBLOCK), 0, st, S, cfg, bt.buf, offs,
                             first, c0, nth, tout, d_cnt + 1, tn.refill);
        else
          hipLaunchKernelGGL((k_trace_pf<TBLOCK, false, false>), dim3(nb), dim3(TBLOCK), 0, st, S, cfg, bt.buf, offs,
                             first, c0, nth, tout, d_cnt + 1, tn.refill);
      } else {
        hipLaunchKernelGGL(k_trace, dim3(tblocks), dim3(TBLOCK), 0, st, S, cfg, bt.buf, offs, bt.n, c0, nth, tout);
      }
      check(hipGetLastError(), "k_trace launch");
      if (timing) check(hipEventRecord(e[2], st), "event");
      if (adj)
        hipLaunchKernelGGL(k_shade<true>, dim3((nth + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, S, cfg, bt.buf,
                           bt.pbuf, offs, c0, nth, tout, sk);
      else
        hipLaunchKernelGGL(k_shade<false>, dim3((nth + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, S, cfg, bt.buf,
                           bt.pbuf, offs, c0, nth, tout, sk);
      check(hipGetLastError(), "k_shade launch");
      if (timing) check(hipEventRecord(e[3], st), "event");
      if (cfg.nscenes <= 1) {
        // the new batch's child counts scanned on the device right after
        // shading: one read-back (spawn count and child total) per window
        uint64_t* sums = (uint64_t*)arena.get("scan_dev_sums", sizeof(uint64_t) * (cap / SCAN_TILE + 2), st, false);
        hipLaunchKernelGGL(k_scan_tiles_dev, dim3(scan_grid), dim3(SCAN_BLOCK), 0, st, ocnt, ooffs, d_cnt,
                           (uint32_t)cap, sums);
        hipLaunchKernelGGL(k_scan_sums_dev, dim3(1), dim3(SUMS_BLOCK), 0, st, d_cnt, (uint32_t)cap, sums, ooffs,
                           (uint64_t*)(d_cnt + 2));
        hipLaunchKernelGGL(k_scan_add_dev, dim3(scan_grid), dim3(SCAN_BLOCK), 0, st, ooffs, d_cnt, (uint32_t)cap,
                           sums);
        check(hipGetLastError(), "scan launch");
      }
      check(hipMemcpyAsync(h_win + 4 * pair, d_cnt, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "count");
      check(hipEventRecord(e[0], st), "event");
      inflight.push_back({Batch{ob, opb, 0, s}, cap, nth, pair});
    };
    auto finish_window = [&]() {
      const Pending pw = inflight.front();
      inflight.erase(inflight.begin());
      hipEvent_t* e = &evs.ev[pw.pair * ev_per];
      check(hipEventSynchronize(e[0]), "bounce");
      const uint32_t* h = h_win + 4 * pw.pair;
      if (timing) {
        float ms = 0, ms2 = 0;
        check(hipEventElapsedTime(&ms, e[1], e[2]), "event time");
        check(hipEventElapsedTime(&ms2, e[2], e[3]), "event time");
        stats->bounce_ms += ms;
        stats->shade_ms += ms2;
      }
      if (stats) { stats->launches++; stats->threads += pw.nth; }
      const int s = pw.out.slot;
      const uint32_t nout = std::min<uint32_t>(h[0], (uint32_t)pw.cap);
      if (!nout) {
        free_slots.push_back(s);
        return;
      }
      if (cfg.nscenes <= 1) {
        if (++gens > max_batches) throw std::runtime_error("ray tree did not terminate");
        const uint64_t total = (uint64_t)h[2] | ((uint64_t)h[3] << 32);
        frames.push_back({Batch{pw.out.buf, pw.out.pbuf, nout, s}, slot_offs(s, pw.cap), total, 0});
        return;
      }
      // regroup by scene into the slot's second buffer, which the batch then uses
      const size_t cap = pw.cap;
      Spawn* sb = (Spawn*)arena.get("spawn_sorted" + std::to_string(s), sizeof(Spawn) * cap, st, false);
      PathRec* spb = adj ? (PathRec*)arena.get("path_sorted" + std::to_string(s), sizeof(PathRec) * cap, st, false)
                         : nullptr;
      uint32_t* hist = (uint32_t*)arena.get("scene_hist", sizeof(uint32_t) * 3 * NSC, st, true);
      const uint32_t nb = (nout + SORT_BLOCK * SORT_ITEMS - 1) / (SORT_BLOCK * SORT_ITEMS);
      hipLaunchKernelGGL(k_scene_hist, dim3(nb), dim3(SORT_BLOCK), 0, st, pw.out.buf, nout, hist);
      hipLaunchKernelGGL(k_scene_start, dim3(1), dim3(1024), 0, st, hist, hist + NSC);
      hipLaunchKernelGGL(k_scene_scatter, dim3(nb), dim3(SORT_BLOCK), 0, st, pw.out.buf, pw.out.pbuf, nout,
                         hist + NSC, hist + 2 * NSC, sb, spb);
      check(hipGetLastError(), "scene sort launch");
      open_frame(Batch{sb, spb, nout, s}, true);
    };
    while (!frames.empty() || !inflight.empty()) {
      while (!frames.empty() && frames.back().c0 >= frames.back().total) {
        free_slots.push_back(frames.back().bt.slot);
        frames.pop_back();
      }
      if (!frames.empty() && inflight.size() < depth) {
        launch_window();
        continue;
      }
      if (!inflight.empty()) finish_window();
    }
  }
  if (stats) {
    stats->generations = gens;
    stats->node_bytes = (tn.trace_kernel == 1 && tn.wide) ? (int)sizeof(BVH4Node) : (int)sizeof(BVHNode);
  }
}

}  // namespace rg
