// C ABI: contexts, scene upload, rtrace / rcontrib / dctimestep entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>

#include "../../include/radgpu.h"
#include "rg_bvh.hpp"
#include "rg_render.hpp"
#include "rg_scene.hpp"
#include "rg_sky.hpp"
#include "rg_view.hpp"

using namespace rg;

// Scenes up to this many BVH primitives get the host SAH tree by default;
// larger ones the device LBVH (build time grows as n log n on one core).
constexpr uint32_t kSahMaxPrims = 4u << 20;
// Deepest tree the traversal stack (16 LDS + 48 private entries) can walk.
constexpr int kMaxTreeDepth = 60;

struct rg_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  SceneDB db;
  CompiledScene cs;
  bool committed = false;
  DevArena arena;          // per-call scratch
  DevArena scene_mem;      // scene tables + BVH
  DevScene dscene{};
  Params params{};
  std::string err;
  unsigned long long last_nrays = 0;
  int bvh_builder = RG_BVH_AUTO;
  Tuning tuning;
  int profile = 0;              // 0 off, 1 time k_bounce launches, 2 also count visits
  RenderStats last_stats;
  unsigned long long last_visits[2] = {0, 0};
  // shadow cache: filled at the first call after a commit (the reference
  // fills a source's cache at its first shadow test in the process); the
  // cast rays of a source count once, in the call that first tests it
  bool shadow_filled = false;
  std::vector<unsigned long long> cast_count;   // closest-hit queries per source's cast rays
  std::vector<char> cast_counted;
  hipStream_t caller = nullptr;   // device-memory inputs are read after the work queued here
  hipEvent_t caller_ev = nullptr;
  // pinned host staging: a call's inputs go up and its small results come
  // back through it in one transfer each way
  char* stage = nullptr;
  size_t stage_bytes = 0;
  const double* d_cast_rays = nullptr;   // shadow-cache cast rays (in the scene blob)
  const int32_t* d_cast_src = nullptr;
};

static thread_local std::string g_create_err;

// Wavefront scratch (spawn windows, scans, trace records) is shared by all
// contexts of a device: it is sized by the largest call, not multiplied by
// the number of scenes a process keeps committed.  Calls on one device take
// turns on it.  The arenas live until the process ends (never destroyed, so
// no device memory is released after the runtime has shut down).
static std::mutex g_render_mu;
static std::unordered_map<int, DevArena*> g_render_arena;

static DevArena& render_arena(int device) {
  DevArena*& a = g_render_arena[device];
  if (!a) a = new DevArena();
  return *a;
}

#define RG_TRY(ctx, ...)                      \
  try {                                       \
    __VA_ARGS__;                              \
    return 0;                                 \
  } catch (const std::exception& e) {         \
    if (ctx) (ctx)->err = e.what();           \
    return -1;                                \
  }

static void hipck(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}

template <class T>
static T* upload(DevArena& a, const std::string& name, const std::vector<T>& v, hipStream_t st) {
  T* p = (T*)a.get(name, sizeof(T) * std::max<size_t>(v.size(), 1), st, false);
  if (!v.empty()) hipck(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st), "upload");
  return p;
}

// The context's pinned staging buffer, at least `bytes` long.  Every call
// ends with its stream synchronised, so no transfer uses it when it grows.
static char* stage_buf(rg_ctx* c, size_t bytes) {
  if (c->stage_bytes < bytes) {
    if (c->stage) (void)hipHostFree(c->stage);
    c->stage = nullptr;
    c->stage_bytes = 0;
    hipck(hipHostMalloc((void**)&c->stage, bytes, 0), "pinned staging");
    c->stage_bytes = bytes;
  }
  return c->stage;
}

// Host arrays laid out in one device block (256-byte aligned parts), filled
// through the pinned staging buffer and sent in one copy.  A part with no
// source is zero-filled; one with fewer source bytes than its size has a
// zero tail.
struct Pack {
  struct Part {
    size_t off, bytes, src_bytes;
    const void* src;
  };
  std::vector<Part> parts;
  size_t total = 0;
  size_t add(const void* src, size_t src_bytes, size_t bytes) {
    const size_t off = (total + 255) & ~(size_t)255;
    parts.push_back({off, bytes, src_bytes, src});
    total = off + std::max<size_t>(bytes, 16);
    return off;
  }
  template <class T>
  size_t add(const std::vector<T>& v) { return add(v.data(), sizeof(T) * v.size(), sizeof(T) * v.size()); }
  size_t zero(size_t bytes) { return add(nullptr, 0, bytes); }
  // stage the parts and queue the copy to dev (total bytes)
  void send(rg_ctx* c, void* dev) const {
    char* h = stage_buf(c, total);
    size_t at = 0;
    for (const Part& p : parts) {
      if (p.off > at) memset(h + at, 0, p.off - at);
      if (p.src_bytes) memcpy(h + p.off, p.src, p.src_bytes);
      const size_t end = p.off + std::max<size_t>(p.bytes, 16);
      if (end > p.off + p.src_bytes) memset(h + p.off + p.src_bytes, 0, end - p.off - p.src_bytes);
      at = end;
    }
    if (total) hipck(hipMemcpyAsync(dev, h, total, hipMemcpyHostToDevice, c->st), "upload");
  }
};

extern "C" {

void rg_default_params(rg_params* p) {
  memset(p, 0, sizeof *p);
  p->ambounce = 0;
  p->ambdiv = 1024;
  p->maxdepth = -10;
  p->backvis = 1;
  p->minweight = 1e-4;
  p->dstrsrc = 0.0;
  p->srcsizerat = 0.2;
  p->specthresh = 0.15;
  p->specjitter = 1.0;
  p->directvis = 1;
  p->contrib = 0;
  p->seed = 0;
}

void rg_default_params_rcontrib(rg_params* p) {
  rg_default_params(p);
  p->ambounce = 1;
  p->ambdiv = 350;
  p->minweight = 2e-3;
  p->dstrsrc = 0.9;
  p->specthresh = 0.02;
}

int rg_ctx_create(rg_ctx** out, int device) {
  try {
    int n = 0;
    hipck(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) throw std::runtime_error("no HIP device " + std::to_string(device));
    hipck(hipSetDevice(device), "hipSetDevice");
    rg_ctx* c = new rg_ctx();
    c->device = device;
    hipck(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking), "hipStreamCreate");
    rg_params p;
    rg_default_params(&p);
    rg_set_params(c, &p);
    *out = c;
    return 0;
  } catch (const std::exception& e) {
    g_create_err = e.what();
    *out = nullptr;
    return -1;
  }
}

void rg_ctx_destroy(rg_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  if (c->caller_ev) (void)hipEventDestroy(c->caller_ev);
  if (c->stage) (void)hipHostFree(c->stage);
  c->arena.release();
  c->scene_mem.release();
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

const char* rg_last_error(const rg_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int rg_scene_load(rg_ctx* c, const char* path, int allow_exec) {
  RG_TRY(c, {
    c->db.allow_exec = allow_exec != 0;
    c->db.load_any(path);
    c->committed = false;
  });
}

int rg_scene_load_text(rg_ctx* c, const char* text, int allow_exec) {
  RG_TRY(c, {
    c->db.allow_exec = allow_exec != 0;
    c->db.load_rad_text(text, "scene text");
    c->committed = false;
  });
}

int rg_scene_commit(rg_ctx* c) {
  RG_TRY(c, {
    hipck(hipSetDevice(c->device), "hipSetDevice");
    c->db.resolve_instances();
    c->db.resolve_meshes();
    c->cs = compile_scene(c->db);
    CompiledScene& cs = c->cs;
    DevScene& S = c->dscene;
    S.nprims = (int)cs.prims.size();
    S.nmats = (int)cs.mats.size();
    S.nsrcs = (int)cs.srcs.size();
    S.lamb = cs.lamb;
    S.cusize = cs.cusize;
    S.shadow_on = 0;
    c->shadow_filled = false;
    c->cast_count.assign(cs.srcs.size(), 0);
    c->cast_counted.assign(cs.srcs.size(), 0);
    // Room-sized primitives (box diagonal > 1/8 of the scene's) stay out of
    // the BVH when there are at most max_big of them (a room's walls): inside
    // the tree their boxes would inflate every ancestor, and each ray tests
    // them first, nearest plane first.  More of them than that (the floors
    // and facades of a building) all go into the tree: tested on every ray
    // they would cost more than the walk they shorten.
    std::vector<int32_t> big, small_ids;
    {
      double sd2 = 0;
      for (int a = 0; a < 3; ++a) sd2 += (cs.bmax[a] - cs.bmin[a]) * (cs.bmax[a] - cs.bmin[a]);
      std::vector<std::pair<double, int32_t>> cand;
      for (size_t k = 0; k < cs.prims.size(); ++k) {
        double d2 = 0;
        for (int a = 0; a < 3; ++a) d2 += (double)(cs.hi[3 * k + a] - cs.lo[3 * k + a]) * (cs.hi[3 * k + a] - cs.lo[3 * k + a]);
        if (d2 * 64.0 > sd2) cand.push_back({d2, (int32_t)k});
      }
      std::sort(cand.begin(), cand.end(), [](auto& a, auto& b) { return a.first > b.first; });
      if (cand.size() > c->tuning.max_big) cand.clear();
      if (cand.size() == cs.prims.size() && !cand.empty()) cand.pop_back();   // keep the tree non-empty
      std::vector<char> isbig(cs.prims.size(), 0);
      for (auto& pr : cand) { isbig[pr.second] = 1; big.push_back(pr.second); }
      std::sort(big.begin(), big.end());
      for (size_t k = 0; k < cs.prims.size(); ++k)
        if (!isbig[k]) small_ids.push_back((int32_t)k);
    }
    std::vector<BigPlane> bigpl(kBigPlanes, BigPlane{{0.0, 0.0, 0.0}, 0.0, -1, 0, 0, 0});
    for (size_t k = 0; k < big.size() && k < (size_t)kBigPlanes; ++k) {
      const Prim& P = cs.prims[(size_t)big[k]];
      bigpl[k].prim = big[k];
      bigpl[k].kind = P.kind;
      bigpl[k].goff = P.goff;
      bigpl[k].nv = P.nv;
      if (P.kind != PK_POLYGON) continue;
      const double* g = cs.geom.data() + P.goff;
      for (int a = 0; a < 3; ++a) bigpl[k].n[a] = g[a];
      bigpl[k].off = g[3];
    }
    std::vector<float> slo_v, shi_v;
    slo_v.reserve(3 * small_ids.size());
    shi_v.reserve(3 * small_ids.size());
    float slo[3] = {1e30f, 1e30f, 1e30f}, shi[3] = {-1e30f, -1e30f, -1e30f};
    for (int32_t k : small_ids)
      for (int a = 0; a < 3; ++a) {
        slo_v.push_back(cs.lo[3 * k + a]);
        shi_v.push_back(cs.hi[3 * k + a]);
        slo[a] = std::min(slo[a], cs.lo[3 * k + a]);
        shi[a] = std::max(shi[a], cs.hi[3 * k + a]);
      }
    const uint32_t n = (uint32_t)small_ids.size();
    const bool sah = n && (c->bvh_builder == RG_BVH_SAH || (c->bvh_builder == RG_BVH_AUTO && n <= kSahMaxPrims));
    // host SAH tree (leaf refs hold primitive indices) and its four-wide
    // form for the persistent walk, kept when every path's pushes (at most
    // three per four-wide node) fit the traversal stack
    std::vector<BVHNode> hn;
    std::vector<BVH4Node> h4;
    int root = -1;
    bool wide = false;
    if (sah) {
      root = build_sah_bvh(slo_v.data(), shi_v.data(), small_ids.data(), n, hn, kMaxTreeDepth);
      const int d4 = collapse_bvh4(hn, root, h4);
      wide = root == 0 && !h4.empty() && 3 * d4 <= kTraceStackEntries;   // both walks start at node 0
    }
    // every table in one device block, sent in one copy
    Pack pk;
    const size_t o_prims = pk.add(cs.prims), o_geom = pk.add(cs.geom), o_mats = pk.add(cs.mats),
                 o_srcs = pk.add(cs.srcs), o_scache = pk.add(cs.scache),
                 o_obs = pk.zero(sizeof(unsigned long long) * std::max<int32_t>(cs.ncells, 1)),
                 o_tested = pk.zero(sizeof(uint32_t) * std::max<size_t>(cs.srcs.size(), 1)),
                 o_cast = pk.add(cs.cast_rays), o_cast_src = pk.add(cs.cast_src), o_big = pk.add(big), o_bigpl = pk.add(bigpl),
                 o_nodes = pk.add(hn.data(), sizeof(BVHNode) * hn.size(), sizeof(BVHNode) * std::max<uint32_t>(n, 1)),
                 o_leaf = pk.add(sah ? small_ids.data() : nullptr, sah ? sizeof(int32_t) * n : 0,
                                 sizeof(int32_t) * std::max<uint32_t>(n, 1)),
                 o_nodes4 = wide ? pk.add(h4) : 0, o_blo = sah ? 0 : pk.add(slo_v), o_bhi = sah ? 0 : pk.add(shi_v),
                 o_bids = sah ? 0 : pk.add(small_ids);
    char* blob = (char*)c->scene_mem.get("blob", pk.total, c->st, false);
    pk.send(c, blob);
    S.prims = to_global((const Prim*)(blob + o_prims));
    S.geom = to_global((const double*)(blob + o_geom));
    S.mats = to_global((const Material*)(blob + o_mats));
    S.srcs = to_global((const Source*)(blob + o_srcs));
    S.scache = to_global((const SrcCache*)(blob + o_scache));
    S.shadow_obs = to_global((unsigned long long*)(blob + o_obs));
    S.src_tested = (RG_GLOBAL uint32_t*)(blob + o_tested);
    c->d_cast_rays = (const double*)(blob + o_cast);
    c->d_cast_src = (const int32_t*)(blob + o_cast_src);
    S.big = to_global((const int32_t*)(blob + o_big));
    S.bigpl = to_global((const BigPlane*)(blob + o_bigpl));
    S.nbig = (int)big.size();
    S.nbvh = (int)n;
    S.nodes4 = wide ? to_global((const BVH4Node*)(blob + o_nodes4)) : nullptr;
    BVHNode* nodes = (BVHNode*)(blob + o_nodes);
    int32_t* leaf = (int32_t*)(blob + o_leaf);
    if (n == 0) S.root = -1;
    else if (sah) S.root = root;
    else S.root = build_lbvh((float*)(blob + o_blo), (float*)(blob + o_bhi), (int32_t*)(blob + o_bids), n, slo, shi,
                             nodes, leaf, c->arena, c->st);
    S.nodes = to_global(nodes);
    S.leaf_prim = to_global(leaf);
    hipck(hipStreamSynchronize(c->st), "scene commit");
    c->committed = true;
  });
}

int rg_scene_info(rg_ctx* c, long long* info, int n) {
  RG_TRY(c, {
    long long v[6] = {(long long)c->db.objs.size(), (long long)c->cs.prims.size(), (long long)c->cs.mats.size(),
                      (long long)c->cs.srcs.size(), (long long)(c->cs.prims.size() ? c->cs.prims.size() - 1 : 0),
                      (long long)c->cs.warnings.size()};
    for (int i = 0; i < n && i < 6; ++i) info[i] = v[i];
  });
}

int rg_set_params(rg_ctx* c, const rg_params* p) {
  RG_TRY(c, {
    if (p->ambdiv < 0 || p->ambounce < 0) throw std::runtime_error("negative -ad or -ab");
    if (p->maxdepth <= 0 && p->minweight <= 0.0) throw std::runtime_error("zero ray weight in Russian roulette");
    Params& q = c->params;
    q.ambounce = p->ambounce;
    q.ambdiv = p->ambdiv;
    q.maxdepth = p->maxdepth;
    q.backvis = p->backvis;
    q.minweight = p->minweight;
    q.dstrsrc = p->dstrsrc;
    q.srcsizerat = p->srcsizerat;
    q.specthresh = p->specthresh;
    q.specjitter = p->specjitter;
    for (int i = 0; i < 3; ++i) q.ambval[i] = p->ambval[i];
    q.directvis = p->directvis;
    q.contrib = p->contrib;
    q.seed = p->seed;
  });
}

// Adjoint replay request: loss weights per record, radiance per output
// column, material -> gradient row map, host gradient [rows][ADJ_COLS].
struct AdjSpec {
  const double* dlde;
  const double* lcol;
  std::vector<int32_t> row;
  int nrows;
  double* grad;
};

__global__ void k_scale_out(double* v, size_t n, double s) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= s;
}

// Fixed-point bin sums (int64 multiples of 1 / fix) to doubles, in place.
__global__ void k_fixed_out(double* v, size_t n, double inv_fix) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const long long q = ((const long long*)v)[i];
    v[i] = (double)q * inv_fix;
  }
}

// Scale of the fixed-point bin sums of a coefficient-mode call: a ray's
// coefficients over all its bins add up to at most 16 (a Lambertian
// irradiance record reaches pi, and every material below it multiplies by
// at most 1), so a record of `accum` rays stays below 16 accum, and the
// scale 2^(58 - ceil(log2 accum)) keeps that below 2^62 while resolving
// 2^-58 of a single-ray record.
static double fixed_scale(int accum) {
  int e = 0;
  while ((1LL << e) < accum) ++e;
  return ldexp(1.0, 58 - e);
}

// The cast rays of every source of c flagged in `tested` for the first
// time since the commit, as the reference's counter counts them
// (the cache is filled by rt/srcobstr.c:63-183 at a source's first test).
static unsigned long long shadow_count(rg_ctx* c, const uint32_t* tested) {
  unsigned long long add = 0;
  for (size_t k = 0; k < c->cs.srcs.size(); ++k)
    if (tested[k] && !c->cast_counted[k]) {
      c->cast_counted[k] = 1;
      add += c->cast_count[k];
    }
  return add;
}

// device_io: `rays` and `out` are device memory of the context's device
// batch_ctx: the contexts of the batch's scenes (their shadow caches).
static void run(rg_ctx* c, const double* rays, long long nrays, int accum, int flags, int mode,
                const std::vector<BinSpec>& bins, int ncols, double* out,
                std::vector<DevScene>* batch = nullptr, const std::vector<uint32_t>* batch_start = nullptr,
                const AdjSpec* adj = nullptr, bool device_io = false,
                const std::vector<rg_ctx*>* batch_ctx = nullptr) {
  if (!c->committed) throw std::runtime_error("scene not committed");
  if (nrays < 0 || nrays > 0x7fffffffLL) throw std::runtime_error("bad ray count");
  if (accum < 1) throw std::runtime_error("bad accumulation count");
  hipck(hipSetDevice(c->device), "hipSetDevice");
  c->last_nrays = 0;
  long long nrec = nrays / accum;
  size_t outn = (size_t)nrec * ncols;
  double* d_out = (double*)c->arena.get("out", sizeof(double) * std::max<size_t>(outn, 1), c->st, true);
  // the scenes of this call (a batch's, or the context's own as a copy
  // that carries the call's shadow-cache fields)
  DevScene Sd = c->dscene;
  std::vector<rg_ctx*> sc_ctx;
  std::vector<DevScene*> sc_dev;
  if (batch) {
    for (size_t i = 0; i < batch->size(); ++i) {
      sc_ctx.push_back((*batch_ctx)[i]);
      sc_dev.push_back(&(*batch)[i]);
    }
  } else {
    sc_ctx.push_back(c);
    sc_dev.push_back(&Sd);
  }
  // The call block, cleared once and read back once: rays traced (u64),
  // error flags (u32), then per scene with a shadow cache the closest-hit
  // queries of its cast rays per source (u64, when this call fills the
  // cache: the first call after a commit, as the reference fills a source's
  // cache at its first shadow test) and its tested-source flags (u32).
  const bool shadow = c->tuning.shadow_cache != 0;
  const size_t nsc = sc_ctx.size();
  std::vector<size_t> cast_at(nsc, 0), tested_at(nsc, 0);
  std::vector<char> fill(nsc, 0);
  size_t blk = 16;
  for (size_t i = 0; i < nsc; ++i) {
    rg_ctx* bc = sc_ctx[i];
    DevScene& S = *sc_dev[i];
    S.shadow_on = 0;
    if (!shadow || bc->cs.scache.empty()) continue;
    S.shadow_on = 1;
    const size_t ns = bc->cs.srcs.size();
    fill[i] = !bc->shadow_filled;
    cast_at[i] = blk;
    blk += 8 * ns;
    tested_at[i] = blk;
    blk += (4 * ns + 7) & ~(size_t)7;
  }
  char* d_blk = (char*)c->arena.get("call", blk, c->st, true);
  for (size_t i = 0; i < nsc; ++i)
    if (sc_dev[i]->shadow_on) sc_dev[i]->src_tested = (RG_GLOBAL uint32_t*)(d_blk + tested_at[i]);
  // inputs in one copy: bin specs, a batch's scene tables, the rays
  Pack in;
  const size_t o_bins = bins.empty() ? 0 : in.add(bins);
  const size_t o_sc = batch ? in.add(*batch) : 0, o_ss = batch ? in.add(*batch_start) : 0;
  const bool host_rays = nrays && rays && !device_io;
  const size_t o_rays = host_rays ? in.add(rays, sizeof(double) * 6 * nrays, sizeof(double) * 6 * nrays) : 0;
  // results come back through the staging buffer when small enough to
  // stage (large records go straight to the caller's memory)
  const bool stage_out = outn && out && !device_io && outn * sizeof(double) <= ((size_t)64 << 20);
  const size_t o_hout = (blk + 255) & ~(size_t)255;
  stage_buf(c, std::max(in.total, o_hout + (stage_out ? sizeof(double) * outn : 0)));
  char* d_in = nullptr;
  if (in.total) {
    d_in = (char*)c->arena.get("call_in", in.total, c->st, false);
    in.send(c, d_in);
  }
  double* d_rays = host_rays ? (double*)(d_in + o_rays)
                             : (double*)c->arena.get("rays", sizeof(double) * 6 * std::max<long long>(nrays, 1), c->st,
                                                      false);
  if (nrays && rays && device_io)
    hipck(hipMemcpyAsync(d_rays, rays, sizeof(double) * 6 * nrays, hipMemcpyDeviceToDevice, c->st), "rays");
  RunCfg cfg;
  cfg.p = c->params;
  cfg.mode = mode;
  cfg.irrad = (flags & 16) ? 1 : 0;
  cfg.ncols = ncols;
  cfg.nbinspec = (int)bins.size();
  cfg.bins = bins.empty() ? nullptr : (const BinSpec*)(d_in + o_bins);
  cfg.out = d_out;
  // rcontrib coefficients (not -V values, nor rtrace radiances, whose range
  // the host cannot bound) accumulate in fixed point
  cfg.fix = (mode == 1 && !c->params.contrib && !adj) ? fixed_scale(accum) : 0.0;
  cfg.nrays = (unsigned long long*)d_blk;
  cfg.err = (uint32_t*)(d_blk + 8);
  cfg.accum = accum;
  cfg.child_order = c->tuning.child_order;
  cfg.nscenes = batch ? (int)batch->size() : 0;
  cfg.scenes = batch ? (const DevScene*)(d_in + o_sc) : nullptr;
  cfg.scene_start = batch ? (const uint32_t*)(d_in + o_ss) : nullptr;
  // shadow caches filled by this call: their cast rays go first on the stream
  for (size_t i = 0; i < nsc; ++i) {
    if (!fill[i]) continue;
    rg_ctx* bc = sc_ctx[i];
    shadow_cast(*sc_dev[i], bc->d_cast_rays, bc->d_cast_src, (uint32_t)bc->cs.cast_src.size(),
                (unsigned long long*)sc_dev[i]->shadow_obs, (unsigned long long*)(d_blk + cast_at[i]), cfg.err, c->st);
  }
  cfg.adj = 0;
  cfg.adj_w = nullptr;
  cfg.adj_L = nullptr;
  cfg.grad_row = nullptr;
  cfg.grad = nullptr;
  double* d_grad = nullptr;
  if (adj) {
    std::vector<double> w(adj->dlde, adj->dlde + 3 * nrec);
    for (double& v : w) v /= accum;        // the record is the mean of its rays
    double* d_w = (double*)c->arena.get("adj_w", sizeof(double) * std::max<size_t>(w.size(), 1), c->st, false);
    double* d_L = (double*)c->arena.get("adj_L", sizeof(double) * std::max(ncols, 1), c->st, false);
    int32_t* d_row = (int32_t*)c->arena.get("adj_row", sizeof(int32_t) * adj->row.size(), c->st, false);
    d_grad = (double*)c->arena.get("adj_grad", sizeof(double) * ADJ_COLS * std::max(adj->nrows, 1), c->st, true);
    if (!w.empty()) hipck(hipMemcpyAsync(d_w, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, c->st), "adj");
    hipck(hipMemcpyAsync(d_L, adj->lcol, sizeof(double) * ncols, hipMemcpyHostToDevice, c->st), "adj");
    hipck(hipMemcpyAsync(d_row, adj->row.data(), sizeof(int32_t) * adj->row.size(), hipMemcpyHostToDevice, c->st),
          "adj");
    hipck(hipStreamSynchronize(c->st), "adjoint upload");   // `w` is freed on return
    cfg.adj = 1;
    cfg.adj_w = d_w;
    cfg.adj_L = d_L;
    cfg.grad_row = d_row;
    cfg.grad = d_grad;
  }
  cfg.visits = nullptr;
  unsigned long long* d_vis = nullptr;
  if (c->profile >= 2) {
    d_vis = (unsigned long long*)c->arena.get("visits", 16, c->st, true);
    cfg.visits = d_vis;
  }
  c->last_stats = RenderStats();
  c->last_stats.timing = c->profile >= 1;
  int max_srcs = c->dscene.nsrcs;
  if (batch)
    for (const DevScene& b : *batch) max_srcs = std::max(max_srcs, b.nsrcs);
  // the four-wide walk needs the four-wide tree of every scene traced
  Tuning tn = c->tuning;
  if (!c->dscene.nodes4) tn.wide = 0;
  if (batch)
    for (const DevScene& b : *batch)
      if (!b.nodes4) tn.wide = 0;
  if (nrays) {
    std::lock_guard<std::mutex> lock(g_render_mu);
    render(Sd, cfg, d_rays, (uint32_t)nrays, (flags & 8) != 0, max_srcs, tn, render_arena(c->device), c->st,
           &c->last_stats);
  }
  if (d_vis) hipck(hipMemcpyAsync(c->last_visits, d_vis, 16, hipMemcpyDeviceToHost, c->st), "visits");
  else c->last_visits[0] = c->last_visits[1] = 0;
  if (cfg.fix > 0.0 && outn)
    k_fixed_out<<<(unsigned)((outn + 255) / 256), 256, 0, c->st>>>(d_out, outn, 1.0 / cfg.fix);
  // one read-back of the call block (and of staged records)
  char* h = c->stage;
  hipck(hipMemcpyAsync(h, d_blk, blk, hipMemcpyDeviceToHost, c->st), "call block");
  if (outn && out && device_io) {
    if (accum > 1)
      k_scale_out<<<(unsigned)((outn + 255) / 256), 256, 0, c->st>>>(d_out, outn, 1.0 / accum);
    hipck(hipMemcpyAsync(out, d_out, sizeof(double) * outn, hipMemcpyDeviceToDevice, c->st), "out");
  } else if (stage_out) {
    hipck(hipMemcpyAsync(h + o_hout, d_out, sizeof(double) * outn, hipMemcpyDeviceToHost, c->st), "out");
  } else if (outn && out) {
    hipck(hipMemcpyAsync(out, d_out, sizeof(double) * outn, hipMemcpyDeviceToHost, c->st), "out");
  }
  if (adj)
    hipck(hipMemcpyAsync(adj->grad, d_grad, sizeof(double) * ADJ_COLS * adj->nrows, hipMemcpyDeviceToHost, c->st),
          "grad");
  hipck(hipStreamSynchronize(c->st), "render");
  uint32_t herr = 0;
  memcpy(&c->last_nrays, h, 8);
  memcpy(&herr, h + 8, 4);
  for (size_t i = 0; i < nsc; ++i) {
    rg_ctx* bc = sc_ctx[i];
    if (fill[i]) {
      memcpy(bc->cast_count.data(), h + cast_at[i], sizeof(unsigned long long) * bc->cs.srcs.size());
      bc->shadow_filled = !(herr & ERR_STACK);
    }
    if (sc_dev[i]->shadow_on) c->last_nrays += shadow_count(bc, (const uint32_t*)(h + tested_at[i]));
  }
  if (stage_out) memcpy(out, h + o_hout, sizeof(double) * outn);
  if (herr & ERR_STACK) throw std::runtime_error("BVH traversal stack overflow");
  if (herr & ERR_CAPACITY) throw std::runtime_error("spawn buffer overflow");
  if (herr & ERR_NOMAT) throw std::runtime_error("surface without material");
  if (herr & ERR_FIXED)
    throw std::runtime_error("a contribution coefficient beyond the fixed-point range (coefficients above 16 per ray)");
  if (herr & ERR_PATH)
    throw std::runtime_error("adjoint: the differentiated vertices of one path need more than " +
                             std::to_string(PATH_WORDS) + " record words");
  if (!out || device_io) return;
  if (accum > 1) {
    double s = 1.0 / accum;
    for (size_t i = 0; i < outn; ++i) out[i] *= s;
  }
}

int rg_trace(rg_ctx* c, const double* rays, long long n, int flags, double* out) {
  RG_TRY(c, {
    std::vector<BinSpec> none;
    run(c, rays, n, 1, flags, 0, none, 3, out);
  });
}

int rg_trace_view(rg_ctx* c, const rg_view* v, int xres, int yres, double pj, int flags, double* out) {
  RG_TRY(c, {
    if (!v || xres <= 0 || yres <= 0 || (long long)xres * yres > 0x7fffffffLL) throw std::runtime_error("bad resolution");
    if (v->vaft > 1e-6) throw std::runtime_error("aft clipping plane (-va) is not supported on the GPU path");
    if (!strchr("vlcahs", v->type) || !v->type) throw std::runtime_error("unknown view type");
    if (!c->committed) throw std::runtime_error("scene not committed");
    const long long n = (long long)xres * yres;
    hipck(hipSetDevice(c->device), "hipSetDevice");
    double* d_rays = (double*)c->arena.get("rays", sizeof(double) * 6 * n, c->st, false);
    uint8_t* d_valid = (uint8_t*)c->arena.get("view_valid", (size_t)n, c->st, false);
    ViewSpec vs;
    memset(&vs, 0, sizeof vs);
    vs.type = v->type;
    for (int k = 0; k < 3; ++k) {
      vs.vp[k] = v->vp[k]; vs.vdir[k] = v->vdir[k]; vs.hvec[k] = v->hvec[k]; vs.vvec[k] = v->vvec[k];
    }
    vs.hn2 = v->hn2; vs.vn2 = v->vn2; vs.horiz = v->horiz; vs.vert = v->vert;
    vs.hoff = v->hoff; vs.voff = v->voff; vs.vfore = v->vfore; vs.vaft = v->vaft;
    view_rays(vs, xres, yres, pj, c->params.seed, d_rays, d_valid, c->st);
    hipck(hipGetLastError(), "view rays");
    std::vector<BinSpec> none;
    run(c, nullptr, n, 1, flags, 0, none, 3, out);
    std::vector<uint8_t> valid((size_t)n);
    hipck(hipMemcpy(valid.data(), d_valid, (size_t)n, hipMemcpyDeviceToHost), "view mask");
    for (long long i = 0; i < n; ++i)
      if (!valid[(size_t)i]) out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = 0.0;
  });
}

// Bin specs for the modifier list; fills the name -> slot map.
static std::vector<BinSpec> bin_specs(const rg_modifier* mods, int nmods, std::unordered_map<std::string, int>& slot,
                                      int& ncols) {
  std::vector<BinSpec> bins(nmods);
  int col = 0;
  for (int i = 0; i < nmods; ++i) {
    const rg_modifier& m = mods[i];
    if (!m.name || !m.name[0]) throw std::runtime_error("modifier without a name");
    if (!strcmp(m.name, "void")) throw std::runtime_error("cannot track 'void' modifier");
    if (slot.count(m.name)) throw std::runtime_error(std::string("duplicate modifier '") + m.name + "'");
    if (m.nbins < 1) throw std::runtime_error("bad bin count");
    if (m.bin_kind < 0 || m.bin_kind > 7) throw std::runtime_error("bad bin kind");
    if ((m.bin_kind == BK_REINHART || m.bin_kind == BK_SHIRLEY_CHIU) && !(m.mf > 0))
      throw std::runtime_error(std::string("bad MF / SCdim for modifier '") + m.name + "'");
    slot[m.name] = i;
    BinSpec& b = bins[i];
    memset(&b, 0, sizeof b);
    b.kind = m.bin_kind;
    b.nbins = m.nbins;
    b.col0 = col;
    b.cbin = m.cbin;
    b.mf = m.mf;
    b.rhs = m.rhs;
    for (int a = 0; a < 3; ++a) { b.nrm[a] = m.nrm[a]; b.up[a] = m.up[a]; }
    col += m.nbins;
  }
  ncols = 3 * col;
  return bins;
}

// The context's scene with tracked-modifier slots assigned by modifier name
// (rt/rcontrib.c:289) and every source followed (rt/rcontrib.c:235-237).
static DevScene contrib_scene(rg_ctx* c, const std::unordered_map<std::string, int>& slot, hipStream_t st) {
  if (!c->committed) throw std::runtime_error("scene not committed");
  CompiledScene& cs = c->cs;
  std::vector<Prim> prims = cs.prims;
  for (size_t k = 0; k < prims.size(); ++k) {
    auto it = slot.find(cs.prim_modname[k]);
    prims[k].modslot = it == slot.end() ? -1 : it->second;
  }
  std::vector<Source> srcs = cs.srcs;
  for (size_t k = 0; k < srcs.size(); ++k) {
    auto it = slot.find(cs.src_modname[k]);
    srcs[k].modslot = it == slot.end() ? -1 : it->second;
    srcs[k].flags |= SFOLLOW;
  }
  if (c->params.contrib) {
    for (size_t k = 0; k < prims.size(); ++k) {
      int mk = prims[k].mat >= 0 ? cs.mats[prims[k].mat].kind : -1;
      if (prims[k].modslot >= 0 && mk != MK_LIGHT && mk != MK_GLOW && mk != MK_ILLUM)
        throw std::runtime_error("-V+ is supported for light-source modifiers only on the GPU path");
    }
  }
  DevScene S = c->dscene;
  S.prims = to_global(upload(c->arena, "prims_c", prims, st));
  S.srcs = to_global(upload(c->arena, "srcs_c", srcs, st));
  return S;
}

int rg_contrib(rg_ctx* c, const double* rays, long long nrec, int accum, int flags, const rg_modifier* mods,
               int nmods, double* out) {
  RG_TRY(c, {
    std::unordered_map<std::string, int> slot;
    int ncols = 0;
    std::vector<BinSpec> bins = bin_specs(mods, nmods, slot, ncols);
    DevScene S = contrib_scene(c, slot, c->st);
    DevScene saved = c->dscene;
    c->dscene = S;
    try {
      run(c, rays, nrec * (long long)accum, accum, flags, 1, bins, ncols, out);
    } catch (...) {
      c->dscene = saved;
      throw;
    }
    c->dscene = saved;
  });
}

// The context's stream waits for the work queued so far on the caller's
// stream, so device inputs written there are complete before they are read.
static void order_after_caller(rg_ctx* c) {
  hipck(hipSetDevice(c->device), "hipSetDevice");
  if (!c->caller_ev) hipck(hipEventCreateWithFlags(&c->caller_ev, hipEventDisableTiming), "event");
  hipck(hipEventRecord(c->caller_ev, c->caller), "caller event");
  hipck(hipStreamWaitEvent(c->st, c->caller_ev, 0), "caller wait");
}

int rg_set_stream(rg_ctx* c, void* stream) {
  RG_TRY(c, { c->caller = (hipStream_t)stream; });
}

int rg_contrib_device(rg_ctx* c, const double* d_rays, long long nrec, int accum, int flags, const rg_modifier* mods,
                      int nmods, double* d_out) {
  RG_TRY(c, {
    if (nrec > 0 && (!d_rays || !d_out)) throw std::runtime_error("null device buffer");
    order_after_caller(c);
    std::unordered_map<std::string, int> slot;
    int ncols = 0;
    std::vector<BinSpec> bins = bin_specs(mods, nmods, slot, ncols);
    DevScene S = contrib_scene(c, slot, c->st);
    DevScene saved = c->dscene;
    c->dscene = S;
    try {
      run(c, d_rays, nrec * (long long)accum, accum, flags, 1, bins, ncols, d_out, nullptr, nullptr, nullptr, true);
    } catch (...) {
      c->dscene = saved;
      throw;
    }
    c->dscene = saved;
  });
}

int rg_contrib_adjoint(rg_ctx* c, const double* rays, long long nrec, int accum, int flags, const rg_modifier* mods,
                       int nmods, const double* dlde, const double* lcol, const int* mats, int nmats, double* grad) {
  RG_TRY(c, {
    if (!c->committed) throw std::runtime_error("scene not committed");
    if (nmats < 0 || (nmats && (!mats || !grad))) throw std::runtime_error("bad material list");
    if (!dlde || !lcol) throw std::runtime_error("missing loss weights or column radiances");
    std::unordered_map<std::string, int> slot;
    int ncols = 0;
    std::vector<BinSpec> bins = bin_specs(mods, nmods, slot, ncols);
    AdjSpec adj;
    adj.dlde = dlde;
    adj.lcol = lcol;
    adj.row.assign(c->cs.mats.size(), -1);
    adj.nrows = nmats;
    adj.grad = grad;
    for (int i = 0; i < nmats; ++i) {
      int mi = mats[i];
      if (mi < 0 || mi >= (int)c->cs.mats.size() || mi == c->cs.lamb) throw std::runtime_error("bad material index");
      const Material& M = c->cs.mats[mi];
      if (M.kind != MK_PLASTIC && M.kind != MK_METAL && M.kind != MK_TRANS)
        throw std::runtime_error("adjoint: only plastic, metal and trans materials are differentiated");
      // the lobe score assumes untruncated samples (-ss >= 1 or < 0)
      if (M.f[4] * M.f[4] > FTINY && c->params.specjitter >= 0.0 && c->params.specjitter < 1.0)
        throw std::runtime_error("adjoint: a rough material's roughness derivative needs -ss 1 (untruncated lobe samples)");
      if (mi > 0xfff) throw std::runtime_error("adjoint: material index beyond 4095");
      if (adj.row[mi] >= 0) throw std::runtime_error("duplicate material in the gradient list");
      adj.row[mi] = i;
    }
    DevScene S = contrib_scene(c, slot, c->st);
    DevScene saved = c->dscene;
    c->dscene = S;
    try {
      run(c, rays, nrec * (long long)accum, accum, flags, 1, bins, ncols, nullptr, nullptr, nullptr, &adj);
    } catch (...) {
      c->dscene = saved;
      throw;
    }
    c->dscene = saved;
  });
}

int rg_material_index(rg_ctx* c, const char* name, int* index) {
  RG_TRY(c, {
    if (!c->committed) throw std::runtime_error("scene not committed");
    if (!name || !index) throw std::runtime_error("null argument");
    *index = -1;
    for (size_t i = 0; i < c->cs.mat_names.size(); ++i)
      if ((int)i != c->cs.lamb && c->cs.mat_names[i] == name) { *index = (int)i; break; }
    if (*index < 0) throw std::runtime_error(std::string("no material '") + name + "'");
  });
}

int rg_material_get(rg_ctx* c, int index, int* kind, double* f, int n) {
  RG_TRY(c, {
    if (!c->committed) throw std::runtime_error("scene not committed");
    if (index < 0 || index >= (int)c->cs.mats.size()) throw std::runtime_error("bad material index");
    const Material& m = c->cs.mats[index];
    if (kind) *kind = m.kind;
    for (int i = 0; i < n && i < m.nfargs; ++i) f[i] = m.f[i];
  });
}

int rg_material_set(rg_ctx* c, int index, const double* f, int n) {
  RG_TRY(c, {
    if (!c->committed) throw std::runtime_error("scene not committed");
    if (index < 0 || index >= (int)c->cs.mats.size() || index == c->cs.lamb)
      throw std::runtime_error("bad material index");
    Material& m = c->cs.mats[index];
    if (n != m.nfargs) throw std::runtime_error("material argument count differs");
    for (int i = 0; i < n; ++i) m.f[i] = f[i];
    hipck(hipSetDevice(c->device), "hipSetDevice");
    hipck(hipMemcpyAsync((void*)c->dscene.mats, c->cs.mats.data(), sizeof(Material) * c->cs.mats.size(),
                         hipMemcpyHostToDevice, c->st), "material upload");
    hipck(hipStreamSynchronize(c->st), "material upload");
  });
}

int rg_contrib_batch(rg_ctx* const* ctxs, int nctx, const double* rays, const long long* nrec, int accum, int flags,
                     const rg_modifier* mods, int nmods, double* out) {
  rg_ctx* c0 = nctx > 0 ? ctxs[0] : nullptr;
  if (!c0) {
    g_create_err = "empty batch";
    return -1;
  }
  RG_TRY(c0, {
    if (nctx > kMaxBatchScenes)
      throw std::runtime_error("at most " + std::to_string(kMaxBatchScenes) + " scenes per batch");
    std::unordered_map<std::string, int> slot;
    int ncols = 0;
    std::vector<BinSpec> bins = bin_specs(mods, nmods, slot, ncols);
    std::vector<DevScene> scenes;
    std::vector<uint32_t> start(1, 0);
    long long total = 0;
    for (int i = 0; i < nctx; ++i) {
      if (!ctxs[i] || ctxs[i]->device != c0->device) throw std::runtime_error("batch contexts must share a device");
      if (nrec[i] < 0) throw std::runtime_error("bad record count");
      scenes.push_back(contrib_scene(ctxs[i], slot, ctxs[i]->st));
      hipck(hipStreamSynchronize(ctxs[i]->st), "batch upload");
      total += nrec[i] * (long long)accum;
      if (total > 0x7fffffffLL) throw std::runtime_error("bad ray count");
      start.push_back((uint32_t)total);
    }
    std::vector<rg_ctx*> cl(ctxs, ctxs + nctx);
    run(c0, rays, total, accum, flags, 1, bins, ncols, out, &scenes, &start, nullptr, false, &cl);
  });
}

int rg_trace_batch(rg_ctx* const* ctxs, int nctx, const double* rays, const long long* n, int flags, double* out) {
  rg_ctx* c0 = nctx > 0 ? ctxs[0] : nullptr;
  if (!c0) {
    g_create_err = "empty batch";
    return -1;
  }
  RG_TRY(c0, {
    if (nctx > kMaxBatchScenes)
      throw std::runtime_error("at most " + std::to_string(kMaxBatchScenes) + " scenes per batch");
    std::vector<DevScene> scenes;
    std::vector<uint32_t> start(1, 0);
    long long total = 0;
    for (int i = 0; i < nctx; ++i) {
      if (!ctxs[i] || ctxs[i]->device != c0->device) throw std::runtime_error("batch contexts must share a device");
      if (!ctxs[i]->committed) throw std::runtime_error("scene not committed");
      if (n[i] < 0) throw std::runtime_error("bad ray count");
      scenes.push_back(ctxs[i]->dscene);
      total += n[i];
      if (total > 0x7fffffffLL) throw std::runtime_error("bad ray count");
      start.push_back((uint32_t)total);
    }
    std::vector<BinSpec> none;
    std::vector<rg_ctx*> cl(ctxs, ctxs + nctx);
    run(c0, rays, total, 1, flags, 0, none, 3, out, &scenes, &start, nullptr, false, &cl);
  });
}

int rg_first_hit(rg_ctx* c, const double* rays, long long n, double* t, double* nrm, int* prim) {
  RG_TRY(c, {
    if (!c->committed) throw std::runtime_error("scene not committed");
    if (n < 0 || n > 0x7fffffffLL) throw std::runtime_error("bad ray count");
    if (n == 0) return 0;
    hipck(hipSetDevice(c->device), "hipSetDevice");
    double* d_rays = (double*)c->arena.get("fh_rays", sizeof(double) * 6 * n, c->st, false);
    double* d_t = (double*)c->arena.get("fh_t", sizeof(double) * n, c->st, false);
    double* d_n = (double*)c->arena.get("fh_n", sizeof(double) * 3 * n, c->st, false);
    int32_t* d_p = (int32_t*)c->arena.get("fh_p", sizeof(int32_t) * n, c->st, false);
    uint32_t* d_err = (uint32_t*)c->arena.get("err", 8, c->st, true);
    hipck(hipMemcpyAsync(d_rays, rays, sizeof(double) * 6 * n, hipMemcpyHostToDevice, c->st), "rays");
    first_hit(c->dscene, d_rays, (uint32_t)n, d_t, d_n, d_p, d_err, c->st);
    uint32_t herr = 0;
    hipck(hipMemcpyAsync(&herr, d_err, 4, hipMemcpyDeviceToHost, c->st), "err");
    if (t) hipck(hipMemcpyAsync(t, d_t, sizeof(double) * n, hipMemcpyDeviceToHost, c->st), "t");
    if (nrm) hipck(hipMemcpyAsync(nrm, d_n, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->st), "normals");
    if (prim) hipck(hipMemcpyAsync(prim, d_p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->st), "prims");
    hipck(hipStreamSynchronize(c->st), "first hit");
    if (herr & ERR_STACK) throw std::runtime_error("BVH traversal stack overflow");
  });
}

int rg_last_nrays(rg_ctx* c, unsigned long long* n) {
  RG_TRY(c, { *n = c->last_nrays; });
}

int rg_set_tuning(rg_ctx* c, const char* key, long long value) {
  RG_TRY(c, {
    if (!key) throw std::runtime_error("null tuning key");
    std::string k = key;
    Tuning& t = c->tuning;
    if (k == "trace_kernel") {
      if (value != 0 && value != 1) throw std::runtime_error("trace_kernel is 0 or 1");
      t.trace_kernel = (int)value;
    } else if (k == "refill") {
      if (value < 1 || value > 64) throw std::runtime_error("refill is 1..64");
      t.refill = (uint32_t)value;
    } else if (k == "window") {
      if (value < 1024 || value > (1LL << 24)) throw std::runtime_error("window is 1024..16M");
      t.window = (uint32_t)value;
    } else if (k == "max_big") {
      if (value < 0 || value > 32) throw std::runtime_error("max_big is 0..32");
      t.max_big = (uint32_t)value;
    } else if (k == "wide") {
      if (value != 0 && value != 1) throw std::runtime_error("wide is 0 or 1");
      t.wide = (int)value;
    } else if (k == "child_order") {
      if (value != 0 && value != 1) throw std::runtime_error("child_order is 0 or 1");
      t.child_order = (int)value;
    } else if (k == "inflight") {
      if (value != 1 && value != 2) throw std::runtime_error("inflight is 1 or 2");
      t.inflight = (int)value;
    } else if (k == "shadow_cache") {
      if (value != 0 && value != 1) throw std::runtime_error("shadow_cache is 0 or 1");
      t.shadow_cache = (int)value;
    } else if (k == "blocks_per_cu") {
      if (value < 1 || value > 64) throw std::runtime_error("blocks_per_cu is 1..64");
      t.blocks_per_cu = (uint32_t)value;
    } else {
      throw std::runtime_error("unknown tuning key '" + k + "'");
    }
  });
}

int rg_set_bvh_builder(rg_ctx* c, int kind) {
  RG_TRY(c, {
    if (kind != RG_BVH_AUTO && kind != RG_BVH_SAH && kind != RG_BVH_LBVH)
      throw std::runtime_error("unknown BVH builder");
    c->bvh_builder = kind;
  });
}

int rg_set_profiling(rg_ctx* c, int level) {
  RG_TRY(c, {
    if (level < 0 || level > 2) throw std::runtime_error("bad profiling level");
    c->profile = level;
  });
}

int rg_last_stats(rg_ctx* c, double* out, int n) {
  RG_TRY(c, {
    double v[9] = {(double)c->last_nrays, (double)c->last_stats.launches, (double)c->last_stats.threads,
                   c->last_stats.bounce_ms, (double)c->last_visits[0], (double)c->last_visits[1],
                   (double)c->last_stats.generations, c->last_stats.shade_ms, (double)c->last_stats.node_bytes};
    for (int i = 0; i < n && i < 9; ++i) out[i] = v[i];
  });
}

int rg_dcsky_device(const float* A, const float* B, float* C, int M, int K, int N, int mode) {
  try {
    if (M <= 0 || K <= 0 || N <= 0) throw std::runtime_error("bad matrix dimensions");
    if (mode < 0 || mode > 4) throw std::runtime_error("bad mode");
    dcsky(A, B, C, M, K, N, mode, nullptr);
    hipck(hipGetLastError(), "dcsky launch");
    hipck(hipDeviceSynchronize(), "dcsky");
    return 0;
  } catch (const std::exception& e) {
    g_create_err = e.what();
    return -1;
  }
}

int rg_dcsky(const float* A, const float* B, float* C, int M, int K, int N, int mode) {
  float *dA = nullptr, *dB = nullptr, *dC = nullptr;
  try {
    if (M <= 0 || K <= 0 || N <= 0) throw std::runtime_error("bad matrix dimensions");
    size_t na = (size_t)M * K * 3, nb = (size_t)K * N * 3, nc = (size_t)M * N * 3;
    hipck(hipMalloc(&dA, na * 4), "hipMalloc");
    hipck(hipMalloc(&dB, nb * 4), "hipMalloc");
    hipck(hipMalloc(&dC, nc * 4), "hipMalloc");
    hipck(hipMemcpy(dA, A, na * 4, hipMemcpyHostToDevice), "copy A");
    hipck(hipMemcpy(dB, B, nb * 4, hipMemcpyHostToDevice), "copy B");
    if (rg_dcsky_device(dA, dB, dC, M, K, N, mode) != 0) throw std::runtime_error(g_create_err);
    hipck(hipMemcpy(C, dC, nc * 4, hipMemcpyDeviceToHost), "copy C");
    (void)hipFree(dA); (void)hipFree(dB); (void)hipFree(dC);
    return 0;
  } catch (const std::exception& e) {
    g_create_err = e.what();
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    if (dC) (void)hipFree(dC);
    return -1;
  }
}

// ---- gendaymtx ---------------------------------------------------------------

struct rg_sky {
  rg::SkyPlan plan;
  std::string err;
};

int rg_sky_create(rg_sky** out, const char* weather, const rg_skyopts* o) {
  *out = nullptr;
  try {
    if (!weather || !o) throw std::runtime_error("missing weather data or options");
    rg::SkyOpts so;
    so.mfactor = o->mfactor;
    for (int c = 0; c < 3; ++c) {
      so.sky_color[c] = o->sky_color[c];
      so.ground_color[c] = o->ground_color[c];
      so.sun_color[c] = o->sun_color[c];
    }
    so.rotation = o->rotation;
    so.nsuns = o->nsuns;
    so.fixed_sun_sa = o->fixed_sun_sa;
    so.output = o->output;
    so.sun_hours_only = o->sun_hours_only;
    so.average = o->average;
    rg_sky* s = new rg_sky;
    try {
      s->plan = rg::sky_plan(weather, so);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
    return 0;
  } catch (const std::exception& e) {
    g_create_err = e.what();
    return -1;
  }
}

void rg_sky_destroy(rg_sky* s) { delete s; }

const char* rg_sky_error(const rg_sky* s) { return s ? s->err.c_str() : g_create_err.c_str(); }

int rg_sky_shape(const rg_sky* s, int* out) {
  if (!s) return -1;
  out[0] = s->plan.npatch;
  out[1] = s->plan.ncols;
  out[2] = s->plan.nsteps;
  out[3] = s->plan.ntsteps;
  out[4] = (int)s->plan.suns.size();
  return 0;
}

int rg_sky_site(const rg_sky* s, double* latlong) {
  if (!s) return -1;
  latlong[0] = s->plan.latitude;
  latlong[1] = s->plan.longitude;
  return 0;
}

int rg_sky_suns(const rg_sky* s, int* id_good, double* alt_azi_illum) {
  if (!s) return -1;
  for (size_t i = 0; i < s->plan.suns.size(); ++i) {
    const rg::SkySun& u = s->plan.suns[i];
    id_good[2 * i] = u.id;
    id_good[2 * i + 1] = u.good;
    alt_azi_illum[3 * i] = u.altitude;
    alt_azi_illum[3 * i + 1] = u.azimuth;
    alt_azi_illum[3 * i + 2] = u.dir_illum;
  }
  return 0;
}

int rg_sky_matrix_device(rg_sky* s, float* dout) {
  try {
    if (!s) throw std::runtime_error("no sky");
    rg::sky_matrix(s->plan, dout, nullptr);
    return 0;
  } catch (const std::exception& e) {
    s->err = e.what();
    return -1;
  }
}

int rg_sky_matrix_device_noalloc(rg_sky* s, float* dout) {
  try {
    if (!s) throw std::runtime_error("no sky");
    if (!dout) throw std::runtime_error("null device buffer");
    rg::sky_matrix_noalloc(s->plan, dout, nullptr);
    return 0;
  } catch (const std::exception& e) {
    s->err = e.what();
    return -1;
  }
}

int rg_sky_matrix(rg_sky* s, float* out) {
  float* d = nullptr;
  try {
    if (!s) throw std::runtime_error("no sky");
    size_t n = (size_t)s->plan.npatch * s->plan.ncols * 3;
    hipck(hipMalloc(&d, n * sizeof(float)), "hipMalloc");
    rg::sky_matrix(s->plan, d, nullptr);
    hipck(hipMemcpy(out, d, n * sizeof(float), hipMemcpyDeviceToHost), "download");
    (void)hipFree(d);
    return 0;
  } catch (const std::exception& e) {
    s->err = e.what();
    if (d) (void)hipFree(d);
    return -1;
  }
}

}  // extern "C"
