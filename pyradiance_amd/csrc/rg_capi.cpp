// C ABI: contexts, scene upload, rtrace / rcontrib / dctimestep entry points.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <unordered_map>

#include "../../include/radgpu.h"
#include "rg_render.hpp"
#include "rg_scene.hpp"

using namespace rg;

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
  int profile = 0;              // 0 off, 1 time k_bounce launches, 2 also count visits
  RenderStats last_stats;
  unsigned long long last_visits[2] = {0, 0};
};

static thread_local std::string g_create_err;

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
    c->cs = compile_scene(c->db);
    CompiledScene& cs = c->cs;
    DevScene& S = c->dscene;
    S.prims = upload(c->scene_mem, "prims", cs.prims, c->st);
    S.geom = upload(c->scene_mem, "geom", cs.geom, c->st);
    S.mats = upload(c->scene_mem, "mats", cs.mats, c->st);
    S.srcs = upload(c->scene_mem, "srcs", cs.srcs, c->st);
    S.nprims = (int)cs.prims.size();
    S.nmats = (int)cs.mats.size();
    S.nsrcs = (int)cs.srcs.size();
    S.lamb = cs.lamb;
    S.cusize = cs.cusize;
    uint32_t n = (uint32_t)cs.prims.size();
    float* dlo = upload(c->scene_mem, "blo", cs.lo, c->st);
    float* dhi = upload(c->scene_mem, "bhi", cs.hi, c->st);
    BVHNode* nodes = (BVHNode*)c->scene_mem.get("nodes", sizeof(BVHNode) * std::max<uint32_t>(n, 1), c->st, true);
    int32_t* leaf = (int32_t*)c->scene_mem.get("leaf", sizeof(int32_t) * std::max<uint32_t>(n, 1), c->st, false);
    float slo[3], shi[3];
    for (int a = 0; a < 3; ++a) { slo[a] = (float)cs.bmin[a]; shi[a] = (float)cs.bmax[a]; }
    S.root = n ? build_lbvh(dlo, dhi, n, slo, shi, nodes, leaf, c->arena, c->st) : -1;
    S.nodes = nodes;
    S.leaf_prim = leaf;
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

static void run(rg_ctx* c, const double* rays, long long nrays, int accum, int flags, int mode,
                const std::vector<BinSpec>& bins, int ncols, double* out) {
  if (!c->committed) throw std::runtime_error("scene not committed");
  if (nrays < 0 || nrays > 0x7fffffffLL) throw std::runtime_error("bad ray count");
  if (accum < 1) throw std::runtime_error("bad accumulation count");
  hipck(hipSetDevice(c->device), "hipSetDevice");
  c->last_nrays = 0;
  long long nrec = nrays / accum;
  size_t outn = (size_t)nrec * ncols;
  double* d_out = (double*)c->arena.get("out", sizeof(double) * std::max<size_t>(outn, 1), c->st, true);
  double* d_rays = (double*)c->arena.get("rays", sizeof(double) * 6 * std::max<long long>(nrays, 1), c->st, false);
  unsigned long long* d_nr = (unsigned long long*)c->arena.get("nrays", 8, c->st, true);
  uint32_t* d_err = (uint32_t*)c->arena.get("err", 8, c->st, true);
  BinSpec* d_bins = nullptr;
  if (!bins.empty()) {
    d_bins = (BinSpec*)c->arena.get("bins", sizeof(BinSpec) * bins.size(), c->st, false);
    hipck(hipMemcpyAsync(d_bins, bins.data(), sizeof(BinSpec) * bins.size(), hipMemcpyHostToDevice, c->st), "bins");
  }
  if (nrays)
    hipck(hipMemcpyAsync(d_rays, rays, sizeof(double) * 6 * nrays, hipMemcpyHostToDevice, c->st), "rays");
  RunCfg cfg;
  cfg.p = c->params;
  cfg.mode = mode;
  cfg.irrad = (flags & 16) ? 1 : 0;
  cfg.ncols = ncols;
  cfg.nbinspec = (int)bins.size();
  cfg.bins = d_bins;
  cfg.out = d_out;
  cfg.nrays = d_nr;
  cfg.err = d_err;
  cfg.accum = accum;
  cfg.visits = nullptr;
  unsigned long long* d_vis = nullptr;
  if (c->profile >= 2) {
    d_vis = (unsigned long long*)c->arena.get("visits", 16, c->st, true);
    cfg.visits = d_vis;
  }
  c->last_stats = RenderStats();
  c->last_stats.timing = c->profile >= 1;
  if (nrays) render(c->dscene, cfg, d_rays, (uint32_t)nrays, (flags & 8) != 0, c->arena, c->st, &c->last_stats);
  if (d_vis) hipck(hipMemcpyAsync(c->last_visits, d_vis, 16, hipMemcpyDeviceToHost, c->st), "visits");
  else c->last_visits[0] = c->last_visits[1] = 0;
  uint32_t herr = 0;
  hipck(hipMemcpyAsync(&herr, d_err, 4, hipMemcpyDeviceToHost, c->st), "err");
  hipck(hipMemcpyAsync(&c->last_nrays, d_nr, 8, hipMemcpyDeviceToHost, c->st), "nrays");
  if (outn) hipck(hipMemcpyAsync(out, d_out, sizeof(double) * outn, hipMemcpyDeviceToHost, c->st), "out");
  hipck(hipStreamSynchronize(c->st), "render");
  if (herr & ERR_STACK) throw std::runtime_error("BVH traversal stack overflow");
  if (herr & ERR_CAPACITY) throw std::runtime_error("spawn buffer overflow");
  if (herr & ERR_NOMAT) throw std::runtime_error("surface without material");
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

int rg_contrib(rg_ctx* c, const double* rays, long long nrec, int accum, int flags, const rg_modifier* mods,
               int nmods, double* out) {
  RG_TRY(c, {
    if (!c->committed) throw std::runtime_error("scene not committed");
    std::vector<BinSpec> bins(nmods);
    std::unordered_map<std::string, int> slot;
    int col = 0;
    for (int i = 0; i < nmods; ++i) {
      const rg_modifier& m = mods[i];
      if (!m.name || !m.name[0]) throw std::runtime_error("modifier without a name");
      if (!strcmp(m.name, "void")) throw std::runtime_error("cannot track 'void' modifier");
      if (slot.count(m.name)) throw std::runtime_error(std::string("duplicate modifier '") + m.name + "'");
      if (m.nbins < 1) throw std::runtime_error("bad bin count");
      if (m.bin_kind < 0 || m.bin_kind > 4) throw std::runtime_error("bad bin kind");
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
    // assign tracked-modifier slots by modifier name (rt/rcontrib.c:289)
    CompiledScene& cs = c->cs;
    std::vector<Prim> prims = cs.prims;
    bool any = false;
    for (size_t k = 0; k < prims.size(); ++k) {
      auto it = slot.find(cs.prim_modname[k]);
      prims[k].modslot = it == slot.end() ? -1 : it->second;
      any |= prims[k].modslot >= 0;
    }
    std::vector<Source> srcs = cs.srcs;
    for (size_t k = 0; k < srcs.size(); ++k) {
      auto it = slot.find(cs.src_modname[k]);
      srcs[k].modslot = it == slot.end() ? -1 : it->second;
      srcs[k].flags |= SFOLLOW;      // rt/rcontrib.c:235-237
    }
    if (c->params.contrib) {
      for (size_t k = 0; k < prims.size(); ++k) {
        int mk = prims[k].mat >= 0 ? cs.mats[prims[k].mat].kind : -1;
        if (prims[k].modslot >= 0 && mk != MK_LIGHT && mk != MK_GLOW && mk != MK_ILLUM)
          throw std::runtime_error("-V+ is supported for light-source modifiers only on the GPU path");
      }
    }
    (void)any;
    DevScene S = c->dscene;
    S.prims = upload(c->arena, "prims_c", prims, c->st);
    S.srcs = upload(c->arena, "srcs_c", srcs, c->st);
    DevScene saved = c->dscene;
    c->dscene = S;
    try {
      run(c, rays, nrec * (long long)accum, accum, flags, 1, bins, 3 * col, out);
    } catch (...) {
      c->dscene = saved;
      throw;
    }
    c->dscene = saved;
  });
}

int rg_last_nrays(rg_ctx* c, unsigned long long* n) {
  RG_TRY(c, { *n = c->last_nrays; });
}

int rg_set_profiling(rg_ctx* c, int level) {
  RG_TRY(c, {
    if (level < 0 || level > 2) throw std::runtime_error("bad profiling level");
    c->profile = level;
  });
}

int rg_last_stats(rg_ctx* c, double* out, int n) {
  RG_TRY(c, {
    double v[8] = {(double)c->last_nrays, (double)c->last_stats.launches, (double)c->last_stats.threads,
                   c->last_stats.bounce_ms, (double)c->last_visits[0], (double)c->last_visits[1],
                   (double)c->last_stats.generations, c->last_stats.shade_ms};
    for (int i = 0; i < n && i < 8; ++i) out[i] = v[i];
  });
}

int rg_dcsky_device(const float* A, const float* B, float* C, int M, int K, int N, int mode) {
  try {
    if (M <= 0 || K <= 0 || N <= 0) throw std::runtime_error("bad matrix dimensions");
    if (mode < 0 || mode > 1) throw std::runtime_error("bad mode");
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

}  // extern "C"
