// Host-side declarations for the renderer, the BVH builder and the device
// memory arena that keeps per-context buffers alive between calls.
#pragma once
#include <hip/hip_runtime.h>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>
#include "rg_shade.hpp"

namespace rg {

// Named device buffers owned by a context; grown on demand, freed on destroy.
class DevArena {
 public:
  ~DevArena() { release(); }
  void* get(const std::string& name, size_t bytes, hipStream_t st, bool zero) {
    auto it = bufs_.find(name);
    if (it == bufs_.end() || it->second.bytes < bytes) {
      if (it != bufs_.end()) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(it->second.p);
        bufs_.erase(it);
      }
      void* p = nullptr;
      size_t want = bytes ? bytes : 16;
      hipError_t e = hipMalloc(&p, want);
      if (e != hipSuccess)
        throw std::runtime_error("device allocation of " + std::to_string(want) + " bytes failed: " +
                                 hipGetErrorString(e));
      bufs_[name] = {p, want};
      it = bufs_.find(name);
    }
    if (zero) (void)hipMemsetAsync(it->second.p, 0, bytes ? bytes : 16, st);
    return it->second.p;
  }
  uint32_t* host_u32(size_t n) {
    if (host_n_ < n) {
      if (host_) (void)hipHostFree(host_);
      host_ = nullptr;
      if (hipHostMalloc((void**)&host_, n * sizeof(uint32_t), 0) != hipSuccess)
        throw std::runtime_error("pinned host allocation failed");
      host_n_ = n;
    }
    return host_;
  }
  void release() {
    for (auto& kv : bufs_) (void)hipFree(kv.second.p);
    bufs_.clear();
    if (host_) (void)hipHostFree(host_);
    host_ = nullptr;
    host_n_ = 0;
  }
  size_t bytes() const {
    size_t t = 0;
    for (auto& kv : bufs_) t += kv.second.bytes;
    return t;
  }

 private:
  struct Buf { void* p; size_t bytes; };
  std::map<std::string, Buf> bufs_;
  uint32_t* host_ = nullptr;
  size_t host_n_ = 0;
};

struct RenderStats {
  long generations = 0;
  long launches = 0;
  unsigned long long threads = 0;
  bool timing = false;       // time every generation's kernels with HIP events
  double bounce_ms = 0;      // summed traversal-kernel (k_trace) durations
  double shade_ms = 0;       // summed shading-kernel (k_shade) durations
  int node_bytes = 64;       // bytes of one tree node the traversal fetched (64 binary, 128 four-wide)
};

// Per-lane traversal stack entries (LDS and private memory together).
constexpr int kTraceStackEntries = 64;

void exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t n, DevArena& arena, hipStream_t st, int level);

// Launch shape of the wavefront renderer (set per context, chosen by
// measurement: see rg_set_tuning).
struct Tuning {
  int trace_kernel = 1;          // 0: one child per thread, 1: persistent with dynamic ray fetch
  uint32_t refill = 16;          // idle lanes that make a persistent wavefront fetch new rays
  uint32_t window = 1u << 24;    // children per generation launch (at most a spawn slot's share of device memory)
  uint32_t blocks_per_cu = 4;    // persistent grid: blocks of 256 per compute unit
  uint32_t max_big = 8;          // room-sized primitives kept out of the BVH when there are at most this many (applies at commit)
  int wide = 1;                  // persistent traversal walks the four-wide tree when the scene has one
  int shadow_cache = 1;          // pre-empt shadow rays by the reference's shadow cache (rt/srcobstr.c)
  int child_order = 1;           // hemisphere children in 8 x 8 blocks of strata (1) or by stratum row (0)
  int inflight = 1;              // windows queued ahead of the host's read-back (1 or 2)
};

// max_srcs: the largest source count over the scenes of a batch (sizes the
// spawn buffers); S.nsrcs is used when that is larger.
void render(const DevScene& S, RunCfg cfg, const double* d_rays, uint32_t nsens, bool imm_irrad, int max_srcs,
            const Tuning& tn, DevArena& arena, hipStream_t st, RenderStats* stats);

// First intersection of n rays (n x 6 doubles): distance (FHUGE on a miss),
// unflipped normal (n x 3), primitive index (-1 on a miss).
void first_hit(const DevScene& S, const double* d_rays, uint32_t n, double* d_t, double* d_n, int32_t* d_p,
               uint32_t* d_err, hipStream_t st);

// Shadow-cache fill: n cast rays (n x 6 doubles, d_src their sources) file
// their first recordable blocker into d_obs; d_counts[source] += the
// closest-hit queries they took.
void shadow_cast(const DevScene& S, const double* d_rays, const int32_t* d_src, uint32_t n, unsigned long long* d_obs,
                 unsigned long long* d_counts, uint32_t* d_err, hipStream_t st);

// BVH over the primitives' FP32 bounds (lo/hi: 3*n floats, already widened).
// Fills nodes (n-1 entries) and leaf_prim (n entries); returns the root ref.
// d_ids maps the n boxes to primitive indices written into leaf_prim.
int build_lbvh(const float* d_lo, const float* d_hi, const int32_t* d_ids, uint32_t n, const float scene_lo[3],
               const float scene_hi[3], BVHNode* d_nodes, int32_t* d_leaf_prim, DevArena& arena, hipStream_t st);

// DC x sky product C[M,N,3] = A[M,K,3] x B[K,N,3], per channel.
// mode 0: FP32 FMA, mode 1: BF16 hi/lo split on MFMA with FP32 accumulation
// (LDS-staged 128 x 128 tiles for long K), 2 / 3: the split product forced to
// per-wavefront operand loads / LDS-staged tiles.
void dcsky(const float* A, const float* B, float* C, int M, int K, int N, int mode, hipStream_t st);

}  // namespace rg
