// LBVH construction on the GPU (Karras 2012, "Maximizing parallelism in the
// construction of BVHs, octrees and k-d trees"):
//   1. 30-bit Morton codes of primitive centroids,
//   2. LSD radix sort with 4-bit digits, ranks computed per wavefront with
//      ballot/mbcnt and combined across the block's waves in LDS,
//   3. one thread per internal node finds its key range and split,
//   4. bottom-up refit: the second thread to reach a node writes both child
//      boxes into it (one 64-byte node holds both children's bounds).
#include <hip/hip_runtime.h>
#include "rg_render.hpp"

namespace rg {

constexpr int BB = 256;              // threads per block (4 wavefronts)
constexpr int NWAVE = BB / 64;
constexpr int RADIX_BITS = 4;
constexpr int RADIX = 1 << RADIX_BITS;

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void __launch_bounds__(BB) k_morton(const float* lo, const float* hi, uint32_t n, float3 slo, float3 sinv,
                                               uint32_t* keys, uint32_t* vals) {
  uint32_t i = blockIdx.x * BB + threadIdx.x;
  if (i >= n) return;
  float cx = 0.5f * (lo[3 * i + 0] + hi[3 * i + 0]);
  float cy = 0.5f * (lo[3 * i + 1] + hi[3 * i + 1]);
  float cz = 0.5f * (lo[3 * i + 2] + hi[3 * i + 2]);
  auto q = [](float c, float l, float inv) -> uint32_t {
    float f = (c - l) * inv * 1024.0f;
    f = fminf(fmaxf(f, 0.0f), 1023.0f);
    return (uint32_t)f;
  };
  uint32_t x = q(cx, slo.x, sinv.x), y = q(cy, slo.y, sinv.y), z = q(cz, slo.z, sinv.z);
  keys[i] = (expand_bits(x) << 2) | (expand_bits(y) << 1) | expand_bits(z);
  vals[i] = i;
}

__device__ __forceinline__ uint32_t mbcnt(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

// Per-block digit histogram, written digit-major: counts[d * nblocks + b].
__global__ void __launch_bounds__(BB) k_digit_count(const uint32_t* keys, uint32_t n, int shift, uint32_t* counts,
                                                    uint32_t nblocks) {
  __shared__ uint32_t wc[NWAVE][RADIX];
  uint32_t i = blockIdx.x * BB + threadIdx.x;
  int w = threadIdx.x >> 6;
  bool valid = i < n;
  uint32_t dig = valid ? (keys[i] >> shift) & (RADIX - 1) : RADIX;
  for (int d = 0; d < RADIX; ++d) {
    unsigned long long m = __ballot(dig == (uint32_t)d);
    if ((threadIdx.x & 63) == 0) wc[w][d] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (threadIdx.x < RADIX) {
    uint32_t s = 0;
    for (int k = 0; k < NWAVE; ++k) s += wc[k][threadIdx.x];
    counts[threadIdx.x * nblocks + blockIdx.x] = s;
  }
}

// Stable scatter: rank within the block = earlier waves' count + lane rank.
__global__ void __launch_bounds__(BB) k_digit_scatter(const uint32_t* keys, const uint32_t* vals, uint32_t n, int shift,
                                                      const uint32_t* offs, uint32_t nblocks, uint32_t* okeys,
                                                      uint32_t* ovals) {
  __shared__ uint32_t wc[NWAVE][RADIX];
  uint32_t i = blockIdx.x * BB + threadIdx.x;
  int w = threadIdx.x >> 6;
  bool valid = i < n;
  uint32_t key = valid ? keys[i] : 0;
  uint32_t dig = valid ? (key >> shift) & (RADIX - 1) : RADIX;
  uint32_t rank = 0;
  for (int d = 0; d < RADIX; ++d) {
    unsigned long long m = __ballot(dig == (uint32_t)d);
    if (dig == (uint32_t)d) rank = mbcnt(m);
    if ((threadIdx.x & 63) == 0) wc[w][d] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (!valid) return;
  uint32_t before = 0;
  for (int k = 0; k < w; ++k) before += wc[k][dig];
  uint32_t dst = offs[dig * nblocks + blockIdx.x] + before + rank;
  okeys[dst] = key;
  ovals[dst] = vals[i];
}

__device__ __forceinline__ int delta(const uint32_t* keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = keys[i], b = keys[j];
  if (a == b) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(a ^ b);
}

__global__ void __launch_bounds__(BB) k_karras(const uint32_t* keys, int n, BVHNode* nodes, int32_t* parent_int,
                                               int32_t* parent_leaf, int32_t* chl, int32_t* chr) {
  int i = blockIdx.x * BB + threadIdx.x;
  if (i >= n - 1) return;
  int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta(keys, n, i, j);
  int s = 0;
  for (int div = 2;; div <<= 1) {
    int t = (l + div - 1) / div;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  int gamma = i + s * d + (d < 0 ? -1 : 0);
  int left, right;
  if (min(i, j) == gamma) { left = ~gamma; parent_leaf[gamma] = i; }
  else { left = gamma; parent_int[gamma] = i; }
  if (max(i, j) == gamma + 1) { right = ~(gamma + 1); parent_leaf[gamma + 1] = i; }
  else { right = gamma + 1; parent_int[gamma + 1] = i; }
  nodes[i].left = left;
  nodes[i].right = right;
  chl[i] = left;
  chr[i] = right;
  nodes[i].pad0 = 0;
  nodes[i].pad1 = 0;
}

__global__ void __launch_bounds__(BB) k_refit(const float* lo, const float* hi, const uint32_t* vals, int n,
                                              BVHNode* nodes, const int32_t* parent_int, const int32_t* parent_leaf,
                                              const int32_t* chl, const int32_t* chr, float* nbox, uint32_t* flags,
                                              const int32_t* ids, int32_t* leaf_prim) {
  int k = blockIdx.x * BB + threadIdx.x;
  if (k >= n) return;
  leaf_prim[k] = ids[vals[k]];
  int p = parent_leaf[k];
  // the parent refers to this leaf by primitive index from now on
  if (chl[p] == ~k) nodes[p].left = ~ids[vals[k]];
  else nodes[p].right = ~ids[vals[k]];
  while (p >= 0) {
    __threadfence();
    if (atomicAdd(&flags[p], 1u) == 0) return;   // first arrival: the sibling finishes
    __threadfence();
    BVHNode& N = nodes[p];
    float b[2][6];
    int ch[2] = {chl[p], chr[p]};
    for (int c = 0; c < 2; ++c) {
      int r = ch[c];
      if (r < 0) {
        uint32_t pi = vals[~r];
        for (int a = 0; a < 3; ++a) { b[c][a] = lo[3 * pi + a]; b[c][3 + a] = hi[3 * pi + a]; }
      } else {
        const volatile float* nb = nbox + 6 * r;
        for (int a = 0; a < 6; ++a) b[c][a] = nb[a];
      }
    }
    for (int a = 0; a < 3; ++a) {
      N.lmin[a] = b[0][a]; N.lmax[a] = b[0][3 + a];
      N.rmin[a] = b[1][a]; N.rmax[a] = b[1][3 + a];
      nbox[6 * p + a] = fminf(b[0][a], b[1][a]);
      nbox[6 * p + 3 + a] = fmaxf(b[0][3 + a], b[1][3 + a]);
    }
    p = parent_int[p];
  }
}

static void chk(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}

__global__ void k_leaf_one(const int32_t* ids, int32_t* leaf_prim) { leaf_prim[0] = ids[0]; }

int build_lbvh(const float* d_lo, const float* d_hi, const int32_t* d_ids, uint32_t n, const float scene_lo[3],
               const float scene_hi[3], BVHNode* d_nodes, int32_t* d_leaf_prim, DevArena& arena, hipStream_t st) {
  if (n == 0) return -1;
  uint32_t* ka = (uint32_t*)arena.get("bvh_ka", 4 * n, st, false);
  uint32_t* va = (uint32_t*)arena.get("bvh_va", 4 * n, st, false);
  uint32_t* kb = (uint32_t*)arena.get("bvh_kb", 4 * n, st, false);
  uint32_t* vb = (uint32_t*)arena.get("bvh_vb", 4 * n, st, false);
  uint32_t nblocks = (n + BB - 1) / BB;
  float3 slo = make_float3(scene_lo[0], scene_lo[1], scene_lo[2]);
  float3 sinv;
  sinv.x = 1.0f / fmaxf(scene_hi[0] - scene_lo[0], 1e-12f);
  sinv.y = 1.0f / fmaxf(scene_hi[1] - scene_lo[1], 1e-12f);
  sinv.z = 1.0f / fmaxf(scene_hi[2] - scene_lo[2], 1e-12f);
  hipLaunchKernelGGL(k_morton, dim3(nblocks), dim3(BB), 0, st, d_lo, d_hi, n, slo, sinv, ka, va);
  uint32_t ncnt = RADIX * nblocks;
  uint32_t* counts = (uint32_t*)arena.get("bvh_counts", 4 * (size_t)ncnt, st, false);
  uint32_t* offs = (uint32_t*)arena.get("bvh_offs", 4 * ((size_t)ncnt + 1), st, false);
  for (int shift = 0; shift < 32; shift += RADIX_BITS) {
    hipLaunchKernelGGL(k_digit_count, dim3(nblocks), dim3(BB), 0, st, ka, n, shift, counts, nblocks);
    exclusive_scan(counts, offs, ncnt, arena, st, 8);
    hipLaunchKernelGGL(k_digit_scatter, dim3(nblocks), dim3(BB), 0, st, ka, va, n, shift, offs, nblocks, kb, vb);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  chk(hipGetLastError(), "lbvh sort");
  if (n == 1) {
    hipLaunchKernelGGL(k_leaf_one, dim3(1), dim3(1), 0, st, d_ids, d_leaf_prim);
    return ~0;
  }
  int32_t* pint = (int32_t*)arena.get("bvh_pint", 4 * (size_t)n, st, false);
  int32_t* pleaf = (int32_t*)arena.get("bvh_pleaf", 4 * (size_t)n, st, false);
  float* nbox = (float*)arena.get("bvh_nbox", 24 * (size_t)n, st, false);
  uint32_t* flags = (uint32_t*)arena.get("bvh_flags", 4 * (size_t)n, st, true);
  chk(hipMemsetAsync(pint, 0xff, 4 * (size_t)n, st), "memset");
  int32_t* chl = (int32_t*)arena.get("bvh_chl", 4 * (size_t)n, st, false);
  int32_t* chr = (int32_t*)arena.get("bvh_chr", 4 * (size_t)n, st, false);
  hipLaunchKernelGGL(k_karras, dim3((n - 1 + BB - 1) / BB), dim3(BB), 0, st, ka, (int)n, d_nodes, pint, pleaf, chl,
                     chr);
  hipLaunchKernelGGL(k_refit, dim3(nblocks), dim3(BB), 0, st, d_lo, d_hi, va, (int)n, d_nodes, pint, pleaf, chl, chr,
                     nbox, flags, d_ids, d_leaf_prim);
  chk(hipGetLastError(), "lbvh build");
  return 0;
}

}  // namespace rg
