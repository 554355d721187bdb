// Host-side structural check of the SAH BVH builder (pyradiance_amd/csrc/rg_bvh.cpp):
// random boxes -> tree -> every primitive referenced once, every child box
// encloses its subtree, depth within the limit, and a stack walk that finds
// the nearest box entry for random rays agrees with brute force.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "rg_bvh.hpp"

using rg::BVHNode;

namespace {

bool slab(const float* lo, const float* hi, const float* o, const float* inv, float& tn) {
  float t0 = -1e30f, t1 = 1e30f;
  for (int a = 0; a < 3; ++a) {
    float u = (lo[a] - o[a]) * inv[a], v = (hi[a] - o[a]) * inv[a];
    t0 = std::fmax(t0, std::fmin(u, v));
    t1 = std::fmin(t1, std::fmax(u, v));
  }
  tn = t0;
  return t0 <= t1 && t1 >= 0.0f;
}

struct Walk {
  const std::vector<BVHNode>& nodes;
  const std::vector<float>& lo;
  const std::vector<float>& hi;
  int* seen;
  int n;
  int bad = 0;
  int maxdepth = 0;
  void encloses(const float* mn, const float* mx, int ref, int depth) {
    if (ref < 0) {
      int p = ~ref;
      if (p < 0 || p >= n) { bad |= 1; return; }
      seen[p]++;
      for (int a = 0; a < 3; ++a)
        if (lo[3 * p + a] < mn[a] || hi[3 * p + a] > mx[a]) bad |= 2;
      return;
    }
    if (depth > maxdepth) maxdepth = depth;
    const BVHNode& N = nodes[ref];
    for (int a = 0; a < 3; ++a) {
      if (N.lmin[a] < mn[a] || N.lmax[a] > mx[a] || N.rmin[a] < mn[a] || N.rmax[a] > mx[a]) bad |= 4;
    }
    encloses(N.lmin, N.lmax, N.left, depth + 1);
    encloses(N.rmin, N.rmax, N.right, depth + 1);
  }
};

}  // namespace

// returns 0 when every check passes, else a bit mask of failed checks
extern "C" int check_sah(int n, int nrays, unsigned seed, int max_depth, int clustered, int* depth_out) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> U(0.0f, 1.0f);
  std::vector<float> lo(3 * n), hi(3 * n);
  std::vector<int32_t> ids(n);
  for (int i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      // clustered: many coincident centroids exercise the median fallback
      float c = clustered ? (float)(i % 7) : U(rng) * 10.0f;
      float h = 0.001f + U(rng) * 0.2f;
      lo[3 * i + a] = c - h;
      hi[3 * i + a] = c + h;
    }
    ids[i] = i;
  }
  std::vector<BVHNode> nodes;
  int32_t root = rg::build_sah_bvh(lo.data(), hi.data(), ids.data(), (uint32_t)n, nodes, max_depth);
  int bad = 0;
  if (n >= 2 && (root != 0 || (int)nodes.size() != n - 1)) bad |= 8;
  if (n == 1 && root != ~0) bad |= 8;
  std::vector<int> seen(n, 0);
  Walk w{nodes, lo, hi, seen.data(), n};
  if (n >= 2) {
    float big[3] = {-1e30f, -1e30f, -1e30f}, bigh[3] = {1e30f, 1e30f, 1e30f};
    w.encloses(big, bigh, root, 0);
    bad |= w.bad;
    for (int i = 0; i < n; ++i)
      if (seen[i] != 1) bad |= 16;
  }
  if (depth_out) *depth_out = w.maxdepth;
  if (w.maxdepth > max_depth) bad |= 32;
  // nearest box entry along random rays: tree walk vs brute force
  for (int r = 0; r < nrays && n >= 2; ++r) {
    float o[3], d[3], inv[3];
    for (int a = 0; a < 3; ++a) { o[a] = U(rng) * 12.0f - 1.0f; d[a] = U(rng) * 2.0f - 1.0f; }
    for (int a = 0; a < 3; ++a) inv[a] = 1.0f / (std::fabs(d[a]) < 1e-20f ? 1e-20f : d[a]);
    float best = 1e30f;
    int bi = -1;
    for (int i = 0; i < n; ++i) {
      float tn;
      if (slab(&lo[3 * i], &hi[3 * i], o, inv, tn)) {
        tn = std::fmax(tn, 0.0f);
        if (tn < best || (tn == best && i < bi)) { best = tn; bi = i; }
      }
    }
    float tb = 1e30f;
    int ti = -1;
    std::vector<int> stk{root};
    while (!stk.empty()) {
      int ref = stk.back();
      stk.pop_back();
      const BVHNode& N = nodes[ref];
      const float* mn[2] = {N.lmin, N.rmin};
      const float* mx[2] = {N.lmax, N.rmax};
      int ch[2] = {N.left, N.right};
      for (int k = 0; k < 2; ++k) {
        float tn;
        if (!slab(mn[k], mx[k], o, inv, tn)) continue;
        if (ch[k] >= 0) { stk.push_back(ch[k]); continue; }
        int p = ~ch[k];
        float tp;
        if (slab(&lo[3 * p], &hi[3 * p], o, inv, tp)) {
          tp = std::fmax(tp, 0.0f);
          if (tp < tb || (tp == tb && p < ti)) { tb = tp; ti = p; }
        }
      }
    }
    if (ti != bi) bad |= 64;
  }
  return bad;
}
