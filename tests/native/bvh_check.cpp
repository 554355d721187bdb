// Host-side structural check of the SAH BVH builder (pyradiance_amd/csrc/rg_bvh.cpp):
// random boxes -> tree -> every primitive referenced once, every child box
// encloses its subtree, depth within the limit, and a stack walk that finds
// the nearest box entry for random rays agrees with brute force.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <tuple>
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

// Four-wide form of the same tree (rg::collapse_bvh4): each four-wide node,
// met in a walk from the root alongside the binary tree, holds the binary
// node's grandchildren (or its leaf child) with their boxes in slot order;
// every four-wide node is reached exactly once and numbered after its
// parent; the walk reaches every primitive once and finds the same nearest
// box entry as brute force.  Returns 0 or a bit mask.
extern "C" int check_bvh4(int n, int nrays, unsigned seed, int max_depth, int clustered, int* depth4_out) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> U(0.0f, 1.0f);
  std::vector<float> lo(3 * n), hi(3 * n);
  std::vector<int32_t> ids(n);
  for (int i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      float c = clustered ? (float)(i % 7) : U(rng) * 10.0f;
      float h = 0.001f + U(rng) * 0.2f;
      lo[3 * i + a] = c - h;
      hi[3 * i + a] = c + h;
    }
    ids[i] = i;
  }
  std::vector<BVHNode> nodes;
  int32_t root = rg::build_sah_bvh(lo.data(), hi.data(), ids.data(), (uint32_t)n, nodes, max_depth);
  std::vector<rg::BVH4Node> q;
  int d4 = rg::collapse_bvh4(nodes, root, q);
  if (depth4_out) *depth4_out = d4;
  if (n < 2) return (d4 == 0 && q.empty()) ? 0 : 2;
  int bad = 0;
  // parallel walk: binary node b <-> four-wide node j
  std::vector<int> hits(q.size(), 0), seen(n, 0);
  int deep = 0;
  std::vector<std::tuple<int32_t, int32_t, int>> st{{root, 0, 1}};
  while (!st.empty()) {
    auto [b, j, d] = st.back();
    st.pop_back();
    if (j < 0 || j >= (int)q.size()) { bad |= 1; continue; }
    hits[(size_t)j]++;
    deep = std::max(deep, d);
    const BVHNode& N = nodes[(size_t)b];
    int32_t wref[4];
    const float *wlo[4], *whi[4];
    int k = 0;
    for (int side = 0; side < 2; ++side) {
      int32_t ref = side ? N.right : N.left;
      if (ref >= 0) {
        const BVHNode& Cn = nodes[(size_t)ref];
        wref[k] = Cn.left; wlo[k] = Cn.lmin; whi[k] = Cn.lmax; ++k;
        wref[k] = Cn.right; wlo[k] = Cn.rmin; whi[k] = Cn.rmax; ++k;
      } else {
        wref[k] = ref; wlo[k] = side ? N.rmin : N.lmin; whi[k] = side ? N.rmax : N.lmax; ++k;
      }
    }
    const rg::BVH4Node& Q = q[(size_t)j];
    for (int s2 = 0; s2 < 4; ++s2) {
      if (s2 >= k) { if (Q.child[s2] != INT32_MIN) bad |= 4; continue; }
      const float got_lo[3] = {Q.lox[s2], Q.loy[s2], Q.loz[s2]}, got_hi[3] = {Q.hix[s2], Q.hiy[s2], Q.hiz[s2]};
      for (int a = 0; a < 3; ++a)
        if (got_lo[a] != wlo[s2][a] || got_hi[a] != whi[s2][a]) bad |= 8;
      if (wref[s2] < 0) {
        if (Q.child[s2] != wref[s2]) bad |= 4;
        else if (~wref[s2] < n) seen[~wref[s2]]++;
        else bad |= 16;
      } else {
        if (Q.child[s2] <= j) bad |= 4;                 // children numbered after their parent
        st.push_back({wref[s2], Q.child[s2], d + 1});
      }
    }
  }
  for (size_t j = 0; j < q.size(); ++j)
    if (hits[j] != 1) bad |= 2;
  for (int i = 0; i < n; ++i)
    if (seen[i] != 1) bad |= 16;
  if (deep != d4) bad |= 32;
  // nearest box entry: four-wide walk vs brute force
  for (int r = 0; r < nrays; ++r) {
    float o[3], d[3], inv[3];
    for (int a = 0; a < 3; ++a) { o[a] = U(rng) * 12.0f - 1.0f; d[a] = U(rng) * 2.0f - 1.0f; }
    for (int a = 0; a < 3; ++a) inv[a] = 1.0f / (std::fabs(d[a]) < 1e-20f ? 1e-20f : d[a]);
    float best = 1e30f, tb = 1e30f;
    int bi = -1, ti = -1;
    for (int i = 0; i < n; ++i) {
      float tn;
      if (slab(&lo[3 * i], &hi[3 * i], o, inv, tn)) {
        tn = std::fmax(tn, 0.0f);
        if (tn < best || (tn == best && i < bi)) { best = tn; bi = i; }
      }
    }
    std::vector<int> stk{0};
    while (!stk.empty()) {
      const rg::BVH4Node& Q = q[(size_t)stk.back()];
      stk.pop_back();
      for (int k = 0; k < 4; ++k) {
        if (Q.child[k] == INT32_MIN) continue;
        const float mn[3] = {Q.lox[k], Q.loy[k], Q.loz[k]}, mx[3] = {Q.hix[k], Q.hiy[k], Q.hiz[k]};
        float tn;
        if (!slab(mn, mx, o, inv, tn)) continue;
        if (Q.child[k] >= 0) { stk.push_back(Q.child[k]); continue; }
        int p = ~Q.child[k];
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

// The tree built by one host thread and by up to 16 (RG_BVH_THREADS) for the
// same boxes: returns 0 when the node arrays are identical byte for byte.
extern "C" int check_sah_threads(int n, unsigned seed, int clustered) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> U(0.0f, 1.0f);
  std::vector<float> lo(3 * n), hi(3 * n);
  std::vector<int32_t> ids(n);
  for (int i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      float c = clustered ? (float)(i % 13) : U(rng) * 10.0f;
      float h = 0.001f + U(rng) * 0.2f;
      lo[3 * i + a] = c - h;
      hi[3 * i + a] = c + h;
    }
    ids[i] = i;
  }
  std::vector<BVHNode> one, many;
  setenv("RG_BVH_THREADS", "1", 1);
  const int32_t r1 = rg::build_sah_bvh(lo.data(), hi.data(), ids.data(), (uint32_t)n, one, 60);
  setenv("RG_BVH_THREADS", "16", 1);
  const int32_t r2 = rg::build_sah_bvh(lo.data(), hi.data(), ids.data(), (uint32_t)n, many, 60);
  unsetenv("RG_BVH_THREADS");
  if (r1 != r2 || one.size() != many.size()) return 1;
  return std::memcmp(one.data(), many.data(), one.size() * sizeof(BVHNode)) != 0 ? 2 : 0;
}
