// Binned SAH builder (host).  Each split chooses, over three axes and 32
// centroid bins, the partition minimising area(L)*n(L) + area(R)*n(R); ranges
// whose centroids coincide, and ranges below the median-split depth, are split
// at the object median of the longest centroid axis.  Nodes are written in
// depth-first order (left subtree right after its parent) so a walk down the
// near side stays within a few cache lines.
#include "rg_bvh.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>

namespace rg {

namespace {

struct Box {
  float lo[3], hi[3];
  void reset() {
    for (int a = 0; a < 3; ++a) { lo[a] = 3.0e38f; hi[a] = -3.0e38f; }
  }
  void grow(const float* l, const float* h) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], l[a]);
      hi[a] = std::max(hi[a], h[a]);
    }
  }
  void grow(const Box& b) { grow(b.lo, b.hi); }
  double area() const {
    double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return dx * dy + dy * dz + dz * dx;
  }
};

struct Ref {
  float lo[3], hi[3];
  float c[3];
  int32_t id;
};

struct Task {
  uint32_t b, e;       // ref range
  int32_t parent;      // parent node (-1: root)
  int side;            // 0: left slot, 1: right slot
  int depth;
  int32_t node;        // this range's node (ranges of two or more refs)
};

constexpr int NBINS = 32;

// bin of a centroid coordinate, clamped in float before the conversion
inline int bin_of(float c, float lo, float scale) {
  float k = (c - lo) * scale;
  k = std::min(std::max(k, 0.0f), (float)(NBINS - 1));
  return (int)k;
}

void set_child(BVHNode& N, int side, const Box& bx, int32_t ref) {
  float* mn = side ? N.rmin : N.lmin;
  float* mx = side ? N.rmax : N.lmax;
  for (int a = 0; a < 3; ++a) { mn[a] = bx.lo[a]; mx[a] = bx.hi[a]; }
  if (side) N.right = ref; else N.left = ref;
}

}  // namespace

// One range: its node (boxes of both children are written by the children
// themselves into their slots of this node), and the two child ranges pushed
// onto `out`.  Nodes are numbered in depth-first preorder with the left
// subtree first: a range of k refs holds k - 1 internal nodes, so the left
// child's node follows its parent and the right child's follows the left
// subtree.  Ranges are disjoint, so ranges of different subtrees can be
// processed by different threads.
static void process_range(std::vector<Ref>& refs, std::vector<BVHNode>& nodes, const Task& t, int max_depth,
                          int median_depth, std::vector<Task>& out) {
  Box bx, cb;
  bx.reset();
  cb.reset();
  for (uint32_t i = t.b; i < t.e; ++i) {
    bx.grow(refs[i].lo, refs[i].hi);
    cb.grow(refs[i].c, refs[i].c);
  }
  if (t.e - t.b == 1) {
    set_child(nodes[t.parent], t.side, bx, ~refs[t.b].id);
    return;
  }
  const int32_t me = t.node;
  if (t.depth > max_depth) throw std::runtime_error("BVH deeper than the traversal stack");
  std::memset(&nodes[me], 0, sizeof(BVHNode));
  if (t.parent >= 0) set_child(nodes[t.parent], t.side, bx, me);

  int axis = 0;
  float ext[3];
  for (int a = 0; a < 3; ++a) ext[a] = cb.hi[a] - cb.lo[a];
  if (ext[1] > ext[axis]) axis = 1;
  if (ext[2] > ext[axis]) axis = 2;
  uint32_t mid = t.b + (t.e - t.b) / 2;
  bool done = false;
  if (ext[axis] > 0.0f && t.depth < median_depth) {
    // binned SAH over all three axes
    double best = 1e300;
    int best_axis = -1, best_split = 0;
    for (int a = 0; a < 3; ++a) {
      if (!(ext[a] > 0.0f)) continue;
      Box bins[NBINS];
      uint32_t cnt[NBINS] = {};
      for (int k = 0; k < NBINS; ++k) bins[k].reset();
      const float scale = NBINS / ext[a];
      for (uint32_t i = t.b; i < t.e; ++i) {
        int k = bin_of(refs[i].c[a], cb.lo[a], scale);
        cnt[k]++;
        bins[k].grow(refs[i].lo, refs[i].hi);
      }
      double rarea[NBINS];
      uint32_t rcnt[NBINS];
      Box acc;
      acc.reset();
      uint32_t c = 0;
      for (int k = NBINS - 1; k > 0; --k) {
        acc.grow(bins[k]);
        c += cnt[k];
        rarea[k] = acc.area();
        rcnt[k] = c;
      }
      acc.reset();
      c = 0;
      for (int k = 0; k < NBINS - 1; ++k) {
        acc.grow(bins[k]);
        c += cnt[k];
        if (c == 0 || rcnt[k + 1] == 0) continue;
        double cost = acc.area() * c + rarea[k + 1] * rcnt[k + 1];
        if (cost < best) { best = cost; best_axis = a; best_split = k + 1; }
      }
    }
    if (best_axis >= 0) {
      const int a = best_axis;
      const float scale = NBINS / ext[a];
      const float l0 = cb.lo[a];
      auto it = std::partition(refs.begin() + t.b, refs.begin() + t.e,
                               [&](const Ref& r) { return bin_of(r.c[a], l0, scale) < best_split; });
      mid = (uint32_t)(it - refs.begin());
      done = mid > t.b && mid < t.e;
    }
  }
  if (!done) {
    mid = t.b + (t.e - t.b) / 2;
    std::nth_element(refs.begin() + t.b, refs.begin() + mid, refs.begin() + t.e,
                     [axis](const Ref& x, const Ref& y) { return x.c[axis] < y.c[axis]; });
  }
  // right first so that a stack pops the left subtree first
  out.push_back({mid, t.e, me, 1, t.depth + 1, me + (int32_t)(mid - t.b)});
  out.push_back({t.b, mid, me, 0, t.depth + 1, me + 1});
}

int32_t build_sah_bvh(const float* lo, const float* hi, const int32_t* ids, uint32_t n, std::vector<BVHNode>& nodes,
                      int max_depth) {
  nodes.clear();
  if (n == 0) return -1;
  if (n == 1) return ~0;
  std::vector<Ref> refs(n);
  for (uint32_t i = 0; i < n; ++i) {
    Ref& r = refs[i];
    for (int a = 0; a < 3; ++a) {
      r.lo[a] = lo[3 * i + a];
      r.hi[a] = hi[3 * i + a];
      r.c[a] = 0.5f * (r.lo[a] + r.hi[a]);
    }
    r.id = ids[i];
  }
  nodes.resize(n - 1);
  const int median_depth = std::max(1, max_depth - 26);   // leaves room for log2(n) median levels
  // The top of the tree is split on this thread, largest range first, until
  // there are enough ranges for the host threads; each remaining range's
  // subtree is then built by whichever thread takes it.  Every range is split
  // as the one-thread build splits it and its nodes land at the indices the
  // one-thread build gives them, so the tree is the same for any thread count.
  // RG_BVH_THREADS caps the thread count (1: the one-thread build)
  unsigned nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (const char* env = std::getenv("RG_BVH_THREADS")) {
    const long v = std::strtol(env, nullptr, 10);
    if (v >= 1) nthr = std::min<unsigned>(nthr, (unsigned)v);
  }
  const uint32_t kSerial = 1u << 16;     // ranges below this size are built by one thread
  std::vector<Task> front{{0, n, -1, 0, 0, 0}}, kids;
  if (nthr > 1 && n > kSerial) {
    for (;;) {
      auto big = std::max_element(front.begin(), front.end(),
                                  [](const Task& x, const Task& y) { return x.e - x.b < y.e - y.b; });
      if (big->e - big->b < kSerial || front.size() >= 8 * (size_t)nthr) break;
      const Task t = *big;
      front.erase(big);
      kids.clear();
      process_range(refs, nodes, t, max_depth, median_depth, kids);
      front.insert(front.end(), kids.begin(), kids.end());
    }
    std::sort(front.begin(), front.end(), [](const Task& x, const Task& y) { return x.e - x.b > y.e - y.b; });
  } else {
    nthr = 1;
  }
  std::atomic<size_t> next{0};
  std::mutex err_mu;
  std::string err;
  auto worker = [&]() {
    std::vector<Task> stack;
    try {
      for (size_t k; (k = next.fetch_add(1)) < front.size();) {
        stack.assign(1, front[k]);
        while (!stack.empty()) {
          const Task t = stack.back();
          stack.pop_back();
          process_range(refs, nodes, t, max_depth, median_depth, stack);
        }
      }
    } catch (const std::exception& e) {
      std::lock_guard<std::mutex> lk(err_mu);
      if (err.empty()) err = e.what();
      next.store(front.size());
    }
  };
  std::vector<std::thread> pool;
  for (unsigned i = 1; i < nthr; ++i) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  if (!err.empty()) throw std::runtime_error(err);
  return 0;
}

int collapse_bvh4(const std::vector<BVHNode>& nodes, int32_t root, std::vector<BVH4Node>& out) {
  out.clear();
  if (root < 0 || nodes.empty()) return 0;
  // the binary nodes a four-wide walk visits (the root, then grandchildren
  // of visited nodes), numbered in depth-first preorder with the first slot's
  // subtree first, so a node's first child follows it in memory
  auto slots = [&](int32_t b, int32_t* ref, const float** lo, const float** hi) {
    const BVHNode& N = nodes[(size_t)b];
    int k = 0;
    for (int side = 0; side < 2; ++side) {
      const int32_t r = side ? N.right : N.left;
      if (r >= 0) {
        const BVHNode& C = nodes[(size_t)r];
        ref[k] = C.left; lo[k] = C.lmin; hi[k] = C.lmax; ++k;
        ref[k] = C.right; lo[k] = C.rmin; hi[k] = C.rmax; ++k;
      } else {
        ref[k] = r; lo[k] = side ? N.rmin : N.lmin; hi[k] = side ? N.rmax : N.lmax; ++k;
      }
    }
    return k;
  };
  std::vector<int32_t> order, newidx(nodes.size(), -1), st{root};
  while (!st.empty()) {
    const int32_t b = st.back();
    st.pop_back();
    newidx[(size_t)b] = (int32_t)order.size();
    order.push_back(b);
    int32_t ref[4];
    const float *lo[4], *hi[4];
    const int n = slots(b, ref, lo, hi);
    for (int k = n; k--;)
      if (ref[k] >= 0) st.push_back(ref[k]);
  }
  out.assign(order.size(), BVH4Node{});
  std::vector<int> depth(order.size(), 1);
  int dmax = 0;
  for (size_t i = 0; i < order.size(); ++i) {
    int32_t ref[4];
    const float *lo[4], *hi[4];
    const int n = slots(order[i], ref, lo, hi);
    BVH4Node& q = out[i];
    dmax = std::max(dmax, depth[i]);
    for (int k = 0; k < 4; ++k) {
      if (k < n) {
        q.lox[k] = lo[k][0]; q.loy[k] = lo[k][1]; q.loz[k] = lo[k][2];
        q.hix[k] = hi[k][0]; q.hiy[k] = hi[k][1]; q.hiz[k] = hi[k][2];
        q.child[k] = ref[k] >= 0 ? newidx[(size_t)ref[k]] : ref[k];
        if (ref[k] >= 0) depth[(size_t)q.child[k]] = depth[i] + 1;   // preorder: children come later
      } else {
        q.lox[k] = q.loy[k] = q.loz[k] = 0.f;
        q.hix[k] = q.hiy[k] = q.hiz[k] = 0.f;
        q.child[k] = INT32_MIN;
      }
    }
  }
  return dmax;
}

}  // namespace rg
