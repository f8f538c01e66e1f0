// omr_bvh.cpp -- binned SAH build + collapse to 128-byte 4-wide nodes.
#include "omr_bvh.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <thread>

namespace omr {
namespace {

struct Box {
  float lo[3] = {kInf, kInf, kInf};
  float hi[3] = {-kInf, -kInf, -kInf};
  void grow(const Box &b) {
    for (int a = 0; a < 3; a++) {
      lo[a] = std::min(lo[a], b.lo[a]);
      hi[a] = std::max(hi[a], b.hi[a]);
    }
  }
  float area() const {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
    return dx * dy + dy * dz + dz * dx;
  }
};

struct PrimInfo {
  Box box;
  float c[3];
};

// Binary node of the intermediate tree.
struct BNode {
  Box box;
  int32_t left = -1, right = -1;  // children (inner) or -1 (leaf)
  uint32_t first = 0, count = 0;  // leaf range into the index array
};

constexpr int kBins = 32;
constexpr uint32_t kLeafTarget = 4;
constexpr float kTraversalCost = 1.0f;  // relative to one primitive test

struct Builder {
  const std::vector<PrimInfo> &info;
  std::vector<uint32_t> &idx;
  std::vector<BNode> nodes;          // preallocated: references stay valid
  std::atomic<uint32_t> used{1};

  Builder(const std::vector<PrimInfo> &i, std::vector<uint32_t> &x) : info(i), idx(x) {
    nodes.resize(2 * info.size() + 1);
  }

  int32_t alloc(uint32_t n) { return (int32_t)used.fetch_add(n); }
  BNode &nodes_at(int32_t i) { return nodes[(size_t)i]; }

  // Builds the subtree over idx[first, first+count) into node `me`.
  void build(int32_t me, uint32_t first, uint32_t count, int par_depth) {
    Box box, cbox;
    for (uint32_t i = first; i < first + count; i++) {
      const PrimInfo &p = info[idx[i]];
      box.grow(p.box);
      for (int a = 0; a < 3; a++) {
        cbox.lo[a] = std::min(cbox.lo[a], p.c[a]);
        cbox.hi[a] = std::max(cbox.hi[a], p.c[a]);
      }
    }
    int best_axis = -1, best_split = -1;
    float leaf_cost = (float)count;
    float best_cost = kInf;
    if (count > 1) {
      float parea = box.area();
      float inv_parea = parea > 0.0f ? 1.0f / parea : 0.0f;
      for (int a = 0; a < 3; a++) {
        float ext = cbox.hi[a] - cbox.lo[a];
        if (!(ext > 0.0f)) continue;
        Box bb[kBins];
        uint32_t bn[kBins] = {0};
        float k = (float)kBins / ext;
        for (uint32_t i = first; i < first + count; i++) {
          const PrimInfo &p = info[idx[i]];
          int b = std::min(kBins - 1, std::max(0, (int)((p.c[a] - cbox.lo[a]) * k)));
          bn[b]++;
          bb[b].grow(p.box);
        }
        float rarea[kBins];
        uint32_t rcnt[kBins];
        Box acc;
        uint32_t n = 0;
        for (int b = kBins - 1; b > 0; b--) {
          acc.grow(bb[b]);
          n += bn[b];
          rarea[b] = acc.area();
          rcnt[b] = n;
        }
        Box lacc;
        uint32_t ln = 0;
        for (int s = 1; s < kBins; s++) {
          lacc.grow(bb[s - 1]);
          ln += bn[s - 1];
          if (ln == 0 || rcnt[s] == 0) continue;
          float cost = kTraversalCost + (lacc.area() * (float)ln + rarea[s] * (float)rcnt[s]) * inv_parea;
          if (cost < best_cost) {
            best_cost = cost;
            best_axis = a;
            best_split = s;
          }
        }
      }
    }
    bool make_leaf = count <= 1 || (count <= kLeafTarget && best_cost >= leaf_cost) ||
                     (count <= kMaxLeafPrims && best_cost >= leaf_cost);
    if (make_leaf && count <= kMaxLeafPrims) {
      BNode &n = nodes_at(me);
      n.box = box;
      n.left = n.right = -1;
      n.first = first;
      n.count = count;
      return;
    }
    uint32_t mid;
    if (best_axis < 0) {
      mid = first + count / 2;  // coincident centroids: split the range
    } else {
      float ext = cbox.hi[best_axis] - cbox.lo[best_axis];
      float k = (float)kBins / ext;
      auto it = std::partition(idx.begin() + first, idx.begin() + first + count, [&](uint32_t id) {
        int b = std::min(kBins - 1, std::max(0, (int)((info[id].c[best_axis] - cbox.lo[best_axis]) * k)));
        return b < best_split;
      });
      mid = (uint32_t)(it - idx.begin());
      if (mid == first || mid == first + count) mid = first + count / 2;
    }
    int32_t kids = alloc(2);
    {
      BNode &n = nodes_at(me);
      n.box = box;
      n.left = kids;
      n.right = kids + 1;
    }
    uint32_t nl = mid - first, nr = first + count - mid;
    if (par_depth > 0 && count > 20000) {
      auto f = std::async(std::launch::async, [&] { build(kids, first, nl, par_depth - 1); });
      build(kids + 1, mid, nr, par_depth - 1);
      f.get();
    } else {
      build(kids, first, nl, 0);
      build(kids + 1, mid, nr, 0);
    }
  }

};

template <int W> struct NodeOf;
template <> struct NodeOf<4> { using T = WNode4; };
template <> struct NodeOf<8> { using T = WNode8; };

// Collapses the binary tree into W-wide nodes: each node gathers up to W
// children by repeatedly opening its inner child of largest surface area;
// leaf(bi) gives the leaf code of binary leaf bi.  Child boxes grow by `pad`.
// Reports the depth in W-wide levels and the traversal stack bound (a visit
// pushes at most W - 1 entries).
template <int W, class Leaf>
void collapse(const std::vector<BNode> &bn, float pad, Leaf leaf, std::vector<typename NodeOf<W>::T> &nodes,
              uint32_t &depth, uint32_t &stack_bound) {
  using Node = typename NodeOf<W>::T;
  struct Pending {
    int32_t bnode;
    uint32_t wnode;
    uint32_t depth;
    uint32_t stack_above;
  };
  std::vector<Pending> work;
  nodes.clear();
  nodes.reserve(bn.size() / (W / 2) + 2);
  nodes.emplace_back();
  work.push_back({0, 0, 1, 0});
  uint32_t max_depth = 0, max_stack = 0;
  while (!work.empty()) {
    Pending pw = work.back();
    work.pop_back();
    max_depth = std::max(max_depth, pw.depth);
    int32_t kids[W];
    int nk = 0;
    const BNode &root = bn[(size_t)pw.bnode];
    if (root.left < 0) {
      kids[nk++] = pw.bnode;  // single-leaf scene: root node holds one leaf
    } else {
      kids[nk++] = root.left;
      kids[nk++] = root.right;
      while (nk < W) {
        int best = -1;
        float ba = -1.0f;
        for (int k = 0; k < nk; k++) {
          const BNode &c = bn[(size_t)kids[k]];
          if (c.left >= 0 && c.box.area() > ba) { ba = c.box.area(); best = k; }
        }
        if (best < 0) break;
        int32_t open = kids[best];
        kids[best] = bn[(size_t)open].left;
        kids[nk++] = bn[(size_t)open].right;
      }
    }
    Node w;
    std::memset(&w, 0, sizeof(w));
    // empty slots: code -1 and an inverted box
    for (int k = 0; k < W; k++) {
      w.child[k] = -1;
      w.lox[k] = w.loy[k] = w.loz[k] = kInf;
      w.hix[k] = w.hiy[k] = w.hiz[k] = -kInf;
    }
    w.meta[0] = (int32_t)pw.depth;
    uint32_t stack_here = pw.stack_above + (uint32_t)(nk > 0 ? nk - 1 : 0);
    max_stack = std::max(max_stack, stack_here);
    for (int k = 0; k < nk; k++) {
      const BNode &c = bn[(size_t)kids[k]];
      w.lox[k] = c.box.lo[0] - pad; w.hix[k] = c.box.hi[0] + pad;
      w.loy[k] = c.box.lo[1] - pad; w.hiy[k] = c.box.hi[1] + pad;
      w.loz[k] = c.box.lo[2] - pad; w.hiz[k] = c.box.hi[2] + pad;
      if (c.left < 0) {
        w.child[k] = leaf(kids[k]);
      } else {
        uint32_t wi = (uint32_t)nodes.size();
        nodes.emplace_back();
        w.child[k] = (int32_t)wi;
        work.push_back({kids[k], wi, pw.depth + 1, stack_here});
      }
    }
    nodes[pw.wnode] = w;
  }
  depth = max_depth;
  stack_bound = max_stack + 1;
}

}  // namespace

bool build_wide_bvh(const BuildInput &in, BuildOutput &out, std::string &err, int threads) {
  auto t0 = std::chrono::steady_clock::now();
  size_t n = in.n_tri + in.n_sph;
  out = BuildOutput();
  if (n == 0) {
    err = "scene has no primitives";
    return false;
  }
  if (n > kMaxLeafFirst) {
    err = "too many primitives for the leaf encoding (max 134217727)";
    return false;
  }
  std::vector<PrimInfo> info(n);
  float scale = 1.0f;
  for (size_t i = 0; i < n; i++) {
    PrimInfo &p = info[i];
    if (i < in.n_tri) {
      const float *t = in.tri + 9 * i;
      for (int a = 0; a < 3; a++) {
        float p0 = t[a], p1 = t[a] + t[3 + a], p2 = t[a] + t[6 + a];
        if (!std::isfinite(p0) || !std::isfinite(p1) || !std::isfinite(p2)) {
          err = "non-finite triangle coordinate at triangle " + std::to_string(i);
          return false;
        }
        p.box.lo[a] = std::min(std::min(p0, p1), p2);
        p.box.hi[a] = std::max(std::max(p0, p1), p2);
      }
    } else {
      const float *s = in.sph + 4 * (i - in.n_tri);
      if (!(s[3] > 0.0f) || !std::isfinite(s[0]) || !std::isfinite(s[1]) || !std::isfinite(s[2]) || !std::isfinite(s[3])) {
        err = "invalid sphere at sphere " + std::to_string(i - in.n_tri);
        return false;
      }
      for (int a = 0; a < 3; a++) {
        p.box.lo[a] = s[a] - s[3];
        p.box.hi[a] = s[a] + s[3];
      }
    }
    for (int a = 0; a < 3; a++) {
      p.c[a] = 0.5f * (p.box.lo[a] + p.box.hi[a]);
      scale = std::max(scale, std::max(std::fabs(p.box.lo[a]), std::fabs(p.box.hi[a])));
    }
  }
  // Padding: every child box grows by 2^-17 of the largest coordinate
  // magnitude, far more than the binary32 primitive tests can stray outside
  // a primitive's exact bounds.
  const float pad = scale * 7.62939453125e-06f;
  out.pad = pad;

  std::vector<uint32_t> idx(n);
  for (size_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
  Builder B(info, idx);
  unsigned hw = threads > 0 ? (unsigned)threads : std::max(1u, std::thread::hardware_concurrency());
  int par_depth = 0;
  while ((1u << par_depth) < hw * 2 && par_depth < 8) par_depth++;
  B.build(0, 0, (uint32_t)n, par_depth);
  std::vector<BNode> &bn = B.nodes;

  // Collapse into 4-wide nodes, then the same tree into 8-wide nodes.
  // Leaves are written in the depth-first order of the 4-wide collapse, so a
  // leaf's primitives are contiguous in `prims`; the 8-wide nodes reuse the
  // leaf codes.
  out.prims.reserve(n);
  out.prim_order.reserve(n);
  std::vector<int32_t> leaf_code((size_t)B.used.load(), -1);
  auto emit_leaf = [&](int32_t bi) -> int32_t {
    const BNode &b = bn[(size_t)bi];
    uint32_t first = (uint32_t)out.prims.size();
    for (uint32_t i = b.first; i < b.first + b.count; i++) {
      uint32_t id = idx[i];
      PrimRec r;
      std::memset(&r, 0, sizeof(r));
      uint32_t kind;
      if (id < in.n_tri) {
        const float *t = in.tri + 9 * (size_t)id;
        for (int a = 0; a < 3; a++) { r.a[a] = t[a]; r.b[a] = t[3 + a]; r.c[a] = t[6 + a]; }
        kind = 0;
      } else {
        const float *s = in.sph + 4 * (size_t)(id - in.n_tri);
        for (int a = 0; a < 3; a++) r.a[a] = s[a];
        r.b[0] = s[3];
        kind = 1;
      }
      std::memcpy(&r.a[3], &id, 4);
      std::memcpy(&r.b[3], &kind, 4);
      out.prims.push_back(r);
      out.prim_order.push_back(id);
    }
    return leaf_code[(size_t)bi] = make_leaf(first, b.count);
  };
  collapse<4>(bn, pad, emit_leaf, out.nodes, out.depth, out.stack_bound);
  collapse<8>(bn, pad, [&](int32_t bi) { return leaf_code[(size_t)bi]; }, out.nodes8, out.depth8,
              out.stack_bound8);
  out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return true;
}

}  // namespace omr
