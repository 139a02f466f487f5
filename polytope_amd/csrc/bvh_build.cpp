// Host BVH builder: makes exactly the decisions of the reference's binned-SAH
// builder and DFS flattening (src/cpu/acceleration/bvh.cpp:36-395, 652-698),
// so the GPU traverses the reference's own topology, leaf order and split axes
// (closest-hit ids on ties depend on them).
//
// What differs is only how the work is organised:
//   * no per-node index-vector copies: each node owns a range of one index
//     array, partitioned stably in bucket order (the order the reference
//     obtains by concatenating its bucket vectors, bvh.cpp:352-360);
//   * large nodes run their per-face passes on all threads in ordered chunks,
//     and subtrees below a size threshold are built by worker threads.
// Chunked reductions give identical bits: the reference's min/max/union use the
// selection `a < b ? a : b` (ties take the right operand), which is
// associative, and the SAH face counts are summed per bucket in bucket order as
// in the reference.  Per-node float expressions keep the reference's operation
// order; this file is compiled with -ffp-contract=off.

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <thread>
#include <utility>

#include "ptx_internal.h"

namespace ptx {
namespace {

constexpr float kInf = std::numeric_limits<float>::infinity();
constexpr unsigned kBuckets = 12;
constexpr size_t kParallelNode = 1u << 16;   // nodes this large use chunked passes
constexpr size_t kTaskNode = 1u << 13;       // subtrees at most this large become one task

struct Box {
   float lo[3], hi[3];
};

static inline void box_union(Box &a, const Box &b) {   // BoundingBox::UnionInPlace (BoundingBox.cpp:136-144)
   for (int k = 0; k < 3; k++) {
      a.lo[k] = a.lo[k] < b.lo[k] ? a.lo[k] : b.lo[k];
      a.hi[k] = a.hi[k] > b.hi[k] ? a.hi[k] : b.hi[k];
   }
}

static inline float surface_area(const Box &b) {        // BoundingBox::surface_area (BoundingBox.cpp:146-152)
   const float x = b.hi[0] - b.lo[0];
   const float y = b.hi[1] - b.lo[1];
   const float z = b.hi[2] - b.lo[2];
   return (2 * x * y) + (2 * y * z) + (2 * z * x);
}

struct FaceInfo {
   Box bb;
   float c[3];
};

struct Node {
   Box bb;
   uint32_t lo = 0, n = 0;      // index range
   int32_t high = -1;           // high child; low child is high + 1
   uint16_t axis = 0;
   uint16_t depth = 0;
};

class Pool {
 public:
   explicit Pool(unsigned threads) : nthreads_(threads ? threads : 1) {}
   unsigned size() const { return nthreads_; }
   // Runs fn(chunk, lo, hi) over `chunks` ordered chunks of [0, n).
   void chunks(size_t n, unsigned nch, const std::function<void(unsigned, size_t, size_t)> &fn) {
      if (nch <= 1) { fn(0, 0, n); return; }
      std::vector<std::thread> ts;
      ts.reserve(nch);
      for (unsigned c = 0; c < nch; c++) {
         const size_t lo = n * c / nch, hi = n * (c + 1) / nch;
         ts.emplace_back([=, &fn]() { fn(c, lo, hi); });
      }
      for (auto &t : ts) t.join();
   }

 private:
   unsigned nthreads_;
};

class Builder {
 public:
   Builder(const std::vector<WorldMesh> &meshes, unsigned threads) : meshes_(meshes), pool_(threads) {}

   void run(HostBvh &out);

 private:
   // Returns true if the node was split (children appended).
   bool split(uint32_t ni, bool parallel);
   uint32_t alloc_pair();

   const std::vector<WorldMesh> &meshes_;
   Pool pool_;
   std::vector<FaceInfo> info_;
   std::vector<uint32_t> mesh_of_, face_of_;     // per global face id
   std::vector<uint32_t> idx_, tmp_;
   std::vector<uint8_t> bucket_;
   std::vector<Node> nodes_;
   std::atomic<uint32_t> next_node_{1};
};

uint32_t Builder::alloc_pair() { return next_node_.fetch_add(2); }

bool Builder::split(uint32_t ni, bool parallel) {
   Node &node = nodes_[ni];
   const size_t lo = node.lo, n = node.n;
   if (n == 1) return false;                                  // bvh.cpp:141
   const float node_sa = surface_area(node.bb);
   if (node_sa == 0.f) return false;                          // bvh.cpp:145-148
   const float inv_sa = 1.f / node_sa;
   const unsigned nch = parallel ? pool_.size() : 1;

   // centroid bounds (bvh.cpp:152-173)
   std::vector<Box> cb(nch);
   pool_.chunks(n, nch, [&](unsigned c, size_t a, size_t b) {
      Box r = {{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}};
      for (size_t i = lo + a; i < lo + b; i++) {
         const float *cc = info_[idx_[i]].c;
         for (int k = 0; k < 3; k++) {
            r.lo[k] = r.lo[k] < cc[k] ? r.lo[k] : cc[k];
            r.hi[k] = r.hi[k] > cc[k] ? r.hi[k] : cc[k];
         }
      }
      cb[c] = r;
   });
   Box cbox = cb[0];
   for (unsigned c = 1; c < nch; c++)
      for (int k = 0; k < 3; k++) {
         cbox.lo[k] = cbox.lo[k] < cb[c].lo[k] ? cbox.lo[k] : cb[c].lo[k];
         cbox.hi[k] = cbox.hi[k] > cb[c].hi[k] ? cbox.hi[k] : cb[c].hi[k];
      }
   const float ex = cbox.hi[0] - cbox.lo[0];
   const float ey = cbox.hi[1] - cbox.lo[1];
   const float ez = cbox.hi[2] - cbox.lo[2];
   if (ex == ey && ey == ez && ez == 0.f) return false;       // bvh.cpp:184-189
   int axis;                                                  // bvh.cpp:197-214
   if (ex >= ey && ex >= ez) axis = 0;
   else if (ey >= ex && ey >= ez) axis = 1;
   else axis = 2;
   const float extent = axis == 0 ? ex : (axis == 1 ? ey : ez);
   const float amin = cbox.lo[axis];
   const float bucket_width = extent / (float)kBuckets;       // bvh.cpp:221

   // bucket assignment, per-chunk counts and bucket box unions (bvh.cpp:232-244)
   std::vector<std::array<uint32_t, kBuckets>> cnt(nch);
   std::vector<std::array<Box, kBuckets>> bbs(nch);
   std::vector<std::array<uint8_t, kBuckets>> seen(nch);
   pool_.chunks(n, nch, [&](unsigned c, size_t a, size_t b) {
      auto &ct = cnt[c];
      auto &bx = bbs[c];
      auto &sn = seen[c];
      ct.fill(0);
      sn.fill(0);
      for (size_t i = lo + a; i < lo + b; i++) {
         const FaceInfo &fi = info_[idx_[i]];
         const float raw = (fi.c[axis] - amin) / (float)bucket_width;
         unsigned bk = (unsigned)raw;
         if (bk == kBuckets) bk--;
         bucket_[i] = (uint8_t)bk;
         ct[bk]++;
         if (!sn[bk]) { bx[bk] = fi.bb; sn[bk] = 1; }
         else box_union(bx[bk], fi.bb);
      }
   });
   uint32_t counts[kBuckets] = {0};
   Box bucket_box[kBuckets];
   bool has[kBuckets] = {false};
   for (unsigned c = 0; c < nch; c++)
      for (unsigned k = 0; k < kBuckets; k++) {
         if (!seen[c][k]) continue;
         counts[k] += cnt[c][k];
         if (!has[k]) { bucket_box[k] = bbs[c][k]; has[k] = true; }
         else box_union(bucket_box[k], bbs[c][k]);
      }

   // SAH scans (bvh.cpp:246-326)
   float fwd[kBuckets], bwd[kBuckets];
   Box low_bbs[kBuckets], high_bbs[kBuckets];
   Box low_bb = bucket_box[0];            // bucket 0 is never empty: it holds the min centroid
   float num_low = 0.f;
   for (unsigned k = 0; k < kBuckets; k++) {
      if (counts[k]) box_union(low_bb, bucket_box[k]);
      low_bbs[k] = low_bb;
      num_low += counts[k];
      const float cost = surface_area(low_bb) * inv_sa * num_low * 1.f;
      fwd[k] = cost == 0 ? kInf : cost;
   }
   Box high_bb = bucket_box[kBuckets - 1];   // bucket 11 holds the max centroid
   float num_high = 0.f;
   for (int k = (int)kBuckets - 1; k >= 0; k--) {
      if (counts[k]) box_union(high_bb, bucket_box[k]);
      high_bbs[k] = high_bb;
      num_high += counts[k];
      const float cost = surface_area(high_bb) * inv_sa * num_high * 1.f;
      bwd[kBuckets - 1 - k] = cost == 0 ? kInf : cost;
   }
   float total[kBuckets - 1];
   for (unsigned k = 0; k < kBuckets - 1; k++) total[k] = 1.f + fwd[k] + bwd[kBuckets - 2 - k];
   const float leaf_cost = n * 1.f;
   bool make_leaf = true;
   for (unsigned k = 0; k < kBuckets - 1; k++)
      if (leaf_cost > total[k]) make_leaf = false;
   if (make_leaf && n <= 4) return false;                     // bvh.cpp:338-340
   unsigned best = 0;
   float best_cost = std::numeric_limits<float>::max();
   for (unsigned k = 0; k < kBuckets - 1; k++)
      if (total[k] < best_cost) { best_cost = total[k]; best = k; }
   size_t n_low = 0;
   for (unsigned k = 0; k <= best; k++) n_low += counts[k];
   node.axis = (uint16_t)axis;                                // bvh.cpp:362
   if (n_low == n || n_low == 0) return false;                // bvh.cpp:367-369

   // stable partition into bucket order
   std::vector<std::array<size_t, kBuckets>> pos(nch);
   {
      size_t start[kBuckets];
      size_t acc = 0;
      for (unsigned k = 0; k < kBuckets; k++) { start[k] = acc; acc += counts[k]; }
      for (unsigned c = 0; c < nch; c++)
         for (unsigned k = 0; k < kBuckets; k++) {
            pos[c][k] = start[k];
            start[k] += seen[c][k] ? cnt[c][k] : 0;
         }
   }
   pool_.chunks(n, nch, [&](unsigned c, size_t a, size_t b) {
      auto p = pos[c];
      for (size_t i = lo + a; i < lo + b; i++) tmp_[lo + p[bucket_[i]]++] = idx_[i];
   });
   pool_.chunks(n, nch, [&](unsigned, size_t a, size_t b) {
      std::memcpy(&idx_[lo + a], &tmp_[lo + a], (b - a) * sizeof(uint32_t));
   });

   const uint32_t hc = alloc_pair();
   Node &h = nodes_[hc];
   Node &l = nodes_[hc + 1];
   h.bb = high_bbs[best + 1]; h.lo = (uint32_t)(lo + n_low); h.n = (uint32_t)(n - n_low);
   l.bb = low_bbs[best];      l.lo = (uint32_t)lo;            l.n = (uint32_t)n_low;
   h.depth = l.depth = (uint16_t)std::min<int>(nodes_[ni].depth + 1, 65535);
   nodes_[ni].high = (int32_t)hc;
   return true;
}

void Builder::run(HostBvh &out) {
   const bool trace = std::getenv("PTX_BUILD_TRACE") != nullptr;
   auto t_last = std::chrono::steady_clock::now();
   auto mark = [&](const char *what) {
      if (!trace) return;
      const auto t = std::chrono::steady_clock::now();
      std::fprintf(stderr, "[bvh] %s %.3f s\n", what, std::chrono::duration<double>(t - t_last).count());
      t_last = t;
   };
   size_t total = 0;
   for (const auto &m : meshes_) total += m.nf;
   info_.resize(total);
   mesh_of_.resize(total);
   face_of_.resize(total);
   idx_.resize(total);
   tmp_.resize(total);
   bucket_.resize(total);
   nodes_.assign(total ? 2 * total : 1, Node{});

   // per-face world boxes and centroids (bvh.cpp:53-121)
   {
      size_t base = 0;
      std::vector<size_t> bases;
      for (const auto &m : meshes_) { bases.push_back(base); base += m.nf; }
      for (size_t mi = 0; mi < meshes_.size(); mi++) {
         const WorldMesh &m = meshes_[mi];
         const unsigned nch = m.nf > kParallelNode ? pool_.size() : 1;
         pool_.chunks(m.nf, nch, [&](unsigned, size_t a, size_t b) {
            for (size_t f = a; f < b; f++) {
               const float *v = &m.wv[9 * f];
               FaceInfo fi;
               for (int k = 0; k < 3; k++) {
                  float mn = v[k] < v[3 + k] ? v[k] : v[3 + k];
                  mn = mn < v[6 + k] ? mn : v[6 + k];
                  float mx = v[k] > v[3 + k] ? v[k] : v[3 + k];
                  mx = mx > v[6 + k] ? mx : v[6 + k];
                  fi.bb.lo[k] = mn;
                  fi.bb.hi[k] = mx;
               }
               for (int k = 0; k < 3; k++) fi.c[k] = (fi.bb.lo[k] + fi.bb.hi[k]) * .5f;
               const size_t g = bases[mi] + f;
               info_[g] = fi;
               mesh_of_[g] = (uint32_t)mi;
               face_of_[g] = (uint32_t)f;
               idx_[g] = (uint32_t)g;
            }
         });
      }
   }
   Box root = {{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}};
   for (size_t g = 0; g < total; g++) {           // root box, in mesh-major order (bvh.cpp:100-106)
      for (int k = 0; k < 3; k++) {
         root.lo[k] = root.lo[k] < info_[g].bb.lo[k] ? root.lo[k] : info_[g].bb.lo[k];
         root.hi[k] = root.hi[k] > info_[g].bb.hi[k] ? root.hi[k] : info_[g].bb.hi[k];
      }
   }
   nodes_[0].bb = root;
   nodes_[0].lo = 0;
   nodes_[0].n = (uint32_t)total;
   mark("face boxes");

   // Phase 1: breadth-first over the nodes too large to be one worker's subtree,
   // each split with its per-face passes on all threads.
   const size_t task_limit = std::max<size_t>(kTaskNode, total / (size_t(pool_.size()) * 16));
   std::vector<uint32_t> small;
   {
      std::vector<uint32_t> frontier = {0};
      while (!frontier.empty()) {
         std::vector<uint32_t> next;
         for (uint32_t ni : frontier) {
            if (nodes_[ni].n <= task_limit || pool_.size() == 1) { small.push_back(ni); continue; }
            if (split(ni, nodes_[ni].n >= kParallelNode)) {
               next.push_back((uint32_t)nodes_[ni].high);
               next.push_back((uint32_t)nodes_[ni].high + 1);
            }
         }
         frontier.swap(next);
      }
   }
   // largest subtrees first for load balance
   std::sort(small.begin(), small.end(), [&](uint32_t a, uint32_t b) { return nodes_[a].n > nodes_[b].n; });
   mark("large nodes");
   // Phase 2: whole subtrees per worker.
   {
      std::atomic<size_t> cursor{0};
      auto worker = [&]() {
         std::vector<uint32_t> stack;
         while (true) {
            const size_t t = cursor.fetch_add(1);
            if (t >= small.size()) break;
            stack.push_back(small[t]);
            while (!stack.empty()) {
               const uint32_t ni = stack.back();
               stack.pop_back();
               if (split(ni, false)) {
                  stack.push_back((uint32_t)nodes_[ni].high);
                  stack.push_back((uint32_t)nodes_[ni].high + 1);
               }
            }
         }
      };
      const unsigned nt = std::max(1u, std::min<unsigned>(pool_.size(), (unsigned)small.size()));
      std::vector<std::thread> ts;
      for (unsigned t = 0; t + 1 < nt; t++) ts.emplace_back(worker);
      worker();
      for (auto &t : ts) t.join();
   }

   mark("subtrees");
   // Flatten: DFS pre-order, low child first (bvh.cpp:652-693).  Children are
   // allocated after their parent, so a reverse sweep gives subtree sizes and a
   // forward sweep gives each child's DFS index from its parent's.
   const uint32_t used = next_node_.load();
   std::vector<uint32_t> subtree(used, 1);
   for (int64_t i = (int64_t)used - 1; i >= 0; i--) {
      const Node &nd = nodes_[(size_t)i];
      if (nd.high >= 0) subtree[i] = 1 + subtree[nd.high] + subtree[nd.high + 1];
   }
   const uint32_t nn = total ? subtree[0] : 0;
   out.nodes.assign(nn, ptx_bvh_node{});
   out.pairs.assign(2 * total, 0);
   if (!total) return;
   std::vector<uint32_t> dfs(used, 0);
   for (uint32_t i = 0; i < used; i++) {
      const Node &nd = nodes_[i];
      if (nd.high < 0) continue;
      dfs[nd.high + 1] = dfs[i] + 1;                             // low child
      dfs[nd.high] = dfs[i] + 1 + subtree[nd.high + 1];         // high child
   }
   // faces per DFS position (0 for interior nodes), prefix-summed in DFS order
   std::vector<uint32_t> leaf_node(nn, 0xFFFFFFFFu);
   std::vector<uint32_t> offs(nn, 0);
   pool_.chunks(used, pool_.size(), [&](unsigned, size_t a, size_t b) {
      for (size_t i = a; i < b; i++) {
         const Node &nd = nodes_[i];
         ptx_bvh_node &cn = out.nodes[dfs[i]];
         for (int k = 0; k < 3; k++) { cn.p0[k] = nd.bb.lo[k]; cn.p1[k] = nd.bb.hi[k]; }
         cn.flags = nd.axis;
         if (nd.high < 0) {
            cn.num_faces = (uint16_t)nd.n;
            leaf_node[dfs[i]] = (uint32_t)i;
         } else {
            cn.offset = subtree[nd.high + 1] + 1;
            cn.num_faces = 0;
         }
      }
   });
   {
      uint64_t acc = 0;
      for (uint32_t d = 0; d < nn; d++) {
         offs[d] = (uint32_t)acc;
         if (leaf_node[d] != 0xFFFFFFFFu) acc += nodes_[leaf_node[d]].n;
      }
   }
   struct Item { uint32_t node, dfs; };
   std::vector<Item> order;           // leaves in DFS order
   order.reserve(used / 2 + 1);
   for (uint32_t d = 0; d < nn; d++)
      if (leaf_node[d] != 0xFFFFFFFFu) order.push_back({leaf_node[d], d});
   mark("dfs order");
   // pairs sorted with the reference's comparator (bvh.cpp:666-671) by the same std::sort
   const unsigned nch = order.size() > 4096 ? pool_.size() : 1;
   pool_.chunks(order.size(), nch, [&](unsigned, size_t a, size_t b) {
      std::vector<std::pair<unsigned int, unsigned int>> leaf;
      for (size_t i = a; i < b; i++) {
         const Node &nd = nodes_[order[i].node];
         const uint32_t off = offs[order[i].dfs];
         out.nodes[order[i].dfs].offset = off;
         leaf.clear();
         for (uint32_t k = 0; k < nd.n; k++) {
            const uint32_t g = idx_[nd.lo + k];
            leaf.emplace_back(mesh_of_[g], face_of_[g]);
         }
         std::sort(leaf.begin(), leaf.end(),
                   [](const std::pair<unsigned int, unsigned int> &lhs,
                      const std::pair<unsigned int, unsigned int> &rhs) {
                      if (lhs.first == rhs.first) return lhs.second < rhs.second;
                      return lhs.first < lhs.second;
                   });
         for (uint32_t k = 0; k < nd.n; k++) {
            out.pairs[2 * (size_t(off) + k)] = leaf[k].first;
            out.pairs[2 * (size_t(off) + k) + 1] = leaf[k].second;
         }
      }
   });
   mark("leaves");
}

}  // namespace

void build_reference_bvh(const std::vector<WorldMesh> &meshes, unsigned threads, HostBvh &out) {
   if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
   Builder b(meshes, threads);
   b.run(out);
}

}  // namespace ptx
