// cg_lbvh.h — building a BLAS on the device: a linear BVH (Karras, "Maximizing
// Parallelism in the Construction of BVHs, Octrees, and k-d Trees", HPG 2012)
// over 30-bit Morton codes of the triangle box centres, bottom-up boxes and
// collapse prices, then a top-down collapse into 8-wide nodes (the
// SAH-optimal collapse the host build uses: each node's children follow the
// prices' choices; subtrees of at most CG_LEAF_MAX triangles may become
// leaves, which are contiguous ranges of the sorted triangles) quantised as
// the host build's nodes are.
//
// Each step is a per-thread function over plain arrays: the kernels in
// cg_device.hip run them one thread per item, and lbvh_build_host runs the
// same functions in loops, so the layout check on the host covers the exact
// code the device runs.  The device sorts with a stable radix sort and the
// host with a stable sort on the same keys: the same order.
#ifndef CG_LBVH_H_
#define CG_LBVH_H_

#include "cg_layout.h"
#include "cg_quant.h"

namespace cg {

struct PrimBox {
  float lo[3], hi[3];
};

// The SAH-optimal collapse's price of an internal node n (Ylitie, Karras and
// Laine, HPG 2017; the host build's CollapseWideSah): cost[i - 1] is the least
// cost of n's subtree given as at most i children of the wide node above, a
// node visit costing one step and a leaf one step per two triangles, each
// weighted by its box's half area.  kl/kr: the two subtrees' budgets (both 0:
// n itself, a wide node or, with `leaf`, one leaf); il/ir: their budgets
// when n is a wide node.
struct Price {
  float cost[CG_WIDE];
  uint8_t kl[CG_WIDE], kr[CG_WIDE];
  uint8_t il, ir, leaf, pad;
};

// Child encoding in the binary tree: internal node index, or leaf (sorted
// triangle position) with the top bit set.
constexpr uint32_t kLbvhLeaf = 0x80000000u;
constexpr uint32_t kLbvhNone = 0xFFFFFFFFu;

struct LbvhArrays {
  uint32_t n;                 // triangles, >= 2
  const PrimBox* prims;       // unsorted triangle boxes (unpadded)
  float pad;                  // added to every triangle box
  float cmin[3], cscale[3];   // centre -> [0, 1024) per axis
  uint32_t* keys;             // n Morton codes (sorted in place by the driver)
  uint32_t* ids;              // n triangle ids (sorted with the keys)
  uint32_t* left;             // n - 1 internal nodes: children
  uint32_t* right;
  uint32_t* first;            // first sorted triangle of the node's range
  uint32_t* count;            // triangles in the node's range
  uint32_t* parent;           // n - 1 internal + n leaves (leaf i at n - 1 + i)
  PrimBox* box;               // n - 1 internal node boxes (padded)
  uint32_t* flag;             // n - 1 arrival counters, zero before the bounds pass
  Price* price;               // n - 1 collapse prices (null: boxes only, as a refit)
  // collapse
  uint2* front_in;            // (binary internal node, wide node) pairs
  uint2* front_out;
  uint32_t* front_out_n;
  uint32_t* wide_n;           // wide nodes allocated so far
  uint32_t* max_children;     // most children of a node built in this pass
  Node8Q* nodes;              // wide nodes, local indices
  Node8* exact;               // nullable: the unquantised nodes too (host checks)
  uint32_t* kids;             // nullable: per wide node its 8 binary children (kLbvhNone: empty), for refits
  uint32_t node_base;         // global index of local wide node 0
  uint32_t slot_base;         // global triangle slot of sorted triangle 0
};

CG_QHD uint32_t lbvh_clz(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return x == 0u ? 32u : static_cast<uint32_t>(__clz(static_cast<int>(x)));
#else
  return x == 0u ? 32u : static_cast<uint32_t>(__builtin_clz(x));
#endif
}
CG_QHD uint32_t spread3(uint32_t v) {  // 10 bits -> every third bit of 30
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// Step 1 (per triangle i): Morton code of the box centre.
CG_QHD void lbvh_morton(const LbvhArrays& A, uint32_t i) {
  const PrimBox& p = A.prims[i];
  uint32_t q[3];
  for (int a = 0; a < 3; ++a) {
    const float c = 0.5f * (p.lo[a] + p.hi[a]);
    const float x = fminf(fmaxf((c - A.cmin[a]) * A.cscale[a], 0.0f), 1023.0f);
    q[a] = static_cast<uint32_t>(x);
  }
  A.keys[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
  A.ids[i] = i;
}

// Common-prefix length of sorted keys i and j (-1 outside the array; equal
// keys are told apart by their positions).
CG_QHD int lbvh_delta(const LbvhArrays& A, int i, int j) {
  if (j < 0 || j >= static_cast<int>(A.n)) return -1;
  const uint32_t ki = A.keys[i], kj = A.keys[j];
  if (ki == kj) return 32 + static_cast<int>(lbvh_clz(static_cast<uint32_t>(i) ^ static_cast<uint32_t>(j)));
  return static_cast<int>(lbvh_clz(ki ^ kj));
}

// Step 2 (per internal node i < n - 1): its range, split and children.
CG_QHD void lbvh_node(const LbvhArrays& A, uint32_t iu) {
  const int i = static_cast<int>(iu);
  const int d = (lbvh_delta(A, i, i + 1) - lbvh_delta(A, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = lbvh_delta(A, i, i - d);
  int lmax = 2;
  while (lbvh_delta(A, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (lbvh_delta(A, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = lbvh_delta(A, i, j);
  int s = 0;
  for (int div = 2;; div *= 2) {
    const int t = (l + div - 1) / div;
    if (lbvh_delta(A, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  const int gamma = i + s * d + (d < 0 ? -1 : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const uint32_t lc = lo == gamma ? (kLbvhLeaf | static_cast<uint32_t>(gamma)) : static_cast<uint32_t>(gamma);
  const uint32_t rc = hi == gamma + 1 ? (kLbvhLeaf | static_cast<uint32_t>(gamma + 1)) : static_cast<uint32_t>(gamma + 1);
  A.left[i] = lc;
  A.right[i] = rc;
  A.first[i] = static_cast<uint32_t>(lo);
  A.count[i] = static_cast<uint32_t>(hi - lo + 1);
  const uint32_t leaf0 = A.n - 1u;
  A.parent[(lc & kLbvhLeaf) ? leaf0 + (lc & ~kLbvhLeaf) : lc] = iu;
  A.parent[(rc & kLbvhLeaf) ? leaf0 + (rc & ~kLbvhLeaf) : rc] = iu;
}

CG_QHD PrimBox lbvh_leaf_box(const LbvhArrays& A, uint32_t k) {
  const PrimBox& p = A.prims[A.ids[k]];
  PrimBox b;
  for (int a = 0; a < 3; ++a) {
    b.lo[a] = p.lo[a] - A.pad;
    b.hi[a] = p.hi[a] + A.pad;
  }
  return b;
}
CG_QHD PrimBox lbvh_child_box(const LbvhArrays& A, uint32_t c) {
  return (c & kLbvhLeaf) ? lbvh_leaf_box(A, c & ~kLbvhLeaf) : A.box[c];
}
CG_QHD PrimBox lbvh_union(const PrimBox& x, const PrimBox& y) {
  PrimBox b;
  for (int a = 0; a < 3; ++a) {
    b.lo[a] = fminf(x.lo[a], y.lo[a]);
    b.hi[a] = fmaxf(x.hi[a], y.hi[a]);
  }
  return b;
}

// Step 3: bounds, from every leaf towards the root; the second arrival at a
// node computes its box.  The atomic increment orders the first arrival's
// box store before the second's loads (acquire-release, device scope).
CG_QHD uint32_t lbvh_arrive(uint32_t* flag) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __hip_atomic_fetch_add(flag, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
#else
  return (*flag)++;
#endif
}
CG_QHD float lbvh_half_area(const PrimBox& b) {
  const float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
  return x * y + y * z + z * x;
}
CG_QHD float lbvh_child_price(const LbvhArrays& A, uint32_t c, int budget) {
  if (c & kLbvhLeaf) return lbvh_half_area(lbvh_leaf_box(A, c & ~kLbvhLeaf));  // one triangle: one step
  return A.price[c].cost[budget - 1];
}
// The price of internal node p, from its box and its children's prices.
CG_QHD void lbvh_price(const LbvhArrays& A, uint32_t p) {
  const uint32_t l = A.left[p], r = A.right[p];
  float dist[CG_WIDE + 1];
  uint8_t dl[CG_WIDE + 1], dr[CG_WIDE + 1];
  for (int j = 2; j <= CG_WIDE; ++j) {
    dist[j] = __builtin_inff();
    dl[j] = 1u;  // a valid split even when every cost overflows
    dr[j] = static_cast<uint8_t>(j - 1);
    for (int k = 1; k < j; ++k) {
      const float c = lbvh_child_price(A, l, k) + lbvh_child_price(A, r, j - k);
      if (c < dist[j]) {
        dist[j] = c;
        dl[j] = static_cast<uint8_t>(k);
        dr[j] = static_cast<uint8_t>(j - k);
      }
    }
  }
  Price x;
  const float area = lbvh_half_area(A.box[p]);
  const uint32_t cnt = A.count[p];
  const float internal = area + dist[CG_WIDE];
  const bool can_leaf = cnt <= CG_LEAF_MAX;
  const float leaf = can_leaf ? area * static_cast<float>((cnt + 1u) / 2u) : __builtin_inff();
  x.il = dl[CG_WIDE];
  x.ir = dr[CG_WIDE];
  x.leaf = can_leaf && leaf <= internal ? 1u : 0u;
  x.pad = 0u;
  x.cost[0] = fminf(internal, leaf);
  x.kl[0] = x.kr[0] = 0u;
  for (int i = 2; i <= CG_WIDE; ++i) {
    if (dist[i] < x.cost[i - 2]) {
      x.cost[i - 1] = dist[i];
      x.kl[i - 1] = dl[i];
      x.kr[i - 1] = dr[i];
    } else {
      x.cost[i - 1] = x.cost[i - 2];
      x.kl[i - 1] = x.kl[i - 2];
      x.kr[i - 1] = x.kr[i - 2];
    }
  }
  A.price[p] = x;
}
CG_QHD void lbvh_bounds(const LbvhArrays& A, uint32_t leaf) {
  uint32_t p = A.parent[A.n - 1u + leaf];
  // a Karras tree is at most 62 levels deep (30 key bits, 32 position bits)
  for (int level = 0; p != kLbvhNone && level < 64; ++level) {
    if (lbvh_arrive(&A.flag[p]) == 0u) return;
    // the release half of the next node's arrival publishes this box and price
    A.box[p] = lbvh_union(lbvh_child_box(A, A.left[p]), lbvh_child_box(A, A.right[p]));
    if (A.price) lbvh_price(A, p);
    p = A.parent[p];
  }
}

// Step 4 (per frontier entry): one 8-wide node.  The driver sizes the node
// array for n - 1 nodes: every wide node comes from a distinct internal node.
CG_QHD uint32_t lbvh_alloc(uint32_t* counter) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicAdd(counter, 1u);
#else
  return (*counter)++;
#endif
}
CG_QHD void lbvh_collapse(const LbvhArrays& A, uint2 entry) {
  // The children: the node's two subtrees with the budgets its price chose,
  // each given as itself or split again by its own price.  The budgets on the
  // stack and the children taken sum to at most eight, so neither overflows.
  uint32_t ch[CG_WIDE];
  PrimBox bx[CG_WIDE];
  int n = 0;
  uint32_t sn[CG_WIDE];
  uint8_t sb[CG_WIDE];
  int sp = 0;
  const Price& root = A.price[entry.x];
  sn[sp] = A.right[entry.x];
  sb[sp++] = root.ir;
  sn[sp] = A.left[entry.x];
  sb[sp++] = root.il;
  while (sp > 0) {
    const uint32_t c = sn[--sp];
    const int b = sb[sp];
    if ((c & kLbvhLeaf) == 0u && A.price[c].kl[b - 1] != 0u) {
      const Price& pc = A.price[c];
      sn[sp] = A.right[c];
      sb[sp++] = pc.kr[b - 1];
      sn[sp] = A.left[c];
      sb[sp++] = pc.kl[b - 1];
      continue;
    }
    ch[n] = c;
    bx[n] = lbvh_child_box(A, c);
    ++n;
  }
  Node8 nd;
  for (int k = 0; k < CG_WIDE; ++k) {
    for (int a = 0; a < 3; ++a) {
      nd.lo[a][k] = k < n ? bx[k].lo[a] : __builtin_inff();
      nd.hi[a][k] = k < n ? bx[k].hi[a] : -__builtin_inff();
    }
    nd.ref[k] = CG_REF_NONE;
  }
  for (int k = 0; k < n; ++k) {
    const uint32_t c = ch[k];
    if (c & kLbvhLeaf) {
      nd.ref[k] = CG_MAKE_LEAF(A.slot_base + (c & ~kLbvhLeaf), 1u);
    } else if (A.price[c].leaf) {
      nd.ref[k] = CG_MAKE_LEAF(A.slot_base + A.first[c], A.count[c]);
    } else {
      const uint32_t w = lbvh_alloc(A.wide_n);
      nd.ref[k] = A.node_base + w;
      A.front_out[lbvh_alloc(A.front_out_n)] = make_uint2(c, w);
    }
  }
  A.nodes[entry.y] = quantize_node(nd);
  if (A.exact) A.exact[entry.y] = nd;
  if (A.kids)
    for (int k = 0; k < CG_WIDE; ++k) A.kids[entry.y * CG_WIDE + k] = k < n ? ch[k] : kLbvhNone;
#if defined(__HIP_DEVICE_COMPILE__)
  atomicMax(A.max_children, static_cast<uint32_t>(n));
#else
  if (static_cast<uint32_t>(n) > *A.max_children) *A.max_children = static_cast<uint32_t>(n);
#endif
}

// Refit (per wide node w): the same children and refs, boxes from the
// current binary boxes (after lbvh_bounds over new triangle boxes).  Reads
// the node's refs before writing it, so it may run in place.
CG_QHD void lbvh_refit_wide(const LbvhArrays& A, uint32_t w) {
  Node8 nd;
  for (int k = 0; k < CG_WIDE; ++k) {
    const uint32_t c = A.kids[w * CG_WIDE + k];
    nd.ref[k] = c == kLbvhNone ? CG_REF_NONE : A.nodes[w].ref[k];
    PrimBox b;
    if (c != kLbvhNone) b = lbvh_child_box(A, c);
    for (int a = 0; a < 3; ++a) {
      nd.lo[a][k] = c != kLbvhNone ? b.lo[a] : __builtin_inff();
      nd.hi[a][k] = c != kLbvhNone ? b.hi[a] : -__builtin_inff();
    }
  }
  A.nodes[w] = quantize_node(nd);
}

}  // namespace cg

#endif  // CG_LBVH_H_
