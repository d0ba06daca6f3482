// vxp_tree.hpp — the broad phase's dynamic AABB trees (src/broad_phase.zig:40-529),
// node for node: node indices, the free list, the descent costs and the
// rotations all follow the reference, because the pair order it emits
// (findOverlappingPairs: leaves in node-index order, each querying the tree
// depth first with child2 popped before child1) decides which pairs the
// 512-pair filter keeps and the order the brick-pair cache sees them.
//
// The same functions run on the host (body creation and destruction) and on
// the device (one thread per world re-inserting the leaves whose fat AABB
// moved, in slot order, as broadPhase does).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace vxd {

constexpr uint32_t kTreeNull = 0xFFFFFFFFu;  // NULL_NODE
constexpr uint32_t kTreeStack = 64;          // query stack entries per lane

struct TreeNode {  // TreeNode (src/broad_phase.zig:40-62)
  float lo[3], hi[3];
  uint32_t user;  // body id (leaves)
  uint32_t parent;  // free nodes: next free node
  uint32_t c1, c2;
  int32_t height;  // -1: on the free list
  uint32_t pad;
};
static_assert(sizeof(TreeNode) == 48, "tree node");

struct TreeHead {
  uint32_t root, free_list, len, count;  // len = nodes.items.len
};

struct Tree {
  TreeNode* n;
  TreeHead* h;
  uint32_t cap;  // node slots available
};

#define VXP_TREE_FN inline __host__ __device__

VXP_TREE_FN bool treeLeaf(const TreeNode& a) { return a.c1 == kTreeNull; }

// AABB.surfaceArea (src/math.zig:628-631)
VXP_TREE_FN float treeArea(const float lo[3], const float hi[3]) {
  const float sx = hi[0] - lo[0], sy = hi[1] - lo[1], sz = hi[2] - lo[2];
  return 2.0f * (sx * sy + sy * sz + sz * sx);
}
// AABB.merge (src/math.zig:640-645): a's components first
VXP_TREE_FN void treeMerge(const float alo[3], const float ahi[3], const float blo[3],
                           const float bhi[3], float lo[3], float hi[3]) {
  for (int k = 0; k < 3; k++) {
    lo[k] = fminf(alo[k], blo[k]);
    hi[k] = fmaxf(ahi[k], bhi[k]);
  }
}
VXP_TREE_FN bool treeOverlaps(const float alo[3], const float ahi[3], const float blo[3],
                              const float bhi[3]) {  // AABB.overlaps (src/math.zig:655-659)
  return alo[0] <= bhi[0] && ahi[0] >= blo[0] && alo[1] <= bhi[1] && ahi[1] >= blo[1] &&
         alo[2] <= bhi[2] && ahi[2] >= blo[2];
}
VXP_TREE_FN void treeSetBox(TreeNode& d, const float lo[3], const float hi[3]) {
  for (int k = 0; k < 3; k++) {
    d.lo[k] = lo[k];
    d.hi[k] = hi[k];
  }
}
VXP_TREE_FN void treeFit(Tree t, uint32_t i) {  // aabb = child1.merge(child2)
  TreeNode& a = t.n[i];
  treeMerge(t.n[a.c1].lo, t.n[a.c1].hi, t.n[a.c2].lo, t.n[a.c2].hi, a.lo, a.hi);
}
VXP_TREE_FN int32_t treeMaxH(Tree t, uint32_t a, uint32_t b) {
  return t.n[a].height > t.n[b].height ? t.n[a].height : t.n[b].height;
}

// allocateNode (src/broad_phase.zig:112-140); kTreeNull when full
VXP_TREE_FN uint32_t treeAlloc(Tree t) {
  uint32_t id;
  if (t.h->free_list != kTreeNull) {
    id = t.h->free_list;
    t.h->free_list = t.n[id].parent;
  } else {
    if (t.h->len >= t.cap) return kTreeNull;
    id = t.h->len++;
  }
  TreeNode& d = t.n[id];
  for (int k = 0; k < 3; k++) {
    d.lo[k] = INFINITY;  // AABB.empty
    d.hi[k] = -INFINITY;
  }
  d.user = 0;
  d.parent = d.c1 = d.c2 = kTreeNull;
  d.height = 0;
  t.h->count++;
  return id;
}
// freeNode (src/broad_phase.zig:142-147)
VXP_TREE_FN void treeFree(Tree t, uint32_t id) {
  t.n[id].parent = t.h->free_list;
  t.n[id].height = -1;
  t.h->free_list = id;
  t.h->count--;
}

// balance (src/broad_phase.zig:365-458): one rotation toward the taller child.
VXP_TREE_FN uint32_t treeBalance(Tree t, uint32_t ia) {
  TreeNode* N = t.n;
  if (treeLeaf(N[ia]) || N[ia].height < 2) return ia;
  const uint32_t ib = N[ia].c1, ic = N[ia].c2;
  const int32_t bal = N[ic].height - N[ib].height;
  // rotate the taller child `up` into ia's place; `keep` is ia's other child
  auto rotate = [&](uint32_t up, uint32_t keep, bool up_is_c2) -> uint32_t {
    const uint32_t f = N[up].c1, g = N[up].c2;
    N[up].c1 = ia;
    N[up].parent = N[ia].parent;
    N[ia].parent = up;
    if (N[up].parent != kTreeNull) {
      if (N[N[up].parent].c1 == ia)
        N[N[up].parent].c1 = up;
      else
        N[N[up].parent].c2 = up;
    } else {
      t.h->root = up;
    }
    // the taller grandchild stays under `up`, the other moves under ia
    const bool f_taller = N[f].height > N[g].height;
    const uint32_t stay = f_taller ? f : g, move = f_taller ? g : f;
    N[up].c2 = stay;
    if (up_is_c2)
      N[ia].c2 = move;
    else
      N[ia].c1 = move;
    N[move].parent = ia;
    treeMerge(N[keep].lo, N[keep].hi, N[move].lo, N[move].hi, N[ia].lo, N[ia].hi);
    treeMerge(N[ia].lo, N[ia].hi, N[stay].lo, N[stay].hi, N[up].lo, N[up].hi);
    N[ia].height = 1 + treeMaxH(t, keep, move);
    N[up].height = 1 + treeMaxH(t, ia, stay);
    return up;
  };
  if (bal > 1) return rotate(ic, ib, true);
  if (bal < -1) return rotate(ib, ic, false);
  return ia;
}

// Rebalance and refit from `i` up to the root (the walks that end insertLeaf
// and removeLeaf; they refit box and height in either order, to the same result).
VXP_TREE_FN void treeWalkUp(Tree t, uint32_t i) {
  while (i != kTreeNull) {
    i = treeBalance(t, i);
    const uint32_t a = t.n[i].c1, b = t.n[i].c2;
    t.n[i].height = 1 + treeMaxH(t, a, b);
    treeFit(t, i);
    i = t.n[i].parent;
  }
}

// insertLeaf (src/broad_phase.zig:215-314)
VXP_TREE_FN void treeInsertLeaf(Tree t, uint32_t leaf) {
  TreeNode* N = t.n;
  if (t.h->root == kTreeNull) {
    t.h->root = leaf;
    N[leaf].parent = kTreeNull;
    return;
  }
  float llo[3], lhi[3];
  for (int k = 0; k < 3; k++) {
    llo[k] = N[leaf].lo[k];
    lhi[k] = N[leaf].hi[k];
  }
  uint32_t index = t.h->root;
  while (!treeLeaf(N[index])) {
    const uint32_t a = N[index].c1, b = N[index].c2;
    const float area = treeArea(N[index].lo, N[index].hi);
    float clo[3], chi[3];
    treeMerge(N[index].lo, N[index].hi, llo, lhi, clo, chi);
    const float carea = treeArea(clo, chi);
    const float cost = 2.0f * carea;
    const float inherit = 2.0f * (carea - area);
    float cost_c[2];
    const uint32_t ch[2] = {a, b};
    for (int j = 0; j < 2; j++) {
      float mlo[3], mhi[3];
      treeMerge(llo, lhi, N[ch[j]].lo, N[ch[j]].hi, mlo, mhi);
      if (treeLeaf(N[ch[j]]))
        cost_c[j] = treeArea(mlo, mhi) + inherit;
      else
        cost_c[j] = (treeArea(mlo, mhi) - treeArea(N[ch[j]].lo, N[ch[j]].hi)) + inherit;
    }
    if (cost < cost_c[0] && cost < cost_c[1]) break;
    index = cost_c[0] < cost_c[1] ? a : b;
  }
  const uint32_t sib = index, old_parent = N[sib].parent;
  const uint32_t np = treeAlloc(t);
  if (np == kTreeNull) return;  // allocateNode failure: the leaf stays detached
  N[np].parent = old_parent;
  N[np].user = 0;
  treeMerge(llo, lhi, N[sib].lo, N[sib].hi, N[np].lo, N[np].hi);
  N[np].height = N[sib].height + 1;
  if (old_parent != kTreeNull) {
    if (N[old_parent].c1 == sib)
      N[old_parent].c1 = np;
    else
      N[old_parent].c2 = np;
  } else {
    t.h->root = np;
  }
  N[np].c1 = sib;
  N[np].c2 = leaf;
  N[sib].parent = np;
  N[leaf].parent = np;
  treeWalkUp(t, N[leaf].parent);
}

// removeLeaf (src/broad_phase.zig:317-362)
VXP_TREE_FN void treeRemoveLeaf(Tree t, uint32_t leaf) {
  TreeNode* N = t.n;
  if (leaf == t.h->root) {
    t.h->root = kTreeNull;
    return;
  }
  const uint32_t parent = N[leaf].parent, grand = N[parent].parent;
  const uint32_t sib = N[parent].c1 == leaf ? N[parent].c2 : N[parent].c1;
  if (grand != kTreeNull) {
    if (N[grand].c1 == parent)
      N[grand].c1 = sib;
    else
      N[grand].c2 = sib;
    N[sib].parent = grand;
    treeFree(t, parent);
    treeWalkUp(t, grand);
  } else {
    t.h->root = sib;
    N[sib].parent = kTreeNull;
    treeFree(t, parent);
  }
}

// insertBody (src/broad_phase.zig:150-163) with the fat box already inflated;
// returns the leaf's node, kTreeNull when the node pool is full.
VXP_TREE_FN uint32_t treeInsertBody(Tree t, uint32_t body_id, const float lo[3], const float hi[3]) {
  const uint32_t id = treeAlloc(t);
  if (id == kTreeNull) return kTreeNull;
  treeSetBox(t.n[id], lo, hi);
  t.n[id].user = body_id;
  t.n[id].height = 0;
  treeInsertLeaf(t, id);
  return id;
}
// removeBody (src/broad_phase.zig:166-172)
VXP_TREE_FN void treeRemoveBody(Tree t, uint32_t leaf) {
  treeRemoveLeaf(t, leaf);
  treeFree(t, leaf);
}
// updateBody's re-insertion (src/broad_phase.zig:190-212) with the new box
VXP_TREE_FN void treeMoveLeaf(Tree t, uint32_t leaf, const float lo[3], const float hi[3]) {
  treeRemoveLeaf(t, leaf);
  treeSetBox(t.n[leaf], lo, hi);
  treeInsertLeaf(t, leaf);
}

// AABBTree.query (src/broad_phase.zig:461-485) for the leaf `q` of `src`
// against tree `t`: visits overlapping leaves in the reference's order and
// calls emit(body id).  Returns false when the stack overflows.
template <class F>
VXP_TREE_FN bool treeQuery(const TreeNode* N, uint32_t root, const float lo[3], const float hi[3],
                           F&& emit) {
  if (root == kTreeNull) return true;
  uint32_t st[kTreeStack];
  uint32_t sp = 0;
  st[sp++] = root;
  while (sp) {
    const uint32_t id = st[--sp];
    if (id == kTreeNull) continue;
    const TreeNode& d = N[id];
    if (!treeOverlaps(d.lo, d.hi, lo, hi)) continue;
    if (treeLeaf(d)) {
      emit(d.user);
    } else {
      if (sp + 2 > kTreeStack) return false;
      st[sp++] = d.c1;
      st[sp++] = d.c2;
    }
  }
  return true;
}

}  // namespace vxd
