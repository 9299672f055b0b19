// Host-side adaptive octree for the FMM plan (C++17).
//
// Semantics follow the reference partitioner exactly (icyspirit/SimpleFMM
// box.hpp:84-157, partition.hpp:46-103,168-366, tree.hpp:24-174, zorder.hpp:26-58):
//   * the bounding cube and its quirks (width grows upward in z only, 1e-3 margin),
//   * cell codes clamp(int(u*2^l)) interleaved as Morton bits (x bit0, y bit1, z bit2),
//   * a node subdivides while level < max_level and count > cap,
//   * nodes of a level are kept in code order, leaves are nodes that hold particles,
//   * same-level interaction lists from level 2, coarse-leaf + descendant near lists.
// The construction itself is different: one Morton sort at max_level replaces the
// per-level recursive sorts (level-l codes are prefixes of level-L codes), and the
// interaction and near lists come from a same-level neighbour table propagated
// top-down (a node's neighbours are children of its parent's neighbours) instead
// of per-candidate binary searches.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sfmm {

struct BoxGeom {
    int X = 0, Y = 1, Z = 2;  // axis permutation for vertical_axis
    double phi = 0;           // rotation about the vertical axis (radians)
    double cphi = 1, sphi = 0;
    double width = 1;
    double corner[3] = {0, 0, 0};

    void rotate(const double* v, double* r) const;
    void centre(int level, int64_t code, double* out) const;
};

// Box(positions, vertical_axis, phideg), box.hpp:84-107
BoxGeom make_box(int64_t n, const double* pos, int vertical_axis, double phideg);
// Box(corner, width, vertical_axis, phideg), box.hpp:73-82
BoxGeom make_box_explicit(const double* corner, double width, int vertical_axis, double phideg);

int32_t morton_axis(int64_t code, int axis);
int64_t morton_code(int32_t i, int32_t j, int32_t k);

struct TreeLevel {
    std::vector<int64_t> code;         // node codes, ascending (= reference c_node)
    std::vector<int64_t> pbegin, pend;  // particle range of the node's subtree (sorted order)
    std::vector<int32_t> parent;       // index at level-1 (-1 at level 0)
    std::vector<int32_t> child0;       // first child index at level+1 (-1 if leaf)
    std::vector<int8_t> nchild;        // number of children
    std::vector<int32_t> node2leaf;    // leaf index at this level or -1
    std::vector<int32_t> leaf2node;    // reference _i_leaf
    // same-level neighbours (3x3x3 window, existing nodes, code order, self included)
    std::vector<int64_t> nb_row;
    std::vector<int32_t> nb;
    // interaction list (levels >= 2): sources sorted by index, offset id
    // ((dk+3)*7 + (dj+3))*7 + (di+3) with d = source - target
    std::vector<int64_t> il_row;
    std::vector<int32_t> il_src;
    std::vector<int16_t> il_off;

    int32_t n_node() const { return static_cast<int32_t>(code.size()); }
    int32_t n_leaf() const { return static_cast<int32_t>(leaf2node.size()); }
    int32_t find(int64_t c) const;  // node index of code c, or -1
};

struct Tree {
    int64_t n = 0;
    int max_level = 20;
    int cap = 64;
    int depth = 0;
    BoxGeom box;
    std::vector<int32_t> perm;  // sorted slot -> original particle index
    std::vector<TreeLevel> lv;
    // leaves over all levels, in (level, leaf) order
    std::vector<int32_t> leaf_level, leaf_node;
    std::vector<int64_t> leaf_row;  // first global leaf id of each level (size depth+2)
    // near-field sources of each global leaf as merged particle ranges
    std::vector<int64_t> near_row;
    std::vector<int64_t> near_begin, near_end;
};

// Builds the tree; returns an empty string or an error message.
std::string build_tree(Tree& t, int64_t n, const double* pos, const BoxGeom& box, int max_level, int cap,
                       int threads);

// The same tree built on the current GPU (codes, sort, levels, leaves, lists and
// near ranges in HBM) and copied into `t`; every array equals build_tree's.
std::string build_tree_device(Tree& t, int64_t n, const double* pos, const BoxGeom& box, int max_level, int cap);

// Reference-format lists of one level (partition.hpp:155-166), for parity checks.
struct RefLists {
    std::vector<int64_t> c_node, leaf2node, idx_row, idx_val, ilist_row, ilist_val, clist_row, clist_val, plist,
        nlist_row, nlist_val;
};
void reference_lists(const Tree& t, int level, RefLists& out);

// Pair counts used for the algorithmic work figures.
struct TreeStats {
    int64_t n_node = 0, n_leaf = 0, m2l_pairs = 0, near_pairs = 0, m2m_edges = 0, far_leaf_particles = 0;
};
TreeStats tree_stats(const Tree& t);

}  // namespace sfmm
