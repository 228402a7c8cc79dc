// SPDX-License-Identifier: CeCILL-2.1
//
// tree_layout.hpp -- 64-byte node records of the compressed-leaf BVH (one record per cell,
// internal cells first, then leaves, as in the reference's total cell numbering
// KarrasTreeTraverser.hpp:55-66).
//   [0..2] aabb min  [3..5] aabb max  [6] rint = max h * htol
//   [7] internal: (left total id, right total id) ; leaf: (first, end) in sorted order
#pragma once

#include "common.hpp"

namespace shamx {

    constexpr int NODE_F64  = 8;
    constexpr u32 NO_PARENT = 0xffffffffu;

    __host__ __device__ inline void node_pack(f64 *nodes, u32 node, u32 a, u32 b) {
        u32 *w = (u32 *) (nodes + (u64) node * NODE_F64 + 7);
        w[0]   = a;
        w[1]   = b;
    }
    __host__ __device__ inline void node_set_children(f64 *nodes, u32 node, u32 l, u32 r) {
        node_pack(nodes, node, l, r);
    }
    __host__ __device__ inline void node_set_range(f64 *nodes, u32 node, u32 s0, u32 s1) {
        node_pack(nodes, node, s0, s1);
    }
    __host__ __device__ inline void node_children(const f64 *nodes, u32 node, u32 &l, u32 &r) {
        const u32 *w = (const u32 *) (nodes + (u64) node * NODE_F64 + 7);
        l            = w[0];
        r            = w[1];
    }

} // namespace shamx
