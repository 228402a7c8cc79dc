// SPDX-License-Identifier: CeCILL-2.1
//
// neigh.hip -- two-stage neighbour cache (NeighbourCache::start_neighbors_cache_2stages,
// shammodels/sph/src/modules/NeighbourCache.cpp:222-604).
//   stage 1: per leaf, the list of interacting leaves (tree walk, :276-404)
//   owner  : per local particle, the LAST leaf in depth-first order whose box contains it
//            (inclusive max, :406-451)
//   stage 2: per local particle, candidates of the owner's neighbour leaves kept when
//            r_ab^2 <= max(rint_a, rint_b)^2 R^2 with rint = h * htol (:465-590)
// Each list is built with a count pass, an exclusive scan and a fill pass that walks the same
// candidates in the same order, so lists (and their order) are deterministic and identical to
// the reference algorithm.  Traversal: KarrasTreeTraverser.hpp:69-118 (left child first).
#include "launch.hpp"
#include "tree_layout.hpp"

namespace shamx {

    namespace {
        constexpr int STACK = 40; // Karras tree on unique 30-bit codes has depth <= 31

        struct Box {
            f64 lo[3], hi[3];
        };
        __device__ inline Box node_box(const f64 *__restrict rec) {
            return Box{{rec[0], rec[1], rec[2]}, {rec[3], rec[4], rec[5]}};
        }
        // BBAA::cella_neigh_b<f64_3> (geometry_utils.hpp:125-135)
        __device__ inline bool cella_neigh_b(const f64 *alo, const f64 *ahi, const f64 *blo,
                                             const f64 *bhi) {
            return (fmax(alo[0], blo[0]) <= fmin(ahi[0], bhi[0])) &&
                   (fmax(alo[1], blo[1]) <= fmin(ahi[1], bhi[1])) &&
                   (fmax(alo[2], blo[2]) <= fmin(ahi[2], bhi[2]));
        }

        template<class Crit, class Found>
        __device__ inline bool traverse(const f64 *__restrict nodes, u32 I, u32 ncell, Crit &&crit,
                                        Found &&found) {
            u32 stack[STACK];
            int sp      = 0;
            stack[sp++] = 0;
            while (sp > 0) {
                u32 node       = stack[--sp];
                const f64 *rec = nodes + (u64) node * NODE_F64;
                if (crit(rec)) {
                    if (node >= I) {
                        found(node);
                    } else {
                        if (sp + 2 > STACK) return false;
                        u32 l, r;
                        node_children(nodes, node, l, r);
                        if (l >= ncell || r >= ncell) return false;
                        stack[sp++] = r;
                        stack[sp++] = l;
                    }
                }
            }
            return true;
        }

        // stage 1 count / fill
        template<bool FILL>
        __global__ void k_leaf_neigh(u32 L, u32 I, const f64 *__restrict nodes, f64 Rkern,
                                     u32 *__restrict cnt, const u32 *__restrict offset,
                                     u32 *__restrict ids, u32 *__restrict err) {
            u32 g = blockIdx.x * blockDim.x + threadIdx.x;
            if (g >= L) return;
            const f64 *ra = nodes + (u64) (I + g) * NODE_F64;
            f64 rint_a    = ra[6] * Rkern;
            f64 alo[3] = {ra[0], ra[1], ra[2]}, ahi[3] = {ra[3], ra[4], ra[5]};
            f64 elo[3] = {alo[0] - rint_a, alo[1] - rint_a, alo[2] - rint_a};
            f64 ehi[3] = {ahi[0] + rint_a, ahi[1] + rint_a, ahi[2] + rint_a};
            u32 c      = FILL ? offset[g] : 0u;
            bool ok    = traverse(
                nodes, I, I + L,
                [&](const f64 *rec) {
                    f64 r      = rec[6] * Rkern;
                    f64 nlo[3] = {rec[0], rec[1], rec[2]}, nhi[3] = {rec[3], rec[4], rec[5]};
                    f64 xlo[3] = {nlo[0] - r, nlo[1] - r, nlo[2] - r};
                    f64 xhi[3] = {nhi[0] + r, nhi[1] + r, nhi[2] + r};
                    return cella_neigh_b(alo, ahi, xlo, xhi) || cella_neigh_b(elo, ehi, nlo, nhi);
                },
                [&](u32 node) {
                    if (FILL) ids[c] = node;
                    c++;
                });
            if (!ok) atomicOr(err, 1u);
            if (!FILL) cnt[g] = c;
        }

        __global__ void k_owner(u32 n, u32 I, const f64 *__restrict nodes,
                                const f64 *__restrict x, const f64 *__restrict y,
                                const f64 *__restrict z, u32 L, u32 *__restrict owner,
                                u32 *__restrict err) {
            u32 a = blockIdx.x * blockDim.x + threadIdx.x;
            if (a >= n) return;
            f64 p[3]  = {x[a], y[a], z[a]};
            u32 found = 0x7fffffffu;
            bool ok   = traverse(
                nodes, I, I + L,
                [&](const f64 *rec) { // is_coord_in_range_incl_max (geometry_utils.hpp:74-81)
                    return (rec[0] <= p[0]) && (p[0] <= rec[3]) && (rec[1] <= p[1]) &&
                           (p[1] <= rec[4]) && (rec[2] <= p[2]) && (p[2] <= rec[5]);
                },
                [&](u32 node) { found = node - I; });
            if (!ok) atomicOr(err, 1u);
            if (found >= L) atomicOr(err, 2u);
            owner[a] = (found < L) ? found : 0u;
        }

        // stage 2 count / fill
        template<bool FILL>
        __global__ void k_part_neigh(u32 n, u32 I, const f64 *__restrict nodes,
                                     const u32 *__restrict sort_index,
                                     const u32 *__restrict owner,
                                     const u32 *__restrict leaf_cnt,
                                     const u32 *__restrict leaf_offset,
                                     const u32 *__restrict leaf_ids, const f64 *__restrict x,
                                     const f64 *__restrict y, const f64 *__restrict z,
                                     const f64 *__restrict h, f64 htol, f64 Rker2,
                                     u32 *__restrict cnt, const u32 *__restrict offset,
                                     u32 *__restrict ids) {
            u32 a = blockIdx.x * blockDim.x + threadIdx.x;
            if (a >= n) return;
            f64 rint_a = h[a] * htol;
            f64 xa = x[a], ya = y[a], za = z[a];
            f64 ca = rint_a * rint_a * Rker2;
            u32 lo = owner[a];
            u32 k0 = leaf_offset[lo], k1 = k0 + leaf_cnt[lo];
            u32 c  = FILL ? offset[a] : 0u;
            for (u32 k = k0; k < k1; k++) {
                u32 s0, s1;
                node_children(nodes, leaf_ids[k], s0, s1); // leaf record: sorted range
                for (u32 s = s0; s < s1; s++) {
                    u32 b    = sort_index[s];
                    f64 dx   = xa - x[b], dy = ya - y[b], dz = za - z[b];
                    f64 rab2 = dx * dx + dy * dy + dz * dz;
                    f64 rb   = h[b] * htol;
                    bool no_interact = rab2 > ca && rab2 > rb * rb * Rker2;
                    if (!no_interact) {
                        if (FILL) ids[c] = b;
                        c++;
                    }
                }
            }
            if (!FILL) cnt[a] = c;
        }
    } // namespace

    void neigh_build(hipStream_t s, NeighBuffers &nb, const TreeBuffers &t, u32 n_local,
                     const f64 *x, const f64 *y, const f64 *z, const f64 *h, f64 htol,
                     f64 Rkern) {
        u32 L = t.L, I = t.I;
        nb.n_local = n_local;
        nb.counters.resize(8);
        nb.scratch.resize(std::max(scan_scratch_bytes(L), scan_scratch_bytes(n_local)));
        nb.leaf_cnt.resize(L);
        nb.leaf_offset.resize(L);
        nb.owner.resize(std::max<u32>(n_local, 1));
        nb.cnt.resize(std::max<u32>(n_local, 1));
        nb.offset.resize(std::max<u32>(n_local, 1));
        u32 *err = nb.counters.data() + 4;
        fill_u32(s, nb.counters.data(), 0, 8);

        // stage 1
        k_leaf_neigh<false><<<div_up(L, 128), 128, 0, s>>>(L, I, t.nodes.data(), Rkern,
                                                           nb.leaf_cnt.data(), nullptr, nullptr,
                                                           err);
        exclusive_scan_u32(s, nb.leaf_cnt.data(), nb.leaf_offset.data(), L, nb.counters.data(),
                           nb.scratch.data());
        u32 hc[8];
        SHAMX_HIP(hipMemcpyAsync(hc, nb.counters.data(), 8 * sizeof(u32), hipMemcpyDeviceToHost, s));
        SHAMX_HIP(hipStreamSynchronize(s));
        if (hc[4]) throw std::runtime_error("neighbour cache: tree traversal stack overflow");
        nb.n_leaf_pairs = hc[0];
        nb.leaf_ids.resize(std::max<u64>(nb.n_leaf_pairs, 1));
        k_leaf_neigh<true><<<div_up(L, 128), 128, 0, s>>>(L, I, t.nodes.data(), Rkern,
                                                          nullptr, nb.leaf_offset.data(),
                                                          nb.leaf_ids.data(), err);
        if (n_local == 0) {
            nb.n_pairs = 0;
            return;
        }
        // owner leaf
        k_owner<<<div_up(n_local, 128), 128, 0, s>>>(n_local, I, t.nodes.data(), x, y, z, L,
                                                     nb.owner.data(), err);
        // stage 2
        f64 Rker2 = Rkern * Rkern;
        k_part_neigh<false><<<div_up(n_local, 128), 128, 0, s>>>(
            n_local, I, t.nodes.data(), t.sort_index.data(), nb.owner.data(), nb.leaf_cnt.data(),
            nb.leaf_offset.data(), nb.leaf_ids.data(), x, y, z, h, htol, Rker2, nb.cnt.data(),
            nullptr, nullptr);
        exclusive_scan_u32(s, nb.cnt.data(), nb.offset.data(), n_local, nb.counters.data() + 1,
                           nb.scratch.data());
        SHAMX_HIP(hipMemcpyAsync(hc, nb.counters.data(), 8 * sizeof(u32), hipMemcpyDeviceToHost, s));
        SHAMX_HIP(hipStreamSynchronize(s));
        if (hc[4] & 1u) throw std::runtime_error("neighbour cache: tree traversal stack overflow");
        if (hc[4] & 2u)
            throw std::runtime_error("neighbour cache: a particle is outside every leaf box");
        nb.n_pairs = hc[1];
        nb.ids.resize(std::max<u64>(nb.n_pairs, 1));
        k_part_neigh<true><<<div_up(n_local, 128), 128, 0, s>>>(
            n_local, I, t.nodes.data(), t.sort_index.data(), nb.owner.data(), nb.leaf_cnt.data(),
            nb.leaf_offset.data(), nb.leaf_ids.data(), x, y, z, h, htol, Rker2, nullptr,
            nb.offset.data(), nb.ids.data());
        SHAMX_HIP(hipGetLastError());
    }

} // namespace shamx
