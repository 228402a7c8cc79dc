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

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>

namespace shamx {

    namespace {
        struct Box {
            f64 lo[3], hi[3];
        };
        __device__ inline Box node_box(const f64 *__restrict rec) {
            return Box{{rec[0], rec[1], rec[2]}, {rec[3], rec[4], rec[5]}};
        }
        // BBAA::cella_neigh_b<f64_3> (geometry_utils.hpp:125-135).  The three comparisons are
        // combined with & rather than &&: the same value (they have no side effects), without
        // the branch and exec-mask updates the short-circuit form compiles to in a walk loop.
        __device__ inline bool cella_neigh_b(const f64 *alo, const f64 *ahi, const f64 *blo,
                                             const f64 *bhi) {
            return (fmax(alo[0], blo[0]) <= fmin(ahi[0], bhi[0])) &
                   (fmax(alo[1], blo[1]) <= fmin(ahi[1], bhi[1])) &
                   (fmax(alo[2], blo[2]) <= fmin(ahi[2], bhi[2]));
        }

        // orders the LDS accesses of one wave's lanes around this point (LDS executes a wave's
        // operations in issue order; the fences keep the compiler from moving them across)
        __device__ inline void wave_lds_sync() {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }

        // Depth-first traversal shared by the 64 lanes of a wave.  The stack lives in lanes
        // (entry i in lane i: node id, the mask of lanes whose walk reaches it, its two child
        // ids).  A node is tested before it is pushed, by the lanes of its parent's mask, so
        // only nodes that some lane enters are pushed, and the two children of a popped node
        // are loaded and tested together (one dependent load per tree level).  Each lane still
        // visits exactly the nodes of its own depth-first walk (KarrasTreeTraverser.hpp:69-118,
        // left child first), in the same order.  crit(rec) is evaluated per lane; found(node)
        // is called in the lanes whose walk reaches leaf node.  Returns false on stack overflow
        // or a child id out of range (wave-uniform).
        template<class Crit, class Found>
        __device__ inline bool wave_traverse(const f64 *__restrict nodes, u32 I, u32 ncell,
                                             u64 live, Crit &&crit, Found &&found) {
            const int lane = __lane_id();
            u32 st_node = 0, st_lo = 0, st_hi = 0, st_l = 0, st_r = 0; // lane i: entry i
            int sp      = 0;
            auto push   = [&](u32 node, u64 mask, u32 cl, u32 cr) { // sp is wave-uniform
                if (lane == sp) {
                    st_node = node;
                    st_lo   = (u32) mask;
                    st_hi   = (u32) (mask >> 32);
                    st_l    = cl;
                    st_r    = cr;
                }
                sp++;
            };
            // a node record (box, rint, child ids), loaded whole so that the two children's
            // loads are in flight together
            struct Rec {
                f64 v[7];
                u32 c0, c1;
            };
            auto load = [&](u32 node) {
                const f64 *p = nodes + (u64) node * NODE_F64;
                Rec r;
                for (int i = 0; i < 7; i++) r.v[i] = p[i];
                node_children(nodes, node, r.c0, r.c1);
                return r;
            };
            // lanes of mask that enter the node
            auto test = [&](const Rec &r, u64 mask) -> u64 {
                bool pass = ((mask >> lane) & 1ull) && crit(r.v);
                return __ballot(pass);
            };
            {
                Rec r0 = load(0);
                u64 pm = test(r0, live);
                if (pm) push(0, pm, r0.c0, r0.c1);
            }
            auto entry_mask = [&](int e) {
                return (u64) (u32) __builtin_amdgcn_readlane((int) st_lo, e) |
                       ((u64) (u32) __builtin_amdgcn_readlane((int) st_hi, e) << 32);
            };
            while (sp > 0) {
                u32 node = (u32) __builtin_amdgcn_readlane((int) st_node, sp - 1);
                u64 mask = entry_mask(sp - 1);
                if (node >= I) { // a leaf on top: this is its turn in every lane's walk
                    sp--;
                    if ((mask >> lane) & 1ull) found(node);
                    continue;
                }
                u32 l = (u32) __builtin_amdgcn_readlane((int) st_l, sp - 1);
                u32 r = (u32) __builtin_amdgcn_readlane((int) st_r, sp - 1);
                if (l >= ncell || r >= ncell) return false;
                // the entry below is expanded in the same step when it is internal too: its
                // children go under the top entry's, so every walk keeps its order, and four
                // child records are in flight at once
                u32 node2 = sp >= 2 ? (u32) __builtin_amdgcn_readlane((int) st_node, sp - 2) : I;
                if (node2 < I && sp + 2 <= WAVE - 8) {
                    u64 mask2 = entry_mask(sp - 2);
                    u32 l2    = (u32) __builtin_amdgcn_readlane((int) st_l, sp - 2);
                    u32 r2    = (u32) __builtin_amdgcn_readlane((int) st_r, sp - 2);
                    if (l2 >= ncell || r2 >= ncell) return false;
                    sp -= 2;
                    const Rec nl = load(l), nr = load(r), nl2 = load(l2), nr2 = load(r2);
                    u64 pl = test(nl, mask), pr = test(nr, mask);
                    u64 pl2 = test(nl2, mask2), pr2 = test(nr2, mask2);
                    if (pr2) push(r2, pr2, nr2.c0, nr2.c1);
                    if (pl2) push(l2, pl2, nl2.c0, nl2.c1);
                    if (pr) push(r, pr, nr.c0, nr.c1);
                    if (pl) push(l, pl, nl.c0, nl.c1);
                    continue;
                }
                sp--;
                if (sp + 2 > WAVE) return false;
                const Rec nl = load(l), nr = load(r);
                u64 pl = test(nl, mask);
                u64 pr = test(nr, mask);
                if (pr) push(r, pr, nr.c0, nr.c1);
                if (pl) push(l, pl, nl.c0, nl.c1);
            }
            return true;
        }

        // stage 1 count / fill, one lane per leaf, 64 consecutive leaves per wave.
        // STRIDE: single pass, leaf g owns ids[g * cap, (g + 1) * cap); the count goes on past
        // cap (entries beyond it are dropped) so the caller can detect the overflow.
        // LIST: fill only the n_list leaves of list (overflowed ones), each at its offset[g].
        template<bool FILL, bool STRIDE, bool LIST = false>
        __global__ void __launch_bounds__(256)
            k_leaf_neigh_w(u32 L, u32 I, const f64 *__restrict nodes, f64 Rkern,
                           u32 *__restrict cnt, u32 *__restrict offset, u32 *__restrict ids,
                           u32 cap, u32 *__restrict err, u32 lpw,
                           const u32 *__restrict list = nullptr, u32 n_list = 0) {
            // lpw leaves per wave (64, 32 or 16): fewer leaves give more, shorter walks
            const u32 lane = (u32) __lane_id();
            const u32 j = (blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) * lpw + lane;
            bool valid  = lane < lpw && j < (LIST ? n_list : L);
            u32 g       = LIST ? (valid ? list[j] : 0u) : j;
            u64 live    = __ballot(valid);
            if (!live) return;
            const f64 *ra = nodes + (u64) (I + (valid ? g : 0u)) * NODE_F64;
            f64 rint_a    = ra[6] * Rkern;
            f64 alo[3] = {ra[0], ra[1], ra[2]}, ahi[3] = {ra[3], ra[4], ra[5]};
            f64 elo[3] = {alo[0] - rint_a, alo[1] - rint_a, alo[2] - rint_a};
            f64 ehi[3] = {ahi[0] + rint_a, ahi[1] + rint_a, ahi[2] + rint_a};
            const u32 c0 = STRIDE ? (valid ? g : 0u) * cap : (FILL && valid ? offset[g] : 0u);
            const u32 ce = c0 + cap;
            u32 c        = c0;
            bool ok      = wave_traverse(
                nodes, I, I + L, live,
                [&](const f64 *rec) -> bool {
                    // STRIDE: a lane whose list is past its capacity stops entering cells (its
                    // count, then above cap, only flags the overflow; the true count comes
                    // from the overflow walk), so its wave is not held for the whole list
                    if (STRIDE && c - c0 > cap) return false;
                    f64 r      = rec[6] * Rkern;
                    f64 nlo[3] = {rec[0], rec[1], rec[2]}, nhi[3] = {rec[3], rec[4], rec[5]};
                    f64 xlo[3] = {nlo[0] - r, nlo[1] - r, nlo[2] - r};
                    f64 xhi[3] = {nhi[0] + r, nhi[1] + r, nhi[2] + r};
                    return cella_neigh_b(alo, ahi, xlo, xhi) | cella_neigh_b(elo, ehi, nlo, nhi);
                },
                [&](u32 node) {
                    if (FILL && (!STRIDE || c < ce)) ids[c] = node;
                    c++;
                });
            if (!ok && __lane_id() == __builtin_ctzll(live)) atomicOr(err, 1u);
            if ((!FILL || STRIDE) && valid) cnt[g] = c - c0;
            if (STRIDE && valid) offset[g] = c0;
        }

        // The lists longer than their fixed capacity, one listed leaf per block of WIDE_W
        // waves, its walk spread over the block's threads: the first wave expands the tree
        // breadth first into an ordered frontier of at least one subtree per thread (or of all
        // its accepted leaves), then each thread walks a contiguous run of frontier entries
        // depth first, left child first, with its own stack.  The frontier stays in
        // left-to-right order and every thread's walk is left-first, so the threads' outputs
        // in thread order are the leaf's list in the order of the reference's walk.
        // COUNT pass: each thread's count to tcnt[j * WIDE_T + thread] and the block's total to
        // cnt[g] (the fixed-capacity pass stops counting past the capacity).  FILL pass: the
        // same frontier again, the threads' write positions from those counts, each list
        // walked once more and written at offset[g]; the block's total must equal cnt[g]
        // (err bit 4 otherwise).  The kernel lasts as long as its longest walk, so a block per
        // leaf spreads that walk over several waves: 8 for the few long lists of a settled
        // build, 1 when many leaves pass a capacity set too low (the first build), whose lists
        // are only a little longer.
        constexpr int WIDE_ST = 40;
        template<int WIDE_W, bool COUNT, int WIDE_T = WIDE_W * WAVE, int WIDE_FR = 4 * WIDE_T>
        __global__ void __launch_bounds__(WIDE_T)
            k_leaf_walk_wide(u32 L, u32 I, const f64 *__restrict nodes, f64 Rkern,
                             const u32 *__restrict list, u32 n_list, u32 *__restrict cnt,
                             u32 *__restrict tcnt, const u32 *__restrict offset,
                             u32 *__restrict ids, u32 *__restrict err) {
            __shared__ u32 fr[2][WIDE_FR];
            __shared__ u32 stk[WIDE_ST][WIDE_T]; // [depth][thread]
            __shared__ u32 wsum[WIDE_W];
            __shared__ u32 sh_F, sh_cur;
            const u32 tid  = threadIdx.x;
            const u32 lane = (u32) __lane_id();
            const u32 wv   = tid / WAVE;
            const u32 j    = blockIdx.x;
            if (j >= n_list) return; // block-uniform
            const u32 g     = list[j];
            const f64 *ra   = nodes + (u64) (I + g) * NODE_F64;
            const f64 rint_a = ra[6] * Rkern;
            const f64 alo[3] = {ra[0], ra[1], ra[2]}, ahi[3] = {ra[3], ra[4], ra[5]};
            const f64 elo[3] = {alo[0] - rint_a, alo[1] - rint_a, alo[2] - rint_a};
            const f64 ehi[3] = {ahi[0] + rint_a, ahi[1] + rint_a, ahi[2] + rint_a};
            auto crit = [&](u32 node) {
                const f64 *rec = nodes + (u64) node * NODE_F64;
                f64 r      = rec[6] * Rkern;
                f64 nlo[3] = {rec[0], rec[1], rec[2]}, nhi[3] = {rec[3], rec[4], rec[5]};
                f64 xlo[3] = {nlo[0] - r, nlo[1] - r, nlo[2] - r};
                f64 xhi[3] = {nhi[0] + r, nhi[1] + r, nhi[2] + r};
                return cella_neigh_b(alo, ahi, xlo, xhi) | cella_neigh_b(elo, ehi, nlo, nhi);
            };
            const u32 ncell = I + L;
            constexpr u32 ACC = 0x80000000u; // frontier entry: a leaf already accepted
            bool bad = false;
            if (wv == 0) {
                // breadth-first expansion (wave-uniform control)
                u32 F = 1, cur = 0;
                if (lane == 0) fr[0][0] = 0; // the root
                wave_lds_sync();
                for (int round = 0;; round++) {
                    if (round == 128) { // deeper than any tree of these codes: a broken tree
                        bad = true;
                        break;
                    }
                    u32 total = 0;
                    bool internal = false, full = false;
                    for (u32 b = 0; b < F; b += WAVE) {
                        const u32 e = b + lane;
                        u32 no = 0, o0 = 0, o1 = 0;
                        if (e < F) {
                            const u32 node = fr[cur][e];
                            if (node & ACC) {
                                no = 1;
                                o0 = node;
                            } else if (node < ncell && crit(node)) {
                                if (node >= I) {
                                    no = 1;
                                    o0 = node | ACC;
                                } else {
                                    node_children(nodes, node, o0, o1);
                                    no = (o0 < ncell && o1 < ncell) ? 2u : 0u;
                                    bad |= no == 0;
                                }
                            }
                        }
                        u32 inc = no;
                        for (int o = 1; o < WAVE; o <<= 1) {
                            u32 t = __shfl_up(inc, o, WAVE);
                            if ((int) lane >= o) inc += t;
                        }
                        const u32 ct = __shfl(inc, WAVE - 1, WAVE);
                        if (total + ct > WIDE_FR) {
                            full = true;
                            break;
                        }
                        const u32 pos = total + inc - no;
                        if (no >= 1) fr[cur ^ 1][pos] = o0;
                        if (no == 2) fr[cur ^ 1][pos + 1] = o1;
                        internal |= __ballot(no == 2) != 0;
                        total += ct;
                    }
                    wave_lds_sync();
                    if (full) break; // keep the current frontier
                    cur ^= 1;
                    F = total;
                    if (!internal || F >= WIDE_T) break;
                }
                if (lane == 0) {
                    sh_F   = F;
                    sh_cur = cur;
                }
            }
            __syncthreads();
            const u32 F = sh_F, cur = sh_cur;
            // thread t: entries [t F / WIDE_T, (t + 1) F / WIDE_T) in order, each walked depth
            // first
            const u32 b0 = (tid * F) / WIDE_T, b1 = ((tid + 1) * F) / WIDE_T;
            auto walk = [&](auto emit) {
                for (u32 e = b0; e < b1; e++) {
                    const u32 node = fr[cur][e];
                    if (node & ACC) {
                        emit(node & ~ACC);
                        continue;
                    }
                    int sp            = 0;
                    stk[sp++][tid]    = node;
                    while (sp > 0) {
                        const u32 nd = stk[--sp][tid];
                        if (nd >= ncell || !crit(nd)) continue;
                        if (nd >= I) {
                            emit(nd);
                            continue;
                        }
                        u32 cl, cr;
                        node_children(nodes, nd, cl, cr);
                        if (sp + 2 > WIDE_ST) {
                            bad = true;
                            break;
                        }
                        stk[sp++][tid] = cr; // the left child is popped first
                        stk[sp++][tid] = cl;
                    }
                }
            };
            u32 c = 0;
            if constexpr (COUNT) {
                walk([&](u32) { c++; });
                tcnt[(u64) j * WIDE_T + tid] = c;
            } else {
                c = tcnt[(u64) j * WIDE_T + tid];
            }
            // block-wide exclusive scan of the counts
            u32 inc = c;
            for (int o = 1; o < WAVE; o <<= 1) {
                u32 t = __shfl_up(inc, o, WAVE);
                if ((int) lane >= o) inc += t;
            }
            if (lane == WAVE - 1) wsum[wv] = inc;
            const bool block_bad = __syncthreads_or(bad);
            u32 before = 0, tot = 0;
            for (u32 w = 0; w < (u32) WIDE_W; w++) {
                if (w < wv) before += wsum[w];
                tot += wsum[w];
            }
            if (COUNT) {
                if (tid == 0) {
                    if (block_bad) atomicOr(err, 1u);
                    else cnt[g] = tot;
                }
                return;
            }
            if (block_bad || tot != cnt[g]) {
                if (tid == 0) atomicOr(err, block_bad ? 1u : 4u);
                return;
            }
            u32 pos = offset[g] + before + inc - c;
            walk([&](u32 nd) { ids[pos++] = nd; });
        }

        // rows (leaf lists or particle rows) longer than their fixed capacity: flag and count
        __global__ void k_over_mark(u32 n, const u32 *__restrict cnt, u32 cap,
                                    u32 *__restrict flag, u32 *__restrict ocnt) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            const bool o = cnt[i] > cap;
            flag[i]      = o ? 1u : 0u;
            ocnt[i]      = o ? cnt[i] : 0u;
        }
        // the flagged rows in order into list (as map[i] when map is given), their offsets in
        // the overflow region after the fixed-capacity rows
        __global__ void k_over_list(u32 n, const u32 *__restrict flag, const u32 *__restrict fpos,
                                    const u32 *__restrict opos, u32 base,
                                    const u32 *__restrict map, u32 *__restrict list,
                                    u32 *__restrict offset) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n || !flag[i]) return;
            list[fpos[i]] = map ? map[i] : i;
            offset[i]     = base + opos[i];
        }

        // owner leaf, lanes over sorted positions (spatially coherent waves)
        __global__ void __launch_bounds__(256)
            k_owner_w(u32 nm, u32 n, u32 I, const f64 *__restrict nodes,
                      const u32 *__restrict sort_index, const double4 *__restrict ps, u32 L,
                      u32 *__restrict owner, u32 *__restrict err) {
            u32 s      = blockIdx.x * blockDim.x + threadIdx.x;
            u32 a      = s < nm ? sort_index[s] : 0xffffffffu;
            bool valid = a < n;
            u64 live   = __ballot(valid);
            if (!live) return;
            double4 q = ps[valid ? s : 0u];
            f64 p[3]  = {q.x, q.y, q.z};
            u32 found = 0x7fffffffu;
            bool ok   = wave_traverse(
                nodes, I, I + L, live,
                [&](const f64 *rec) { // is_coord_in_range_incl_max (geometry_utils.hpp:74-81)
                    return (rec[0] <= p[0]) & (p[0] <= rec[3]) & (rec[1] <= p[1]) &
                           (p[1] <= rec[4]) & (rec[2] <= p[2]) & (p[2] <= rec[5]);
                },
                [&](u32 node) { found = node - I; });
            if (!ok && __lane_id() == __builtin_ctzll(live)) atomicOr(err, 1u);
            if (valid) {
                if (found >= L) atomicOr(err, 2u);
                owner[a] = (found < L) ? found : 0u;
            }
        }

        // stage 2 count / fill.  SORTED layout: the CSR is indexed by the sorted position of
        // the particle and holds sorted positions of its neighbours; otherwise both are
        // particle ids (the reference layout).
        template<bool FILL, bool SORTED, bool STRIDE>
        __global__ void k_part_neigh(u32 n, const u32 *__restrict list,
                                     const u32 *__restrict list_len, u32 I,
                                     const f64 *__restrict nodes,
                                     const u32 *__restrict sort_index,
                                     const u32 *__restrict owner,
                                     const u32 *__restrict leaf_cnt,
                                     const u32 *__restrict leaf_offset,
                                     const u32 *__restrict leaf_ids, const f64 *__restrict x,
                                     const f64 *__restrict y, const f64 *__restrict z,
                                     const f64 *__restrict h, const u32 *__restrict inv,
                                     f64 htol, f64 Rker2, u32 *__restrict cnt,
                                     u32 *__restrict offset, u32 *__restrict ids, u32 cap,
                                     u32 leaf_cap) {
            u32 t = blockIdx.x * blockDim.x + threadIdx.x;
            if (t >= (list ? *list_len : n)) return;
            u32 a      = list ? list[t] : t;
            u32 ia     = SORTED ? inv[a] : a;
            f64 rint_a = h[a] * htol;
            f64 xa = x[a], ya = y[a], za = z[a];
            f64 ca = rint_a * rint_a * Rker2;
            u32 lo = owner[a];
            u32 k0 = leaf_offset[lo],
                k1 = k0 + (leaf_cap ? min(leaf_cnt[lo], leaf_cap) : leaf_cnt[lo]);
            const u32 c0 = STRIDE ? ia * cap : (FILL ? offset[ia] : 0u);
            const u32 ce = c0 + cap;
            u32 c        = c0;
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
                        if (FILL && (!STRIDE || c < ce)) ids[c] = SORTED ? s : b;
                        c++;
                    }
                }
            }
            if (!FILL || STRIDE) cnt[ia] = c - c0;
            if (STRIDE) offset[ia] = c0;
        }

        // sorted copy of positions and h (AoS, 32 B per object) + inverse permutation
        __global__ void k_sorted_copy(u32 nm, const u32 *__restrict sort_index,
                                      const f64 *__restrict x, const f64 *__restrict y,
                                      const f64 *__restrict z, const f64 *__restrict h,
                                      double4 *__restrict ps, u32 *__restrict inv) {
            u32 s = blockIdx.x * blockDim.x + threadIdx.x;
            if (s >= nm) return;
            u32 b  = sort_index[s];
            ps[s]  = make_double4(x[b], y[b], z[b], h[b]);
            inv[b] = s;
        }

        // particles whose owner leaf is not the leaf holding them go to the per-particle path
        __global__ void k_mismatch(u32 n, u32 I, const f64 *__restrict nodes,
                                   const u32 *__restrict owner, const u32 *__restrict inv,
                                   u32 *__restrict list, u32 *__restrict list_len) {
            u32 a = blockIdx.x * blockDim.x + threadIdx.x;
            if (a >= n) return;
            u32 s0, s1;
            node_children(nodes, I + owner[a], s0, s1);
            u32 sa = inv[a];
            if (!(s0 <= sa && sa < s1)) list[atomicAdd(list_len, 1u)] = a;
        }

        // local particles in sorted order: flag, then scatter after the scan
        __global__ void k_local_flag(u32 nm, u32 n, const u32 *__restrict sort_index,
                                     u32 *__restrict flag) {
            u32 s = blockIdx.x * blockDim.x + threadIdx.x;
            if (s < nm) flag[s] = sort_index[s] < n ? 1u : 0u;
        }
        __global__ void k_local_scatter(u32 nm, const u32 *__restrict flag,
                                        const u32 *__restrict pos, u32 *__restrict lpos) {
            u32 s = blockIdx.x * blockDim.x + threadIdx.x;
            if (s < nm && flag[s]) lpos[pos[s]] = s;
        }


        __device__ inline f64 wave_bcast(f64 v, int lane) {
            i64 b  = __double_as_longlong(v);
            int lo = __builtin_amdgcn_readlane((int) (b & 0xffffffff), lane);
            int hi = __builtin_amdgcn_readlane((int) (b >> 32), lane);
            return __longlong_as_double(((i64) hi << 32) | (u32) lo);
        }

        constexpr int GRP_WAVES = 4;

        // stage 2, one wave per member batch: a batch is 64 consecutive sorted positions of a
        // leaf (item_leaf / bstart map the wave to its leaf and batch), its members the local
        // particles among them whose owner is this leaf; the candidates are the particles of
        // the neighbour leaves, in list order then sorted order.  Candidates are taken 64 at a
        // time; those farther from the members' bounding box than both the members' largest
        // cut radius and their own are dropped (they cannot pass the pair test: the box
        // distance bounds every member distance from below, kept with a 1e-12 margin so that
        // rounding cannot drop a pair), and the rest are compacted into LDS in order (sorted
        // index, position, cut radius).  The wave then tests P candidates x M members per
        // instruction, lane = candidate * M + member, so the kept lanes of one member are in
        // candidate order and its row is the compaction of a ballot restricted to its lanes.
        // LIST: one wave per particle of list (item_leaf holds particle ids, n_items their
        // count): its whole row at offset[its sorted position], the path of rows longer than
        // their fixed capacity.
        template<bool FILL, bool SORTED, bool STRIDE, bool LIST = false, bool BCULL = false>
        __global__ void __launch_bounds__(GRP_WAVES * WAVE)
            k_group_neigh(u32 n, u32 L, u32 I, const f64 *__restrict nodes,
                          const u32 *__restrict sort_index, const double4 *__restrict ps,
                          const u32 *__restrict owner, const u32 *__restrict leaf_cnt,
                          const u32 *__restrict leaf_offset, const u32 *__restrict leaf_ids,
                          const u32 *__restrict item_leaf, const u32 *__restrict bstart,
                          const u32 *__restrict n_items, f64 htol, f64 Rker2,
                          u32 *__restrict cnt, u32 *__restrict offset, u32 *__restrict ids,
                          u32 cap, u32 leaf_cap, const u32 *__restrict inv = nullptr,
                          const f64 *__restrict bbox = nullptr) {
            __shared__ u32 rstart[GRP_WAVES][WAVE], rexcl[GRP_WAVES][WAVE];
            // BCULL: the neighbour leaves of one pass (range, first batch, batches before it)
            __shared__ u32 lr0[GRP_WAVES][BCULL ? WAVE : 1], lr1[GRP_WAVES][BCULL ? WAVE : 1];
            __shared__ u32 lbs[GRP_WAVES][BCULL ? WAVE : 1], lbx[GRP_WAVES][BCULL ? WAVE : 1];
            __shared__ double4 cpos[GRP_WAVES][WAVE]; // candidate x, y, z, cut radius^2
            __shared__ u32 cid[GRP_WAVES][WAVE];      // candidate id written to the rows
            __shared__ double4 mpos[GRP_WAVES][WAVE]; // member x, y, z, cut radius^2
            __shared__ u32 mrow[GRP_WAVES][WAVE];     // member row start
            const int lane = __lane_id(), wid = threadIdx.x / WAVE;
            const u32 item = blockIdx.x * GRP_WAVES + wid;
            if (item >= *n_items) return;
            u32 l, mb, m1;
            if constexpr (LIST) { // the batch is the listed particle alone (lane 0)
                const u32 a0 = item_leaf[item];
                l            = owner[a0];
                mb           = inv[a0];
                m1           = mb + 1;
            } else {
                l = item_leaf[item];
                u32 m0;
                node_children(nodes, I + l, m0, m1);
                mb = m0 + (item - bstart[l]) * WAVE;
            }
            const u64 lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
            // leaf_cap (fixed-capacity leaf rows): a row that overflowed holds only leaf_cap
            // entries; the build is repeated then, but nothing may be read past the row
            const u32 k0 = leaf_offset[l],
                      k1 = k0 + (leaf_cap ? min(leaf_cnt[l], leaf_cap) : leaf_cnt[l]);
            u32 sm   = mb + lane;
            bool mem = false;
            u32 a    = 0;
            double4 pa{0, 0, 0, 0};
            if (sm < m1) {
                a = sort_index[sm];
                if (a < n && owner[a] == l) {
                    mem       = true;
                    double4 p = ps[sm];
                    f64 ra    = p.w * htol;
                    pa        = make_double4(p.x, p.y, p.z, ra * ra * Rker2);
                }
            }
            const u64 members = __ballot(mem);
            if (!members) return;
            const u32 M    = (u32) __popcll(members);
            const u32 rank = (u32) __popcll(members & lt); // member index of this lane
            const u32 ia   = SORTED ? sm : a;
            const u32 row0 = STRIDE ? (mem ? ia * cap : 0u) : ((FILL && mem) ? offset[ia] : 0u);
            if (mem) {
                mpos[wid][rank] = pa;
                mrow[wid][rank] = row0;
            }
            // the members' bounding box and largest cut radius^2
            constexpr f64 BIG = 1.7976931348623157e308;
            auto fmn = [](f64 p, f64 q) { return fmin(p, q); };
            auto fmx = [](f64 p, f64 q) { return fmax(p, q); };
            const f64 blx = wave_reduce(mem ? pa.x : BIG, fmn), bhx = wave_reduce(mem ? pa.x : -BIG, fmx);
            const f64 bly = wave_reduce(mem ? pa.y : BIG, fmn), bhy = wave_reduce(mem ? pa.y : -BIG, fmx);
            const f64 blz = wave_reduce(mem ? pa.z : BIG, fmn), bhz = wave_reduce(mem ? pa.z : -BIG, fmx);
            const f64 br2 = wave_reduce(mem ? pa.w : 0.0, fmx);
            // lane = c * M + mi: P candidates per pass; lanes past P * M stay idle
            const u32 P  = WAVE / M;
            const u32 mi = (u32) lane % M, ci = (u32) lane / M;
            const bool on = ci < P;
            u64 mine = 0; // this member's lanes
            for (u32 c = 0; c < P; c++) mine |= 1ull << (c * M + mi);
            wave_lds_sync();
            const double4 me = mpos[wid][on ? mi : 0];
            const u32 mrow0  = mrow[wid][on ? mi : 0];
            u32 cursor = mrow0, count = 0;
            // the candidates of the ranges rstart / rexcl [0, nr), `total` in all: staged 64 at
            // a time, tested against the members
            auto candidates = [&](const u32 nr, const u32 total) {
                for (u32 c0 = 0; c0 < total; c0 += WAVE) {
                    // stage 64 candidates: lane j takes candidate c0 + j; the ones that can
                    // reach a member are compacted to the front of cpos / cid
                    u32 cn;
                    {
                        const u32 c = c0 + lane;
                        // the range holding candidate c: largest j with rexcl[j] <= c
                        // (a dropped range is empty and shares its rexcl with the next one, so
                        // the largest such j is a range that holds c)
                        u32 j = 0;
                        for (u32 st = WAVE / 2; st > 0; st >>= 1) {
                            u32 t = j + st;
                            if (t < nr && rexcl[wid][t] <= c) j = t;
                        }
                        const bool valid = c < total;
                        u32 sb    = valid ? rstart[wid][j] + (c - rexcl[wid][j]) : 0u;
                        double4 q = ps[sb];
                        f64 rb    = q.w * htol;
                        f64 rb2   = rb * rb * Rker2;
                        f64 ex = q.x < blx ? blx - q.x : (q.x > bhx ? q.x - bhx : 0.0);
                        f64 ey = q.y < bly ? bly - q.y : (q.y > bhy ? q.y - bhy : 0.0);
                        f64 ez = q.z < blz ? blz - q.z : (q.z > bhz ? q.z - bhz : 0.0);
                        f64 d2 = (ex * ex + ey * ey + ez * ez) * (1.0 - 1e-12);
                        const bool near = valid & !((d2 > br2) & (d2 > rb2));
                        const u64 nm    = __ballot(near);
                        const u32 pos   = (u32) __popcll(nm & lt);
                        cn              = (u32) __popcll(nm);
                        wave_lds_sync();
                        if (near) {
                            cpos[wid][pos] = make_double4(q.x, q.y, q.z, rb2);
                            cid[wid][pos]  = FILL ? (SORTED ? sb : sort_index[sb]) : 0u;
                        }
                        wave_lds_sync();
                    }
                    for (u32 cb0 = 0; cb0 < cn; cb0 += P) {
                        const u32 cl = cb0 + ci; // candidate of this lane
                        bool keep    = false;
                        if (on && cl < cn) {
                            const double4 q = cpos[wid][cl];
                            f64 dx = me.x - q.x, dy = me.y - q.y, dz = me.z - q.z;
                            f64 rab2 = dx * dx + dy * dy + dz * dz;
                            keep     = !((rab2 > me.w) & (rab2 > q.w));
                        }
                        const u64 km = __ballot(keep);
                        const u64 kt = km & mine;
                        if (FILL && keep) {
                            u32 pos = cursor + (u32) __popcll(kt & lt);
                            if (STRIDE) {
                                if (pos < mrow0 + cap) ids[pos] = cid[wid][cl];
                            } else {
                                ids[pos] = cid[wid][cl];
                            }
                        }
                        const u32 kc = (u32) __popcll(kt);
                        cursor += kc;
                        count += kc;
                    }
                }
            };
            for (u32 kb = k0; kb < k1; kb += WAVE) {
                // neighbour leaves in batches of up to 64: lane i loads the range of the i-th
                // leaf of the batch; candidates are the concatenated ranges in order
                u32 k  = kb + lane;
                u32 r0 = 0, r1 = 0, nd = 0;
                if (k < k1) {
                    nd = leaf_ids[k];
                    node_children(nodes, nd, r0, r1); // leaf: sorted range
                    // the whole range is dropped when none of its particles can reach a
                    // member: the leaf's box is farther from the members' box than both the
                    // members' largest cut radius and the leaf's (rint = max h * htol).  The
                    // box distance bounds each particle's from below, so this drops only
                    // candidates the per-particle test below would drop
                    const f64 *rec = nodes + (u64) nd * NODE_F64;
                    const f64 ex   = fmax(fmax(rec[0] - bhx, blx - rec[3]), 0.);
                    const f64 ey   = fmax(fmax(rec[1] - bhy, bly - rec[4]), 0.);
                    const f64 ez   = fmax(fmax(rec[2] - bhz, blz - rec[5]), 0.);
                    const f64 d2   = (ex * ex + ey * ey + ez * ez) * (1.0 - 1e-12);
                    const f64 rr2  = rec[6] * rec[6] * Rker2;
                    if (d2 > br2 && d2 > rr2) r1 = r0;
                }
                const u32 nr = min(k1 - kb, (u32) WAVE);
                if constexpr (!BCULL) {
                    u32 sz  = r1 - r0;
                    u32 inc = sz;
                    for (int o = 1; o < WAVE; o <<= 1) {
                        u32 t = __shfl_up(inc, o, WAVE);
                        if (lane >= o) inc += t;
                    }
                    const u32 total = __shfl(inc, WAVE - 1, WAVE);
                    wave_lds_sync();
                    rstart[wid][lane] = r0;
                    rexcl[wid][lane]  = inc - sz;
                    wave_lds_sync();
                    candidates(nr, total);
                } else {
                    // each range split into the member batches of its leaf (64 sorted
                    // positions from the leaf's start; bbox: their boxes and largest cut
                    // radius^2), every batch dropped whole by the test above
                    const bool okl = nd >= I && nd - I < L;
                    const u32 nbl  = okl ? (r1 - r0 + WAVE - 1) / WAVE : 0u;
                    const u32 bs0  = nbl ? bstart[nd - I] : 0u;
                    u32 binc       = nbl;
                    for (int o = 1; o < WAVE; o <<= 1) {
                        u32 t = __shfl_up(binc, o, WAVE);
                        if (lane >= o) binc += t;
                    }
                    const u32 totb = __shfl(binc, WAVE - 1, WAVE);
                    wave_lds_sync();
                    lr0[wid][lane] = r0;
                    lr1[wid][lane] = r1;
                    lbs[wid][lane] = bs0;
                    lbx[wid][lane] = binc - nbl;
                    wave_lds_sync();
                    for (u32 t0 = 0; t0 < totb; t0 += WAVE) {
                        const u32 tb = t0 + lane;
                        u32 s0 = 0, s1 = 0;
                        if (tb < totb) {
                            // the range holding batch tb: largest j with lbx[j] <= tb
                            u32 j = 0;
                            for (u32 st = WAVE / 2; st > 0; st >>= 1) {
                                u32 q = j + st;
                                if (q < nr && lbx[wid][q] <= tb) j = q;
                            }
                            const u32 bi = tb - lbx[wid][j];
                            s0           = lr0[wid][j] + bi * WAVE;
                            s1           = min(s0 + WAVE, lr1[wid][j]);
                            const f64 *bx = bbox + (u64) (lbs[wid][j] + bi) * 8;
                            const f64 ex  = fmax(fmax(bx[0] - bhx, blx - bx[3]), 0.);
                            const f64 ey  = fmax(fmax(bx[1] - bhy, bly - bx[4]), 0.);
                            const f64 ez  = fmax(fmax(bx[2] - bhz, blz - bx[5]), 0.);
                            const f64 d2  = (ex * ex + ey * ey + ez * ez) * (1.0 - 1e-12);
                            if (d2 > br2 && d2 > bx[6]) s1 = s0;
                        }
                        const u32 sz = s1 - s0;
                        u32 inc      = sz;
                        for (int o = 1; o < WAVE; o <<= 1) {
                            u32 t = __shfl_up(inc, o, WAVE);
                            if (lane >= o) inc += t;
                        }
                        const u32 total = __shfl(inc, WAVE - 1, WAVE);
                        wave_lds_sync();
                        rstart[wid][lane] = s0;
                        rexcl[wid][lane]  = inc - sz;
                        wave_lds_sync();
                        candidates(min(totb - t0, (u32) WAVE), total);
                    }
                }
            }
            // lane mi (< M) holds member mi's count; the member lane of rank mi writes it
            wave_lds_sync();
            if (on && ci == 0) mrow[wid][mi] = count;
            wave_lds_sync();
            if (mem) {
                u32 cm = mrow[wid][rank];
                if (!FILL || STRIDE) cnt[ia] = cm;
                if (STRIDE) offset[ia] = row0;
            }
        }

        // member batches: nb[l] = batches of 64 sorted positions in leaf l (>= 1)
        __global__ void k_leaf_batches(u32 L, u32 I, const f64 *__restrict nodes,
                                       u32 *__restrict nbatch) {
            u32 l = blockIdx.x * blockDim.x + threadIdx.x;
            if (l >= L) return;
            u32 m0, m1;
            node_children(nodes, I + l, m0, m1);
            nbatch[l] = max(1u, (m1 - m0 + WAVE - 1) / WAVE);
        }
        // item -> leaf for the batches of every leaf (bstart: exclusive scan of nbatch)
        __global__ void k_item_leaf(u32 L, const u32 *__restrict nbatch,
                                    const u32 *__restrict bstart, u32 *__restrict item_leaf) {
            u32 l = blockIdx.x * blockDim.x + threadIdx.x;
            if (l >= L) return;
            const u32 b0 = bstart[l], nbl = nbatch[l];
            for (u32 j = 0; j < nbl; j++) item_leaf[b0 + j] = l;
        }
        // box of every member batch (all its sorted positions, images included: they are
        // candidates too) and the largest cut radius^2 among them, 8 doubles per batch:
        // lo x, y, z, hi x, y, z, (h htol)^2 R^2, unused.  One wave per batch.
        __global__ void __launch_bounds__(GRP_WAVES * WAVE)
            k_batch_boxes(u32 I, const f64 *__restrict nodes, const u32 *__restrict item_leaf,
                          const u32 *__restrict bstart, const u32 *__restrict n_items,
                          const double4 *__restrict ps, f64 htol, f64 Rker2,
                          f64 *__restrict bbox) {
            const int lane = __lane_id();
            const u32 item = blockIdx.x * GRP_WAVES + threadIdx.x / WAVE;
            if (item >= *n_items) return;
            const u32 l = item_leaf[item];
            u32 m0, m1;
            node_children(nodes, I + l, m0, m1);
            const u32 sm = m0 + (item - bstart[l]) * WAVE + lane;
            constexpr f64 BIG = 1.7976931348623157e308;
            f64 lx = BIG, ly = BIG, lz = BIG, hx = -BIG, hy = -BIG, hz = -BIG, r2 = 0;
            if (sm < m1) {
                const double4 q = ps[sm];
                const f64 rb    = q.w * htol;
                lx = hx = q.x, ly = hy = q.y, lz = hz = q.z;
                r2 = rb * rb * Rker2;
            }
            auto fmn = [](f64 p, f64 q) { return fmin(p, q); };
            auto fmx = [](f64 p, f64 q) { return fmax(p, q); };
            lx = wave_reduce(lx, fmn), ly = wave_reduce(ly, fmn), lz = wave_reduce(lz, fmn);
            hx = wave_reduce(hx, fmx), hy = wave_reduce(hy, fmx), hz = wave_reduce(hz, fmx);
            r2 = wave_reduce(r2, fmx);
            if (lane < 8) {
                const f64 v[8] = {lx, ly, lz, hx, hy, hz, r2, 0.};
                bbox[(u64) item * 8 + lane] = v[lane];
            }
        }
    } // namespace

    namespace {
        // max of n u32 values into *out (device; *out must start at 0)
        __global__ void __launch_bounds__(256)
            k_max_u32(u32 n, const u32 *__restrict v, u32 *__restrict out) {
            __shared__ u32 lds[256 / WAVE];
            u32 m = 0;
            for (u32 i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
                m = max(m, v[i]);
            m = wave_reduce(m, [](u32 p, u32 q) { return p > q ? p : q; });
            if (__lane_id() == 0) lds[threadIdx.x / WAVE] = m;
            __syncthreads();
            if (threadIdx.x == 0) {
                for (int w = 1; w < 256 / WAVE; w++) m = max(m, lds[w]);
                atomicMax(out, m);
            }
        }
        void max_u32(hipStream_t s, u32 n, const u32 *v, u32 *out) {
            if (n) k_max_u32<<<std::min<u32>(div_up(n, 256), 256), 256, 0, s>>>(n, v, out);
        }

        // the stage-2 wave items (one per 64 sorted positions of every leaf): item -> leaf map,
        // first item of every leaf, item count at counters[7]; returns the launch's block
        // count, sized for the bound L + nm / 64 (waves past the count return at once)
        u32 batch_items(hipStream_t s, NeighBuffers &nb, const TreeBuffers &t) {
            const u32 L = t.L, bound = L + div_up(t.n_obj, (u32) WAVE);
            nb.nbatch.resize(L);
            nb.bstart.resize(L);
            nb.item_leaf.resize(bound);
            k_leaf_batches<<<div_up(L, 256), 256, 0, s>>>(L, t.I, t.nodes.data(),
                                                           nb.nbatch.data());
            exclusive_scan_u32(s, nb.nbatch.data(), nb.bstart.data(), L, nb.counters.data() + 7,
                               nb.scratch.data());
            k_item_leaf<<<div_up(L, 256), 256, 0, s>>>(L, nb.nbatch.data(), nb.bstart.data(),
                                                        nb.item_leaf.data());
            return div_up(bound, (u32) GRP_WAVES);
        }

        // error flags of the walks (counters[4]) into exceptions
        void check_walk_flags(u32 f) {
            if (f & 1u) throw std::runtime_error("neighbour cache: tree traversal stack overflow");
            if (f & 2u)
                throw std::runtime_error("neighbour cache: a particle is outside every leaf box");
            if (f & 4u)
                throw std::runtime_error("neighbour cache: a leaf walk disagrees with its count");
        }

        // boxes of the member batches for the batch-level cull of stage 2, on sets whose tree
        // sort took the spatial tie order (runs of equal codes: leaves of hundreds of
        // particles, where a leaf-level test keeps most candidates); SHAMX_BATCH_CULL=1 forces
        // it (tests).  nullptr: no batch cull.
        const f64 *batch_boxes(hipStream_t s, NeighBuffers &nb, const TreeBuffers &t,
                               u32 gblocks, f64 htol, f64 Rker2) {
            const char *e = std::getenv("SHAMX_BATCH_CULL");
            if (!t.fine_ties && !(e && std::atoi(e) == 1)) return nullptr;
            nb.bbox.resize((size_t) gblocks * GRP_WAVES * 8);
            k_batch_boxes<<<gblocks, GRP_WAVES * WAVE, 0, s>>>(
                t.I, t.nodes.data(), nb.item_leaf.data(), nb.bstart.data(),
                nb.counters.data() + 7, nb.ps.data(), htol, Rker2, nb.bbox.data());
            return nb.bbox.data();
        }
        // stage-2 launch, with the batch-level cull when bb holds the batch boxes
        template<bool FILL, bool SORTED, bool STRIDE, class... A>
        void launch_group(hipStream_t s, u32 gblocks, const f64 *bb, A... a) {
            if (bb)
                k_group_neigh<FILL, SORTED, STRIDE, false, true>
                    <<<gblocks, GRP_WAVES * WAVE, 0, s>>>(a..., nullptr, bb);
            else
                k_group_neigh<FILL, SORTED, STRIDE><<<gblocks, GRP_WAVES * WAVE, 0, s>>>(a...);
        }

        // packed rows, two passes (count, scan, fill); SORTED: rows by sorted position
        // (leaf_cap: capacity of fixed-capacity leaf lists, 0 for packed ones)
        template<bool SORTED>
        void stage2_packed(hipStream_t s, NeighBuffers &nb, const TreeBuffers &t, u32 n_local,
                           const f64 *x, const f64 *y, const f64 *z, const f64 *h, f64 htol,
                           f64 Rker2, u32 *hc, u32 leaf_cap = 0) {
            u32 L = t.L, I = t.I;
            u32 rows      = SORTED ? t.n_obj : n_local;
            u32 *mism_len = nb.counters.data() + 5;
            u32 gblocks   = batch_items(s, nb, t);
            const u32 *n_items = nb.counters.data() + 7;
            if (SORTED) fill_u32(s, nb.cnt.data(), 0, rows); // ghost rows stay empty
            const f64 *bb = batch_boxes(s, nb, t, gblocks, htol, Rker2);
            launch_group<false, SORTED, false>(s, gblocks, bb,
                n_local, L, I, t.nodes.data(), t.sort_index.data(), nb.ps.data(),
                nb.owner.data(), nb.leaf_cnt.data(), nb.leaf_offset.data(), nb.leaf_ids.data(),
                nb.item_leaf.data(), nb.bstart.data(), n_items, htol, Rker2, nb.cnt.data(),
                nullptr, nullptr, 0, leaf_cap);
            k_part_neigh<false, SORTED, false><<<div_up(n_local, 128), 128, 0, s>>>(
                n_local, nb.mism.data(), mism_len, I, t.nodes.data(), t.sort_index.data(),
                nb.owner.data(), nb.leaf_cnt.data(), nb.leaf_offset.data(), nb.leaf_ids.data(), x,
                y, z, h, nb.inv.data(), htol, Rker2, nb.cnt.data(), nullptr, nullptr, 0,
                leaf_cap);
            exclusive_scan_u32(s, nb.cnt.data(), nb.offset.data(), rows, nb.counters.data() + 1,
                               nb.scratch.data());
            // longest leaf list and row (counters[8], [9])
            fill_u32(s, nb.counters.data() + 8, 0, 2);
            max_u32(s, L, nb.leaf_cnt.data(), nb.counters.data() + 8);
            max_u32(s, rows, nb.cnt.data(), nb.counters.data() + 9);
            SHAMX_HIP(hipMemcpyAsync(hc, nb.counters.data(), 16 * sizeof(u32),
                                     hipMemcpyDeviceToHost, s));
            SHAMX_HIP(hipStreamSynchronize(s));
            check_walk_flags(hc[4]);
            nb.n_pairs    = hc[1];
            nb.n_mismatch = hc[5];
            nb.ids.resize(std::max<u64>(nb.n_pairs, 1));
            launch_group<true, SORTED, false>(s, gblocks, bb,
                n_local, L, I, t.nodes.data(), t.sort_index.data(), nb.ps.data(),
                nb.owner.data(), nb.leaf_cnt.data(), nb.leaf_offset.data(), nb.leaf_ids.data(),
                nb.item_leaf.data(), nb.bstart.data(), n_items, htol, Rker2, nullptr,
                nb.offset.data(), nb.ids.data(), 0, leaf_cap);
            k_part_neigh<true, SORTED, false><<<div_up(n_local, 128), 128, 0, s>>>(
                n_local, nb.mism.data(), mism_len, I, t.nodes.data(), t.sort_index.data(),
                nb.owner.data(), nb.leaf_cnt.data(), nb.leaf_offset.data(), nb.leaf_ids.data(), x,
                y, z, h, nb.inv.data(), htol, Rker2, nullptr, nb.offset.data(), nb.ids.data(), 0,
                leaf_cap);
        }

        // leaves per wave in the stage-1 walk (SHAMX_LEAF_LANES: 64, 32 or 16)
        u32 leaf_lanes() {
            static const u32 v = [] {
                const char *e = std::getenv("SHAMX_LEAF_LANES");
                int x         = e ? std::atoi(e) : 64;
                return (u32) ((x == 16 || x == 32) ? x : 64);
            }();
            return v;
        }

        // SHAMX_NEIGH_ROWS=packed (read at every build) takes packed particle rows even where
        // fixed-capacity rows fit, so that path can be tested at small sizes
        bool packed_rows_forced() {
            const char *e = std::getenv("SHAMX_NEIGH_ROWS");
            return e && std::strcmp(e, "packed") == 0;
        }

        // row capacity for a measured maximum: 1/8 headroom, rounded up to 16 entries so that
        // rows start on 64-byte boundaries
        u32 grow_cap(u32 need) { return (need + need / 8 + 16 + 15) & ~15u; }
    } // namespace

    void neigh_build(hipStream_t s, NeighBuffers &nb, TreeBuffers &t, u32 n_local,
                     const f64 *x, const f64 *y, const f64 *z, const f64 *h, f64 htol,
                     f64 Rkern, bool sorted_layout) {
        u32 L = t.L, I = t.I, nm = t.n_obj;
        nb.n_local       = n_local;
        nb.sorted_layout = sorted_layout;
        u32 rows         = sorted_layout ? nm : n_local;
        nb.counters.resize(16);
        nb.scratch.resize(std::max({scan_scratch_bytes(L), scan_scratch_bytes(rows),
                                    scan_scratch_bytes(nm)}));
        nb.leaf_cnt.resize(L);
        nb.leaf_offset.resize(L);
        nb.owner.resize(std::max<u32>(n_local, 1));
        nb.cnt.resize(std::max<u32>(rows, 1));
        nb.offset.resize(std::max<u32>(rows, 1));
        nb.psum.resize(std::max<u32>(std::max(rows, L), 1));
        u32 *err = nb.counters.data() + 4;
        fill_u32(s, nb.counters.data(), 0, 16);
        f64 Rker2 = Rkern * Rkern;
        u32 *hc = nb.host.data(); // 16 pinned words: the reads queue on the stream

        // sorted copy and inverse permutation (independent of the leaf lists)
        nb.ps.resize(nm);
        nb.inv.resize(nm);
        nb.mism.resize(std::max<u32>(n_local, 1));
        k_sorted_copy<<<div_up(nm, 256), 256, 0, s>>>(nm, t.sort_index.data(), x, y, z, h,
                                                      nb.ps.data(), nb.inv.data());

        if (!sorted_layout) {
            // reference layout: stage 1 in two passes
            k_leaf_neigh_w<false, false><<<div_up(L, 256), 256, 0, s>>>(
                L, I, t.nodes.data(), Rkern, nb.leaf_cnt.data(), nullptr, nullptr, 0, err, 64);
            exclusive_scan_u32(s, nb.leaf_cnt.data(), nb.leaf_offset.data(), L,
                               nb.counters.data(), nb.scratch.data());
            SHAMX_HIP(hipMemcpyAsync(hc, nb.counters.data(), 8 * sizeof(u32),
                                     hipMemcpyDeviceToHost, s));
            SHAMX_HIP(hipStreamSynchronize(s));
            tree_check(t);
            if (hc[4]) throw std::runtime_error("neighbour cache: tree traversal stack overflow");
            nb.n_leaf_pairs = hc[0];
            nb.leaf_stride  = 0;
            nb.leaf_ids.resize(std::max<u64>(nb.n_leaf_pairs, 1));
            k_leaf_neigh_w<true, false><<<div_up(L, 256), 256, 0, s>>>(
                L, I, t.nodes.data(), Rkern, nullptr, nb.leaf_offset.data(), nb.leaf_ids.data(),
                0, err, 64);
            if (n_local == 0) {
                nb.n_pairs = 0;
                return;
            }
            k_owner_w<<<div_up(nm, 256), 256, 0, s>>>(nm, n_local, I, t.nodes.data(),
                                                      t.sort_index.data(), nb.ps.data(), L,
                                                      nb.owner.data(), err);
            k_mismatch<<<div_up(n_local, 256), 256, 0, s>>>(n_local, I, t.nodes.data(),
                                                            nb.owner.data(), nb.inv.data(),
                                                            nb.mism.data(),
                                                            nb.counters.data() + 5);
            stage2_packed<false>(s, nb, t, n_local, x, y, z, h, htol, Rker2, hc);
            nb.row_stride = 0;
            SHAMX_HIP(hipGetLastError());
            return;
        }

        // sorted layout: owner leaves, mismatch list, local rows in sorted order
        k_owner_w<<<div_up(nm, 256), 256, 0, s>>>(nm, n_local, I, t.nodes.data(),
                                                  t.sort_index.data(), nb.ps.data(), L,
                                                  nb.owner.data(), err);
        k_mismatch<<<div_up(n_local, 256), 256, 0, s>>>(n_local, I, t.nodes.data(),
                                                        nb.owner.data(), nb.inv.data(),
                                                        nb.mism.data(), nb.counters.data() + 5);
        nb.lflag.resize(nm);
        nb.lscan.resize(nm);
        nb.lpos.resize(std::max<u32>(n_local, 1));
        k_local_flag<<<div_up(nm, 256), 256, 0, s>>>(nm, n_local, t.sort_index.data(),
                                                     nb.lflag.data());
        exclusive_scan_u32(s, nb.lflag.data(), nb.lscan.data(), nm, nb.counters.data() + 6,
                           nb.scratch.data());
        k_local_scatter<<<div_up(nm, 256), 256, 0, s>>>(nm, nb.lflag.data(), nb.lscan.data(),
                                                        nb.lpos.data());

        // Both stages in one pass each, rows of fixed capacity, chosen from the previous build
        // (about four times the mean length, at most the longest row plus headroom).  Rows
        // longer than their capacity (the counts run on past it) are listed and filled again,
        // whole, into an overflow region after the fixed-capacity rows, their offsets pointing
        // there: consumers read (offset, count) and see one layout.  Leaf lists that overflow
        // must be complete before stage 2 reads them: once a build has seen leaf overflow,
        // the next ones resolve it between the stages (one more host sync); the first time, the
        // build is repeated.
        const u32 gblocks = batch_items(s, nb, t);
        // capacity for rows of measured longest `mx` and total `tot` over `rows_` rows, at most
        // what u32 entry indices address for rows_ rows (a multiple of 16: 64-byte rows)
        auto fit = [](u32 rows_, u32 cap) {
            u64 lim = 0xffffffffull / std::max<u32>(rows_, 1);
            return (u32) std::max<u64>(16, std::min<u64>(cap, lim) & ~15ull);
        };
        auto choose = [&](u32 mx, u64 tot, u32 used_rows, u32 rows_) {
            u64 mean4 = (4 * div_up(tot, (u64) std::max<u32>(used_rows, 1)) + 15) & ~15ull;
            return fit(rows_, (u32) std::min<u64>(grow_cap(mx), std::max<u64>(64, mean4)));
        };
        if (nb.cap_leaf == 0) nb.cap_leaf = fit(L, 128);
        if (nb.cap_part == 0) nb.cap_part = fit(nm, 256);
        if (const char *e = std::getenv("SHAMX_NEIGH_CAP")) { // tests: force overflow
            nb.cap_leaf = nb.cap_part = std::max(16, std::atoi(e)) & ~15;
        }
        u32 *mism_len = nb.counters.data() + 5;
        const u32 big = std::max(L, nm);
        for (auto *a_ : {&nb.oflag, &nb.ocnt, &nb.ofpos, &nb.opos, &nb.olist}) a_->resize(big);
        // list the rows of cnt[0, n) longer than cap: count at counters[c], total at c + 1
        auto mark = [&](u32 n, const u32 *cnt, u32 cap, int c) {
            k_over_mark<<<div_up(n, 256), 256, 0, s>>>(n, cnt, cap, nb.oflag.data(),
                                                        nb.ocnt.data());
            exclusive_scan_u32(s, nb.oflag.data(), nb.ofpos.data(), n, nb.counters.data() + c,
                               nb.scratch.data());
            exclusive_scan_u32(s, nb.ocnt.data(), nb.opos.data(), n, nb.counters.data() + c + 1,
                               nb.scratch.data());
        };
        // blocks of 8 waves per listed leaf while there are few of them, 1 wave beyond
        constexpr u32 WIDE_FEW = 8192;
        // the listed leaves (n_over, from the fixed-capacity pass's counts: above the capacity
        // for exactly these) walked again into the overflow region of leaf_ids: their true
        // list lengths first (one leaf per block, its walk spread over the threads: the listed
        // leaves are far apart, and a wave walks the union of its lanes' walks), then the
        // maximum, the total and the overflow offsets over the true lengths (counters 8, 10,
        // 12, 13, read back into hc: one more host read), then the lists
        auto resolve_leaves = [&](u32 cl, u32 n_over) {
            const u64 base = (u64) L * cl;
            k_over_list<<<div_up(L, 256), 256, 0, s>>>(L, nb.oflag.data(), nb.ofpos.data(),
                                                        nb.opos.data(), (u32) base, nullptr,
                                                        nb.olist.data(), nb.leaf_offset.data());
            const bool few = n_over <= WIDE_FEW;
            nb.wide_cnt.resize((size_t) n_over * (few ? 8 : 1) * WAVE);
            auto walk_wide = [&](auto count, u32 *offset, u32 *ids_) {
                constexpr bool C = decltype(count)::value;
                if (few)
                    k_leaf_walk_wide<8, C><<<n_over, 8 * WAVE, 0, s>>>(
                        L, I, t.nodes.data(), Rkern, nb.olist.data(), n_over, nb.leaf_cnt.data(),
                        nb.wide_cnt.data(), offset, ids_, err);
                else
                    k_leaf_walk_wide<1, C><<<n_over, WAVE, 0, s>>>(
                        L, I, t.nodes.data(), Rkern, nb.olist.data(), n_over, nb.leaf_cnt.data(),
                        nb.wide_cnt.data(), offset, ids_, err);
            };
            walk_wide(std::true_type{}, nullptr, nullptr);
            fill_u32(s, nb.counters.data() + 8, 0, 1);
            max_u32(s, L, nb.leaf_cnt.data(), nb.counters.data() + 8);
            exclusive_scan_u32(s, nb.leaf_cnt.data(), nb.psum.data(), L, nb.counters.data() + 10,
                               nb.scratch.data());
            mark(L, nb.leaf_cnt.data(), cl, 12);
            SHAMX_HIP(hipMemcpyAsync(hc, nb.counters.data(), 16 * sizeof(u32),
                                     hipMemcpyDeviceToHost, s));
            SHAMX_HIP(hipStreamSynchronize(s));
            check_walk_flags(hc[4]);
            const u32 tot = hc[13];
            if (hc[12] != n_over)
                throw std::runtime_error("neighbour cache: overflowing leaf lists changed");
            if (base + tot > 0xffffffffull)
                throw std::runtime_error("neighbour cache: leaf lists exceed u32 entry indices");
            nb.leaf_ids.resize_keep(std::max<u64>(base + tot, 1), s);
            k_over_list<<<div_up(L, 256), 256, 0, s>>>(L, nb.oflag.data(), nb.ofpos.data(),
                                                        nb.opos.data(), (u32) base, nullptr,
                                                        nb.olist.data(), nb.leaf_offset.data());
            walk_wide(std::false_type{}, nb.leaf_offset.data(), nb.leaf_ids.data());
        };
        for (int attempt = 0;; attempt++) {
            if (attempt == 4) throw std::runtime_error("neighbour cache: capacities do not settle");
            u32 cl = nb.cap_leaf, cp = nb.cap_part;
            if ((u64) L * cl > 0xffffffffull) {
                // u32 entry indices cannot address fixed-capacity leaf lists: two passes each
                k_leaf_neigh_w<false, false><<<div_up(L, 256), 256, 0, s>>>(
                    L, I, t.nodes.data(), Rkern, nb.leaf_cnt.data(), nullptr, nullptr, 0, err, 64);
                exclusive_scan_u32(s, nb.leaf_cnt.data(), nb.leaf_offset.data(), L,
                                   nb.counters.data(), nb.scratch.data());
                SHAMX_HIP(hipMemcpyAsync(hc, nb.counters.data(), 8 * sizeof(u32),
                                         hipMemcpyDeviceToHost, s));
                SHAMX_HIP(hipStreamSynchronize(s));
                tree_check(t);
                if (hc[4]) throw std::runtime_error("neighbour cache: tree traversal stack overflow");
                nb.n_leaf_pairs = hc[0];
                nb.leaf_ids.resize(std::max<u64>(nb.n_leaf_pairs, 1));
                k_leaf_neigh_w<true, false><<<div_up(L, 256), 256, 0, s>>>(
                    L, I, t.nodes.data(), Rkern, nullptr, nb.leaf_offset.data(),
                    nb.leaf_ids.data(), 0, err, 64);
                stage2_packed<true>(s, nb, t, n_local, x, y, z, h, htol, Rker2, hc);
                nb.leaf_stride = 0;
                nb.row_stride  = 0;
                nb.cap_leaf    = choose(hc[8], nb.n_leaf_pairs, L, L);
                nb.cap_part    = choose(hc[9], nb.n_pairs, n_local, nm);
                break;
            }
            // stage 1: fixed-capacity leaf lists
            nb.leaf_ids.resize(std::max<u64>((u64) L * cl, 1));
            fill_u32(s, nb.counters.data() + 8, 0, 8);
            const u32 lpw = leaf_lanes();
            k_leaf_neigh_w<true, true><<<div_up(div_up(L, lpw), 256 / WAVE), 256, 0, s>>>(
                L, I, t.nodes.data(), Rkern, nb.leaf_cnt.data(), nb.leaf_offset.data(),
                nb.leaf_ids.data(), cl, err, lpw);
            max_u32(s, L, nb.leaf_cnt.data(), nb.counters.data() + 8);
            exclusive_scan_u32(s, nb.leaf_cnt.data(), nb.psum.data(), L, nb.counters.data() + 10,
                               nb.scratch.data());
            mark(L, nb.leaf_cnt.data(), cl, 12);
            const bool resolve_first = nb.leaf_overflowed;
            if (resolve_first) {
                SHAMX_HIP(hipMemcpyAsync(hc, nb.counters.data(), 16 * sizeof(u32),
                                         hipMemcpyDeviceToHost, s));
                SHAMX_HIP(hipStreamSynchronize(s));
                tree_check(t);
                if (hc[4]) throw std::runtime_error("neighbour cache: tree traversal stack overflow");
                if (hc[12]) resolve_leaves(cl, hc[12]);
            }
            const u32 leaf_cap = resolve_first ? 0u : cl; // lists past cl are truncated
            if ((u64) nm * cp > 0xffffffffull || packed_rows_forced()) {
                // particle rows packed (count, scan, fill)
                if (!resolve_first) {
                    SHAMX_HIP(hipMemcpyAsync(hc, nb.counters.data(), 16 * sizeof(u32),
                                             hipMemcpyDeviceToHost, s));
                    SHAMX_HIP(hipStreamSynchronize(s));
                    tree_check(t);
                    if (hc[4])
                        throw std::runtime_error("neighbour cache: tree traversal stack overflow");
                    if (hc[12]) resolve_leaves(cl, hc[12]);
                }
                const u32 n_leaf_pairs = hc[10];
                nb.leaf_overflowed     = hc[12] > 0;
                const u32 mxl          = hc[8];
                stage2_packed<true>(s, nb, t, n_local, x, y, z, h, htol, Rker2, hc, 0);
                nb.n_leaf_pairs = n_leaf_pairs;
                nb.leaf_stride  = cl;
                nb.row_stride   = 0;
                nb.cap_leaf     = choose(mxl, n_leaf_pairs, L, L);
                nb.cap_part     = choose(hc[9], nb.n_pairs, n_local, nm);
                break;
            }
            // stage 2: fixed-capacity particle rows
            nb.ids.resize(std::max<u64>((u64) nm * cp, 1));
            fill_u32(s, nb.cnt.data(), 0, nm); // ghost rows stay empty
            launch_group<true, true, true>(s, gblocks, batch_boxes(s, nb, t, gblocks, htol, Rker2),
                n_local, L, I, t.nodes.data(), t.sort_index.data(), nb.ps.data(),
                nb.owner.data(), nb.leaf_cnt.data(), nb.leaf_offset.data(), nb.leaf_ids.data(),
                nb.item_leaf.data(), nb.bstart.data(), nb.counters.data() + 7, htol, Rker2,
                nb.cnt.data(), nb.offset.data(), nb.ids.data(), cp, leaf_cap);
            k_part_neigh<true, true, true><<<div_up(n_local, 128), 128, 0, s>>>(
                n_local, nb.mism.data(), mism_len, I, t.nodes.data(), t.sort_index.data(),
                nb.owner.data(), nb.leaf_cnt.data(), nb.leaf_offset.data(), nb.leaf_ids.data(), x,
                y, z, h, nb.inv.data(), htol, Rker2, nb.cnt.data(), nb.offset.data(),
                nb.ids.data(), cp, leaf_cap);
            max_u32(s, nm, nb.cnt.data(), nb.counters.data() + 9);
            exclusive_scan_u32(s, nb.cnt.data(), nb.psum.data(), nm, nb.counters.data() + 11,
                               nb.scratch.data());
            SHAMX_HIP(hipMemcpyAsync(hc, nb.counters.data(), 16 * sizeof(u32),
                                     hipMemcpyDeviceToHost, s));
            SHAMX_HIP(hipStreamSynchronize(s));
            tree_check(t);
            check_walk_flags(hc[4]);
            if (!resolve_first && hc[12]) {
                // leaf lists overflowed for the first time: stage 2 read truncated lists
                nb.leaf_overflowed = true;
                continue;
            }
            nb.leaf_overflowed = hc[12] > 0;
            nb.n_leaf_pairs    = hc[10];
            nb.n_pairs         = hc[11];
            nb.n_mismatch      = hc[5];
            nb.leaf_stride     = cl;
            nb.row_stride      = cp;
            const u32 mxl = hc[8], mxp = hc[9];
            // particle rows past cp: list them (sorted positions -> particle ids) and fill each
            // again, one wave per row, into the overflow region of ids
            fill_u32(s, nb.counters.data() + 14, 0, 2);
            mark(nm, nb.cnt.data(), cp, 14);
            SHAMX_HIP(hipMemcpyAsync(hc + 14, nb.counters.data() + 14, 2 * sizeof(u32),
                                     hipMemcpyDeviceToHost, s));
            SHAMX_HIP(hipStreamSynchronize(s));
            if (hc[14]) {
                const u64 base = (u64) nm * cp;
                if (base + hc[15] > 0xffffffffull)
                    throw std::runtime_error("neighbour cache: rows exceed u32 entry indices");
                nb.ids.resize_keep(base + hc[15], s);
                k_over_list<<<div_up(nm, 256), 256, 0, s>>>(nm, nb.oflag.data(), nb.ofpos.data(),
                                                             nb.opos.data(), (u32) base,
                                                             t.sort_index.data(), nb.olist.data(),
                                                             nb.offset.data());
                k_group_neigh<true, true, false, true>
                    <<<div_up(hc[14], (u32) GRP_WAVES), GRP_WAVES * WAVE, 0, s>>>(
                        n_local, L, I, t.nodes.data(), t.sort_index.data(), nb.ps.data(),
                        nb.owner.data(), nb.leaf_cnt.data(), nb.leaf_offset.data(),
                        nb.leaf_ids.data(), nb.olist.data(), nullptr, nb.counters.data() + 14,
                        htol, Rker2, nullptr, nb.offset.data(), nb.ids.data(), 0, 0,
                        nb.inv.data());
            }
            nb.cap_leaf = choose(mxl, nb.n_leaf_pairs, L, L);
            nb.cap_part = choose(mxp, nb.n_pairs, n_local, nm);
            break;
        }
        SHAMX_HIP(hipGetLastError());
    }

} // namespace shamx
