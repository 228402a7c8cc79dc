// SPDX-License-Identifier: CeCILL-2.1
//
// tree.hip -- compressed-leaf BVH build on the merged (local + ghost) particle set:
//   bbox with one-ulp outward widening   (BuildTrees.cpp:38-55)
//   30-bit Morton codes                  (MortonCodeSet.cpp:62-118, morton.hpp:28-230)
//   stable onesweep LSD radix sort        (replaces bitonicSort_updated_usm.cpp:284-396)
//   leaf reduction                        (reduction_alg.cpp:62-585, MortonReducedSet.cpp:24-73)
//   Karras 2012 radix tree                (KarrasRadixTree.cpp:45-180)
//   leaf AABB + max-h field, then the internal cells level by level from the deepest up,
//   one launch per depth (replaces the tree_depth = 32 full passes of
//   KarrasRadixTreeAABB.cpp:32-74 and KarrasRadixTreeField.hpp:133-169; min / max are exact
//   so the result is identical).  A bottom-up climb with arrival counters would need a
//   device-scope fence per step, which on a multi-XCD GPU writes back and invalidates L2.
//
// Node records: every cell (I internal cells then L leaves) owns one 64-byte record
//   [0..2] aabb min, [3..5] aabb max, [6] rint = max h * htol,
//   [7]    internal: (left total id, right total id); leaf: (first, end) in sorted order
// so a traversal step reads exactly one record.
#include "launch.hpp"
#include "tree_layout.hpp"

#include <cstdlib>
#include <mutex>

namespace shamx {

    namespace {
        // --------------------------------------------------------------------- bbox
        constexpr int BB_BLOCK = 256;
        constexpr u32 BB_MAX_BLOCKS = 1024;

        __device__ inline f64 next_up(f64 v) {
            if (v != v || v == __builtin_inf()) return v;
            if (v == 0.) return __longlong_as_double(1ll);
            i64 b = __double_as_longlong(v);
            return __longlong_as_double(v > 0 ? b + 1 : b - 1);
        }
        __device__ inline f64 next_down(f64 v) {
            if (v != v || v == -__builtin_inf()) return v;
            if (v == 0.) return __longlong_as_double((i64) 0x8000000000000001ull);
            i64 b = __double_as_longlong(v);
            return __longlong_as_double(v > 0 ? b - 1 : b + 1);
        }

        __global__ void __launch_bounds__(BB_BLOCK)
            k_bbox_partial(u32 n, const f64 *__restrict x, const f64 *__restrict y,
                           const f64 *__restrict z, f64 *__restrict part) {
            f64 v[6] = {__builtin_inf(), __builtin_inf(), __builtin_inf(),
                        -__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
            for (u32 i = blockIdx.x * BB_BLOCK + threadIdx.x; i < n; i += gridDim.x * BB_BLOCK) {
                f64 a = x[i], b = y[i], c = z[i];
                v[0] = fmin(v[0], a);
                v[1] = fmin(v[1], b);
                v[2] = fmin(v[2], c);
                v[3] = fmax(v[3], a);
                v[4] = fmax(v[4], b);
                v[5] = fmax(v[5], c);
            }
            __shared__ f64 lds[6][BB_BLOCK / WAVE];
            int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
#pragma unroll
            for (int k = 0; k < 6; k++) {
                v[k] = (k < 3) ? wave_reduce(v[k], [](f64 a, f64 b) { return fmin(a, b); })
                               : wave_reduce(v[k], [](f64 a, f64 b) { return fmax(a, b); });
                if (lane == 0) lds[k][wid] = v[k];
            }
            __syncthreads();
            if (threadIdx.x < 6) {
                int k = threadIdx.x;
                f64 r = lds[k][0];
                for (int w = 1; w < BB_BLOCK / WAVE; w++)
                    r = (k < 3) ? fmin(r, lds[k][w]) : fmax(r, lds[k][w]);
                part[blockIdx.x * 6 + k] = r;
            }
        }

        // one wave per component: lanes stride over the block partials, then a wave reduction
        __global__ void __launch_bounds__(6 * WAVE)
            k_bbox_final(u32 nb, const f64 *__restrict part, f64 *__restrict bb) {
            const int k = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
            f64 r       = (k < 3) ? __builtin_inf() : -__builtin_inf();
            for (u32 b = lane; b < nb; b += WAVE)
                r = (k < 3) ? fmin(r, part[b * 6 + k]) : fmax(r, part[b * 6 + k]);
            if (k < 3) r = wave_reduce(r, [](f64 p, f64 q) { return fmin(p, q); });
            else r = wave_reduce(r, [](f64 p, f64 q) { return fmax(p, q); });
            if (lane == 0) bb[k] = (k < 3) ? next_down(r) : next_up(r);
        }

        // --------------------------------------------------------------------- Morton
        __device__ inline u32 expand_bits_10(u32 v) { // bmi.hpp:42-51
            v &= 0x3ffU;
            v = (v | v << 16U) & 0x30000ffU;
            v = (v | v << 8U) & 0x300f00fU;
            v = (v | v << 4U) & 0x30c30c3U;
            v = (v | v << 2U) & 0x9249249U;
            return v;
        }
        __device__ inline u32 to_grid(f64 c, f64 lo, f64 hi) {
            c        = fmin(fmax(c, lo), hi);        // AABB::clamp_coord
            f64 fact = (hi - lo) / 1024.;            // CoordRangeTransform<u16_3,f64_3>
            u16 g    = (u16) ((c - lo) / fact);      // reverse_transform, static_cast<u16>
            return (u32) min(max(g, (u16) 0), (u16) 1023); // MortonConverter::to_morton_grid
        }
        __global__ void k_morton(u32 n, const f64 *__restrict x, const f64 *__restrict y,
                                 const f64 *__restrict z, const f64 *__restrict bb,
                                 u32 *__restrict codes) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            u32 gx = to_grid(x[i], bb[0], bb[3]);
            u32 gy = to_grid(y[i], bb[1], bb[4]);
            u32 gz = to_grid(z[i], bb[2], bb[5]);
            codes[i] = expand_bits_10(gx) * 4 + expand_bits_10(gy) * 2 + expand_bits_10(gz);
        }

        // position inside its Morton cell, as a 30-bit Morton code of a 1024^3 sub-grid of the
        // cell (the cell from the same transform as to_grid): orders particles that share a
        // code spatially
        __device__ inline u32 to_subgrid(f64 c, f64 lo, f64 hi) {
            c        = fmin(fmax(c, lo), hi);
            f64 fact = (hi - lo) / 1024.;
            f64 t    = (c - lo) / fact;
            u32 g    = (u32) min(max((u16) t, (u16) 0), (u16) 1023);
            f64 fr   = t - (f64) g; // in [0, 1] (1 only where the coordinate was clamped)
            return (u32) fmin(fmax(fr * 1024., 0.), 1023.);
        }
        __global__ void k_fine_key(u32 n, const f64 *__restrict x, const f64 *__restrict y,
                                   const f64 *__restrict z, const f64 *__restrict bb,
                                   u32 *__restrict key) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            key[i] = expand_bits_10(to_subgrid(x[i], bb[0], bb[3])) * 4 +
                     expand_bits_10(to_subgrid(y[i], bb[1], bb[4])) * 2 +
                     expand_bits_10(to_subgrid(z[i], bb[2], bb[5]));
        }
        __global__ void k_gather_u32(u32 n, const u32 *__restrict src, const u32 *__restrict idx,
                                     u32 *__restrict dst) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i < n) dst[i] = src[idx[i]];
        }

        // --------------------------------------------------------------------- radix sort
        // Onesweep (one global histogram pass, then one scatter kernel per 8-bit digit with
        // decoupled look-back across tiles).  Ranks inside a tile are computed with wave-wide
        // peer masks built from ballots, in (wave, item, lane) order, so the sort is stable;
        // each tile is staged in LDS in digit order before it is written out.
        constexpr int RS_BLOCK = 256;
        constexpr int RS_WAVES = RS_BLOCK / WAVE;
        // keys per thread: 16 (8 measured slower at 2.5 M and 16.7 M keys: 206 vs 168 us and
        // 804 vs 607 us per sort; kept as a tuning knob)
        constexpr int RS_ITEMS       = 16;
        constexpr int RS_ITEMS_SMALL = 8;
        constexpr int RS_TILE_SMALL  = RS_BLOCK * RS_ITEMS_SMALL;
        constexpr int RS_BINS  = 256;
        constexpr u32 ST_AGG = 1u << 30, ST_PREFIX = 2u << 30, ST_VALUE = (1u << 30) - 1;
        constexpr u64 RS_SPIN_LIMIT = 200000000ull; // 2 s of the 100 MHz constant clock
        constexpr int RS_LOOKBACK   = 8; // predecessors read per look-back window

        // digit counts of every pass: each wave counts into its own LDS histograms (a quarter
        // of the contention on the atomics), 4 keys per load where the keys are 16-byte aligned
        __global__ void __launch_bounds__(RS_BLOCK)
            k_rs_hist(u32 n, const u32 *__restrict keys, u32 passes, u32 *__restrict hist,
                      u32 *__restrict status0, u32 n_status0) {
            __shared__ u32 h[RS_WAVES][4][RS_BINS];
            // the first scatter pass's status words start at zero
            for (u32 i = blockIdx.x * RS_BLOCK + threadIdx.x; i < n_status0;
                 i += gridDim.x * RS_BLOCK)
                status0[i] = 0u;
            for (int i = threadIdx.x; i < RS_WAVES * 4 * RS_BINS; i += RS_BLOCK)
                (&h[0][0][0])[i] = 0;
            __syncthreads();
            u32(*hw)[RS_BINS] = h[threadIdx.x / WAVE];
            auto count        = [&](u32 k) {
#pragma unroll
                for (u32 p = 0; p < 4; p++)
                    if (p < passes) atomicAdd(&hw[p][(k >> (8 * p)) & 0xffu], 1u);
            };
            const u32 stride = gridDim.x * RS_BLOCK;
            const u32 n4     = ((uintptr_t) keys & 15) == 0 ? n / 4 : 0;
            const uint4 *k4  = (const uint4 *) keys;
            for (u32 i = blockIdx.x * RS_BLOCK + threadIdx.x; i < n4; i += stride) {
                uint4 q = k4[i];
                count(q.x);
                count(q.y);
                count(q.z);
                count(q.w);
            }
            for (u32 i = 4 * n4 + blockIdx.x * RS_BLOCK + threadIdx.x; i < n; i += stride)
                count(keys[i]);
            __syncthreads();
            for (int i = threadIdx.x; i < (int) passes * RS_BINS; i += RS_BLOCK) {
                u32 v = 0;
#pragma unroll
                for (int w = 0; w < RS_WAVES; w++) v += (&h[w][0][0])[i];
                if (v) atomicAdd(&hist[i], v);
            }
        }

        // one wave per pass: exclusive scan of the 256 digit counts (4 per lane)
        __global__ void __launch_bounds__(WAVE) k_rs_hist_scan(u32 *__restrict hist) {
            u32 *h         = hist + blockIdx.x * RS_BINS;
            const u32 lane = threadIdx.x;
            u32 v[4], sum = 0;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                v[q] = h[4 * lane + q];
                sum += v[q];
            }
            u32 inc = sum;
#pragma unroll
            for (int o = 1; o < WAVE; o <<= 1) {
                u32 t = __shfl_up(inc, o, WAVE);
                if (lane >= (u32) o) inc += t;
            }
            u32 run = inc - sum;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                h[4 * lane + q] = run;
                run += v[q];
            }
        }

        __device__ inline u64 lanemask_lt() {
            u32 lane = __lane_id();
            return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        }

        // ITEMS keys per thread: tiles of RS_BLOCK * ITEMS keys.  Each tile ranks its keys by
        // digit (wave peer masks, (wave, item, lane) order: stable), publishes its digit counts,
        // stages its keys and values in LDS in digit order, looks back for the counts of the
        // earlier tiles, then writes each digit's run out with consecutive lanes on consecutive
        // addresses (whole 64-byte lines instead of one scattered 4-byte store per key).
        template<int ITEMS>
        __global__ void __launch_bounds__(RS_BLOCK)
            k_rs_onesweep(u32 n, u32 shift, const u32 *__restrict keys_in,
                          const u32 *__restrict vals_in, u32 *__restrict keys_out,
                          u32 *__restrict vals_out, const u32 *__restrict digit_base,
                          u32 *__restrict status, u32 *__restrict next_status,
                          u32 *__restrict tile_counter, int first, u32 nbits,
                          u32 *__restrict err) {
            constexpr u32 TILE = RS_BLOCK * ITEMS;
            __shared__ u32 s_tile;
            __shared__ u32 wave_hist[RS_WAVES][RS_BINS];
            __shared__ u32 s_start[RS_BINS]; // tile-local start of each digit's run
            __shared__ u32 s_dst[RS_BINS];   // output position of the run's key 0, minus s_start
            __shared__ u32 s_wsum[RS_WAVES];
            __shared__ u64 s_kv[TILE];       // (value << 32 | key) in digit order
            const u32 lane = __lane_id(), wid = threadIdx.x / WAVE;
            for (int i = threadIdx.x; i < RS_WAVES * RS_BINS; i += RS_BLOCK)
                (&wave_hist[0][0])[i] = 0;
            if (threadIdx.x == 0) s_tile = atomicAdd(tile_counter, 1u);
            __syncthreads();
            const u32 tile = s_tile;
            // the next pass's status words of this tile start at zero (read only after this
            // launch has ended)
            if (next_status) next_status[(u64) tile * RS_BINS + threadIdx.x] = 0u;
            const u64 base = (u64) tile * TILE + (u64) wid * (WAVE * ITEMS);
            const u64 lt   = lanemask_lt();

            // ranks inside the wave: peers (valid lanes holding the same digit) from one ballot
            // per digit bit; the lowest peer adds the group's count to the wave's digit counter
            // with one LDS atomic whose return value is the count of the earlier items (LDS
            // operations of a wave complete in issue order, so the atomics of all items are
            // issued back to back and waited for once), then broadcasts it to its peers
            u32 key[ITEMS], val[ITEMS], rank[ITEMS], src[ITEMS];
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                u64 idx    = base + (u64) k * WAVE + lane;
                bool valid = idx < n;
                key[k]     = valid ? keys_in[idx] : 0u;
                val[k]     = valid ? (first ? (u32) idx : vals_in[idx]) : 0u;
            }
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                u64 idx    = base + (u64) k * WAVE + lane;
                bool valid = idx < n;
                u32 d      = (key[k] >> shift) & 0xffu;
                u64 peers  = __ballot(valid);
#pragma unroll
                for (u32 b = 0; b < 8; b++) {
                    if (b >= nbits) break; // digit bits above the key's width are all 0
                    u64 bb = __ballot((d >> b) & 1u);
                    peers &= ((d >> b) & 1u) ? bb : ~bb;
                }
                u32 before = (u32) __popcll(peers & lt);
                u32 prev   = 0;
                if (valid && before == 0) prev = atomicAdd(&wave_hist[wid][d], (u32) __popcll(peers));
                rank[k] = prev + before;
                src[k]  = peers ? (u32) __builtin_ctzll(peers) : lane;
            }
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                // mine: the leader's count of its digit in the earlier items, or a peer's
                // position among its peers; the leader's value is added to its peers'
                u32 mine = rank[k];
                u32 lead = (u32) __shfl((int) mine, (int) src[k], WAVE);
                rank[k]  = (lane == src[k]) ? mine : lead + mine;
            }
            __syncthreads();

            // one thread per digit (RS_BLOCK == RS_BINS): offsets of the waves inside the
            // digit's run, the tile's count published at once, then the runs' starts
            const u32 d = threadIdx.x;
            u32 run     = 0;
#pragma unroll
            for (int w = 0; w < RS_WAVES; w++) {
                u32 t           = wave_hist[w][d];
                wave_hist[w][d] = run;
                run += t;
            }
            u32 *st = status + (u64) tile * RS_BINS + d;
            __hip_atomic_store(st, (tile == 0 ? ST_PREFIX : ST_AGG) | run, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            u32 inc = run;
#pragma unroll
            for (int o = 1; o < WAVE; o <<= 1) {
                u32 t = __shfl_up(inc, o, WAVE);
                if (lane >= (u32) o) inc += t;
            }
            if (lane == WAVE - 1) s_wsum[wid] = inc;
            __syncthreads();
            u32 start = inc - run;
            for (u32 w = 0; w < wid; w++) start += s_wsum[w];
            s_start[d] = start;
            __syncthreads();

            // stage the tile in digit order (while earlier tiles finish publishing)
#pragma unroll
            for (int k = 0; k < ITEMS; k++) {
                u64 idx = base + (u64) k * WAVE + lane;
                if (idx < n) {
                    u32 dk = (key[k] >> shift) & 0xffu;
                    s_kv[s_start[dk] + wave_hist[wid][dk] + rank[k]] =
                        ((u64) val[k] << 32) | key[k];
                }
            }

            // decoupled look-back for the counts of digit d in the earlier tiles, RS_LOOKBACK
            // predecessors read at once (independent loads: one memory latency per window)
            u32 excl = 0;
            if (tile > 0) {
                i64 j  = (i64) tile - 1;
                u64 t0 = wall_clock64();
                while (j >= 0) {
                    u32 sv[RS_LOOKBACK];
#pragma unroll
                    for (int q = 0; q < RS_LOOKBACK; q++)
                        sv[q] = (j - q >= 0)
                                    ? __hip_atomic_load(status + (u64) (j - q) * RS_BINS + d,
                                                        __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT)
                                    : ST_PREFIX; // before tile 0: a prefix of 0
                    // take the window in order up to the first prefix or unpublished entry
                    int used = 0;
                    bool done = false, wait = false;
#pragma unroll
                    for (int q = 0; q < RS_LOOKBACK; q++) {
                        if (done || wait) continue;
                        const u32 flag = sv[q] & ~ST_VALUE;
                        if (flag == 0) {
                            wait = true;
                        } else {
                            excl += sv[q] & ST_VALUE;
                            used++;
                            done = flag == ST_PREFIX;
                        }
                    }
                    if (done) break;
                    j -= used;
                    // predecessors always publish; a wait this long means a bug: flag it and
                    // leave instead of hanging the device
                    if (wait && wall_clock64() - t0 > RS_SPIN_LIMIT) {
                        atomicOr(err, 1u);
                        break;
                    }
                }
                __hip_atomic_store(st, ST_PREFIX | (excl + run), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            s_dst[d] = digit_base[d] + excl - start;
            __syncthreads();

            // write the runs out: consecutive positions of the tile go to consecutive
            // addresses of their digit's output run
            const u64 t0    = (u64) tile * TILE;
            const u32 count = (u32) min((u64) TILE, (u64) n - t0);
#pragma unroll 4
            for (u32 i = threadIdx.x; i < count; i += RS_BLOCK) {
                u64 kv  = s_kv[i];
                u32 k   = (u32) kv;
                u32 dst = s_dst[(k >> shift) & 0xffu] + i;
                if (dst < n) {
                    keys_out[dst] = k;
                    vals_out[dst] = (u32) (kv >> 32);
                } else {
                    atomicOr(err, 2u);
                }
            }
        }

        // --------------------------------------------------------------------- reduction
        __device__ inline i32 clz32(u32 v) { return v ? __clz(v) : 32; }
        // karras_delta: shambackends/math.hpp:825-828
        __device__ inline i32 kdelta(i32 x, i32 y, u32 len, const u32 *__restrict m) {
            return ((u32) y > len - 1 || y < 0) ? -1 : clz32(m[x] ^ m[y]);
        }

        __global__ void k_split(u32 n, const u32 *__restrict m, u32 *__restrict s) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            s[i] = (i > 0) ? (m[i - 1] != m[i] ? 1u : 0u) : 1u;
        }

        // sycl_reduction_iteration (NEW_BEHAVIOR), reduction_alg.cpp:196-244.  The reference
        // walks from i to the nearest set flags on either side, which is linear in the run
        // length (runs of equal codes are long in clustered sets); here the flagged indices
        // are compacted first (F, with cnt[i] = flags before i and *total flags in all), so
        // the walks become lookups with the same results, including the wrapped u32 values
        // (-1, -2) the walks stop at when no flag is left.
        __global__ void k_reduc_iter(u32 n, const u32 *__restrict m, const u32 *__restrict s_in,
                                     const u32 *__restrict cnt, const u32 *__restrict F,
                                     const u32 *__restrict total, u32 *__restrict s_out) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            const u32 c = cnt[i], K = *total;
            u32 before1 = c >= 1 ? F[c - 1] : 0xFFFFFFFFu;
            u32 before2 = c >= 2 ? F[c - 2] : (c == 1 ? 0xFFFFFFFFu : 0xFFFFFFFEu);
            const u32 ci = c + (s_in[i] ? 1u : 0u);
            u32 next1    = ci < K ? F[ci] : n;
            i32 d0  = kdelta((i32) i, (i32) next1, n, m);
            i32 dm  = kdelta((i32) i, (i32) before1, n, m);
            i32 dmm = kdelta((i32) before1, (i32) before2, n, m);
            s_out[i] = (!(d0 < dm && dmm < dm) && s_in[i]) ? 1u : 0u;
        }

        // sycl_reduction_iteration on the compacted list of kept positions: F[0..K) are the
        // positions whose split flag is set (K on the device); keep[j] is F[j]'s new flag.  The
        // reference walks from F[j] to the nearest set flags on either side (reduction_alg.cpp:
        // 196-244): those are F[j - 1], F[j - 2] and F[j + 1], with the wrapped u32 values
        // (-1, -2) the walks stop at when no flag is left before, and n when none is after.
        __global__ void k_reduc_keep(u32 n, const u32 *__restrict m, const u32 *__restrict F,
                                     const u32 *__restrict K_dev, u32 *__restrict keep) {
            const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
            const u32 K = *K_dev;
            if (j >= K) return;
            const u32 i       = F[j];
            const u32 before1 = j >= 1 ? F[j - 1] : 0xFFFFFFFFu;
            const u32 before2 = j >= 2 ? F[j - 2] : (j == 1 ? 0xFFFFFFFFu : 0xFFFFFFFEu);
            const u32 next1   = j + 1 < K ? F[j + 1] : n;
            i32 d0  = kdelta((i32) i, (i32) next1, n, m);
            i32 dm  = kdelta((i32) i, (i32) before1, n, m);
            i32 dmm = kdelta((i32) before1, (i32) before2, n, m);
            keep[j] = !(d0 < dm && dmm < dm) ? 1u : 0u;
        }
        // F_out[off[j]] = F_in[j] for the kept j < K (K on the device)
        __global__ void k_compact_list(const u32 *__restrict K_dev, const u32 *__restrict keep,
                                       const u32 *__restrict off, const u32 *__restrict F_in,
                                       u32 *__restrict F_out) {
            const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
            if (j < *K_dev && keep[j]) F_out[off[j]] = F_in[j];
        }

        __global__ void k_compact(u32 n, const u32 *__restrict flag, const u32 *__restrict off,
                                  u32 *__restrict map) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            if (flag[i]) map[off[i]] = i;
        }

        // update_morton_buf tail (reduction_alg.cpp:274-295) + remap (:564-585)
        __global__ void k_leaf_codes(u32 n, u32 L, u32 *__restrict map, const u32 *__restrict m,
                                     u32 *__restrict leaf_codes) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i < L) leaf_codes[i] = m[map[i]];
            if (i == 0) {
                map[L]     = n;
                map[L + 1] = 0;
            }
        }

        // --------------------------------------------------------------------- Karras
        __global__ void k_karras(u32 L, const u32 *__restrict morton, u32 *__restrict lchild,
                                 u32 *__restrict rchild, u8 *__restrict lflag,
                                 u8 *__restrict rflag, u32 *__restrict endrange,
                                 u32 *__restrict parent, f64 *__restrict nodes) {
            u32 I  = L - 1;
            u32 iu = blockIdx.x * blockDim.x + threadIdx.x;
            if (iu >= I) return;
            i32 i      = (i32) iu;
            auto D     = [&](i32 a, i32 b) { return kdelta(a, b, L, morton); };
            int ddelta = D(i, i + 1) - D(i, i - 1);
            int d      = (ddelta == 0) ? 0 : ((ddelta > 0) ? 1 : -1);
            int delta_min = D(i, i - d);
            int lmax      = 2;
            while (D(i, i + lmax * d) > delta_min) lmax *= 2;
            int l = 0, t = lmax / 2;
            while (t > 0) {
                if (D(i, i + (l + t) * d) > delta_min) l = l + t;
                t = t / 2;
            }
            u32 j          = (u32) (i + l * d);
            endrange[iu]   = j;
            int delta_node = D(i, (i32) j);
            int s          = 0;
            float div      = 2;
            t              = (int) ceilf((float) l / div);
            while (true) {
                int tmp_ = i + (s + t) * d;
                if (D(i, tmp_) > delta_node) s = s + t;
                if (t <= 1) break;
                div *= 2;
                t = (int) ceilf((float) l / div);
            }
            int gamma = i + s * d + min(d, 0);
            u8 lf     = (min(iu, j) == (u32) gamma) ? 1 : 0;
            u8 rf     = (max(iu, j) == (u32) (gamma + 1)) ? 1 : 0;
            lchild[iu] = (u32) gamma;
            rchild[iu] = (u32) (gamma + 1);
            lflag[iu]  = lf;
            rflag[iu]  = rf;
            u32 lt     = (u32) gamma + (lf ? I : 0u);
            u32 rt     = (u32) (gamma + 1) + (rf ? I : 0u);
            parent[lt] = iu;
            parent[rt] = iu;
            node_set_children(nodes, iu, lt, rt);
            if (iu == 0) parent[0] = NO_PARENT;
        }

        // leaf boxes + max h of the leaf's sorted range
        // Leaf boxes and max h: LEAF_LANES lanes per leaf stride through its particles, then
        // combine (min and max are exact, so the order does not matter).  Leaves of more than
        // BIG_LEAF particles (clustered sets put thousands in one cell) are listed instead and
        // done by k_big_leaf_boxes, one block each.
        constexpr u32 LEAF_LANES = 16, BIG_LEAF = 512, BIG_BLOCK = 256;
        struct Box7 {
            f64 lo0, lo1, lo2, hi0, hi1, hi2, hm;
            __device__ Box7() {
                const f64 mx = 1.7976931348623157e308;
                lo0 = lo1 = lo2 = mx;
                hi0 = hi1 = hi2 = hm = -mx;
            }
            __device__ void add(f64 a, f64 b, f64 c, f64 hh) {
                lo0 = fmin(lo0, a);
                lo1 = fmin(lo1, b);
                lo2 = fmin(lo2, c);
                hi0 = fmax(hi0, a);
                hi1 = fmax(hi1, b);
                hi2 = fmax(hi2, c);
                hm  = fmax(hm, hh);
            }
            __device__ void merge(const Box7 &o) {
                lo0 = fmin(lo0, o.lo0);
                lo1 = fmin(lo1, o.lo1);
                lo2 = fmin(lo2, o.lo2);
                hi0 = fmax(hi0, o.hi0);
                hi1 = fmax(hi1, o.hi1);
                hi2 = fmax(hi2, o.hi2);
                hm  = fmax(hm, o.hm);
            }
            template<int W>
            __device__ void xor_reduce() {
                for (int o = W / 2; o > 0; o >>= 1) {
                    Box7 t;
                    t.lo0 = __shfl_xor(lo0, o, W);
                    t.lo1 = __shfl_xor(lo1, o, W);
                    t.lo2 = __shfl_xor(lo2, o, W);
                    t.hi0 = __shfl_xor(hi0, o, W);
                    t.hi1 = __shfl_xor(hi1, o, W);
                    t.hi2 = __shfl_xor(hi2, o, W);
                    t.hm  = __shfl_xor(hm, o, W);
                    merge(t);
                }
            }
            __device__ void store(f64 *nodes, u32 node, f64 htol, u32 s0, u32 s1) const {
                f64 *rec = nodes + (u64) node * NODE_F64;
                rec[0]   = lo0;
                rec[1]   = lo1;
                rec[2]   = lo2;
                rec[3]   = hi0;
                rec[4]   = hi1;
                rec[5]   = hi2;
                // tree field max then *= htol (Solver.cpp:1345-1352): x -> x*htol is monotone,
                // so scaling before the max-propagation gives the identical doubles
                rec[6] = hm * htol;
                node_set_range(nodes, node, s0, s1);
            }
        };

        __global__ void k_leaf_boxes(u32 L, u32 I, const u32 *__restrict map,
                                     const u32 *__restrict sort_index, const f64 *__restrict x,
                                     const f64 *__restrict y, const f64 *__restrict z,
                                     const f64 *__restrict h, f64 htol, f64 *__restrict nodes,
                                     u32 *__restrict big, u32 *__restrict n_big) {
            const u32 l  = (blockIdx.x * blockDim.x + threadIdx.x) / LEAF_LANES;
            const u32 lg = threadIdx.x % LEAF_LANES;
            const bool act = l < L;
            u32 s0 = act ? map[l] : 0u, s1 = act ? map[l + 1] : 0u;
            const bool is_big = s1 - s0 > BIG_LEAF; // uniform in the group
            if (is_big) {
                if (lg == 0) big[atomicAdd(n_big, 1u)] = l;
                s1 = s0; // nothing to do here
            }
            Box7 b;
            for (u32 s = s0 + lg; s < s1; s += LEAF_LANES) {
                u32 o = sort_index[s];
                b.add(x[o], y[o], z[o], h[o]);
            }
            b.xor_reduce<LEAF_LANES>();
            if (!act || is_big || lg != 0) return;
            b.store(nodes, I + l, htol, s0, s1);
        }

        // the listed big leaves, one block of BIG_BLOCK threads each; the grid strides through
        // the list (its length is read on the device)
        __global__ void __launch_bounds__(BIG_BLOCK)
            k_big_leaf_boxes(u32 I, const u32 *__restrict map, const u32 *__restrict sort_index,
                             const f64 *__restrict x, const f64 *__restrict y,
                             const f64 *__restrict z, const f64 *__restrict h, f64 htol,
                             f64 *__restrict nodes, const u32 *__restrict big,
                             const u32 *__restrict n_big) {
            __shared__ f64 part[BIG_BLOCK / WAVE][7];
            const u32 nb = *n_big;
            for (u32 k = blockIdx.x; k < nb; k += gridDim.x) {
                const u32 l = big[k], s0 = map[l], s1 = map[l + 1];
                Box7 b;
                for (u32 s = s0 + threadIdx.x; s < s1; s += BIG_BLOCK) {
                    u32 o = sort_index[s];
                    b.add(x[o], y[o], z[o], h[o]);
                }
                b.xor_reduce<WAVE>();
                if (threadIdx.x % WAVE == 0) {
                    f64 *p = part[threadIdx.x / WAVE];
                    p[0] = b.lo0, p[1] = b.lo1, p[2] = b.lo2, p[3] = b.hi0, p[4] = b.hi1;
                    p[5] = b.hi2, p[6] = b.hm;
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    for (u32 w = 1; w < BIG_BLOCK / WAVE; w++) {
                        const f64 *p = part[w];
                        b.lo0 = fmin(b.lo0, p[0]);
                        b.lo1 = fmin(b.lo1, p[1]);
                        b.lo2 = fmin(b.lo2, p[2]);
                        b.hi0 = fmax(b.hi0, p[3]);
                        b.hi1 = fmax(b.hi1, p[4]);
                        b.hi2 = fmax(b.hi2, p[5]);
                        b.hm  = fmax(b.hm, p[6]);
                    }
                    b.store(nodes, I + l, htol, s0, s1);
                }
                __syncthreads();
            }
        }

        // Internal cells are merged in bands of levels, deepest first, over MERGE_LEVELS levels
        // with no host read of the actual depth: the leaf codes are distinct 32-bit codes (runs of
        // equal codes are merged into one leaf), every child's range shares a strictly longer
        // prefix than its parent's, so no internal cell lies deeper than 31.
        constexpr u32 MERGE_LEVELS = 32;
        // cycle guard of the parent walk
        constexpr u32 DEPTH_MAX = 1024;

        // depth of every internal node (root 0) by walking the parent links; err bit 0: broken
        // links, bit 2: a cell at depth >= MERGE_LEVELS
        __global__ void __launch_bounds__(256)
            k_node_depth(u32 I, const u32 *__restrict parent, u16 *__restrict depth,
                         u32 *__restrict err) {
            bool broken = false, deep = false;
            for (u32 i = blockIdx.x * 256 + threadIdx.x; i < I; i += gridDim.x * 256) {
                u32 d = 0, p = parent[i];
                while (p < I && d < DEPTH_MAX) {
                    d++;
                    p = parent[p];
                }
                broken |= p != NO_PARENT;
                deep |= d >= MERGE_LEVELS;
                depth[i] = (u16) min(d, DEPTH_MAX - 1);
            }
            const bool any_broken = __ballot(broken) != 0, any_deep = __ballot(deep) != 0;
            if (__lane_id() == 0 && (any_broken || any_deep))
                atomicOr(err, (any_broken ? 1u : 0u) | (any_deep ? 4u : 0u));
        }

        // box and h max of a cell: [0..2] min, [3..5] max, [6] max rint
        struct Rec7 {
            f64 v[7];
        };
        __device__ inline Rec7 rec7_load(const f64 *__restrict nodes, u32 c) {
            const f64 *p = nodes + (u64) c * NODE_F64;
            Rec7 r;
#pragma unroll
            for (int k = 0; k < 7; k++) r.v[k] = p[k];
            return r;
        }
        __device__ inline Rec7 rec7_merge(const Rec7 &a, const Rec7 &b) {
            Rec7 r;
#pragma unroll
            for (int k = 0; k < 3; k++) r.v[k] = fmin(a.v[k], b.v[k]);
#pragma unroll
            for (int k = 3; k < 7; k++) r.v[k] = fmax(a.v[k], b.v[k]);
            return r;
        }
        // the record of cell c as merged from the complete cells rem levels below it (or from
        // leaves above that): cells at depth >= the band's end were merged by an earlier launch.
        // Only the children links of the cells on the way are read; this kernel never writes
        // them.  R bounds the recursion at compile time (rem <= R).
        template<int R>
        __device__ inline Rec7 band_subtree(const f64 *__restrict nodes, u32 I, u32 L, u32 c,
                                            u32 rem, bool &bad) {
            if constexpr (R == 0) {
                return rec7_load(nodes, c);
            } else {
                if (c >= I || rem == 0) return rec7_load(nodes, c);
                u32 lt, rt;
                node_children(nodes, c, lt, rt);
                if (lt >= I + L || rt >= I + L) {
                    bad = true;
                    return rec7_load(nodes, c);
                }
                return rec7_merge(band_subtree<R - 1>(nodes, I, L, lt, rem - 1, bad),
                                  band_subtree<R - 1>(nodes, I, L, rt, rem - 1, bad));
            }
        }

        // boxes and h max of the internal cells of depth [b0, b0 + BAND), each from its
        // descendants of depth b0 + BAND (merged by the previous launch) and the leaves above
        // them.  min / max are exact, so this equals any other merge order bit for bit.
        constexpr u32 BAND = 4;
        __global__ void k_merge_band(u32 I, u32 L, u32 b0, const u16 *__restrict depth,
                                     f64 *__restrict nodes, u32 *__restrict err) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= I) return;
            const u32 d = depth[i];
            if (d < b0 || d >= b0 + BAND) return;
            bool bad        = false;
            const Rec7 m    = band_subtree<BAND>(nodes, I, L, i, b0 + BAND - d, bad);
            if (bad) {
                atomicOr(err, 2u);
                return;
            }
            f64 *P = nodes + (u64) i * NODE_F64;
#pragma unroll
            for (int k = 0; k < 7; k++) P[k] = m.v[k];
        }
    } // namespace

    size_t bbox_scratch_bytes(u32) { return BB_MAX_BLOCKS * 6 * sizeof(f64); }

    void bbox_nextafter(hipStream_t s, u32 n, const f64 *x, const f64 *y, const f64 *z,
                        f64 *bbox6, void *scratch) {
        u32 nb = std::max<u32>(1, std::min<u32>(BB_MAX_BLOCKS, div_up(n, BB_BLOCK * 8)));
        k_bbox_partial<<<nb, BB_BLOCK, 0, s>>>(n, x, y, z, (f64 *) scratch);
        k_bbox_final<<<1, 6 * WAVE, 0, s>>>(nb, (const f64 *) scratch, bbox6);
        SHAMX_HIP(hipGetLastError());
    }

    void morton_codes(hipStream_t s, u32 n, const f64 *x, const f64 *y, const f64 *z,
                      const f64 *bbox6, u32 *codes) {
        if (n == 0) return;
        k_morton<<<div_up(n, 256), 256, 0, s>>>(n, x, y, z, bbox6, codes);
        SHAMX_HIP(hipGetLastError());
    }

    namespace {
        // scratch of radix_sort_pairs for n keys: 4 digit histograms, 8 counters / error
        // words, then 4 passes of per-tile digit status words
        constexpr size_t rs_scratch_bytes(u32 n) {
            // tiles counted at the smaller tile size: enough status words for either
            size_t nt = ((size_t) (n > 0 ? n : 1) + RS_TILE_SMALL - 1) / RS_TILE_SMALL;
            return (4 * RS_BINS + 8 + 4 * nt * RS_BINS) * sizeof(u32);
        }
    } // namespace

    size_t radix_sort_scratch_bytes(u32 n) { return rs_scratch_bytes(n); }

    void radix_sort_pairs(hipStream_t s, u32 n, u32 key_bits, u32 *keys, u32 *vals,
                          u32 *keys_alt, u32 *vals_alt, void *scratch, bool vals_given) {
        if (n == 0) return;
        if (key_bits == 0 || key_bits > 32) throw std::invalid_argument("radix sort key_bits");
        if (n >= (1u << 30)) throw std::invalid_argument("radix sort: n must be below 2^30");
        // SHAMX_SORT_ITEMS = 8 or 16 forces the keys per thread (a test and tuning knob)
        const char *ie   = std::getenv("SHAMX_SORT_ITEMS");
        const int items  = ie ? std::atoi(ie) : 0;
        const bool small = items == RS_ITEMS_SMALL;
        u32 passes   = div_up(key_bits, 8);
        u32 nt       = div_up(n, RS_BLOCK * (small ? RS_ITEMS_SMALL : RS_ITEMS));
        u32 *hist    = (u32 *) scratch;
        u32 *ctr     = hist + 4 * RS_BINS;
        u32 *err     = ctr + 4;
        u32 *status  = ctr + 8;
        // digit counts and tile counters start at zero; the status words of pass 0 are zeroed
        // by the histogram kernel, those of pass p + 1 by pass p's tiles
        fill_u32(s, hist, 0, 4 * RS_BINS + 8);
        u32 hb = std::max<u32>(1, std::min<u32>(1024, div_up(n, RS_BLOCK * 16)));
        k_rs_hist<<<hb, RS_BLOCK, 0, s>>>(n, keys, passes, hist, status, nt * RS_BINS);
        k_rs_hist_scan<<<passes, WAVE, 0, s>>>(hist);
        u32 *kin = keys, *vin = vals, *kout = keys_alt, *vout = vals_alt;
        for (u32 p = 0; p < passes; p++) {
            u32 *st      = status + (u64) p * nt * RS_BINS;
            u32 *st_next = p + 1 < passes ? st + (u64) nt * RS_BINS : nullptr;
            const u32 nbits = std::min<u32>(8, key_bits - 8 * p);
            const int first = (p == 0 && !vals_given) ? 1 : 0;
            if (small)
                k_rs_onesweep<RS_ITEMS_SMALL><<<nt, RS_BLOCK, 0, s>>>(
                    n, 8 * p, kin, vin, kout, vout, hist + p * RS_BINS, st, st_next, ctr + p,
                    first, nbits, err);
            else
                k_rs_onesweep<RS_ITEMS><<<nt, RS_BLOCK, 0, s>>>(
                    n, 8 * p, kin, vin, kout, vout, hist + p * RS_BINS, st, st_next, ctr + p,
                    first, nbits, err);
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
        if (kin != keys) {
            SHAMX_HIP(hipMemcpyAsync(keys, kin, n * sizeof(u32), hipMemcpyDeviceToDevice, s));
            SHAMX_HIP(hipMemcpyAsync(vals, vin, n * sizeof(u32), hipMemcpyDeviceToDevice, s));
        }
        SHAMX_HIP(hipGetLastError());
    }

    // ---- the sort without allocation: its buffers live in one file-scope device array
    // (one copy per GPU, held until the process ends) whose address is looked up once per
    // device.  Layout: the alternate keys, the alternate values (NOALLOC_SORT_MAX_KEYS each),
    // then the scratch of radix_sort_pairs for that many keys.
    namespace {
        constexpr size_t NOALLOC_ALT_BYTES = (size_t) NOALLOC_SORT_MAX_KEYS * sizeof(u32);
        constexpr size_t NOALLOC_SCRATCH_BYTES =
            2 * NOALLOC_ALT_BYTES + ((rs_scratch_bytes(NOALLOC_SORT_MAX_KEYS) + 255) / 256) * 256;
        static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

        constexpr int NOALLOC_MAX_DEVICES = 64;
        std::mutex g_noalloc_mutex;
        void *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

        // device address of `scratch` on the current device (the host-side name is only a
        // stand-in for it)
        unsigned char *noalloc_scratch_ptr() {
            int dev = -1;
            SHAMX_HIP(hipGetDevice(&dev));
            if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
                throw std::runtime_error("shamx_sort_noalloc: device id out of range");
            std::lock_guard<std::mutex> lock(g_noalloc_mutex);
            if (!g_noalloc_ptr[dev]) {
                void *p = nullptr;
                SHAMX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
                if (!p) throw std::runtime_error("shamx_sort_noalloc: no address for scratch");
                g_noalloc_ptr[dev] = p;
            }
            return (unsigned char *) g_noalloc_ptr[dev];
        }
    } // namespace

    size_t radix_sort_noalloc_bytes() { return NOALLOC_SCRATCH_BYTES; }

    void radix_sort_noalloc(hipStream_t s, u32 n, u32 key_bits, u32 *keys, u32 *perm) {
        if (n > NOALLOC_SORT_MAX_KEYS)
            throw std::invalid_argument("shamx_sort_noalloc: n exceeds the maximum of 4194304 "
                                        "keys of the library's static sort buffer");
        if (key_bits == 0 || key_bits > 32)
            throw std::invalid_argument("shamx_sort_noalloc: key_bits must be in [1, 32]");
        if (n == 0) return;
        unsigned char *base = noalloc_scratch_ptr();
        u32 *alt_k = (u32 *) base;
        u32 *alt_v = (u32 *) (base + NOALLOC_ALT_BYTES);
        radix_sort_pairs(s, n, key_bits, keys, perm, alt_k, alt_v, base + 2 * NOALLOC_ALT_BYTES);
    }

    void tree_stage_codes(hipStream_t s, TreeBuffers &t, u32 n, const f64 *x, const f64 *y,
                          const f64 *z) {
        if (n == 0) throw std::invalid_argument("obj_cnt is 0, cannot build a CompressedLeafBVH");
        t.n_obj = n;
        size_t scratch = std::max({bbox_scratch_bytes(n), radix_sort_scratch_bytes(n),
                                   scan_scratch_bytes(n)});
        t.scratch.resize(scratch);
        t.bbox6.resize(6);
        t.counters.resize(12);
        for (auto *a : {&t.codes, &t.sorted, &t.sort_index, &t.keys_alt, &t.vals_alt, &t.split_a,
                        &t.split_b})
            a->resize(n);
        t.reduc_map.resize((size_t) n + 2);
        fill_u32(s, t.counters.data(), 0, 8);
        bbox_nextafter(s, n, x, y, z, t.bbox6.data(), t.scratch.data());
        morton_codes(s, n, x, y, z, t.bbox6.data(), t.codes.data());
        SHAMX_HIP(hipMemcpyAsync(t.sorted.data(), t.codes.data(), n * sizeof(u32),
                                 hipMemcpyDeviceToDevice, s));
    }

    u32 *radix_sort_error_word(void *scratch) { return (u32 *) scratch + 4 * RS_BINS + 4; }

    void tree_stage_sort(hipStream_t s, TreeBuffers &t, const f64 *x, const f64 *y,
                         const f64 *z) {
        const u32 n = t.n_obj;
        if (t.fine_ties) {
            // runs of equal codes ordered by the position inside the cell: a stable sort by
            // the sub-cell code, then a stable sort of that order by the code
            k_fine_key<<<div_up(n, 256), 256, 0, s>>>(n, x, y, z, t.bbox6.data(),
                                                       t.split_a.data());
            radix_sort_pairs(s, n, 30, t.split_a.data(), t.split_b.data(), t.keys_alt.data(),
                             t.vals_alt.data(), t.scratch.data(), false);
            SHAMX_HIP(hipMemcpyAsync(t.counters.data() + 3,
                                     radix_sort_error_word(t.scratch.data()), sizeof(u32),
                                     hipMemcpyDeviceToDevice, s));
            k_gather_u32<<<div_up(n, 256), 256, 0, s>>>(n, t.codes.data(), t.split_b.data(),
                                                         t.sorted.data());
            SHAMX_HIP(hipMemcpyAsync(t.sort_index.data(), t.split_b.data(), n * sizeof(u32),
                                     hipMemcpyDeviceToDevice, s));
        }
        radix_sort_pairs(s, n, 30, t.sorted.data(), t.sort_index.data(), t.keys_alt.data(),
                         t.vals_alt.data(), t.scratch.data(), t.fine_ties);
        // keep the sort's error word before the scratch is reused by the scans
        SHAMX_HIP(hipMemcpyAsync(t.counters.data() + 2, radix_sort_error_word(t.scratch.data()),
                                 sizeof(u32), hipMemcpyDeviceToDevice, s));
    }

    void tree_stage_karras(hipStream_t s, TreeBuffers &t, u32 reduction_level) {
        u32 n = t.n_obj;
        // leaf reduction (reduction_alg.cpp:413-456) on the compacted list of split positions:
        // each iteration tests the positions still kept and compacts them, so it costs the
        // number of kept positions, not n (the lengths stay on the device: no host round trip)
        u32 g = div_up(n, 256);
        k_split<<<g, 256, 0, s>>>(n, t.sorted.data(), t.split_a.data());
        t.red_cnt.resize(std::max<u32>(n, 1));
        t.red_pos.resize(std::max<u32>(n, 1));
        u32 *distinct = t.counters.data() + 1; // distinct codes (positions where codes change)
        u32 *L_dev    = t.counters.data() + 0;
        u32 *K_tmp[2] = {t.counters.data() + 8, t.counters.data() + 9};
        exclusive_scan_u32(s, t.split_a.data(), t.red_cnt.data(), n, distinct, t.scratch.data());
        u32 *F_in = reduction_level > 0 ? t.red_pos.data() : t.reduc_map.data();
        k_compact<<<g, 256, 0, s>>>(n, t.split_a.data(), t.red_cnt.data(), F_in);
        const u32 *K_in = distinct;
        if (reduction_level == 0)
            SHAMX_HIP(hipMemcpyAsync(L_dev, distinct, sizeof(u32), hipMemcpyDeviceToDevice, s));
        for (u32 it = 1; it <= reduction_level; it++) {
            const bool last = it == reduction_level;
            u32 *F_out = last ? t.reduc_map.data()
                              : (F_in == t.red_pos.data() ? t.split_b.data() : t.red_pos.data());
            u32 *K_out = last ? L_dev : K_tmp[it % 2];
            k_reduc_keep<<<g, 256, 0, s>>>(n, t.sorted.data(), F_in, K_in, t.split_a.data());
            exclusive_scan_u32_devn(s, t.split_a.data(), t.red_cnt.data(), n, K_in, K_out,
                                    t.scratch.data());
            k_compact_list<<<g, 256, 0, s>>>(K_in, t.split_a.data(), t.red_cnt.data(), F_in,
                                             F_out);
            F_in = F_out;
            K_in = K_out;
        }
        u32 *hc = t.host.data();
        t.host.read_async(s, t.counters.data(), 4);
        SHAMX_HIP(hipStreamSynchronize(s));
        tree_check(t); // the previous build's, if nothing has read it since
        if (hc[2] | hc[3]) throw std::runtime_error("radix sort: look-back or scatter inconsistency");
        // the next builds order runs of equal codes spatially when codes repeat on average
        // more than 1.5 times (clustered sets whose cells hold many particles); unique codes
        // leave nothing to order
        if (reduction_level > 0)
            t.fine_ties = t.allow_fine_ties && (u64) hc[1] * 3 < (u64) n * 2;
        u32 L = hc[0];
        if (L == 0) throw std::runtime_error("0 leaf tree cannot exists");
        t.L = L;
        t.I = L - 1;
        t.leaf_codes.resize(L);
        k_leaf_codes<<<div_up(L, 256), 256, 0, s>>>(n, L, t.reduc_map.data(), t.sorted.data(),
                                                   t.leaf_codes.data());
        u32 I = t.I;
        for (auto *a : {&t.lchild, &t.rchild, &t.endrange}) a->resize(std::max<u32>(I, 1));
        t.depth.resize(std::max<u32>(I, 1));
        t.lflag.resize(std::max<u32>(I, 1));
        t.rflag.resize(std::max<u32>(I, 1));
        t.parent.resize((size_t) I + L);
        t.nodes.resize((size_t) (I + L) * NODE_F64);
        fill_u32(s, t.parent.data(), NO_PARENT, (size_t) I + L);
        if (I > 0) {
            k_karras<<<div_up(I, 256), 256, 0, s>>>(L, t.leaf_codes.data(), t.lchild.data(),
                                                   t.rchild.data(), t.lflag.data(),
                                                   t.rflag.data(), t.endrange.data(),
                                                   t.parent.data(), t.nodes.data());
        } else {
            fill_u32(s, t.parent.data(), NO_PARENT, 1);
        }
        SHAMX_HIP(hipGetLastError());
    }

    void tree_stage_aabb(hipStream_t s, TreeBuffers &t, const f64 *x, const f64 *y,
                         const f64 *z, const f64 *h, f64 htol) {
        u32 L = t.L, I = t.I;
        if (L == 0) throw std::invalid_argument("tree_stage_aabb: build the tree first");
        t.big_leaves.resize(std::max<u32>(L, 1));
        u32 *n_big = t.counters.data() + 7;
        fill_u32(s, n_big, 0, 1);
        k_leaf_boxes<<<div_up((u64) L * LEAF_LANES, 256), 256, 0, s>>>(
            L, I, t.reduc_map.data(), t.sort_index.data(), x, y, z, h, htol, t.nodes.data(),
            t.big_leaves.data(), n_big);
        k_big_leaf_boxes<<<std::min<u32>(L, 1024), BIG_BLOCK, 0, s>>>(
            I, t.reduc_map.data(), t.sort_index.data(), x, y, z, h, htol, t.nodes.data(),
            t.big_leaves.data(), n_big);
        if (I > 0) {
            // internal boxes band by band, deepest first (the kernel boundary between two
            // bands orders the writes of one band before the reads of the next)
            u32 *err = t.counters.data() + 5;
            k_node_depth<<<std::min<u32>(1024, div_up(I, 256)), 256, 0, s>>>(
                I, t.parent.data(), t.depth.data(), err);
            static_assert(MERGE_LEVELS % BAND == 0, "bands tile the merge levels");
            for (u32 b0 = MERGE_LEVELS; b0 > 0;) {
                b0 -= BAND;
                k_merge_band<<<div_up(I, 256), 256, 0, s>>>(I, L, b0, t.depth.data(),
                                                           t.nodes.data(), err);
            }
        }
        // the flags reach the host with the next synchronisation of the stream (tree_check)
        t.host.read_async(s, t.counters.data() + 5, 1, 5);
        t.err_pending = true;
        SHAMX_HIP(hipGetLastError());
    }

    void tree_check(TreeBuffers &t) {
        if (!t.err_pending) return;
        t.err_pending = false;
        const u32 e   = t.host[5];
        if (e & 1u) throw std::runtime_error("tree build: broken parent links");
        if (e & 2u) throw std::runtime_error("tree build: child index out of range");
        if (e & 4u) throw std::runtime_error("tree build: internal cell deeper than 31");
    }

    void tree_stage_finish(hipStream_t s, TreeBuffers &t, const f64 *x, const f64 *y,
                           const f64 *z, const f64 *h, u32 reduction_level, f64 htol) {
        tree_stage_karras(s, t, reduction_level);
        tree_stage_aabb(s, t, x, y, z, h, htol);
    }

    void tree_build(hipStream_t s, TreeBuffers &t, u32 n, const f64 *x, const f64 *y,
                    const f64 *z, const f64 *h, u32 reduction_level, f64 htol) {
        tree_stage_codes(s, t, n, x, y, z);
        tree_stage_sort(s, t, x, y, z);
        tree_stage_finish(s, t, x, y, z, h, reduction_level, htol);
    }

} // namespace shamx
