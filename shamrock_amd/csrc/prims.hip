// SPDX-License-Identifier: CeCILL-2.1
//
// prims.hip -- device-wide primitives: fills and a two-kernel exclusive scan (one kernel for a
// single tile, three past FUSED_TILES tiles).
// Replaces shamalgs::numeric::scan_exclusive (shamalgs/src/details/numeric/numeric.cpp:47-59,
// decoupled look-back implementation scanDecoupledLookback.hpp:391) for the u32 CSR offsets
// and stream compactions of the step.  Scans are exact integer sums, so any scan order gives
// the same result.
#include "launch.hpp"

namespace shamx {

    namespace {
        constexpr int SCAN_BLOCK = 256;
        constexpr int SCAN_ITEMS = 16;
        constexpr int SCAN_TILE  = SCAN_BLOCK * SCAN_ITEMS; // 4096

        template<class T>
        __global__ void k_fill(T *__restrict p, T v, size_t n) {
            size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
            size_t stride = (size_t) gridDim.x * blockDim.x;
            for (; i < n; i += stride) p[i] = v;
        }

        // block-wide exclusive scan of one value per thread; returns the block total
        __device__ inline u32 block_exclusive_scan(u32 v, u32 &excl, u32 *lds_waves) {
            const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
            u32 inc = v;
            for (int o = 1; o < WAVE; o <<= 1) {
                u32 t = __shfl_up(inc, o, WAVE);
                if (lane >= o) inc += t;
            }
            if (lane == WAVE - 1) lds_waves[wid] = inc;
            __syncthreads();
            u32 wave_base = 0, total = 0;
            for (int w = 0; w < SCAN_BLOCK / WAVE; w++) {
                u32 t = lds_waves[w];
                if (w < wid) wave_base += t;
                total += t;
            }
            __syncthreads();
            excl = wave_base + inc - v;
            return total;
        }

        __global__ void __launch_bounds__(SCAN_BLOCK)
            k_tile_sums(const u32 *__restrict in, u32 n, u32 *__restrict tile_sums,
                        const u32 *__restrict n_dev) {
            __shared__ u32 lds_waves[SCAN_BLOCK / WAVE];
            if (n_dev) n = min(n, *n_dev); // the length known on the device only
            u64 base = (u64) blockIdx.x * SCAN_TILE;
            u32 s = 0;
#pragma unroll
            for (int k = 0; k < SCAN_ITEMS; k++) {
                u64 i = base + (u64) k * SCAN_BLOCK + threadIdx.x;
                if (i < n) s += in[i];
            }
            u32 ex;
            u32 tot = block_exclusive_scan(s, ex, lds_waves);
            if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
        }

        // single block: exclusive scan of the tile sums in place, total to *total
        __global__ void __launch_bounds__(SCAN_BLOCK)
            k_scan_tiles(u32 *__restrict tile_sums, u32 nt, u32 *__restrict total) {
            __shared__ u32 lds_waves[SCAN_BLOCK / WAVE];
            u32 carry = 0;
            for (u32 b = 0; b < nt; b += SCAN_BLOCK) {
                u32 i = b + threadIdx.x;
                u32 v = (i < nt) ? tile_sums[i] : 0;
                u32 ex;
                u32 tot = block_exclusive_scan(v, ex, lds_waves);
                if (i < nt) tile_sums[i] = carry + ex;
                carry += tot;
            }
            if (threadIdx.x == 0 && total) *total = carry;
        }

        __global__ void __launch_bounds__(SCAN_BLOCK)
            k_scan_tiles_apply(const u32 *__restrict in, u32 *__restrict out, u32 n,
                               const u32 *__restrict tile_base, const u32 *__restrict n_dev) {
            __shared__ u32 tile[SCAN_TILE];
            __shared__ u32 lds_waves[SCAN_BLOCK / WAVE];
            if (n_dev) n = min(n, *n_dev);
            u64 base = (u64) blockIdx.x * SCAN_TILE;
            if (base >= n) return; // block-uniform
#pragma unroll
            for (int k = 0; k < SCAN_ITEMS; k++) {
                u64 i                               = base + (u64) k * SCAN_BLOCK + threadIdx.x;
                tile[k * SCAN_BLOCK + threadIdx.x] = (i < n) ? in[i] : 0;
            }
            __syncthreads();
            u32 loc[SCAN_ITEMS];
            u32 s = 0;
#pragma unroll
            for (int k = 0; k < SCAN_ITEMS; k++) {
                loc[k] = tile[threadIdx.x * SCAN_ITEMS + k];
                s += loc[k];
            }
            u32 ex;
            block_exclusive_scan(s, ex, lds_waves);
            u32 run = tile_base[blockIdx.x] + ex;
#pragma unroll
            for (int k = 0; k < SCAN_ITEMS; k++) {
                tile[threadIdx.x * SCAN_ITEMS + k] = run;
                run += loc[k];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < SCAN_ITEMS; k++) {
                u64 i = base + (u64) k * SCAN_BLOCK + threadIdx.x;
                if (i < n) out[i] = tile[k * SCAN_BLOCK + threadIdx.x];
            }
        }

        // k_scan_tiles and k_scan_tiles_apply in one launch, for up to FUSED_TILES tiles: each
        // block sums the tile sums before its own (a few loads per thread from L2) instead of
        // reading a base scanned by a single-block launch.  Block 0 sums all of them and writes
        // the total.  tile_sums == nullptr: a single tile (n <= SCAN_TILE), no k_tile_sums.
        constexpr u32 FUSED_TILES = 4096;
        __global__ void __launch_bounds__(SCAN_BLOCK)
            k_scan_apply_fused(const u32 *__restrict in, u32 *__restrict out, u32 n,
                               const u32 *__restrict tile_sums, u32 nt, u32 *__restrict total,
                               const u32 *__restrict n_dev) {
            __shared__ u32 tile[SCAN_TILE];
            __shared__ u32 lds_waves[SCAN_BLOCK / WAVE];
            if (n_dev) n = min(n, *n_dev);
            u32 base = 0;
            if (tile_sums) {
                const u32 lim = blockIdx.x == 0 ? nt : blockIdx.x;
                u32 part = 0;
#pragma unroll 4
                for (u32 i = threadIdx.x; i < lim; i += SCAN_BLOCK) part += tile_sums[i];
                u32 ex;
                const u32 sum = block_exclusive_scan(part, ex, lds_waves);
                if (blockIdx.x == 0) {
                    if (threadIdx.x == 0 && total) *total = sum;
                } else
                    base = sum;
            }
            const u64 first = (u64) blockIdx.x * SCAN_TILE;
            if (first >= n) { // block-uniform
                if (!tile_sums && threadIdx.x == 0 && total) *total = 0;
                return;
            }
#pragma unroll
            for (int k = 0; k < SCAN_ITEMS; k++) {
                u64 i                               = first + (u64) k * SCAN_BLOCK + threadIdx.x;
                tile[k * SCAN_BLOCK + threadIdx.x] = (i < n) ? in[i] : 0;
            }
            __syncthreads();
            u32 loc[SCAN_ITEMS];
            u32 s = 0;
#pragma unroll
            for (int k = 0; k < SCAN_ITEMS; k++) {
                loc[k] = tile[threadIdx.x * SCAN_ITEMS + k];
                s += loc[k];
            }
            u32 ex;
            const u32 tile_total = block_exclusive_scan(s, ex, lds_waves);
            if (!tile_sums && threadIdx.x == 0 && total) *total = tile_total;
            u32 run = base + ex;
#pragma unroll
            for (int k = 0; k < SCAN_ITEMS; k++) {
                tile[threadIdx.x * SCAN_ITEMS + k] = run;
                run += loc[k];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < SCAN_ITEMS; k++) {
                u64 i = first + (u64) k * SCAN_BLOCK + threadIdx.x;
                if (i < n) out[i] = tile[k * SCAN_BLOCK + threadIdx.x];
            }
        }

        // the scan of in[0, n) with n <= n_max, n read from n_dev when it is not null
        void scan_impl(hipStream_t s, const u32 *in, u32 *out, u32 n_max, const u32 *n_dev,
                       u32 *total, void *scratch) {
            u32 nt         = div_up(n_max, SCAN_TILE);
            u32 *tile_sums = (u32 *) scratch;
            if (nt == 1) {
                k_scan_apply_fused<<<1, SCAN_BLOCK, 0, s>>>(in, out, n_max, nullptr, 1, total,
                                                            n_dev);
            } else if (nt <= FUSED_TILES) {
                k_tile_sums<<<nt, SCAN_BLOCK, 0, s>>>(in, n_max, tile_sums, n_dev);
                k_scan_apply_fused<<<nt, SCAN_BLOCK, 0, s>>>(in, out, n_max, tile_sums, nt,
                                                             total, n_dev);
            } else {
                k_tile_sums<<<nt, SCAN_BLOCK, 0, s>>>(in, n_max, tile_sums, n_dev);
                k_scan_tiles<<<1, SCAN_BLOCK, 0, s>>>(tile_sums, nt, total);
                k_scan_tiles_apply<<<nt, SCAN_BLOCK, 0, s>>>(in, out, n_max, tile_sums, n_dev);
            }
            SHAMX_HIP(hipGetLastError());
        }

        template<class T>
        void fill_impl(hipStream_t s, T *p, T v, size_t n) {
            if (n == 0) return;
            u32 blocks = (u32) std::min<size_t>(div_up(n, 256), 2048);
            k_fill<T><<<blocks, 256, 0, s>>>(p, v, n);
            SHAMX_HIP(hipGetLastError());
        }
    } // namespace

    void fill_u32(hipStream_t s, u32 *p, u32 v, size_t n) { fill_impl(s, p, v, n); }
    void fill_u64(hipStream_t s, u64 *p, u64 v, size_t n) { fill_impl(s, p, v, n); }
    void fill_f64(hipStream_t s, f64 *p, f64 v, size_t n) { fill_impl(s, p, v, n); }

    size_t scan_scratch_bytes(u32 n) { return (size_t) (div_up(n, SCAN_TILE) + 1) * sizeof(u32); }

    void exclusive_scan_u32(hipStream_t s, const u32 *in, u32 *out, u32 n, u32 *total,
                            void *scratch) {
        if (n == 0) {
            if (total) fill_u32(s, total, 0, 1);
            return;
        }
        scan_impl(s, in, out, n, nullptr, total, scratch);
    }

    void exclusive_scan_u32_devn(hipStream_t s, const u32 *in, u32 *out, u32 n_max,
                                 const u32 *n_dev, u32 *total, void *scratch) {
        if (n_max == 0) {
            if (total) fill_u32(s, total, 0, 1);
            return;
        }
        scan_impl(s, in, out, n_max, n_dev, total, scratch);
    }

} // namespace shamx
