// SPDX-License-Identifier: CeCILL-2.1
//
// sph.hip -- per-particle SPH kernels of the step, all SoA FP64 on HBM:
//   smoothing-length Newton sweep   IterateSmoothingLengthDensity.cpp:28-120, density.hpp:22-41
//   grad-h term Omega               ComputeOmega.cpp:25-74
//   equation of state               ComputeEos.cpp:152-247, 656-665, 726-798; eos.hpp:31-98
//   force + du/dt, constant AV      UpdateDerivs.cpp:88-288; forces.hpp:34-226; q_ab.hpp:36-39
//     fused with: a_old/du_old save (Solver.cpp:1667-1710), vsig (Solver.cpp:2716-2789) and
//     the per-particle CFL dt (ComputeCFLCourant.hpp, ComputeCFLForce.hpp:50-62)
//   leapfrog predictor + periodic wrap  Solver.cpp:390-524, integrators.cpp:206-236
//   leapfrog corrector              integrators.cpp:87-119, Solver.cpp:2580-2616
//   periodic images                 BasicSPHGhosts.cpp:260-579, Solver.cpp:1394-1633
// FP64 expressions keep the reference operation order (see common.hpp).  The mixed variant of
// the force loop computes r_ab in FP64, then the pair terms in FP32, and accumulates with
// Kahan compensation.
#include "launch.hpp"

#include <cstdlib>
#include <type_traits>

namespace shamx {

    namespace {
        constexpr int BLK = 256;

        __device__ inline void atomic_max_key(u64 *p, u64 v) { atomicMax((unsigned long long *) p, v); }
        __device__ inline void atomic_min_key(u64 *p, u64 v) { atomicMin((unsigned long long *) p, v); }
        __device__ inline u64 umax64(u64 a, u64 b) { return a > b ? a : b; }
        __device__ inline u64 umin64(u64 a, u64 b) { return a < b ? a : b; }

        // ------------------------------------------------------------- launch helpers
        // Orders the LDS accesses of the lanes of one wave around this point: LDS executes one
        // wave's operations in issue order, the fences keep the compiler from moving them across.
        __device__ inline void wave_lds_sync() {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }

        // x / y from the correctly rounded reciprocal ry = RN(1 / y): q = RN(x ry) is within one
        // ulp of x / y and the residual x - q y is exact (one FMA), so RN(q + (x - q y) ry) is
        // the correctly rounded quotient (Markstein's theorem, IBM J. Res. Dev. 34(1), 1990;
        // tests/test_division_identity.py checks it on 10^7 adversarial operand pairs).  Valid
        // while no operand, quotient or residual leaves the normal range; the records mark the
        // particles outside [2^-200, 2^200] (ry = 0), which take the IEEE division instead.
        __device__ inline f64 div_rcp(f64 x, f64 y, f64 ry) {
            f64 q = x * ry;
            f64 e = __builtin_fma(-q, y, x);
            return __builtin_fma(e, ry, q);
        }

        __device__ inline bool rcp_safe(f64 v) { return v > 0x1p-200 && v < 0x1p200; }

        // W, dW/dr and dW/dh of kernel K at one particle's h (sphkernels.hpp:2329-2343): the same
        // expressions as Kern<K, f64>, the divisions by h, h^3 and h^4 through their
        // reciprocals.  FAST = false is the plain IEEE division path, for h outside
        // [2^-200, 2^200]^(1/4) or so (`fast` says which one applies).
        template<class K>
        struct KernAt {
            f64 h, h3, h4, rh, rh3, rh4;
            bool fast;
            __device__ explicit KernAt(f64 hh) : h(hh) {
                h3   = hh * hh * hh;
                h4   = hh * hh * hh * hh;
                fast = rcp_safe(hh) && rcp_safe(h3) && rcp_safe(h4);
                rh   = fast ? 1. / hh : 0.;
                rh3  = fast ? 1. / h3 : 0.;
                rh4  = fast ? 1. / h4 : 0.;
            }
            template<bool FAST>
            __device__ f64 W(f64 r) const {
                if constexpr (FAST)
                    return div_rcp(Kern<K, f64>::norm * K::f(div_rcp(r, h, rh)), h3, rh3);
                else
                    return Kern<K, f64>::W(r, h);
            }
            template<bool FAST>
            __device__ f64 dhW(f64 r) const {
                if constexpr (FAST) {
                    f64 q = div_rcp(r, h, rh);
                    return div_rcp(-(Kern<K, f64>::norm) * (3. * K::f(q) + q * K::df(q)), h4, rh4);
                } else
                    return Kern<K, f64>::dhW(r, h);
            }
        };

        // The dispatcher hands block b to XCD b % 8.  Renumbering so that every XCD walks one
        // contiguous run of logical blocks keeps particles that are close in sorted order, and
        // the neighbour records they share, in the same L2.
        __device__ inline u32 xcd_block(u32 b, u32 nb) {
            constexpr u32 NX = 8;
            u32 per = nb / NX, rem = nb % NX, x = b % NX, i = b / NX;
            return x * per + (x < rem ? x : rem) + i;
        }

        // block-wide reduction; every thread gets the result
        template<class T, class Op>
        __device__ inline T block_reduce(T v, Op op, T *lds) {
            v = wave_reduce(v, op);
            if (__lane_id() == 0) lds[threadIdx.x / WAVE] = v;
            __syncthreads();
            T r = lds[0];
            for (int w = 1; w < BLK / WAVE; w++) r = op(r, lds[w]);
            __syncthreads();
            return r;
        }

        // value held by lane L of this lane's group of G lanes (G = 4: DPP quad_perm
        // broadcast, G = 16: DPP row_newbcast)
        template<int G, int L>
        __device__ inline f64 group_lane(f64 v) {
            static_assert(G == 2 || G == 4 || G == 16, "DPP broadcasts cover quads and rows");
            // quad_perm selects within quads (pairs: lanes L and 2 + L), row_newbcast rows of 16
            constexpr int ctrl = (G == 16)  ? 0x150 + L
                                 : (G == 4) ? (L | (L << 2) | (L << 4) | (L << 6))
                                            : (L | (L << 2) | ((2 + L) << 4) | ((2 + L) << 6));
            i64 b  = __double_as_longlong(v);
            int lo = __builtin_amdgcn_mov_dpp((int) b, ctrl, 0xf, 0xf, false);
            int hi = __builtin_amdgcn_mov_dpp((int) (b >> 32), ctrl, 0xf, 0xf, false);
            return __longlong_as_double(((i64) hi << 32) | (u32) lo);
        }
        // acc + t[lane 0] + t[lane 1] + ... + t[lane G-1], left to right, in the first lane of
        // the group: the G terms of one iteration enter the sum in list order, one IEEE add
        // each, exactly as in a sequential loop over the row.  The first lane adds its own term
        // and then the broadcasts of the others (the other lanes carry sums that are not used:
        // only the group's first lane may read the result).  Must be reached by the whole group.
        template<int G, int L = 0>
        __device__ inline f64 ordered_fold(f64 acc, f64 t) {
            if constexpr (G == 1) {
                return acc + t;
            } else {
                acc = acc + (L == 0 ? t : group_lane<G, L>(t));
                if constexpr (L + 1 < G) return ordered_fold<G, L + 1>(acc, t);
                else return acc;
            }
        }
        // the same with two terms per list entry, added in the order t1, t2
        template<int G, int L = 0>
        __device__ inline f64 ordered_fold2(f64 acc, f64 t1, f64 t2) {
            if constexpr (G == 1) {
                acc = acc + t1;
                return acc + t2;
            } else {
                acc = acc + (L == 0 ? t1 : group_lane<G, L>(t1));
                acc = acc + (L == 0 ? t2 : group_lane<G, L>(t2));
                if constexpr (L + 1 < G) return ordered_fold2<G, L + 1>(acc, t1, t2);
                else return acc;
            }
        }

        template<int G>
        __device__ inline f64 group_sum(f64 v) {
            for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
            return v;
        }
        template<int G>
        __device__ inline f64 group_max(f64 v) {
            for (int o = G / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, G));
            return v;
        }

        // rows of the neighbour CSR processed by a kernel: the j-th processed particle has row
        // sel[j] (identity when null) and particle id outi[row] (identity when null)
        struct Rows {
            const u32 *sel, *outi, *cnt, *off, *ids;
        };

        // neighbour positions: sorted (x, y, z, h) records or the SoA arrays
        struct PosRec {
            const double4 *__restrict ps;
            __device__ double3 pos(u32 b) const {
                double4 q = ps[b];
                return make_double3(q.x, q.y, q.z);
            }
            __device__ double4 pos4(u32 b) const { return ps[b]; }
        };
        struct PosSoA {
            const f64 *__restrict x, *__restrict y, *__restrict z;
            __device__ double3 pos(u32 b) const { return make_double3(x[b], y[b], z[b]); }
        };

        __device__ inline void kahan_add(f32 &s, f32 &c, f32 v) {
            f32 y = v - c;
            f32 t = s + y;
            c     = (t - s) - y;
            s     = t;
        }

        // mixed-precision pair arithmetic: hardware reciprocal and square root (about 1 ulp),
        // kernel values from q = r / h evaluated as r * (1 / h) with 1 / h per particle
        __device__ inline f32 rcp_approx(f32 v) { return __builtin_amdgcn_rcpf(v); }
        __device__ inline f32 sqrt_approx(f32 v) { return __builtin_amdgcn_sqrtf(v); }
        template<class K>
        struct KernInv {
            static constexpr f32 norm = f32(K::norm_3d);
            // dW/dr and dW/dh (Kern::dW, Kern::dhW) with the powers of 1 / h precomputed
            __device__ static f32 dW(f32 r, f32 inv_h, f32 inv_h4) {
                return norm * K::df(r * inv_h) * inv_h4;
            }
            __device__ static f32 W(f32 r, f32 inv_h, f32 inv_h3) {
                return norm * K::f(r * inv_h) * inv_h3;
            }
            __device__ static f32 dhW(f32 r, f32 inv_h, f32 inv_h4) {
                f32 q = r * inv_h;
                return -norm * (f32(3) * K::f(q) + q * K::df(q)) * inv_h4;
            }
        };

        // ------------------------------------------------------------- h iteration
        // One Newton sweep (IterateSmoothingLengthDensity.cpp:28-120).  G lanes share one
        // particle and stride through its row; the density sums are combined across the group.
        // MIXED: sums in FP32 with Kahan compensation on FP64 r_ab; Newton step in FP64.
        template<class K, int G, bool MIXED, class Src>
        __global__ void __launch_bounds__(BLK)
            k_h_iterate(u32 n, Rows r, Src src, const f64 *__restrict h_old,
                        f64 *__restrict h_new, f64 *__restrict eps, f64 mass, f64 h_evol_max,
                        f64 h_evol_iter_max, u64 *__restrict part, const u64 *__restrict stop) {
            using KK = Kern<K, f64>;
            using KF = Kern<K, f32>;
            // a sweep queued after the loop has converged does nothing (same value in every
            // thread: the whole grid leaves)
            if (stop && *stop) return;
            __shared__ u64 lds[3][BLK / WAVE];
            const u32 j    = (xcd_block(blockIdx.x, gridDim.x) * BLK + threadIdx.x) / G;
            const u32 lg   = threadIdx.x % G;
            const bool act = j < n;
            u32 row = 0, a = 0;
            f64 e = 0;
            if (act) {
                row = r.sel ? r.sel[j] : j;
                a   = r.outi ? r.outi[row] : row;
                e   = eps[a];
            }
            const bool work = act && e > 1e-6;
            const f64 Rkern = K::Rkern;
            f64 h_a = 0, rho_sum = 0, sumdWdh = 0;
            if (work) {
                h_a         = h_new[a];
                double3 pa  = src.pos(row);
                f64 dint    = h_a * h_a * Rkern * Rkern;
                u32 k0 = r.off[row], k1 = k0 + r.cnt[row];
                if constexpr (MIXED) {
                    using KI   = KernInv<K>;
                    f32 ih     = (f32) (1. / h_a);
                    f32 ih3    = ih * ih * ih, ih4 = ih3 * ih;
                    f32 rs = 0, rc = 0, ds = 0, dc = 0;
                    // the next entry's position is loaded while this one is used, its id a
                    // pass earlier still (each lane sums its own entries in order)
                    double3 qn{0, 0, 0};
                    u32 idn = 0;
                    if (k0 + lg < k1) qn = src.pos(r.ids[k0 + lg]);
                    if (k0 + lg + G < k1) idn = r.ids[k0 + lg + G];
                    for (u32 k = k0 + lg; k < k1; k += G) {
                        const double3 q = qn;
                        if (k + G < k1) qn = src.pos(idn);
                        if (k + 2 * G < k1) idn = r.ids[k + 2 * G];
                        f64 dx = pa.x - q.x, dy = pa.y - q.y, dz = pa.z - q.z;
                        f64 rab2 = dx * dx + dy * dy + dz * dz;
                        if (rab2 > dint) continue;
                        f32 rab = sqrt_approx((f32) rab2);
                        kahan_add(rs, rc, KI::W(rab, ih, ih3));
                        kahan_add(ds, dc, KI::dhW(rab, ih, ih4));
                    }
                    rho_sum = (f64) rs - (f64) rc;
                    sumdWdh = (f64) ds - (f64) dc;
                } else {
                    const KernAt<K> ka(h_a);
                    // group-uniform trip count; terms folded in list order
                    auto sweep = [&](auto fast) {
                        constexpr bool FAST = decltype(fast)::value;
                        // the next neighbour's position is loaded while this one is used,
                        // its id a pass earlier still
                        double3 qn{0, 0, 0};
                        u32 idn = 0;
                        if (k0 + lg < k1) qn = src.pos(r.ids[k0 + lg]);
                        if (k0 + lg + G < k1) idn = r.ids[k0 + lg + G];
                        for (u32 base = k0; base < k1; base += G) {
                            u32 k  = base + lg;
                            f64 tw = 0, th = 0;
                            const double3 qc = qn;
                            if (k + G < k1) qn = src.pos(idn);
                            if (k + 2 * G < k1) idn = r.ids[k + 2 * G];
                            if (k < k1) {
                                double3 q = qc;
                                f64 dx = pa.x - q.x, dy = pa.y - q.y, dz = pa.z - q.z;
                                f64 rab2 = dx * dx + dy * dy + dz * dz;
                                if (!(rab2 > dint)) {
                                    f64 rab = sqrt(rab2);
                                    tw      = mass * ka.template W<FAST>(rab);
                                    th      = mass * ka.template dhW<FAST>(rab);
                                }
                            }
                            rho_sum = ordered_fold<G>(rho_sum, tw);
                            sumdWdh = ordered_fold<G>(sumdWdh, th);
                        }
                    };
                    if (ka.fast) sweep(std::true_type{});
                    else sweep(std::false_type{});
                }
            }
            if (MIXED) {
                rho_sum = mass * group_sum<G>(rho_sum);
                sumdWdh = mass * group_sum<G>(sumdWdh);
            }
            if (work && lg == 0) {
                f64 h_max_evol_p = h_evol_iter_max, h_max_evol_m = 1 / h_evol_iter_max;
                f64 rho_ha  = rho_h(mass, h_a, KK::hfactd);
                f64 f_iter  = rho_sum - rho_ha;
                f64 df_iter = sumdWdh + 3 * rho_ha / h_a;
                f64 new_h   = h_a - f_iter / df_iter;
                if (new_h < h_a * h_max_evol_m) new_h = h_max_evol_m * h_a;
                if (new_h > h_a * h_max_evol_p) new_h = h_max_evol_p * h_a;
                f64 ha_0 = h_old[a];
                if (new_h < ha_0 * h_evol_max) {
                    h_new[a] = new_h;
                    e        = fabs(new_h - h_a) / ha_0;
                } else {
                    h_new[a] = ha_0 * h_evol_max;
                    e        = -1;
                }
                eps[a] = e;
            }
            u64 kmax = ORDERED_NEG_INF, kmin = ORDERED_POS_INF, kact = 0;
            if (act && lg == 0) {
                kmax = kmin = f64_to_ordered(e);
                kact         = e > 1e-6 ? 1u : 0u; // still iterating in the next sweep
            }
            kmax = block_reduce(kmax, umax64, lds[0]);
            kmin = block_reduce(kmin, umin64, lds[1]);
            kact = block_reduce(kact, [](u64 p, u64 q) { return p + q; }, lds[2]);
            if (threadIdx.x == 0) {
                part[blockIdx.x]                 = kmax;
                part[gridDim.x + blockIdx.x]     = kmin;
                part[2 * gridDim.x + blockIdx.x] = kact;
            }
        }

        // ---- the Newton loop on the device (LoopSmoothingLengthIter.cpp:28-80)
        // Loop state, u64 words: 0 stop, 1 sweeps done, 2/3 max/min eps key of the particles
        // dropped from the sweep list, 4/5/6 max/min eps key and active count of the last
        // sweep (over its list), 8/9 max/min eps key over all particles, 10 active count.
        enum : int {
            HL_STOP = 0, HL_ITERS, HL_FROZEN_MAX, HL_FROZEN_MIN, HL_SWEEP_MAX, HL_SWEEP_MIN,
            HL_SWEEP_ACT, HL_PAD, HL_MAX, HL_MIN, HL_ACT, HL_WORDS = 16
        };
        __global__ void k_h_loop_init(u64 *__restrict st) {
            if (threadIdx.x < HL_WORDS) {
                u64 v = 0;
                if (threadIdx.x == HL_FROZEN_MAX) v = ORDERED_NEG_INF;
                if (threadIdx.x == HL_FROZEN_MIN) v = ORDERED_POS_INF;
                st[threadIdx.x] = v;
            }
        }
        // after each sweep: count it and stop the loop once max eps < epsilon_h, the
        // reference's break condition
        __global__ void k_h_loop_ctl(u64 *__restrict st, f64 epsilon_h) {
            if (threadIdx.x != 0 || st[HL_STOP]) return;
            st[HL_ITERS] += 1;
            u64 mx = umax64(st[HL_FROZEN_MAX], st[HL_SWEEP_MAX]);
            u64 mn = umin64(st[HL_FROZEN_MIN], st[HL_SWEEP_MIN]);
            st[HL_MAX] = mx;
            st[HL_MIN] = mn;
            st[HL_ACT] = st[HL_SWEEP_ACT];
            if (ordered_to_f64(mx) < epsilon_h) st[HL_STOP] = 1;
        }
        // rows of the sweep list whose particle still iterates (eps > 1e-6: the sweep's own
        // test); the eps of the rows dropped is final for the subcycle and goes into the
        // frozen max / min
        // (grid-stride over few blocks: one pair of same-address atomics per block)
        __global__ void __launch_bounds__(BLK)
            k_h_active_flags(u32 n, const u32 *__restrict sel, const u32 *__restrict outi,
                             const f64 *__restrict eps, u32 *__restrict flag,
                             u64 *__restrict st) {
            __shared__ u64 lds[2][BLK / WAVE];
            u64 kmax = ORDERED_NEG_INF, kmin = ORDERED_POS_INF;
            for (u32 j = blockIdx.x * BLK + threadIdx.x; j < n; j += gridDim.x * BLK) {
                u32 row = sel ? sel[j] : j;
                f64 e   = eps[outi ? outi[row] : row];
                bool keep = e > 1e-6;
                flag[j]   = keep ? 1u : 0u;
                if (!keep) {
                    u64 k = f64_to_ordered(e);
                    kmax  = umax64(kmax, k);
                    kmin  = umin64(kmin, k);
                }
            }
            kmax = block_reduce(kmax, umax64, lds[0]);
            kmin = block_reduce(kmin, umin64, lds[1]);
            if (threadIdx.x == 0) {
                if (kmax != ORDERED_NEG_INF) atomicMax((unsigned long long *) &st[HL_FROZEN_MAX],
                                                       (unsigned long long) kmax);
                if (kmin != ORDERED_POS_INF) atomicMin((unsigned long long *) &st[HL_FROZEN_MIN],
                                                       (unsigned long long) kmin);
            }
        }
        __global__ void k_h_compact_rows(u32 n, const u32 *__restrict sel,
                                         const u32 *__restrict flag, const u32 *__restrict off,
                                         u32 *__restrict out) {
            u32 j = blockIdx.x * blockDim.x + threadIdx.x;
            if (j < n && flag[j]) out[off[j]] = sel ? sel[j] : j;
        }

        // ------------------------------------------------------------- omega
        // ComputeOmega.cpp:25-74, same group layout as the h iteration.  FILTER: h and the
        // positions are final for the rest of the step, so the same pass also writes, for the
        // force loop, the subsequence of each row that passes the force criterion
        // rab2 <= max(h_a, h_b)^2 R^2 (UpdateDerivs.cpp:160-167), in row order, at the row's
        // offset in fids, with its length in fcnt (ps.w must hold the final h then).
        // MIXED: the dW/dh sum in FP32 with Kahan compensation per lane on FP64 r_ab (lanes
        // combined in FP64), like the mixed density sum; the force-row criterion stays FP64.
        template<class K, int G, bool FILTER, bool MIXED, class Src>
        __global__ void __launch_bounds__(BLK)
            k_omega(u32 n, Rows r, Src src, const f64 *__restrict h, f64 *__restrict omega,
                    f64 mass, u32 *__restrict fcnt, u32 *__restrict fids) {
            using KK = Kern<K, f64>;
            using KI = KernInv<K>;
            // the kept ids of a row go through a ring in LDS and leave in runs of FL entries
            // (64 bytes), so every line of the force rows is written whole and at once
            constexpr u32 RING = 32, FL = 16, PER = FL / G;
            __shared__ u32 ring[FILTER ? BLK / G : 1][RING];
            const u32 j    = (xcd_block(blockIdx.x, gridDim.x) * BLK + threadIdx.x) / G;
            const u32 lg   = threadIdx.x % G;
            const u32 grp  = FILTER ? threadIdx.x / G : 0;
            const bool act = j < n;
            const f64 Rkern = K::Rkern, Rker2 = K::Rkern * K::Rkern;
            u32 row = 0, a = 0, fc = 0, flushed = 0;
            f64 h_a = 0, part_omega_sum = 0;
            f32 ms = 0, mc = 0; // MIXED: per-lane Kahan sum of dW/dh
            if (act) {
                row        = r.sel ? r.sel[j] : j;
                a          = r.outi ? r.outi[row] : row;
                h_a        = h[a];
                double3 pa = src.pos(row);
                f64 dint   = h_a * h_a * Rkern * Rkern;
                f64 ca     = h_a * h_a * Rker2;
                u32 k0 = r.off[row], k1 = k0 + r.cnt[row];
                const KernAt<K> ka(h_a);
                f32 ih = 0, ih4 = 0;
                if constexpr (MIXED) {
                    ih  = (f32) (1. / h_a);
                    ih4 = ih * ih * ih * ih;
                }
                // one entry of the row: the Omega term and the force-row keep flag
                auto entry = [&](bool v, u32 id, double4 q, f64 &th, bool &keep) {
                    if (!v) return;
                    f64 dx = pa.x - q.x, dy = pa.y - q.y, dz = pa.z - q.z;
                    f64 rab2 = dx * dx + dy * dy + dz * dz;
                    if (!(rab2 > dint)) {
                        if constexpr (MIXED) {
                            kahan_add(ms, mc, KI::dhW(sqrt_approx((f32) rab2), ih, ih4));
                        } else {
                            f64 rab = sqrt(rab2);
                            if (ka.fast) th = mass * ka.template dhW<true>(rab);
                            else th = mass * ka.template dhW<false>(rab);
                        }
                    }
                    if constexpr (FILTER) keep = !((rab2 > ca) & (rab2 > q.w * q.w * Rker2));
                };
                // fold the G terms of one pass, then append the kept ids in order
                auto emit = [&](f64 th, bool keep, u32 id) {
                    if constexpr (!MIXED) part_omega_sum = ordered_fold<G>(part_omega_sum, th);
                    if constexpr (FILTER) {
                        u64 bal   = __ballot(keep);
                        u32 gbits = (u32) (bal >> (__lane_id() & ~(G - 1))) & ((1u << G) - 1u);
                        if (keep)
                            ring[grp][(fc + (u32) __popc(gbits & ((1u << lg) - 1u))) & (RING - 1)] =
                                id;
                        fc += (u32) __popc(gbits);
                        if (fc - flushed >= FL) { // uniform in the group
                            wave_lds_sync();
                            u32 p = flushed + lg * PER;
                            u32 *dst = fids + k0 + p;
                            if constexpr (PER == 4) {
                                uint4 v = make_uint4(ring[grp][p & (RING - 1)],
                                                     ring[grp][(p + 1) & (RING - 1)],
                                                     ring[grp][(p + 2) & (RING - 1)],
                                                     ring[grp][(p + 3) & (RING - 1)]);
                                if (((k0 + p) & 3u) == 0) *(uint4 *) dst = v;
                                else {
                                    dst[0] = v.x;
                                    dst[1] = v.y;
                                    dst[2] = v.z;
                                    dst[3] = v.w;
                                }
                            } else {
                                for (u32 t = 0; t < PER; t++) dst[t] = ring[grp][(p + t) & (RING - 1)];
                            }
                            flushed += FL;
                            wave_lds_sync();
                        }
                    }
                };
                // four entries per lane per pass (k, k + G, k + 2G, k + 3G), all loaded before
                // any is used, folded in that (list) order
                constexpr int U = 4;
                for (u32 base = k0; base < k1; base += U * G) {
                    auto ld = [&](u32 id) -> double4 { // h only needed by the filter
                        if constexpr (FILTER) return src.pos4(id);
                        else {
                            double3 q = src.pos(id);
                            return make_double4(q.x, q.y, q.z, 0.);
                        }
                    };
                    bool v[U];
                    u32 iu[U];
                    double4 q[U];
#pragma unroll
                    for (int t = 0; t < U; t++) {
                        const u32 k = base + lg + (u32) t * G;
                        v[t]        = k < k1;
                        iu[t]       = r.ids[v[t] ? k : k0];
                    }
#pragma unroll
                    for (int t = 0; t < U; t++) q[t] = ld(iu[t]);
#pragma unroll
                    for (int t = 0; t < U; t++) {
                        f64 th    = 0;
                        bool keep = false;
                        entry(v[t], iu[t], q[t], th, keep);
                        emit(th, keep, iu[t]);
                    }
                }
                if constexpr (FILTER) { // the last partial run
                    wave_lds_sync();
                    for (u32 p = flushed + lg; p < fc; p += G) fids[k0 + p] = ring[grp][p & (RING - 1)];
                }
            }
            if constexpr (MIXED) part_omega_sum = mass * group_sum<G>((f64) ms - (f64) mc);
            if (act && lg == 0) {
                f64 rho_ha = rho_h(mass, h_a, KK::hfactd);
                omega[a]   = 1 + (h_a / (3 * rho_ha)) * part_omega_sum;
                if (FILTER) fcnt[row] = fc;
            }
        }

        // ps[s].w = h[sort_index[s]] (final h into the sorted position records)
        __global__ void k_ps_refresh_h(u32 nm, const u32 *__restrict sort_index,
                                       const f64 *__restrict h, double4 *__restrict ps) {
            u32 s = blockIdx.x * blockDim.x + threadIdx.x;
            if (s < nm) ps[s].w = h[sort_index[s]];
        }

        // ------------------------------------------------------------- eos
        // ComputeEos.cpp:152-247 (adiabatic), 656-665 (isothermal), 726-798 (LP07)
        __device__ inline void eos_eval(const EosArgs &e, f64 rho, f64 u, f64 x, f64 y, f64 z,
                                        f64 &P, f64 &cs) {
            if (e.kind == 0) {
                P  = (e.gamma - 1) * rho * u;
                cs = sqrt(e.gamma * P / rho);
            } else if (e.kind == 1) {
                P  = e.cs * e.cs * rho;
                cs = e.cs;
            } else {
                f64 r0sq  = e.r0 * e.r0;
                f64 mq    = -e.q;
                f64 Rsq   = x * x + y * y + z * z;
                f64 cs_sq = (e.cs0 * e.cs0) * pow(Rsq / r0sq, mq);
                cs        = sqrt(cs_sq);
                P         = cs_sq * rho;
            }
        }

        template<class K>
        __global__ void k_eos(u32 nm, const f64 *__restrict h, const f64 *__restrict u,
                              const f64 *__restrict x, const f64 *__restrict y,
                              const f64 *__restrict z, f64 *__restrict P, f64 *__restrict cs,
                              f64 mass, EosArgs e) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= nm) return;
            f64 rho = rho_h(mass, h[i], (f64) K::hfactd);
            f64 p, c;
            eos_eval(e, rho, u[i], x[i], y[i], z[i], p, c);
            P[i]  = p;
            cs[i] = c;
        }

        // ------------------------------------------------------------- force records
        __device__ inline f64 inv_sat_positive(f64 v) { return (v >= 1e-9) ? 1. / v : 0.; }
        __device__ inline f64 inv_sat_zero(f64 v) { return (v != 0. && v == v) ? 1. / v : 0.; }

        // everything the pair loop reads of a neighbour (144 B): the per-particle factors of
        // UpdateDerivs.cpp:140-240 evaluated once (the same IEEE operations as the pair loop,
        // so the values are bit-identical) and the reciprocals of its three per-particle
        // divisors (h, h^4, rho Omega) for div_rcp.  (A 128-byte record that recomputes rho
        // and (h R)^2 from h measured 23 % slower in the force loop.)
        struct alignas(16) FRec {
            f64 x, y, z, h, vx, vy, vz, u, P, cs, rho, inv_sub, romega, hR2, alpha;
            f64 rh, rh4, rromega; // 1 / h, 1 / h^4, 1 / (rho Omega); 0: use the IEEE division
        };
        static_assert(sizeof(FRec) == 144, "force record is 144 B");
        template<class K>
        __device__ inline f64 rec_rho(const FRec &r, f64) {
            return r.rho;
        }
        template<class K>
        __device__ inline f64 rec_hR2(const FRec &r) {
            return r.hR2;
        }
        // mixed: FP64 positions for r_ab, FP32 fields (64 B)
        struct alignas(16) FRec32 {
            f64 x, y, z;
            f32 inv_h, vx, vy, vz, u, P, cs, rho, inv_sub, inv_romega;
        };
        static_assert(sizeof(FRec32) == 64, "mixed force record is 64 B");

        template<class K>
        __device__ inline FRec make_frec(f64 x, f64 y, f64 z, f64 h, f64 vx, f64 vy, f64 vz,
                                         f64 u, f64 P, f64 cs, f64 omega, f64 mass) {
            const f64 Rker2 = K::Rkern * K::Rkern;
            FRec r;
            r.x = x, r.y = y, r.z = z, r.h = h;
            r.vx = vx, r.vy = vy, r.vz = vz, r.u = u;
            r.P = P, r.cs = cs;
            r.rho     = rho_h(mass, h, (f64) K::hfactd);
            r.inv_sub = inv_sat_zero((r.rho * r.rho) * omega);
            r.romega  = r.rho * omega;
            r.hR2     = h * h * Rker2;
            r.alpha   = 0;
            f64 h4    = h * h * h * h;
            bool fast = rcp_safe(h) && rcp_safe(h4) && rcp_safe(r.romega);
            r.rh      = fast ? 1. / h : 0.;
            r.rh4     = fast ? 1. / h4 : 0.;
            r.rromega = fast ? 1. / r.romega : 0.;
            return r;
        }
        template<class K>
        __device__ inline FRec32 make_frec32(f64 x, f64 y, f64 z, f64 h, f64 vx, f64 vy, f64 vz,
                                             f64 u, f64 P, f64 cs, f64 omega, f64 mass) {
            FRec32 r;
            r.x = x, r.y = y, r.z = z;
            r.inv_h = (f32) (1. / h);
            r.vx = (f32) vx, r.vy = (f32) vy, r.vz = (f32) vz, r.u = (f32) u;
            r.P = (f32) P, r.cs = (f32) cs;
            f32 om       = (f32) omega;
            r.rho        = (f32) rho_h(mass, h, (f64) K::hfactd);
            r.inv_sub    = 1.f / ((r.rho * r.rho) * om);
            r.inv_romega = 1.f / (om * r.rho);
            return r;
        }

        // record sources: packed records, or the SoA arrays (reference-layout entry points)
        template<class Rec>
        struct RecSrc {
            const Rec *__restrict rec;
            __device__ Rec get(u32 b) const { return rec[b]; }
        };
        template<class K, class Rec>
        struct SoASrc {
            const f64 *__restrict x, *__restrict y, *__restrict z, *__restrict vx, *__restrict vy,
                *__restrict vz, *__restrict h, *__restrict omega, *__restrict u, *__restrict P,
                *__restrict cs;
            f64 mass;
            __device__ Rec get(u32 b) const {
                if constexpr (sizeof(Rec) == sizeof(FRec))
                    return make_frec<K>(x[b], y[b], z[b], h[b], vx[b], vy[b], vz[b], u[b], P[b],
                                        cs[b], omega[b], mass);
                else
                    return make_frec32<K>(x[b], y[b], z[b], h[b], vx[b], vy[b], vz[b], u[b], P[b],
                                          cs[b], omega[b], mass);
            }
        };

        // EOS on the merged set fused with the record build: record s takes particle perm[s].
        // REFRESH: rewrite only the fields that a corrector pass changes
        template<class K, bool MIXED, bool REFRESH = false>
        __global__ void k_pack_force(u32 nm, const u32 *__restrict perm, const f64 *__restrict x,
                                     const f64 *__restrict y, const f64 *__restrict z,
                                     const f64 *__restrict h, const f64 *__restrict vx,
                                     const f64 *__restrict vy, const f64 *__restrict vz,
                                     const f64 *__restrict u, const f64 *__restrict omega,
                                     f64 *__restrict P, f64 *__restrict cs, void *__restrict rec,
                                     f64 mass, EosArgs e) {
            u32 s = blockIdx.x * blockDim.x + threadIdx.x;
            if (s >= nm) return;
            u32 b  = perm ? perm[s] : s;
            f64 hb = h[b], ub = u[b], xb = 0, yb = 0, zb = 0;
            if (!REFRESH || e.kind == 2) xb = x[b], yb = y[b], zb = z[b];
            f64 rho = rho_h(mass, hb, (f64) K::hfactd);
            f64 p, c;
            eos_eval(e, rho, ub, xb, yb, zb, p, c);
            P[b]  = p;
            cs[b] = c;
            if constexpr (REFRESH && MIXED) {
                // corrector reruns: positions, h and omega are those of the full build, only
                // v and u (and with u the EOS) changed; the other fields are already in place
                FRec32 *r = (FRec32 *) rec + s;
                r->vx = (f32) vx[b], r->vy = (f32) vy[b], r->vz = (f32) vz[b];
                r->u = (f32) ub, r->P = (f32) p, r->cs = (f32) c;
            } else if constexpr (REFRESH) {
                FRec *r = (FRec *) rec + s;
                r->vx = vx[b], r->vy = vy[b], r->vz = vz[b], r->u = ub, r->P = p, r->cs = c;
            } else if constexpr (MIXED)
                ((FRec32 *) rec)[s] =
                    make_frec32<K>(xb, yb, zb, hb, vx[b], vy[b], vz[b], ub, p, c, omega[b], mass);
            else
                ((FRec *) rec)[s] =
                    make_frec<K>(xb, yb, zb, hb, vx[b], vy[b], vz[b], ub, p, c, omega[b], mass);
        }

        struct ForceOut {
            f64 *ax, *ay, *az, *du, *ax_old, *ay_old, *az_old, *du_old, *vsig, *dtc;
        };

        // a_old/du_old save, outputs, CFL dt (ComputeCFLCourant.hpp, ComputeCFLForce.hpp:50-62)
        __device__ inline f64 force_store(const ForceOut &o, u32 a, f64 fx, f64 fy, f64 fz,
                                          f64 tdu, f64 vsig_max, f64 h_a, const ForceArgs &fa) {
            o.ax_old[a] = o.ax[a];
            o.ay_old[a] = o.ay[a];
            o.az_old[a] = o.az[a];
            o.du_old[a] = o.du[a];
            if (fa.ext_x) { // ExternalForces.cpp:324-360
                fx = fx + fa.ext_x[a];
                fy = fy + fa.ext_y[a];
                fz = fz + fa.ext_z[a];
            }
            o.ax[a]     = fx;
            o.ay[a]     = fy;
            o.az[a]     = fz;
            o.du[a]     = tdu;
            o.vsig[a]   = vsig_max;
            f64 dt_c    = fa.C_cour * h_a / vsig_max;
            f64 abs_a   = sqrt(fx * fx + fy * fy + fz * fz);
            f64 dt_f    = fa.C_force * sqrt(h_a / abs_a);
            f64 cfl_dt  = fmin(__builtin_inf(), dt_c);
            cfl_dt      = fmin(cfl_dt, dt_f);
            o.dtc[a]    = cfl_dt;
            return cfl_dt;
        }

        // ------------------------------------------------------------- force (FP64)
        // UpdateDerivs.cpp:88-288 (constant AV) fused with the CFL signal speed
        // (Solver.cpp:2749-2783).  G lanes per particle stride through its row.
        template<class K, int G, class Src>
        __global__ void __launch_bounds__(BLK)
            k_force(u32 n, Rows r, Src src, ForceOut o, ForceArgs fa, u64 *__restrict part) {
            using KK = Kern<K, f64>;
            __shared__ u64 lds[2][BLK / WAVE];
            // G = 4: the four lanes of a particle share the ordered folds out through LDS, lane
            // c folding component c (x, y, z, du) over the group's terms in lane order; slot 5
            // holds +0.0, added by lanes 0-2 where lane 3 adds du's second term (a sum that
            // starts at +0.0 never becomes -0.0 under round-to-nearest, so adding +0.0 leaves it
            // unchanged).  Other G: every lane folds all four and the first lane's sums count.
            constexpr bool SHARE = G == 4;
            __shared__ f64 fold[SHARE ? BLK : 1][6];
            const u32 j    = (xcd_block(blockIdx.x, gridDim.x) * BLK + threadIdx.x) / G;
            const u32 lg   = threadIdx.x % G;
            const bool act = j < n;
            u64 npair = 0;
            f64 fx = 0, fy = 0, fz = 0, tdu = 0, vsig_max = 0;
            f64 facc = 0; // SHARE: this lane's component
            if constexpr (SHARE) fold[threadIdx.x][5] = 0.;
            u32 row = 0;
            FRec A{};
            if (act) {
                row             = r.sel ? r.sel[j] : j;
                A               = src.get(row);
                const f64 pmass = fa.mass;
                const f64 h_a = A.h, xa = A.x, ya = A.y, za = A.z;
                const f64 vxa = A.vx, vya = A.vy, vza = A.vz;
                const f64 P_a = A.P, u_a = A.u, cs_a = A.cs;
                const f64 rho_a = rec_rho<K>(A, fa.mass);
                const f64 rho_a_inv = 1. / rho_a;
                // 1 / (omega_a rho_a): romega is rho_a omega_a, the same product
                const f64 omega_a_rho_a_inv = 1 / A.romega;
                const f64 ia_fac = A.inv_sub;
                const f64 ca     = rec_hR2<K>(A);
                // varying AV: per-particle alpha (NodeUpdateDerivsVaryingAlphaAV.cpp:60-150)
                const f64 alpha_a = fa.varying ? A.alpha : fa.alpha_AV;
                const f64 h4_a    = h_a * h_a * h_a * h_a;
                u32 k0 = r.off[row], k1 = k0 + r.cnt[row];
                // group-uniform trip count; the sums take the terms in list order.  FA: the
                // particle's own divisions go through its reciprocals (one loop per case)
                auto sweep = [&](auto fa_tag) {
                constexpr bool FA = decltype(fa_tag)::value;
                // the pair terms of entry k with neighbour record B; a skipped entry leaves
                // +0.0 terms, which leave a sum that starts at +0.0 unchanged (it can never
                // become -0.0 under round-to-nearest)
                auto pair = [&](const FRec &B, u32 k, f64 &tx, f64 &ty, f64 &tz, f64 &tu1,
                                f64 &tu2) {
                    if (k >= k1) return;
                    f64 dx = xa - B.x, dy = ya - B.y, dz = za - B.z;
                    f64 rab2 = dx * dx + dy * dy + dz * dz;
                    if ((rab2 > ca) & (rab2 > rec_hR2<K>(B))) return;
                    npair++;
                    f64 rab = sqrt(rab2);
                    // dW = norm f'(r / h) / h^4 (sphkernels.hpp:2329-2343) and
                    // Fab_b / (rho_b omega_b) of the conductivity term, the divisions
                    // by way of the particles' reciprocals
                    f64 Fab_a, Fab_b, rho_b, fb_rom;
                    if constexpr (FA)
                        Fab_a = div_rcp(KK::norm * K::df(div_rcp(rab, h_a, A.rh)), h4_a, A.rh4);
                    else
                        Fab_a = KK::dW(rab, h_a);
                    rho_b = B.rho;
                    if (B.rh != 0) {
                        f64 h4_b = B.h * B.h * B.h * B.h;
                        Fab_b  = div_rcp(KK::norm * K::df(div_rcp(rab, B.h, B.rh)), h4_b, B.rh4);
                        fb_rom = div_rcp(Fab_b, B.romega, B.rromega);
                    } else {
                        Fab_b  = KK::dW(rab, B.h);
                        fb_rom = Fab_b / B.romega;
                    }
                    f64 vabx = vxa - B.vx, vaby = vya - B.vy, vabz = vza - B.vz;
                    f64 inv  = inv_sat_positive(rab);
                    f64 rux = dx * inv, ruy = dy * inv, ruz = dz * inv;
                    f64 v_ab_r_ab = vabx * rux + vaby * ruy + vabz * ruz;
                    f64 abs_v     = fabs(v_ab_r_ab);
                    f64 vsig_a = alpha_a * cs_a + fa.beta_AV * abs_v;
                    f64 vsig_b = (fa.varying ? B.alpha : fa.alpha_AV) * B.cs + fa.beta_AV * abs_v;
                    f64 rho_avg = (rho_a + rho_b) * 0.5;
                    f64 vsig_u  = sqrt(fabs(P_a - B.P) / rho_avg);
                    f64 qa_ab = fmax(-0.5 * rho_a * vsig_a * v_ab_r_ab, 0.);
                    f64 qb_ab = fmax(-0.5 * rho_b * vsig_b * v_ab_r_ab, 0.);
                    f64 AV_P_a = P_a + qa_ab, AV_P_b = B.P + qb_ab;
                    f64 nax = rux * Fab_a, nay = ruy * Fab_a, naz = ruz * Fab_a;
                    f64 nbx = rux * Fab_b, nby = ruy * Fab_b, nbz = ruz * Fab_b;
                    f64 ia = AV_P_a * ia_fac;
                    f64 ib = AV_P_b * B.inv_sub;
                    f64 mm = -pmass;
                    tx     = mm * (ia * nax + ib * nbx);
                    ty     = mm * (ia * nay + ib * nby);
                    tz     = mm * (ia * naz + ib * nbz);
                    tu1    = AV_P_a * (omega_a_rho_a_inv * rho_a_inv) * pmass *
                          (vabx * nax + vaby * nay + vabz * naz);
                    tu2 = pmass * fa.alpha_u * vsig_u * (u_a - B.u) * 0.5 *
                          (Fab_a * omega_a_rho_a_inv + fb_rom);
                    // CFL signal speed: r_ab / |r_ab|, alpha 1, beta 2.  inv is RN(1 / rab)
                    // from 1e-9 on and 0 below, where r_ab / |r_ab| is taken as 0: div_rcp
                    // gives both
                    f64 r2x = div_rcp(dx, rab, inv), r2y = div_rcp(dy, rab, inv),
                        r2z = div_rcp(dz, rab, inv);
                    f64 vdot = vabx * r2x + vaby * r2y + vabz * r2z;
                    vsig_max = fmax(vsig_max, 1.0 * cs_a + 2.0 * fabs(vdot));
                };
                auto step = [&](const FRec &B, u32 k) {
                    f64 tx = 0, ty = 0, tz = 0, tu1 = 0, tu2 = 0;
                    pair(B, k, tx, ty, tz, tu1, tu2);
                    if constexpr (SHARE) {
                        f64 *m = fold[threadIdx.x];
                        m[0]   = tx;
                        m[1]   = ty;
                        m[2]   = tz;
                        m[3]   = tu1;
                        m[4]   = tu2;
                        wave_lds_sync();
                        const f64(*grp)[6] = fold + (threadIdx.x & ~3u);
                        const u32 sa = lg, sb = lg == 3 ? 4u : 5u;
#pragma unroll
                        for (int L = 0; L < 4; L++) {
                            facc = facc + grp[L][sa];
                            facc = facc + grp[L][sb];
                        }
                        wave_lds_sync();
                    } else {
                        fx  = ordered_fold<G>(fx, tx);
                        fy  = ordered_fold<G>(fy, ty);
                        fz  = ordered_fold<G>(fz, tz);
                        tdu = ordered_fold2<G>(tdu, tu1, tu2);
                    }
                };
                // two records in flight, used in turn: the next neighbour's record is loaded
                // while the current one is used (3.21 -> 2.98 ms per evaluation at Sedov
                // 128^3), and unrolling by two keeps the records in place instead of copying
                // 144 bytes of registers per entry.  Trip counts are uniform in the group.
                // Loads are unconditional (entries past the row end read its last entry again,
                // a cached line), so no record has to be kept across a skipped load.
                if (k0 < k1) {
                    const u32 last = k1 - 1;
                    // the id of the next B1 is loaded an iteration ahead, so its record load
                    // does not wait for an id load first
                    u32 id1 = r.ids[min(k0 + lg + G, last)];
                    FRec B0 = src.get(r.ids[min(k0 + lg, last)]);
                    for (u32 base = k0; base < k1; base += 2 * G) {
                        u32 k          = base + lg;
                        const FRec B1  = src.get(id1);
                        const u32 id0  = r.ids[min(k + 2 * G, last)];
                        id1            = r.ids[min(k + 3 * G, last)];
                        step(B0, k);
                        B0 = src.get(id0);
                        if (base + G < k1) step(B1, k + G);
                    }
                }
                };
                if (A.rh != 0) sweep(std::true_type{});
                else sweep(std::false_type{});
            }
            if constexpr (SHARE) {
                fx  = facc;
                fy  = group_lane<G, 1>(facc);
                fz  = group_lane<G, 2>(facc);
                tdu = group_lane<G, 3>(facc);
            }
            vsig_max = group_max<G>(vsig_max);
            u64 key  = ORDERED_POS_INF;
            if (act && lg == 0) {
                u32 a = r.outi ? r.outi[row] : row;
                key   = f64_to_ordered(force_store(o, a, fx, fy, fz, tdu, vsig_max, A.h, fa));
            }
            key   = block_reduce(key, umin64, lds[0]);
            npair = block_reduce(npair, [](u64 p, u64 q) { return p + q; }, lds[1]);
            if (threadIdx.x == 0) {
                part[blockIdx.x]             = key;
                part[gridDim.x + blockIdx.x] = npair;
            }
        }

        // ------------------------------------------------------------- force (mixed)
        // r_ab and the neighbour criterion in FP64, pair terms in FP32, Kahan-compensated
        // per-lane sums combined across the group in FP64
        template<class K, int G, class Src>
        __global__ void __launch_bounds__(BLK)
            k_force_mixed(u32 n, Rows r, Src src, const f64 *__restrict h64, ForceOut o,
                          ForceArgs fa, u64 *__restrict part) {
            using KI = KernInv<K>;
            __shared__ u64 lds[2][BLK / WAVE];
            const u32 j    = (xcd_block(blockIdx.x, gridDim.x) * BLK + threadIdx.x) / G;
            const u32 lg   = threadIdx.x % G;
            const bool act = j < n;
            const f64 Rker2 = K::Rkern * K::Rkern;
            u64 npair = 0;
            f32 sx = 0, cx = 0, sy = 0, cy = 0, sz = 0, cz = 0, su = 0, cu = 0, vsig_max = 0;
            u32 row = 0, a = 0;
            f64 h_a64 = 0;
            if (act) {
                row = r.sel ? r.sel[j] : j;
                a   = r.outi ? r.outi[row] : row;
                const FRec32 A = src.get(row);
                h_a64 = h64[a];
                const f32 pmass = (f32) fa.mass;
                const f32 ih_a = A.inv_h, ih4_a = ih_a * ih_a * ih_a * ih_a;
                const f32 vxa = A.vx, vya = A.vy, vza = A.vz, P_a = A.P, u_a = A.u, cs_a = A.cs;
                const f32 rho_a = A.rho, rho_a_inv = rcp_approx(rho_a);
                const f32 omega_a_rho_a_inv = A.inv_romega, inv_sub_a = A.inv_sub;
                const f32 alpha_AV = (f32) fa.alpha_AV, beta_AV = (f32) fa.beta_AV,
                          alpha_u = (f32) fa.alpha_u;
                const f32 c_du1 = omega_a_rho_a_inv * rho_a_inv * pmass;
                const f64 ca = h_a64 * h_a64 * Rker2;
                u32 k0 = r.off[row], k1 = k0 + r.cnt[row];
                auto pair = [&](const FRec32 &B) {
                    f64 dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
                    f64 rab2 = dx * dx + dy * dy + dz * dz;
                    // neighbour-side criterion in FP32 (the solver's rows are prefiltered in
                    // FP64 already; this only matters for unfiltered reference-layout rows)
                    if ((rab2 > ca) & ((f32) rab2 * (B.inv_h * B.inv_h) > (f32) Rker2)) return;
                    npair++;
                    f32 rab  = sqrt_approx((f32) rab2);
                    f32 inv  = (rab >= 1e-9f) ? rcp_approx(rab) : 0.f;
                    f32 rux = (f32) dx * inv, ruy = (f32) dy * inv, ruz = (f32) dz * inv;
                    f32 vabx = vxa - B.vx, vaby = vya - B.vy, vabz = vza - B.vz;
                    f32 rho_b = B.rho;
                    f32 ih_b  = B.inv_h, ih4_b = ih_b * ih_b * ih_b * ih_b;
                    f32 Fab_a = KI::dW(rab, ih_a, ih4_a), Fab_b = KI::dW(rab, ih_b, ih4_b);
                    f32 v_ab_r_ab = vabx * rux + vaby * ruy + vabz * ruz;
                    f32 abs_v     = fabsf(v_ab_r_ab);
                    f32 vsig_a = alpha_AV * cs_a + beta_AV * abs_v;
                    f32 vsig_b = alpha_AV * B.cs + beta_AV * abs_v;
                    f32 rho_avg = (rho_a + rho_b) * 0.5f;
                    f32 vsig_u  = sqrt_approx(fabsf(P_a - B.P) * rcp_approx(rho_avg));
                    f32 qa_ab = fmaxf(-0.5f * rho_a * vsig_a * v_ab_r_ab, 0.f);
                    f32 qb_ab = fmaxf(-0.5f * rho_b * vsig_b * v_ab_r_ab, 0.f);
                    f32 AV_P_a = P_a + qa_ab, AV_P_b = B.P + qb_ab;
                    f32 ia = AV_P_a * inv_sub_a * Fab_a;
                    f32 ib = AV_P_b * B.inv_sub * Fab_b;
                    f32 s  = -pmass * (ia + ib);
                    kahan_add(sx, cx, s * rux);
                    kahan_add(sy, cy, s * ruy);
                    kahan_add(sz, cz, s * ruz);
                    f32 t1 = AV_P_a * c_du1 * Fab_a * v_ab_r_ab;
                    f32 t2 = pmass * alpha_u * vsig_u * (u_a - B.u) * 0.5f *
                             (Fab_a * omega_a_rho_a_inv + Fab_b * B.inv_romega);
                    kahan_add(su, cu, t1 + t2);
                    vsig_max = fmaxf(vsig_max, cs_a + 2.f * abs_v);
                };
                // two records in flight, used in turn (no register copies between entries);
                // loads past the row end read its last entry again (a cached line)
                if (k0 < k1) {
                    const u32 last = k1 - 1;
                    u32 id1   = r.ids[min(k0 + lg + G, last)]; // an iteration ahead
                    FRec32 B0 = src.get(r.ids[min(k0 + lg, last)]);
                    for (u32 base = k0; base < k1; base += 2 * G) {
                        u32 k           = base + lg;
                        const FRec32 B1 = src.get(id1);
                        const u32 id0   = r.ids[min(k + 2 * G, last)];
                        id1             = r.ids[min(k + 3 * G, last)];
                        if (k < k1) pair(B0);
                        B0 = src.get(id0);
                        if (k + G < k1) pair(B1);
                    }
                }
            }
            f64 fx   = group_sum<G>((f64) sx - (f64) cx);
            f64 fy   = group_sum<G>((f64) sy - (f64) cy);
            f64 fz   = group_sum<G>((f64) sz - (f64) cz);
            f64 tdu  = group_sum<G>((f64) su - (f64) cu);
            f64 vs   = group_max<G>((f64) vsig_max);
            u64 key  = ORDERED_POS_INF;
            if (act && lg == 0)
                key = f64_to_ordered(force_store(o, a, fx, fy, fz, tdu, vs, h_a64, fa));
            key   = block_reduce(key, umin64, lds[0]);
            npair = block_reduce(npair, [](u64 p, u64 q) { return p + q; }, lds[1]);
            if (threadIdx.x == 0) {
                part[blockIdx.x]             = key;
                part[gridDim.x + blockIdx.x] = npair;
            }
        }

        // ------------------------------------------------------------- varying AV (CD10)
        __device__ inline void cross3(const f64 a[3], const f64 b[3], f64 c[3]) {
            c[0] = a[1] * b[2] - a[2] * b[1];
            c[1] = a[2] * b[0] - a[0] * b[2];
            c[2] = a[0] * b[1] - a[1] * b[0];
        }

        // div v and curl v (DiffOperator.cpp:24-268) over the force rows (their criterion is the
        // operators' own), one lane per particle, sums in row order
        template<class K>
        __global__ void __launch_bounds__(BLK)
            k_divv_curlv(u32 n, Rows r, const FRec *__restrict rec, f64 mass,
                         f64 *__restrict divv, f64 *__restrict curlx, f64 *__restrict curly,
                         f64 *__restrict curlz) {
            using KK = Kern<K, f64>;
            u32 j    = xcd_block(blockIdx.x, gridDim.x) * BLK + threadIdx.x;
            if (j >= n) return;
            u32 row = r.sel[j], a = r.outi[row];
            const FRec A = rec[row];
            f64 inv_rho_omega_a = 1. / A.romega; // 1 / (omega_a rho_a)
            f64 sum_nabla_v = 0, sc[3] = {0, 0, 0};
            u32 k0 = r.off[row], k1 = k0 + r.cnt[row];
            for (u32 k = k0; k < k1; k++) {
                const FRec B = rec[r.ids[k]];
                f64 dr[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
                f64 rab2  = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                if (rab2 > rec_hR2<K>(A) && rab2 > rec_hR2<K>(B)) continue;
                f64 rab     = sqrt(rab2);
                f64 v_ab[3] = {A.vx - B.vx, A.vy - B.vy, A.vz - B.vz};
                f64 ru[3]   = {dr[0] / rab, dr[1] / rab, dr[2] / rab};
                if (rab < 1e-9) ru[0] = ru[1] = ru[2] = 0;
                f64 dW      = KK::dW(rab, A.h);
                f64 dWa[3]  = {dW * ru[0], dW * ru[1], dW * ru[2]};
                sum_nabla_v += mass * (v_ab[0] * dWa[0] + v_ab[1] * dWa[1] + v_ab[2] * dWa[2]);
                f64 c[3];
                cross3(v_ab, dWa, c);
                for (int t = 0; t < 3; t++) sc[t] = sc[t] + mass * c[t];
            }
            divv[a]  = -inv_rho_omega_a * sum_nabla_v;
            curlx[a] = -inv_rho_omega_a * sc[0];
            curly[a] = -inv_rho_omega_a * sc[1];
            curlz[a] = -inv_rho_omega_a * sc[2];
        }

        // d(div v)/dt (DiffOperatorDtDivv.cpp:89-175) with the merged accelerations in sorted
        // order; 3x3 inverse and products as in matrix_legacy.hpp:25-95 (rows stored)
        template<class K>
        __global__ void __launch_bounds__(BLK)
            k_dtdivv(u32 n, Rows r, const FRec *__restrict rec, const f64 *__restrict sax,
                     const f64 *__restrict say, const f64 *__restrict saz, f64 mass,
                     f64 *__restrict dtdivv) {
            using KK = Kern<K, f64>;
            u32 j    = xcd_block(blockIdx.x, gridDim.x) * BLK + threadIdx.x;
            if (j >= n) return;
            u32 row = r.sel[j], a = r.outi[row];
            const FRec A = rec[row];
            const f64 aa[3] = {sax[row], say[row], saz[row]};
            f64 R[3][3] = {}, Rv[3][3] = {}, Ra[3][3] = {};
            u32 k0 = r.off[row], k1 = k0 + r.cnt[row];
            for (u32 k = k0; k < k1; k++) {
                u32 b        = r.ids[k];
                const FRec B = rec[b];
                f64 rr[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
                f64 rab2  = rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
                if (rab2 > rec_hR2<K>(A) && rab2 > rec_hR2<K>(B)) continue;
                f64 rab     = sqrt(rab2);
                f64 v_ab[3] = {A.vx - B.vx, A.vy - B.vy, A.vz - B.vz};
                f64 a_ab[3] = {aa[0] - sax[b], aa[1] - say[b], aa[2] - saz[b]};
                f64 ru[3]   = {rr[0] / rab, rr[1] / rab, rr[2] / rab};
                if (rab < 1e-9) ru[0] = ru[1] = ru[2] = 0;
                f64 dW = KK::dW(rab, A.h);
                f64 m[3] = {dW * ru[0] * mass, dW * ru[1] * mass, dW * ru[2] * mass};
                for (int i = 0; i < 3; i++)
                    for (int c = 0; c < 3; c++) {
                        R[i][c]  = R[i][c] - rr[i] * m[c];
                        Rv[i][c] = Rv[i][c] - v_ab[c] * m[i];
                        Ra[i][c] = Ra[i][c] - a_ab[c] * m[i];
                    }
            }
            // inverse (matrix_legacy.hpp:25-50): aij = R[i][j] (element j of stored row i)
            f64 a00 = R[0][0], a10 = R[1][0], a20 = R[2][0];
            f64 a01 = R[0][1], a11 = R[1][1], a21 = R[2][1];
            f64 a02 = R[0][2], a12 = R[1][2], a22 = R[2][2];
            f64 det = (-a02 * a11 * a20 + a01 * a12 * a20 + a02 * a10 * a21 - a00 * a12 * a21 -
                       a01 * a10 * a22 + a00 * a11 * a22);
            f64 iv[3][3] = {{(-a12 * a21 + a11 * a22) / det, (a02 * a21 - a01 * a22) / det,
                             (-a02 * a11 + a01 * a12) / det},
                            {(a12 * a20 - a10 * a22) / det, (-a02 * a20 + a00 * a22) / det,
                             (a02 * a10 - a00 * a12) / det},
                            {(-a11 * a20 + a10 * a21) / det, (a01 * a20 - a00 * a21) / det,
                             (-a01 * a10 + a00 * a11) / det}};
            // products (matrix_legacy.hpp:53-95): out[i][c] = sum_k iv[i][k] B[k][c], k order
            auto prod = [&](const f64 B[3][3], f64 O[3][3]) {
                for (int i = 0; i < 3; i++)
                    for (int c = 0; c < 3; c++)
                        O[i][c] = iv[i][0] * B[0][c] + iv[i][1] * B[1][c] + iv[i][2] * B[2][c];
            };
            f64 dv[3][3], da[3][3];
            prod(Rv, dv);
            prod(Ra, da);
            f64 div_a = da[0][0] + da[1][1] + da[2][2];
            f64 tens  = dv[0][0] * dv[0][0] + dv[1][0] * dv[0][1] + dv[2][0] * dv[0][2] +
                       dv[0][1] * dv[1][0] + dv[1][1] * dv[1][1] + dv[2][1] * dv[1][2] +
                       dv[0][2] * dv[2][0] + dv[1][2] * dv[2][1] + dv[2][2] * dv[2][2];
            dtdivv[a] = div_a - tens;
        }

        // alpha update (UpdateViscosity.cpp:120-219), local particles
        __global__ void k_alpha_cd10(u32 n, f64 dt, const f64 *__restrict cs,
                                     const f64 *__restrict h, const f64 *__restrict alpha,
                                     const f64 *__restrict divv, const f64 *__restrict curlx,
                                     const f64 *__restrict curly, const f64 *__restrict curlz,
                                     const f64 *__restrict dtdivv, f64 alpha_min, f64 alpha_max,
                                     f64 sigma_decay, f64 *__restrict alpha_new) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            const f64 eps = 2.220446049250313080847e-16; // DBL_EPSILON
            f64 cs_a = cs[i], h_a = h[i], alpha_a = alpha[i], divv_a = divv[i];
            f64 cx = curlx[i], cy = curly[i], cz = curlz[i], dtdivv_a = dtdivv[i];
            f64 vsig            = cs_a;
            f64 inv_tau_a       = vsig * sigma_decay / h_a;
            f64 fact_t          = dt * inv_tau_a;
            f64 euler_impl_fact = 1 / (1 + fact_t);
            f64 fac = fmax(-divv_a, 0.);
            fac *= fac;
            f64 traceS  = cx * cx + cy * cy + cz * cz;
            f64 balsara = (fac + traceS > eps) ? fac / (fac + traceS) : 1.;
            f64 A_a     = balsara * fmax(-dtdivv_a, 0.);
            f64 temp    = cs_a * cs_a;
            f64 alpha_loc_a =
                fmin((cs_a > 0) ? 10 * h_a * h_a * A_a / (temp) : alpha_min, alpha_max);
            alpha_loc_a   = (temp > eps) ? alpha_loc_a : alpha_min;
            f64 new_alpha = (alpha_a + alpha_loc_a * fact_t) * euler_impl_fact;
            if (alpha_loc_a > alpha_a) new_alpha = alpha_loc_a;
            alpha_new[i] = new_alpha;
        }

        // alpha of every merged particle (images take their source's) into the force records
        __global__ void k_record_alpha(u32 nm, u32 N, const u32 *__restrict sort_index,
                                       const u32 *__restrict ghost_src,
                                       const f64 *__restrict alpha, FRec *__restrict rec) {
            u32 s = blockIdx.x * blockDim.x + threadIdx.x;
            if (s >= nm) return;
            u32 b    = sort_index[s];
            rec[s].alpha = alpha[b < N ? b : ghost_src[b - N]];
        }

        // ------------------------------------------------------------- sink-gas gravity
        __global__ void __launch_bounds__(BLK)
            k_sink_gas(u32 n, const f64 *__restrict x, const f64 *__restrict y,
                       const f64 *__restrict z, SinkArgs sk, int first, f64 *__restrict ex,
                       f64 *__restrict ey, f64 *__restrict ez, u64 *__restrict part) {
            u32 i  = blockIdx.x * BLK + threadIdx.x;
            f64 sx = 0, sy = 0, sz = 0;
            if (i < n) {
                f64 dx = x[i] - sk.x, dy = y[i] - sk.y, dz = z[i] - sk.z;
                f64 r2 = dx * dx + dy * dy + dz * dz;
                f64 d  = sqrt(r2);
                f64 ds = (sk.soft == 0) ? d : sqrt(r2 + sk.soft * sk.soft);
                f64 d3 = ds * ds * ds;
                f64 fx = sk.G * dx / d3, fy = sk.G * dy / d3, fz = sk.G * dz / d3;
                if (d < sk.racc) fx = fy = fz = 0;
                sx = fx * sk.gpart_mass;
                sy = fy * sk.gpart_mass;
                sz = fz * sk.gpart_mass;
                f64 bx = first ? 0. : ex[i], by = first ? 0. : ey[i], bz = first ? 0. : ez[i];
                ex[i] = bx + (-fx) * sk.mass;
                ey[i] = by + (-fy) * sk.mass;
                ez[i] = bz + (-fz) * sk.mass;
            }
            __shared__ f64 ldf[3][BLK / WAVE];
            auto add = [](f64 p, f64 q) { return p + q; };
            sx = block_reduce(sx, add, ldf[0]);
            sy = block_reduce(sy, add, ldf[1]);
            sz = block_reduce(sz, add, ldf[2]);
            if (threadIdx.x == 0) {
                part[blockIdx.x]                 = (u64) __double_as_longlong(sx);
                part[gridDim.x + blockIdx.x]     = (u64) __double_as_longlong(sy);
                part[2 * gridDim.x + blockIdx.x] = (u64) __double_as_longlong(sz);
            }
        }

        // ------------------------------------------------------------- partial reductions
        // slot k of the nb block partials is reduced in a fixed order (fixed chunks, then one
        // block), so the result, an FP64 sum included, does not depend on scheduling
        template<int OP>
        __device__ inline u64 red_combine(u64 p, u64 q) {
            if constexpr (OP == RED_MAX_KEY) return umax64(p, q);
            else if constexpr (OP == RED_MIN_KEY) return umin64(p, q);
            else if constexpr (OP == RED_SUM_U64) return p + q;
            else {
                f64 r = __longlong_as_double((long long) p) + __longlong_as_double((long long) q);
                return (u64) __double_as_longlong(r);
            }
        }
        template<int OP>
        __device__ inline u64 red_identity() {
            if constexpr (OP == RED_MAX_KEY) return ORDERED_NEG_INF;
            else if constexpr (OP == RED_MIN_KEY) return ORDERED_POS_INF;
            else return 0; // 0 (u64) and +0.0 (f64 bits)
        }
        constexpr int FIN_BLK   = 1024;
        constexpr u32 FIN_CHUNK = 8192; // partials per block of the first pass

        // block b reduces p[b * chunk, min((b + 1) * chunk, n)) into out[b]
        template<int OP>
        __device__ inline void fin_block(const u64 *__restrict p, u32 n, u32 chunk,
                                         u64 *__restrict out, u64 *lds) {
            u32 i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
            u64 acc = red_identity<OP>();
            for (u32 i = i0 + threadIdx.x; i < i1; i += FIN_BLK) acc = red_combine<OP>(acc, p[i]);
            for (int o = 32; o > 0; o >>= 1) acc = red_combine<OP>(acc, __shfl_xor(acc, o, WAVE));
            if (__lane_id() == 0) lds[threadIdx.x / WAVE] = acc;
            __syncthreads();
            if (threadIdx.x == 0) {
                u64 r = lds[0];
                for (int w = 1; w < FIN_BLK / WAVE; w++) r = red_combine<OP>(r, lds[w]);
                out[blockIdx.x] = r;
            }
            __syncthreads();
        }
        __global__ void __launch_bounds__(FIN_BLK)
            k_finalize(u32 nb, u32 chunk, const u64 *__restrict part, u64 stride, RedSpec sp,
                       u64 *__restrict out, u64 out_stride) {
            __shared__ u64 lds[FIN_BLK / WAVE];
            for (int k = 0; k < sp.nslot; k++) {
                const u64 *p = part + (u64) k * stride;
                u64 *o       = out + (u64) sp.out_slot[k] * out_stride;
                switch (sp.op[k]) {
                case RED_MAX_KEY: fin_block<RED_MAX_KEY>(p, nb, chunk, o, lds); break;
                case RED_MIN_KEY: fin_block<RED_MIN_KEY>(p, nb, chunk, o, lds); break;
                case RED_SUM_U64: fin_block<RED_SUM_U64>(p, nb, chunk, o, lds); break;
                default: fin_block<RED_SUM_F64>(p, nb, chunk, o, lds); break;
                }
            }
        }

        // ------------------------------------------------------------- integrators
        __device__ inline f64 pos_mod(f64 v, f64 bmin, f64 delt) {
            f64 r = v - bmin;
            r     = fmod(r, delt);
            r += delt;
            r = fmod(r, delt);
            r += bmin;
            return r;
        }

        __global__ void k_predictor(u32 n, f64 dt, f64 dt_half, f64 *__restrict x,
                                    f64 *__restrict y, f64 *__restrict z, f64 *__restrict vx,
                                    f64 *__restrict vy, f64 *__restrict vz,
                                    const f64 *__restrict ax, const f64 *__restrict ay,
                                    const f64 *__restrict az, f64 *__restrict u,
                                    const f64 *__restrict du, int periodic,
                                    const f64 *__restrict box6) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            f64 a0 = ax[i], a1 = ay[i], a2 = az[i], d = du[i];
            f64 v0 = vx[i] + dt_half * a0, v1 = vy[i] + dt_half * a1, v2 = vz[i] + dt_half * a2;
            f64 uu = u[i] + dt_half * d;
            f64 p0 = x[i] + dt * v0, p1 = y[i] + dt * v1, p2 = z[i] + dt * v2;
            v0 = v0 + dt_half * a0;
            v1 = v1 + dt_half * a1;
            v2 = v2 + dt_half * a2;
            uu = uu + dt_half * d;
            if (periodic) {
                p0 = pos_mod(p0, box6[0], box6[3] - box6[0]);
                p1 = pos_mod(p1, box6[1], box6[4] - box6[1]);
                p2 = pos_mod(p2, box6[2], box6[5] - box6[2]);
            }
            x[i]  = p0;
            y[i]  = p1;
            z[i]  = p2;
            vx[i] = v0;
            vy[i] = v1;
            vz[i] = v2;
            u[i]  = uu;
        }

        // periodic wrap alone (integrators.cpp:206-236), for steps whose sink forces must see
        // the drifted positions before the wrap
        __global__ void k_wrap(u32 n, f64 *__restrict x, f64 *__restrict y, f64 *__restrict z,
                               const f64 *__restrict box6) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= n) return;
            x[i] = pos_mod(x[i], box6[0], box6[3] - box6[0]);
            y[i] = pos_mod(y[i], box6[1], box6[4] - box6[1]);
            z[i] = pos_mod(z[i], box6[2], box6[5] - box6[2]);
        }

        __global__ void __launch_bounds__(BLK)
            k_corrector(u32 n, f64 hdt, f64 *__restrict vx, f64 *__restrict vy,
                        f64 *__restrict vz, const f64 *__restrict ax, const f64 *__restrict ay,
                        const f64 *__restrict az, const f64 *__restrict ax_old,
                        const f64 *__restrict ay_old, const f64 *__restrict az_old,
                        f64 *__restrict u, const f64 *__restrict du,
                        const f64 *__restrict du_old, f64 mass, u64 *__restrict part) {
            u32 i    = blockIdx.x * BLK + threadIdx.x;
            u64 kmax = ORDERED_NEG_INF;
            f64 vsq = 0, en = 0;
            if (i < n) {
                f64 ix = hdt * (ax[i] - ax_old[i]), iy = hdt * (ay[i] - ay_old[i]),
                    iz = hdt * (az[i] - az_old[i]);
                f64 v0 = vx[i] + ix, v1 = vy[i] + iy, v2 = vz[i] + iz;
                vx[i]  = v0;
                vy[i]  = v1;
                vz[i]  = v2;
                kmax   = f64_to_ordered(ix * ix + iy * iy + iz * iz);
                f64 ui = hdt * (du[i] - du_old[i]);
                f64 uu = u[i] + ui;
                u[i]   = uu;
                vsq    = v0 * v0 + v1 * v1 + v2 * v2;
                en     = mass * (uu + 0.5 * vsq);
            }
            __shared__ u64 lds[BLK / WAVE];
            __shared__ f64 ldf[2][BLK / WAVE];
            auto add = [](f64 p, f64 q) { return p + q; };
            kmax = block_reduce(kmax, umax64, lds);
            vsq  = block_reduce(vsq, add, ldf[0]);
            en   = block_reduce(en, add, ldf[1]);
            if (threadIdx.x == 0) {
                part[blockIdx.x]                 = kmax;
                part[gridDim.x + blockIdx.x]     = (u64) __double_as_longlong(vsq);
                part[2 * gridDim.x + blockIdx.x] = (u64) __double_as_longlong(en);
            }
        }

        // CFL dt per particle and per-block minima (the expressions of force_store)
        __global__ void __launch_bounds__(BLK)
            k_cfl(u32 n, const f64 *__restrict h, const f64 *__restrict vsig,
                  const f64 *__restrict ax, const f64 *__restrict ay, const f64 *__restrict az,
                  f64 C_cour, f64 C_force, f64 *__restrict dt_cfl, u64 *__restrict part) {
            __shared__ u64 lds[BLK / WAVE];
            u32 i   = blockIdx.x * BLK + threadIdx.x;
            u64 key = ORDERED_POS_INF;
            if (i < n) {
                f64 fx = ax[i], fy = ay[i], fz = az[i], h_a = h[i];
                f64 dt_c   = C_cour * h_a / vsig[i];
                f64 abs_a  = sqrt(fx * fx + fy * fy + fz * fz);
                f64 dt_f   = C_force * sqrt(h_a / abs_a);
                f64 cfl    = fmin(__builtin_inf(), dt_c);
                cfl        = fmin(cfl, dt_f);
                dt_cfl[i]  = cfl;
                key        = f64_to_ordered(cfl);
            }
            key = block_reduce(key, umin64, lds);
            if (threadIdx.x == 0) part[blockIdx.x] = key;
        }

        // (one same-address atomic per block: a few hundred blocks)
        __global__ void __launch_bounds__(BLK) k_max_key(u32 n, const f64 *__restrict v,
                                                         u64 *__restrict key) {
            __shared__ u64 lds[BLK / WAVE];
            u64 k = ORDERED_NEG_INF;
            for (u32 i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK)
                k = umax64(k, f64_to_ordered(v[i]));
            k = block_reduce(k, umax64, lds);
            if (threadIdx.x == 0) atomic_max_key(key, k);
        }

        // ------------------------------------------------------------- periodic images
        struct CutBoxes {
            f64 lo[26][3], hi[26][3], off[26][3];
            u32 active[26];
        };

        // the 26 image offsets and cut volumes of the periodic box (BasicSPHGhosts.cpp:260-509),
        // offset o in (x, y, z) nesting order without (0, 0, 0), from the interaction range
        // (hmax * f1) * f2 with hmax read from the device (no host round trip)
        struct Box6 {
            f64 v[6];
        };
        __global__ void k_cut_boxes(Box6 box, const u64 *__restrict hmax_key, f64 f1, f64 f2,
                                    CutBoxes *__restrict cb) {
            const int o = threadIdx.x;
            if (o >= 26) return;
            const int q     = o < 13 ? o : o + 1;
            const int io[3] = {q / 9 - 1, (q / 3) % 3 - 1, q % 3 - 1};
            const f64 int_range = (ordered_to_f64(*hmax_key) * f1) * f2;
            bool nonempty = true;
            for (int k = 0; k < 3; k++) {
                f64 bmin = box.v[k], bmax = box.v[3 + k];
                f64 bsize = bmax - bmin, pmin = bmin, pmax = bsize + bmin;
                f64 off = io[k] * bsize;
                f64 rlo = pmin - int_range, rhi = pmax + int_range;
                f64 ilo = fmax(rlo, pmin + off), ihi = fmin(rhi, pmax + off);
                if (!(ihi >= ilo)) nonempty = false;
                cb->off[o][k] = off;
                cb->lo[o][k]  = fmax(pmin, rlo - off);
                cb->hi[o][k]  = fmin(pmax, rhi - off);
            }
            cb->active[o] = nonempty ? 1u : 0u;
        }

        __device__ inline bool in_cut(const CutBoxes &cb, int o, f64 px, f64 py, f64 pz) {
            // cb.active[o] is wave-uniform (its branch skips empty offsets); the coordinate
            // tests are combined without short-circuit branches
            return cb.active[o] && ((cb.lo[o][0] <= px) & (px < cb.hi[o][0]) &
                                    (cb.lo[o][1] <= py) & (py < cb.hi[o][1]) &
                                    (cb.lo[o][2] <= pz) & (pz < cb.hi[o][2]));
        }

        __global__ void __launch_bounds__(BLK)
            k_ghost_count(u32 n, const f64 *__restrict x, const f64 *__restrict y,
                          const f64 *__restrict z, const CutBoxes *__restrict cbp,
                          u32 *__restrict block_counts, u32 nb) {
            __shared__ u32 cnt[26];
            if (threadIdx.x < 26) cnt[threadIdx.x] = 0;
            __syncthreads();
            const CutBoxes &cb = *cbp;
            u32 i      = blockIdx.x * BLK + threadIdx.x;
            bool valid = i < n;
            f64 px = valid ? x[i] : 0, py = valid ? y[i] : 0, pz = valid ? z[i] : 0;
            for (int o = 0; o < 26; o++) {
                u64 m = __ballot(valid && in_cut(cb, o, px, py, pz));
                if (__lane_id() == 0 && m) atomicAdd(&cnt[o], (u32) __popcll(m));
            }
            __syncthreads();
            if (threadIdx.x < 26) block_counts[(u64) threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
        }

        __global__ void __launch_bounds__(BLK)
            k_ghost_scatter(u32 n, f64 *__restrict x, f64 *__restrict y, f64 *__restrict z,
                            f64 *__restrict h, const CutBoxes *__restrict cbp,
                            const u32 *__restrict block_base, u32 nb, u32 *__restrict src,
                            u8 *__restrict oid) {
            __shared__ u32 wave_cnt[26][BLK / WAVE];
            const CutBoxes &cb = *cbp;
            u32 i      = blockIdx.x * BLK + threadIdx.x;
            u32 lane = __lane_id(), wid = threadIdx.x / WAVE;
            bool valid = i < n;
            f64 px = valid ? x[i] : 0, py = valid ? y[i] : 0, pz = valid ? z[i] : 0;
            f64 ph = valid ? h[i] : 0;
            u64 lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
            u32 in_mask = 0, rank[26];
            for (int o = 0; o < 26; o++) {
                bool in = valid && in_cut(cb, o, px, py, pz);
                u64 m   = __ballot(in);
                if (in) in_mask |= 1u << o;
                rank[o] = (u32) __popcll(m & lt);
                if (lane == 0) wave_cnt[o][wid] = (u32) __popcll(m);
            }
            __syncthreads();
            for (int o = 0; o < 26; o++) {
                if (!(in_mask & (1u << o))) continue;
                u32 pos = block_base[(u64) o * nb + blockIdx.x] + rank[o];
                for (u32 w = 0; w < wid; w++) pos += wave_cnt[o][w];
                src[pos]       = i;
                oid[pos]       = (u8) o;
                u32 dst        = n + pos;
                x[dst]         = px + cb.off[o][0];
                y[dst]         = py + cb.off[o][1];
                z[dst]         = pz + cb.off[o][2];
                h[dst]         = ph;
            }
        }

        __global__ void k_gather(u32 n, u32 G, const u32 *__restrict src, FieldPtrs fp) {
            u32 g = blockIdx.x * blockDim.x + threadIdx.x;
            if (g >= G) return;
            u32 s = src[g];
            for (int f = 0; f < fp.n; f++) fp.p[f][n + g] = fp.p[f][s];
        }
    } // namespace

    namespace {
        // lanes per particle in the neighbour loops: 1, 2, 4 or 16 (SHAMX_PAIR_LANES, read at
        // every call so a test can compare the settings in one process; the FP64 sums are taken
        // in list order whatever the choice, so results do not depend on it)
        int pair_lanes() {
            const char *e = std::getenv("SHAMX_PAIR_LANES");
            int v         = e ? std::atoi(e) : 4;
            return (v == 1 || v == 2 || v == 4 || v == 16) ? v : 4;
        }
        template<class F>
        void dispatch_lanes(F &&f) {
            switch (pair_lanes()) {
            case 1: f(std::integral_constant<int, 1>{}); break;
            case 2: f(std::integral_constant<int, 2>{}); break;
            case 4: f(std::integral_constant<int, 4>{}); break;
            case 16: f(std::integral_constant<int, 16>{}); break;
            default: f(std::integral_constant<int, 4>{}); break;
            }
        }

        Rows rows_of(const Csr &c) { return Rows{c.sel, c.outi, c.cnt, c.offset, c.ids}; }

        // partials of slot k at part[k * nb + b]; results to out[sp.out_slot[k]].  Large
        // counts go through a first pass of FIN_CHUNK-sized chunks whose results go to mid
        // (sp.nslot * div_up(nb, FIN_CHUNK) words that must not overlap the partials).
        void finalize(hipStream_t s, u32 nb, const u64 *part, RedSpec sp, u64 *out, u64 *mid) {
            if (nb <= FIN_CHUNK) {
                k_finalize<<<1, FIN_BLK, 0, s>>>(nb, nb, part, nb, sp, out, 1);
            } else {
                u32 nb2     = div_up(nb, FIN_CHUNK);
                RedSpec sp1 = sp;
                for (int k = 0; k < sp.nslot; k++) sp1.out_slot[k] = k;
                k_finalize<<<nb2, FIN_BLK, 0, s>>>(nb, FIN_CHUNK, part, nb, sp1, mid, nb2);
                k_finalize<<<1, FIN_BLK, 0, s>>>(nb2, nb2, mid, nb2, sp, out, 1);
            }
            SHAMX_HIP(hipGetLastError());
        }
        // words of partial storage for nslot slots of nb blocks, followed by the first-pass room
        size_t part_words(int nslot, u32 nb) { return (size_t) nslot * (nb + div_up(nb, FIN_CHUNK)); }
    } // namespace

    namespace {
        // one Newton sweep over n rows; partials of max eps, min eps and the active count in
        // part (3 slots of div_up(n G, BLK) blocks, then the finalize room)
        void h_sweep(hipStream_t s, int kernel, u32 n, Rows r, const double4 *ps, const f64 *x,
                     const f64 *y, const f64 *z, const f64 *h_old, f64 *h_new, f64 *eps,
                     f64 mass, f64 h_evol_max, f64 h_evol_iter_max, int mixed,
                     DevArray<u64> &part, const u64 *stop) {
            u32 nb = div_up((u64) n * pair_lanes(), BLK);
            part.resize(std::max<size_t>(part.size(), part_words(3, std::max<u32>(nb, 1))));
            if (n == 0) return;
            dispatch_kernel(kernel, [&](auto K) {
                using KT = decltype(K);
                dispatch_lanes([&](auto GC) {
                    constexpr int G = decltype(GC)::value;
                    auto go         = [&](auto src) {
                        if (mixed)
                            k_h_iterate<KT, G, true><<<nb, BLK, 0, s>>>(
                                n, r, src, h_old, h_new, eps, mass, h_evol_max, h_evol_iter_max,
                                part.data(), stop);
                        else
                            k_h_iterate<KT, G, false><<<nb, BLK, 0, s>>>(
                                n, r, src, h_old, h_new, eps, mass, h_evol_max, h_evol_iter_max,
                                part.data(), stop);
                    };
                    if (ps) go(PosRec{ps});
                    else go(PosSoA{x, y, z});
                });
            });
            SHAMX_HIP(hipGetLastError());
        }
    } // namespace

    void h_iterate(hipStream_t s, int kernel, u32 n, Csr c, const double4 *ps, const f64 *x,
                   const f64 *y, const f64 *z, const f64 *h_old, f64 *h_new, f64 *eps, f64 mass,
                   f64 h_evol_max, f64 h_evol_iter_max, u64 *eps_minmax, int mixed,
                   DevArray<u64> &part) {
        if (n == 0) return;
        h_sweep(s, kernel, n, rows_of(c), ps, x, y, z, h_old, h_new, eps, mass, h_evol_max,
                h_evol_iter_max, mixed, part, nullptr);
        finalize(s, div_up((u64) n * pair_lanes(), BLK), part.data(),
                 RedSpec{2, {RED_MAX_KEY, RED_MIN_KEY}, {0, 1}}, eps_minmax,
                 part.data() + 3 * (size_t) div_up((u64) n * pair_lanes(), BLK));
    }

    HLoopResult h_iterate_loop(hipStream_t s, int kernel, u32 n, Csr c, const double4 *ps,
                               const f64 *h_old, f64 *h_new, f64 *eps, f64 mass,
                               f64 h_evol_max, f64 h_evol_iter_max, f64 epsilon_h,
                               u32 max_iter, int mixed, HLoopBuffers &b) {
        HLoopResult res{};
        if (n == 0 || max_iter == 0) return res;
        b.st.resize(HL_WORDS);
        k_h_loop_init<<<1, 64, 0, s>>>(b.st.data());
        const u32 *sel = c.sel;
        u32 n_act      = n;
        u32 done = 0, batch = 1;
        static_assert(HL_WORDS <= decltype(b.host)::size(), "h-loop words exceed the host copy");
        u64 *host = b.host.data();
        while (done < max_iter) {
            u32 k = std::min(batch, max_iter - done);
            for (u32 i = 0; i < k; i++) {
                Rows r{sel, c.outi, c.cnt, c.offset, c.ids};
                h_sweep(s, kernel, n_act, r, ps, nullptr, nullptr, nullptr, h_old, h_new, eps,
                        mass, h_evol_max, h_evol_iter_max, mixed, b.part, b.st.data());
                u32 nb = div_up((u64) n_act * pair_lanes(), BLK);
                finalize(s, nb, b.part.data(),
                         RedSpec{3, {RED_MAX_KEY, RED_MIN_KEY, RED_SUM_U64},
                                 {HL_SWEEP_MAX, HL_SWEEP_MIN, HL_SWEEP_ACT}},
                         b.st.data(), b.part.data() + 3 * (size_t) nb);
                k_h_loop_ctl<<<1, 64, 0, s>>>(b.st.data(), epsilon_h);
            }
            SHAMX_HIP(hipGetLastError());
            b.host.read_async(s, b.st.data(), HL_WORDS);
            SHAMX_HIP(hipStreamSynchronize(s));
            done += k;
            res.host_reads++;
            if (host[HL_STOP]) break;
            // the particles still iterating are few: sweep only their rows from now on
            u32 keep = (u32) host[HL_ACT];
            if (done < max_iter && (u64) keep * 4 < n_act) {
                b.flag.resize(std::max<u32>(n_act, 1));
                b.off.resize(std::max<u32>(n_act, 1));
                DevArray<u32> &dst = (sel == b.sel_a.data()) ? b.sel_b : b.sel_a;
                dst.resize(std::max<u32>(keep, 1));
                b.scan.resize(std::max(b.scan.size(), scan_scratch_bytes(n_act)));
                k_h_active_flags<<<std::min<u32>(256, div_up(n_act, BLK)), BLK, 0, s>>>(
                    n_act, sel, c.outi, eps, b.flag.data(), b.st.data());
                exclusive_scan_u32(s, b.flag.data(), b.off.data(), n_act, nullptr, b.scan.data());
                k_h_compact_rows<<<div_up(n_act, 256), 256, 0, s>>>(n_act, sel, b.flag.data(),
                                                                   b.off.data(), dst.data());
                SHAMX_HIP(hipGetLastError());
                sel   = dst.data();
                n_act = keep;
                // (with no row left the remaining sweeps launch nothing; their reduction over
                // no partials leaves max / min to the frozen extrema, and they still count,
                // as the reference's sweeps over converged particles do)
                res.compactions++;
            }
            batch = std::min<u32>(2 * batch, 8);
        }
        res.iterations = (u32) host[HL_ITERS];
        res.max_eps    = ordered_to_f64(host[HL_MAX]);
        res.min_eps    = ordered_to_f64(host[HL_MIN]);
        return res;
    }

    void compute_omega(hipStream_t s, int kernel, u32 n, Csr c, const double4 *ps, const f64 *x,
                       const f64 *y, const f64 *z, const f64 *h, f64 *omega, f64 mass,
                       u32 *fcnt, u32 *fids, int mixed) {
        if (n == 0) return;
        if (fids && !ps) throw std::invalid_argument("compute_omega: the filter needs records");
        u32 nb = div_up((u64) n * pair_lanes(), BLK);
        Rows r = rows_of(c);
        dispatch_kernel(kernel, [&](auto K) {
            using KT = decltype(K);
            dispatch_lanes([&](auto GC) {
                constexpr int G = decltype(GC)::value;
                if (ps && fids && mixed)
                    k_omega<KT, G, true, true><<<nb, BLK, 0, s>>>(n, r, PosRec{ps}, h, omega,
                                                                  mass, fcnt, fids);
                else if (ps && fids)
                    k_omega<KT, G, true, false><<<nb, BLK, 0, s>>>(n, r, PosRec{ps}, h, omega,
                                                                   mass, fcnt, fids);
                else if (ps)
                    k_omega<KT, G, false, false><<<nb, BLK, 0, s>>>(n, r, PosRec{ps}, h, omega,
                                                                    mass, nullptr, nullptr);
                else
                    k_omega<KT, G, false, false><<<nb, BLK, 0, s>>>(n, r, PosSoA{x, y, z}, h,
                                                                    omega, mass, nullptr,
                                                                    nullptr);
            });
        });
        SHAMX_HIP(hipGetLastError());
    }

    void ps_refresh_h(hipStream_t s, u32 nm, const u32 *sort_index, const f64 *h, double4 *ps) {
        if (nm == 0) return;
        k_ps_refresh_h<<<div_up(nm, 256), 256, 0, s>>>(nm, sort_index, h, ps);
        SHAMX_HIP(hipGetLastError());
    }

    void compute_eos(hipStream_t s, int kernel, u32 nm, const f64 *h, const f64 *u,
                     const f64 *x, const f64 *y, const f64 *z, f64 *P, f64 *cs, f64 mass,
                     EosArgs e) {
        if (nm == 0) return;
        dispatch_kernel(kernel, [&](auto K) {
            k_eos<decltype(K)><<<div_up(nm, 256), 256, 0, s>>>(nm, h, u, x, y, z, P, cs, mass, e);
        });
        SHAMX_HIP(hipGetLastError());
    }

    size_t force_record_bytes(int mixed) { return mixed ? sizeof(FRec32) : sizeof(FRec); }

    void pack_force(hipStream_t s, int kernel, u32 nm, const u32 *perm, const f64 *x,
                    const f64 *y, const f64 *z, const f64 *h, const f64 *vx, const f64 *vy,
                    const f64 *vz, const f64 *u, const f64 *omega, f64 *P, f64 *cs, void *rec,
                    f64 mass, EosArgs e, int mixed, bool refresh) {
        if (nm == 0) return;
        dispatch_kernel(kernel, [&](auto K) {
            using KT = decltype(K);
            if (refresh && mixed)
                k_pack_force<KT, true, true><<<div_up(nm, 256), 256, 0, s>>>(
                    nm, perm, x, y, z, h, vx, vy, vz, u, omega, P, cs, rec, mass, e);
            else if (refresh)
                k_pack_force<KT, false, true><<<div_up(nm, 256), 256, 0, s>>>(
                    nm, perm, x, y, z, h, vx, vy, vz, u, omega, P, cs, rec, mass, e);
            else if (mixed)
                k_pack_force<KT, true><<<div_up(nm, 256), 256, 0, s>>>(
                    nm, perm, x, y, z, h, vx, vy, vz, u, omega, P, cs, rec, mass, e);
            else
                k_pack_force<KT, false><<<div_up(nm, 256), 256, 0, s>>>(
                    nm, perm, x, y, z, h, vx, vy, vz, u, omega, P, cs, rec, mass, e);
        });
        SHAMX_HIP(hipGetLastError());
    }

    void compute_force(hipStream_t s, int kernel, u32 n, Csr c, const void *rec, const f64 *x,
                       const f64 *y, const f64 *z, const f64 *vx, const f64 *vy, const f64 *vz,
                       const f64 *h, const f64 *omega, const f64 *u, const f64 *P, const f64 *cs,
                       f64 *ax, f64 *ay, f64 *az, f64 *du, f64 *ax_old, f64 *ay_old,
                       f64 *az_old, f64 *du_old, f64 *vsig, f64 *dt_cfl, ForceArgs fa,
                       u64 *dt_min_key, u64 *pair_count, DevArray<u64> &part) {
        if (n == 0) return;
        u32 nb = div_up((u64) n * pair_lanes(), BLK);
        part.resize(part_words(2, nb));
        Rows r = rows_of(c);
        ForceOut o{ax, ay, az, du, ax_old, ay_old, az_old, du_old, vsig, dt_cfl};
        dispatch_kernel(kernel, [&](auto K) {
            using KT = decltype(K);
            dispatch_lanes([&](auto GC) {
                constexpr int G = decltype(GC)::value;
                if (fa.mixed) {
                    if (rec)
                        k_force_mixed<KT, G><<<nb, BLK, 0, s>>>(
                            n, r, RecSrc<FRec32>{(const FRec32 *) rec}, h, o, fa, part.data());
                    else
                        k_force_mixed<KT, G><<<nb, BLK, 0, s>>>(
                            n, r,
                            SoASrc<KT, FRec32>{x, y, z, vx, vy, vz, h, omega, u, P, cs, fa.mass},
                            h, o, fa, part.data());
                } else {
                    if (rec)
                        k_force<KT, G><<<nb, BLK, 0, s>>>(
                            n, r, RecSrc<FRec>{(const FRec *) rec}, o, fa, part.data());
                    else
                        k_force<KT, G><<<nb, BLK, 0, s>>>(
                            n, r,
                            SoASrc<KT, FRec>{x, y, z, vx, vy, vz, h, omega, u, P, cs, fa.mass}, o,
                            fa, part.data());
                }
            });
        });
        SHAMX_HIP(hipGetLastError());
        // dt_min and the pair count may live in different buffers: reduce each slot separately;
        // both calls use the room after the two slots in turn (stream order)
        u64 *mid = part.data() + 2 * (size_t) nb;
        finalize(s, nb, part.data(), RedSpec{1, {RED_MIN_KEY}, {0}}, dt_min_key, mid);
        if (pair_count)
            finalize(s, nb, part.data() + nb, RedSpec{1, {RED_SUM_U64}, {0}}, pair_count, mid);
    }

    void predictor(hipStream_t s, u32 n, f64 dt, f64 *x, f64 *y, f64 *z, f64 *vx, f64 *vy,
                   f64 *vz, const f64 *ax, const f64 *ay, const f64 *az, f64 *u, const f64 *du,
                   int periodic, const f64 *box6) {
        if (n == 0) return;
        k_predictor<<<div_up(n, 256), 256, 0, s>>>(n, dt, dt / 2, x, y, z, vx, vy, vz, ax, ay,
                                                   az, u, du, periodic, box6);
        SHAMX_HIP(hipGetLastError());
    }

    void corrector(hipStream_t s, u32 n, f64 hdt, f64 *vx, f64 *vy, f64 *vz, const f64 *ax,
                   const f64 *ay, const f64 *az, const f64 *ax_old, const f64 *ay_old,
                   const f64 *az_old, f64 *u, const f64 *du, const f64 *du_old, f64 mass,
                   u64 *red, DevArray<u64> &part) {
        if (n == 0) return;
        u32 nb = div_up(n, BLK);
        part.resize(part_words(3, nb));
        k_corrector<<<nb, BLK, 0, s>>>(n, hdt, vx, vy, vz, ax, ay, az, ax_old, ay_old, az_old, u,
                                       du, du_old, mass, part.data());
        SHAMX_HIP(hipGetLastError());
        finalize(s, nb, part.data(),
                 RedSpec{3, {RED_MAX_KEY, RED_SUM_F64, RED_SUM_F64}, {0, 1, 2}}, red,
                 part.data() + 3 * (size_t) nb);
    }

    void wrap_positions(hipStream_t s, u32 n, f64 *x, f64 *y, f64 *z, const f64 *box6) {
        if (n == 0) return;
        k_wrap<<<div_up(n, 256), 256, 0, s>>>(n, x, y, z, box6);
        SHAMX_HIP(hipGetLastError());
    }

    void sink_gas(hipStream_t s, u32 n, const f64 *x, const f64 *y, const f64 *z, SinkArgs sk,
                  int first, f64 *ex, f64 *ey, f64 *ez, u64 *sum_out, DevArray<u64> &part) {
        if (n == 0) {
            fill_u64(s, sum_out, 0, 3);
            return;
        }
        u32 nb = div_up(n, BLK);
        part.resize(part_words(3, nb));
        k_sink_gas<<<nb, BLK, 0, s>>>(n, x, y, z, sk, first, ex, ey, ez, part.data());
        SHAMX_HIP(hipGetLastError());
        finalize(s, nb, part.data(), RedSpec{3, {RED_SUM_F64, RED_SUM_F64, RED_SUM_F64}, {0, 1, 2}},
                 sum_out, part.data() + 3 * (size_t) nb);
    }

    void cd10_alpha(hipStream_t s, int kernel, u32 n, u32 nm, Csr c, const void *rec,
                    const u32 *sort_index, const u32 *ghost_src, const f64 *const *sorted_acc,
                    const f64 *cs_saved, const f64 *h, const f64 *alpha, f64 dt, f64 mass,
                    f64 alpha_min, f64 alpha_max, f64 sigma_decay, f64 *const *work6,
                    f64 *alpha_new) {
        if (n == 0) return;
        u32 nb = div_up(n, BLK);
        Rows r = rows_of(c);
        const FRec *fr = (const FRec *) rec;
        dispatch_kernel(kernel, [&](auto K) {
            using KT = decltype(K);
            k_divv_curlv<KT><<<nb, BLK, 0, s>>>(n, r, fr, mass, work6[0], work6[1], work6[2],
                                                work6[3]);
            k_dtdivv<KT><<<nb, BLK, 0, s>>>(n, r, fr, sorted_acc[0], sorted_acc[1],
                                            sorted_acc[2], mass, work6[4]);
        });
        k_alpha_cd10<<<div_up(n, 256), 256, 0, s>>>(n, dt, cs_saved, h, alpha, work6[0],
                                                    work6[1], work6[2], work6[3], work6[4],
                                                    alpha_min, alpha_max, sigma_decay,
                                                    alpha_new);
        k_record_alpha<<<div_up(nm, 256), 256, 0, s>>>(nm, n, sort_index, ghost_src, alpha_new,
                                                       (FRec *) rec);
        SHAMX_HIP(hipGetLastError());
    }

    void cfl_dt(hipStream_t s, u32 n, const f64 *h, const f64 *vsig, const f64 *ax,
                const f64 *ay, const f64 *az, f64 C_cour, f64 C_force, f64 *dt_cfl,
                u64 *dt_min_key, DevArray<u64> &part) {
        if (n == 0) return;
        u32 nb = div_up(n, BLK);
        part.resize(part_words(1, nb));
        k_cfl<<<nb, BLK, 0, s>>>(n, h, vsig, ax, ay, az, C_cour, C_force, dt_cfl, part.data());
        SHAMX_HIP(hipGetLastError());
        finalize(s, nb, part.data(), RedSpec{1, {RED_MIN_KEY}, {0}}, dt_min_key,
                 part.data() + nb);
    }

    void max_key_f64(hipStream_t s, u32 n, const f64 *v, u64 *key) {
        if (n == 0) return;
        u32 nb = std::min<u32>(div_up(n, BLK), 256);
        k_max_key<<<nb, BLK, 0, s>>>(n, v, key);
        SHAMX_HIP(hipGetLastError());
    }

    // cut volumes are computed on the host exactly as find_interfaces does for one patch:
    // offset o = (xo, yo, zo) in lexicographic order without (0,0,0); interf volume =
    // sender box ∩ (receiver box expanded by int_range - offset); membership half-open.
    u32 periodic_ghosts(hipStream_t s, u32 n, f64 *x, f64 *y, f64 *z, f64 *h, const f64 *box6,
                        const u64 *hmax_key, f64 f1, f64 f2, DevArray<u32> &ghost_src,
                        DevArray<u8> &ghost_off_id, DevArray<u32> &work,
                        DevArray<unsigned char> &scratch, u32 cap, u32 *count_host) {
        Box6 box{{box6[0], box6[1], box6[2], box6[3], box6[4], box6[5]}};
        u32 nb = div_up(std::max<u32>(n, 1), BLK);
        // work: [cutboxes (as u32 words)] [26*nb counts] [26*nb bases] [total]
        size_t cb_words = (sizeof(CutBoxes) + 3) / 4;
        work.resize(cb_words + 2 * (size_t) 26 * nb + 4);
        u32 *w           = work.data();
        CutBoxes *cbd    = (CutBoxes *) w;
        u32 *counts      = w + cb_words;
        u32 *bases       = counts + (size_t) 26 * nb;
        u32 *total       = bases + (size_t) 26 * nb;
        scratch.resize(std::max<size_t>(scratch.size(), scan_scratch_bytes(26 * nb)));
        k_cut_boxes<<<1, 32, 0, s>>>(box, hmax_key, f1, f2, cbd);
        k_ghost_count<<<nb, BLK, 0, s>>>(n, x, y, z, cbd, counts, nb);
        exclusive_scan_u32(s, counts, bases, 26 * nb, total, scratch.data());
        SHAMX_HIP(hipMemcpyAsync(count_host, total, sizeof(u32), hipMemcpyDeviceToHost, s));
        SHAMX_HIP(hipStreamSynchronize(s));
        const u32 G = *count_host;
        if (G > cap) return G; // caller grows the merged arrays and calls again
        ghost_src.resize(std::max<u32>(G, 1));
        ghost_off_id.resize(std::max<u32>(G, 1));
        k_ghost_scatter<<<nb, BLK, 0, s>>>(n, x, y, z, h, cbd, bases, nb, ghost_src.data(),
                                           ghost_off_id.data());
        SHAMX_HIP(hipGetLastError());
        return G;
    }

    // ---- storage order: local particles kept in the order of the last tree's sort
    namespace {
        // flag[k] = 1 when sorted position k holds a local particle (index < n)
        __global__ void k_local_flags(u32 nm, u32 n, const u32 *__restrict sort_index,
                                      u32 *__restrict flag) {
            u32 k = blockIdx.x * blockDim.x + threadIdx.x;
            if (k < nm) flag[k] = sort_index[k] < n ? 1u : 0u;
        }
        __global__ void k_local_compact(u32 nm, const u32 *__restrict sort_index,
                                        const u32 *__restrict flag, const u32 *__restrict off,
                                        u32 *__restrict perm) {
            u32 k = blockIdx.x * blockDim.x + threadIdx.x;
            if (k < nm && flag[k]) perm[off[k]] = sort_index[k];
        }
        // entries of perm that are not the identity (few blocks: one atomic per block)
        __global__ void __launch_bounds__(256)
            k_perm_moved(u32 n, const u32 *__restrict perm, u32 *__restrict count) {
            u32 c = 0;
            for (u32 j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256)
                c += perm[j] != j ? 1u : 0u;
            c = wave_reduce(c, [](u32 a, u32 b) { return a + b; });
            __shared__ u32 lds[256 / WAVE];
            if (__lane_id() == 0) lds[threadIdx.x / WAVE] = c;
            __syncthreads();
            if (threadIdx.x == 0) {
                u32 t = 0;
                for (int w = 0; w < 256 / WAVE; w++) t += lds[w];
                if (t) atomicAdd(count, t);
            }
        }
        template<class T>
        __global__ void k_perm_gather(u32 n, const u32 *__restrict perm, const T *__restrict src,
                                      T *__restrict dst) {
            u32 j = blockIdx.x * blockDim.x + threadIdx.x;
            if (j < n) dst[j] = src[perm[j]];
        }
        template<class T>
        __global__ void k_perm_scatter(u32 n, const u32 *__restrict perm, const T *__restrict src,
                                       T *__restrict dst) {
            u32 j = blockIdx.x * blockDim.x + threadIdx.x;
            if (j < n) dst[perm[j]] = src[j];
        }
    } // namespace

    void perm_moved_async(hipStream_t s, u32 n, const u32 *perm, u32 *count_dev) {
        fill_u32(s, count_dev, 0, 1);
        if (n) k_perm_moved<<<std::min<u32>(512, div_up(n, 256)), 256, 0, s>>>(n, perm, count_dev);
        SHAMX_HIP(hipGetLastError());
    }

    u32 perm_moved(hipStream_t s, u32 n, const u32 *perm, u32 *count_dev, u32 *count_host) {
        if (n == 0) return 0;
        perm_moved_async(s, n, perm, count_dev);
        SHAMX_HIP(hipMemcpyAsync(count_host, count_dev, sizeof(u32), hipMemcpyDeviceToHost, s));
        SHAMX_HIP(hipStreamSynchronize(s));
        return *count_host;
    }

    void local_sorted_order(hipStream_t s, u32 nm, u32 n, const u32 *sort_index, u32 *perm,
                            u32 *flag, u32 *off, void *scan_scratch) {
        if (nm == 0) return;
        k_local_flags<<<div_up(nm, 256), 256, 0, s>>>(nm, n, sort_index, flag);
        exclusive_scan_u32(s, flag, off, nm, nullptr, scan_scratch);
        k_local_compact<<<div_up(nm, 256), 256, 0, s>>>(nm, sort_index, flag, off, perm);
        SHAMX_HIP(hipGetLastError());
    }
    void perm_gather_f64(hipStream_t s, u32 n, const u32 *perm, const f64 *src, f64 *dst) {
        if (n == 0) return;
        k_perm_gather<f64><<<div_up(n, 256), 256, 0, s>>>(n, perm, src, dst);
        SHAMX_HIP(hipGetLastError());
    }
    void perm_scatter_f64(hipStream_t s, u32 n, const u32 *perm, const f64 *src, f64 *dst) {
        if (n == 0) return;
        k_perm_scatter<f64><<<div_up(n, 256), 256, 0, s>>>(n, perm, src, dst);
        SHAMX_HIP(hipGetLastError());
    }
    void perm_gather_u32(hipStream_t s, u32 n, const u32 *perm, const u32 *src, u32 *dst) {
        if (n == 0) return;
        k_perm_gather<u32><<<div_up(n, 256), 256, 0, s>>>(n, perm, src, dst);
        SHAMX_HIP(hipGetLastError());
    }

    void gather_ghosts(hipStream_t s, u32 n, u32 G, const u32 *ghost_src, FieldPtrs fields) {
        if (G == 0) return;
        k_gather<<<div_up(G, 256), 256, 0, s>>>(n, G, ghost_src, fields);
        SHAMX_HIP(hipGetLastError());
    }

} // namespace shamx
