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

namespace shamx {

    namespace {
        constexpr int BLK = 256;

        __device__ inline void atomic_max_key(u64 *p, u64 v) { atomicMax((unsigned long long *) p, v); }
        __device__ inline void atomic_min_key(u64 *p, u64 v) { atomicMin((unsigned long long *) p, v); }
        __device__ inline u64 umax64(u64 a, u64 b) { return a > b ? a : b; }
        __device__ inline u64 umin64(u64 a, u64 b) { return a < b ? a : b; }

        // ------------------------------------------------------------- h iteration
        template<class K>
        __global__ void __launch_bounds__(BLK)
            k_h_iterate(u32 n, const u32 *__restrict cnt, const u32 *__restrict offset,
                        const u32 *__restrict ids, const f64 *__restrict x,
                        const f64 *__restrict y, const f64 *__restrict z,
                        const f64 *__restrict h_old, f64 *__restrict h_new, f64 *__restrict eps,
                        f64 mass, f64 h_evol_max, f64 h_evol_iter_max, u64 *__restrict minmax) {
            using KK = Kern<K, f64>;
            u32 a    = blockIdx.x * BLK + threadIdx.x;
            u64 kmax = ORDERED_NEG_INF, kmin = ORDERED_POS_INF;
            if (a < n) {
                f64 e = eps[a];
                if (e > 1e-6) {
                    const f64 Rkern = K::Rkern;
                    f64 h_max_evol_p = h_evol_iter_max, h_max_evol_m = 1 / h_evol_iter_max;
                    f64 xa = x[a], ya = y[a], za = z[a];
                    f64 h_a  = h_new[a];
                    f64 dint = h_a * h_a * Rkern * Rkern;
                    f64 rho_sum = 0, sumdWdh = 0;
                    u32 k0 = offset[a], k1 = k0 + cnt[a];
                    for (u32 k = k0; k < k1; k++) {
                        u32 b    = ids[k];
                        f64 dx   = xa - x[b], dy = ya - y[b], dz = za - z[b];
                        f64 rab2 = dx * dx + dy * dy + dz * dz;
                        if (rab2 > dint) continue;
                        f64 rab = sqrt(rab2);
                        rho_sum += mass * KK::W(rab, h_a);
                        sumdWdh += mass * KK::dhW(rab, h_a);
                    }
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
                kmax = kmin = f64_to_ordered(e);
            }
            kmax = wave_reduce(kmax, umax64);
            kmin = wave_reduce(kmin, umin64);
            if (__lane_id() == 0) {
                atomic_max_key(&minmax[0], kmax);
                atomic_min_key(&minmax[1], kmin);
            }
        }

        // mixed: density sums in FP32 (Kahan) on FP64 r_ab; Newton step and test in FP64
        __device__ inline void kahan_add(f32 &s, f32 &c, f32 v) {
            f32 y = v - c;
            f32 t = s + y;
            c     = (t - s) - y;
            s     = t;
        }

        template<class K>
        __global__ void __launch_bounds__(BLK)
            k_h_iterate_mixed(u32 n, const u32 *__restrict cnt, const u32 *__restrict offset,
                              const u32 *__restrict ids, const f64 *__restrict x,
                              const f64 *__restrict y, const f64 *__restrict z,
                              const f64 *__restrict h_old, f64 *__restrict h_new,
                              f64 *__restrict eps, f64 mass, f64 h_evol_max,
                              f64 h_evol_iter_max, u64 *__restrict minmax) {
            using KF = Kern<K, f32>;
            u32 a    = blockIdx.x * BLK + threadIdx.x;
            u64 kmax = ORDERED_NEG_INF, kmin = ORDERED_POS_INF;
            if (a < n) {
                f64 e = eps[a];
                if (e > 1e-6) {
                    const f64 Rkern = K::Rkern;
                    f64 h_max_evol_p = h_evol_iter_max, h_max_evol_m = 1 / h_evol_iter_max;
                    f64 xa = x[a], ya = y[a], za = z[a];
                    f64 h_a  = h_new[a];
                    f64 dint = h_a * h_a * Rkern * Rkern;
                    f32 hf = (f32) h_a;
                    f32 rs = 0, rc = 0, ds = 0, dc = 0;
                    u32 k0 = offset[a], k1 = k0 + cnt[a];
                    for (u32 k = k0; k < k1; k++) {
                        u32 b    = ids[k];
                        f64 dx   = xa - x[b], dy = ya - y[b], dz = za - z[b];
                        f64 rab2 = dx * dx + dy * dy + dz * dz;
                        if (rab2 > dint) continue;
                        f32 rab = sqrtf((f32) rab2);
                        kahan_add(rs, rc, KF::W(rab, hf));
                        kahan_add(ds, dc, KF::dhW(rab, hf));
                    }
                    f64 rho_sum = mass * ((f64) rs - (f64) rc);
                    f64 sumdWdh = mass * ((f64) ds - (f64) dc);
                    f64 rho_ha  = rho_h(mass, h_a, (f64) K::hfactd);
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
                kmax = kmin = f64_to_ordered(e);
            }
            kmax = wave_reduce(kmax, umax64);
            kmin = wave_reduce(kmin, umin64);
            if (__lane_id() == 0) {
                atomic_max_key(&minmax[0], kmax);
                atomic_min_key(&minmax[1], kmin);
            }
        }

        // ------------------------------------------------------------- omega
        template<class K>
        __global__ void __launch_bounds__(BLK)
            k_omega(u32 n, const u32 *__restrict cnt, const u32 *__restrict offset,
                    const u32 *__restrict ids, const f64 *__restrict x, const f64 *__restrict y,
                    const f64 *__restrict z, const f64 *__restrict h, f64 *__restrict omega,
                    f64 mass) {
            using KK = Kern<K, f64>;
            u32 a    = blockIdx.x * BLK + threadIdx.x;
            if (a >= n) return;
            const f64 Rkern = K::Rkern;
            f64 xa = x[a], ya = y[a], za = z[a];
            f64 h_a  = h[a];
            f64 dint = h_a * h_a * Rkern * Rkern;
            f64 rho_sum = 0, part_omega_sum = 0;
            u32 k0 = offset[a], k1 = k0 + cnt[a];
            for (u32 k = k0; k < k1; k++) {
                u32 b    = ids[k];
                f64 dx   = xa - x[b], dy = ya - y[b], dz = za - z[b];
                f64 rab2 = dx * dx + dy * dy + dz * dz;
                if (rab2 > dint) continue;
                f64 rab = sqrt(rab2);
                rho_sum += mass * KK::W(rab, h_a);
                part_omega_sum += mass * KK::dhW(rab, h_a);
            }
            f64 rho_ha = rho_h(mass, h_a, KK::hfactd);
            omega[a]   = 1 + (h_a / (3 * rho_ha)) * part_omega_sum;
        }

        // ------------------------------------------------------------- eos
        template<class K>
        __global__ void k_eos(u32 nm, const f64 *__restrict h, const f64 *__restrict u,
                              const f64 *__restrict x, const f64 *__restrict y,
                              const f64 *__restrict z, f64 *__restrict P, f64 *__restrict cs,
                              f64 mass, EosArgs e) {
            u32 i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= nm) return;
            f64 rho = rho_h(mass, h[i], (f64) K::hfactd);
            if (e.kind == 0) {
                f64 P_a = (e.gamma - 1) * rho * u[i];
                P[i]    = P_a;
                cs[i]   = sqrt(e.gamma * P_a / rho);
            } else if (e.kind == 1) {
                P[i]  = e.cs * e.cs * rho;
                cs[i] = e.cs;
            } else {
                f64 r0sq  = e.r0 * e.r0;
                f64 mq    = -e.q;
                f64 Rsq   = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
                f64 cs_sq = (e.cs0 * e.cs0) * pow(Rsq / r0sq, mq);
                cs[i]     = sqrt(cs_sq);
                P[i]      = cs_sq * rho;
            }
        }

        // ------------------------------------------------------------- force (FP64)
        __device__ inline f64 inv_sat_positive(f64 v) { return (v >= 1e-9) ? 1. / v : 0.; }
        __device__ inline f64 inv_sat_zero(f64 v) { return (v != 0. && v == v) ? 1. / v : 0.; }

        template<class K>
        __global__ void __launch_bounds__(BLK)
            k_force(u32 n, const u32 *__restrict cnt, const u32 *__restrict offset,
                    const u32 *__restrict ids, const f64 *__restrict x, const f64 *__restrict y,
                    const f64 *__restrict z, const f64 *__restrict vx, const f64 *__restrict vy,
                    const f64 *__restrict vz, const f64 *__restrict h,
                    const f64 *__restrict omega, const f64 *__restrict u,
                    const f64 *__restrict P, const f64 *__restrict cs, f64 *__restrict ax,
                    f64 *__restrict ay, f64 *__restrict az, f64 *__restrict du,
                    f64 *__restrict ax_old, f64 *__restrict ay_old, f64 *__restrict az_old,
                    f64 *__restrict du_old, f64 *__restrict vsig_out, f64 *__restrict dt_out,
                    ForceArgs fa, u64 *__restrict dt_min, u64 *__restrict pair_count) {
            using KK = Kern<K, f64>;
            u32 a    = blockIdx.x * BLK + threadIdx.x;
            u64 key  = ORDERED_POS_INF;
            u64 npair = 0;
            if (a < n) {
                const f64 pmass = fa.mass;
                const f64 Rker2 = K::Rkern * K::Rkern;
                f64 h_a = h[a];
                f64 xa = x[a], ya = y[a], za = z[a];
                f64 vxa = vx[a], vya = vy[a], vza = vz[a];
                f64 P_a = P[a], omega_a = omega[a];
                const f64 u_a = u[a];
                f64 rho_a     = rho_h(pmass, h_a, KK::hfactd);
                f64 rho_a_sq  = rho_a * rho_a;
                f64 rho_a_inv = 1. / rho_a;
                f64 omega_a_rho_a_inv = 1 / (omega_a * rho_a);
                f64 cs_a = cs[a];
                f64 fx = 0, fy = 0, fz = 0, tdu = 0, vsig_max = 0;
                f64 ca = h_a * h_a * Rker2;
                u32 k0 = offset[a], k1 = k0 + cnt[a];
                for (u32 k = k0; k < k1; k++) {
                    u32 b    = ids[k];
                    f64 dx   = xa - x[b], dy = ya - y[b], dz = za - z[b];
                    f64 rab2 = dx * dx + dy * dy + dz * dz;
                    f64 h_b  = h[b];
                    if (rab2 > ca && rab2 > h_b * h_b * Rker2) continue;
                    npair++;
                    f64 rab = sqrt(rab2);
                    f64 vxb = vx[b], vyb = vy[b], vzb = vz[b];
                    const f64 u_b = u[b];
                    f64 rho_b = rho_h(pmass, h_b, KK::hfactd);
                    f64 P_b = P[b], omega_b = omega[b], cs_b = cs[b];
                    f64 Fab_a = KK::dW(rab, h_a), Fab_b = KK::dW(rab, h_b);
                    f64 vabx = vxa - vxb, vaby = vya - vyb, vabz = vza - vzb;
                    f64 inv  = inv_sat_positive(rab);
                    f64 rux = dx * inv, ruy = dy * inv, ruz = dz * inv;
                    f64 v_ab_r_ab = vabx * rux + vaby * ruy + vabz * ruz;
                    f64 abs_v     = fabs(v_ab_r_ab);
                    f64 vsig_a = fa.alpha_AV * cs_a + fa.beta_AV * abs_v;
                    f64 vsig_b = fa.alpha_AV * cs_b + fa.beta_AV * abs_v;
                    f64 rho_avg = (rho_a + rho_b) * 0.5;
                    f64 vsig_u  = sqrt(fabs(P_a - P_b) / rho_avg);
                    f64 qa_ab = fmax(-0.5 * rho_a * vsig_a * v_ab_r_ab, 0.);
                    f64 qb_ab = fmax(-0.5 * rho_b * vsig_b * v_ab_r_ab, 0.);
                    f64 AV_P_a = P_a + qa_ab, AV_P_b = P_b + qb_ab;
                    f64 nax = rux * Fab_a, nay = ruy * Fab_a, naz = ruz * Fab_a;
                    f64 nbx = rux * Fab_b, nby = ruy * Fab_b, nbz = ruz * Fab_b;
                    f64 ia = AV_P_a * inv_sat_zero(rho_a_sq * omega_a);
                    f64 ib = AV_P_b * inv_sat_zero((rho_b * rho_b) * omega_b);
                    f64 mm = -pmass;
                    fx     = fx + mm * (ia * nax + ib * nbx);
                    fy     = fy + mm * (ia * nay + ib * nby);
                    fz     = fz + mm * (ia * naz + ib * nbz);
                    tdu += AV_P_a * (omega_a_rho_a_inv * rho_a_inv) * pmass *
                           (vabx * nax + vaby * nay + vabz * naz);
                    tdu += pmass * fa.alpha_u * vsig_u * (u_a - u_b) * 0.5 *
                           (Fab_a * omega_a_rho_a_inv + Fab_b / (rho_b * omega_b));
                    // CFL signal speed (Solver.cpp:2749-2783): r_ab / |r_ab|, alpha 1, beta 2
                    f64 r2x = dx / rab, r2y = dy / rab, r2z = dz / rab;
                    if (rab < 1e-9) r2x = r2y = r2z = 0;
                    f64 vdot = vabx * r2x + vaby * r2y + vabz * r2z;
                    vsig_max = fmax(vsig_max, 1.0 * cs_a + 2.0 * fabs(vdot));
                }
                ax_old[a] = ax[a];
                ay_old[a] = ay[a];
                az_old[a] = az[a];
                du_old[a] = du[a];
                ax[a]     = fx;
                ay[a]     = fy;
                az[a]     = fz;
                du[a]     = tdu;
                vsig_out[a] = vsig_max;
                f64 dt_c   = fa.C_cour * h_a / vsig_max;
                f64 abs_a  = sqrt(fx * fx + fy * fy + fz * fz);
                f64 dt_f   = fa.C_force * sqrt(h_a / abs_a);
                f64 cfl_dt = fmin(__builtin_inf(), dt_c);
                cfl_dt     = fmin(cfl_dt, dt_f);
                dt_out[a]  = cfl_dt;
                key        = f64_to_ordered(cfl_dt);
            }
            key   = wave_reduce(key, umin64);
            npair = wave_reduce(npair, [](u64 p, u64 q) { return p + q; });
            if (__lane_id() == 0) {
                atomic_min_key(dt_min, key);
                if (pair_count) atomicAdd((unsigned long long *) pair_count, npair);
            }
        }

        // ------------------------------------------------------------- force (mixed)
        template<class K>
        __global__ void __launch_bounds__(BLK)
            k_force_mixed(u32 n, const u32 *__restrict cnt, const u32 *__restrict offset,
                          const u32 *__restrict ids, const f64 *__restrict x,
                          const f64 *__restrict y, const f64 *__restrict z,
                          const f64 *__restrict vx, const f64 *__restrict vy,
                          const f64 *__restrict vz, const f64 *__restrict h,
                          const f64 *__restrict omega, const f64 *__restrict u,
                          const f64 *__restrict P, const f64 *__restrict cs, f64 *__restrict ax,
                          f64 *__restrict ay, f64 *__restrict az, f64 *__restrict du,
                          f64 *__restrict ax_old, f64 *__restrict ay_old,
                          f64 *__restrict az_old, f64 *__restrict du_old,
                          f64 *__restrict vsig_out, f64 *__restrict dt_out, ForceArgs fa,
                          u64 *__restrict dt_min, u64 *__restrict pair_count) {
            using KF = Kern<K, f32>;
            u32 a    = blockIdx.x * BLK + threadIdx.x;
            u64 key  = ORDERED_POS_INF;
            u64 npair = 0;
            if (a < n) {
                const f32 pmass = (f32) fa.mass;
                const f64 Rker2 = K::Rkern * K::Rkern;
                const f32 hfact = (f32) K::hfactd;
                f64 h_a64 = h[a];
                f32 h_a   = (f32) h_a64;
                f64 xa = x[a], ya = y[a], za = z[a];
                f32 vxa = (f32) vx[a], vya = (f32) vy[a], vza = (f32) vz[a];
                f32 P_a = (f32) P[a], omega_a = (f32) omega[a];
                f32 u_a = (f32) u[a];
                f32 rho_a = rho_h(pmass, h_a, hfact);
                f32 rho_a_sq = rho_a * rho_a;
                f32 rho_a_inv = 1.f / rho_a;
                f32 omega_a_rho_a_inv = 1.f / (omega_a * rho_a);
                f32 inv_sub_a = 1.f / (rho_a_sq * omega_a);
                f32 cs_a = (f32) cs[a];
                f32 alpha_AV = (f32) fa.alpha_AV, beta_AV = (f32) fa.beta_AV,
                    alpha_u = (f32) fa.alpha_u;
                f32 sx = 0, cx = 0, sy = 0, cy = 0, sz = 0, cz = 0, su = 0, cu = 0;
                f32 vsig_max = 0;
                f64 ca = h_a64 * h_a64 * Rker2;
                u32 k0 = offset[a], k1 = k0 + cnt[a];
                for (u32 k = k0; k < k1; k++) {
                    u32 b    = ids[k];
                    f64 dx   = xa - x[b], dy = ya - y[b], dz = za - z[b];
                    f64 rab2 = dx * dx + dy * dy + dz * dz;
                    f64 h_b64 = h[b];
                    if (rab2 > ca && rab2 > h_b64 * h_b64 * Rker2) continue;
                    npair++;
                    f32 fdx = (f32) dx, fdy = (f32) dy, fdz = (f32) dz;
                    f32 rab = sqrtf((f32) rab2);
                    f32 h_b = (f32) h_b64;
                    f32 vabx = vxa - (f32) vx[b], vaby = vya - (f32) vy[b],
                        vabz = vza - (f32) vz[b];
                    f32 u_b = (f32) u[b];
                    f32 rho_b = rho_h(pmass, h_b, hfact);
                    f32 P_b = (f32) P[b], omega_b = (f32) omega[b], cs_b = (f32) cs[b];
                    f32 Fab_a = KF::dW(rab, h_a), Fab_b = KF::dW(rab, h_b);
                    f32 inv = (rab >= 1e-9f) ? 1.f / rab : 0.f;
                    f32 rux = fdx * inv, ruy = fdy * inv, ruz = fdz * inv;
                    f32 v_ab_r_ab = vabx * rux + vaby * ruy + vabz * ruz;
                    f32 abs_v     = fabsf(v_ab_r_ab);
                    f32 vsig_a = alpha_AV * cs_a + beta_AV * abs_v;
                    f32 vsig_b = alpha_AV * cs_b + beta_AV * abs_v;
                    f32 rho_avg = (rho_a + rho_b) * 0.5f;
                    f32 vsig_u  = sqrtf(fabsf(P_a - P_b) / rho_avg);
                    f32 qa_ab = fmaxf(-0.5f * rho_a * vsig_a * v_ab_r_ab, 0.f);
                    f32 qb_ab = fmaxf(-0.5f * rho_b * vsig_b * v_ab_r_ab, 0.f);
                    f32 AV_P_a = P_a + qa_ab, AV_P_b = P_b + qb_ab;
                    f32 ia = AV_P_a * inv_sub_a * Fab_a;
                    f32 ib = AV_P_b * (1.f / ((rho_b * rho_b) * omega_b)) * Fab_b;
                    f32 s  = -pmass * (ia + ib);
                    kahan_add(sx, cx, s * rux);
                    kahan_add(sy, cy, s * ruy);
                    kahan_add(sz, cz, s * ruz);
                    f32 t1 = AV_P_a * (omega_a_rho_a_inv * rho_a_inv) * pmass * Fab_a *
                             (vabx * rux + vaby * ruy + vabz * ruz);
                    f32 t2 = pmass * alpha_u * vsig_u * (u_a - u_b) * 0.5f *
                             (Fab_a * omega_a_rho_a_inv + Fab_b / (rho_b * omega_b));
                    kahan_add(su, cu, t1 + t2);
                    vsig_max = fmaxf(vsig_max, cs_a + 2.f * abs_v);
                }
                f64 fx = (f64) sx - (f64) cx, fy = (f64) sy - (f64) cy, fz = (f64) sz - (f64) cz;
                f64 tdu = (f64) su - (f64) cu;
                ax_old[a] = ax[a];
                ay_old[a] = ay[a];
                az_old[a] = az[a];
                du_old[a] = du[a];
                ax[a]     = fx;
                ay[a]     = fy;
                az[a]     = fz;
                du[a]     = tdu;
                f64 vs    = (f64) vsig_max;
                vsig_out[a] = vs;
                f64 dt_c   = fa.C_cour * h_a64 / vs;
                f64 abs_a  = sqrt(fx * fx + fy * fy + fz * fz);
                f64 dt_f   = fa.C_force * sqrt(h_a64 / abs_a);
                f64 cfl_dt = fmin(__builtin_inf(), dt_c);
                cfl_dt     = fmin(cfl_dt, dt_f);
                dt_out[a]  = cfl_dt;
                key        = f64_to_ordered(cfl_dt);
            }
            key   = wave_reduce(key, umin64);
            npair = wave_reduce(npair, [](u64 p, u64 q) { return p + q; });
            if (__lane_id() == 0) {
                atomic_min_key(dt_min, key);
                if (pair_count) atomicAdd((unsigned long long *) pair_count, npair);
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

        __global__ void __launch_bounds__(BLK)
            k_corrector(u32 n, f64 hdt, f64 *__restrict vx, f64 *__restrict vy,
                        f64 *__restrict vz, const f64 *__restrict ax, const f64 *__restrict ay,
                        const f64 *__restrict az, const f64 *__restrict ax_old,
                        const f64 *__restrict ay_old, const f64 *__restrict az_old,
                        f64 *__restrict u, const f64 *__restrict du,
                        const f64 *__restrict du_old, f64 mass, u64 *__restrict red) {
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
            kmax = wave_reduce(kmax, umax64);
            vsq  = wave_reduce(vsq, [](f64 p, f64 q) { return p + q; });
            en   = wave_reduce(en, [](f64 p, f64 q) { return p + q; });
            if (__lane_id() == 0) {
                atomic_max_key(&red[0], kmax);
                atomicAdd((f64 *) &red[1], vsq);
                atomicAdd((f64 *) &red[2], en);
            }
        }

        __global__ void __launch_bounds__(BLK) k_max_key(u32 n, const f64 *__restrict v,
                                                         u64 *__restrict key) {
            u64 k = ORDERED_NEG_INF;
            for (u32 i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK)
                k = umax64(k, f64_to_ordered(v[i]));
            k = wave_reduce(k, umax64);
            if (__lane_id() == 0) atomic_max_key(key, k);
        }

        // ------------------------------------------------------------- periodic images
        struct CutBoxes {
            f64 lo[26][3], hi[26][3], off[26][3];
            u32 active[26];
        };

        __device__ inline bool in_cut(const CutBoxes &cb, int o, f64 px, f64 py, f64 pz) {
            return cb.active[o] && (cb.lo[o][0] <= px) && (px < cb.hi[o][0]) &&
                   (cb.lo[o][1] <= py) && (py < cb.hi[o][1]) && (cb.lo[o][2] <= pz) &&
                   (pz < cb.hi[o][2]);
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

    void h_iterate(hipStream_t s, int kernel, u32 n, Csr c, const f64 *x, const f64 *y,
                   const f64 *z, const f64 *h_old, f64 *h_new, f64 *eps, f64 mass,
                   f64 h_evol_max, f64 h_evol_iter_max, u64 *eps_minmax, int mixed) {
        if (n == 0) return;
        dispatch_kernel(kernel, [&](auto K) {
            using KT = decltype(K);
            if (mixed)
                k_h_iterate_mixed<KT><<<div_up(n, BLK), BLK, 0, s>>>(
                    n, c.cnt, c.offset, c.ids, x, y, z, h_old, h_new, eps, mass, h_evol_max,
                    h_evol_iter_max, eps_minmax);
            else
                k_h_iterate<KT><<<div_up(n, BLK), BLK, 0, s>>>(n, c.cnt, c.offset, c.ids, x, y,
                                                               z, h_old, h_new, eps, mass,
                                                               h_evol_max, h_evol_iter_max,
                                                               eps_minmax);
        });
        SHAMX_HIP(hipGetLastError());
    }

    void compute_omega(hipStream_t s, int kernel, u32 n, Csr c, const f64 *x, const f64 *y,
                       const f64 *z, const f64 *h, f64 *omega, f64 mass) {
        if (n == 0) return;
        dispatch_kernel(kernel, [&](auto K) {
            k_omega<decltype(K)><<<div_up(n, BLK), BLK, 0, s>>>(n, c.cnt, c.offset, c.ids, x, y,
                                                                z, h, omega, mass);
        });
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

    void compute_force(hipStream_t s, int kernel, u32 n, Csr c, const f64 *x, const f64 *y,
                       const f64 *z, const f64 *vx, const f64 *vy, const f64 *vz, const f64 *h,
                       const f64 *omega, const f64 *u, const f64 *P, const f64 *cs, f64 *ax,
                       f64 *ay, f64 *az, f64 *du, f64 *ax_old, f64 *ay_old, f64 *az_old,
                       f64 *du_old, f64 *vsig, f64 *dt_cfl, ForceArgs fa, u64 *dt_min_key,
                       u64 *pair_count) {
        if (n == 0) return;
        dispatch_kernel(kernel, [&](auto K) {
            using KT = decltype(K);
            if (fa.mixed)
                k_force_mixed<KT><<<div_up(n, BLK), BLK, 0, s>>>(
                    n, c.cnt, c.offset, c.ids, x, y, z, vx, vy, vz, h, omega, u, P, cs, ax, ay,
                    az, du, ax_old, ay_old, az_old, du_old, vsig, dt_cfl, fa, dt_min_key, pair_count);
            else
                k_force<KT><<<div_up(n, BLK), BLK, 0, s>>>(
                    n, c.cnt, c.offset, c.ids, x, y, z, vx, vy, vz, h, omega, u, P, cs, ax, ay,
                    az, du, ax_old, ay_old, az_old, du_old, vsig, dt_cfl, fa, dt_min_key, pair_count);
        });
        SHAMX_HIP(hipGetLastError());
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
                   u64 *red) {
        if (n == 0) return;
        k_corrector<<<div_up(n, BLK), BLK, 0, s>>>(n, hdt, vx, vy, vz, ax, ay, az, ax_old, ay_old,
                                                   az_old, u, du, du_old, mass, red);
        SHAMX_HIP(hipGetLastError());
    }

    void max_key_f64(hipStream_t s, u32 n, const f64 *v, u64 *key) {
        if (n == 0) return;
        u32 nb = std::min<u32>(div_up(n, BLK), 1024);
        k_max_key<<<nb, BLK, 0, s>>>(n, v, key);
        SHAMX_HIP(hipGetLastError());
    }

    // cut volumes are computed on the host exactly as find_interfaces does for one patch:
    // offset o = (xo, yo, zo) in lexicographic order without (0,0,0); interf volume =
    // sender box ∩ (receiver box expanded by int_range - offset); membership half-open.
    u32 periodic_ghosts(hipStream_t s, u32 n, f64 *x, f64 *y, f64 *z, f64 *h, const f64 *box6,
                        f64 int_range, DevArray<u32> &ghost_src, DevArray<u8> &ghost_off_id,
                        DevArray<u32> &work, DevArray<unsigned char> &scratch, u32 cap) {
        CutBoxes cb{};
        f64 bmin[3] = {box6[0], box6[1], box6[2]}, bmax[3] = {box6[3], box6[4], box6[5]};
        f64 bsize[3], pmin[3], pmax[3];
        for (int k = 0; k < 3; k++) {
            bsize[k] = bmax[k] - bmin[k];
            pmin[k]  = bmin[k];
            pmax[k]  = bsize[k] + bmin[k];
        }
        int o = 0;
        for (int xo = -1; xo <= 1; xo++)
            for (int yo = -1; yo <= 1; yo++)
                for (int zo = -1; zo <= 1; zo++) {
                    if (xo == 0 && yo == 0 && zo == 0) continue;
                    int io[3] = {xo, yo, zo};
                    bool nonempty = true;
                    for (int k = 0; k < 3; k++) {
                        f64 off = io[k] * bsize[k];
                        f64 rlo = pmin[k] - int_range, rhi = pmax[k] + int_range;
                        f64 ilo = std::fmax(rlo, pmin[k] + off), ihi = std::fmin(rhi, pmax[k] + off);
                        if (!(ihi >= ilo)) nonempty = false;
                        cb.off[o][k] = off;
                        cb.lo[o][k]  = std::fmax(pmin[k], rlo - off);
                        cb.hi[o][k]  = std::fmin(pmax[k], rhi - off);
                    }
                    cb.active[o] = nonempty ? 1u : 0u;
                    o++;
                }
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
        SHAMX_HIP(hipMemcpyAsync(cbd, &cb, sizeof(CutBoxes), hipMemcpyHostToDevice, s));
        k_ghost_count<<<nb, BLK, 0, s>>>(n, x, y, z, cbd, counts, nb);
        exclusive_scan_u32(s, counts, bases, 26 * nb, total, scratch.data());
        u32 G = 0;
        SHAMX_HIP(hipMemcpyAsync(&G, total, sizeof(u32), hipMemcpyDeviceToHost, s));
        SHAMX_HIP(hipStreamSynchronize(s));
        if (G > cap) return G; // caller grows the merged arrays and calls again
        ghost_src.resize(std::max<u32>(G, 1));
        ghost_off_id.resize(std::max<u32>(G, 1));
        k_ghost_scatter<<<nb, BLK, 0, s>>>(n, x, y, z, h, cbd, bases, nb, ghost_src.data(),
                                           ghost_off_id.data());
        SHAMX_HIP(hipGetLastError());
        return G;
    }

    void gather_ghosts(hipStream_t s, u32 n, u32 G, const u32 *ghost_src, FieldPtrs fields) {
        if (G == 0) return;
        k_gather<<<div_up(G, 256), 256, 0, s>>>(n, G, ghost_src, fields);
        SHAMX_HIP(hipGetLastError());
    }

} // namespace shamx
