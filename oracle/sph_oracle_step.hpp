// SPDX-License-Identifier: CeCILL-2.1
//
// sph_oracle_step.hpp -- CPU restatement of one full SPH timestep of the reference
// (shammodels/sph/src/Solver.cpp:1942-3272, Solver<f64_3, Kern>::evolve_once) for one patch.
// Test oracle / CPU baseline only; the HIP product never uses it.
#pragma once

#include "sph_oracle.hpp"
#include <chrono>

namespace orc {

    // C layout shared with the Python test harness (tests/oracle_lib.py)
    struct OrcConfig {
        int32_t kernel;               // 0 M4, 1 M6, 2 M8, 3 C2, 4 C4, 5 C6
        int32_t periodic;             // 1 periodic box, 0 free boundaries
        int32_t eos_kind;             // EosKind
        int32_t tree_reduction_level; // SolverConfig.hpp:584 (default 3)
        int32_t h_iter_per_subcycles; // SolverConfig.hpp:608 (default 50)
        int32_t h_max_subcycles;      // SolverConfig.hpp:609 (default 100)
        double box_min[3], box_max[3];
        double gpart_mass;
        double gamma, eos_cs, eos_cs0, eos_q, eos_r0;
        double alpha_u, alpha_AV, beta_AV;
        double htol_up_coarse, htol_up_fine, epsilon_h; // SolverConfig.hpp:605-607
        double cfl_cour, cfl_force, cfl_multiplier_stiffness;
        // artificial viscosity: 0 constant (alpha_AV, beta_AV, alpha_u), 1 varying Cullen &
        // Dehnen 2010 (AVConfig.hpp VaryingCD10: alpha_min, alpha_max, sigma_decay)
        int32_t av_kind;
        double alpha_min, alpha_max, sigma_decay;
        // gravitational constant of the sink forces (SolverConfig.hpp:469-477) and a Plummer
        // softening of the sink-gas force (0 = the reference's unsoftened force)
        double constant_G, sink_softening;
    };

    // one sink (sink_edges_helper.hpp:234-248)
    struct Sink {
        vec3 pos, vel, acc_sph, acc_ext;
        f64 mass;
        vec3 angmom;
        f64 racc;
    };

    // sink-gas gravity: ExternalForces.cpp:48-70 (axyz_ext reset by the caller) and
    // SinkParticlesUpdate.cpp:69-150 (per sink: the particles gain -G M d/|d|^3, zero inside
    // the accretion radius; the sink's acc_sph becomes the sum of G m d/|d|^3).  soft > 0 is a
    // Plummer softening of the distance cubed (no counterpart in the reference).
    inline void sink_gas(u32 n, const f64 *x, const f64 *y, const f64 *z,
                         std::vector<Sink> &sinks, f64 G, f64 gpart_mass, f64 soft, f64 *axe,
                         f64 *aye, f64 *aze) {
        for (auto &s : sinks) {
            vec3 acc{0, 0, 0};
            for (u32 a = 0; a < n; a++) {
                vec3 delta{x[a] - s.pos.x, y[a] - s.pos.y, z[a] - s.pos.z};
                f64 r2     = dot(delta, delta);
                f64 d      = std::sqrt(r2);
                f64 ds     = (soft == 0) ? d : std::sqrt(r2 + soft * soft);
                vec3 force = G * delta / (ds * ds * ds);
                if (d < s.racc) force = {0, 0, 0};
                acc    = acc + force * gpart_mass;
                axe[a] = axe[a] + (-force.x) * s.mass;
                aye[a] = aye[a] + (-force.y) * s.mass;
                aze[a] = aze[a] + (-force.z) * s.mass;
            }
            s.acc_sph = acc;
        }
    }

    struct StepStats {
        double t_total, t_ghost, t_tree, t_neigh, t_hiter, t_omega, t_eos, t_force, t_corr, t_cfl;
        uint32_t n_merged, n_leaf, subcycles, h_iterations, corrector_iterations;
        uint64_t n_pairs;
        double energy, eps_v, next_dt;
    };

    struct Timer {
        std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
        double sec() const {
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    };

    template<class K>
    struct Solver {
        OrcConfig cfg;
        u32 N = 0;
        std::vector<f64> x, y, z, vx, vy, vz, ax, ay, az, h, u, du;
        // varying AV: per-particle alpha and the sound speed saved at the end of each step
        std::vector<f64> alpha_av, soundspeed;
        // sinks and the external acceleration of the particles (axyz_ext)
        std::vector<Sink> sinks;
        std::vector<f64> axe, aye, aze;
        f64 t = 0, dt = 0, cfl_mult = 1e-2; // Solver.hpp:124-147 time-state edges
        StepStats stats{};

        // transient state of the last step, kept for inspection by the parity tests
        u32 Nm = 0;
        std::vector<u32> ghost_src;
        std::vector<vec3> ghost_off;
        std::vector<f64> mx, my, mz, mh, mvx, mvy, mvz, mu, momega, P, cs, vsig, omega;
        std::vector<f64> divv, curlx, curly, curlz, dtdivv, alpha_new, malpha, max_, may_, maz_;
        vec3 tree_bmin{}, tree_bmax{};
        CLBVH<u32> tree;
        std::vector<f64> rint_tree;
        NeighOut neigh;

        EosParams eos() const {
            EosParams e;
            e.kind  = cfg.eos_kind;
            e.gamma = cfg.gamma;
            e.cs    = cfg.eos_cs;
            e.cs0   = cfg.eos_cs0;
            e.q     = cfg.eos_q;
            e.r0    = cfg.eos_r0;
            return e;
        }

        void resize(u32 n) {
            N = n;
            for (auto *v : {&x, &y, &z, &vx, &vy, &vz, &ax, &ay, &az, &h, &u, &du, &alpha_av,
                            &soundspeed})
                v->assign(n, 0.);
        }

        // ghosts + merge of positions and h: Solver.cpp:984-1046, BasicSPHGhosts.cpp:260-579
        void build_ghosts_and_merge() {
            ghost_src.clear();
            ghost_off.clear();
            if (cfg.periodic) {
                f64 hmax = std::numeric_limits<f64>::lowest();
                for (u32 i = 0; i < N; i++) hmax = std::fmax(hmax, h[i]);
                f64 int_range = hmax * cfg.htol_up_coarse * K::Rkern;
                vec3 bmin{cfg.box_min[0], cfg.box_min[1], cfg.box_min[2]};
                vec3 bmax{cfg.box_max[0], cfg.box_max[1], cfg.box_max[2]};
                ghost_src = periodic_ghost_ids(x.data(), y.data(), z.data(), N, bmin, bmax,
                                               int_range, ghost_off);
            }
            u32 G = (u32) ghost_src.size();
            Nm    = N + G;
            mx.resize(Nm);
            my.resize(Nm);
            mz.resize(Nm);
            mh.resize(Nm);
            for (u32 i = 0; i < N; i++) {
                mx[i] = x[i];
                my[i] = y[i];
                mz[i] = z[i];
                mh[i] = h[i];
            }
            for (u32 g = 0; g < G; g++) {
                u32 s     = ghost_src[g];
                mx[N + g] = x[s] + ghost_off[g].x;
                my[N + g] = y[s] + ghost_off[g].y;
                mz[N + g] = z[s] + ghost_off[g].z;
                mh[N + g] = h[s];
            }
        }

        // BuildTrees.cpp:27-65 + Solver.cpp:1322-1356 (rint field)
        void build_tree() {
            bbox_nextafter(mx.data(), my.data(), mz.data(), Nm, tree_bmin, tree_bmax);
            tree.build_no_aabb(mx.data(), my.data(), mz.data(), Nm, tree_bmin, tree_bmax,
                               (u32) cfg.tree_reduction_level);
            tree.compute_aabb(mx.data(), my.data(), mz.data());
            rint_tree = tree.field_max(mh.data());
            for (auto &v : rint_tree) v *= cfg.htol_up_coarse;
        }

        // ghost field exchange: Solver.cpp:1394-1633 (h, u, v, omega of ghost sources)
        void merge_fields() {
            u32 G = (u32) ghost_src.size();
            for (auto *v : {&mh, &mvx, &mvy, &mvz, &mu, &momega}) v->resize(Nm);
            for (u32 i = 0; i < N; i++) {
                mh[i]     = h[i];
                mvx[i]    = vx[i];
                mvy[i]    = vy[i];
                mvz[i]    = vz[i];
                mu[i]     = u[i];
                momega[i] = omega[i];
            }
            for (u32 g = 0; g < G; g++) {
                u32 s         = ghost_src[g];
                mh[N + g]     = h[s];
                mvx[N + g]    = vx[s];
                mvy[N + g]    = vy[s];
                mvz[N + g]    = vz[s];
                mu[N + g]     = u[s];
                momega[N + g] = omega[s];
            }
        }

        // varying AV (Solver.cpp:2262-2361): div v, curl v and d(div v)/dt on the merged fields
        // (accelerations of the images from their sources), the Cullen & Dehnen update of alpha
        // (not yet written back), then the merged alpha for the force
        void cd10_alpha(f64 dtc) {
            u32 G = (u32) ghost_src.size();
            for (auto *v : {&max_, &may_, &maz_, &malpha}) v->resize(Nm);
            for (u32 i = 0; i < N; i++) {
                max_[i] = ax[i];
                may_[i] = ay[i];
                maz_[i] = az[i];
            }
            for (u32 g = 0; g < G; g++) {
                u32 s        = ghost_src[g];
                max_[N + g] = ax[s];
                may_[N + g] = ay[s];
                maz_[N + g] = az[s];
            }
            for (auto *v : {&divv, &curlx, &curly, &curlz, &dtdivv, &alpha_new}) v->assign(N, 0.);
            compute_divv_curlv<K>(N, neigh.part_cache, mx.data(), my.data(), mz.data(),
                                  mvx.data(), mvy.data(), mvz.data(), mh.data(), momega.data(),
                                  cfg.gpart_mass, divv.data(), curlx.data(), curly.data(),
                                  curlz.data());
            compute_dtdivv<K>(N, neigh.part_cache, mx.data(), my.data(), mz.data(), mvx.data(),
                              mvy.data(), mvz.data(), max_.data(), may_.data(), maz_.data(),
                              mh.data(), cfg.gpart_mass, dtdivv.data());
            update_alpha_cd10(N, dtc, soundspeed.data(), h.data(), alpha_av.data(), divv.data(),
                              curlx.data(), curly.data(), curlz.data(), dtdivv.data(),
                              CD10Params{cfg.alpha_min, cfg.alpha_max, cfg.sigma_decay},
                              alpha_new.data());
            for (u32 i = 0; i < N; i++) malpha[i] = alpha_new[i];
            for (u32 g = 0; g < G; g++) malpha[N + g] = alpha_new[ghost_src[g]];
        }

        // Solver.cpp:1060-1304.  max_sub subcycles; with require_converged false a last
        // subcycle that did not converge is kept (its clamped h) instead of failing.
        void prestep(u32 max_sub, bool require_converged) {
            u32 hstep = 0, hmax = max_sub;
            std::vector<f64> eps(N), h_old(N);
            stats.h_iterations = 0;
            for (; hstep < hmax; hstep++) {
                Timer tg;
                build_ghosts_and_merge();
                stats.t_ghost += tg.sec();
                Timer tt;
                build_tree();
                stats.t_tree += tt.sec();
                Timer tn;
                neigh = neigh_cache_2stages(tree, rint_tree, mx.data(), my.data(), mz.data(),
                                            mh.data(), N, cfg.htol_up_coarse, K::Rkern);
                stats.t_neigh += tn.sec();
                if (cfg.gpart_mass == 0)
                    throw std::runtime_error("invalid gpart_mass 0, this configuration can not converge");
                Timer th;
                std::fill(eps.begin(), eps.end(), 100.);
                h_old = h;
                f64 max_eps = std::numeric_limits<f64>::max();
                for (int it = 0; it < cfg.h_iter_per_subcycles; it++) {
                    iterate_h<K>(N, neigh.part_cache, mx.data(), my.data(), mz.data(),
                                 h_old.data(), h.data(), eps.data(), cfg.gpart_mass,
                                 cfg.htol_up_coarse, cfg.htol_up_fine);
                    stats.h_iterations++;
                    max_eps = std::numeric_limits<f64>::lowest();
                    for (u32 i = 0; i < N; i++) max_eps = std::fmax(max_eps, eps[i]);
                    if (max_eps < cfg.epsilon_h) break;
                }
                f64 min_eps = std::numeric_limits<f64>::max();
                for (u32 i = 0; i < N; i++) min_eps = std::fmin(min_eps, eps[i]);
                stats.t_hiter += th.sec();
                bool converged = (max_eps < cfg.epsilon_h) && !(min_eps < 0);
                if (!converged && (require_converged || hstep + 1 < hmax)) continue;
                break;
            }
            stats.subcycles = std::min(hstep + 1, hmax);
            if (hstep == hmax) throw std::runtime_error("the h iterator is not converged");
            Timer to;
            omega.assign(N, 0.);
            compute_omega<K>(N, neigh.part_cache, mx.data(), my.data(), mz.data(), h.data(),
                             omega.data(), cfg.gpart_mass);
            stats.t_omega += to.sec();
        }

        // sink-sink gravity: SinkParticlesUpdate.cpp:155-200 (compute_ext_forces)
        void sink_ext_forces() {
            const f64 eps = 1e-9;
            for (auto &s : sinks) s.acc_ext = {0, 0, 0};
            for (size_t i = 0; i < sinks.size(); i++) {
                vec3 sum{0, 0, 0};
                for (size_t j = 0; j < sinks.size(); j++) {
                    vec3 rij = sinks[i].pos - sinks[j].pos;
                    f64 r    = std::sqrt(dot(rij, rij));
                    sum      = sum - (cfg.constant_G * sinks[j].mass) * rij / (r * r * r + eps);
                }
                sinks[i].acc_ext = sum;
            }
        }
        void sink_gas_forces() {
            axe.assign(N, 0.);
            aye.assign(N, 0.);
            aze.assign(N, 0.);
            sink_gas(N, x.data(), y.data(), z.data(), sinks, cfg.constant_G, cfg.gpart_mass,
                     cfg.sink_softening, axe.data(), aye.data(), aze.data());
        }

        void evolve_once() {
            stats = StepStats{};
            Timer tstep;
            f64 t_cur = t, dtc = dt;

            // sink predictor: SinkParticlesUpdate.cpp:23-46 (Solver.cpp:2017)
            if (!sinks.empty()) {
                sink_ext_forces();
                for (auto &s : sinks) s.vel = s.vel + (dtc / 2) * (s.acc_sph + s.acc_ext);
                for (auto &s : sinks) s.pos = s.pos + dtc * s.vel;
            }

            // leapfrog predictor: Solver.cpp:390-524, 582-608; ForwardEuler.hpp:62-69
            f64 dt_half = dtc / 2;
            for (u32 i = 0; i < N; i++) {
                vx[i] = vx[i] + dt_half * ax[i];
                vy[i] = vy[i] + dt_half * ay[i];
                vz[i] = vz[i] + dt_half * az[i];
                u[i]  = u[i] + dt_half * du[i];
                x[i]  = x[i] + dtc * vx[i];
                y[i]  = y[i] + dtc * vy[i];
                z[i]  = z[i] + dtc * vz[i];
                vx[i] = vx[i] + dt_half * ax[i];
                vy[i] = vy[i] + dt_half * ay[i];
                vz[i] = vz[i] + dt_half * az[i];
                u[i]  = u[i] + dt_half * du[i];
            }
            // sink-sink forces at the drifted sink positions, then the sink-gas forces
            // (Solver.cpp:2033-2035), before the position boundary
            if (!sinks.empty()) {
                sink_ext_forces();
                sink_gas_forces();
            }
            // apply_position_boundary: Solver.cpp:938-981
            if (cfg.periodic) {
                f64 dx = cfg.box_max[0] - cfg.box_min[0], dy = cfg.box_max[1] - cfg.box_min[1],
                    dz = cfg.box_max[2] - cfg.box_min[2];
                for (u32 i = 0; i < N; i++) {
                    x[i] = position_modulo(x[i], cfg.box_min[0], dx);
                    y[i] = position_modulo(y[i], cfg.box_min[1], dy);
                    z[i] = position_modulo(z[i], cfg.box_min[2], dz);
                }
            }

            prestep((u32) cfg.h_max_subcycles, true);

            // corrector loop: Solver.cpp:2247-3171
            std::vector<f64> ax_old(N), ay_old(N), az_old(N), du_old(N), eps_v2(N);
            f64 next_cfl   = 0;
            u32 corr_iter  = 0;
            bool rerun     = false;
            do {
                if (corr_iter == 50)
                    throw std::runtime_error("the corrector has made over 50 loops");
                merge_fields();
                if (cfg.av_kind == 1) cd10_alpha(dtc);
                Timer te;
                P.assign(Nm, 0.);
                cs.assign(Nm, 0.);
                compute_eos<K>(Nm, mh.data(), mu.data(), mx.data(), my.data(), mz.data(),
                               P.data(), cs.data(), cfg.gpart_mass, eos());
                stats.t_eos += te.sec();
                ax_old = ax;
                ay_old = ay;
                az_old = az;
                du_old = du;
                Timer tf;
                AVParams av{cfg.alpha_u, cfg.alpha_AV, cfg.beta_AV};
                update_derivs_constant_av<K>(N, neigh.part_cache, mx.data(), my.data(), mz.data(),
                                             mvx.data(), mvy.data(), mvz.data(), mh.data(),
                                             momega.data(), mu.data(), P.data(), cs.data(),
                                             ax.data(), ay.data(), az.data(), du.data(),
                                             cfg.gpart_mass, av,
                                             cfg.av_kind == 1 ? malpha.data() : nullptr);
                // add_ext_forces: ExternalForces.cpp:324-360
                if (!sinks.empty())
                    for (u32 i = 0; i < N; i++) {
                        ax[i] = ax[i] + axe[i];
                        ay[i] = ay[i] + aye[i];
                        az[i] = az[i] + aze[i];
                    }
                stats.t_force += tf.sec();
                Timer tc;
                // leapfrog_corrector: integrators.cpp:87-119 with hdt = dt/2
                f64 hdt = dtc / 2;
                f64 max_e2 = std::numeric_limits<f64>::lowest();
                for (u32 i = 0; i < N; i++) {
                    vec3 incr{hdt * (ax[i] - ax_old[i]), hdt * (ay[i] - ay_old[i]),
                              hdt * (az[i] - az_old[i])};
                    vx[i] = vx[i] + incr.x;
                    vy[i] = vy[i] + incr.y;
                    vz[i] = vz[i] + incr.z;
                    max_e2 = std::fmax(max_e2, dot(incr, incr));
                    f64 ui = hdt * (du[i] - du_old[i]);
                    u[i]   = u[i] + ui;
                }
                f64 rank_veps_v = std::sqrt(max_e2);
                f64 sum_vsq     = 0;
                for (u32 i = 0; i < N; i++)
                    sum_vsq += dot({vx[i], vy[i], vz[i]}, {vx[i], vy[i], vz[i]});
                f64 vmean = std::sqrt(sum_vsq / f64(N));
                f64 eps_v = rank_veps_v / vmean;
                if (vmean <= 0) eps_v = 0;
                stats.eps_v = eps_v;
                stats.t_corr += tc.sec();
                if (eps_v > 1e-2) {
                    rerun = true;
                    cfl_mult /= 2;
                } else {
                    rerun = false;
                    // sink corrector: SinkParticlesUpdate.cpp:48-67 (Solver.cpp:2620)
                    for (auto &s : sinks) s.vel = s.vel + (dtc / 2) * (s.acc_sph + s.acc_ext);
                    // write back the updated alpha (Solver.cpp:2622-2650)
                    if (cfg.av_kind == 1) alpha_av = alpha_new;
                    Timer tl;
                    vsig.assign(N, 0.);
                    compute_vsig<K>(N, neigh.part_cache, mx.data(), my.data(), mz.data(),
                                    mvx.data(), mvy.data(), mvz.data(), mh.data(), cs.data(),
                                    vsig.data());
                    // ComputeCFLCourant.hpp:40-60, ComputeCFLForce.hpp:50-62, Solver.cpp:3095-3124
                    f64 C_cour  = cfg.cfl_cour * cfl_mult;
                    f64 C_force = cfg.cfl_force * cfl_mult;
                    f64 rank_dt = std::numeric_limits<f64>::infinity();
                    for (u32 i = 0; i < N; i++) {
                        f64 h_a    = mh[i];
                        f64 dt_c   = C_cour * h_a / vsig[i];
                        f64 abs_a  = std::sqrt(dot({ax[i], ay[i], az[i]}, {ax[i], ay[i], az[i]}));
                        f64 dt_f   = C_force * std::sqrt(h_a / abs_a);
                        f64 cfl_dt = std::fmin(std::numeric_limits<f64>::infinity(), dt_c);
                        cfl_dt     = std::fmin(cfl_dt, dt_f);
                        rank_dt    = std::fmin(rank_dt, cfl_dt);
                    }
                    next_cfl = rank_dt;
                    stats.t_cfl += tl.sec();
                }
                corr_iter++;
            } while (rerun);
            stats.corrector_iterations = corr_iter;
            // sound speed kept for the next step's viscosity update (Solver.cpp:3126-3160)
            if (cfg.av_kind == 1)
                for (u32 i = 0; i < N; i++) soundspeed[i] = cs[i];

            stats.t_total  = tstep.sec();
            stats.n_merged = Nm;
            stats.n_leaf   = tree.L;
            stats.n_pairs  = neigh.part_cache.ids.size();
            f64 E          = 0;
            for (u32 i = 0; i < N; i++)
                E += cfg.gpart_mass * (u[i] + 0.5 * dot({vx[i], vy[i], vz[i]}, {vx[i], vy[i], vz[i]}));
            stats.energy  = E;
            stats.next_dt = next_cfl;

            // Solver.cpp:3235-3263
            dt = next_cfl;
            t  = t_cur + dtc;
            f64 stiff = cfg.cfl_multiplier_stiffness;
            cfl_mult  = (cfl_mult * stiff + 1.) / (stiff + 1.);
        }

        // The phases of evolve_once between the predictor and the corrector, on the current
        // state: sink forces, one h subcycle kept whether or not it converged, Omega, EOS,
        // forces (+ external), vsig and the CFL step.  Constant viscosity only.
        void density_force() {
            if (cfg.av_kind != 0)
                throw std::runtime_error("the density/force pass takes the constant viscosity only");
            stats = StepStats{};
            Timer tstep;
            if (!sinks.empty()) {
                sink_ext_forces();
                sink_gas_forces();
            }
            prestep(1, false);
            merge_fields();
            Timer te;
            P.assign(Nm, 0.);
            cs.assign(Nm, 0.);
            compute_eos<K>(Nm, mh.data(), mu.data(), mx.data(), my.data(), mz.data(), P.data(),
                           cs.data(), cfg.gpart_mass, eos());
            stats.t_eos += te.sec();
            Timer tf;
            AVParams av{cfg.alpha_u, cfg.alpha_AV, cfg.beta_AV};
            update_derivs_constant_av<K>(N, neigh.part_cache, mx.data(), my.data(), mz.data(),
                                         mvx.data(), mvy.data(), mvz.data(), mh.data(),
                                         momega.data(), mu.data(), P.data(), cs.data(), ax.data(),
                                         ay.data(), az.data(), du.data(), cfg.gpart_mass, av,
                                         nullptr);
            if (!sinks.empty())
                for (u32 i = 0; i < N; i++) {
                    ax[i] = ax[i] + axe[i];
                    ay[i] = ay[i] + aye[i];
                    az[i] = az[i] + aze[i];
                }
            stats.t_force += tf.sec();
            Timer tl;
            vsig.assign(N, 0.);
            compute_vsig<K>(N, neigh.part_cache, mx.data(), my.data(), mz.data(), mvx.data(),
                            mvy.data(), mvz.data(), mh.data(), cs.data(), vsig.data());
            f64 C_cour = cfg.cfl_cour * cfl_mult, C_force = cfg.cfl_force * cfl_mult;
            f64 rank_dt = std::numeric_limits<f64>::infinity();
            for (u32 i = 0; i < N; i++) {
                f64 h_a    = mh[i];
                f64 dt_c   = C_cour * h_a / vsig[i];
                f64 abs_a  = std::sqrt(dot({ax[i], ay[i], az[i]}, {ax[i], ay[i], az[i]}));
                f64 dt_f   = C_force * std::sqrt(h_a / abs_a);
                f64 cfl_dt = std::fmin(std::numeric_limits<f64>::infinity(), dt_c);
                cfl_dt     = std::fmin(cfl_dt, dt_f);
                rank_dt    = std::fmin(rank_dt, cfl_dt);
            }
            stats.t_cfl += tl.sec();
            stats.t_total  = tstep.sec();
            stats.n_merged = Nm;
            stats.n_leaf   = tree.L;
            stats.n_pairs  = neigh.part_cache.ids.size();
            stats.next_dt  = rank_dt;
        }
    };

} // namespace orc
