/* SPDX-License-Identifier: CeCILL-2.1
 *
 * shamx_solver.hpp -- C++20 face of the composite step, with the method names of the
 * reference's shammodels::sph::Solver<Tvec, Kern> (Solver.hpp:70-434) for one patch of f64_3
 * particles on one GPU.  Header only, over the C ABI of shamx_sph.h (link libshamx_sph.so):
 * a C++ caller that drove the reference solver keeps its calls (time edges, evolve_once,
 * evolve_once_time_expl, evolve_until) and gets C++ exceptions where the reference throws.
 *
 * The kernel and the solver configuration are runtime values of shamx_config instead of
 * template arguments; fields are set and read through the layout's field ids (shamx_sph.h).
 */
#ifndef SHAMX_SOLVER_HPP
#define SHAMX_SOLVER_HPP

#include "shamx_sph.h"

#include <chrono>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace shamx::sph {

    /* Solver.hpp:43-54 for one rank: rate = particles / compute time of the step */
    struct TimestepLog {
        int32_t rank;
        double rate;
        uint64_t npart;
        double tcompute;
        double rate_sum() const { return rate; }
        uint64_t npart_sum() const { return npart; }
        double tcompute_max() const { return tcompute; }
    };

    /* Solver.hpp:56-62 */
    struct EvolveUntilResults {
        bool reach_target_time;
        bool reach_niter_max;
        bool reach_max_walltime;
        int32_t iter_count;
    };

    /* a failed call of the C ABI: the status code and shamx_last_error()'s message */
    class Error : public std::runtime_error {
        public:
        Error(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
        int code() const { return code_; }

        private:
        int code_;
    };

    inline void check(int rc) {
        if (rc == SHAMX_OK) return;
        std::string msg = shamx_last_error();
        if (rc == SHAMX_E_INVALID_ARG) throw std::invalid_argument(msg);
        throw Error(rc, msg);
    }

    class Solver {
        public:
        using Tscal = double;

        Solver(const shamx_config &cfg, uint32_t n) : n_(n) {
            check(shamx_solver_create(&cfg, n, &s_));
        }
        ~Solver() {
            if (s_) (void) shamx_solver_destroy(s_);
        }
        Solver(const Solver &)            = delete;
        Solver &operator=(const Solver &) = delete;
        Solver(Solver &&o) noexcept : s_(std::exchange(o.s_, nullptr)), n_(o.n_), last_(o.last_) {}
        Solver &operator=(Solver &&o) noexcept {
            std::swap(s_, o.s_);
            std::swap(n_, o.n_);
            std::swap(last_, o.last_);
            return *this;
        }

        shamx_solver *handle() { return s_; }
        uint32_t part_count() const { return n_; }

        /* time state edges (Solver.hpp:116-121) */
        Tscal get_time() { return edges()[0]; }
        void set_time(Tscal t) { set_edge(0, t); }
        Tscal get_dt_sph() { return edges()[1]; }
        void set_next_dt(Tscal dt) { set_edge(1, dt); }
        Tscal get_cfl_multipler() { return edges()[2]; }
        void set_cfl_multipler(Tscal lambda) { set_edge(2, lambda); }

        /* fields of the layout (shamx_sph.h field ids), n values each, host or device memory */
        void set_field(int field, const double *src, bool on_device = false) {
            check(shamx_solver_set_field(s_, field, src, on_device ? 1 : 0));
        }
        void get_field(int field, double *dst, bool on_device = false) {
            check(shamx_solver_get_field(s_, field, dst, on_device ? 1 : 0));
        }
        std::vector<double> get_field(int field) {
            std::vector<double> v(n_);
            get_field(field, v.data(), false);
            return v;
        }

        /* Model::add_sink (Model.hpp:183-200) */
        void add_sink(double mass, const double pos[3], const double vel[3],
                      double accretion_radius) {
            check(shamx_solver_add_sink(s_, mass, pos, vel, accretion_radius));
        }

        /* Solver::evolve_once (Solver.cpp:1942-3272) */
        TimestepLog evolve_once() {
            check(shamx_solver_evolve_once(s_, &last_));
            return TimestepLog{0, last_.rate, last_.n_local, last_.t_total};
        }
        /* the full per-step log of the last step (phase times, iteration counts, energy) */
        const shamx_step_log &last_step_log() const { return last_; }

        /* Solver.hpp:297-303 */
        Tscal evolve_once_time_expl(Tscal t_current, Tscal dt_input) {
            set_time(t_current);
            set_next_dt(dt_input);
            evolve_once();
            return get_dt_sph();
        }

        /* Solver.hpp:305-433: step until target_time (the last dt clipped to reach it), at most
         * niter_max steps when niter_max >= 0, and no further step once max_walltime seconds
         * (>= 0) have passed since the call (one rank: no collective time) */
        EvolveUntilResults evolve_until(Tscal target_time, int32_t niter_max,
                                        double max_walltime = -1) {
            const bool niter_limit    = niter_max >= 0;
            const bool walltime_limit = max_walltime >= 0;
            const auto start          = std::chrono::steady_clock::now();
            int32_t iter_count        = 0;
            while (get_time() < target_time) {
                Tscal dt = get_dt_sph(), t = get_time();
                if (t > target_time)
                    throw std::invalid_argument("the target time is higher than the current time");
                if (t + dt > target_time) set_next_dt(target_time - t);
                evolve_once();
                iter_count++;
                if (niter_limit && iter_count >= niter_max) return {false, true, false, iter_count};
                if (walltime_limit) {
                    std::chrono::duration<double> el = std::chrono::steady_clock::now() - start;
                    if (el.count() >= max_walltime) return {false, false, true, iter_count};
                }
            }
            return {true, false, false, iter_count};
        }

        private:
        shamx_solver *s_ = nullptr;
        uint32_t n_      = 0;
        shamx_step_log last_{};

        std::vector<double> edges() {
            std::vector<double> t(3);
            check(shamx_solver_get_time(s_, t.data()));
            return t;
        }
        void set_edge(int k, double v) {
            std::vector<double> t = edges();
            t[k]                  = v;
            check(shamx_solver_set_time(s_, t.data()));
        }
    };

} // namespace shamx::sph

#endif
