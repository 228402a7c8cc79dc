// Test harness for include/shamx_solver.hpp (the C++ face of the composite step): a C entry the
// Python tests load with ctypes, which builds a shamx::sph::Solver from a config and 12 host
// fields, runs evolve_until, and returns the fields, the time edges and the loop's results.
#include "shamx_solver.hpp"

#include <cstring>
#include <exception>

extern "C" int facade_evolve_until(const shamx_config *cfg, uint32_t n,
                                   const double *const *fields12, double target_time,
                                   int32_t niter_max, double *const *out12, double *out_tdc3,
                                   int32_t *out_results4, char *err, int errlen) {
    try {
        shamx::sph::Solver s(*cfg, n);
        for (int k = 0; k < 12; k++) s.set_field(k, fields12[k]);
        shamx::sph::EvolveUntilResults r = s.evolve_until(target_time, niter_max);
        for (int k = 0; k < 12; k++) s.get_field(k, out12[k]);
        out_tdc3[0]     = s.get_time();
        out_tdc3[1]     = s.get_dt_sph();
        out_tdc3[2]     = s.get_cfl_multipler();
        out_results4[0] = r.reach_target_time;
        out_results4[1] = r.reach_niter_max;
        out_results4[2] = r.reach_max_walltime;
        out_results4[3] = r.iter_count;
        return 0;
    } catch (const std::invalid_argument &e) {
        std::strncpy(err, e.what(), errlen - 1);
        return 1;
    } catch (const std::exception &e) {
        std::strncpy(err, e.what(), errlen - 1);
        return 2;
    }
}
