// SPDX-License-Identifier: CeCILL-2.1
//
// C entry points of the CPU oracle (test infrastructure only; loaded by tests/ and by the
// cpu_baseline leg of bench.py through ctypes).  Errors are reported as a non-zero return code
// and a thread-local message (orc_last_error).
#include "sph_oracle_step.hpp"

#include <memory>
#include <string>

using namespace orc;

namespace {
    thread_local std::string g_err;

    template<class F>
    int guard(F &&f) {
        try {
            f();
            return 0;
        } catch (std::exception &e) {
            g_err = e.what();
            return 1;
        }
    }

    // kernel dispatch on the reference's template argument (Solver.cpp:3276-3282)
    template<class F>
    void with_kernel(int k, F &&f) {
        switch (k) {
        case 0: f(KM4{}); break;
        case 1: f(KM6{}); break;
        case 2: f(KM8{}); break;
        case 3: f(KC2{}); break;
        case 4: f(KC4{}); break;
        case 5: f(KC6{}); break;
        default: throw std::invalid_argument("unknown SPH kernel id");
        }
    }

    struct TreeHandle {
        int bits;
        CLBVH<u32> t32;
        CLBVH<u64> t64;
    };

    struct SolverBase {
        virtual ~SolverBase() = default;
    };
    template<class K>
    struct SolverT : SolverBase {
        Solver<K> s;
    };
} // namespace

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }

// ---- kernels -----------------------------------------------------------------------------
int orc_kernel_eval(int kernel, uint32_t n, const double *r, const double *h, double *W,
                    double *dW, double *dhW, double *f, double *df, double *out_consts) {
    return guard([&] {
        with_kernel(kernel, [&](auto K) {
            using KT = decltype(K);
            using KK = Kern<KT>;
            for (uint32_t i = 0; i < n; i++) {
                W[i]   = KK::W_3d(r[i], h[i]);
                dW[i]  = KK::dW_3d(r[i], h[i]);
                dhW[i] = KK::dhW_3d(r[i], h[i]);
                f[i]   = KT::f(r[i]);
                df[i]  = KT::df(r[i]);
            }
            out_consts[0] = KT::Rkern;
            out_consts[1] = KT::hfactd;
            out_consts[2] = KT::norm_3d;
        });
    });
}

// ---- morton / bbox / sort ----------------------------------------------------------------
int orc_bbox(uint32_t n, const double *x, const double *y, const double *z, double *out6) {
    return guard([&] {
        vec3 lo, hi;
        bbox_nextafter(x, y, z, n, lo, hi);
        out6[0] = lo.x;
        out6[1] = lo.y;
        out6[2] = lo.z;
        out6[3] = hi.x;
        out6[4] = hi.y;
        out6[5] = hi.z;
    });
}

int orc_morton_u32(uint32_t n, const double *x, const double *y, const double *z,
                   const double *bb6, uint32_t morton_count, uint32_t *out) {
    return guard([&] {
        if (morton_count < n) throw std::invalid_argument("morton_count < cnt_obj");
        vec3 lo{bb6[0], bb6[1], bb6[2]}, hi{bb6[3], bb6[4], bb6[5]};
        for (uint32_t i = 0; i < n; i++) out[i] = morton_of<u32>({x[i], y[i], z[i]}, lo, hi);
        for (uint32_t i = n; i < morton_count; i++) out[i] = MortonTraits<u32>::err_code;
    });
}
int orc_morton_u64(uint32_t n, const double *x, const double *y, const double *z,
                   const double *bb6, uint32_t morton_count, uint64_t *out) {
    return guard([&] {
        if (morton_count < n) throw std::invalid_argument("morton_count < cnt_obj");
        vec3 lo{bb6[0], bb6[1], bb6[2]}, hi{bb6[3], bb6[4], bb6[5]};
        for (uint32_t i = 0; i < n; i++) out[i] = morton_of<u64>({x[i], y[i], z[i]}, lo, hi);
        for (uint32_t i = n; i < morton_count; i++) out[i] = MortonTraits<u64>::err_code;
    });
}
int orc_bitonic_u32(uint32_t *keys, uint32_t *vals, uint32_t len) {
    return guard([&] { bitonic_sort_by_key(keys, vals, len); });
}
int orc_bitonic_u64(uint64_t *keys, uint32_t *vals, uint32_t len) {
    return guard([&] { bitonic_sort_by_key(keys, vals, len); });
}

// ---- reduction / karras ------------------------------------------------------------------
int orc_reduce_u32(const uint32_t *sorted, uint32_t cnt, uint32_t level, uint32_t *out_map,
                   uint32_t *out_leaf_count) {
    return guard([&] {
        auto m = reduction_index_map<u32>(sorted, cnt, level);
        std::copy(m.begin(), m.end(), out_map);
        *out_leaf_count = (uint32_t) m.size() - 2;
    });
}
int orc_reduce_u64(const uint64_t *sorted, uint32_t cnt, uint32_t level, uint32_t *out_map,
                   uint32_t *out_leaf_count) {
    return guard([&] {
        auto m = reduction_index_map<u64>(sorted, cnt, level);
        std::copy(m.begin(), m.end(), out_map);
        *out_leaf_count = (uint32_t) m.size() - 2;
    });
}
int orc_karras_u32(const uint32_t *codes, uint32_t L, uint32_t *lc, uint32_t *rc, uint8_t *lf,
                   uint8_t *rf, uint32_t *er) {
    return guard([&] { karras<u32>(codes, L, lc, rc, lf, rf, er); });
}
int orc_karras_u64(const uint64_t *codes, uint32_t L, uint32_t *lc, uint32_t *rc, uint8_t *lf,
                   uint8_t *rf, uint32_t *er) {
    return guard([&] { karras<u64>(codes, L, lc, rc, lf, rf, er); });
}

// ---- CLBVH handle -------------------------------------------------------------------------
void *orc_tree_build(int bits, uint32_t n, const double *x, const double *y, const double *z,
                     const double *bb6, uint32_t level, uint32_t morton_count) {
    auto *t = new TreeHandle();
    t->bits = bits;
    int rc  = guard([&] {
        vec3 lo{bb6[0], bb6[1], bb6[2]}, hi{bb6[3], bb6[4], bb6[5]};
        if (bits == 32) {
            t->t32.build_no_aabb(x, y, z, n, lo, hi, level, morton_count);
            t->t32.compute_aabb(x, y, z);
        } else {
            t->t64.build_no_aabb(x, y, z, n, lo, hi, level, morton_count);
            t->t64.compute_aabb(x, y, z);
        }
    });
    if (rc) {
        delete t;
        return nullptr;
    }
    return t;
}
void orc_tree_free(void *h) { delete (TreeHandle *) h; }

// sizes: [morton_count, L, I]
void orc_tree_sizes(void *h, uint32_t *out3) {
    auto *t = (TreeHandle *) h;
    if (t->bits == 32) {
        out3[0] = t->t32.morton_count;
        out3[1] = t->t32.L;
        out3[2] = t->t32.I;
    } else {
        out3[0] = t->t64.morton_count;
        out3[1] = t->t64.L;
        out3[2] = t->t64.I;
    }
}

} // extern "C"
template<class M>
static void tree_get(const CLBVH<M> &t, M *codes, M *sorted, uint32_t *sort_index,
                     uint32_t *reduc_map, M *leaf_codes, uint32_t *lc, uint32_t *rc, uint8_t *lf,
                     uint8_t *rf, uint32_t *er, double *aabb_min, double *aabb_max) {
    std::copy(t.codes.begin(), t.codes.end(), codes);
    std::copy(t.sorted_codes.begin(), t.sorted_codes.end(), sorted);
    std::copy(t.sort_index.begin(), t.sort_index.end(), sort_index);
    std::copy(t.reduc_map.begin(), t.reduc_map.end(), reduc_map);
    std::copy(t.leaf_codes.begin(), t.leaf_codes.end(), leaf_codes);
    std::copy(t.lchild.begin(), t.lchild.end(), lc);
    std::copy(t.rchild.begin(), t.rchild.end(), rc);
    std::copy(t.lflag.begin(), t.lflag.end(), lf);
    std::copy(t.rflag.begin(), t.rflag.end(), rf);
    std::copy(t.endrange.begin(), t.endrange.end(), er);
    for (size_t i = 0; i < t.aabb_min.size(); i++) {
        aabb_min[3 * i + 0] = t.aabb_min[i].x;
        aabb_min[3 * i + 1] = t.aabb_min[i].y;
        aabb_min[3 * i + 2] = t.aabb_min[i].z;
        aabb_max[3 * i + 0] = t.aabb_max[i].x;
        aabb_max[3 * i + 1] = t.aabb_max[i].y;
        aabb_max[3 * i + 2] = t.aabb_max[i].z;
    }
}
extern "C" {
void orc_tree_get_u32(void *h, uint32_t *codes, uint32_t *sorted, uint32_t *sort_index,
                      uint32_t *reduc_map, uint32_t *leaf_codes, uint32_t *lc, uint32_t *rc,
                      uint8_t *lf, uint8_t *rf, uint32_t *er, double *amin, double *amax) {
    tree_get(((TreeHandle *) h)->t32, codes, sorted, sort_index, reduc_map, leaf_codes, lc, rc, lf,
             rf, er, amin, amax);
}
void orc_tree_get_u64(void *h, uint64_t *codes, uint64_t *sorted, uint32_t *sort_index,
                      uint32_t *reduc_map, uint64_t *leaf_codes, uint32_t *lc, uint32_t *rc,
                      uint8_t *lf, uint8_t *rf, uint32_t *er, double *amin, double *amax) {
    tree_get(((TreeHandle *) h)->t64, codes, sorted, sort_index, reduc_map, leaf_codes, lc, rc, lf,
             rf, er, amin, amax);
}
int orc_tree_field_max(void *h, const double *field, double *out) {
    return guard([&] {
        auto *t = (TreeHandle *) h;
        auto v  = (t->bits == 32) ? t->t32.field_max(field) : t->t64.field_max(field);
        std::copy(v.begin(), v.end(), out);
    });
}

} // extern "C"
// box query of half-width s around each object (CLBVHObjectIteratorTests.cpp:160-230);
// s < 0 means "accept every node" (:57-110).  Two-pass count/fill like the reference test.
template<class M>
static void box_query(const CLBVH<M> &t, uint32_t n, const double *x, const double *y,
                      const double *z, double s, uint32_t *counts, uint32_t *neigh) {
    uint32_t off = t.offset_leaf();
    uint32_t cursor = 0;
    for (uint32_t i = 0; i < n; i++) {
        vec3 r{x[i], y[i], z[i]};
        vec3 lo{r.x - s, r.y - s, r.z - s}, hi{r.x + s, r.y + s, r.z + s};
        uint32_t c = 0;
        t.traverse(
            [&](uint32_t, vec3 nlo, vec3 nhi) {
                if (s < 0) return true;
                vec3 ilo = vmax(nlo, lo), ihi = vmin(nhi, hi);
                return ihi.x >= ilo.x && ihi.y >= ilo.y && ihi.z >= ilo.z;
            },
            [&](uint32_t node) {
                t.for_each_in_leaf(node - off, [&](uint32_t o) {
                    bool ok = true;
                    if (s >= 0) {
                        vec3 p{x[o], y[o], z[o]};
                        vec3 ilo = vmax(p, lo), ihi = vmin(p, hi);
                        ok       = ihi.x >= ilo.x && ihi.y >= ilo.y && ihi.z >= ilo.z;
                    }
                    if (ok) {
                        neigh[cursor++] = o;
                        c++;
                    }
                });
            });
        counts[i] = c;
    }
}
extern "C" {
int orc_tree_box_query(void *h, uint32_t n, const double *x, const double *y, const double *z,
                       double s, uint32_t *counts, uint32_t *neigh) {
    return guard([&] {
        auto *t = (TreeHandle *) h;
        if (t->bits == 32) box_query(t->t32, n, x, y, z, s, counts, neigh);
        else box_query(t->t64, n, x, y, z, s, counts, neigh);
    });
}

// ---- neighbour cache on arbitrary positions ----------------------------------------------
// builds the u32 CLBVH on all n objects (bbox widened by nextafter), the rint field
// (max h * htol) and the two-stage cache of the first n_local objects; out_sizes =
// {L, I, P, PL}; the handle keeps the result for orc_neigh_get
struct NeighHandle {
    CLBVH<u32> t;
    std::vector<double> rint;
    NeighOut o;
};
} // extern "C"
extern "C" {
void *orc_neigh_build(uint32_t n, const double *x, const double *y, const double *z,
                      const double *h, uint32_t n_local, uint32_t level, double htol,
                      double Rkern, uint64_t *out_sizes) {
    auto *nh = new NeighHandle();
    int rc   = guard([&] {
        vec3 lo, hi;
        bbox_nextafter(x, y, z, n, lo, hi);
        nh->t.build_no_aabb(x, y, z, n, lo, hi, level);
        nh->t.compute_aabb(x, y, z);
        nh->rint = nh->t.field_max(h);
        for (auto &v : nh->rint) v *= htol;
        nh->o = neigh_cache_2stages(nh->t, nh->rint, x, y, z, h, n_local, htol, Rkern);
        out_sizes[0] = nh->t.L;
        out_sizes[1] = nh->t.I;
        out_sizes[2] = nh->o.part_cache.ids.size();
        out_sizes[3] = nh->o.leaf_cache.ids.size();
    });
    if (rc) {
        delete nh;
        return nullptr;
    }
    return nh;
}
void orc_neigh_get(void *hnd, uint32_t *cnt, uint32_t *offset, uint32_t *ids, uint32_t *owner,
                   uint32_t *leaf_cnt, uint32_t *leaf_ids) {
    auto *nh = (NeighHandle *) hnd;
    std::copy(nh->o.part_cache.cnt.begin(), nh->o.part_cache.cnt.end(), cnt);
    std::copy(nh->o.part_cache.offset.begin(), nh->o.part_cache.offset.end(), offset);
    std::copy(nh->o.part_cache.ids.begin(), nh->o.part_cache.ids.end(), ids);
    std::copy(nh->o.owner.begin(), nh->o.owner.end(), owner);
    std::copy(nh->o.leaf_cache.cnt.begin(), nh->o.leaf_cache.cnt.end(), leaf_cnt);
    std::copy(nh->o.leaf_cache.ids.begin(), nh->o.leaf_cache.ids.end(), leaf_ids);
}
void orc_neigh_free(void *hnd) { delete (NeighHandle *) hnd; }

// ---- h iteration on an explicit cache ----------------------------------------------------
int orc_iterate_h(int kernel, uint32_t n, const uint32_t *cnt, const uint32_t *offset,
                  const uint32_t *ids, uint32_t nids, const double *x, const double *y,
                  const double *z, const double *h_old, double *h_new, double *eps, double mass,
                  double h_evol_max, double h_evol_iter_max) {
    return guard([&] {
        ObjectCache c;
        c.cnt.assign(cnt, cnt + n);
        c.offset.assign(offset, offset + n);
        c.ids.assign(ids, ids + nids);
        with_kernel(kernel, [&](auto K) {
            iterate_h<decltype(K)>(n, c, x, y, z, h_old, h_new, eps, mass, h_evol_max,
                                   h_evol_iter_max);
        });
    });
}

// ---- varying-AV operators on an explicit cache ------------------------------------------
// div v, curl v (DiffOperator.cpp:24-268) and d(div v)/dt (DiffOperatorDtDivv.cpp:89-175) of
// the n rows of the cache over the positions / velocities / accelerations / h / Omega of all
// the particles the rows name
int orc_cd10_operators(int kernel, uint32_t n, const uint32_t *cnt, const uint32_t *offset,
                       const uint32_t *ids, uint32_t nids, const double *x, const double *y,
                       const double *z, const double *vx, const double *vy, const double *vz,
                       const double *ax, const double *ay, const double *az, const double *h,
                       const double *omega, double mass, double *divv, double *curlx,
                       double *curly, double *curlz, double *dtdivv) {
    return guard([&] {
        ObjectCache c;
        c.cnt.assign(cnt, cnt + n);
        c.offset.assign(offset, offset + n);
        c.ids.assign(ids, ids + nids);
        with_kernel(kernel, [&](auto K) {
            using KT = decltype(K);
            compute_divv_curlv<KT>(n, c, x, y, z, vx, vy, vz, h, omega, mass, divv, curlx,
                                   curly, curlz);
            compute_dtdivv<KT>(n, c, x, y, z, vx, vy, vz, ax, ay, az, h, mass, dtdivv);
        });
    });
}
// alpha update (UpdateViscosity.cpp:120-219)
int orc_cd10_alpha(uint32_t n, double dt, const double *cs, const double *h,
                   const double *alpha, const double *divv, const double *curlx,
                   const double *curly, const double *curlz, const double *dtdivv,
                   double alpha_min, double alpha_max, double sigma_decay, double *alpha_new) {
    return guard([&] {
        update_alpha_cd10(n, dt, cs, h, alpha, divv, curlx, curly, curlz, dtdivv,
                          CD10Params{alpha_min, alpha_max, sigma_decay}, alpha_new);
    });
}

// ---- sink-gas gravity on explicit arrays ---------------------------------------------------
// sinks as 17-double records (get_sinks layout); ext_out (3 n) starts from 0; acc_sph of each
// sink written back into its record
int orc_sink_gas(uint32_t n, const double *x, const double *y, const double *z, uint32_t ns,
                 double *sinks17, double G, double gpart_mass, double soft, double *ext_out) {
    return guard([&] {
        std::vector<Sink> s(ns);
        for (uint32_t i = 0; i < ns; i++) {
            const double *o = sinks17 + 17 * (size_t) i;
            s[i].pos        = {o[0], o[1], o[2]};
            s[i].mass       = o[12];
            s[i].racc       = o[16];
        }
        std::fill(ext_out, ext_out + 3 * (size_t) n, 0.);
        sink_gas(n, x, y, z, s, G, gpart_mass, soft, ext_out, ext_out + n, ext_out + 2 * (size_t) n);
        for (uint32_t i = 0; i < ns; i++) {
            sinks17[17 * (size_t) i + 6] = s[i].acc_sph.x;
            sinks17[17 * (size_t) i + 7] = s[i].acc_sph.y;
            sinks17[17 * (size_t) i + 8] = s[i].acc_sph.z;
        }
    });
}

// ---- periodic ghosts ---------------------------------------------------------------------
int orc_periodic_ghosts(uint32_t n, const double *x, const double *y, const double *z,
                        const double *box6, double int_range, uint32_t cap, uint32_t *out_src,
                        double *out_off, uint32_t *out_count) {
    return guard([&] {
        std::vector<vec3> offs;
        auto ids = periodic_ghost_ids(x, y, z, n, {box6[0], box6[1], box6[2]},
                                      {box6[3], box6[4], box6[5]}, int_range, offs);
        *out_count = (uint32_t) ids.size();
        if (ids.size() > cap) return;
        for (size_t i = 0; i < ids.size(); i++) {
            out_src[i]         = ids[i];
            out_off[3 * i + 0] = offs[i].x;
            out_off[3 * i + 1] = offs[i].y;
            out_off[3 * i + 2] = offs[i].z;
        }
    });
}

// ---- full solver -------------------------------------------------------------------------
void *orc_solver_new(const OrcConfig *cfg, uint32_t n) {
    SolverBase *b = nullptr;
    int rc        = guard([&] {
        with_kernel(cfg->kernel, [&](auto K) {
            auto *s  = new SolverT<decltype(K)>();
            s->s.cfg = *cfg;
            s->s.resize(n);
            b = s;
        });
    });
    return rc ? nullptr : b;
}
void orc_solver_free(void *h) { delete (SolverBase *) h; }

} // extern "C"
template<class F>
static int with_solver(void *h, int kernel, F &&f) {
    return guard([&] {
        with_kernel(kernel, [&](auto K) { f(static_cast<SolverT<decltype(K)> *>((SolverBase *) h)->s); });
    });
}

extern "C" {
// field ids: 0 x 1 y 2 z 3 vx 4 vy 5 vz 6 ax 7 ay 8 az 9 h 10 u 11 du 12 alpha_AV
// 13 soundspeed (saved sound speed of the varying AV) 14 divv 15-17 curlv 18 dtdivv
// 19-21 axyz_ext
int orc_solver_field(void *h, int kernel, int field, double *buf, int set) {
    return with_solver(h, kernel, [&](auto &s) {
        std::vector<double> *f[22] = {&s.x,     &s.y,     &s.z,     &s.vx,        &s.vy,
                                      &s.vz,    &s.ax,    &s.ay,    &s.az,        &s.h,
                                      &s.u,     &s.du,    &s.alpha_av, &s.soundspeed, &s.divv,
                                      &s.curlx, &s.curly, &s.curlz, &s.dtdivv,    &s.axe,
                                      &s.aye,   &s.aze};
        if (field < 0 || field > 21) throw std::invalid_argument("field id");
        if (f[field]->size() != s.N) f[field]->assign(s.N, 0.);
        if (set) std::copy(buf, buf + s.N, f[field]->begin());
        else std::copy(f[field]->begin(), f[field]->end(), buf);
    });
}
// time state: [t, dt, cfl_mult]
int orc_solver_time(void *h, int kernel, double *tdc, int set) {
    return with_solver(h, kernel, [&](auto &s) {
        if (set) {
            s.t        = tdc[0];
            s.dt       = tdc[1];
            s.cfl_mult = tdc[2];
        } else {
            tdc[0] = s.t;
            tdc[1] = s.dt;
            tdc[2] = s.cfl_mult;
        }
    });
}
// Model.hpp:183-200 (add_sink) and pySPHModel.cpp:1047-1065 (get_sinks): 17 doubles per sink,
// pos, velocity, sph_acceleration, ext_acceleration, mass, angular_momentum, accretion_radius
int orc_solver_add_sink(void *h, int kernel, const double *d17) {
    return with_solver(h, kernel, [&](auto &s) {
        Sink k{};
        k.pos  = {d17[0], d17[1], d17[2]};
        k.vel  = {d17[3], d17[4], d17[5]};
        k.mass = d17[12];
        k.racc = d17[16];
        s.sinks.push_back(k);
    });
}
int orc_solver_get_sinks(void *h, int kernel, uint32_t *count, double *out) {
    return with_solver(h, kernel, [&](auto &s) {
        *count = (uint32_t) s.sinks.size();
        if (!out) return;
        for (size_t i = 0; i < s.sinks.size(); i++) {
            const Sink &k = s.sinks[i];
            double *o     = out + 17 * i;
            vec3 v[4]     = {k.pos, k.vel, k.acc_sph, k.acc_ext};
            for (int j = 0; j < 4; j++) {
                o[3 * j]     = v[j].x;
                o[3 * j + 1] = v[j].y;
                o[3 * j + 2] = v[j].z;
            }
            o[12] = k.mass;
            o[13] = k.angmom.x;
            o[14] = k.angmom.y;
            o[15] = k.angmom.z;
            o[16] = k.racc;
        }
    });
}
int orc_solver_evolve_once(void *h, int kernel, StepStats *stats) {
    return with_solver(h, kernel, [&](auto &s) {
        s.evolve_once();
        if (stats) *stats = s.stats;
    });
}

// the density/force pass of Solver::density_force (no time integration)
int orc_solver_density_force(void *h, int kernel, StepStats *stats) {
    return with_solver(h, kernel, [&](auto &s) {
        s.density_force();
        if (stats) *stats = s.stats;
    });
}

// transient products of the last step; sizes: [Nm, G, L, I, n_pairs, n_leaf_pairs]
int orc_solver_sizes(void *h, int kernel, uint64_t *out6) {
    return with_solver(h, kernel, [&](auto &s) {
        out6[0] = s.Nm;
        out6[1] = s.ghost_src.size();
        out6[2] = s.tree.L;
        out6[3] = s.tree.I;
        out6[4] = s.neigh.part_cache.ids.size();
        out6[5] = s.neigh.leaf_cache.ids.size();
    });
}
// merged arrays: 0 mx 1 my 2 mz 3 mh 4 mvx 5 mvy 6 mvz 7 mu 8 momega 9 P 10 cs
// local arrays: 11 omega 12 vsig ; 13 tree bbox (6 values); 14 rint_tree (I+L)
int orc_solver_get_f64(void *h, int kernel, int what, double *out) {
    return with_solver(h, kernel, [&](auto &s) {
        const std::vector<double> *v[13] = {&s.mx, &s.my, &s.mz, &s.mh, &s.mvx, &s.mvy, &s.mvz,
                                            &s.mu, &s.momega, &s.P, &s.cs, &s.omega, &s.vsig};
        if (what >= 0 && what <= 12) std::copy(v[what]->begin(), v[what]->end(), out);
        else if (what == 13) {
            out[0] = s.tree_bmin.x;
            out[1] = s.tree_bmin.y;
            out[2] = s.tree_bmin.z;
            out[3] = s.tree_bmax.x;
            out[4] = s.tree_bmax.y;
            out[5] = s.tree_bmax.z;
        } else if (what == 14) std::copy(s.rint_tree.begin(), s.rint_tree.end(), out);
        else throw std::invalid_argument("what");
    });
}
// 0 ghost_src (G) ; 1 sorted codes (Nm, first cnt_obj) ; 2 sort index (Nm) ; 3 reduc map (L+2)
// 4 owner (N) ; 5 part cnt (N) ; 6 part offset (N) ; 7 part ids (P) ; 8 leaf cnt (L)
// 9 leaf offset (L) ; 10 leaf ids ; 11 lchild 12 rchild 13 endrange (I) ; 14 codes (Nm)
int orc_solver_get_u32(void *h, int kernel, int what, uint32_t *out) {
    return with_solver(h, kernel, [&](auto &s) {
        auto cp = [&](auto &v, size_t n) { std::copy(v.begin(), v.begin() + n, out); };
        switch (what) {
        case 0: cp(s.ghost_src, s.ghost_src.size()); break;
        case 1: cp(s.tree.sorted_codes, s.Nm); break;
        case 2: cp(s.tree.sort_index, s.Nm); break;
        case 3: cp(s.tree.reduc_map, s.tree.reduc_map.size()); break;
        case 4: cp(s.neigh.owner, s.neigh.owner.size()); break;
        case 5: cp(s.neigh.part_cache.cnt, s.N); break;
        case 6: cp(s.neigh.part_cache.offset, s.N); break;
        case 7: cp(s.neigh.part_cache.ids, s.neigh.part_cache.ids.size()); break;
        case 8: cp(s.neigh.leaf_cache.cnt, s.tree.L); break;
        case 9: cp(s.neigh.leaf_cache.offset, s.tree.L); break;
        case 10: cp(s.neigh.leaf_cache.ids, s.neigh.leaf_cache.ids.size()); break;
        case 11: cp(s.tree.lchild, s.tree.I); break;
        case 12: cp(s.tree.rchild, s.tree.I); break;
        case 13: cp(s.tree.endrange, s.tree.I); break;
        case 14: cp(s.tree.codes, s.Nm); break;
        default: throw std::invalid_argument("what");
        }
    });
}
int orc_solver_get_u8(void *h, int kernel, int what, uint8_t *out) {
    return with_solver(h, kernel, [&](auto &s) {
        if (what == 0) std::copy(s.tree.lflag.begin(), s.tree.lflag.end(), out);
        else if (what == 1) std::copy(s.tree.rflag.begin(), s.tree.rflag.end(), out);
        else throw std::invalid_argument("what");
    });
}
int orc_solver_get_aabb(void *h, int kernel, double *amin, double *amax) {
    return with_solver(h, kernel, [&](auto &s) {
        for (size_t i = 0; i < s.tree.aabb_min.size(); i++) {
            amin[3 * i + 0] = s.tree.aabb_min[i].x;
            amin[3 * i + 1] = s.tree.aabb_min[i].y;
            amin[3 * i + 2] = s.tree.aabb_min[i].z;
            amax[3 * i + 0] = s.tree.aabb_max[i].x;
            amax[3 * i + 1] = s.tree.aabb_max[i].y;
            amax[3 * i + 2] = s.tree.aabb_max[i].z;
        }
    });
}

} // extern "C"
