// SPDX-License-Identifier: CeCILL-2.1
//
// capi.cpp -- C++20 host side of the library: the Solver that runs one SPH timestep on the
// GPU (restating shammodels::sph::Solver<f64_3,Kern>::evolve_once, Solver.cpp:1942-3272, for
// one patch) and the extern "C" entry points declared in include/shamx_sph.h.
#include "../../include/shamx_sph.h"
#include "launch.hpp"
#include "tree_layout.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

using namespace shamx;

namespace {

    thread_local std::string g_last_error;

    template<class F>
    int guarded(F &&f) {
        try {
            f();
            return SHAMX_OK;
        } catch (const std::invalid_argument &e) {
            g_last_error = e.what();
            return SHAMX_E_INVALID_ARG;
        } catch (const HipError &e) {
            g_last_error = e.what();
            return SHAMX_E_HIP;
        } catch (const std::bad_alloc &e) {
            g_last_error = e.what();
            return SHAMX_E_OOM;
        } catch (const std::exception &e) {
            g_last_error = e.what();
            return SHAMX_E_INTERNAL;
        }
    }

    void need(bool c, const char *msg) {
        if (!c) throw std::invalid_argument(msg);
    }

    f64 kernel_rkern(int k) {
        f64 r = 0;
        dispatch_kernel(k, [&](auto K) { r = decltype(K)::Rkern; });
        return r;
    }

    // small device scratch for the phase entries that do not take a context
    struct Slots {
        DevArray<u64> v;
        DevArray<u64> part; // per-block partials of the reductions
        Slots() { v.resize(8); }
    };
    thread_local std::unique_ptr<Slots> g_slots;
    Slots &slot_state() {
        if (!g_slots) g_slots = std::make_unique<Slots>();
        return *g_slots;
    }
    u64 *slots() { return slot_state().v.data(); }

    // tests: SHAMX_FULL_PACK=1 rebuilds the force records on corrector reruns instead of
    // refreshing their v, u, P and c_s in place (the two must give the same bytes)
    bool full_pack_forced() {
        const char *e = std::getenv("SHAMX_FULL_PACK");
        return e && e[0] == '1';
    }

    // phase timing with HIP events, read back once the step has synchronised
    struct PhaseTimer {
        std::vector<hipEvent_t> pool;
        std::vector<std::pair<int, std::pair<int, int>>> marks; // phase, (ev start, ev end)
        size_t used = 0;
        hipEvent_t next() {
            if (used == pool.size()) {
                hipEvent_t e;
                SHAMX_HIP(hipEventCreate(&e));
                pool.push_back(e);
            }
            return pool[used++];
        }
        int begin(hipStream_t s) {
            int i = (int) used;
            SHAMX_HIP(hipEventRecord(next(), s));
            return i;
        }
        void end(hipStream_t s, int phase, int b) {
            int i = (int) used;
            SHAMX_HIP(hipEventRecord(next(), s));
            marks.push_back({phase, {b, i}});
        }
        void reset() {
            marks.clear();
            used = 0;
        }
        void collect(double *acc) {
            for (auto &m : marks) {
                float ms = 0;
                SHAMX_HIP(hipEventElapsedTime(&ms, pool[m.second.first], pool[m.second.second]));
                acc[m.first] += ms * 1e-3;
            }
        }
        ~PhaseTimer() {
            for (auto e : pool) (void) hipEventDestroy(e);
        }
    };
    enum Phase {
        PH_GHOST, PH_TREE, PH_SORT, PH_NEIGH, PH_HITER, PH_OMEGA, PH_EOS, PH_FORCE, PH_CORR,
        PH_REORDER, PH_N
    };

} // namespace

struct shamx_ctx {
    TreeBuffers tree;
    NeighBuffers neigh;
    DevArray<u32> alt_k, alt_v;
    DevArray<unsigned char> scratch;
};

struct shamx_solver {
    shamx_config cfg{};
    u32 N = 0, G = 0, cap = 0;
    hipStream_t stream = nullptr;
    f64 Rkern = 0;
    // merged arrays (local particles then ghosts), capacity cap
    DevArray<f64> x, y, z, h, vx, vy, vz, u, omega, P, cs;
    // local arrays
    DevArray<f64> ax, ay, az, du, ax_old, ay_old, az_old, du_old, eps, h_old, vsig, dtc;
    // varying AV (Cullen & Dehnen 2010): alpha and the sound speed saved at the end of the
    // previous step (local), the updated alpha, 5 operator arrays (local), sorted accelerations
    DevArray<f64> alpha_av, soundspeed, alpha_new, sacc;
    DevArray<f64> divv, curlx, curly, curlz, dtdivv; // the operators of the last force pass
    // sinks (sink_edges_helper.hpp:234-248), host-side like the reference's synchronized data,
    // and the particles' external acceleration (axyz_ext, 3 N)
    struct Sink {
        f64 pos[3], vel[3], acc_sph[3], acc_ext[3], mass, angmom[3], racc;
    };
    std::vector<Sink> sinks;
    std::vector<f64> sink_pos_start; // sink positions at the start of the last step
    // adjoint seeds of the sinks carried between steps of a backward pass: 9 per sink,
    // (position, velocity, acceleration from the gas)
    std::vector<f64> adj_sink;
    DevArray<f64> axe, aye, aze;
    DevArray<u64> sink_red;
    DevArray<u32> ghost_src, gwork;
    DevArray<u8> ghost_oid;
    DevArray<unsigned char> gscratch;
    DevArray<u64> red, part;
    // host copies of the step's reductions and counts (pinned, so the reads of one decision
    // queue on the stream and share one synchronisation)
    PinnedWords<u64, 8> hostw;
    PinnedWords<u32, 4> hostc;
    PinnedWords<u64, 3 * 64> hosts; // the sinks' gas accelerations (up to 64 sinks)
    HLoopBuffers hloop;
    // storage order: the local particles are kept in the order of the last tree's sort (the
    // reference's ParticleReordering, Solver.cpp:2043-2048, here at every step), so that the
    // per-particle gathers of the tree, records and outputs read and write nearly contiguous
    // memory.  pid[j] = particle id (the caller's order) of storage slot j; field getters and
    // setters and the inspection getters translate, so every result keeps the caller's order.
    DevArray<u32> pid, pid_tmp, reo_perm, reo_flag, reo_off, reo_count;
    // the next step's reorder, prepared after the step's last tree build: reo_perm holds the
    // permutation and reo_moved (read with the corrector's reductions, no sync of its own) how
    // many particles it moves; reo_known says both describe the current storage and tree
    u32 reo_moved  = 0;
    bool reo_known = false;
    DevArray<unsigned char> reo_scan;
    DevArray<f64> tmp_f64;
    bool pid_identity = true; // pid[j] = j (nothing reordered yet)
    bool tree_valid   = false; // tree.sort_index describes the current storage
    DevArray<unsigned char> frec; // force records, sorted order
    TreeBuffers tree;
    NeighBuffers nb;
    f64 t = 0, dt = 0, cfl_mult = 1e-2; // Solver.hpp:124-147
    shamx_step_log last{};
    PhaseTimer timer;

    DevArray<f64> *field(int id) {
        DevArray<f64> *f[22] = {&x,     &y,     &z,      &vx,       &vy,         &vz,
                                &ax,    &ay,    &az,     &h,        &u,          &du,
                                &alpha_av, &soundspeed, &divv, &curlx, &curly, &curlz,
                                &dtdivv, &axe,  &aye,    &aze};
        need(id >= 0 && id < 22, "field id must be in [0, 21]");
        return f[id];
    }

    void alloc(u32 n) {
        N = n;
        // runs of equal Morton codes ordered spatially on clustered sets (the tie order only;
        // the reference's bitonic network has its own)
        tree.allow_fine_ties = true;
        cap                  = n + n / 3 + 1024;
        for (auto *a : {&x, &y, &z, &h, &vx, &vy, &vz, &u, &omega, &P, &cs}) a->resize(cap);
        for (auto *a : {&ax, &ay, &az, &du, &ax_old, &ay_old, &az_old, &du_old, &eps, &h_old,
                        &vsig, &dtc, &alpha_av, &soundspeed, &alpha_new, &divv, &curlx, &curly,
                        &curlz, &dtdivv, &axe, &aye, &aze})
            a->resize(std::max<u32>(n, 1));
        red.resize(8);
        for (auto *a : {&x, &y, &z, &h, &vx, &vy, &vz, &u, &omega, &ax, &ay, &az, &du,
                        &alpha_av, &soundspeed, &divv, &curlx, &curly, &curlz, &dtdivv, &axe,
                        &aye, &aze})
            fill_f64(stream, a->data(), 0., a->capacity());
        SHAMX_HIP(hipStreamSynchronize(stream));
    }

    void grow(u32 need_cap) {
        u32 c = need_cap + need_cap / 8;
        for (auto *a : {&x, &y, &z, &h, &vx, &vy, &vz, &u, &omega}) a->resize_keep(c, stream);
        P.resize(c);
        cs.resize(c);
        cap = c;
    }

    f64 read_key(int slot) {
        hostw.read_async(stream, red.data() + slot, 1);
        SHAMX_HIP(hipStreamSynchronize(stream));
        return ordered_to_f64(hostw[0]);
    }

    // sorted-layout CSR: local particles in sorted order, records and rows by sorted position
    Csr csr() const {
        return Csr{nb.cnt.data(), nb.offset.data(), nb.ids.data(), nb.lpos.data(),
                   tree.sort_index.data()};
    }
    // the force-interacting subsequences of the rows (written by the Omega pass)
    Csr csr_force() const {
        return Csr{nb.fcnt.data(), nb.offset.data(), nb.fids.data(), nb.lpos.data(),
                   tree.sort_index.data()};
    }

    // the permutation from the storage to the last tree's sorted order of the local particles
    // (in reo_perm), and whether the storage is far enough from it to reorder: the reorder
    // moves every field (22 gathers, about 0.2 ms at 2.1 M particles), so it waits until at
    // least 1/32 of the particles have changed places since the last one
    bool pending_reorder() {
        if (cfg.keep_storage_order || !tree_valid || tree.n_obj != N + G || N == 0) return false;
        if (reo_known) return (u64) reo_moved * 32 >= N;
        const u32 Nm = N + G;
        reo_perm.resize(N);
        reo_flag.resize(Nm);
        reo_off.resize(Nm);
        reo_scan.resize(std::max(reo_scan.size(), scan_scratch_bytes(Nm)));
        local_sorted_order(stream, Nm, N, tree.sort_index.data(), reo_perm.data(),
                           reo_flag.data(), reo_off.data(), reo_scan.data());
        const u32 moved = perm_moved(stream, N, reo_perm.data(), reo_flag.data(), hostc.data());
        return (u64) moved * 32 >= N;
    }

    // after the step's last tree build: the permutation of the next reorder and its count of
    // moved particles on the device (the count is read with the corrector's reductions)
    bool prepare_reorder() {
        reo_known = false;
        if (cfg.keep_storage_order || !tree_valid || tree.n_obj != N + G || N == 0) return false;
        const u32 Nm = N + G;
        reo_perm.resize(N);
        reo_flag.resize(Nm);
        reo_off.resize(Nm);
        reo_count.resize(1);
        reo_scan.resize(std::max(reo_scan.size(), scan_scratch_bytes(Nm)));
        local_sorted_order(stream, Nm, N, tree.sort_index.data(), reo_perm.data(),
                           reo_flag.data(), reo_off.data(), reo_scan.data());
        perm_moved_async(stream, N, reo_perm.data(), reo_count.data());
        return true;
    }

    // permute every per-particle field of the local particles into the sorted order of the
    // last tree build (one gather per field; the images are rebuilt by the step anyway)
    void reorder_storage() {
        const bool go = pending_reorder();
        reo_known     = false;
        if (!go) return;
        const u32 *perm = reo_perm.data();
        for (auto *f : {&x, &y, &z, &h, &vx, &vy, &vz, &u, &ax, &ay, &az, &du, &alpha_av,
                        &soundspeed, &divv, &curlx, &curly, &curlz, &dtdivv, &axe, &aye, &aze}) {
            tmp_f64.resize(f->size()); // same element count: no allocation once sizes settle
            perm_gather_f64(stream, N, perm, f->data(), tmp_f64.data());
            f->swap(tmp_f64);
        }
        if (pid_identity) {
            std::vector<u32> iota(N);
            for (u32 j = 0; j < N; j++) iota[j] = j;
            pid.resize(N);
            SHAMX_HIP(hipMemcpyAsync(pid.data(), iota.data(), N * sizeof(u32),
                                     hipMemcpyHostToDevice, stream));
            SHAMX_HIP(hipStreamSynchronize(stream));
        }
        pid_tmp.resize(N);
        perm_gather_u32(stream, N, perm, pid.data(), pid_tmp.data());
        pid.swap(pid_tmp);
        pid_identity = false;
        tree_valid   = false;
    }

    // the storage order the next step will run in (the pending reorder applied to pid), as
    // particle ids of the caller's order; the storage itself is left as it is
    std::vector<u32> next_storage_order() {
        if (!pending_reorder()) return host_pid();
        std::vector<u32> perm(N), p = host_pid(), out(N);
        SHAMX_HIP(hipMemcpyAsync(perm.data(), reo_perm.data(), N * sizeof(u32),
                                 hipMemcpyDeviceToHost, stream));
        SHAMX_HIP(hipStreamSynchronize(stream));
        for (u32 j = 0; j < N; j++) out[j] = p[perm[j]];
        return out;
    }

    // store the local particles in the given order (order[j] = caller id held by slot j); every
    // field keeps its values per particle, and the next step does not reorder again
    void set_storage_order(const u32 *order) {
        std::vector<u8> seen(N, 0);
        bool identity = true;
        for (u32 j = 0; j < N; j++) {
            need(order[j] < N && !seen[order[j]],
                 "storage order must be a permutation of the particle ids");
            seen[order[j]] = 1;
            identity &= order[j] == j;
        }
        if (identity && pid_identity) return;
        need(!cfg.keep_storage_order, "shamx_solver_set_storage_order: the solver keeps the "
                                      "caller's storage order (keep_storage_order = 1)");
        DevArray<f64> caller;
        caller.resize(std::max<u32>(N, 1));
        pid_tmp.resize(std::max<u32>(N, 1));
        SHAMX_HIP(hipMemcpyAsync(pid_tmp.data(), order, N * sizeof(u32), hipMemcpyHostToDevice,
                                 stream));
        for (auto *f : {&x, &y, &z, &h, &vx, &vy, &vz, &u, &ax, &ay, &az, &du, &alpha_av,
                        &soundspeed, &divv, &curlx, &curly, &curlz, &dtdivv, &axe, &aye, &aze}) {
            field_to_caller(*f, caller.data(), true);
            perm_gather_f64(stream, N, pid_tmp.data(), caller.data(), f->data());
        }
        pid.swap(pid_tmp);
        pid_identity = identity;
        tree_valid   = false;
        SHAMX_HIP(hipStreamSynchronize(stream));
    }

    // field in the caller's particle order <-> storage order
    void field_from_caller(DevArray<f64> &f, const f64 *src, bool on_device) {
        if (pid_identity) {
            SHAMX_HIP(hipMemcpyAsync(f.data(), src, (size_t) N * sizeof(f64),
                                     on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                     stream));
            return;
        }
        const f64 *dsrc = src;
        if (!on_device) {
            tmp_f64.resize(std::max<u32>(N, 1));
            SHAMX_HIP(hipMemcpyAsync(tmp_f64.data(), src, (size_t) N * sizeof(f64),
                                     hipMemcpyHostToDevice, stream));
            dsrc = tmp_f64.data();
        }
        perm_gather_f64(stream, N, pid.data(), dsrc, f.data());
    }
    void field_to_caller(const DevArray<f64> &f, f64 *dst, bool on_device) {
        if (pid_identity) {
            SHAMX_HIP(hipMemcpyAsync(dst, f.data(), (size_t) N * sizeof(f64),
                                     on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                     stream));
            return;
        }
        if (on_device) {
            perm_scatter_f64(stream, N, pid.data(), f.data(), dst);
            return;
        }
        tmp_f64.resize(std::max<u32>(N, 1));
        perm_scatter_f64(stream, N, pid.data(), f.data(), tmp_f64.data());
        SHAMX_HIP(hipMemcpyAsync(dst, tmp_f64.data(), (size_t) N * sizeof(f64),
                                 hipMemcpyDeviceToHost, stream));
    }

    // particle id of every storage slot (host copy)
    std::vector<u32> host_pid() {
        std::vector<u32> p(N);
        if (pid_identity) {
            for (u32 j = 0; j < N; j++) p[j] = j;
        } else if (N) {
            SHAMX_HIP(hipMemcpyAsync(p.data(), pid.data(), N * sizeof(u32),
                                     hipMemcpyDeviceToHost, stream));
            SHAMX_HIP(hipStreamSynchronize(stream));
        }
        return p;
    }
    // the label, in the caller's order, of every merged index: local slots through pid, the
    // periodic images in the order the reference generates them (offset by offset, then by
    // source particle id: BasicSPHGhosts.cpp:260-579), which is the order of generation when
    // the storage holds the caller's order
    std::vector<u32> external_labels() {
        const u32 Nm = N + G;
        std::vector<u32> p = host_pid(), ext(Nm);
        for (u32 j = 0; j < N; j++) ext[j] = p[j];
        if (G) {
            std::vector<u32> src(G);
            std::vector<u8> oid(G);
            SHAMX_HIP(hipMemcpyAsync(src.data(), ghost_src.data(), G * sizeof(u32),
                                     hipMemcpyDeviceToHost, stream));
            SHAMX_HIP(hipMemcpyAsync(oid.data(), ghost_oid.data(), G, hipMemcpyDeviceToHost,
                                     stream));
            SHAMX_HIP(hipStreamSynchronize(stream));
            std::vector<u32> order(G);
            for (u32 g = 0; g < G; g++) order[g] = g;
            std::stable_sort(order.begin(), order.end(), [&](u32 a, u32 b) {
                if (oid[a] != oid[b]) return oid[a] < oid[b];
                return p[src[a]] < p[src[b]];
            });
            for (u32 r = 0; r < G; r++) ext[N + order[r]] = N + r;
        }
        return ext;
    }

    // Solver.cpp:984-1046 + BasicSPHGhosts: images of the local particles after them
    void ghosts() {
        if (!cfg.periodic) {
            G = 0;
            return;
        }
        // interaction range max h * htol_up_coarse * R, from the max key on the device
        fill_u64(stream, red.data(), ORDERED_NEG_INF, 1);
        max_key_f64(stream, N, h.data(), red.data());
        f64 box6[6] = {cfg.box_min[0], cfg.box_min[1], cfg.box_min[2],
                       cfg.box_max[0], cfg.box_max[1], cfg.box_max[2]};
        u32 g = periodic_ghosts(stream, N, x.data(), y.data(), z.data(), h.data(), box6,
                                red.data(), cfg.htol_up_coarse, Rkern, ghost_src, ghost_oid,
                                gwork, gscratch, cap - N, hostc.data());
        if (g > cap - N) {
            grow(N + g);
            g = periodic_ghosts(stream, N, x.data(), y.data(), z.data(), h.data(), box6,
                                red.data(), cfg.htol_up_coarse, Rkern, ghost_src, ghost_oid,
                                gwork, gscratch, cap - N, hostc.data());
        }
        G = g;
    }

    // Solver.cpp:1060-1304.  max_sub subcycles (the configured count in a step); with
    // require_converged false, a last subcycle that did not converge is kept as it is (h of
    // the unconverged particles clamped to the cache radius) instead of failing the step.
    void prestep(u32 max_sub, bool require_converged) {
        reo_known = false; // the tree built here replaces the one a prepared reorder follows
        u32 hstep = 0, hmax = max_sub;
        last.h_iterations = 0;
        for (; hstep < hmax; hstep++) {
            int b = timer.begin(stream);
            ghosts();
            timer.end(stream, PH_GHOST, b);
            u32 Nm = N + G;
            b      = timer.begin(stream);
            tree_stage_codes(stream, tree, Nm, x.data(), y.data(), z.data());
            timer.end(stream, PH_TREE, b);
            b = timer.begin(stream);
            tree_stage_sort(stream, tree, x.data(), y.data(), z.data());
            timer.end(stream, PH_SORT, b);
            b = timer.begin(stream);
            tree_stage_finish(stream, tree, x.data(), y.data(), z.data(), h.data(),
                              (u32) cfg.tree_reduction_level, cfg.htol_up_coarse);
            timer.end(stream, PH_TREE, b);
            tree_valid = true;
            last.n_sorts++;
            b = timer.begin(stream);
            neigh_build(stream, nb, tree, N, x.data(), y.data(), z.data(), h.data(),
                        cfg.htol_up_coarse, Rkern, true);
            timer.end(stream, PH_NEIGH, b);
            need(cfg.gpart_mass != 0, "invalid gpart_mass 0, this configuration can not converge");
            b = timer.begin(stream);
            fill_f64(stream, eps.data(), 100., N);
            SHAMX_HIP(hipMemcpyAsync(h_old.data(), h.data(), N * sizeof(f64),
                                     hipMemcpyDeviceToDevice, stream));
            f64 max_eps = std::numeric_limits<f64>::max(), min_eps = 0;
            if (cfg.h_iter_per_subcycles > 0) {
                HLoopResult hl = h_iterate_loop(
                    stream, cfg.kernel, N, csr(), nb.ps.data(), h_old.data(), h.data(),
                    eps.data(), cfg.gpart_mass, cfg.htol_up_coarse, cfg.htol_up_fine,
                    cfg.epsilon_h, (u32) cfg.h_iter_per_subcycles, cfg.mixed_precision, hloop);
                max_eps = hl.max_eps;
                min_eps = hl.min_eps;
                last.h_iterations += hl.iterations;
            }
            timer.end(stream, PH_HITER, b);
            bool converged = (max_eps < cfg.epsilon_h) && !(min_eps < 0);
            if (!converged && (require_converged || hstep + 1 < hmax)) continue;
            break;
        }
        last.subcycles = std::min(hstep + 1, hmax);
        if (hstep == hmax) throw std::runtime_error("the h iterator is not converged");
        int b = timer.begin(stream);
        // final h of the images and in the sorted records, then Omega fused with the force
        // list filter (h and positions stay fixed for the rest of the step)
        FieldPtrs fh{{h.data()}, 1};
        gather_ghosts(stream, N, G, ghost_src.data(), fh);
        ps_refresh_h(stream, N + G, tree.sort_index.data(), h.data(), nb.ps.data());
        nb.fcnt.resize(std::max<u32>(N + G, 1));
        nb.fids.resize(std::max<size_t>(nb.ids.size(), 1)); // same row offsets as ids
        compute_omega(stream, cfg.kernel, N, csr(), nb.ps.data(), x.data(), y.data(), z.data(),
                      h.data(), omega.data(), cfg.gpart_mass, nb.fcnt.data(), nb.fids.data(),
                      cfg.mixed_precision);
        timer.end(stream, PH_OMEGA, b);
    }

    // sink-sink gravity: SinkParticlesUpdate.cpp:155-200
    void sink_ext_forces() {
        const f64 eps = 1e-9;
        for (auto &s : sinks) s.acc_ext[0] = s.acc_ext[1] = s.acc_ext[2] = 0;
        for (size_t i = 0; i < sinks.size(); i++) {
            f64 sum[3] = {0, 0, 0};
            for (size_t j = 0; j < sinks.size(); j++) {
                f64 r[3];
                for (int k = 0; k < 3; k++) r[k] = sinks[i].pos[k] - sinks[j].pos[k];
                f64 rs = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                f64 gm = cfg.constant_G * sinks[j].mass;
                for (int k = 0; k < 3; k++) sum[k] = sum[k] - gm * r[k] / (rs * rs * rs + eps);
            }
            for (int k = 0; k < 3; k++) sinks[i].acc_ext[k] = sum[k];
        }
    }
    // sink-gas gravity: axyz_ext of the particles (ExternalForces.cpp:48-70) and the sinks'
    // acceleration from the gas (SinkParticlesUpdate.cpp:69-150)
    void sink_gas_forces() {
        u32 S = (u32) sinks.size();
        sink_red.resize(3 * (size_t) S);
        for (u32 k = 0; k < S; k++) {
            const Sink &sk = sinks[k];
            sink_gas(stream, N, x.data(), y.data(), z.data(),
                     SinkArgs{sk.pos[0], sk.pos[1], sk.pos[2], sk.mass, sk.racc, cfg.constant_G,
                              cfg.gpart_mass, cfg.sink_softening},
                     k == 0, axe.data(), aye.data(), aze.data(),
                     sink_red.data() + 3 * (size_t) k, part);
        }
        std::vector<u64> big;
        u64 *r = hosts.data();
        if (3 * (size_t) S > hosts.size()) {
            big.resize(3 * (size_t) S);
            r = big.data();
        }
        SHAMX_HIP(hipMemcpyAsync(r, sink_red.data(), 3 * (size_t) S * 8, hipMemcpyDeviceToHost,
                                 stream));
        SHAMX_HIP(hipStreamSynchronize(stream));
        for (u32 k = 0; k < S; k++)
            for (int c = 0; c < 3; c++) std::memcpy(&sinks[k].acc_sph[c], &r[3 * k + c], 8);
    }
    void sink_kick(f64 dt) { // SinkParticlesUpdate.cpp:42-44, 63-65
        for (auto &s : sinks)
            for (int k = 0; k < 3; k++) s.vel[k] = s.vel[k] + (dt / 2) * (s.acc_sph[k] + s.acc_ext[k]);
    }

    void evolve_once() {
        need(N > 0, "evolve_once on an empty particle set");
        timer.reset();
        last = shamx_step_log{};
        SHAMX_HIP(hipStreamSynchronize(stream));
        auto t0   = std::chrono::steady_clock::now();
        f64 t_cur = t, dtc_ = dt;
        f64 box6[6] = {cfg.box_min[0], cfg.box_min[1], cfg.box_min[2],
                       cfg.box_max[0], cfg.box_max[1], cfg.box_max[2]};
        const bool has_sinks = !sinks.empty();
        int b = timer.begin(stream);
        reorder_storage();
        timer.end(stream, PH_REORDER, b);
        sink_pos_start.resize(3 * sinks.size());
        for (size_t k = 0; k < sinks.size(); k++)
            for (int c = 0; c < 3; c++) sink_pos_start[3 * k + c] = sinks[k].pos[c];
        if (has_sinks) { // sink predictor: SinkParticlesUpdate.cpp:23-46 (Solver.cpp:2017)
            sink_ext_forces();
            sink_kick(dtc_);
            for (auto &s : sinks)
                for (int k = 0; k < 3; k++) s.pos[k] = s.pos[k] + dtc_ * s.vel[k];
        }
        b = timer.begin(stream);
        // with sinks, the sink forces see the drifted positions before the periodic wrap
        // (Solver.cpp:2033-2040)
        int wrap_now = cfg.periodic && !has_sinks;
        predictor(stream, N, dtc_, x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(),
                  ax.data(), ay.data(), az.data(), u.data(), du.data(), wrap_now,
                  wrap_now ? nullptr_box() : nullptr);
        timer.end(stream, PH_CORR, b);
        if (has_sinks) {
            sink_ext_forces();
            sink_gas_forces();
            if (cfg.periodic) wrap_positions(stream, N, x.data(), y.data(), z.data(), nullptr_box());
        }
        (void) box6;
        prestep((u32) cfg.h_max_subcycles, true);
        const bool count_pending = prepare_reorder();

        f64 next_cfl = 0;
        u32 corr_iter = 0;
        bool rerun    = false;
        do {
            if (corr_iter == 50)
                throw std::runtime_error("the corrector has made over 50 loops, either their is "
                                         "a bug, either you are using a dt that is too large");
            b = timer.begin(stream);
            if (cfg.keep_storage_order) record_pass_inputs(corr_iter);
            FieldPtrs fp{{h.data(), u.data(), vx.data(), vy.data(), vz.data(), omega.data()}, 6};
            gather_ghosts(stream, N, G, ghost_src.data(), fp);
            // EOS of the merged set fused with the sorted force records (a corrector rerun
            // rewrites only the velocity, u, P and c_s of the records)
            frec.resize(force_record_bytes(cfg.mixed_precision) * (size_t) (N + G));
            pack_force(stream, cfg.kernel, N + G, tree.sort_index.data(), x.data(), y.data(),
                       z.data(), h.data(), vx.data(), vy.data(), vz.data(), u.data(),
                       omega.data(), P.data(), cs.data(), frec.data(), cfg.gpart_mass,
                       EosArgs{cfg.eos_kind, cfg.gamma, cfg.eos_cs, cfg.eos_cs0, cfg.eos_q,
                               cfg.eos_r0},
                       cfg.mixed_precision, corr_iter > 0 && !full_pack_forced());
            timer.end(stream, PH_EOS, b);
            if (cfg.av_kind == 1) {
                // Solver.cpp:2370-2470: div v, curl v, d(div v)/dt from the previous
                // accelerations, alpha update over this dt, alpha into the records
                b = timer.begin(stream);
                u32 Nm = N + G;
                sacc.resize(3 * (size_t) std::max<u32>(Nm, 1));
                const f64 *acc_in[3] = {ax.data(), ay.data(), az.data()};
                f64 *acc_out[3] = {sacc.data(), sacc.data() + Nm, sacc.data() + 2 * (size_t) Nm};
                adjoint_gather_sorted(stream, Nm, N, tree.sort_index.data(), ghost_src.data(), 3,
                                      acc_in, acc_out);
                f64 *w5[5] = {divv.data(), curlx.data(), curly.data(), curlz.data(),
                              dtdivv.data()};
                cd10_alpha(stream, cfg.kernel, N, Nm, csr_force(), frec.data(),
                           tree.sort_index.data(), ghost_src.data(), acc_out, soundspeed.data(),
                           h.data(), alpha_av.data(), dtc_, cfg.gpart_mass, cfg.alpha_min,
                           cfg.alpha_max, cfg.sigma_decay, w5, alpha_new.data());
                timer.end(stream, PH_EOS, b);
            }
            b = timer.begin(stream);
            ForceArgs fa{cfg.gpart_mass,
                         cfg.alpha_u,
                         cfg.alpha_AV,
                         cfg.beta_AV,
                         cfg.cfl_cour * cfl_mult,
                         cfg.cfl_force * cfl_mult,
                         cfg.mixed_precision,
                         cfg.av_kind == 1};
            if (has_sinks) {
                fa.ext_x = axe.data();
                fa.ext_y = aye.data();
                fa.ext_z = aze.data();
            }
            compute_force(stream, cfg.kernel, N, csr_force(), frec.data(), x.data(), y.data(),
                          z.data(), vx.data(), vy.data(), vz.data(), h.data(), omega.data(),
                          u.data(), P.data(), cs.data(), ax.data(), ay.data(), az.data(),
                          du.data(), ax_old.data(), ay_old.data(), az_old.data(), du_old.data(),
                          vsig.data(), dtc.data(), fa, red.data() + 3, red.data() + 7, part);
            timer.end(stream, PH_FORCE, b);
            last.n_force_evals++;
            b = timer.begin(stream);
            corrector(stream, N, dtc_ / 2, vx.data(), vy.data(), vz.data(), ax.data(), ay.data(),
                      az.data(), ax_old.data(), ay_old.data(), az_old.data(), u.data(), du.data(),
                      du_old.data(), cfg.gpart_mass, red.data() + 4, part);
            timer.end(stream, PH_CORR, b);
            const u64 *k = hostw.data();
            hostw.read_async(stream, red.data() + 3, 5);
            if (count_pending) hostc.read_async(stream, reo_count.data(), 1, 2);
            SHAMX_HIP(hipStreamSynchronize(stream));
            if (count_pending) {
                reo_moved = hostc[2];
                reo_known = true;
            }
            f64 dt_min         = ordered_to_f64(k[0]);
            last.n_interacting = k[4];
            f64 max_e2 = ordered_to_f64(k[1]);
            f64 sum_vsq, energy;
            std::memcpy(&sum_vsq, &k[2], 8);
            std::memcpy(&energy, &k[3], 8);
            f64 rank_veps_v = std::sqrt(max_e2);
            f64 vmean       = std::sqrt(sum_vsq / f64(N));
            f64 eps_v       = rank_veps_v / vmean;
            if (vmean <= 0) eps_v = 0;
            last.eps_v  = eps_v;
            last.energy = energy;
            if (eps_v > 1e-2) {
                rerun = true;
                cfl_mult /= 2;
            } else {
                rerun    = false;
                next_cfl = dt_min;
                if (has_sinks) sink_kick(dtc_); // sink corrector (Solver.cpp:2620)
                // the updated alpha is kept once the step is accepted (Solver.cpp:2622-2650)
                if (cfg.av_kind == 1)
                    SHAMX_HIP(hipMemcpyAsync(alpha_av.data(), alpha_new.data(), N * sizeof(f64),
                                             hipMemcpyDeviceToDevice, stream));
            }
            corr_iter++;
        } while (rerun);
        // sound speed kept for the next step's viscosity update (Solver.cpp:3126-3160)
        if (cfg.av_kind == 1)
            SHAMX_HIP(hipMemcpyAsync(soundspeed.data(), cs.data(), N * sizeof(f64),
                                     hipMemcpyDeviceToDevice, stream));
        SHAMX_HIP(hipStreamSynchronize(stream));
        auto t1 = std::chrono::steady_clock::now();

        f64 ph[PH_N] = {0};
        timer.collect(ph);
        last.t_total  = std::chrono::duration<f64>(t1 - t0).count();
        last.rate     = f64(N) / last.t_total;
        last.t_ghost  = ph[PH_GHOST];
        last.t_tree   = ph[PH_TREE];
        last.t_sort   = ph[PH_SORT];
        last.t_neigh  = ph[PH_NEIGH];
        last.t_hiter  = ph[PH_HITER];
        last.t_omega  = ph[PH_OMEGA];
        last.t_eos    = ph[PH_EOS];
        last.t_force  = ph[PH_FORCE];
        last.t_corr   = ph[PH_CORR];
        last.t_reorder = ph[PH_REORDER];
        last.n_local  = N;
        last.n_merged = N + G;
        last.n_leaf   = tree.L;
        last.n_pairs  = nb.n_pairs;
        last.corrector_iterations = corr_iter;
        last.next_dt  = next_cfl;
        last.time     = t_cur;
        last.dt       = dtc_;
        // Solver.cpp:3235-3263
        dt        = next_cfl;
        t         = t_cur + dtc_;
        f64 stiff = cfg.cfl_multiplier_stiffness;
        cfl_mult  = (cfl_mult * stiff + 1.) / (stiff + 1.);
    }

    // One density and force pass on the current state, with no time integration: the
    // prestep with a single subcycle (Solver.cpp:1060-1304; particles whose h does not
    // converge in it keep the value clamped to the cache radius), the EOS, the force loop
    // with the external sink forces (Solver.cpp:2247-2400, UpdateDerivs.cpp:88-288) and the
    // CFL reduction (Solver.cpp:3095-3124).  a and du are overwritten, their previous values
    // saved as in a step; positions, velocities, u and the time are left alone.
    void density_force() {
        need(N > 0, "density/force pass on an empty particle set");
        need(cfg.av_kind == 0, "the density/force pass takes the constant viscosity only");
        timer.reset();
        last = shamx_step_log{};
        SHAMX_HIP(hipStreamSynchronize(stream));
        auto t0              = std::chrono::steady_clock::now();
        const bool has_sinks = !sinks.empty();
        int b                = timer.begin(stream);
        reorder_storage();
        timer.end(stream, PH_REORDER, b);
        if (has_sinks) {
            sink_ext_forces();
            sink_gas_forces();
        }
        prestep(1, false);
        b = timer.begin(stream);
        FieldPtrs fp{{h.data(), u.data(), vx.data(), vy.data(), vz.data(), omega.data()}, 6};
        gather_ghosts(stream, N, G, ghost_src.data(), fp);
        frec.resize(force_record_bytes(cfg.mixed_precision) * (size_t) (N + G));
        pack_force(stream, cfg.kernel, N + G, tree.sort_index.data(), x.data(), y.data(),
                   z.data(), h.data(), vx.data(), vy.data(), vz.data(), u.data(), omega.data(),
                   P.data(), cs.data(), frec.data(), cfg.gpart_mass,
                   EosArgs{cfg.eos_kind, cfg.gamma, cfg.eos_cs, cfg.eos_cs0, cfg.eos_q,
                           cfg.eos_r0},
                   cfg.mixed_precision);
        timer.end(stream, PH_EOS, b);
        b = timer.begin(stream);
        ForceArgs fa{cfg.gpart_mass,           cfg.alpha_u, cfg.alpha_AV, cfg.beta_AV,
                     cfg.cfl_cour * cfl_mult, cfg.cfl_force * cfl_mult, cfg.mixed_precision,
                     false};
        if (has_sinks) {
            fa.ext_x = axe.data();
            fa.ext_y = aye.data();
            fa.ext_z = aze.data();
        }
        compute_force(stream, cfg.kernel, N, csr_force(), frec.data(), x.data(), y.data(),
                      z.data(), vx.data(), vy.data(), vz.data(), h.data(), omega.data(),
                      u.data(), P.data(), cs.data(), ax.data(), ay.data(), az.data(), du.data(),
                      ax_old.data(), ay_old.data(), az_old.data(), du_old.data(), vsig.data(),
                      dtc.data(), fa, red.data() + 3, red.data() + 7, part);
        timer.end(stream, PH_FORCE, b);
        const u64 *k = hostw.data();
        hostw.read_async(stream, red.data() + 3, 5);
        SHAMX_HIP(hipStreamSynchronize(stream));
        auto t1 = std::chrono::steady_clock::now();
        f64 ph[PH_N] = {0};
        timer.collect(ph);
        last.t_total       = std::chrono::duration<f64>(t1 - t0).count();
        last.rate          = f64(N) / last.t_total;
        last.t_ghost       = ph[PH_GHOST];
        last.t_tree        = ph[PH_TREE];
        last.t_sort        = ph[PH_SORT];
        last.t_neigh       = ph[PH_NEIGH];
        last.t_hiter       = ph[PH_HITER];
        last.t_omega       = ph[PH_OMEGA];
        last.t_eos         = ph[PH_EOS];
        last.t_force       = ph[PH_FORCE];
        last.t_reorder     = ph[PH_REORDER];
        last.n_local       = N;
        last.n_merged      = N + G;
        last.n_leaf        = tree.L;
        last.n_pairs       = nb.n_pairs;
        last.n_force_evals = 1;
        last.n_interacting = k[4];
        last.next_dt       = ordered_to_f64(k[0]);
        last.time          = t;
        last.dt            = dt;
    }

    // periodic box for the predictor wrap, kept in device memory
    DevArray<f64> box_dev;
    const f64 *nullptr_box() {
        if (!cfg.periodic) return nullptr;
        if (box_dev.size() != 6) {
            box_dev.resize(6);
            f64 b[6] = {cfg.box_min[0], cfg.box_min[1], cfg.box_min[2],
                        cfg.box_max[0], cfg.box_max[1], cfg.box_max[2]};
            SHAMX_HIP(hipMemcpy(box_dev.data(), b, sizeof(b), hipMemcpyHostToDevice));
        }
        return box_dev.data();
    }

    // the neighbour cache in the reference layout (row = particle id, entries = particle ids),
    // rebuilt on the host from the sorted layout: 5 counts, 6 offsets, 7 ids
    void reference_layout_csr(int what, u32 *dst) {
        u32 Nm = N + G;
        std::vector<u32> ext = external_labels(), pslot(N);
        for (u32 j = 0; j < N; j++) pslot[ext[j]] = j; // particle id -> storage slot
        std::vector<u32> cnt_s(Nm), off_s(Nm), ids_s(nb.ids.size()), sidx(Nm), inv(Nm);
        SHAMX_HIP(hipMemcpyAsync(cnt_s.data(), nb.cnt.data(), Nm * 4, hipMemcpyDeviceToHost, stream));
        SHAMX_HIP(hipMemcpyAsync(off_s.data(), nb.offset.data(), Nm * 4, hipMemcpyDeviceToHost,
                                 stream));
        SHAMX_HIP(hipMemcpyAsync(sidx.data(), tree.sort_index.data(), Nm * 4,
                                 hipMemcpyDeviceToHost, stream));
        if (what == 7 && ids_s.size())
            SHAMX_HIP(hipMemcpyAsync(ids_s.data(), nb.ids.data(), ids_s.size() * 4,
                                     hipMemcpyDeviceToHost, stream));
        SHAMX_HIP(hipStreamSynchronize(stream));
        for (u32 k = 0; k < Nm; k++) inv[sidx[k]] = k;
        u64 acc = 0;
        for (u32 a = 0; a < N; a++) {
            u32 row = inv[pslot[a]], c = cnt_s[row];
            if (what == 5) dst[a] = c;
            if (what == 6) dst[a] = (u32) acc;
            if (what == 7)
                for (u32 k = 0; k < c; k++) dst[acc + k] = ext[sidx[ids_s[off_s[row] + k]]];
            acc += c;
        }
    }

    // leaf lists in the packed layout: 8 counts, 9 offsets, 10 ids
    void reference_layout_leaves(int what, u32 *dst) {
        u32 L = tree.L;
        std::vector<u32> cnt(L), off(L), ids(nb.leaf_ids.size());
        SHAMX_HIP(hipMemcpyAsync(cnt.data(), nb.leaf_cnt.data(), L * 4, hipMemcpyDeviceToHost,
                                 stream));
        SHAMX_HIP(hipMemcpyAsync(off.data(), nb.leaf_offset.data(), L * 4, hipMemcpyDeviceToHost,
                                 stream));
        if (what == 10 && ids.size())
            SHAMX_HIP(hipMemcpyAsync(ids.data(), nb.leaf_ids.data(), ids.size() * 4,
                                     hipMemcpyDeviceToHost, stream));
        SHAMX_HIP(hipStreamSynchronize(stream));
        u64 acc = 0;
        for (u32 l = 0; l < L; l++) {
            if (what == 8) dst[l] = cnt[l];
            if (what == 9) dst[l] = (u32) acc;
            if (what == 10)
                for (u32 k = 0; k < cnt[l]; k++) dst[acc + k] = ids[off[l] + k];
            acc += cnt[l];
        }
    }

    // the adjoint of the last step reads the step's own intermediate products
    DevArray<f64> adj_work;
    // the adjoint's tape: the inputs (local v and u) of every corrector pass of the last step,
    // kept when the storage keeps the caller's order (the adjoint's mode); the FP64 force
    // records of one pass rebuilt from them, and the step's end state (v, u) kept meanwhile
    std::vector<DevArray<f64>> tape;
    DevArray<unsigned char> adj_frec;
    DevArray<f64> adj_keep;

    // The adjoint's product store: what the adjoint of a step reads (its tree order, images,
    // neighbour rows, sorted positions, end-of-step positions, h and Omega, the tape of its
    // corrector passes, its log and sinks), saved after the step and loaded back before its
    // adjoint, so a backward pass runs every step of a segment once instead of twice.
    struct ProductSlot {
        bool valid = false;
        u32 G = 0, L = 0, n_obj = 0, passes = 0;
        shamx_step_log last{};
        std::vector<Sink> sinks;
        std::vector<f64> sink_pos_start;
        DevArray<u32> sort_index, ghost_src, lpos, cnt, offset, ids, fcnt, fids;
        DevArray<double4> ps;
        DevArray<f64> x, y, z, h, omega;
        std::vector<DevArray<f64>> tape;
    };
    std::vector<ProductSlot> slots;

    template<class T>
    void copy_dev(DevArray<T> &dst, const DevArray<T> &src, size_t n) {
        dst.resize(std::max<size_t>(n, 1));
        if (n)
            SHAMX_HIP(hipMemcpyAsync(dst.data(), src.data(), n * sizeof(T),
                                     hipMemcpyDeviceToDevice, stream));
    }

    void products_save(u32 slot) {
        need(cfg.keep_storage_order, "products: the solver must keep the caller's storage order "
                                     "(keep_storage_order = 1)");
        need(N > 0 && tree.L > 0 && last.corrector_iterations >= 1 &&
                 tape.size() >= last.corrector_iterations,
             "products: run evolve_once first");
        if (slots.size() <= slot) slots.resize(slot + 1);
        ProductSlot &p = slots[slot];
        const u32 Nm   = N + G;
        p.G = G;
        p.L = tree.L;
        p.n_obj  = tree.n_obj;
        p.passes = last.corrector_iterations;
        p.last   = last;
        p.sinks  = sinks;
        p.sink_pos_start = sink_pos_start;
        copy_dev(p.sort_index, tree.sort_index, tree.n_obj);
        copy_dev(p.ghost_src, ghost_src, G);
        using U32Pair = std::pair<DevArray<u32> *, DevArray<u32> *>;
        for (auto [d, sarr] : {U32Pair{&p.lpos, &nb.lpos}, U32Pair{&p.cnt, &nb.cnt},
                               U32Pair{&p.offset, &nb.offset}, U32Pair{&p.ids, &nb.ids},
                               U32Pair{&p.fcnt, &nb.fcnt}, U32Pair{&p.fids, &nb.fids}})
            copy_dev(*d, *sarr, sarr->size());
        copy_dev(p.ps, nb.ps, nb.ps.size());
        using F64Pair = std::pair<DevArray<f64> *, DevArray<f64> *>;
        for (auto [d, sarr] : {F64Pair{&p.x, &x}, F64Pair{&p.y, &y}, F64Pair{&p.z, &z},
                               F64Pair{&p.h, &h}, F64Pair{&p.omega, &omega}})
            copy_dev(*d, *sarr, Nm);
        if (p.tape.size() < p.passes) p.tape.resize(p.passes);
        for (u32 i = 0; i < p.passes; i++) copy_dev(p.tape[i], tape[i], 4 * (size_t) N);
        p.valid = true;
        SHAMX_HIP(hipStreamSynchronize(stream));
    }

    // the solver then holds the saved step's products (its fields x, y, z, h, Omega included:
    // the forward state is left for the adjoint, and a restore of the fields must follow
    // before the next step)
    void products_load(u32 slot) {
        need(slot < slots.size() && slots[slot].valid, "products: nothing saved in this slot");
        ProductSlot &p = slots[slot];
        const u32 Nm   = N + p.G;
        need(Nm <= cap, "products: the saved step has more images than the solver holds");
        G           = p.G;
        tree.L      = p.L;
        tree.n_obj  = p.n_obj;
        last        = p.last;
        sinks       = p.sinks;
        sink_pos_start = p.sink_pos_start;
        copy_dev(tree.sort_index, p.sort_index, p.n_obj);
        copy_dev(ghost_src, p.ghost_src, p.G);
        using U32Pair = std::pair<DevArray<u32> *, DevArray<u32> *>;
        for (auto [d, sarr] : {U32Pair{&nb.lpos, &p.lpos}, U32Pair{&nb.cnt, &p.cnt},
                               U32Pair{&nb.offset, &p.offset}, U32Pair{&nb.ids, &p.ids},
                               U32Pair{&nb.fcnt, &p.fcnt}, U32Pair{&nb.fids, &p.fids}})
            copy_dev(*d, *sarr, sarr->size());
        copy_dev(nb.ps, p.ps, p.ps.size());
        using F64Pair = std::pair<DevArray<f64> *, DevArray<f64> *>;
        for (auto [d, sarr] : {F64Pair{&x, &p.x}, F64Pair{&y, &p.y}, F64Pair{&z, &p.z},
                               F64Pair{&h, &p.h}, F64Pair{&omega, &p.omega}}) {
            // the merged arrays keep their capacity (cap >= N + G): copy into place
            if (Nm)
                SHAMX_HIP(hipMemcpyAsync(d->data(), sarr->data(), Nm * sizeof(f64),
                                         hipMemcpyDeviceToDevice, stream));
        }
        if (tape.size() < p.passes) tape.resize(p.passes);
        for (u32 i = 0; i < p.passes; i++) copy_dev(tape[i], p.tape[i], 4 * (size_t) N);
        tree_valid = false; // the tree no longer describes the storage of the forward state
        SHAMX_HIP(hipStreamSynchronize(stream));
    }

    void record_pass_inputs(u32 pass) {
        if (tape.size() <= pass) tape.resize(pass + 1);
        DevArray<f64> &t = tape[pass];
        t.resize(4 * (size_t) std::max<u32>(N, 1));
        const f64 *src[4] = {vx.data(), vy.data(), vz.data(), u.data()};
        for (int k = 0; k < 4; k++)
            SHAMX_HIP(hipMemcpyAsync(t.data() + k * (size_t) N, src[k], N * sizeof(f64),
                                     hipMemcpyDeviceToDevice, stream));
    }

    // the step's state as corrector pass `pass` (1-based) saw it: local v and u from the tape,
    // the images' from their sources, P, c_s and the FP64 force records (adj_frec) from them
    void adjoint_records(u32 pass) {
        const DevArray<f64> &t = tape[pass - 1];
        f64 *dst[4] = {vx.data(), vy.data(), vz.data(), u.data()};
        for (int k = 0; k < 4; k++)
            SHAMX_HIP(hipMemcpyAsync(dst[k], t.data() + k * (size_t) N, N * sizeof(f64),
                                     hipMemcpyDeviceToDevice, stream));
        FieldPtrs fp{{h.data(), u.data(), vx.data(), vy.data(), vz.data(), omega.data()}, 6};
        gather_ghosts(stream, N, G, ghost_src.data(), fp);
        adj_frec.resize(force_record_bytes(0) * (size_t) (N + G));
        pack_force(stream, cfg.kernel, N + G, tree.sort_index.data(), x.data(), y.data(),
                   z.data(), h.data(), vx.data(), vy.data(), vz.data(), u.data(), omega.data(),
                   P.data(), cs.data(), adj_frec.data(), cfg.gpart_mass,
                   EosArgs{cfg.eos_kind, cfg.gamma, cfg.eos_cs, cfg.eos_cs0, cfg.eos_q,
                           cfg.eos_r0},
                   0, false);
    }

    // the adjoint of the last step over all its corrector passes, on FP64 records rebuilt
    // from the tape (so a mixed-precision step is differentiated as the FP64 step map at its
    // own state); the step's end state is put back afterwards
    void adjoint_passes(f64 *const *seeds11, const SinkAdjoint *sk) {
        AdjointArgs A = adjoint_args(); // the configuration checks come first
        const u32 r   = last.corrector_iterations;
        need(r >= 1 && tape.size() >= r,
             "adjoint: the last step kept no corrector inputs (run it with keep_storage_order = 1)");
        adj_keep.resize(4 * (size_t) std::max<u32>(N, 1));
        f64 *cur[4] = {vx.data(), vy.data(), vz.data(), u.data()};
        for (int k = 0; k < 4; k++)
            SHAMX_HIP(hipMemcpyAsync(adj_keep.data() + k * (size_t) N, cur[k], N * sizeof(f64),
                                     hipMemcpyDeviceToDevice, stream));
        adj_frec.resize(force_record_bytes(0) * (size_t) (N + G));
        A.frec = adj_frec.data();
        std::function<void(u32)> prep = [this](u32 pass) { adjoint_records(pass); };
        adjoint_step(stream, A, seeds11, adj_work, sk, r, &prep);
        // end state: images, P and c_s as the last pass left them, then the local v and u
        adjoint_records(r);
        for (int k = 0; k < 4; k++)
            SHAMX_HIP(hipMemcpyAsync(cur[k], adj_keep.data() + k * (size_t) N, N * sizeof(f64),
                                     hipMemcpyDeviceToDevice, stream));
    }
    AdjointArgs adjoint_args() {
        need(N > 0 && tree.L > 0, "adjoint: run evolve_once first");
        need(cfg.eos_kind == 0 || cfg.eos_kind == 1 || (cfg.eos_kind == 2 && !cfg.periodic),
             "adjoint: available for the adiabatic and isothermal equations of state, and for "
             "LP07 with free boundaries (periodic images sit at other radii)");
        need(cfg.av_kind == 0, "adjoint: available for the constant artificial viscosity");
        need(pid_identity, "adjoint: the solver must keep the caller's storage order "
                           "(keep_storage_order = 1)");
        need(sinks.empty() || !cfg.periodic,
             "adjoint: sinks only with free boundaries (the sink forces see unwrapped positions)");
        f64 hfact = 0;
        dispatch_kernel(cfg.kernel, [&](auto K) { hfact = decltype(K)::hfactd; });
        return AdjointArgs{N,
                           N + G,
                           cfg.kernel,
                           last.dt,
                           cfg.gpart_mass,
                           hfact,
                           cfg.gamma,
                           cfg.alpha_u,
                           cfg.alpha_AV,
                           cfg.beta_AV,
                           cfg.eos_kind,
                           cfg.eos_cs,
                           cfg.eos_q,
                           tree.sort_index.data(),
                           G ? ghost_src.data() : nullptr,
                           nb.lpos.data(),
                           nb.cnt.data(),
                           nb.offset.data(),
                           nb.ids.data(),
                           nb.fcnt.data(),
                           nb.fids.data(),
                           frec.data(),
                           nb.ps.data(),
                           h.data(),
                           omega.data(),
                           cfg.adjoint_fp32 ? 1 : 0};
    }

    // sink-sink gravity reversed (SinkParticlesUpdate.cpp:155-200): a_ext,i = -sum_j G M_j
    // r_ij / (|r_ij|^3 + 1e-9); pbar += (d a_ext / d p)^T abar at the positions pos
    void sink_ext_vjp(const f64 *pos, const f64 *abar, f64 *pbar) const {
        const f64 eps = 1e-9;
        const size_t S = sinks.size();
        for (size_t i = 0; i < S; i++)
            for (size_t j = 0; j < S; j++) {
                if (i == j) continue;
                f64 r[3], v[3];
                for (int c = 0; c < 3; c++) {
                    r[c] = pos[3 * i + c] - pos[3 * j + c];
                    v[c] = abar[3 * i + c];
                }
                f64 d   = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                f64 den = d * d * d + eps;
                f64 rv  = r[0] * v[0] + r[1] * v[1] + r[2] * v[2];
                f64 gm  = cfg.constant_G * sinks[j].mass;
                for (int c = 0; c < 3; c++) {
                    f64 jv = v[c] / den - 3 * d * r[c] * rv / (den * den);
                    pbar[3 * i + c] -= gm * jv;
                    pbar[3 * j + c] += gm * jv;
                }
            }
    }

    // the adjoint of the last step with its sinks: the sink kick of the corrector, the
    // sink-gas and sink-sink forces at the drifted positions, the sink predictor (kick, drift),
    // around the gas adjoint; adj_sink carries (pbar, wbar, abar_gas) from step to step
    void adjoint_step_with_sinks(f64 *const *seeds11) {
        const size_t S = sinks.size();
        const f64 dt   = last.dt;
        if (adj_sink.size() != 9 * S) adj_sink.assign(9 * S, 0.);
        std::vector<f64> pmid(3 * S), mass(S), racc(S), ab_gas(3 * S), ab_ext(3 * S),
            pbar_gas(3 * S, 0.), pbar(3 * S), wbar(3 * S);
        for (size_t k = 0; k < S; k++) {
            mass[k] = sinks[k].mass;
            racc[k] = sinks[k].racc;
            for (int c = 0; c < 3; c++) {
                pmid[3 * k + c] = sinks[k].pos[c];
                f64 wb          = adj_sink[9 * k + 3 + c];
                // corrector kick w_end = w_half + dt/2 (a_gas + a_ext)
                ab_gas[3 * k + c] = adj_sink[9 * k + 6 + c] + dt / 2 * wb;
                ab_ext[3 * k + c] = dt / 2 * wb;
                pbar[3 * k + c]   = adj_sink[9 * k + c];
                wbar[3 * k + c]   = wb;
            }
        }
        SinkAdjoint sa{(u32) S,         pmid.data(), mass.data(), racc.data(),
                       cfg.constant_G,  cfg.sink_softening,       ab_gas.data(),
                       pbar_gas.data(), x.data(),    y.data(),    z.data()};
        adjoint_passes(seeds11, &sa);
        SHAMX_HIP(hipStreamSynchronize(stream));
        for (size_t i = 0; i < 3 * S; i++) pbar[i] += pbar_gas[i];
        sink_ext_vjp(pmid.data(), ab_ext.data(), pbar.data());
        // predictor: p_mid = p_old + dt w_half, w_half = w_old + dt/2 (a_gas,old + a_ext,old)
        std::vector<f64> ab_old(3 * S);
        for (size_t i = 0; i < 3 * S; i++) {
            wbar[i] += dt * pbar[i];
            ab_old[i] = dt / 2 * wbar[i];
        }
        sink_ext_vjp(sink_pos_start.data(), ab_old.data(), pbar.data());
        for (size_t k = 0; k < S; k++)
            for (int c = 0; c < 3; c++) {
                adj_sink[9 * k + c]     = pbar[3 * k + c];
                adj_sink[9 * k + 3 + c] = wbar[3 * k + c];
                adj_sink[9 * k + 6 + c] = ab_old[3 * k + c];
            }
    }

    ~shamx_solver() {
        if (stream) (void) hipStreamDestroy(stream);
    }
};

namespace {
    template<class T>
    void d2h(hipStream_t s, T *dst, const T *src, size_t n) {
        if (n) SHAMX_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    template<class T>
    void d2d(hipStream_t s, T *dst, const T *src, size_t n) {
        if (dst && n) SHAMX_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    }
    // node records -> interleaved aabb (3 per cell) and rint
    void unpack_nodes(const std::vector<f64> &nodes, u32 cells, f64 *amin, f64 *amax, f64 *rint) {
        for (u32 c = 0; c < cells; c++) {
            const f64 *r = nodes.data() + (size_t) c * NODE_F64;
            if (amin)
                for (int k = 0; k < 3; k++) amin[3 * c + k] = r[k];
            if (amax)
                for (int k = 0; k < 3; k++) amax[3 * c + k] = r[3 + k];
            if (rint) rint[c] = r[6];
        }
    }
} // namespace

extern "C" {

const char *shamx_last_error(void) { return g_last_error.c_str(); }

int shamx_get_device_count(int *count) {
    return guarded([&] { SHAMX_HIP(hipGetDeviceCount(count)); });
}

int shamx_ctx_create(shamx_ctx **out) {
    return guarded([&] {
        need(out != nullptr, "out is NULL");
        *out = new shamx_ctx();
    });
}
int shamx_ctx_destroy(shamx_ctx *ctx) {
    return guarded([&] { delete ctx; });
}

int shamx_bbox(shamx_ctx *ctx, void *stream, uint32_t n, const double *x, const double *y,
               const double *z, double *bbox6) {
    return guarded([&] {
        need(ctx && n > 0, "shamx_bbox: null context or n == 0");
        ctx->scratch.resize(bbox_scratch_bytes(n));
        bbox_nextafter((hipStream_t) stream, n, x, y, z, bbox6, ctx->scratch.data());
    });
}

int shamx_morton(shamx_ctx *ctx, void *stream, uint32_t n, const double *x, const double *y,
                 const double *z, const double *bbox6, uint32_t *codes) {
    return guarded([&] {
        need(ctx != nullptr, "shamx_morton: null context");
        morton_codes((hipStream_t) stream, n, x, y, z, bbox6, codes);
    });
}

int shamx_sort(shamx_ctx *ctx, void *stream, uint32_t n, uint32_t key_bits, uint32_t *keys,
               uint32_t *perm) {
    return guarded([&] {
        need(ctx != nullptr, "shamx_sort: null context");
        need(key_bits >= 1 && key_bits <= 32, "shamx_sort: key_bits must be in [1, 32]");
        ctx->alt_k.resize(std::max<u32>(n, 1));
        ctx->alt_v.resize(std::max<u32>(n, 1));
        ctx->scratch.resize(radix_sort_scratch_bytes(n));
        radix_sort_pairs((hipStream_t) stream, n, key_bits, keys, perm, ctx->alt_k.data(),
                         ctx->alt_v.data(), ctx->scratch.data());
    });
}

int shamx_sort_noalloc(void *stream, uint32_t n, uint32_t key_bits, uint32_t *keys,
                       uint32_t *perm) {
    return guarded([&] {
        radix_sort_noalloc((hipStream_t) stream, n, key_bits, keys, perm);
    });
}

int shamx_tree_build(shamx_ctx *ctx, void *stream, uint32_t n, const double *x, const double *y,
                     const double *z, const double *h, uint32_t reduction_level, double htol) {
    return guarded([&] {
        need(ctx != nullptr, "shamx_tree_build: null context");
        tree_build((hipStream_t) stream, ctx->tree, n, x, y, z, h, reduction_level, htol);
        SHAMX_HIP(hipStreamSynchronize((hipStream_t) stream));
        tree_check(ctx->tree);
    });
}

int shamx_tree_reduce_karras(shamx_ctx *ctx, void *stream, uint32_t n, const double *x,
                             const double *y, const double *z, uint32_t reduction_level) {
    return guarded([&] {
        need(ctx != nullptr, "shamx_tree_reduce_karras: null context");
        hipStream_t s = (hipStream_t) stream;
        tree_stage_codes(s, ctx->tree, n, x, y, z);
        tree_stage_sort(s, ctx->tree, x, y, z);
        tree_stage_karras(s, ctx->tree, reduction_level);
    });
}

int shamx_tree_aabb_hmax(shamx_ctx *ctx, void *stream, const double *x, const double *y,
                         const double *z, const double *h, double htol) {
    return guarded([&] {
        need(ctx && ctx->tree.L > 0, "shamx_tree_aabb_hmax: build the tree first");
        tree_stage_aabb((hipStream_t) stream, ctx->tree, x, y, z, h, htol);
        SHAMX_HIP(hipStreamSynchronize((hipStream_t) stream));
        tree_check(ctx->tree);
    });
}

int shamx_tree_sizes(shamx_ctx *ctx, uint32_t *out4) {
    return guarded([&] {
        out4[0] = ctx->tree.n_obj;
        out4[1] = ctx->tree.L;
        out4[2] = ctx->tree.I;
        out4[3] = 0;
    });
}

int shamx_tree_get(shamx_ctx *ctx, void *stream, uint32_t *sorted, uint32_t *sort_index,
                   uint32_t *reduc_map, uint32_t *lchild, uint32_t *rchild, uint8_t *lflag,
                   uint8_t *rflag, uint32_t *endrange, double *aabb_min, double *aabb_max,
                   double *rint) {
    return guarded([&] {
        hipStream_t s = (hipStream_t) stream;
        TreeBuffers &t = ctx->tree;
        d2d(s, sorted, t.sorted.data(), t.n_obj);
        d2d(s, sort_index, t.sort_index.data(), t.n_obj);
        d2d(s, reduc_map, t.reduc_map.data(), (size_t) t.L + 2);
        d2d(s, lchild, t.lchild.data(), t.I);
        d2d(s, rchild, t.rchild.data(), t.I);
        d2d(s, lflag, t.lflag.data(), t.I);
        d2d(s, rflag, t.rflag.data(), t.I);
        d2d(s, endrange, t.endrange.data(), t.I);
        if (aabb_min || aabb_max || rint) {
            u32 cells = t.I + t.L;
            std::vector<f64> nodes((size_t) cells * NODE_F64), a(3 * (size_t) cells),
                bb(3 * (size_t) cells), r(cells);
            d2h(s, nodes.data(), t.nodes.data(), nodes.size());
            SHAMX_HIP(hipStreamSynchronize(s));
            unpack_nodes(nodes, cells, a.data(), bb.data(), r.data());
            if (aabb_min)
                SHAMX_HIP(hipMemcpyAsync(aabb_min, a.data(), a.size() * 8, hipMemcpyHostToDevice, s));
            if (aabb_max)
                SHAMX_HIP(hipMemcpyAsync(aabb_max, bb.data(), bb.size() * 8, hipMemcpyHostToDevice, s));
            if (rint)
                SHAMX_HIP(hipMemcpyAsync(rint, r.data(), r.size() * 8, hipMemcpyHostToDevice, s));
            SHAMX_HIP(hipStreamSynchronize(s));
        }
    });
}

int shamx_neigh_build(shamx_ctx *ctx, void *stream, uint32_t n_local, const double *x,
                      const double *y, const double *z, const double *h, double htol,
                      double Rkern) {
    return guarded([&] {
        need(ctx && ctx->tree.L > 0, "shamx_neigh_build: build the tree first");
        need(n_local <= ctx->tree.n_obj, "shamx_neigh_build: n_local > tree object count");
        neigh_build((hipStream_t) stream, ctx->neigh, ctx->tree, n_local, x, y, z, h, htol, Rkern,
                    false);
    });
}

int shamx_neigh_sizes(shamx_ctx *ctx, uint32_t *out4) {
    return guarded([&] {
        out4[0] = ctx->neigh.n_local;
        out4[1] = (u32) (ctx->neigh.n_pairs & 0xffffffffu);
        out4[2] = (u32) (ctx->neigh.n_pairs >> 32);
        out4[3] = (u32) ctx->neigh.n_leaf_pairs;
    });
}

int shamx_neigh_get(shamx_ctx *ctx, void *stream, uint32_t *cnt, uint32_t *offset, uint32_t *ids,
                    uint32_t *owner, uint32_t *leaf_cnt, uint32_t *leaf_offset,
                    uint32_t *leaf_ids) {
    return guarded([&] {
        hipStream_t s   = (hipStream_t) stream;
        NeighBuffers &n = ctx->neigh;
        d2d(s, cnt, n.cnt.data(), n.n_local);
        d2d(s, offset, n.offset.data(), n.n_local);
        d2d(s, ids, n.ids.data(), n.n_pairs);
        d2d(s, owner, n.owner.data(), n.n_local);
        d2d(s, leaf_cnt, n.leaf_cnt.data(), ctx->tree.L);
        d2d(s, leaf_offset, n.leaf_offset.data(), ctx->tree.L);
        d2d(s, leaf_ids, n.leaf_ids.data(), n.n_leaf_pairs);
    });
}

int shamx_h_iterate(void *stream, int kernel, uint32_t n, const uint32_t *cnt,
                    const uint32_t *offset, const uint32_t *ids, const double *x,
                    const double *y, const double *z, const double *h_old, double *h_new,
                    double *eps, double gpart_mass, double h_evol_max, double h_evol_iter_max,
                    int mixed, double *eps_minmax_host) {
    return guarded([&] {
        hipStream_t s = (hipStream_t) stream;
        u64 *sl       = slots();
        h_iterate(s, kernel, n, Csr{cnt, offset, ids}, nullptr, x, y, z, h_old, h_new, eps,
                  gpart_mass, h_evol_max, h_evol_iter_max, sl, mixed, slot_state().part);
        if (eps_minmax_host) {
            u64 k[2];
            SHAMX_HIP(hipMemcpyAsync(k, sl, 16, hipMemcpyDeviceToHost, s));
            SHAMX_HIP(hipStreamSynchronize(s));
            eps_minmax_host[0] = ordered_to_f64(k[0]);
            eps_minmax_host[1] = ordered_to_f64(k[1]);
        }
    });
}

int shamx_omega(void *stream, int kernel, uint32_t n, const uint32_t *cnt, const uint32_t *offset,
                const uint32_t *ids, const double *x, const double *y, const double *z,
                const double *h, double *omega, double gpart_mass) {
    return guarded([&] {
        compute_omega((hipStream_t) stream, kernel, n, Csr{cnt, offset, ids}, nullptr, x, y, z, h,
                      omega, gpart_mass, nullptr, nullptr);
    });
}

int shamx_eos(void *stream, int kernel, uint32_t nm, int eos_kind, double gamma, double cs,
              double cs0, double q, double r0, const double *h, const double *u, const double *x,
              const double *y, const double *z, double *P, double *soundspeed,
              double gpart_mass) {
    return guarded([&] {
        need(eos_kind >= 0 && eos_kind <= 2, "shamx_eos: unknown eos kind");
        compute_eos((hipStream_t) stream, kernel, nm, h, u, x, y, z, P, soundspeed, gpart_mass,
                    EosArgs{eos_kind, gamma, cs, cs0, q, r0});
    });
}

int shamx_force(void *stream, int kernel, uint32_t n, const uint32_t *cnt, const uint32_t *offset,
                const uint32_t *ids, const double *x, const double *y, const double *z,
                const double *vx, const double *vy, const double *vz, const double *h,
                const double *omega, const double *u, const double *P, const double *cs,
                double *ax, double *ay, double *az, double *du, double *ax_old, double *ay_old,
                double *az_old, double *du_old, double *vsig, double *dt_cfl, double gpart_mass,
                double alpha_u, double alpha_AV, double beta_AV, double C_cour, double C_force,
                int mixed, double *dt_min_host) {
    return guarded([&] {
        hipStream_t s = (hipStream_t) stream;
        u64 *sl       = slots();
        compute_force(s, kernel, n, Csr{cnt, offset, ids}, nullptr, x, y, z, vx, vy, vz, h,
                      omega, u, P, cs, ax, ay, az, du, ax_old, ay_old, az_old, du_old, vsig,
                      dt_cfl,
                      ForceArgs{gpart_mass, alpha_u, alpha_AV, beta_AV, C_cour, C_force, mixed},
                      sl, nullptr, slot_state().part);
        if (dt_min_host) {
            u64 k;
            SHAMX_HIP(hipMemcpyAsync(&k, sl, 8, hipMemcpyDeviceToHost, s));
            SHAMX_HIP(hipStreamSynchronize(s));
            *dt_min_host = ordered_to_f64(k);
        }
    });
}

int shamx_cfl(void *stream, uint32_t n, const double *h, const double *vsig, const double *ax,
              const double *ay, const double *az, double C_cour, double C_force, double *dt_cfl,
              double *dt_min_host) {
    return guarded([&] {
        hipStream_t s = (hipStream_t) stream;
        u64 *sl       = slots();
        fill_u64(s, sl, ORDERED_POS_INF, 1);
        cfl_dt(s, n, h, vsig, ax, ay, az, C_cour, C_force, dt_cfl, sl, slot_state().part);
        if (dt_min_host) {
            u64 k;
            SHAMX_HIP(hipMemcpyAsync(&k, sl, 8, hipMemcpyDeviceToHost, s));
            SHAMX_HIP(hipStreamSynchronize(s));
            *dt_min_host = ordered_to_f64(k);
        }
    });
}

int shamx_predictor(void *stream, uint32_t n, double dt, double *x, double *y, double *z,
                    double *vx, double *vy, double *vz, const double *ax, const double *ay,
                    const double *az, double *u, const double *du, int periodic,
                    const double *box6_host) {
    return guarded([&] {
        hipStream_t s = (hipStream_t) stream;
        DevArray<f64> box(6);
        if (periodic) {
            need(box6_host != nullptr, "shamx_predictor: periodic needs box6");
            SHAMX_HIP(hipMemcpyAsync(box.data(), box6_host, 48, hipMemcpyHostToDevice, s));
        }
        predictor(s, n, dt, x, y, z, vx, vy, vz, ax, ay, az, u, du, periodic, box.data());
        SHAMX_HIP(hipStreamSynchronize(s));
    });
}

int shamx_corrector(void *stream, uint32_t n, double hdt, double *vx, double *vy, double *vz,
                    const double *ax, const double *ay, const double *az, const double *ax_old,
                    const double *ay_old, const double *az_old, double *u, const double *du,
                    const double *du_old, double gpart_mass, double *out3_host) {
    return guarded([&] {
        hipStream_t s = (hipStream_t) stream;
        u64 *sl       = slots();
        corrector(s, n, hdt, vx, vy, vz, ax, ay, az, ax_old, ay_old, az_old, u, du, du_old,
                  gpart_mass, sl, slot_state().part);
        u64 k[3];
        SHAMX_HIP(hipMemcpyAsync(k, sl, 24, hipMemcpyDeviceToHost, s));
        SHAMX_HIP(hipStreamSynchronize(s));
        if (out3_host) {
            out3_host[0] = ordered_to_f64(k[0]);
            std::memcpy(&out3_host[1], &k[1], 8);
            std::memcpy(&out3_host[2], &k[2], 8);
        }
    });
}

// ---------------------------------------------------------------------------- solver
int shamx_solver_create(const shamx_config *cfg, uint32_t n, shamx_solver **out) {
    return guarded([&] {
        need(cfg && out, "shamx_solver_create: null argument");
        need(cfg->kernel >= 0 && cfg->kernel <= 5, "unknown SPH kernel");
        need(cfg->eos_kind >= 0 && cfg->eos_kind <= 2, "unknown equation of state");
        need(cfg->tree_reduction_level >= 0, "tree_reduction_level must be >= 0");
        need(cfg->av_kind == 0 || cfg->av_kind == 1,
             "av_kind must be 0 (constant) or 1 (varying, Cullen & Dehnen 2010)");
        need(!(cfg->av_kind == 1 && cfg->mixed_precision),
             "the varying artificial viscosity runs in FP64 only");
        need(cfg->htol_up_coarse >= cfg->htol_up_fine,
             "htol_up_coarse_cycle must be larger or equal to htol_up_fine_cycle");
        if (cfg->periodic)
            for (int k = 0; k < 3; k++)
                need(cfg->box_max[k] > cfg->box_min[k], "periodic box must have positive size");
        auto s = std::make_unique<shamx_solver>();
        s->cfg = *cfg;
        SHAMX_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->Rkern = kernel_rkern(cfg->kernel);
        s->alloc(n);
        *out = s.release();
    });
}

int shamx_solver_destroy(shamx_solver *s) {
    return guarded([&] { delete s; });
}

int shamx_solver_set_field(shamx_solver *s, int field, const double *src, int on_device) {
    return guarded([&] {
        auto *f = s->field(field);
        // a device source may still be being written on another stream
        if (on_device) SHAMX_HIP(hipDeviceSynchronize());
        s->field_from_caller(*f, src, on_device != 0);
        SHAMX_HIP(hipStreamSynchronize(s->stream));
    });
}

int shamx_solver_get_field(shamx_solver *s, int field, double *dst, int on_device) {
    return guarded([&] {
        auto *f = s->field(field);
        if (on_device) SHAMX_HIP(hipDeviceSynchronize());
        s->field_to_caller(*f, dst, on_device != 0);
        SHAMX_HIP(hipStreamSynchronize(s->stream));
    });
}

int shamx_solver_get_storage_order(shamx_solver *s, uint32_t *order) {
    return guarded([&] {
        need(s != nullptr && order != nullptr, "shamx_solver_get_storage_order: null argument");
        std::vector<u32> p = s->next_storage_order();
        std::copy(p.begin(), p.end(), order);
    });
}

int shamx_solver_set_storage_order(shamx_solver *s, const uint32_t *order) {
    return guarded([&] {
        need(s != nullptr && order != nullptr, "shamx_solver_set_storage_order: null argument");
        s->set_storage_order(order);
    });
}

int shamx_solver_get_time(shamx_solver *s, double *tdc3) {
    return guarded([&] {
        tdc3[0] = s->t;
        tdc3[1] = s->dt;
        tdc3[2] = s->cfl_mult;
    });
}

int shamx_solver_set_time(shamx_solver *s, const double *tdc3) {
    return guarded([&] {
        s->t        = tdc3[0];
        s->dt       = tdc3[1];
        s->cfl_mult = tdc3[2];
    });
}

int shamx_solver_add_sink(shamx_solver *s, double mass, const double *pos3, const double *vel3,
                          double accretion_radius) {
    return guarded([&] {
        need(s && pos3 && vel3, "shamx_solver_add_sink: null argument");
        shamx_solver::Sink k{};
        for (int c = 0; c < 3; c++) {
            k.pos[c] = pos3[c];
            k.vel[c] = vel3[c];
        }
        k.mass = mass;
        k.racc = accretion_radius;
        s->sinks.push_back(k);
    });
}

int shamx_solver_get_sinks(shamx_solver *s, uint32_t *count, double *out17) {
    return guarded([&] {
        need(s && count, "shamx_solver_get_sinks: null argument");
        *count = (uint32_t) s->sinks.size();
        if (!out17) return;
        for (size_t i = 0; i < s->sinks.size(); i++) {
            const auto &k = s->sinks[i];
            double *o     = out17 + 17 * i;
            for (int c = 0; c < 3; c++) {
                o[c]      = k.pos[c];
                o[3 + c]  = k.vel[c];
                o[6 + c]  = k.acc_sph[c];
                o[9 + c]  = k.acc_ext[c];
                o[13 + c] = k.angmom[c];
            }
            o[12] = k.mass;
            o[16] = k.racc;
        }
    });
}

int shamx_solver_set_sinks(shamx_solver *s, uint32_t count, const double *in17) {
    return guarded([&] {
        need(s && (count == 0 || in17), "shamx_solver_set_sinks: null argument");
        s->sinks.assign(count, shamx_solver::Sink{});
        for (uint32_t i = 0; i < count; i++) {
            auto &k         = s->sinks[i];
            const double *o = in17 + 17 * (size_t) i;
            for (int c = 0; c < 3; c++) {
                k.pos[c]     = o[c];
                k.vel[c]     = o[3 + c];
                k.acc_sph[c] = o[6 + c];
                k.acc_ext[c] = o[9 + c];
                k.angmom[c]  = o[13 + c];
            }
            k.mass = o[12];
            k.racc = o[16];
        }
    });
}

int shamx_solver_evolve_once(shamx_solver *s, shamx_step_log *log) {
    return guarded([&] {
        s->evolve_once();
        if (log) *log = s->last;
    });
}

int shamx_solver_density_force(shamx_solver *s, shamx_step_log *log) {
    return guarded([&] {
        need(s != nullptr, "shamx_solver_density_force: null solver");
        s->density_force();
        if (log) *log = s->last;
    });
}

int shamx_solver_adjoint_step(shamx_solver *s, double *const *seeds11) {
    return guarded([&] {
        need(s && seeds11, "shamx_solver_adjoint_step: null argument");
        for (int k = 0; k < 11; k++) need(seeds11[k] != nullptr, "adjoint: null seed array");
        SHAMX_HIP(hipDeviceSynchronize()); // seeds may still be being written on another stream
        if (s->sinks.empty()) s->adjoint_passes(seeds11, nullptr);
        else s->adjoint_step_with_sinks(seeds11);
        SHAMX_HIP(hipStreamSynchronize(s->stream));
    });
}

int shamx_solver_products_save(shamx_solver *s, uint32_t slot) {
    return guarded([&] {
        need(s != nullptr, "shamx_solver_products_save: null solver");
        s->products_save(slot);
    });
}

int shamx_solver_products_load(shamx_solver *s, uint32_t slot) {
    return guarded([&] {
        need(s != nullptr, "shamx_solver_products_load: null solver");
        s->products_load(slot);
    });
}

int shamx_solver_products_free(shamx_solver *s) {
    return guarded([&] {
        need(s != nullptr, "shamx_solver_products_free: null solver");
        SHAMX_HIP(hipStreamSynchronize(s->stream));
        s->slots.clear();
    });
}

int shamx_solver_adjoint_density(shamx_solver *s, const double *rho_bar, double *xbar,
                                 double *ybar, double *zbar) {
    return guarded([&] {
        need(s && rho_bar && xbar && ybar && zbar, "shamx_solver_adjoint_density: null argument");
        SHAMX_HIP(hipDeviceSynchronize()); // inputs may still be being written on another stream
        adjoint_density(s->stream, s->adjoint_args(), rho_bar, xbar, ybar, zbar, s->adj_work);
        s->adj_sink.assign(9 * s->sinks.size(), 0.); // a backward pass starts here
        SHAMX_HIP(hipStreamSynchronize(s->stream));
    });
}

int shamx_solver_stream(shamx_solver *s, void **stream) {
    return guarded([&] { *stream = (void *) s->stream; });
}

int shamx_solver_sizes(shamx_solver *s, uint64_t *out6) {
    return guarded([&] {
        out6[0] = s->N + s->G;
        out6[1] = s->G;
        out6[2] = s->tree.L;
        out6[3] = s->tree.I;
        out6[4] = s->nb.n_pairs;
        out6[5] = s->nb.n_leaf_pairs;
    });
}

int shamx_solver_get_f64(shamx_solver *s, int what, double *dst) {
    return guarded([&] {
        hipStream_t st = s->stream;
        u32 Nm = s->N + s->G, cells = s->tree.I + s->tree.L;
        DevArray<f64> *m[11] = {&s->x, &s->y, &s->z, &s->h, &s->vx, &s->vy, &s->vz,
                                &s->u, &s->omega, &s->P, &s->cs};
        // per-particle arrays come back in the caller's particle order (images in the
        // reference's order after them)
        auto relabel = [&](const f64 *src, u32 n, bool merged) {
            std::vector<f64> tmp(n);
            d2h(st, tmp.data(), src, n);
            SHAMX_HIP(hipStreamSynchronize(st));
            std::vector<u32> lab = merged ? s->external_labels() : s->host_pid();
            for (u32 j = 0; j < n; j++) dst[lab[j]] = tmp[j];
        };
        if (what >= 0 && what <= 10) relabel(m[what]->data(), Nm, true);
        else if (what == 11) relabel(s->omega.data(), s->N, false);
        else if (what == 12) relabel(s->vsig.data(), s->N, false);
        else if (what == 13) d2h(st, dst, s->tree.bbox6.data(), 6);
        else if (what >= 14 && what <= 16) {
            std::vector<f64> nodes((size_t) cells * NODE_F64);
            d2h(st, nodes.data(), s->tree.nodes.data(), nodes.size());
            SHAMX_HIP(hipStreamSynchronize(st));
            unpack_nodes(nodes, cells, what == 15 ? dst : nullptr, what == 16 ? dst : nullptr,
                         what == 14 ? dst : nullptr);
        } else throw std::invalid_argument("shamx_solver_get_f64: unknown id");
        SHAMX_HIP(hipStreamSynchronize(st));
    });
}

int shamx_solver_get_u32(shamx_solver *s, int what, uint32_t *dst) {
    return guarded([&] {
        hipStream_t st = s->stream;
        u32 Nm = s->N + s->G, L = s->tree.L, I = s->tree.I, N = s->N, G = s->G;
        // arrays indexed by, or holding, particle indices are translated to the caller's
        // particle order (images: the reference's order)
        auto host = [&](const u32 *src, u32 n) {
            std::vector<u32> v(n);
            d2h(st, v.data(), src, n);
            SHAMX_HIP(hipStreamSynchronize(st));
            return v;
        };
        switch (what) {
        case 0: {
            std::vector<u32> ext = s->external_labels(), src = host(s->ghost_src.data(), G);
            for (u32 g = 0; g < G; g++) dst[ext[N + g] - N] = ext[src[g]];
            break;
        }
        case 1: d2h(st, dst, s->tree.sorted.data(), Nm); break;
        case 2: {
            std::vector<u32> ext = s->external_labels(), si = host(s->tree.sort_index.data(), Nm);
            for (u32 k = 0; k < Nm; k++) dst[k] = ext[si[k]];
            break;
        }
        case 3: d2h(st, dst, s->tree.reduc_map.data(), (size_t) L + 2); break;
        case 4: {
            std::vector<u32> p = s->host_pid(), o = host(s->nb.owner.data(), N);
            for (u32 j = 0; j < N; j++) dst[p[j]] = o[j];
            break;
        }
        case 5:
        case 6:
        case 7: s->reference_layout_csr(what, dst); break;
        case 8:
        case 9:
        case 10: s->reference_layout_leaves(what, dst); break;
        case 11: d2h(st, dst, s->tree.lchild.data(), I); break;
        case 12: d2h(st, dst, s->tree.rchild.data(), I); break;
        case 13: d2h(st, dst, s->tree.endrange.data(), I); break;
        case 14: {
            std::vector<u32> ext = s->external_labels(), c = host(s->tree.codes.data(), Nm);
            for (u32 m = 0; m < Nm; m++) dst[ext[m]] = c[m];
            break;
        }
        default: throw std::invalid_argument("shamx_solver_get_u32: unknown id");
        }
        SHAMX_HIP(hipStreamSynchronize(st));
    });
}

int shamx_solver_get_u8(shamx_solver *s, int what, uint8_t *dst) {
    return guarded([&] {
        hipStream_t st = s->stream;
        if (what == 0) d2h(st, dst, s->tree.lflag.data(), s->tree.I);
        else if (what == 1) d2h(st, dst, s->tree.rflag.data(), s->tree.I);
        else if (what == 2)
            d2h(st, dst, s->frec.data(),
                force_record_bytes(s->cfg.mixed_precision) * (size_t) (s->N + s->G));
        else throw std::invalid_argument("shamx_solver_get_u8: unknown id");
        SHAMX_HIP(hipStreamSynchronize(st));
    });
}

} // extern "C"
