/* SPDX-License-Identifier: CeCILL-2.1
 *
 * shamx_sph.h -- C ABI of the MI355X-native SPH timestep library (libshamx_sph.so).
 *
 * Drop-in boundary for the reference's SPH step (shammodels::sph, f64_3 vectors).  The
 * reference drives each phase as a solver-graph node (shamrock::solvergraph::INode, INode.hpp)
 * whose kernels are launched through sham::distributed_data_kernel_call
 * (shambackends/include/shambackends/kernel_call_distrib.hpp:78-108) on per-patch spans; here
 * each phase is one extern "C" entry on ONE patch of SoA device arrays, and the composite step
 * is shamx_solver_evolve_once.
 *
 * Conventions
 *  - every function returns SHAMX_OK (0) or an error code; no exception crosses the ABI;
 *    shamx_last_error() returns the thread-local message of the last failure
 *    (reference: shambase::make_except_with_loc / throw_with_loc, e.g.
 *    CompressedLeafBVH.cpp:39-42, Solver.cpp:1096-1102, 2252-2256).
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).  Device pointers are
 *    plain `double*` / `uint32_t*` on the current HIP device.
 *  - a context (shamx_ctx) owns the transient buffers of the step (tree, neighbour cache,
 *    sort and scan scratch); the caller owns field arrays.  One context per host thread.
 */
#ifndef SHAMX_SPH_H
#define SHAMX_SPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum shamx_status {
    SHAMX_OK              = 0,
    SHAMX_E_INVALID_ARG   = 1,
    SHAMX_E_HIP           = 2,
    SHAMX_E_NOT_CONVERGED = 3,
    SHAMX_E_OOM           = 4,
    SHAMX_E_INTERNAL      = 5
};

/* SPH kernels, as the reference's template argument (Solver.cpp:3276-3282) */
enum shamx_kernel { SHAMX_M4 = 0, SHAMX_M6 = 1, SHAMX_M8 = 2, SHAMX_C2 = 3, SHAMX_C4 = 4, SHAMX_C6 = 5 };
/* equations of state (SolverConfig EOSConfig: Adiabatic, Isothermal, LocallyIsothermalLP07) */
enum shamx_eos { SHAMX_EOS_ADIABATIC = 0, SHAMX_EOS_ISOTHERMAL = 1, SHAMX_EOS_LP07 = 2 };

/* Subset of shammodels::sph::SolverConfig (SolverConfig.hpp:432-1400) used by the hot path. */
typedef struct shamx_config {
    int32_t kernel;               /* shamx_kernel */
    int32_t periodic;             /* 1: BCConfig::Periodic, 0: BCConfig::Free */
    int32_t eos_kind;             /* shamx_eos */
    int32_t tree_reduction_level; /* SolverConfig.hpp:584, default 3 */
    int32_t h_iter_per_subcycles; /* SolverConfig.hpp:608, default 50 */
    int32_t h_max_subcycles;      /* SolverConfig.hpp:609, default 100 */
    double box_min[3], box_max[3];/* simulation box (Model::resize_simulation_box) */
    double gpart_mass;
    double gamma, eos_cs, eos_cs0, eos_q, eos_r0;
    double alpha_u, alpha_AV, beta_AV;            /* AVConfig_Constant (AVConfig.hpp:49-55) */
    double htol_up_coarse, htol_up_fine, epsilon_h; /* SolverConfig.hpp:605-607 */
    double cfl_cour, cfl_force, cfl_multiplier_stiffness;
    int32_t mixed_precision;      /* 0: FP64 step; 1: FP32 pair sums (Kahan) on FP64 r_ab */
    int32_t av_kind;              /* 0: AVConfig::Constant, 1: AVConfig::VaryingCD10 */
    double alpha_min, alpha_max, sigma_decay; /* VaryingCD10 (AVConfig.hpp); alpha_u, beta_AV
                                                 above apply to both kinds */
    double constant_G;     /* gravitational constant of the sink forces (SolverConfig.hpp:469-477:
                              6.6743015e-11 with no unit system, 1 in code units with G = 1) */
    double sink_softening; /* Plummer softening of the sink-gas force; 0 = the reference's
                              unsoftened G M d / |d|^3 (SinkParticlesUpdate.cpp:127-141) */
    int32_t keep_storage_order; /* 0: at the start of each step the solver stores the local
                              particles in the order of the last tree's Morton sort (the
                              reference's ParticleReordering, Solver.cpp:2043-2048, done every
                              step); every field and result keeps the caller's particle order
                              at the interface.  1: keep the caller's order in storage too
                              (required by the adjoint entries) */
    int32_t adjoint_fp32;  /* no reference counterpart: 0 = the adjoint entries compute in FP64;
                              1 = their pair terms in FP32 with Kahan-compensated sums, on FP64
                              r_ab and FP64 neighbour criteria (the forward step is unchanged) */
} shamx_config;

/* Per-step log (reference TimestepLog Solver.hpp:43-54 + per-phase timers).  Times in s. */
typedef struct shamx_step_log {
    double rate;       /* particle-updates/s = N / t_total (Solver.cpp:3199-3200) */
    double t_total, t_ghost, t_tree, t_sort, t_neigh, t_hiter, t_omega, t_eos, t_force, t_corr;
    uint32_t n_local, n_merged, n_leaf, subcycles, h_iterations, corrector_iterations;
    uint32_t n_sorts, n_force_evals;
    uint64_t n_pairs;       /* neighbour-cache entries */
    uint64_t n_interacting; /* pairs inside max(h_a, h_b) R (last force evaluation) */
    double energy, eps_v, next_dt, time, dt;
    double t_reorder;       /* the storage reorder at the start of the step */
} shamx_step_log;

typedef struct shamx_ctx shamx_ctx;

const char *shamx_last_error(void);
int shamx_get_device_count(int *count);

/* ---- context ----------------------------------------------------------------------------- */
int shamx_ctx_create(shamx_ctx **out);
int shamx_ctx_destroy(shamx_ctx *ctx);

/* ---- single phases on caller-owned device arrays ----------------------------------------
 * bbox6 = {xmin, ymin, zmin, xmax, ymax, zmax} (device) */

/* replaces PatchDataField::compute_min/max + nextafter widening (BuildTrees.cpp:38-55) */
int shamx_bbox(shamx_ctx *ctx, void *stream, uint32_t n, const double *x, const double *y,
               const double *z, double *bbox6);
/* replaces morton_code_set_from_positions<u32,f64_3,3> (MortonCodeSet.cpp:62-118) without the
 * 0xFFFFFFFF padding (the radix sort needs no power-of-two length) */
int shamx_morton(shamx_ctx *ctx, void *stream, uint32_t n, const double *x, const double *y,
                 const double *z, const double *bbox6, uint32_t *codes);
/* replaces sort_by_key (sort_by_keys.hpp:158-167 -> bitonicSort_updated_usm.cpp:284-396):
 * stable LSD radix sort; keys sorted in place, perm[i] = original index of sorted key i */
int shamx_sort(shamx_ctx *ctx, void *stream, uint32_t n, uint32_t key_bits, uint32_t *keys,
               uint32_t *perm);
/* No counterpart in the reference: shamx_sort for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and stream semantics as shamx_sort, without a context: the alternate key
 * and value buffers and the sort's scratch live in one static device array of the library
 * (about 40 MB per GPU, held until the process ends), so n is at most 4194304 keys
 * (SHAMX_E_INVALID_ARG beyond).  The call allocates nothing.  Every call on one GPU uses that
 * GPU's array: calls that overlap on one GPU (two streams, two threads) race. */
int shamx_sort_noalloc(void *stream, uint32_t n, uint32_t key_bits, uint32_t *keys,
                       uint32_t *perm);
/* replaces CompressedLeafBVH::rebuild_from_positions (CompressedLeafBVH.cpp:78-95) +
 * compute_presteps_rint (Solver.cpp:1322-1356); tree kept in ctx */
int shamx_tree_build(shamx_ctx *ctx, void *stream, uint32_t n, const double *x, const double *y,
                     const double *z, const double *h, uint32_t reduction_level, double htol);
/* the first half of shamx_tree_build: replaces
 * CompressedLeafBVH::internal_rebuild_from_positions_no_aabb (CompressedLeafBVH.cpp:31-76):
 * bbox, Morton codes, sort, leaf reduction (reduction_alg.cpp:413-456) and the Karras tree
 * (KarrasRadixTree.cpp:152-180); the cell boxes are left unset */
int shamx_tree_reduce_karras(shamx_ctx *ctx, void *stream, uint32_t n, const double *x,
                             const double *y, const double *z, uint32_t reduction_level);
/* the second half: replaces compute_tree_aabb_from_positions (KarrasRadixTreeAABB.cpp:92-140)
 * + compute_presteps_rint (Solver.cpp:1322-1356, KarrasRadixTreeField.hpp:188-222) on the
 * tree in ctx: leaf and cell boxes, rint = max h * htol per cell */
int shamx_tree_aabb_hmax(shamx_ctx *ctx, void *stream, const double *x, const double *y,
                         const double *z, const double *h, double htol);
/* sizes of the tree in ctx: out4 = {n_obj, leaf_count, internal_count, 0} */
int shamx_tree_sizes(shamx_ctx *ctx, uint32_t *out4);
/* copies of the tree arrays (device pointers; any may be NULL):
 *  sorted[n], sort_index[n], reduc_map[L+2], lchild/rchild/endrange[I], lflag/rflag[I],
 *  aabb_min/aabb_max[3*(I+L)] (xyz interleaved), rint[I+L] (max h * htol) */
int shamx_tree_get(shamx_ctx *ctx, void *stream, uint32_t *sorted, uint32_t *sort_index,
                   uint32_t *reduc_map, uint32_t *lchild, uint32_t *rchild, uint8_t *lflag,
                   uint8_t *rflag, uint32_t *endrange, double *aabb_min, double *aabb_max,
                   double *rint);
/* replaces NeighbourCache::start_neighbors_cache_2stages (NeighbourCache.cpp:222-604) on the
 * tree in ctx: neighbours of the first n_local particles among all tree objects */
int shamx_neigh_build(shamx_ctx *ctx, void *stream, uint32_t n_local, const double *x,
                      const double *y, const double *z, const double *h, double htol,
                      double Rkern);
/* out4 = {n_local, n_pairs (low 32), n_pairs (high 32), n_leaf_pairs} */
int shamx_neigh_sizes(shamx_ctx *ctx, uint32_t *out4);
/* CSR copies (device): cnt[n], offset[n], ids[P], owner[n], leaf_cnt[L], leaf_offset[L],
 * leaf_ids[PL]; any may be NULL */
int shamx_neigh_get(shamx_ctx *ctx, void *stream, uint32_t *cnt, uint32_t *offset, uint32_t *ids,
                    uint32_t *owner, uint32_t *leaf_cnt, uint32_t *leaf_offset,
                    uint32_t *leaf_ids);
/* replaces IterateSmoothingLengthDensity::_impl_evaluate_internal
 * (IterateSmoothingLengthDensity.cpp:28-120): one Newton sweep on an explicit CSR neighbour
 * list (cnt/offset/ids, device).  eps_minmax_host receives {max eps, min eps}. */
int shamx_h_iterate(void *stream, int kernel, uint32_t n, const uint32_t *cnt,
                    const uint32_t *offset, const uint32_t *ids, const double *x,
                    const double *y, const double *z, const double *h_old, double *h_new,
                    double *eps, double gpart_mass, double h_evol_max, double h_evol_iter_max,
                    int mixed, double *eps_minmax_host);
/* replaces NodeComputeOmega (ComputeOmega.cpp:25-74) */
int shamx_omega(void *stream, int kernel, uint32_t n, const uint32_t *cnt, const uint32_t *offset,
                const uint32_t *ids, const double *x, const double *y, const double *z,
                const double *h, double *omega, double gpart_mass);
/* replaces ComputeEos::compute_eos (ComputeEos.cpp:1137-1311) for the adiabatic, isothermal
 * and LP07 variants, on nm merged particles */
int shamx_eos(void *stream, int kernel, uint32_t nm, int eos_kind, double gamma, double cs,
              double cs0, double q, double r0, const double *h, const double *u, const double *x,
              const double *y, const double *z, double *P, double *soundspeed, double gpart_mass);
/* replaces UpdateDerivs::update_derivs_constantAV (UpdateDerivs.cpp:88-288) fused with
 * prepare_corrector (Solver.cpp:1667-1710), vsig (Solver.cpp:2716-2789) and the CFL
 * (ComputeCFLCourant.hpp, ComputeCFLForce.hpp:50-62).  Fields x..cs are merged arrays
 * (local particles first); a/du are overwritten, their previous values saved to *_old.
 * dt_min_host receives min over particles of min(C_cour h/vsig, C_force sqrt(h/|a|)). */
int shamx_force(void *stream, int kernel, uint32_t n, const uint32_t *cnt, const uint32_t *offset,
                const uint32_t *ids, const double *x, const double *y, const double *z,
                const double *vx, const double *vy, const double *vz, const double *h,
                const double *omega, const double *u, const double *P, const double *cs,
                double *ax, double *ay, double *az, double *du, double *ax_old, double *ay_old,
                double *az_old, double *du_old, double *vsig, double *dt_cfl, double gpart_mass,
                double alpha_u, double alpha_AV, double beta_AV, double C_cour, double C_force,
                int mixed, double *dt_min_host);
/* replaces ComputeCFLCourant (ComputeCFLCourant.hpp:25-70) + ComputeCFLForce
 * (ComputeCFLForce.hpp:50-62) and the next-dt reduction (Solver.cpp:3095-3124):
 * dt_cfl[i] = min(C_cour h/vsig, C_force sqrt(h/|a|)) (device), dt_min_host = min over i */
int shamx_cfl(void *stream, uint32_t n, const double *h, const double *vsig, const double *ax,
              const double *ay, const double *az, double C_cour, double C_force, double *dt_cfl,
              double *dt_min_host);
/* replaces the leapfrog predictor (ForwardEuler nodes, Solver.cpp:390-524) and
 * apply_position_boundary periodic wrap (integrators.cpp:206-236); box6 is a HOST array */
int shamx_predictor(void *stream, uint32_t n, double dt, double *x, double *y, double *z,
                    double *vx, double *vy, double *vz, const double *ax, const double *ay,
                    const double *az, double *u, const double *du, int periodic,
                    const double *box6_host);
/* replaces fields_leapfrog_corrector (SchedulerUtility.hpp:80-99, integrators.cpp:87-119);
 * out3_host = {max |dv|^2, sum |v|^2, sum m (u + v^2/2)} */
int shamx_corrector(void *stream, uint32_t n, double hdt, double *vx, double *vy, double *vz,
                    const double *ax, const double *ay, const double *az, const double *ax_old,
                    const double *ay_old, const double *az_old, double *u, const double *du,
                    const double *du_old, double gpart_mass, double *out3_host);

/* ---- composite step: shammodels::sph::Solver<f64_3,Kern> (Solver.hpp:70-434) -------------
 * The solver owns SoA device fields (local particles followed by their periodic images). */
typedef struct shamx_solver shamx_solver;

int shamx_solver_create(const shamx_config *cfg, uint32_t n, shamx_solver **out);
int shamx_solver_destroy(shamx_solver *s);
/* The solver runs on its own stream.  Entries that take device pointers (set/get_field with
 * on_device, the adjoint entries) first wait for the whole device, so their inputs may come
 * from work on any stream; they return once the solver's stream has finished with them.
 * field ids: 0 x 1 y 2 z 3 vx 4 vy 5 vz 6 ax 7 ay 8 az 9 hpart 10 uint 11 duint
 * 12 alpha_AV 13 soundspeed 14 divv 15 curlv.x 16 curlv.y 17 curlv.z 18 dtdivv (varying AV)
 * 19 axyz_ext.x 20 axyz_ext.y 21 axyz_ext.z (SolverConfig::set_layout, SolverConfig.cpp:24-121);
 * `on_device` selects a device or host source/destination of n doubles */
int shamx_solver_set_field(shamx_solver *s, int field, const double *src, int on_device);
int shamx_solver_get_field(shamx_solver *s, int field, double *dst, int on_device);
/* Storage order of the local particles, for restart dumps (the reference's dump holds its
 * patch data in storage order, Model.hpp:959-991, so a continued run sums in the same order).
 * get: host array of n caller particle ids, the order the next step runs in (a pending
 * reorder included); set: store the particles in that order (a permutation of 0..n-1; only
 * the identity when keep_storage_order = 1).  Field values per particle do not change. */
int shamx_solver_get_storage_order(shamx_solver *s, uint32_t *order);
int shamx_solver_set_storage_order(shamx_solver *s, const uint32_t *order);
/* time state edges (Solver.hpp:92-148): tdc = {t, dt, cfl_multiplier} */
int shamx_solver_get_time(shamx_solver *s, double *tdc3);
int shamx_solver_set_time(shamx_solver *s, const double *tdc3);
/* sinks: Model::add_sink (Model.hpp:183-200, sink_edges_helper.hpp:234-248) and the records of
 * get_sinks (pySPHModel.cpp:1047-1065), 17 doubles per sink: pos[3], velocity[3],
 * sph_acceleration[3], ext_acceleration[3], mass, angular_momentum[3], accretion_radius.
 * Sink gravity (SinkParticlesUpdate.cpp:23-200, ExternalForces.cpp:48-360) runs in every step
 * once a sink exists; accretion (SinkParticlesFlagAccreteHard & co.) is not part of this
 * library.  set_sinks replaces the whole set (restart from saved records). */
int shamx_solver_add_sink(shamx_solver *s, double mass, const double *pos3, const double *vel3,
                          double accretion_radius);
int shamx_solver_get_sinks(shamx_solver *s, uint32_t *count, double *out17);
int shamx_solver_set_sinks(shamx_solver *s, uint32_t count, const double *in17);
/* Solver::evolve_once (Solver.cpp:1942-3272) */
int shamx_solver_evolve_once(shamx_solver *s, shamx_step_log *log);
/* One density and force pass of the step on the current state, without time integration:
 * ghosts, tree, neighbour cache and ONE smoothing-length subcycle (Solver.cpp:1060-1304; its
 * unconverged particles keep h clamped to the cache radius rather than failing), Omega, EOS,
 * the force loop and the CFL reduction (Solver.cpp:2247-2400, UpdateDerivs.cpp:88-288,
 * Solver.cpp:3095-3124).  ax/ay/az/du are overwritten (previous values kept as the *_old of a
 * step), h takes the subcycle's values; x, v, u and the time are untouched.  log->next_dt is
 * the CFL step.  Constant viscosity only (av_kind 0).  The reference runs these phases only
 * inside evolve_once; this entry exposes them for the clustered-tree timing workload, whose
 * h iteration does not converge in a full step. */
int shamx_solver_density_force(shamx_solver *s, shamx_step_log *log);
/* stream used by the solver (hipStream_t as void*) */
int shamx_solver_stream(shamx_solver *s, void **stream);

/* ---- adjoint (no reference counterpart: the reference has no adjoint) ----------------------
 * Vector-Jacobian product of the last evolve_once (FP64, constant AV, adiabatic or isothermal
 * EOS, or LP07 with free boundaries, M4 / M6 / M8 / C2 kernels, one corrector pass, dt held
 * fixed): seeds11 holds 11 device arrays of N doubles, the adjoints of (x, y, z, vx, vy, vz,
 * u, ax, ay, az, du) at the end of the step, replaced in place by those at its start.  The
 * smoothing length is differentiated through its density equation (implicit function
 * theorem).  With sinks (free boundaries), the sink-gas and sink-sink forces and the sinks'
 * kicks and drift are reversed too; the sinks' own seeds (position, velocity, acceleration
 * from the gas) are held by the solver from step to step of a backward pass, which
 * shamx_solver_adjoint_density starts (it sets them to zero): the gradients of the particle
 * seeds then include the paths through the sinks. */
int shamx_solver_adjoint_step(shamx_solver *s, double *const *seeds11);
/* No counterpart in the reference (the adjoint's product store): products_save keeps what the
 * adjoint of the last step reads (tree order, images, neighbour rows, sorted positions, the
 * positions, h and Omega at the end of the step, the velocities and u every corrector pass
 * started from, the step log and the sinks) in slot `slot`; products_load puts them back, so
 * that shamx_solver_adjoint_step can run for that step without running it again.  After a load
 * the solver's fields hold the saved step's end positions: set the fields (a restore) before
 * the next evolve_once.  products_free releases every slot.  keep_storage_order = 1 only. */
int shamx_solver_products_save(shamx_solver *s, uint32_t slot);
int shamx_solver_products_load(shamx_solver *s, uint32_t slot);
int shamx_solver_products_free(shamx_solver *s);
/* adds to xbar, ybar, zbar (device, N) the position gradient of sum_a rho_bar_a rho_a at the
 * current state, rho_a = m (hfact / h_a)^3 with h_a from the density equation; starts a
 * backward pass (the sinks' seeds are set to zero) */
int shamx_solver_adjoint_density(shamx_solver *s, const double *rho_bar, double *xbar,
                                 double *ybar, double *zbar);
/* transient products of the last step (copies to HOST arrays sized by shamx_solver_sizes):
 * f64 ids: 0..10 merged {x,y,z,h,vx,vy,vz,u,omega,P,cs} (n_merged), 11 omega (n), 12 vsig (n),
 *          13 tree bbox (6), 14 rint (I+L), 15 aabb_min (3(I+L)), 16 aabb_max (3(I+L))
 * u32 ids: 0 ghost_src (G), 1 sorted codes (Nm), 2 sort_index (Nm), 3 reduc_map (L+2),
 *          4 owner (n), 5 cnt (n), 6 offset (n), 7 ids (P), 8 leaf_cnt (L), 9 leaf_offset (L),
 *          10 leaf_ids (PL), 11 lchild (I), 12 rchild (I), 13 endrange (I), 14 codes (Nm)
 * u8 ids: 0 lflag (I), 1 rflag (I), 2 the force records of the last force evaluation, in
 *          sorted order (n_merged records of 144 B, FP64, or 64 B, mixed precision) */
int shamx_solver_sizes(shamx_solver *s, uint64_t *out6); /* {Nm, G, L, I, P, PL} */
int shamx_solver_get_f64(shamx_solver *s, int what, double *dst);
int shamx_solver_get_u32(shamx_solver *s, int what, uint32_t *dst);
int shamx_solver_get_u8(shamx_solver *s, int what, uint8_t *dst);

#ifdef __cplusplus
}
#endif
#endif /* SHAMX_SPH_H */
