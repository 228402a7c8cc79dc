// SPDX-License-Identifier: CeCILL-2.1
//
// launch.hpp -- host-side launchers of the HIP kernels (implemented in prims.hip, tree.hip,
// neigh.hip, sph.hip).  All launches are asynchronous on the given stream; functions that must
// return a count to the host say so and synchronise the stream.
#pragma once

#include <functional>

#include "common.hpp"
#include "devbuf.hpp"

namespace shamx {

    // ------------------------------------------------------------------ prims.hip
    void fill_u32(hipStream_t s, u32 *p, u32 v, size_t n);
    void fill_u64(hipStream_t s, u64 *p, u64 v, size_t n);
    void fill_f64(hipStream_t s, f64 *p, f64 v, size_t n);
    // exclusive prefix sum of n u32 values; writes the grand total to *total (device)
    size_t scan_scratch_bytes(u32 n);
    void exclusive_scan_u32(hipStream_t s, const u32 *in, u32 *out, u32 n, u32 *total,
                            void *scratch);
    // the same over the first min(n_max, *n_dev) values, a length known on the device only
    // (no synchronisation; scratch sized for n_max)
    void exclusive_scan_u32_devn(hipStream_t s, const u32 *in, u32 *out, u32 n_max,
                                 const u32 *n_dev, u32 *total, void *scratch);

    // ------------------------------------------------------------------ tree.hip
    // bounding box of n positions widened outward by one ulp (BuildTrees.cpp:38-55);
    // bbox6 = {xmin, ymin, zmin, xmax, ymax, zmax} in device memory
    size_t bbox_scratch_bytes(u32 n);
    void bbox_nextafter(hipStream_t s, u32 n, const f64 *x, const f64 *y, const f64 *z,
                        f64 *bbox6, void *scratch);
    // 30-bit Morton codes (MortonCodeSet.cpp:87-99)
    void morton_codes(hipStream_t s, u32 n, const f64 *x, const f64 *y, const f64 *z,
                      const f64 *bbox6, u32 *codes);
    // stable LSD radix sort of (key, index) pairs; sorts keys[0..n) and returns the permutation
    // in vals (vals are generated as iota in the first pass unless vals_given: then the
    // values travel with their keys).  key_bits <= 32.
    size_t radix_sort_scratch_bytes(u32 n);
    void radix_sort_pairs(hipStream_t s, u32 n, u32 key_bits, u32 *keys, u32 *vals,
                          u32 *keys_alt, u32 *vals_alt, void *scratch, bool vals_given = false);
    // the same sort on buffers held in a static device array (no allocation), n at most
    // NOALLOC_SORT_MAX_KEYS; calls that overlap on one GPU share that array and race
    constexpr u32 NOALLOC_SORT_MAX_KEYS = 1u << 22;
    void radix_sort_noalloc(hipStream_t s, u32 n, u32 key_bits, u32 *keys, u32 *perm);
    size_t radix_sort_noalloc_bytes();

    // Tree state of a compressed-leaf BVH built on the merged particle set
    struct TreeBuffers {
        u32 n_obj = 0, L = 0, I = 0;
        DevArray<f64> bbox6;           // 6
        DevArray<u32> codes;           // n_obj (unsorted Morton codes)
        DevArray<u32> sorted;          // n_obj sorted codes
        DevArray<u32> sort_index;      // n_obj permutation (map_morton_id_to_obj_id)
        DevArray<u32> keys_alt, vals_alt;
        DevArray<u32> split_a, split_b;
        DevArray<u32> red_cnt, red_pos; // leaf reduction: flags before i, flagged indices
        DevArray<u32> big_leaves;       // leaves too large for the lane-group box kernel
        DevArray<u32> reduc_map;       // L + 2
        DevArray<u32> leaf_codes;      // L
        DevArray<u32> lchild, rchild, endrange; // I
        DevArray<u8> lflag, rflag;     // I
        DevArray<u32> parent;          // I + L
        DevArray<u16> depth;           // I (root 0)
        DevArray<f64> nodes;           // (I + L) * NODE_F64 records (tree_layout.hpp)
        DevArray<u32> counters;        // small device scalars
        PinnedWords<u32, 8> host;      // their host copies
        bool err_pending = false;      // host[5] holds the last box pass's flags (tree_check)
        DevArray<unsigned char> scratch;
        // order runs of equal codes by the position inside the cell (set from the previous
        // build's distinct-code count when allow_fine_ties; the phase entries keep the plain
        // stable order)
        bool fine_ties = false, allow_fine_ties = false;
    };
    // full CompressedLeafBVH::rebuild_from_positions + rint field (BuildTrees.cpp:27-65,
    // CompressedLeafBVH.cpp:31-95, Solver.cpp:1322-1356).  Synchronises once to read L.
    void tree_build(hipStream_t s, TreeBuffers &t, u32 n, const f64 *x, const f64 *y,
                    const f64 *z, const f64 *h, u32 reduction_level, f64 htol);
    // the same build in three stages (for per-stage timing): bbox + Morton codes, radix sort,
    // then reduction + Karras + boxes (synchronises once to read the leaf count)
    void tree_stage_codes(hipStream_t s, TreeBuffers &t, u32 n, const f64 *x, const f64 *y,
                          const f64 *z);
    void tree_stage_sort(hipStream_t s, TreeBuffers &t, const f64 *x, const f64 *y,
                         const f64 *z);
    void tree_stage_finish(hipStream_t s, TreeBuffers &t, const f64 *x, const f64 *y,
                           const f64 *z, const f64 *h, u32 reduction_level, f64 htol);
    // the two halves of tree_stage_finish: leaf reduction + Karras tree (synchronises once for
    // the leaf count), then leaf boxes, max h * htol and the internal cells
    void tree_stage_karras(hipStream_t s, TreeBuffers &t, u32 reduction_level);
    void tree_stage_aabb(hipStream_t s, TreeBuffers &t, const f64 *x, const f64 *y,
                         const f64 *z, const f64 *h, f64 htol);
    // throws if the last box pass flagged the tree (broken parent links, a child index out of
    // range, a cell deeper than the merge levels); valid once the stream has been synchronised
    // after that pass (the neighbour build and the next tree build do it)
    void tree_check(TreeBuffers &t);

    // ------------------------------------------------------------------ neigh.hip
    struct NeighBuffers {
        u32 n_local = 0;
        u64 n_pairs = 0, n_leaf_pairs = 0;
        bool sorted_layout = false;
        DevArray<u32> leaf_cnt, leaf_offset, leaf_ids; // stage 1 (per leaf)
        DevArray<u32> owner;                            // per local particle
        // stage 2 CSR.  Reference layout: one row per local particle id, entries are particle
        // ids.  Sorted layout: one row per sorted position (ghost rows empty), entries are
        // sorted positions, so a row's candidates are runs of adjacent sorted records.
        DevArray<u32> cnt, offset, ids;
        DevArray<u32> lpos;                             // sorted layout: local rows in order
        DevArray<u32> fcnt, fids;                       // force-interacting subsequences
        DevArray<u32> lflag, lscan, psum;
        // sorted layout: one pass per stage into rows of fixed capacity; leaf l's list at
        // leaf_ids[l * leaf_stride], row r at ids[r * row_stride] (0: packed layout).  The
        // capacities persist from build to build and grow on overflow.
        u32 cap_leaf = 0, cap_part = 0, leaf_stride = 0, row_stride = 0;
        DevArray<double4> ps;                           // (x, y, z, h) in sorted order
        DevArray<u32> inv;                              // sorted position of each object
        DevArray<u32> mism;                             // owner leaf != own leaf
        DevArray<u32> nbatch, bstart, item_leaf;        // stage-2 wave items (member batches)
        DevArray<f64> bbox; // member batch boxes (batch-level cull on clustered sets)
        // rows past their fixed capacity: flags, counts, positions, list (overflow region)
        DevArray<u32> oflag, ocnt, ofpos, opos, olist;
        bool leaf_overflowed = false; // the last build had leaf lists past their capacity
        u32 n_mismatch = 0;
        DevArray<u32> counters;
        PinnedWords<u32, 16> host; // host copies of counters
        DevArray<u32> wide_cnt;    // per-thread list counts of the overflow walk
        DevArray<unsigned char> scratch;
    };
    // NeighbourCache::start_neighbors_cache_2stages (NeighbourCache.cpp:222-604).
    // Synchronises twice to size the leaf list and the particle list.
    void neigh_build(hipStream_t s, NeighBuffers &nb, TreeBuffers &t, u32 n_local,
                     const f64 *x, const f64 *y, const f64 *z, const f64 *h, f64 htol,
                     f64 Rkern, bool sorted_layout);

    // ------------------------------------------------------------------ sph.hip
    // Neighbour CSR as the pair loops read it.  Particle j of a launch (j < n) reads row
    // sel[j] and writes particle outi[row]; null sel / outi mean the identity (reference
    // layout).  In the sorted layout row == sorted position and ids are sorted positions.
    struct Csr {
        const u32 *cnt, *offset, *ids;
        const u32 *sel  = nullptr;
        const u32 *outi = nullptr;
    };
    // reductions of per-block partials (one u64 per block and slot) by k_finalize
    enum RedOp : int { RED_MAX_KEY = 0, RED_MIN_KEY = 1, RED_SUM_U64 = 2, RED_SUM_F64 = 3 };
    struct RedSpec {
        int nslot;
        int op[4];
        int out_slot[4];
    };
    // IterateSmoothingLengthDensity (one Newton sweep); eps_minmax (device, 2 ordered keys)
    // receives max and min of eps over all n particles.  Neighbour positions come from the
    // sorted records ps (row-indexed) when given, else from x, y, z.
    void h_iterate(hipStream_t s, int kernel, u32 n, Csr c, const double4 *ps, const f64 *x,
                   const f64 *y, const f64 *z, const f64 *h_old, f64 *h_new, f64 *eps, f64 mass,
                   f64 h_evol_max, f64 h_evol_iter_max, u64 *eps_minmax, int mixed,
                   DevArray<u64> &part);
    // The Newton loop of one subcycle on the device (LoopSmoothingLengthIter.cpp:28-80): up to
    // max_iter sweeps over the rows c (sel / outi select and map them), stopping after the
    // first sweep whose max eps over all rows is below epsilon_h.  Sweeps are queued several
    // at a time with one host read per batch (a queued sweep after the stop does nothing), and
    // once few particles still iterate (eps > 1e-6) only their rows are swept.  The h values,
    // eps values and sweep count equal those of one sweep per host read.
    struct HLoopBuffers {
        DevArray<u64> st, part;
        PinnedWords<u64, 16> host; // host copy of st
        DevArray<u32> sel_a, sel_b, flag, off;
        DevArray<unsigned char> scan;
    };
    struct HLoopResult {
        u32 iterations = 0, host_reads = 0, compactions = 0;
        f64 max_eps = 0, min_eps = 0;
    };
    HLoopResult h_iterate_loop(hipStream_t s, int kernel, u32 n, Csr c, const double4 *ps,
                               const f64 *h_old, f64 *h_new, f64 *eps, f64 mass,
                               f64 h_evol_max, f64 h_evol_iter_max, f64 epsilon_h,
                               u32 max_iter, int mixed, HLoopBuffers &b);
    // ComputeOmega; with records ps (row-indexed, .w = final h) and non-null fids, also writes
    // the force-interacting subsequence of every row to fids (same offsets) and fcnt
    // (mixed: the dW/dh sum in FP32 with Kahan compensation, with the filter only)
    void compute_omega(hipStream_t s, int kernel, u32 n, Csr c, const double4 *ps, const f64 *x,
                       const f64 *y, const f64 *z, const f64 *h, f64 *omega, f64 mass,
                       u32 *fcnt, u32 *fids, int mixed = 0);
    // ps[s].w = h[sort_index[s]] for s < nm
    void ps_refresh_h(hipStream_t s, u32 nm, const u32 *sort_index, const f64 *h, double4 *ps);
    struct EosArgs {
        int kind;
        f64 gamma, cs, cs0, q, r0;
    };
    void compute_eos(hipStream_t s, int kernel, u32 nm, const f64 *h, const f64 *u,
                     const f64 *x, const f64 *y, const f64 *z, f64 *P, f64 *cs, f64 mass,
                     EosArgs e);
    struct ForceArgs {
        f64 mass, alpha_u, alpha_AV, beta_AV;
        f64 C_cour, C_force; // CFL factors (cfl * multiplier)
        int mixed;           // 1: FP32 pair sums with Kahan compensation
        int varying = 0;     // 1: per-particle alpha from the FP64 records (varying AV)
        // external acceleration added to the hydro force (axyz += axyz_ext,
        // ExternalForces.cpp:324-360) before the CFL criterion; null when there is none
        const f64 *ext_x = nullptr, *ext_y = nullptr, *ext_z = nullptr;
    };
    // bytes of one force record (128 FP64, 64 mixed)
    size_t force_record_bytes(int mixed);
    // EOS on nm merged particles fused with the force-record build; record s describes
    // particle perm[s] (identity when null).  P and cs are written per particle id.
    // refresh: the records of the same particles were built by the last call and only
    // v and u changed since (corrector reruns): only those fields, P and c_s are rewritten
    void pack_force(hipStream_t s, int kernel, u32 nm, const u32 *perm, const f64 *x,
                    const f64 *y, const f64 *z, const f64 *h, const f64 *vx, const f64 *vy,
                    const f64 *vz, const f64 *u, const f64 *omega, f64 *P, f64 *cs, void *rec,
                    f64 mass, EosArgs e, int mixed, bool refresh = false);
    // UpdateDerivs constant AV (+ a_old save, vsig, per-particle CFL dt reduced into the
    // dt_min key, interacting-pair count).  Neighbours come from the records rec (row-indexed)
    // when given, else from the SoA arrays.
    void compute_force(hipStream_t s, int kernel, u32 n, Csr c, const void *rec, const f64 *x,
                       const f64 *y, const f64 *z, const f64 *vx, const f64 *vy, const f64 *vz,
                       const f64 *h, const f64 *omega, const f64 *u, const f64 *P, const f64 *cs,
                       f64 *ax, f64 *ay, f64 *az, f64 *du, f64 *ax_old, f64 *ay_old,
                       f64 *az_old, f64 *du_old, f64 *vsig, f64 *dt_cfl, ForceArgs fa,
                       u64 *dt_min_key, u64 *pair_count, DevArray<u64> &part);
    // leapfrog predictor (Solver.cpp:390-524) fused with the periodic wrap (integrators.cpp:206)
    void predictor(hipStream_t s, u32 n, f64 dt, f64 *x, f64 *y, f64 *z, f64 *vx, f64 *vy,
                   f64 *vz, const f64 *ax, const f64 *ay, const f64 *az, f64 *u, const f64 *du,
                   int periodic, const f64 *box6);
    // leapfrog corrector (integrators.cpp:87-119) + reductions for eps_v and energy:
    // red[0] ordered key of max |dv|^2, red[1] sum |v|^2 (f64 bits), red[2] energy (f64 bits);
    // the sums are reduced in a fixed order (deterministic)
    void corrector(hipStream_t s, u32 n, f64 hdt, f64 *vx, f64 *vy, f64 *vz, const f64 *ax,
                   const f64 *ay, const f64 *az, const f64 *ax_old, const f64 *ay_old,
                   const f64 *az_old, f64 *u, const f64 *du, const f64 *du_old, f64 mass,
                   u64 *red, DevArray<u64> &part);
    // max over n values as an ordered key (device)
    void max_key_f64(hipStream_t s, u32 n, const f64 *v, u64 *key);
    // varying AV (Cullen & Dehnen 2010): div v, curl v, d(div v)/dt over the force rows of the
    // n local particles (sorted layout), the alpha update from the saved sound speed into
    // alpha_new (local), and alpha of every merged particle into the FP64 records.
    // sorted_acc: merged accelerations in sorted order; work6: 5 local scratch arrays.
    void cd10_alpha(hipStream_t s, int kernel, u32 n, u32 nm, Csr c, const void *rec,
                    const u32 *sort_index, const u32 *ghost_src, const f64 *const *sorted_acc,
                    const f64 *cs_saved, const f64 *h, const f64 *alpha, f64 dt, f64 mass,
                    f64 alpha_min, f64 alpha_max, f64 sigma_decay, f64 *const *work6,
                    f64 *alpha_new);
    // one sink's gravity on n particles (SinkParticlesUpdate.cpp:69-150): ext += -G M d/|d|^3
    // (ext starts from 0 when first), zero inside the accretion radius, with an optional
    // Plummer softening (0 = the reference's force); sum_out (device, 3 f64 as u64 words)
    // receives sum_a G m d/|d|^3, the sink's acceleration from the gas.
    // periodic wrap of positions into box6 (integrators.cpp:206-236)
    void wrap_positions(hipStream_t s, u32 n, f64 *x, f64 *y, f64 *z, const f64 *box6);
    struct SinkArgs {
        f64 x, y, z, mass, racc, G, gpart_mass, soft;
    };
    void sink_gas(hipStream_t s, u32 n, const f64 *x, const f64 *y, const f64 *z, SinkArgs sk,
                  int first, f64 *ex, f64 *ey, f64 *ez, u64 *sum_out, DevArray<u64> &part);
    // per-particle CFL dt = min(C_cour h / vsig, C_force sqrt(h / |a|)) and its minimum as an
    // ordered key (ComputeCFLCourant.hpp, ComputeCFLForce.hpp:50-62)
    void cfl_dt(hipStream_t s, u32 n, const f64 *h, const f64 *vsig, const f64 *ax,
                const f64 *ay, const f64 *az, f64 C_cour, f64 C_force, f64 *dt_cfl,
                u64 *dt_min_key, DevArray<u64> &part);
    // periodic image selection (BasicSPHGhosts.cpp:260-579): returns G (synchronises) and
    // fills src ids + appends positions/h of the images after the n local entries (count_host:
    // one pinned word for the count).
    // The interaction range is (hmax * f1) * f2 with hmax the ordered key at *hmax_key (device).
    u32 periodic_ghosts(hipStream_t s, u32 n, f64 *x, f64 *y, f64 *z, f64 *h, const f64 *box6,
                        const u64 *hmax_key, f64 f1, f64 f2, DevArray<u32> &ghost_src,
                        DevArray<u8> &ghost_off_id, DevArray<u32> &work,
                        DevArray<unsigned char> &scratch, u32 cap, u32 *count_host);
    // ghost field refresh: dst[n + g] = src[ghost_src[g]] (+ offset for positions)
    struct FieldPtrs {
        f64 *p[8];
        int n;
    };
    void gather_ghosts(hipStream_t s, u32 n, u32 G, const u32 *ghost_src, FieldPtrs fields);
    // perm[j] = the j-th local index (< n) of the sorted order sort_index[0..nm) (flag, off:
    // nm words of work space, scan_scratch: scan_scratch_bytes(nm))
    void local_sorted_order(hipStream_t s, u32 nm, u32 n, const u32 *sort_index, u32 *perm,
                            u32 *flag, u32 *off, void *scan_scratch);
    // dst[j] = src[perm[j]] (gather) and dst[perm[j]] = src[j] (scatter), j < n
    // entries of perm[0..n) that differ from their index (synchronises to return the count;
    // count_host: one pinned word)
    u32 perm_moved(hipStream_t s, u32 n, const u32 *perm, u32 *count_dev, u32 *count_host);
    // the same count left in *count_dev (no synchronisation)
    void perm_moved_async(hipStream_t s, u32 n, const u32 *perm, u32 *count_dev);
    void perm_gather_f64(hipStream_t s, u32 n, const u32 *perm, const f64 *src, f64 *dst);
    void perm_scatter_f64(hipStream_t s, u32 n, const u32 *perm, const f64 *src, f64 *dst);
    void perm_gather_u32(hipStream_t s, u32 n, const u32 *perm, const u32 *src, u32 *dst);

    // ------------------------------------------------------------------ adjoint.hip
    // inputs of the adjoint of one step, all from the solver after that step (sorted layout)
    struct AdjointArgs {
        u32 n, nm;
        int kernel;
        f64 dt, mass, hfact, gamma, alpha_u, alpha_AV, beta_AV;
        int eos_kind;
        f64 cs_iso;
        f64 lp_q; // LP07 exponent: c_s^2 = c_s0^2 (r / r0)^(-2 q), r the distance to the origin
        const u32 *sort_index, *ghost_src, *lpos, *cnt, *offset, *ids, *fcnt, *fids;
        const void *frec;    // FP64 force records of the step's force evaluation
        const double4 *ps;   // (x, y, z, final h), sorted
        const f64 *h;     // final h, local particles
        const f64 *omega; // final Omega, merged (local particles then their images)
        int fp32;         // pair terms of the adjoint in FP32 (Kahan sums) on FP64 r_ab
    };
    // the sinks of a step for its adjoint (free boundaries): the positions the sink-gas forces
    // used, masses, accretion radii and force constants (host), the seeds of the sinks'
    // accelerations from the gas (host, 3 per sink, in), the gradient w.r.t. the sinks'
    // positions from the sink-gas terms (host, 3 per sink, out), the gas positions (device)
    struct SinkAdjoint {
        u32 count = 0;
        const f64 *pos = nullptr, *mass = nullptr, *racc = nullptr;
        f64 G = 0, soft = 0;
        const f64 *abar_gas = nullptr;
        f64 *pbar = nullptr;
        const f64 *x = nullptr, *y = nullptr, *z = nullptr;
    };
    // seeds (11 device arrays of n: x y z vx vy vz u ax ay az du) at the end of the step are
    // replaced by the seeds at its start; with sinks, their sink-gas terms are reversed too
    // passes: the corrector passes of the step (1 + reruns); records_for_pass(i) makes A.frec
    // hold the FP64 records of pass i (1-based) before its force adjoint runs (when null,
    // A.frec holds those of a single pass)
    void adjoint_step(hipStream_t s, const AdjointArgs &A, f64 *const *seeds,
                      DevArray<f64> &work, const SinkAdjoint *sk = nullptr, u32 passes = 1,
                      const std::function<void(u32)> *records_for_pass = nullptr);
    // adds to (gx, gy, gz) the position gradient of sum_a rho_bar_a rho_h(h_a(x))
    void adjoint_density(hipStream_t s, const AdjointArgs &A, const f64 *rho_bar, f64 *gx,
                         f64 *gy, f64 *gz, DevArray<f64> &work);
    void adjoint_gather_sorted(hipStream_t s, u32 nm, u32 N, const u32 *sort_index,
                               const u32 *ghost_src, int nf, const f64 *const *in,
                               f64 *const *out);

} // namespace shamx
