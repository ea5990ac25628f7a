/*
 * beep_hip.h -- device boundary of the BEM mat-vec (C ABI, HIP on gfx950).
 *
 * Drop-in for the reference's hot path:
 *
 *   long BEEP::do_bem_fmm_iteration(unsigned num_patches, double lhs[], double result[])
 *        src/beep/beep.h:276, src/beep/beep.cpp:1485-1583
 *     -> beep_hip_matvec(ctx, lhs, result, precision)
 *        same layout: lhs = [f(0..N-1); h(0..N-1)], result overwritten with
 *        the same layout (near field + far field + diagonal).
 *
 *   BEEP::precalc_neighbour_interactions   src/beep/beep.cpp:768-956
 *     -> beep_hip_precompute(ctx, kappa, precision)
 *
 *   BEEP::gmres(eps, max_iter)             src/beep/beep.cpp:1262-1483
 *     -> beep_hip_solve(ctx, b, x, opts, stats)
 *
 *   BEEP::set_rhs                          src/beep/beep.cpp:1124-1260
 *     -> beep_hip_rhs(ctx, ...)
 *
 *   Mesh::init_energy_precalcs             src/bem/mesh.cpp:730-887
 *     -> beep_hip_energy_coefficients(ctx, ...)
 *
 *   OpenCL_WorkBlob / OpenCL_Handler       src/opencl/opencl_workblob.h:20-52,
 *                                          src/opencl/opencl_handler.h:49-90
 *     -> replaced by the context's own HIP stream; calls are stream-ordered
 *        and return when their results are in caller memory.
 *
 * Inputs that the reference reads from its host objects arrive as flat arrays
 * (beep_hip_geometry, beep_hip_tree; produced by beep_host.h or by the
 * caller's own Mesh / NodePatch / Octree code).  Host arrays are copied; the
 * device copies belong to the context.
 *
 * Conventions: every function returns 0 on success or a negative code; no
 * exception crosses the ABI; beep_hip_last_error(ctx) gives the message.  A
 * context belongs to one host thread at a time; contexts are independent.
 */
#ifndef BEEP_HIP_H
#define BEEP_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* precision of a mat-vec / precompute */
#define BEEP_FP64 0       /* FP64 integrals, FP64 coefficients, FP64 vectors and sums        */
#define BEEP_FP32 1       /* FP32 integrals and coefficients, FP32 vectors and sums          */
#define BEEP_FP32_RESID 2 /* FP32 coefficients promoted to FP64, FP64 vectors, sums and far
                             field: the refinement residual of the mixed-precision solve   */

/* near-field modes (beep_hip_set_near_mode) */
#define BEEP_NEAR_PRECOMPUTE 0 /* coefficients stored by beep_hip_precompute, applied by an SpMV */
#define BEEP_NEAR_ON_THE_FLY 1 /* every mat-vec integrates the explicit pairs (nothing stored) */

/* solve modes */
#define BEEP_SOLVE_FP64 0  /* GMRES with BEEP_FP64 mat-vecs (plain FP64 solve)            */
#define BEEP_SOLVE_MIXED 1 /* FP64 iterative refinement (BEEP_FP32_RESID residuals) around
                              GMRES with BEEP_FP32 mat-vecs and FP64 Arnoldi/Givens        */
#define BEEP_SOLVE_FP32 2  /* GMRES with BEEP_FP32 mat-vecs only (no refinement)          */

typedef struct beep_hip_ctx beep_hip_ctx;

/* Node patches and their point sets (BasicNodePatch, node_patch.h:46-224;
 * QuadPointT, quad_point.h:82-231).  Points are 7 doubles each:
 * x, y, z, nx, ny, nz, weight.  qual = qualocation points of a row patch
 * (weights sum to 1), quad = quadrature points of a column patch (area
 * weights), gal = 4-point galerkin points used for the singular self term. */
typedef struct {
    int n_patches;
    const double* centroid;  /* 3N */
    const double* normal;    /* 3N, unit */
    const double* area;      /* N  */
    const double* eps_ratio; /* N, Dsolvent / Dprotein of the patch's instance */
    const double* gc;        /* N, geometric correction (0.5) */
    const int* qual_ptr;     /* N + 1 */
    const double* qual_pts;
    const int* quad_ptr;     /* N + 1 */
    const double* quad_pts;
    const int* gal_ptr;      /* N + 1 */
    const double* gal_pts;
} beep_hip_geometry;

/* The same node patches given by their surface: the device generates the
 * qualocation, quadrature and galerkin point sets itself (the sub-triangle
 * points of node_patch.cpp:55-273 / triangle.h:387-479) instead of
 * receiving them.  Patch p is vertex p; for every triangle around it, in the
 * order of vt_tri, the two sub-triangles (centre, vertex, edge midpoint)
 * carry the points of the rule (0 = one centroid point per patch). */
typedef struct {
    int n_patches;
    const double* centroid;  /* 3N */
    const double* normal;    /* 3N, unit */
    const double* area;      /* N */
    const double* eps_ratio; /* N */
    const double* gc;        /* N */
    const double* verts;     /* 3N: vertex positions, universe frame */
    int n_tris;
    const int* tris;         /* 3 n_tris: vertex indices (outward orientation) */
    const int* vt_ptr;       /* N + 1 */
    const int* vt_tri;       /* triangles around each vertex */
    const int* vt_corner;    /* corner 0, 1, 2 of the vertex in each of them */
    int qual_rule;           /* rule order of the qualocation points */
    int quad_rule;           /* rule order of the quadrature points */
    /* curved (PNG1) triangles, as beep_scene_surface_curved gives them;
     * n_curved = 0 with NULL arrays for a planar surface.  A curved
     * triangle's points lie on its surface S(u, v), with normal
     * unit(dS/du x dS/dv) and weight w * |dS/du x dS/dv| / 12 (node_patch.cpp
     * with the curved triangle's jacobian, png1.h:135-346). */
    int n_curved;
    const int* ctl_idx;      /* n_tris: index among the curved triangles, or -1 */
    const double* ctl;       /* 63 n_curved: placed control points */
} beep_hip_surface;

/* Explicit-pair structure from the octree: every leaf's rows interact
 * explicitly with the columns of its 27-box neighbourhood
 * (octree.h:627-648, beep.cpp:859-950). */
typedef struct {
    int n_leaves;
    const int* leaf_ptr; /* n_leaves + 1 */
    const int* leaf_idx; /* N: rows of each leaf */
    const int* nb_ptr;   /* n_leaves + 1 */
    const int* nb_idx;   /* columns of each leaf */
} beep_hip_tree;

typedef struct {
    double tol;        /* stop when |residual| <= tol * |b| (beep.cpp:1421) */
    int max_iter;      /* GMRES inner iterations per cycle (m) */
    int mode;          /* BEEP_SOLVE_* */
    int max_refine;    /* refinement steps (mixed mode) */
    double inner_tol;  /* relative tolerance of each inner solve (mixed mode); <= 0: tol / 2 */
    int verbose;
    int resid_fp64;    /* mixed mode: 0 residual with the FP32 coefficient set (FP64 sums and
                          far field, BEEP_FP32_RESID), 1 with the FP64 set (BEEP_FP64; needs its
                          precompute) */
} beep_solve_opts;

typedef struct {
    int iterations;    /* total inner GMRES iterations */
    int matvecs;       /* total mat-vecs of any precision */
    int refinements;   /* refinement steps taken (mixed mode) */
    int converged;
    double residual;   /* final |b - K x| (FP64 residual operator of the mode) */
    double rel_residual;
    double time_ms;
    double matvec_ms;
} beep_solve_stats;

/* Device time of the phases of a mat-vec.  With overlap on (the default) the
 * far field runs on a second stream while the near SpMV runs on the context
 * stream: near_ms and far_ms are then both measured from the start of the
 * mat-vec and total_ms is the whole mat-vec (max of the two + the final add).
 * With overlap off both run on the context stream, the near part first:
 * far_ms is then measured from the end of the near part. */
typedef struct {
    double near_ms;     /* last mat-vec: near-field SpMV + diagonal */
    double far_ms;      /* last mat-vec: far field */
    double total_ms;
    int count;          /* mat-vecs since the last beep_hip_timing_reset */
    double near_ms_sum; /* summed over those mat-vecs */
    double far_ms_sum;
    double total_ms_sum;
    double m2l_ms;      /* last mat-vec: FMM M2L (k_m2l + k_m2l_sum), 0 for the direct far field */
    double m2l_ms_sum;
} beep_matvec_timing;

const char* beep_hip_last_error(const beep_hip_ctx* ctx);

int beep_hip_create(int device, beep_hip_ctx** out);
int beep_hip_destroy(beep_hip_ctx* ctx);

int beep_hip_set_geometry(beep_hip_ctx* ctx, const beep_hip_geometry* geom);
int beep_hip_set_tree(beep_hip_ctx* ctx, const beep_hip_tree* tree);
/* The octree and its explicit-pair structure built on the device instead of
 * received (beep_hip_set_tree with beep_scene_tree's arrays): the same
 * leaves, leaf contents and neighbourhood contents, in the same order, as the
 * host's octree (BEEP::init_fmm_octree, beep.cpp:1585-1641: relaxed
 * insertion at 4 nbsize, build_neighbourhoods, optimize_neighbourhoods,
 * remove_empty_nodes, octree.h:780-1154; contents octree.h:627-648) over the
 * patch centroids of the geometry set, in the bounding cube (centre, edge)
 * (BEEP::calculate_bounding_cube, beep.cpp:1652-1693; beep_scene_get_sizes
 * gives it).  The device computes every patch's descent key, sorts the keys
 * and gathers the lists; the host decides the tree's shape from the sorted
 * keys' box counts.  Returns the number of leaves and of neighbourhood
 * entries. */
int beep_hip_build_tree(beep_hip_ctx* ctx, const double* centre, double edge, int nbsize,
                        int* n_leaves, long long* n_nb);
/* Copies of the tree set (by beep_hip_set_tree or beep_hip_build_tree):
 * leaf_ptr L + 1, leaf_idx N, nb_ptr L + 1, nb_idx n_nb; any may be NULL. */
int beep_hip_get_tree(beep_hip_ctx* ctx, int* leaf_ptr, int* leaf_idx, int* nb_ptr, int* nb_idx);
/* geometry from the surface (point sets generated on the device); replaces
 * beep_hip_set_geometry, after which the tree is set as usual */
int beep_hip_set_surface(beep_hip_ctx* ctx, const beep_hip_surface* surface);

/* Far-field engine for the 1-point interaction of all pairs (the role of
 * FMM_BEM_Octree::solve, fmm_bem_hybrid.h:184-279):
 *   BEEP_FAR_DIRECT  exact O(N^2) sum,
 *   BEEP_FAR_FMM     Chebyshev-interpolation FMM of the given order (3..6),
 *                    leaf boxes holding about leaf_target patches,
 *   BEEP_FAR_PW      plane-wave (Yukawa / Laplace) FMM of expansion order 9
 *                    (9-node plane-wave quadrature, about 3 digits, the
 *                    reference's <9,9,67>), 12 or 18 (18 nodes, about 6
 *                    digits), leaf boxes as for BEEP_FAR_FMM,
 *   BEEP_FAR_AUTO    FMM above 20,000 patches, direct below.
 * Takes effect immediately (kernel tables are built by the first mat-vec or
 * precompute that needs them); default BEEP_FAR_AUTO, order 4, leaf_target 48.
 * The near coefficients depend on the FMM plan (explicit pairs that its leaf
 * pass also covers keep the full rule, and that pass skips them): after a
 * change call beep_hip_precompute again, or the mat-vec returns an error. */
#define BEEP_FAR_AUTO (-1)
#define BEEP_FAR_DIRECT 0
#define BEEP_FAR_FMM 1
#define BEEP_FAR_PW 2
int beep_hip_set_far_field(beep_hip_ctx* ctx, int mode, int order, int leaf_target);

/* Far field on a second HIP stream, concurrent with the near-field SpMV
 * (on = 1, the default), or both on the context stream (on = 0).  The same
 * kernels run in the same order on each stream, so y is bit-identical.  This
 * replaces the reference's "near field, then FMM" sequence
 * (beep.cpp:1517-1535), whose GPU near field finishes before the CPU FMM
 * starts (beep.cpp:1041-1045). */
int beep_hip_set_overlap(beep_hip_ctx* ctx, int on);

/* GMRES Arnoldi steps (mat-vec, Gram-Schmidt, normalisation, coefficient
 * read-back) as one captured HIP graph replayed for every step (on = 1, the
 * default; re-captured when buffers or plans change), or launched kernel by
 * kernel (on = 0).  The same kernels run either way, so results are
 * bit-identical.  The reference launches nothing (CPU BLAS, beep.cpp:
 * 1391-1428). */
int beep_hip_set_graphs(beep_hip_ctx* ctx, int on);

/* Autotuner for the octree leaf size of the FMM far field: times `reps`
 * mat-vecs of the given precision with each candidate leaf target (mean
 * patches per leaf box) at the current order, writes the mean time per
 * mat-vec of each to ms_out (n values, may be NULL), keeps the fastest and
 * returns it in *best.  Needs the geometry, tree and coefficients of that
 * precision; switches the far field to BEEP_FAR_FMM unless it is BEEP_FAR_PW
 * (which then stays).  Each candidate plan
 * gets its own coefficients (precomputed per candidate), and every
 * coefficient set the context held is precomputed again for the chosen one.
 * No counterpart in the reference, whose leaf size follows from nbsize
 * (beep.h:67-74). */
int beep_hip_autotune_leaf(beep_hip_ctx* ctx, int precision, const int* leaf_targets, int n,
                           int reps, double* ms_out, int* best);

/* Residency cap of the near-field SpMV: at most blocks_per_cu blocks per
 * compute unit walk the work items (default 1; 0: one block per work item).
 * A cap leaves room on every CU for the far-field kernels that run on the
 * other stream at the same time.  Sets the cap of every precision. */
int beep_hip_set_spmv_residency(beep_hip_ctx* ctx, int blocks_per_cu);
/* Autotuner for that cap: times `reps` whole mat-vecs of the given precision
 * (far field overlapped) for each candidate, writes the mean times to ms_out
 * (may be NULL), keeps the fastest for mat-vecs of that precision (BEEP_FP64,
 * BEEP_FP32 and BEEP_FP32_RESID each keep their own) and returns it in
 * *best. */
int beep_hip_autotune_spmv(beep_hip_ctx* ctx, int precision, const int* blocks_per_cu, int n,
                           int reps, double* ms_out, int* best);
/* Coefficient loads each SpMV thread keeps in flight: 4 x depth (FP64
 * coefficients) or 8 x depth (FP32), depth 1 (default), 2 or 4.  Deeper
 * lets fewer resident blocks cover the memory latency (with a residency
 * cap).  y is bit-identical for every depth.  Sets the depth of every
 * precision. */
int beep_hip_set_spmv_depth(beep_hip_ctx* ctx, int depth);
/* Autotuner for the depth, as beep_hip_autotune_spmv (per precision). */
int beep_hip_autotune_spmv_depth(beep_hip_ctx* ctx, int precision, const int* depths, int n,
                                 int reps, double* ms_out, int* best);

/* Shape of the current far-field plan: mode in use (BEEP_FAR_DIRECT,
 * BEEP_FAR_FMM or BEEP_FAR_PW), FMM order, leaf level, leaf boxes, all boxes, M2L box pairs,
 * P2P target-source pairs of the leaf pass, the pairs among them that are not
 * explicit near pairs (the only ones the leaf pass evaluates once the near
 * list is set; -1 before), and the host milliseconds spent building the plan
 * and that residual list (zeros for the direct sum). */
typedef struct {
    int mode;
    int order;
    int leaf_level;
    int n_leaf_boxes;
    int n_boxes;
    long long m2l_pairs;
    long long p2p_pairs;
    long long p2p_residual_pairs;
    double plan_ms;
    double bytes_fp64;   /* algorithmic HBM bytes of one far-field evaluation (every far-field */
    double bytes_fp32;   /* array read once and written once), FP64 and FP32 expansions */
} beep_far_info;
int beep_hip_far_info(beep_hip_ctx* ctx, beep_far_info* out);

/* Singular self terms (FP64 always) and near-field coefficients of the given
 * precision (BEEP_FP64 or BEEP_FP32; BEEP_FP32_RESID uses the FP32 set).
 * With the FMM far field, builds its plan first (the coefficients follow
 * it).  The FP32 set computes its near-singular pairs (centroids closer than
 * about two mesh edges, |c_i - c_j|^2 < 2 (a_i + a_j)) in FP64 and rounds
 * them, unless switched off below.  Replaces precalc_neighbour_interactions
 * (beep.cpp:768-956), whose coefficients are all FP32 (bem_kernels.h:134-251
 * on float QuadPoints). */
int beep_hip_precompute(beep_hip_ctx* ctx, double kappa, int precision);

/* Gram-Schmidt passes of every Arnoldi step: 1 (default) is the reference's
 * classical Gram-Schmidt, two dgemv calls (beep.cpp:1395-1396); 2
 * orthogonalises twice (CGS2), for ill-conditioned systems whose Krylov
 * basis loses orthogonality over many iterations. */
int beep_hip_set_gram_schmidt(beep_hip_ctx* ctx, int passes);

/* Coefficient guard, after BEEP::sanity_check (beep.cpp:1066-1122) and
 * LocalIntegrations::insane (local_integrations.h:111-121): a coefficient is
 * bad when |v| > max_value (default 1e6, the reference's) or not finite (the
 * reference's comparison lets NaN through).  Every precompute checks what it
 * computes: a bad pair is recomputed in FP64 and zeroed if still bad; a bad
 * self term is zeroed (the reference's rule for an ill-conditioned patch). */
int beep_hip_set_guard(beep_hip_ctx* ctx, double max_value);
/* Entries repaired (recomputed or zeroed) by the last precompute of the FP64
 * set [0], the FP32 set [1], the diagonal [2], and by the last
 * beep_hip_sanity_check [3].  Waits for the device. */
int beep_hip_guard_stats(beep_hip_ctx* ctx, long long* repaired4);
/* The guard over the stored coefficients (every set held and the diagonal),
 * repairing bad entries as above; *repaired receives their count.  The
 * reference runs its check once, before the first mat-vec; here it is on
 * demand, for memory that may have changed since the precompute. */
int beep_hip_sanity_check(beep_hip_ctx* ctx, long long* repaired);
/* Test hook: overwrite one stored entry (which 0: FP64 set, 1: FP32 set, by
 * global coefficient index; 2: the diagonal of a patch) with v4 = A, B, C,
 * D, so a test can force the guard's condition from the host. */
int beep_hip_debug_poke(beep_hip_ctx* ctx, int which, long long entry, const double* v4);

/* Near field shared with another context, for the instances of a pose
 * (several rigid copies of one mesh): the coefficient sets, diagonal and tree
 * of src (a plain context precomputed with the direct far field, so every
 * explicit pair stores its full rule less the 1-point interaction) apply to
 * n_copies copies of src's patches in this context, copy k at patch offset
 * copy_offset[k] (n_copies x src's patches = this context's patches; the
 * integrals are invariant under the rigid motion of a copy), plus the
 * explicit pairs (pair_i[q], pair_j[q]) of this context (patches of
 * different copies near each other; rows and columns in any order, no
 * repeats).  Takes the place of beep_hip_set_tree; beep_hip_precompute then
 * computes only the listed pairs' coefficients (in FP64, kept in both
 * precisions), for src's kappa.  The FMM leaf pass then covers every
 * neighbouring-leaf pair.  src must outlive this use and not change
 * meanwhile; contexts sharing one src may run concurrently.  Replaces the
 * per-pose LocalIntegrations of the reference's pose loop (tools/
 * run_pair.py:34-52 -> beep.cpp:223-290, 768-956), which recomputes the
 * intra-instance pairs for every pose. */
int beep_hip_share_near(beep_hip_ctx* ctx, beep_hip_ctx* src, int n_copies, const int* copy_offset,
                        long long n_pairs, const int* pair_i, const int* pair_j);
/* The same, with the explicit pairs between copies found on the device:
 * every pair of patches of different copies whose centroids are closer than
 * cutoff (Angstrom), both orders; *n_pairs (may be NULL) receives their
 * count. */
int beep_hip_share_near_cutoff(beep_hip_ctx* ctx, beep_hip_ctx* src, int n_copies, const int* copy_offset,
                               double cutoff, long long* n_pairs);

/* FP64 integration of the near-singular pairs of the FP32 coefficient set
 * (on = 1, the default) or FP32 for every pair (on = 0); applies from the
 * next FP32 precompute.  No counterpart in the reference (all FP32). */
int beep_hip_set_near_singular_fp64(beep_hip_ctx* ctx, int on);

/* Near-field mode.  BEEP_NEAR_PRECOMPUTE (the default): beep_hip_precompute
 * stores the explicit pairs' coefficients (LocalIntegrations, beep.cpp:
 * 768-956) and each mat-vec applies them.  BEEP_NEAR_ON_THE_FLY: nothing is
 * stored; beep_hip_precompute computes the self terms only, and every
 * mat-vec integrates each explicit pair and applies it at once, as the
 * reference's OpenCL path does (BEM_kernels_instant, opencl_bem_kernels.cl:
 * 237-323, driven by BEM_OnDemand_Resources, opencl_bem.cpp:346-747): for
 * memory-bound use (no coefficient memory; a mat-vec costs about one
 * precompute).  Each pair's values are the ones the precompute would store
 * (guard and near-singular FP64 pass included), so both modes give the same
 * operator to rounding.  Changing the mode drops the precomputed sets: call
 * beep_hip_precompute again.  A context sharing a near field
 * (beep_hip_share_near) and its source stay in BEEP_NEAR_PRECOMPUTE. */
int beep_hip_set_near_mode(beep_hip_ctx* ctx, int mode);
/* Near-singular entries of the current FP32 set computed in FP64 (listed),
 * and the entries the list had no room for (computed in FP32): both 0 when
 * the set was built with the FP64 pass off.  Waits for the precompute. */
int beep_hip_near_singular_count(beep_hip_ctx* ctx, long long* listed, long long* overflow);

/* y = K x on host arrays (do_bem_fmm_iteration); blocks until y is written. */
int beep_hip_matvec(beep_hip_ctx* ctx, const double* x, double* y, int precision);
/* y = K x on device arrays; stream-ordered on the context stream. */
int beep_hip_matvec_device(beep_hip_ctx* ctx, const double* d_x, double* d_y, int precision);
/* As beep_hip_matvec_device, but allocating nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Its scratch (source records, the far
 * vector, partial sums) is a 64 MiB array of the library, one per device,
 * held until the process ends; calls whose scratch would exceed it (with the
 * direct far field, about 100,000 patches) fail with a message, and so do
 * contexts with an FMM far field.  Every call on a device shares that
 * device's array: calls that overlap in time on one device (two contexts or
 * streams at once) race and must be serialised by the caller.  Stream-ordered
 * on the context stream; y is bit-identical to beep_hip_matvec_device's. */
int beep_hip_matvec_device_noalloc(beep_hip_ctx* ctx, const double* d_x, double* d_y,
                                   int precision);
/* per-phase device time of mat-vecs (HIP events on the context stream):
 * the last one and the sums since beep_hip_timing_reset */
int beep_hip_matvec_timing(beep_hip_ctx* ctx, beep_matvec_timing* out);
int beep_hip_timing_reset(beep_hip_ctx* ctx);

/* b = [f_rhs; h_rhs] from point charges xyzq (n x 4, universe frame) */
int beep_hip_rhs(beep_hip_ctx* ctx, int n_charges, const double* xyzq, double Dsolvent,
                 double* b);
/* As beep_hip_rhs, into a device vector of 2N doubles (d_b). */
int beep_hip_rhs_device(beep_hip_ctx* ctx, int n_charges, const double* xyzq, double Dsolvent,
                        double* d_b);
/* raw energy coefficients (before the Gauss-law renormalisation) of every
 * patch from the charges of its own instance */
int beep_hip_energy_coefficients(beep_hip_ctx* ctx, int n_instances, const int* inst_patch_ptr,
                                 const int* inst_charge_ptr, const double* xyzq, double* ec_f,
                                 double* ec_h);

/* Solve K x = b; x holds the initial guess on entry (zeros: cold start). */
int beep_hip_solve(beep_hip_ctx* ctx, const double* b, double* x, const beep_solve_opts* opts,
                   beep_solve_stats* stats);
/* The RHS of point charges (as beep_hip_rhs) and the solve from a zero
 * guess, b kept on the device; x (2N doubles, host) receives the solution.
 * One PCIe transfer of a vector instead of four (BEEP::set_rhs then
 * BEEP::gmres, beep.cpp:1124-1483). */
int beep_hip_solve_from_charges(beep_hip_ctx* ctx, int n_charges, const double* xyzq,
                                double Dsolvent, const beep_solve_opts* opts, double* x,
                                beep_solve_stats* stats);
/* As beep_hip_solve on device vectors of 2N doubles (d_b read, d_x guess in
 * and solution out): no PCIe transfer of b or x. */
int beep_hip_solve_device(beep_hip_ctx* ctx, const double* d_b, double* d_x,
                          const beep_solve_opts* opts, beep_solve_stats* stats);

/* Diagnostic: D = A B through one 16x16x4 MFMA (fp32 = 0: FP64 operands,
 * 1: FP32), A 16x4 and B 4x16 row-major host arrays; d[4 l + r] receives
 * accumulator element r of lane l (256 values).  Checks the operand and
 * result layout the M2L kernels rely on. */
int beep_hip_selftest_mfma(beep_hip_ctx* ctx, int fp32, const double* a, const double* b,
                           double* d);

/* Diagnostic (host only): the M2L symmetry used by the FMM.  For a source
 * box at offset t = (tx, ty, tz) (box units, 2 <= max |t_k| <= 3) returns
 * the index of its canonical offset (components in canon_xyz) and the node
 * map psi (order^3 entries, canonical node -> actual node) with
 * K_t[psi(m)][psi(n)] = K_canonical[m][n]; -1 on bad input (message from
 * beep_hip_last_error(NULL)). */
int beep_hip_fmm_symmetry(int tx, int ty, int tz, int order, int* canon_xyz, int* node_map);

/* Diagnostic (host only): the plane-wave tables of expansion order `order`
 * with the `nodes`-node quadrature (9 or 18) at scaled frequency beta (box
 * units), against direct sums of seeded sources and targets.  err5 receives
 * errors relative to the largest value of each quantity: [0] the kernel's
 * plane-wave quadrature over 1 <= z <= 4, [1] P2M -> M2X -> X2X -> X2L ->
 * L2P potential over the six directions, [2] the same through M2M and L2L
 * one level down, [3] the gradient of [1], [4] the orthogonality defect of
 * the direction rotations.  The tables play the role of the reference's
 * Level_Dependent_FMM_Globals (fmm_globals.h:372-493). */
int beep_hip_pw_selftest(int order, int nodes, double beta, double* err5);

/* device memory helpers for callers that keep vectors resident */
int beep_hip_device_alloc(beep_hip_ctx* ctx, size_t bytes, void** out);
int beep_hip_device_free(beep_hip_ctx* ctx, void* p);
/* kind: 1 = host to device, 2 = device to host, 3 = device to device */
int beep_hip_memcpy(beep_hip_ctx* ctx, void* dst, const void* src, size_t bytes, int kind);
int beep_hip_synchronize(beep_hip_ctx* ctx);
/* Measurement aid, no counterpart in the reference: dst row r = src row
 * idx[r] over rows of 64 bytes (device pointers; idx a permutation or any
 * indices below the rows of src), stream-ordered on the context stream.  A
 * scattered read of known size, to check hardware byte counters on gathers
 * as a plain copy checks them on streams. */
int beep_hip_gather_rows64(beep_hip_ctx* ctx, const void* src, const int* idx, size_t rows, void* dst);
int beep_hip_memory_info(beep_hip_ctx* ctx, size_t* free_bytes, size_t* total_bytes);
/* device bytes held by the context (its buffers only grow: a later geometry
 * of no larger size reuses them) */
int beep_hip_memory_held(beep_hip_ctx* ctx, size_t* bytes);
/* Free and total bytes of a device, without a context (the server's admission
 * reads it; beep_hip_last_error(NULL) gives the error text). */
int beep_hip_device_memory(int device, size_t* free_bytes, size_t* total_bytes);

#ifdef __cplusplus
}
#endif

#endif /* BEEP_HIP_H */
