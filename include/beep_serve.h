/*
 * beep_serve.h -- solve server with continuous batching (C ABI).
 *
 * A request is one scene (rigid mesh instances + charges, a beep_scene that
 * is not yet built) to be solved for its surface potentials and solvation
 * energy: the whole of BEEP::solve + Mesh::calculate_energy
 * (beep.cpp:417-512; mesh.cpp:1373-1403) for that scene.  The reference
 * has no server; its accelerator queue is OpenCL_Handler (opencl_handler.h:
 * 49, one background thread feeding one device queue).  Here a pool of
 * worker slots serves requests as they arrive: each slot owns a device
 * context (beep_hip.h) whose buffers stay resident and are reused by its next
 * request; a request is admitted to a slot when a slot is free and the
 * device bytes it needs fit in the HBM budget; the work of all admitted
 * requests overlaps on the device (one stream pair per slot).  Every kernel
 * is deterministic, so a request's result is bit-identical whether it runs
 * alone or in a mix.
 *
 * All functions return 0 on success and a negative code on error
 * (beep_server_last_error gives the text).
 */
#ifndef BEEP_SERVE_H
#define BEEP_SERVE_H

#include <stddef.h>

#include "beep_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct beep_server beep_server;

typedef struct {
    int device;
    int slots;              /* concurrent requests (worker threads, one context each) */
    double hbm_budget;      /* cap on the bytes the slots' contexts hold together (0: none);
                               admission also needs the device's measured free memory plus
                               what the slots hold, less 1 GiB, to cover the request */
    int solve_mode;         /* BEEP_SOLVE_FP64 / MIXED / FP32 (beep_hip.h) */
    double tol;             /* GMRES tolerance (relative to |b|) */
    int max_iter;
    int max_refine;         /* mixed: refinement steps */
    double inner_tol;       /* mixed: inner FP32 GMRES tolerance (<= 0: tol / 2) */
    int far_mode;           /* BEEP_FAR_* */
    int fmm_order;
    int fmm_leaf;
    int keep_fh;            /* keep f and h of each request for beep_server_take_fh */
    int near_mode;          /* BEEP_NEAR_PRECOMPUTE, BEEP_NEAR_ON_THE_FLY (beep_hip.h), or
                               BEEP_SERVE_NEAR_AUTO: stored coefficients when the request
                               fits the memory the budget and the device allow with them,
                               else on the fly (the reference's OpenCL path for blobs that
                               do not fit, opencl_bem.cpp:346-747) */
    int device_tree;        /* 1: each request's octree and near list are built on the
                               device after admission (beep_hip_build_tree), the host
                               builds the patches only; 0: on the host with the scene;
                               BEEP_SERVE_TREE_AUTO: on the device for requests of at
                               least BEEP_SERVE_TREE_DEVICE_MIN patches */
} beep_server_opts;

#define BEEP_SERVE_TREE_AUTO (-1)
/* below this the host's octree is quicker than the device's fixed costs */
#define BEEP_SERVE_TREE_DEVICE_MIN 65536

#define BEEP_SERVE_NEAR_AUTO (-1)

/* fills the defaults: device 0, 4 slots, no budget cap, mixed, tol 1e-6, 100
 * iterations, 4 refinements, inner tol / 2, far field auto, FMM order 5, leaf
 * 64, f and h not kept, near field auto */
void beep_server_default_opts(beep_server_opts* o);
/* The device bytes the server's admission charges a scene (built here if it
 * is not): with stored near coefficients and on the fly.  The scene stays
 * the caller's. */
int beep_server_estimate(const beep_server_opts* o, beep_scene* scene, double* stored_bytes,
                         double* fly_bytes);

typedef struct {
    long long id;
    int status;             /* 0 = solved, < 0 = failed (see message) */
    int n_patches;
    double energy;          /* solvation energy, kJ/mol (sum over instances) */
    int iterations, matvecs, refinements, converged;
    double rel_residual;
    /* server clock (seconds since beep_server_create) */
    double t_submit;        /* submission */
    double t_admit;         /* slot and memory granted (after the host scene build) */
    double t_done;
    double host_ms;         /* scene build (patches, point sets, octree) */
    double device_ms;       /* upload, precompute, RHS, solve, energy */
    double upload_ms;       /* of which: geometry, tree and far-field plan */
    double precompute_ms;   /*           singular terms and near coefficients */
    double solve_ms;        /*           RHS, GMRES, energy */
    double bytes;           /* device bytes held by the slot after the request */
    double device_free;     /* free device bytes read at admission */
    int slot;
    char message[128];
    int near_on_the_fly;    /* 1 when the near field ran on the fly (near_mode) */
} beep_result;

int beep_server_create(const beep_server_opts* opts, beep_server** out);
/* Takes ownership of the scene (destroyed by the server after the solve);
 * the meshes it refers to must outlive the request and must not change
 * (charges included) while it is pending.  Non-blocking. */
int beep_server_submit(beep_server* s, beep_scene* scene, long long* id);
/* Next finished request: 0 when one was returned, 1 when none finished
 * within timeout_s (<= 0: do not wait). */
int beep_server_poll(beep_server* s, double timeout_s, beep_result* out);
/* surface potential f and its normal derivative h (n_patches each) of a
 * request returned by beep_server_poll (needs keep_fh; once per request) */
int beep_server_take_fh(beep_server* s, long long id, double* f, double* h);
double beep_server_clock(const beep_server* s);
/* requests submitted but not yet returned by beep_server_poll */
int beep_server_pending(const beep_server* s);
const char* beep_server_last_error(const beep_server* s);
/* waits for every submitted request, then stops the workers */
int beep_server_destroy(beep_server* s);

/* ---- pose batches: many rigid poses of one mesh pair ----
 *
 * Two instances of one mesh per pose: instance 0 unrotated with its centre at
 * the origin, instance 1 rotated about its centre and moved (the pose loop of
 * tools/run_pair.py).  The reference solves poses one
 * after another (tools/run_pair.py:34-52 -> beep.cpp:223-290), recomputing
 * every explicit pair each time.  Here the mesh's own near field (its
 * octree's explicit pairs, their coefficients and the singular self terms)
 * is built once: these integrals do not change under a rigid motion, so one
 * coefficient set serves both instances of every pose (beep_hip_share_near).
 * Per pose only the pairs between the instances that are closer than
 * pair_cutoff are integrated explicitly; every other pair of the pose takes
 * the 1-point interaction through the far field.  Poses run concurrently on
 * `slots` device contexts. */
typedef struct beep_poses beep_poses;

typedef struct {
    int device;
    int slots;              /* poses at once (worker threads, one device context each) */
    int solve_mode;         /* BEEP_SOLVE_* (beep_hip.h) */
    double tol;
    int max_iter;
    int max_refine;
    double inner_tol;
    int far_mode;           /* BEEP_FAR_* */
    int fmm_order;
    int fmm_leaf;
    double pair_cutoff;     /* inter-instance explicit pairs closer than this (Angstrom);
                               <= 0: the median reach of the mesh's own explicit pairs */
    int gram_schmidt;       /* Gram-Schmidt passes per Arnoldi step (1 or 2) */
    int keep_fh;            /* keep f and h of each pose for beep_poses_take_fh */
} beep_pose_opts;

/* device 0, 8 slots, mixed, tol 1e-6, 100 iterations, 4 refinements, inner
 * tol / 2, FMM order 5 leaf 64, cutoff from the mesh, 1 pass, f and h not
 * kept */
void beep_pose_default_opts(beep_pose_opts* o);

typedef struct {
    int status;             /* 0 = solved, < 0 = failed (see message) */
    double energy;          /* solvation energy of the pose, kJ/mol (both instances) */
    int iterations, matvecs, refinements, converged;
    double rel_residual;
    long long n_pairs;      /* explicit pairs between the instances */
    double host_ms;         /* pose scene and its inter-instance pairs */
    double device_ms;       /* upload, pair coefficients, far-field plan, RHS, solve */
    int slot;
    char message[128];
} beep_pose_result;

/* The mesh's own near field for the given physics (BEEP(Dsolvent, kappa,
 * quad, qual, nbsize), instance dielectric Dprotein), once: canonical octree,
 * coefficients of the precisions the solve mode needs, singular terms,
 * energy coefficients. */
int beep_poses_create(const beep_pose_opts* opts, const beep_mesh* mesh, double Dsolvent,
                      double kappa, int quad_rule, int qual_rule, int nbsize, double Dprotein,
                      beep_poses** out);
/* Solve n poses (quat4: n x (w, x, y, z) rotation of instance 1 about the
 * mesh centre; offset3: n x 3, where its centre goes; instance 0's centre is
 * at the origin) on the slots; blocks
 * until every pose is done; results: n entries (a failed pose has status <
 * 0, the others are still solved).  Returns 0 unless the call itself fails. */
int beep_poses_solve(beep_poses* p, int n, const double* quat4, const double* offset3,
                     beep_pose_result* results);
/* f and h (2 N each, N = patches of the pose) of pose k of the last
 * beep_poses_solve (keep_fh). */
int beep_poses_take_fh(beep_poses* p, int k, double* f, double* h);
/* Set-up: patches of one instance, its explicit pairs, the cutoff in use,
 * and the milliseconds of the one-time set-up (host and device). */
int beep_poses_info(const beep_poses* p, int* n_patches, long long* intra_pairs, double* cutoff,
                    double* setup_ms);
void beep_poses_destroy(beep_poses* p);

#ifdef __cplusplus
}
#endif

#endif /* BEEP_SERVE_H */
