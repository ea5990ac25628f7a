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
} beep_server_opts;

/* fills the defaults: device 0, 4 slots, no budget cap, mixed, tol 1e-6, 100
 * iterations, 4 refinements, inner tol / 2, far field auto, FMM order 5, leaf
 * 40, f and h not kept */
void beep_server_default_opts(beep_server_opts* o);

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

#ifdef __cplusplus
}
#endif

#endif /* BEEP_SERVE_H */
