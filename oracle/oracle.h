/*
 * oracle.h -- declarations of the FP64 CPU oracle (oracle.c) and of the CPU
 * reference path (cpu_ref.cpp).  TEST INFRASTRUCTURE ONLY: used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.
 */
#ifndef BEEP_ORACLE_H
#define BEEP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n;
    const double *centroid, *normal, *area, *eps, *gc;
    const int* qual_ptr;
    const double* qual; /* 7 per point: x y z nx ny nz w */
    const int* quad_ptr;
    const double* quad;
    const int* gal_ptr;
    const double* gal;
    int n_leaves;
    const int *leaf_ptr, *leaf_idx, *nb_ptr, *nb_idx;
} orc_geom;

void orc_pair(const orc_geom* g, double kappa, int i, int j, double* out4);
void orc_pair_1pt(const orc_geom* g, double kappa, int i, int j, double* out4);
void orc_singular(const orc_geom* g, double kappa, double* out);
void orc_rhs(const orc_geom* g, int nch, const double* xyzq, double Dsolvent, double* b);

#ifdef __cplusplus
}
#endif

#endif
