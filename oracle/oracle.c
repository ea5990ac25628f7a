/*
 * oracle.c -- FP64 CPU restatement of the reference BEM mat-vec path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: it is
 * compiled and loaded by tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py, and nothing else.  The product (beep3_amd/) never links,
 * loads or calls it.
 *
 * Every routine restates one function of the reference in plain FP64 C with
 * OpenMP over rows:
 *   orc_pair          evaluate_local_BEM_kernels    src/bem/bem_kernels.h:134-251
 *   orc_singular      singular_BEM_kernels          src/bem/bem_kernels.h:18-132
 *                     + f/hGeometricCorrection      src/bem/bem_kernels.cpp:123-136
 *                     + diagonal assembly           src/beep/beep.cpp:826-842
 *   orc_matvec_rows   do_bem_fmm_iteration          src/beep/beep.cpp:1485-1583
 *                     near pairs = leaf x 27-box neighbourhood (beep.cpp:859-950,
 *                     octree.h:627-648), product rule local_integrations.h:123-130;
 *                     far pairs = the FMM's 1-point centroid interaction
 *                     (node_patch.h:339-394, bem_kernels.cpp:12-68) summed exactly
 *   orc_matvec_all    BEEP::write_matrix semantics   src/beep/beep.cpp:643-728
 *   orc_rhs           BEEP::set_rhs                  src/beep/beep.cpp:1124-1260 (direct sum,
 *                     the exact limit of its kappa=0 FMM; field = +grad phi, eval_pt.h:175-197)
 *   orc_energy_coeffs Mesh::init_energy_precalcs +  src/bem/mesh.cpp:730-925
 *                     renormalise_energy_precalcs (direct sums)
 *   orc_gmres_dense   BEEP::gmres                    src/beep/beep.cpp:1262-1483
 *
 * Pins (tests/test_oracle.py): Born-Kirkwood closed form, the double-layer
 * identity, finite-difference checks of the kernel derivatives, the kappa->0
 * limits, brute-force tiny cases and a dense LU solve for GMRES.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ONE_OVER_4PI 0.07957747154594767

/* Integrand accumulated over qualocation points (src, with weights w_q) and
 * quadrature points (targ).  Follows bem_kernels.h:161-248 term by term. */
static void integrate(double kappa, double inv_eps, const double* src, int nsrc,
                      const double* targ, int ntarg, double* Ao, double* Bo, double* Co,
                      double* Do) {
    const double sigm = 1e-9;
    double A_out = 0, B_out = 0, C_out = 0, D_out = 0;
    for (int s = 0; s < nsrc; ++s) {
        const double* q = src + 7 * s;
        double A = 0, B = 0, C = 0, D = 0;
        for (int t = 0; t < ntarg; ++t) {
            const double* p = targ + 7 * t;
            const double wt = p[6];
            const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (fabs(r2) < sigm) continue;
            const double r = sqrt(r2);
            const double ir2 = 1.0 / r2, ir = 1.0 / r;
            const double ir3 = ir2 * ir, ir4 = ir2 * ir2, ir5 = ir4 * ir;
            const double gr0 = ONE_OVER_4PI;
            const double gr1 = gr0 * ir, gr3 = gr0 * ir3;
            const double ur0 = exp(-kappa * r) * ONE_OVER_4PI;
            const double ur1 = ur0 * ir, ur2 = ur0 * ir2, ur3 = ur0 * ir3;
            const double ur3ur2 = ur3 + kappa * ur2;
            const double gd[3] = {-dx * gr3, -dy * gr3, -dz * gr3};
            const double ud[3] = {-dx * ur3ur2, -dy * ur3ur2, -dz * ur3ur2};
            const double* n = p + 3;
            const double* n0 = q + 3;
            A += wt * (gr1 - ur1);
            B += wt * (n[0] * (inv_eps * gd[0] - ud[0]) + n[1] * (inv_eps * gd[1] - ud[1]) +
                       n[2] * (inv_eps * gd[2] - ud[2]));
            C += -wt * (n0[0] * (gd[0] - ud[0] * inv_eps) + n0[1] * (gd[1] - ud[1] * inv_eps) +
                        n0[2] * (gd[2] - ud[2] * inv_eps));
            if (kappa != 0) {
                const double gr5 = gr0 * ir5, ur4 = ur0 * ir4, ur5 = ur0 * ir5;
                const double pur4ur3 = kappa * (3.0 * ur4 + kappa * ur3) + 3.0 * ur5;
                const double d[3] = {dx, dy, dz};
                for (int jj = 0; jj < 3; ++jj)
                    for (int kk = 0; kk < 3; ++kk) {
                        const double kr = (jj == kk) ? 1.0 : 0.0;
                        const double djk = d[jj] * d[kk];
                        const double gdd = 3.0 * djk * gr5 - kr * gr3;
                        const double udd = djk * pur4ur3 - kr * ur3ur2;
                        D -= wt * n0[kk] * n[jj] * (gdd - udd);
                    }
            }
        }
        const double qw = q[6];
        A_out += A * qw;
        B_out += B * qw;
        C_out += C * qw;
        D_out += D * qw;
    }
    *Ao = A_out;
    *Bo = B_out;
    *Co = C_out;
    *Do = D_out;
}

/* Near-field coefficients A,B,C,D of pair (i, j): qualocation points of i,
 * quadrature points of j, D scaled by 1/eps_j; zero for i == j. */
void orc_pair(const orc_geom* g, double kappa, int i, int j, double* out4) {
    if (i == j) {
        out4[0] = out4[1] = out4[2] = out4[3] = 0;
        return;
    }
    const double inv_eps = 1.0 / g->eps[j];
    double A, B, C, D;
    integrate(kappa, inv_eps, g->qual + 7 * g->qual_ptr[i], g->qual_ptr[i + 1] - g->qual_ptr[i],
              g->quad + 7 * g->quad_ptr[j], g->quad_ptr[j + 1] - g->quad_ptr[j], &A, &B, &C, &D);
    out4[0] = A;
    out4[1] = B;
    out4[2] = C;
    out4[3] = D * inv_eps;
}

/* 1-point centroid interaction of the far field: qualocation point =
 * centroid_i (weight 1, normal n_i), quadrature point = centroid_j (weight
 * area_j, normal n_j). */
void orc_pair_1pt(const orc_geom* g, double kappa, int i, int j, double* out4) {
    if (i == j) {
        out4[0] = out4[1] = out4[2] = out4[3] = 0;
        return;
    }
    double src[7], targ[7];
    for (int k = 0; k < 3; ++k) {
        src[k] = g->centroid[3 * i + k];
        src[3 + k] = g->normal[3 * i + k];
        targ[k] = g->centroid[3 * j + k];
        targ[3 + k] = g->normal[3 * j + k];
    }
    src[6] = 1.0;
    targ[6] = g->area[j];
    const double inv_eps = 1.0 / g->eps[j];
    double A, B, C, D;
    integrate(kappa, inv_eps, src, 1, targ, 1, &A, &B, &C, &D);
    out4[0] = A;
    out4[1] = B;
    out4[2] = C;
    out4[3] = D * inv_eps;
}

/* Diagonal entries per patch: singular integrals over the galerkin points
 * (both sides), normalised by the summed weights, plus the geometric
 * correction.  out: N x 4 = (A, B + fGC, C - hGC, D). */
void orc_singular(const orc_geom* g, double kappa, double* out) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < g->n; ++i) {
        const double* pts = g->gal + 7 * g->gal_ptr[i];
        const int np = g->gal_ptr[i + 1] - g->gal_ptr[i];
        const double inv_eps = 1.0 / g->eps[i];
        double A, B, C, D, wsum = 0;
        integrate(kappa, inv_eps, pts, np, pts, np, &A, &B, &C, &D);
        for (int k = 0; k < np; ++k) wsum += pts[7 * k + 6];
        A /= wsum;
        B /= wsum;
        C /= wsum;
        D *= inv_eps / wsum;
        const double gc = g->gc[i], e = g->eps[i];
        out[4 * i + 0] = A;
        out[4 * i + 1] = B + ((1.0 - gc) + gc / e);
        out[4 * i + 2] = C - (gc + (1.0 - gc) / e);
        out[4 * i + 3] = D;
    }
}

static int* leaf_of(const orc_geom* g) {
    int* lo = (int*)malloc(sizeof(int) * (size_t)g->n);
    for (int l = 0; l < g->n_leaves; ++l)
        for (int k = g->leaf_ptr[l]; k < g->leaf_ptr[l + 1]; ++k) lo[g->leaf_idx[k]] = l;
    return lo;
}

/* Rows of y = K x (x = [f; h], y = [f_out; h_out]) following the octree
 * split.  flags: bit0 = near pairs, bit1 = far pairs, bit2 = diagonal.
 * diag: N x 4 from orc_singular (may be NULL when bit2 is clear).
 * y_rows: 2 x n_rows (f rows then h rows).  rows == NULL means all rows. */
void orc_matvec_rows(const orc_geom* g, double kappa, const double* diag, int flags,
                     const double* x, int n_rows, const int* rows, double* y_rows) {
    const int N = g->n;
    int* lo = leaf_of(g);
    const int nr = rows ? n_rows : N;
#pragma omp parallel
    {
        char* near = (char*)calloc((size_t)N, 1);
#pragma omp for schedule(dynamic, 4)
        for (int r = 0; r < nr; ++r) {
            const int i = rows ? rows[r] : r;
            const int l = lo[i];
            double yf = 0, yh = 0, c[4];
            for (int k = g->nb_ptr[l]; k < g->nb_ptr[l + 1]; ++k) {
                const int j = g->nb_idx[k];
                near[j] = 1;
                if (flags & 1) {
                    orc_pair(g, kappa, i, j, c);
                    yf += c[1] * x[j] - c[0] * x[N + j];
                    yh += c[3] * x[j] - c[2] * x[N + j];
                }
            }
            if (flags & 2) {
                for (int j = 0; j < N; ++j) {
                    if (near[j] || j == i) continue;
                    orc_pair_1pt(g, kappa, i, j, c);
                    yf += c[1] * x[j] - c[0] * x[N + j];
                    yh += c[3] * x[j] - c[2] * x[N + j];
                }
            }
            for (int k = g->nb_ptr[l]; k < g->nb_ptr[l + 1]; ++k) near[g->nb_idx[k]] = 0;
            if (flags & 4) {
                const double* d = diag + 4 * i;
                yf += d[1] * x[i] - d[0] * x[N + i];
                yh += d[3] * x[i] - d[2] * x[N + i];
            }
            y_rows[r] = yf;
            y_rows[nr + r] = yh;
        }
        free(near);
    }
    free(lo);
}

/* Dense 2N x 2N operator (row-major) following the octree split, with the
 * diagonal; column j multiplies f_j, column N + j multiplies h_j. */
void orc_assemble_tree(const orc_geom* g, double kappa, const double* diag, double* K) {
    const int N = g->n;
    const size_t M = 2 * (size_t)N;
    int* lo = leaf_of(g);
#pragma omp parallel
    {
        char* near = (char*)calloc((size_t)N, 1);
#pragma omp for schedule(dynamic, 4)
        for (int i = 0; i < N; ++i) {
            double c[4];
            double* rf = K + (size_t)i * M;
            double* rh = K + (size_t)(N + i) * M;
            memset(rf, 0, sizeof(double) * M);
            memset(rh, 0, sizeof(double) * M);
            const int l = lo[i];
            for (int k = g->nb_ptr[l]; k < g->nb_ptr[l + 1]; ++k) {
                const int j = g->nb_idx[k];
                near[j] = 1;
                orc_pair(g, kappa, i, j, c);
                rf[j] += c[1];
                rf[N + j] -= c[0];
                rh[j] += c[3];
                rh[N + j] -= c[2];
            }
            for (int j = 0; j < N; ++j) {
                if (near[j] || j == i) continue;
                orc_pair_1pt(g, kappa, i, j, c);
                rf[j] += c[1];
                rf[N + j] -= c[0];
                rh[j] += c[3];
                rh[N + j] -= c[2];
            }
            for (int k = g->nb_ptr[l]; k < g->nb_ptr[l + 1]; ++k) near[g->nb_idx[k]] = 0;
            const double* d = diag + 4 * i;
            rf[i] += d[1];
            rf[N + i] -= d[0];
            rh[i] += d[3];
            rh[N + i] -= d[2];
        }
        free(near);
    }
    free(lo);
}

/* All-pairs operator (every pair with the full rules), BEEP::write_matrix. */
void orc_matvec_all(const orc_geom* g, double kappa, const double* diag, const double* x,
                    double* y) {
    const int N = g->n;
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < N; ++i) {
        double yf = 0, yh = 0, c[4];
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            orc_pair(g, kappa, i, j, c);
            yf += c[1] * x[j] - c[0] * x[N + j];
            yh += c[3] * x[j] - c[2] * x[N + j];
        }
        const double* d = diag + 4 * i;
        yf += d[1] * x[i] - d[0] * x[N + i];
        yh += d[3] * x[i] - d[2] * x[N + i];
        y[i] = yf;
        y[N + i] = yh;
    }
}

/* RHS: f_rhs = sum_q w_q phi(x_q) / (4 pi Dsolvent),
 *      h_rhs = sum_q w_q (grad phi . n_q) / (4 pi Dsolvent),
 * phi = sum_k q_k / |x - x_k| over all charges. */
void orc_rhs(const orc_geom* g, int nch, const double* xyzq, double Dsolvent, double* b) {
    const int N = g->n;
    const double eps = 2.220446049250313e-16;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < N; ++i) {
        double fr = 0, hr = 0;
        for (int k = g->qual_ptr[i]; k < g->qual_ptr[i + 1]; ++k) {
            const double* q = g->qual + 7 * k;
            double pot = 0, fx = 0, fy = 0, fz = 0;
            for (int c = 0; c < nch; ++c) {
                const double rx = q[0] - xyzq[4 * c], ry = q[1] - xyzq[4 * c + 1],
                             rz = q[2] - xyzq[4 * c + 2];
                const double d2 = rx * rx + ry * ry + rz * rz;
                if (d2 <= eps) continue;
                const double d = sqrt(d2);
                const double pt = xyzq[4 * c + 3] / d;
                pot += pt;
                const double ft = -pt / d2;
                fx += ft * rx;
                fy += ft * ry;
                fz += ft * rz;
            }
            fr += q[6] * pot;
            hr += q[6] * (fx * q[3] + fy * q[4] + fz * q[5]);
        }
        b[i] = fr * ONE_OVER_4PI / Dsolvent;
        b[N + i] = hr * ONE_OVER_4PI / Dsolvent;
    }
}

/* Energy coefficients of the patches [p0, p1) from charges [c0, c1) of the
 * same instance, over galerkin points, then the Gauss-law renormalisation.
 * ec: 2 x N (ec_f then ec_h) written for the patch range. */
void orc_energy_coeffs(const orc_geom* g, int p0, int p1, const double* xyzq, int c0, int c1,
                       double* ec_f, double* ec_h) {
    const double eps = 2.220446049250313e-16;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = p0; i < p1; ++i) {
        double ef = 0, eh = 0;
        for (int k = g->gal_ptr[i]; k < g->gal_ptr[i + 1]; ++k) {
            const double* q = g->gal + 7 * k;
            double pot = 0, fx = 0, fy = 0, fz = 0;
            for (int c = c0; c < c1; ++c) {
                const double rx = q[0] - xyzq[4 * c], ry = q[1] - xyzq[4 * c + 1],
                             rz = q[2] - xyzq[4 * c + 2];
                const double d2 = rx * rx + ry * ry + rz * rz;
                if (d2 <= eps) continue;
                const double d = sqrt(d2);
                const double pt = xyzq[4 * c + 3] / d;
                pot += pt;
                const double ft = -pt / d2;
                fx += ft * rx;
                fy += ft * ry;
                fz += ft * rz;
            }
            ef += (fx * q[3] + fy * q[4] + fz * q[5]) * q[6] * ONE_OVER_4PI;
            eh += pot * q[6] * ONE_OVER_4PI;
        }
        ec_f[i] = ef;
        ec_h[i] = eh;
    }
    double net = 0, total_f = 0, total_area = 0;
    for (int c = c0; c < c1; ++c) net += xyzq[4 * c + 3];
    for (int i = p0; i < p1; ++i) {
        total_f -= ec_f[i];
        total_area += g->area[i];
    }
    const double gauss_error = net - total_f;
    for (int i = p0; i < p1; ++i) ec_f[i] -= (g->area[i] / total_area) * gauss_error;
}

/* GMRES without restart on a dense row-major operator, following
 * BEEP::gmres (beep.cpp:1343-1432): r = K x - b, classical Gram-Schmidt via
 * two matrix-vector products, Givens rotations, stop when |y_j| <= eps |b|,
 * packed upper-triangular solve, x -= V y.  Returns its + 1 (mat-vec count). */
int orc_gmres_dense(int n, const double* K, const double* b, double* x, double eps, int m,
                    double* resid_hist) {
    double* V = (double*)calloc((size_t)n * (m + 1), sizeof(double));
    double* U = (double*)calloc((size_t)m * (m + 1) / 2 + 1, sizeof(double));
    double* r = (double*)malloc(sizeof(double) * n);
    double* y = (double*)calloc(m + 1, sizeof(double));
    double* c = (double*)calloc(m + 1, sizeof(double));
    double* s = (double*)calloc(m + 1, sizeof(double));
#define NEAR0(v) ((v) == 0 ? 2.2250738585072014e-308 : (v))
    double nrm2b = 0;
    for (int k = 0; k < n; ++k) nrm2b += b[k] * b[k];
    nrm2b = sqrt(nrm2b);
    for (int i = 0; i < n; ++i) {
        double acc = 0;
        for (int k = 0; k < n; ++k) acc += K[(size_t)i * n + k] * x[k];
        r[i] = acc - b[i];
    }
    double beta = 0;
    for (int k = 0; k < n; ++k) beta += r[k] * r[k];
    beta = sqrt(beta);
    for (int k = 0; k < n; ++k) V[k] = r[k] / NEAR0(beta);
    y[0] = beta;
    int j = 0, uij = 0;
    do {
        const int u0j = uij;
        double* vj = V + (size_t)j * n;
        double* vn = V + (size_t)(j + 1) * n;
        for (int i = 0; i < n; ++i) {
            double acc = 0;
            for (int k = 0; k < n; ++k) acc += K[(size_t)i * n + k] * vj[k];
            vn[i] = acc;
        }
        for (int q = 0; q <= j; ++q) {
            double acc = 0;
            for (int k = 0; k < n; ++k) acc += V[(size_t)q * n + k] * vn[k];
            U[u0j + q] = acc;
        }
        for (int q = 0; q <= j; ++q)
            for (int k = 0; k < n; ++k) vn[k] -= U[u0j + q] * V[(size_t)q * n + k];
        double h = 0;
        for (int k = 0; k < n; ++k) h += vn[k] * vn[k];
        h = sqrt(h);
        for (int k = 0; k < n; ++k) vn[k] /= NEAR0(h);
        for (int i = 0; i < j; ++i) {
            const double tmp = c[i] * U[uij] - s[i] * U[uij + 1];
            U[uij + 1] = s[i] * U[uij] + c[i] * U[uij + 1];
            U[uij] = tmp;
            ++uij;
        }
        const double rd = U[uij];
        const double dd = sqrt(rd * rd + h * h);
        c[j] = rd / NEAR0(dd);
        s[j] = -h / NEAR0(dd);
        U[uij] = dd;
        ++uij;
        y[j + 1] = s[j] * y[j];
        y[j] = c[j] * y[j];
        ++j;
        if (resid_hist) resid_hist[j - 1] = fabs(y[j]);
    } while (j < m && fabs(y[j]) > eps * nrm2b);
    /* packed column-major upper triangle: column q starts at q(q+1)/2 */
    for (int i = j - 1; i >= 0; --i) {
        double acc = y[i];
        for (int q = i + 1; q < j; ++q) acc -= U[q * (q + 1) / 2 + i] * y[q];
        y[i] = acc / U[i * (i + 1) / 2 + i];
    }
    for (int q = 0; q < j; ++q)
        for (int k = 0; k < n; ++k) x[k] -= V[(size_t)q * n + k] * y[q];
#undef NEAR0
    free(V);
    free(U);
    free(r);
    free(y);
    free(c);
    free(s);
    return j + 1;
}
