// Plane-wave FMM (Yukawa and Laplace) for the BEM far field: basis functions
// shared by the host table builder (pw_tables.cpp) and the device kernels
// (fmm_pw.hip), and the host table interface.
//
// Expansions are real coefficient vectors over the regular solutions of
// (Laplacian - beta^2) f = 0, in box units (box edge 1) of their level:
//   psi_j(s) = g_n(|s|^2) * R_j(s),   j = (n, m, part), n <= p,
// with R_j the real, Schmidt semi-normalised regular solid harmonics
//   R_{n,0} = S_n^0,  R_{n,m,c} = sqrt2 Re S_n^m,  R_{n,m,s} = sqrt2 Im S_n^m,
//   S_n^m = sqrt((n-m)!/(n+m)!) r^n P_n^m(cos theta) e^{i m phi}  (no Condon-Shortley phase)
// and g_n(r^2) = i_n(beta r) (2n+1)!! / (beta r)^n (g_n = 1 for Laplace), so
// psi_j = r^n Y_j (1 + O(beta^2 r^2)) keeps every coefficient of order one in
// FP32 and FP64 alike.  Index j = n^2 (m = 0), n^2 + 2m - 1 (cos), n^2 + 2m
// (sin): (p + 1)^2 coefficients.
//
//   multipole moments   M_j = sum_src [ q psi_j(s) + d . grad psi_j(s) ]
//   local expansion     Phi(t) = sum_j L_j psi_j(t)
//
// The same quantities as the reference's spherical-harmonic expansions with
// scaled i_n (src/fmm/fmm.h:1426-1570 yformmp, :696-892 local evaluation,
// src/fmm/fmm_math_funcs.h:24-113 i_n), in a real basis.
#pragma once

#include <cmath>
#include <cstddef>
#include <vector>

#ifdef __HIPCC__
#define PW_HD __host__ __device__
#else
#define PW_HD
#endif

namespace beepdev {

constexpr int PW_MAXP = 18;

PW_HD inline int pw_nj(int p) { return (p + 1) * (p + 1); }
PW_HD inline int pw_idx(int n, int m, int part) { return m == 0 ? n * n : n * n + 2 * m - 1 + part; }

// g_n(r^2) for n = 0..P+1 (g[0..P+1]) at x2 = (beta r)^2: series for the top
// two, then the stable downward recurrence
//   g_{n-1} = g_n + x2 g_{n+1} / ((2n+1)(2n+3)).
template <typename T, int P>
PW_HD inline void pw_g(T x2, T* g) {
    if (x2 == T(0)) {
        for (int n = 0; n <= P + 1; ++n) g[n] = T(1);
        return;
    }
    // series g_n = sum_k (x2/2)^k / (k! prod_{i=1..k} (2n + 2i + 1))
    T top[2];
    for (int q = 0; q < 2; ++q) {
        const int n = P + 1 - q;
        T term = T(1), sum = T(1);
        for (int k = 1; k < 200; ++k) {
            term *= (x2 * T(0.5)) / (T(k) * T(2 * n + 2 * k + 1));
            sum += term;
            if (term < sum * T(1e-17)) break;
        }
        top[q] = sum;
    }
    g[P + 1] = top[0];
    g[P] = top[1];
    for (int n = P; n >= 1; --n) g[n - 1] = g[n] + x2 * g[n + 1] / (T(2 * n + 1) * T(2 * n + 3));
}

// Column m of the complex solid harmonics S_n^m, n = m..P, and their
// gradients, at (x, y, z): sr[n], si[n] (real, imaginary), gr[n][3], gi[n][3].
// Entries n < m are left untouched.  Recurrences (Schmidt-normalised):
//   S_m^m = prod_k sqrt((2k-1)/(2k)) (x + i y)^m
//   S_{m+1}^m = sqrt(2m+1) z S_m^m
//   S_n^m = (2n-1)/sqrt(n^2-m^2) z S_{n-1}^m - sqrt(((n-1)^2-m^2)/(n^2-m^2)) r^2 S_{n-2}^m
// with the gradient carried through each step.
template <typename T, int P>
PW_HD inline void pw_column(int m, T x, T y, T z, T* sr, T* si, T (*gr)[3], T (*gi)[3]) {
    // (x + i y)^(m-1) and the sectoral constant
    T pr = T(1), pi = T(0), c = T(1);
    for (int k = 1; k < m; ++k) {
        const T nr = pr * x - pi * y, ni = pr * y + pi * x;
        pr = nr;
        pi = ni;
    }
    for (int k = 1; k <= m; ++k) c *= sqrt(T(2 * k - 1) / T(2 * k));
    T s0r, s0i, g0r[3], g0i[3];
    if (m == 0) {
        s0r = T(1);
        s0i = T(0);
        for (int d = 0; d < 3; ++d) g0r[d] = g0i[d] = T(0);
    } else {
        // S = c (x+iy)^m, dS/dx = c m (x+iy)^(m-1), dS/dy = i c m (x+iy)^(m-1)
        s0r = c * (pr * x - pi * y);
        s0i = c * (pr * y + pi * x);
        g0r[0] = c * T(m) * pr;
        g0i[0] = c * T(m) * pi;
        g0r[1] = -c * T(m) * pi;
        g0i[1] = c * T(m) * pr;
        g0r[2] = g0i[2] = T(0);
    }
    sr[m] = s0r;
    si[m] = s0i;
    for (int d = 0; d < 3; ++d) {
        gr[m][d] = g0r[d];
        gi[m][d] = g0i[d];
    }
    if (m + 1 > P) return;
    const T r2 = x * x + y * y + z * z;
    const T a1 = sqrt(T(2 * m + 1));
    sr[m + 1] = a1 * z * s0r;
    si[m + 1] = a1 * z * s0i;
    for (int d = 0; d < 3; ++d) {
        gr[m + 1][d] = a1 * z * g0r[d];
        gi[m + 1][d] = a1 * z * g0i[d];
    }
    gr[m + 1][2] += a1 * s0r;
    gi[m + 1][2] += a1 * s0i;
    const T xyz[3] = {x, y, z};
    for (int n = m + 2; n <= P; ++n) {
        const T den = T(n * n - m * m);
        const T a = T(2 * n - 1) / sqrt(den);
        const T b = sqrt(T((n - 1) * (n - 1) - m * m) / den);
        const T r1 = sr[n - 1], i1 = si[n - 1], r0 = sr[n - 2], i0 = si[n - 2];
        sr[n] = a * z * r1 - b * r2 * r0;
        si[n] = a * z * i1 - b * r2 * i0;
        for (int d = 0; d < 3; ++d) {
            gr[n][d] = a * z * gr[n - 1][d] - b * (T(2) * xyz[d] * r0 + r2 * gr[n - 2][d]);
            gi[n][d] = a * z * gi[n - 1][d] - b * (T(2) * xyz[d] * i0 + r2 * gi[n - 2][d]);
        }
        gr[n][2] += a * r1;
        gi[n][2] += a * i1;
    }
}

// ---------------------------------------------------------------------------
// Host tables

// Plane-wave quadrature of one kernel at one level (box units): nodes t_k
// with weights w_k (the Laplace generalized Gaussian rule) and, per node,
// nphys_k angles alpha_j = 2 pi j / nphys_k of which the first half are kept
// (the other half are complex conjugates for real fields):
//   e^{-b r}/r = sum_k w_k e^{-(t_k+b) z} (1/nphys_k) sum_j e^{i sqrt(t_k^2+2 b t_k) (x cos a_j + y sin a_j)}
// for 1 <= z <= 4, sqrt(x^2+y^2) <= 4 sqrt2.
struct PwQuad {
    std::vector<double> u, w, wt;  // per node: decay t+b, oscillation, weight w_k / nphys_k
    std::vector<int> nhalf;        // per node: kept angles
    std::vector<double> alpha;     // per wave (node-major)
    std::vector<int> node;         // per wave: its node
    int nw = 0;                    // waves kept (sum of nhalf)
};

// psi_j(v), j < (p+1)^2, at frequency beta (box units), and the gradients
// (grad may be null): the host evaluation of the basis above.
void pw_basis(int p, double beta, const double v[3], double* psi, double (*grad)[3]);

// The six directions of the plane-wave lists: 0 up (+z), 1 down (-z),
// 2 north (+y), 3 south (-y), 4 east (+x), 5 west (-x).  rot[d] maps local to
// global coordinates (columns: local x, y, z axes), a proper rotation.
void pw_direction_frames(double rot[6][3][3]);
// direction of a source box at box-unit offset (target centre - source
// centre) t, max |t_i| in {2, 3}: z first, then y, then x; -1 if adjacent.
int pw_direction(int tx, int ty, int tz);

// Tables of one kernel (Laplace: beta = 0) at one level.  Row-major, padded:
//   jp  = (p+1)^2 rounded up to 16,  nwp = 12 nw rounded up to 64.
//   m2m  [8][jp][jp]    parent moments  M'_j  = sum_o sum_j' m2m[o][j][j'] M_o,j'
//                       (octant o: bit 0 x, bit 1 y, bit 2 z of the child), so
//                       the stacked K = 8 jp operand of a gathered GEMM is
//                       m2mT[(o jp + j')][j] = m2m[o][j][j']
//   l2lT [8][jp][jp]    child local      L_j' = sum_j L'_j l2lT[o][j][j'],
//                       l2lT[o][j][j'] = A_o[j][j'] (the same A_o as M2M)
//   m2x  [jp][nwp]      waves of the six directions W[(d nw + k) 2 + re/im] =
//                       (1/h) sum_j M_j m2x[j][...]   (m2x includes w_k / nphys_k;
//                       the factor 1/h of the level is applied by the caller)
//   x2l  [nwp][jp]      local  L_j = sum_c Wacc[c] x2l[c][j]
//   shift[98][nw][2]    X2X factors e^{-u dz + i w (dx cos a + dy sin a)} for the
//                       local offset code (dz - 2) 49 + (dy + 3) 7 + (dx + 3)
struct PwLevelTables {
    double beta = 0;  // kappa * h (0: Laplace)
    double h = 1;
    PwQuad q;
    std::vector<double> m2mT;  // [8 jp][jp]
    std::vector<double> l2lT;  // [8][jp][jp]
    std::vector<double> m2x;   // [jp][nwp]
    std::vector<double> x2l;   // [nwp][jp]
    std::vector<double> shift; // [98][nw][2]
};

struct PwTableSet {
    int p = 0, jp = 0, nlam = 0;
    int nwp = 0;  // common padded wave width of every level (max)
    std::vector<PwLevelTables> lev;  // index = level (0..D); levels < 2 hold no M2X/X2L
};

// Quadrature of the Laplace generalized Gaussian rule with nlam nodes (9 or
// 18) adapted to the scaled frequency beta (box units).
PwQuad pw_quadrature(int nlam, double beta);
// Tables for every level 0..D of a plan with root edge `edge`, for the given
// kernel frequency kappa (0: Laplace), expansion order p (<= PW_MAXP) and
// quadrature nlam.  Levels 0 and 1 carry only M2M / L2L.
PwTableSet pw_build_tables(int p, int nlam, double kappa, double edge, int D);

// Host diagnostic: errors of the tables against direct sums (relative,
// max over a few seeded cases): [0] plane-wave quadrature of the kernel,
// [1] P2M -> M2X -> X2X -> X2L -> L2P of a source cluster at a far box,
// [2] the same through M2M (child -> parent) and L2L (parent -> child),
// [3] gradient of [1], [4] rotation matrices' orthogonality defect.
void pw_selftest(int p, int nlam, double beta, double* err5);

}  // namespace beepdev
