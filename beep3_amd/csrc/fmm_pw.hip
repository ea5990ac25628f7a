// Plane-wave FMM far field on gfx950 (the reference's Yukawa / Laplace
// "new version" FMM, src/fmm/fmm_bem_hybrid.h:184-642 and src/fmm/fmm.h).
//
// Same quantity as the Chebyshev engine (fmm.hip): the 1-point interaction
// of every pair, as three potentials of monopole + dipole sources (Laplace;
// Yukawa for the f-row; Yukawa for the h-row, which carries 1/eps of the
// source).  Expansions are real solid-harmonic coefficient vectors (pw.hpp);
// per level, all boxes at once:
//   P2M   leaf moments                            VALU   k_pw_p2m
//   M2M   child x A_o by octant, summed per parent MFMA   k_pw_gemm + k_pw_m2m_sum, per level
//   M2X   six-direction plane waves = M x m2x     MFMA   k_pw_gemm, every level at once
//   X2X   diagonal translations, per target and   HBM    k_pw_x2x, every level at once
//         direction in list order
//   X2L   local = waves x x2l                     MFMA   k_pw_gemm, every level at once
//   L2L   child local += parent local x A_o       MFMA   k_pw_gemm (grouped by octant), per level
//   L2P   potential and gradient at the targets   VALU   k_pw_l2p
// and the leaf pass (P2P) of fmm.hip / fmm_residual.hip.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "device.hpp"
#include "fmm.hpp"
#include "pw.hpp"

namespace beepdev {

namespace {

constexpr double PI = 3.14159265358979323846;

typedef double pw64x4 __attribute__((ext_vector_type(4)));
typedef float pw32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct PwMfma;
template <>
struct PwMfma<double> {
    using acc_t = pw64x4;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // accumulator element r of lane l: row l/16 + 4r, column l % 16
    static __device__ __forceinline__ int row(int l, int r) { return l / 16 + 4 * r; }
};
template <>
struct PwMfma<float> {
    using acc_t = pw32x4;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // accumulator element r of lane l: row 4 (l/16) + r, column l % 16
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l / 16) + r; }
};

// recurrence constants a(n,m) = (2n-1)/sqrt(n^2-m^2), b(n,m) =
// sqrt(((n-1)^2-m^2)/(n^2-m^2)) (n >= m+2), a(m+1,m) = sqrt(2m+1), and the
// sectoral constant c(m) = prod_k sqrt((2k-1)/(2k)) at [m][m]
template <typename T, int P>
__device__ void pw_coeffs(T (*ab)[P + 1][2]) {
    for (int q = threadIdx.x; q < (P + 1) * (P + 1); q += blockDim.x) {
        const int n = q / (P + 1), m = q % (P + 1);
        double a = 0, b = 0;
        if (n == m) {
            a = 1;
            for (int k = 1; k <= m; ++k) a *= sqrt((2.0 * k - 1) / (2.0 * k));
        } else if (n == m + 1) {
            a = sqrt(2.0 * m + 1);
        } else if (n > m + 1) {
            const double den = double(n * n - m * m);
            a = (2.0 * n - 1) / sqrt(den);
            b = sqrt(double((n - 1) * (n - 1) - m * m) / den);
        }
        ab[n][m][0] = T(a);
        ab[n][m][1] = T(b);
    }
}

// Walk column m of the complex solid harmonics S_n^m (n = m..P) with their
// gradients at (x, y, z), calling f(n, sr, si, gr, gi) for every n (n a
// compile-time constant of the unrolled loop, so register arrays indexed by
// it stay in registers).  Rolling recurrence of pw.hpp (pw_column).
template <typename T, int P, typename F>
__device__ __forceinline__ void pw_walk(int m, T x, T y, T z, const T (*ab)[P + 1][2], F&& f) {
    T pr = T(1), pi = T(0);
    for (int k = 1; k < m; ++k) {
        const T nr = pr * x - pi * y, ni = pr * y + pi * x;
        pr = nr;
        pi = ni;
    }
    const T c = ab[m][m][0];
    const T r2 = x * x + y * y + z * z;
    T s1r = 0, s1i = 0, g1r[3] = {0, 0, 0}, g1i[3] = {0, 0, 0};  // S_{n-1}
    T s2r = 0, s2i = 0, g2r[3] = {0, 0, 0}, g2i[3] = {0, 0, 0};  // S_{n-2}
#pragma unroll
    for (int n = 0; n <= P; ++n) {
        if (n < m) continue;
        T sr, si, gr[3], gi[3];
        if (n == m) {
            if (m == 0) {
                sr = T(1);
                si = T(0);
                gr[0] = gr[1] = gr[2] = gi[0] = gi[1] = gi[2] = T(0);
            } else {
                sr = c * (pr * x - pi * y);
                si = c * (pr * y + pi * x);
                const T cm = c * T(m);
                gr[0] = cm * pr;
                gi[0] = cm * pi;
                gr[1] = -cm * pi;
                gi[1] = cm * pr;
                gr[2] = gi[2] = T(0);
            }
        } else if (n == m + 1) {
            const T a1 = ab[n][m][0];
            sr = a1 * z * s1r;
            si = a1 * z * s1i;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                gr[d] = a1 * z * g1r[d];
                gi[d] = a1 * z * g1i[d];
            }
            gr[2] += a1 * s1r;
            gi[2] += a1 * s1i;
        } else {
            const T a = ab[n][m][0], b = ab[n][m][1];
            const T az = a * z, br2 = b * r2, b2 = T(2) * b;
            sr = az * s1r - br2 * s2r;
            si = az * s1i - br2 * s2i;
            gr[0] = az * g1r[0] - br2 * g2r[0] - b2 * x * s2r;
            gi[0] = az * g1i[0] - br2 * g2i[0] - b2 * x * s2i;
            gr[1] = az * g1r[1] - br2 * g2r[1] - b2 * y * s2r;
            gi[1] = az * g1i[1] - br2 * g2i[1] - b2 * y * s2i;
            gr[2] = az * g1r[2] - br2 * g2r[2] - b2 * z * s2r + a * s1r;
            gi[2] = az * g1i[2] - br2 * g2i[2] - b2 * z * s2i + a * s1i;
        }
        f(n, sr, si, gr, gi);
        s2r = s1r;
        s2i = s1i;
        s1r = sr;
        s1i = si;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            g2r[d] = g1r[d];
            g2i[d] = g1i[d];
            g1r[d] = gr[d];
            g1i[d] = gi[d];
        }
    }
}

// lanes per source in P2M (one lane per m)
template <int P>
constexpr int pw_mlanes() { return P + 1 <= 16 ? 16 : 32; }

// P2M: one block per leaf box.  Lane m of a group walks column m for every
// source of the group (sources g, g + G, ...); the G group partial sums are
// added in group order through LDS (deterministic).  Channel 0 (Laplace):
// monopole -a h, dipole ie a f n; channels 1, 2 (Yukawa): monopole a h,
// dipole -a f n, and ie times that.
template <typename T, int P>
__global__ __launch_bounds__(256) void k_pw_p2m(const FmmDev d, const Src64* __restrict__ src,
                                               double kappa, int jp, T* __restrict__ M) {
    constexpr int ML = pw_mlanes<P>(), G = 256 / ML;
    constexpr int NJ = (P + 1) * (P + 1);
    __shared__ T ab[P + 1][P + 1][2];
    extern __shared__ char pw_smem[];
    T* red = reinterpret_cast<T*>(pw_smem);  // [G][2][jp]
    const int bl = xcd_remap(blockIdx.x, gridDim.x);
    const int box = d.leaf_first + bl;
    const int s0 = d.leaf_src[bl], s1 = d.leaf_src[bl + 1];
    const double4 bc = d.box_c[box];
    const double h = d.box_h[d.D], ih = 1.0 / h;
    const T beta = T(kappa * h);
    const int t = threadIdx.x, m = t % ML, g = t / ML;
    pw_coeffs<T, P>(ab);
    __syncthreads();
    const T rt2 = T(1.4142135623730950488);
    for (int pass = 0; pass < 2; ++pass) {
        // pass 0: Laplace (channel 0); pass 1: Yukawa (channels 1 and 2)
        T a1r[P + 1], a1i[P + 1], a2r[P + 1], a2i[P + 1];
#pragma unroll
        for (int n = 0; n <= P; ++n) a1r[n] = a1i[n] = a2r[n] = a2i[n] = T(0);
        if (m <= P) {
            for (int k = s0 + g; k < s1; k += G) {
                const Src64 s = src[k];
                const T x = T((s.x - bc.x) * ih), y = T((s.y - bc.y) * ih), z = T((s.z - bc.z) * ih);
                const T sh = T(s.sh), ie = T(s.inv_eps);
                const T dx = T(s.dx * ih), dy = T(s.dy * ih), dz = T(s.dz * ih);
                T gg[P + 2];
                if (pass == 1) {
                    pw_g<T, P>(beta * beta * (x * x + y * y + z * z), gg);
                } else {
#pragma unroll
                    for (int n = 0; n <= P + 1; ++n) gg[n] = T(1);
                }
                // channel 0: q = -sh, d = ie dv; channel 1: q = sh, d = -dv
                const T q = pass == 0 ? -sh : sh;
                const T ddx = pass == 0 ? ie * dx : -dx, ddy = pass == 0 ? ie * dy : -dy,
                        ddz = pass == 0 ? ie * dz : -dz;
                const T b2 = beta * beta;
                pw_walk<T, P>(m, x, y, z, ab, [&](int n, T sr, T si, const T* gr, const T* gi) {
                    const T gn = gg[n];
                    // grad psi = g_n grad S + beta^2 g_{n+1} / (2n+3) S v
                    const T dg = pass == 1 ? b2 * gg[n + 1] / T(2 * n + 3) : T(0);
                    const T vd = ddx * x + ddy * y + ddz * z;
                    const T cr = gn * (q * sr + ddx * gr[0] + ddy * gr[1] + ddz * gr[2]) + dg * sr * vd;
                    const T ci = gn * (q * si + ddx * gi[0] + ddy * gi[1] + ddz * gi[2]) + dg * si * vd;
                    a1r[n] += cr;
                    a1i[n] += ci;
                    if (pass == 1) {
                        a2r[n] += ie * cr;
                        a2i[n] += ie * ci;
                    }
                });
            }
        }
        // group partial sums to LDS (entries of column m), then the sum
        // over groups in order
        for (int q = t; q < G * 2 * jp; q += 256) red[q] = T(0);
        __syncthreads();
        if (m <= P) {
            T* r0 = red + (g * 2) * jp;
            T* r1 = r0 + jp;
#pragma unroll
            for (int n = 0; n <= P; ++n) {
                if (n < m) continue;
                if (m == 0) {
                    r0[n * n] = a1r[n];
                    r1[n * n] = a2r[n];
                } else {
                    r0[n * n + 2 * m - 1] = rt2 * a1r[n];
                    r0[n * n + 2 * m] = rt2 * a1i[n];
                    r1[n * n + 2 * m - 1] = rt2 * a2r[n];
                    r1[n * n + 2 * m] = rt2 * a2i[n];
                }
            }
        }
        __syncthreads();
        const int nch = pass == 0 ? 1 : 2;
        for (int q = t; q < nch * jp; q += 256) {
            const int c = q / jp, j = q - c * jp;
            T sum = T(0);
            if (j < NJ)
                for (int gi = 0; gi < G; ++gi) sum += red[(gi * 2 + c) * jp + j];
            M[(size_t(box) * 3 + (pass == 0 ? 0 : 1 + c)) * jp + j] = sum;
        }
        __syncthreads();
    }
}

// L2P: one wave per leaf box, lane = target; the box's three local
// expansions in LDS.  f += (Phi_L + Phi_1) / 4pi, h += n . grad(Phi_L +
// Phi_2) / 4pi (gradient in physical units: 1/h per box unit).
template <typename T, int P>
__global__ __launch_bounds__(64) void k_pw_l2p(const FmmDev d, const Patch64* __restrict__ patches,
                                              double kappa, int jp, const T* __restrict__ L,
                                              double* __restrict__ y, int n_patch) {
    __shared__ T ab[P + 1][P + 1][2];
    extern __shared__ char pw_smem[];
    T* lb = reinterpret_cast<T*>(pw_smem);  // [3][jp]
    const int bl = xcd_remap(blockIdx.x, gridDim.x);
    const int box = d.leaf_first + bl;
    const int s0 = d.leaf_src[bl], s1 = d.leaf_src[bl + 1];
    const double4 bc = d.box_c[box];
    const double h = d.box_h[d.D], ih = 1.0 / h;
    const T beta = T(kappa * h), b2 = beta * beta;
    const int t = threadIdx.x;
    pw_coeffs<T, P>(ab);
    for (int q = t; q < 3 * jp; q += 64) lb[q] = L[size_t(box) * 3 * jp + q];
    __syncthreads();
    const T rt2 = T(1.4142135623730950488);
    for (int ti = s0 + t; ti < s1; ti += 64) {
        const int i = d.perm[ti];
        const Patch64 p = patches[i];
        const T x = T((p.cx - bc.x) * ih), yv = T((p.cy - bc.y) * ih), z = T((p.cz - bc.z) * ih);
        T gg[P + 2];
        pw_g<T, P>(b2 * (x * x + yv * yv + z * z), gg);
        T pL = 0, p1 = 0, gL[3] = {0, 0, 0}, g2v[3] = {0, 0, 0};
        for (int m = 0; m <= P; ++m) {
            const T sc = m == 0 ? T(1) : rt2;
            pw_walk<T, P>(m, x, yv, z, ab, [&](int n, T sr, T si, const T* gr, const T* gi) {
                const int jc = n * n + (m == 0 ? 0 : 2 * m - 1);
                const T l0c = lb[jc] * sc, l1c = lb[jp + jc] * sc, l2c = lb[2 * jp + jc] * sc;
                const T l0s = m == 0 ? T(0) : lb[jc + 1] * sc;
                const T l1s = m == 0 ? T(0) : lb[jp + jc + 1] * sc;
                const T l2s = m == 0 ? T(0) : lb[2 * jp + jc + 1] * sc;
                // Laplace: psi = S; Yukawa: psi = g_n S, grad = g_n grad S + dg S v
                pL += l0c * sr + l0s * si;
                const T gn = gg[n], dg = b2 * gg[n + 1] / T(2 * n + 3);
                p1 += gn * (l1c * sr + l1s * si);
                const T s2 = l2c * sr + l2s * si;
                const T v[3] = {x, yv, z};
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    gL[a] += l0c * gr[a] + l0s * gi[a];
                    g2v[a] += gn * (l2c * gr[a] + l2s * gi[a]) + dg * s2 * v[a];
                }
            });
        }
        const T nx = T(p.nx), ny = T(p.ny), nz = T(p.nz);
        y[i] += double(pL + p1) * FOUR_PI_INV;
        y[n_patch + i] += double(T(ih) * (nx * (gL[0] + g2v[0]) + ny * (gL[1] + g2v[1]) +
                                           nz * (gL[2] + g2v[2]))) * FOUR_PI_INV;
    }
}

// Batched GEMMs on MFMA (the PwProb problems of one launch, fmm.hpp).  A
// block computes one 64 x 64 tile of one problem: four waves of 16 rows x 64
// columns, K staged through LDS 32 at a time with the next slab's loads
// issued before the current slab's products (register prefetch), so the
// K loop is not one memory latency per step.  Each output is written by one
// lane, so the result does not depend on the launch shape.
constexpr int PW_BM = 64, PW_BN = 64, PW_BK = 32;

template <typename T>
__device__ __forceinline__ T pw_load_a(const PwProb<T>& g, const int* __restrict__ child8,
                                       const int* __restrict__ parent, int box, int ch, int k) {
    if (box < 0 || k >= g.K) return T(0);
    int src = box, kb = k;
    if (g.mode == 1) {
        const int o = k / g.kblk;
        kb = k - o * g.kblk;
        src = child8[size_t(box) * 8 + o];
    } else if (g.mode == 2) {
        src = parent[box];
    }
    return src >= 0 ? g.A[(size_t(src - g.abox0) * 3 + ch) * g.lda + kb] : T(0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_pw_gemm(const PwProb<T>* __restrict__ probs, int nprob,
                                                const int* __restrict__ child8,
                                                const int* __restrict__ parent) {
    using acc_t = typename PwMfma<T>::acc_t;
    constexpr int AE = PW_BM * PW_BK / 256, BE = PW_BK * PW_BN / 256;  // loads per thread
    __shared__ T As[PW_BM][PW_BK + 1];
    __shared__ T Bs[PW_BK][PW_BN + 4];
    __shared__ int rbox[PW_BM], rch[PW_BM];
    // this block's problem and tile
    int b = blockIdx.x, pi = 0;
    while (pi < nprob - 1 && b >= probs[pi].tiles) {
        b -= probs[pi].tiles;
        ++pi;
    }
    const PwProb<T>& g = probs[pi];
    const int r0 = (b / g.ctiles) * PW_BM, n0 = (b % g.ctiles) * PW_BN;
    const int nr = g.nrows * g.nch;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    if (t < PW_BM) {
        const int r = r0 + t;
        rbox[t] = -1;
        rch[t] = 0;
        if (r < nr) {
            const int bi = r / g.nch;
            rbox[t] = g.rows ? g.rows[bi] : g.first + bi;
            rch[t] = g.ch0 + r % g.nch;
        }
    }
    __syncthreads();
    T ra[AE], rb[BE];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < AE; ++i) {
            const int e = t + 256 * i, rr = e / PW_BK, kk = e % PW_BK;
            ra[i] = pw_load_a(g, child8, parent, rbox[rr], rch[rr], k0 + kk);
        }
#pragma unroll
        for (int i = 0; i < BE; ++i) {
            const int e = t + 256 * i, kk = e / PW_BN, nn = e % PW_BN;
            const int k = k0 + kk, n = n0 + nn;
            rb[i] = (k < g.K && n < g.N) ? g.B[size_t(k) * g.ldb + n] : T(0);
        }
    };
    acc_t acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = acc_t{0, 0, 0, 0};
    fetch(0);
    for (int k0 = 0; k0 < g.K; k0 += PW_BK) {
#pragma unroll
        for (int i = 0; i < AE; ++i) {
            const int e = t + 256 * i;
            As[e / PW_BK][e % PW_BK] = ra[i];
        }
#pragma unroll
        for (int i = 0; i < BE; ++i) {
            const int e = t + 256 * i;
            Bs[e / PW_BN][e % PW_BN] = rb[i];
        }
        __syncthreads();
        if (k0 + PW_BK < g.K) fetch(k0 + PW_BK);  // in flight during the products
#pragma unroll
        for (int ks = 0; ks < PW_BK / 4; ++ks) {
            const T av = As[16 * w + (l % 16)][4 * ks + l / 16];
#pragma unroll
            for (int c = 0; c < 4; ++c)
                acc[c] = PwMfma<T>::op(av, Bs[4 * ks + l / 16][16 * c + (l % 16)], acc[c]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int n = n0 + 16 * c + (l % 16);
        if (n >= g.N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int rr = 16 * w + PwMfma<T>::row(l, r);
            const int box = rbox[rr];
            if (box < 0) continue;
            T* cp = g.C + (size_t(box - g.cbox0) * 3 + rch[rr]) * g.ldc + n;
            const T v = g.alpha * acc[c][r];
            *cp = g.accumulate ? *cp + v : v;
        }
    }
}

// X2X: one block per (target box at a level >= 2, direction), every level
// in one launch.  For each channel and wave: acc = sum over the direction's
// list (in list order) of W_src x shift[code]; written to Wacc (every entry
// of the direction, zero for an empty list; direction 0 also clears the
// padding).  W and Wacc are indexed by box - base (base = the first box of
// level 2).  Channel 0 uses the Laplace waves and shifts of the level,
// channels 1 and 2 the Yukawa ones.
template <typename T>
__global__ __launch_bounds__(256) void k_pw_x2x(int base, int nwp, const int* __restrict__ box_level,
                                               const int* __restrict__ xl_ptr,
                                               const int* __restrict__ xl_src,
                                               const int* __restrict__ xl_code,
                                               const int* __restrict__ lv_nw, T* const* __restrict__ lv_sh,
                                               const T* __restrict__ W, T* __restrict__ Wacc) {
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int tb = bid / 6, dir = bid - 6 * tb;
    const int box = base + tb;
    const int lv = box_level[box];
    const int nwL = lv_nw[2 * lv], nwY = lv_nw[2 * lv + 1];
    const T* shL = lv_sh[2 * lv];
    const T* shY = lv_sh[2 * lv + 1];
    const int e0 = xl_ptr[tb * 6 + dir], e1 = xl_ptr[tb * 6 + dir + 1];
    const int nitem = nwL + 2 * nwY;
    for (int it = threadIdx.x; it < nitem; it += 256) {
        const int ch = it < nwL ? 0 : 1 + (it - nwL) / nwY;
        const int wv = ch == 0 ? it : (it - nwL) % nwY;
        const int nw = ch == 0 ? nwL : nwY;
        const T* sh = ch == 0 ? shL : shY;
        const int col = (dir * nw + wv) * 2;
        T ar = 0, ai = 0;
        for (int e = e0; e < e1; ++e) {
            const int sb = xl_src[e] - base, code = xl_code[e];
            const T* wp = W + (size_t(sb) * 3 + ch) * nwp + col;
            const T wr = wp[0], wi = wp[1];
            const T sr = sh[(size_t(code) * nw + wv) * 2], si = sh[(size_t(code) * nw + wv) * 2 + 1];
            ar += wr * sr - wi * si;
            ai += wr * si + wi * sr;
        }
        T* op = Wacc + (size_t(tb) * 3 + ch) * nwp + col;
        op[0] = ar;
        op[1] = ai;
    }
    if (dir == 0) {
        for (int q = threadIdx.x; q < 3 * nwp; q += 256) {
            const int ch = q / nwp, c = q - ch * nwp;
            if (c >= 12 * (ch == 0 ? nwL : nwY)) Wacc[(size_t(tb) * 3 + ch) * nwp + c] = T(0);
        }
    }
}

template <typename T>
__global__ void k_pw_zero(size_t n, T* __restrict__ p) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        p[i] = T(0);
}

// M2M, second half: the moments of every parent box at a level are the sum
// of its children's translated moments T (written by the octant GEMMs), in
// octant order.  One thread per (parent, channel, coefficient).
template <typename T>
__global__ void k_pw_m2m_sum(int first, int count, int jp, const int* __restrict__ child8,
                             const T* __restrict__ Tc, T* __restrict__ M) {
    const size_t q = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
    if (q >= size_t(count) * 3 * jp) return;
    const int box = first + int(q / (size_t(3) * jp));
    const size_t cj = q % (size_t(3) * jp);
    T sum = T(0);
    for (int o = 0; o < 8; ++o) {
        const int c = child8[size_t(box) * 8 + o];
        if (c >= 0) sum += Tc[size_t(c) * 3 * jp + cj];
    }
    M[size_t(box) * 3 * jp + cj] = sum;
}

template <typename T>
T* upload_tab(DevArena& arena, const char* name, const std::vector<double>& v, hipStream_t s) {
    std::vector<T> h(v.begin(), v.end());
    T* p = arena.get<T>(name, std::max<size_t>(h.size(), 1));
    if (!h.empty()) hip_check(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "H2D pw table");
    hip_check(hipStreamSynchronize(s), "sync pw table");  // the host copy dies here
    return p;
}

template <typename T>
void make_pw_tables(FmmPlan* pl, double kappa, PwDevTables<T>& out, hipStream_t s) {
    const int D = pl->D;
    const PwTableSet lap = pw_build_tables(pl->pw_p, pl->pw_nlam, 0.0, pl->edge, D);
    const PwTableSet yuk = pw_build_tables(pl->pw_p, pl->pw_nlam, kappa, pl->edge, D);
    if (lap.jp != pl->pw_jp || yuk.jp != pl->pw_jp) throw std::runtime_error("plane-wave table shape");
    pl->pw_nwp = std::max(lap.nwp, yuk.nwp);
    const int nwp = pl->pw_nwp, jp = pl->pw_jp;
    out.lev.assign(D + 1, PwDevLevel<T>{});
    const std::string tag = sizeof(T) == 8 ? "64" : "32";
    auto repad = [&](const std::vector<double>& m, int rows, int cols_from, int cols_to) {
        std::vector<double> o(size_t(rows) * cols_to, 0.0);
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols_from; ++c) o[size_t(r) * cols_to + c] = m[size_t(r) * cols_from + c];
        return o;
    };
    auto rowpad = [&](const std::vector<double>& m, int rows_from, int rows_to, int cols) {
        std::vector<double> o(size_t(rows_to) * cols, 0.0);
        std::copy(m.begin(), m.begin() + size_t(rows_from) * cols, o.begin());
        return o;
    };
    for (int l = 0; l <= D; ++l) {
        PwDevLevel<T>& dl = out.lev[l];
        for (int k = 0; k < 2; ++k) {
            const PwLevelTables& t = (k == 0 ? lap : yuk).lev[l];
            const int tnwp = (k == 0 ? lap : yuk).nwp;
            const std::string pre = "pw" + tag + (k == 0 ? "L" : "Y") + std::to_string(l) + "_";
            if (!t.m2mT.empty()) {
                dl.m2mT[k] = upload_tab<T>(*pl->arena, (pre + "m2m").c_str(), t.m2mT, s);
                dl.l2lT[k] = upload_tab<T>(*pl->arena, (pre + "l2l").c_str(), t.l2lT, s);
            }
            if (l >= 2) {
                dl.m2x[k] = upload_tab<T>(*pl->arena, (pre + "m2x").c_str(), repad(t.m2x, jp, tnwp, nwp), s);
                dl.x2l[k] = upload_tab<T>(*pl->arena, (pre + "x2l").c_str(), rowpad(t.x2l, tnwp, nwp, jp), s);
                dl.shift[k] = upload_tab<T>(*pl->arena, (pre + "sh").c_str(), t.shift, s);
                dl.nw[k] = t.q.nw;
            }
        }
    }
}

// The problems of every batched GEMM launch of a plan (expansion buffers M,
// L, plane waves W, accumulated waves Wacc), uploaded once per precision,
// and the per-level parameters of the all-level X2X launch.
template <typename T>
void make_pw_launches(FmmPlan* pl, PwDevTables<T>& tb, T* M, T* L, T* W, T* Wacc, T* Tc, hipStream_t s) {
    const int D = pl->D, jp = pl->pw_jp, nwp = pl->pw_nwp;
    const int base = D >= 2 ? pl->lvl_first[2] : pl->nbox;
    std::vector<PwProb<T>> probs;
    auto add = [&](PwProb<T> g) {
        const int rt = (g.nrows * g.nch + PW_BM - 1) / PW_BM;
        g.ctiles = (g.N + PW_BN - 1) / PW_BN;
        g.tiles = rt * g.ctiles;
        if (g.tiles > 0) probs.push_back(g);
    };
    auto launch = [&](size_t first) {
        int blocks = 0;
        for (size_t q = first; q < probs.size(); ++q) blocks += probs[q].tiles;
        return int3{int(first), int(probs.size() - first), blocks};
    };
    // B's row stride is N, except for M2X (the table's padded width)
    auto prob = [&](int mode, int k, int K, int N, int lda, int ldc, int acc, const T* A, const T* B, T* C,
                    T alpha, int abox0, int cbox0) {
        PwProb<T> g{};
        g.mode = mode;
        g.nch = k == 0 ? 1 : 2;
        g.ch0 = k == 0 ? 0 : 1;
        g.K = K;
        g.N = N;
        g.kblk = jp;
        g.lda = lda;
        g.ldb = N;
        g.ldc = ldc;
        g.accumulate = acc;
        g.A = A;
        g.B = B;
        g.C = C;
        g.alpha = alpha;
        g.abox0 = abox0;
        g.cbox0 = cbox0;
        return g;
    };
    // M2M of the parents at level l: every child (level l + 1, grouped by
    // its octant o) translated by its octant's matrix into Tc, then summed
    // per parent (k_pw_m2m_sum): K = jp per product
    tb.launch_m2m.clear();
    for (int l = D - 1; l >= 2; --l) {
        const size_t f = probs.size();
        const int* gp = pl->pw_l2l_gptr_h.data() + size_t(l + 1) * 9;
        for (int k = 0; k < 2; ++k)
            for (int o = 0; o < 8; ++o) {
                PwProb<T> g = prob(0, k, jp, jp, jp, jp, 0, M, tb.lev[l].m2mT[k] + size_t(o) * jp * jp, Tc,
                                   T(1), 0, 0);
                g.rows = pl->pw_l2l_rows + gp[o];
                g.nrows = gp[o + 1] - gp[o];
                add(g);
            }
        tb.launch_m2m.push_back(launch(f));
    }
    {
        const size_t f = probs.size();
        // only the 12 nw real columns of each kernel and level are formed
        // (X2X reads no others)
        for (int l = 2; l <= D; ++l)
            for (int k = 0; k < 2; ++k) {
                PwProb<T> g = prob(0, k, jp, 12 * tb.lev[l].nw[k], jp, nwp, 0, M, tb.lev[l].m2x[k], W,
                                   T(1.0 / pl->box_h[l]), 0, base);
                g.ldb = nwp;
                g.first = pl->lvl_first[l];
                g.nrows = pl->lvl_first[l + 1] - pl->lvl_first[l];
                add(g);
            }
        tb.launch_m2x = launch(f);
    }
    {
        // into zeroed locals: L = X2L, then L2L adds the parent's part
        const size_t f = probs.size();
        for (int l = 2; l <= D; ++l)
            for (int k = 0; k < 2; ++k) {
                PwProb<T> g = prob(0, k, 12 * tb.lev[l].nw[k], jp, nwp, jp, 1, Wacc, tb.lev[l].x2l[k], L, T(1),
                                   base, 0);
                g.first = pl->lvl_first[l];
                g.nrows = pl->lvl_first[l + 1] - pl->lvl_first[l];
                add(g);
            }
        tb.launch_x2l = launch(f);
    }
    tb.launch_l2l.clear();
    for (int l = 3; l <= D; ++l) {
        const size_t f = probs.size();
        const int* gp = pl->pw_l2l_gptr_h.data() + size_t(l) * 9;
        for (int k = 0; k < 2; ++k)
            for (int o = 0; o < 8; ++o) {
                PwProb<T> g = prob(2, k, jp, jp, jp, jp, 1, L, tb.lev[l - 1].l2lT[k] + size_t(o) * jp * jp, L,
                                   T(1), 0, 0);
                g.rows = pl->pw_l2l_rows + gp[o];
                g.nrows = gp[o + 1] - gp[o];
                add(g);
            }
        tb.launch_l2l.push_back(launch(f));
    }
    const std::string tag = sizeof(T) == 8 ? "64" : "32";
    if (!probs.empty()) {
        tb.probs = pl->arena->get<PwProb<T>>("pw_probs" + tag, probs.size());
        hip_check(hipMemcpyAsync(tb.probs, probs.data(), probs.size() * sizeof(PwProb<T>),
                                 hipMemcpyHostToDevice, s), "H2D pw problems");
    }
    std::vector<int> nw(size_t(2) * (D + 1), 0);
    std::vector<T*> sh(size_t(2) * (D + 1), nullptr);
    for (int l = 2; l <= D; ++l)
        for (int k = 0; k < 2; ++k) {
            nw[2 * l + k] = tb.lev[l].nw[k];
            sh[2 * l + k] = tb.lev[l].shift[k];
        }
    tb.x2x_nw = pl->arena->get<int>("pw_x2x_nw" + tag, nw.size());
    tb.x2x_sh = pl->arena->get<T*>("pw_x2x_sh" + tag, sh.size());
    hip_check(hipMemcpyAsync(tb.x2x_nw, nw.data(), nw.size() * sizeof(int), hipMemcpyHostToDevice, s),
              "H2D pw x2x");
    hip_check(hipMemcpyAsync(tb.x2x_sh, sh.data(), sh.size() * sizeof(T*), hipMemcpyHostToDevice, s),
              "H2D pw x2x");
    hip_check(hipStreamSynchronize(s), "sync pw launches");  // the host copies die here
}

template <typename T>
void gemm_launch(hipStream_t s, const FmmPlan* pl, const PwDevTables<T>& tb, int3 lc) {
    if (lc.z <= 0) return;
    hipLaunchKernelGGL(k_pw_gemm<T>, dim3(lc.z), dim3(256), 0, s, tb.probs + lc.x, lc.y, pl->pw_child8,
                       pl->dv.box_parent);
    hip_check(hipGetLastError(), "k_pw_gemm launch");
}

template <typename T, int P>
void pw_run(FmmPlan* pl, hipStream_t s, const Patch64* patches, double kappa, const PwDevTables<T>& tb,
            T* M, T* L, T* W, T* Wacc, T* Tc, double* y, hipEvent_t* m2l_ev) {
    const FmmDev& d = pl->dv;
    const int D = pl->D, jp = pl->pw_jp, nwp = pl->pw_nwp;
    {
        const size_t n = size_t(pl->nbox) * 3 * jp;
        hipLaunchKernelGGL(k_pw_zero<T>, dim3(unsigned(std::min<size_t>((n + 255) / 256, 65535))), dim3(256), 0, s, n, L);
    }
    constexpr int ML = pw_mlanes<P>(), G = 256 / ML;
    hipLaunchKernelGGL((k_pw_p2m<T, P>), dim3(pl->nleaf), dim3(256), size_t(G) * 2 * jp * sizeof(T), s, d,
                       pl->src, kappa, jp, M);
    hip_check(hipGetLastError(), "k_pw_p2m launch");
    // M2M: parents at levels D-1 .. 2 (the children's products in one
    // launch per level, then the sum per parent)
    for (int l = D - 1, q = 0; l >= 2; --l, ++q) {
        gemm_launch(s, pl, tb, tb.launch_m2m[q]);
        const int first = pl->lvl_first[l], cnt = pl->lvl_first[l + 1] - first;
        const size_t n = size_t(cnt) * 3 * jp;
        hipLaunchKernelGGL(k_pw_m2m_sum<T>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, first, cnt, jp,
                           pl->pw_child8, Tc, M);
        hip_check(hipGetLastError(), "k_pw_m2m_sum launch");
    }
    if (m2l_ev) hip_check(hipEventRecord(m2l_ev[0], s), "hipEventRecord");
    if (D >= 2) {
        // M2X, X2X and X2L of every level >= 2, one launch each
        gemm_launch(s, pl, tb, tb.launch_m2x);
        const int base = pl->lvl_first[2], nb = pl->nbox - base;
        hipLaunchKernelGGL(k_pw_x2x<T>, dim3(6 * nb), dim3(256), 0, s, base, nwp, pl->pw_box_level,
                           pl->pw_xl_ptr, pl->pw_xl_src, pl->pw_xl_code, tb.x2x_nw, tb.x2x_sh, W, Wacc);
        hip_check(hipGetLastError(), "k_pw_x2x launch");
        gemm_launch(s, pl, tb, tb.launch_x2l);
        // L2L from each level's parents, down the tree
        for (const int3& lc : tb.launch_l2l) gemm_launch(s, pl, tb, lc);
    }
    if (m2l_ev) hip_check(hipEventRecord(m2l_ev[1], s), "hipEventRecord");
    if (D >= 2) {
        hipLaunchKernelGGL((k_pw_l2p<T, P>), dim3(pl->nleaf), dim3(64), size_t(3) * jp * sizeof(T), s, d,
                           patches, kappa, jp, L, y, pl->N);
        hip_check(hipGetLastError(), "k_pw_l2p launch");
    }
}

template <typename T>
void pw_dispatch(FmmPlan* pl, hipStream_t s, const Patch64* patches, double kappa, const PwDevTables<T>& tb,
                 T* M, T* L, T* W, T* Wacc, T* Tc, double* y, hipEvent_t* m2l_ev) {
    switch (pl->pw_p) {
        case 9: pw_run<T, 9>(pl, s, patches, kappa, tb, M, L, W, Wacc, Tc, y, m2l_ev); break;
        case 12: pw_run<T, 12>(pl, s, patches, kappa, tb, M, L, W, Wacc, Tc, y, m2l_ev); break;
        case 18: pw_run<T, 18>(pl, s, patches, kappa, tb, M, L, W, Wacc, Tc, y, m2l_ev); break;
        default: throw std::runtime_error("plane-wave order must be 9, 12 or 18");
    }
}

}  // namespace

bool pw_order_supported(int p) { return p == 9 || p == 12 || p == 18; }

void pw_prepare(FmmPlan* pl, double kappa, bool fp64, bool fp32, hipStream_t s) {
    const size_t nexp = size_t(pl->nbox) * 3 * pl->pw_jp;
    // plane waves of every box at levels >= 2 (all levels' X2X run at once)
    const int base = pl->D >= 2 ? pl->lvl_first[2] : pl->nbox;
    const size_t nwb = size_t(pl->nbox - base);
    if (fp64 && !pl->have64) {
        make_pw_tables<double>(pl, kappa, pl->pw64, s);
        const size_t nw = nwb * 3 * pl->pw_nwp;
        pl->pw_M64 = pl->arena->get<double>("pwM64", nexp);
        pl->pw_L64 = pl->arena->get<double>("pwL64", nexp);
        pl->pw_W64 = pl->arena->get<double>("pwW64", std::max<size_t>(nw, 1));
        pl->pw_Wacc64 = pl->arena->get<double>("pwA64", std::max<size_t>(nw, 1));
        pl->pw_T64 = pl->arena->get<double>("pwT64", nexp);
        make_pw_launches<double>(pl, pl->pw64, pl->pw_M64, pl->pw_L64, pl->pw_W64, pl->pw_Wacc64, pl->pw_T64, s);
        pl->have64 = true;
    }
    if (fp32 && !pl->have32) {
        make_pw_tables<float>(pl, kappa, pl->pw32, s);
        const size_t nw = nwb * 3 * pl->pw_nwp;
        pl->pw_M32 = pl->arena->get<float>("pwM32", nexp);
        pl->pw_L32 = pl->arena->get<float>("pwL32", nexp);
        pl->pw_W32 = pl->arena->get<float>("pwW32", std::max<size_t>(nw, 1));
        pl->pw_Wacc32 = pl->arena->get<float>("pwA32", std::max<size_t>(nw, 1));
        pl->pw_T32 = pl->arena->get<float>("pwT32", nexp);
        make_pw_launches<float>(pl, pl->pw32, pl->pw_M32, pl->pw_L32, pl->pw_W32, pl->pw_Wacc32, pl->pw_T32, s);
        pl->have32 = true;
    }
}

void pw_evaluate(FmmPlan* pl, hipStream_t s, int fp32, const Patch64* patches, double kappa, double* y,
                 hipEvent_t* m2l_ev) {
    if (fp32) {
        if (!pl->have32) throw std::runtime_error("plane-wave FP32 tables not prepared");
        pw_dispatch<float>(pl, s, patches, kappa, pl->pw32, pl->pw_M32, pl->pw_L32, pl->pw_W32, pl->pw_Wacc32,
                           pl->pw_T32, y, m2l_ev);
    } else {
        if (!pl->have64) throw std::runtime_error("plane-wave FP64 tables not prepared");
        pw_dispatch<double>(pl, s, patches, kappa, pl->pw64, pl->pw_M64, pl->pw_L64, pl->pw_W64, pl->pw_Wacc64,
                            pl->pw_T64, y, m2l_ev);
    }
}

}  // namespace beepdev
