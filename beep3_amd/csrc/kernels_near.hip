// Near-field kernels for gfx950: singular self terms, explicit-pair
// coefficient precompute and the per-leaf block SpMV.
//
// Integrand: evaluate_local_BEM_kernels (src/bem/bem_kernels.h:134-251) in the
// factored form
//   A += w (G0 - u)
//   B += -w (n.d) (gr3/eps - g1)
//   C +=  w (n0.d) (gr3 - g1/eps)
//   D -= w [ (3 gr5 - p) (n.d)(n0.d) - (gr3 - g1)(n.n0) ]      (kappa != 0)
// with d = y - x, gr3 = 1/(4 pi r^3), g1 = e^{-kr}(1 + k r)/(4 pi r^3),
// p = e^{-kr}(k^2 r^2 + 3 k r + 3)/(4 pi r^5), summed over the quadrature points
// y of the column patch and the qualocation points x of the row patch, D
// finally scaled by 1/eps_col.  Algebraically identical to the reference's
// 3x3 loops over gdd - udd.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "device.hpp"

namespace beepdev {

template <typename T>
struct Math;
// e^{-kappa r} = expk(ek * r) with ek = ek_of(kappa) formed once per kernel
template <>
struct Math<float> {
    // hardware reciprocal square root (1 ulp)
    static __device__ __forceinline__ float rsqrt_(float v) { return __builtin_amdgcn_rsqf(v); }
    // hardware 2^x with the factor log2(e) folded into ek
    static __host__ __device__ __forceinline__ float ek_of(float kappa) {
        return -kappa * 1.44269504088896341f;
    }
    static __device__ __forceinline__ float expk(float v) { return __builtin_amdgcn_exp2f(v); }
};
template <>
struct Math<double> {
    // hardware estimate refined by two Newton steps y <- y + y (1/2 - (v/2) y^2)
    static __device__ __forceinline__ double rsqrt_(double v) { return rsqrt_nr(v); }
    static __host__ __device__ __forceinline__ double ek_of(double kappa) { return -kappa; }
    static __device__ __forceinline__ double expk(double v) { return exp(v); }
};

// Accumulate A..D of one (qual point, quad point) pair into the per-qual-point
// partial sums, without the common factor 1/(4 pi) (applied once per patch
// pair by the caller).  Coincident points (r^2 < 1e-9, bem_kernels.h:185)
// contribute nothing: 1/r is set to 0, which zeroes every term, so the loop
// carries no branch.  KAPPA = (kappa != 0), fixed per launch.
template <typename T, bool KAPPA>
__device__ __forceinline__ void integrand(T kappa, T ek, T inv_eps, T dx, T dy, T dz, T n0x,
                                          T n0y, T n0z, T nx, T ny, T nz, T wt, T& A, T& B,
                                          T& C, T& D) {
    const T r2 = dx * dx + dy * dy + dz * dz;
    const T ir = r2 < T(1e-9) ? T(0) : Math<T>::rsqrt_(r2);
    const T ir2 = ir * ir;
    const T ir3 = ir2 * ir;
    const T nd = nx * dx + ny * dy + nz * dz;
    const T n0d = n0x * dx + n0y * dy + n0z * dz;
    if constexpr (KAPPA) {
        const T r = r2 * ir;
        const T e = Math<T>::expk(ek * r);
        const T u3 = e * (ir3 + kappa * ir2);  // e^{-kr}(1 + kr)/r^3
        A += wt * (ir - e * ir);
        B -= wt * nd * (inv_eps * ir3 - u3);
        C += wt * n0d * (ir3 - u3 * inv_eps);
        const T ir5 = ir3 * ir2;
        const T p = e * (kappa * (T(3) * ir2 * ir2 + kappa * ir3) + T(3) * ir5);
        const T nn0 = nx * n0x + ny * n0y + nz * n0z;
        D -= wt * ((T(3) * ir5 - p) * nd * n0d - (ir3 - u3) * nn0);
    } else {
        // e^{-kr} = 1: A and D vanish (D for kappa = 0 is not computed,
        // bem_kernels.h:210)
        B -= wt * nd * (inv_eps * ir3 - ir3);
        C += wt * n0d * (ir3 - ir3 * inv_eps);
    }
}

// ---------------------------------------------------------------------------
// Coefficient guard (BEEP::sanity_check, beep.cpp:1066-1122, with
// LocalIntegrations::insane, local_integrations.h:111-121): a value is bad if
// |v| > maxv, or not finite (the reference's comparison lets NaN through).
__device__ __forceinline__ bool coef_bad(double v, double maxv) { return !(fabs(v) <= maxv); }
__device__ __forceinline__ bool coef4_bad(const double v[4], double maxv) {
    return coef_bad(v[0], maxv) || coef_bad(v[1], maxv) || coef_bad(v[2], maxv) || coef_bad(v[3], maxv);
}

// Singular self terms (singular_BEM_kernels, bem_kernels.h:18-132) in FP64
// over the galerkin points of each patch, normalised by the summed weights,
// plus the geometric correction (bem_kernels.cpp:123-136; beep.cpp:836-841).
// One wave per patch; lane = source point.  A bad self term is zeroed (the
// reference's rule for an ill-conditioned patch, beep.cpp:1097-1100) and
// counted in *repaired.  guard_only: the wave first checks the stored terms
// and recomputes only bad ones (the guard pass over an existing diagonal).
template <bool KAPPA>
__global__ __launch_bounds__(64) void k_singular(int n, const Patch64* __restrict__ patches,
                                                 const double* __restrict__ gc,
                                                 const int* __restrict__ gal_ptr,
                                                 const Pt64* __restrict__ gal, double kappa,
                                                 int guard_only, double maxv,
                                                 unsigned* __restrict__ repaired,
                                                 double* __restrict__ diag4) {
    const int i = blockIdx.x;
    if (i >= n) return;
    if (guard_only) {
        const double v[4] = {diag4[4 * i], diag4[4 * i + 1], diag4[4 * i + 2], diag4[4 * i + 3]};
        if (!coef4_bad(v, maxv)) return;  // the same for every lane of the wave
    }
    const int p0 = gal_ptr[i], np = gal_ptr[i + 1] - p0;
    const double inv_eps = patches[i].inv_eps;
    double A = 0, B = 0, C = 0, D = 0, wsum = 0;
    for (int s = threadIdx.x; s < np; s += 64) {
        const Pt64 q = gal[p0 + s];
        double a = 0, b = 0, c = 0, d = 0;
        for (int t = 0; t < np; ++t) {
            const Pt64 p = gal[p0 + t];
            integrand<double, KAPPA>(kappa, -kappa, inv_eps, p.x - q.x, p.y - q.y, p.z - q.z,
                                     q.nx, q.ny, q.nz, p.nx, p.ny, p.nz, p.w, a, b, c, d);
        }
        A += a * q.w;
        B += b * q.w;
        C += c * q.w;
        D += d * q.w;
        wsum += q.w;
    }
    for (int off = 32; off > 0; off >>= 1) {
        A += __shfl_down(A, off, 64);
        B += __shfl_down(B, off, 64);
        C += __shfl_down(C, off, 64);
        D += __shfl_down(D, off, 64);
        wsum += __shfl_down(wsum, off, 64);
    }
    if (threadIdx.x == 0) {
        const double e = 1.0 / inv_eps, g = gc[i];
        const double k4 = FOUR_PI_INV / wsum;
        double v[4] = {A * k4, B * k4 + ((1.0 - g) + g / e), C * k4 - (g + (1.0 - g) / e),
                       D * inv_eps * k4};
        if (coef4_bad(v, maxv)) {
            v[0] = v[1] = v[2] = v[3] = 0.0;
            atomicAdd(repaired, 1u);
        } else if (guard_only) {
            atomicAdd(repaired, 1u);
        }
        for (int k = 0; k < 4; ++k) diag4[4 * i + k] = v[k];
    }
}

void launch_singular(hipStream_t s, int n, const Patch64* patches, const double* gc,
                     const int* gal_ptr, const Pt64* gal, double kappa, int guard_only,
                     double maxv, unsigned* repaired, double* diag4) {
    if (n <= 0) return;
    if (kappa != 0.0)
        hipLaunchKernelGGL(k_singular<true>, dim3(n), dim3(64), 0, s, n, patches, gc, gal_ptr, gal,
                           kappa, guard_only, maxv, repaired, diag4);
    else
        hipLaunchKernelGGL(k_singular<false>, dim3(n), dim3(64), 0, s, n, patches, gc, gal_ptr,
                           gal, kappa, guard_only, maxv, repaired, diag4);
    hip_check(hipGetLastError(), "k_singular launch");
}

// ---------------------------------------------------------------------------
// Explicit-pair coefficients.  Leaf L owns an R x C block (R rows = patches
// of the leaf, C columns = its neighbourhood) stored column-major at
// coef_off[L]: entry e = c * R + r.  One thread per entry; a block covers 256
// consecutive entries of one leaf (blk_ptr maps blocks to leaves).  The
// thread integrates the qualocation points of row i against the quadrature
// points of column j; with subtract_1pt it stores the difference to the
// 1-point centroid interaction, which the far-field kernels add back for
// every pair, except for pairs whose packed FMM leaf cells (cellkey, when
// given) are adjacent: the FMM leaf pass leaves those out, so they keep the
// full rule.
// Near-singular pairs: centroids closer than about two mesh edges,
// |c_i - c_j|^2 < 2 (a_i + a_j) (the first ring of neighbours has
// |d|^2 ~ 1.15 a, the second 3.5 a - 4.6 a).  There the kernel differences
// G0 - Gk and their derivatives cancel (D's by a factor 6 / (kappa r)^2),
// so the FP32 coefficient set computes these pairs in FP64 and rounds them.
__device__ __forceinline__ bool near_singular(const Patch64& pi, const Patch64& pj) {
    const double dx = pj.cx - pi.cx, dy = pj.cy - pi.cy, dz = pj.cz - pi.cz;
    return dx * dx + dy * dy + dz * dz < 2.0 * (pi.area + pj.area);
}

__device__ __forceinline__ int find_leaf(const int* __restrict__ blk_ptr, int n_leaves, int b) {
    int lo = 0, hi = n_leaves;  // blk_ptr[lo] <= b < blk_ptr[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (blk_ptr[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

template <typename T, bool KAPPA>
__device__ __forceinline__ void pair_coeffs(const Patch64& pi, const Patch64& pj, int qb, int qn,
                                            const Pt<T>* __restrict__ qual, int pb, int pn,
                                            const Pt<T>* __restrict__ quad, T kappa,
                                            int subtract_1pt, T out[4]) {
    // centroid difference in FP64, then the working precision
    const T dcx = T(pj.cx - pi.cx), dcy = T(pj.cy - pi.cy), dcz = T(pj.cz - pi.cz);
    const T inv_eps = T(pj.inv_eps);
    const T ek = Math<T>::ek_of(kappa);
    T A = 0, B = 0, C = 0, D = 0;
    constexpr int UNROLL = sizeof(T) == 4 ? 4 : 2;  // source points per unrolled step
    for (int a = 0; a < qn; ++a) {
        const Pt<T> q = qual[qb + a];
        T a_ = 0, b_ = 0, c_ = 0, d_ = 0;
        const T ox = dcx - q.x, oy = dcy - q.y, oz = dcz - q.z;
#pragma unroll UNROLL
        for (int k = 0; k < pn; ++k) {
            const Pt<T> p = quad[pb + k];
            integrand<T, KAPPA>(kappa, ek, inv_eps, ox + p.x, oy + p.y, oz + p.z, q.nx, q.ny,
                                q.nz, p.nx, p.ny, p.nz, p.w, a_, b_, c_, d_);
        }
        A += a_ * q.w;
        B += b_ * q.w;
        C += c_ * q.w;
        D += d_ * q.w;
    }
    if (subtract_1pt) {
        T a_ = 0, b_ = 0, c_ = 0, d_ = 0;
        integrand<T, KAPPA>(kappa, ek, inv_eps, dcx, dcy, dcz, T(pi.nx), T(pi.ny), T(pi.nz),
                            T(pj.nx), T(pj.ny), T(pj.nz), T(pj.area), a_, b_, c_, d_);
        A -= a_;
        B -= b_;
        C -= c_;
        D -= d_;
    }
    const T k4 = T(FOUR_PI_INV);
    out[0] = A * k4;
    out[1] = B * k4;
    out[2] = C * k4;
    out[3] = D * inv_eps * k4;
}

// At most 128 VGPRs (FP64) or 102 (FP32): 4 or 5 waves per SIMD.  The FP64
// kernel otherwise takes 136 and runs 3 waves per SIMD, 16 % slower.
template <typename T, bool KAPPA>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 4 : 5) void k_near_coeffs(
    int n_leaves, const int* __restrict__ blk_ptr, const int* __restrict__ leaf_ptr,
    const int* __restrict__ leaf_idx, const int* __restrict__ nb_ptr,
    const int* __restrict__ nb_idx, const int64_t* __restrict__ coef_off,
    const Patch64* __restrict__ patches, const int* __restrict__ qual_ptr,
    const Pt<T>* __restrict__ qual, const int* __restrict__ quad_ptr,
    const Pt<T>* __restrict__ quad, T kappa, int subtract_1pt,
    const unsigned* __restrict__ cellkey, NsList ns, NsList guard, double maxv,
    void* __restrict__ coef) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int L = find_leaf(blk_ptr, n_leaves, b);
    const int lp = leaf_ptr[L], R = leaf_ptr[L + 1] - lp;
    const int np = nb_ptr[L], Cn = nb_ptr[L + 1] - np;
    const int64_t e = int64_t(b - blk_ptr[L]) * 256 + threadIdx.x;
    if (e >= int64_t(R) * Cn) return;
    const int c = int(e / R);
    const int r = int(e - int64_t(c) * R);
    const int i = leaf_idx[lp + r];
    const int j = nb_idx[np + c];
    T out[4] = {0, 0, 0, 0};
    if (i != j) {
        const Patch64 pi = patches[i];
        const Patch64 pj = patches[j];
        const int qb = qual_ptr[i], qn = qual_ptr[i + 1] - qb;
        const int pb = quad_ptr[j], pn = quad_ptr[j + 1] - pb;
        // pairs the FMM leaf pass covers keep the full rule (that pass then
        // skips them)
        const bool sub = subtract_1pt && !(cellkey && fmm_cells_adjacent(cellkey[i], cellkey[j]));
        if (sizeof(T) == 4 && ns.rec && near_singular(pi, pj)) {
            // FP64 pass (k_near_fix64) for this entry; past the list's
            // capacity the entry stays in FP32
            const unsigned slot = atomicAdd(ns.count, 1u);
            if (slot < ns.cap) {
                ns.rec[slot] = NsRec{coef_off[L] + e, i, 2 * j + (sub ? 1 : 0)};
                return;
            }
        }
        pair_coeffs<T, KAPPA>(pi, pj, qb, qn, qual, pb, pn, quad, kappa, sub, out);
        const double v[4] = {double(out[0]), double(out[1]), double(out[2]), double(out[3])};
        if (coef4_bad(v, maxv)) {
            // stored as zero; k_coef_repair recomputes it in FP64
            const unsigned slot = atomicAdd(guard.count, 1u);
            if (slot < guard.cap) guard.rec[slot] = NsRec{coef_off[L] + e, i, 2 * j + (sub ? 1 : 0)};
            out[0] = out[1] = out[2] = out[3] = T(0);
        }
    }
    const int64_t idx = coef_off[L] + e;
    if constexpr (sizeof(T) == 4) {
        reinterpret_cast<float4*>(coef)[idx] = make_float4(out[0], out[1], out[2], out[3]);
    } else {
        double2* c2 = reinterpret_cast<double2*>(coef);
        c2[2 * idx] = make_double2(out[0], out[1]);
        c2[2 * idx + 1] = make_double2(out[2], out[3]);
    }
}

void launch_near_coeffs_fp32(hipStream_t s, int n_blocks, int n_leaves, const int* blk_ptr,
                             const int* leaf_ptr, const int* leaf_idx, const int* nb_ptr,
                             const int* nb_idx, const int64_t* coef_off, const Patch64* patches,
                             const int* qual_ptr, const Pt32* qual, const int* quad_ptr,
                             const Pt32* quad, float kappa, int subtract_1pt,
                             const unsigned* cellkey, const NsList& ns, const NsList& guard,
                             double maxv, float4* coef) {
    if (n_blocks <= 0) return;
    if (ns.rec) launch_set_int(s, reinterpret_cast<int*>(ns.count), 0);
    launch_set_int(s, reinterpret_cast<int*>(guard.count), 0);
    auto k = kappa != 0.0f ? k_near_coeffs<float, true> : k_near_coeffs<float, false>;
    hipLaunchKernelGGL(k, dim3(n_blocks), dim3(256), 0, s, n_leaves, blk_ptr, leaf_ptr, leaf_idx,
                       nb_ptr, nb_idx, coef_off, patches, qual_ptr, qual, quad_ptr, quad, kappa,
                       subtract_1pt, cellkey, ns, guard, maxv, (void*)coef);
    hip_check(hipGetLastError(), "k_near_coeffs<float> launch");
}

// The near-singular entries of an FP32 coefficient set, in FP64 on the FP64
// point sets, rounded to FP32.  The record count is read on the device, so
// the grid is fixed and walks the list with a stride.
template <bool KAPPA>
__global__ __launch_bounds__(256) void k_near_fix64(NsList ns, const Patch64* __restrict__ patches,
                                                   const int* __restrict__ qual_ptr,
                                                   const Pt64* __restrict__ qual,
                                                   const int* __restrict__ quad_ptr,
                                                   const Pt64* __restrict__ quad, double kappa,
                                                   double maxv, unsigned* __restrict__ repaired,
                                                   float4* __restrict__ coef) {
    const unsigned n = min(*ns.count, ns.cap);
    for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const NsRec r = ns.rec[k];
        const int i = r.i, j = r.j2 >> 1;
        const Patch64 pi = patches[i], pj = patches[j];
        const int qb = qual_ptr[i], qn = qual_ptr[i + 1] - qb;
        const int pb = quad_ptr[j], pn = quad_ptr[j + 1] - pb;
        double out[4];
        pair_coeffs<double, KAPPA>(pi, pj, qb, qn, qual, pb, pn, quad, kappa, r.j2 & 1, out);
        const float4 o = make_float4(float(out[0]), float(out[1]), float(out[2]), float(out[3]));
        const double v[4] = {o.x, o.y, o.z, o.w};
        if (coef4_bad(v, maxv)) {
            coef[r.idx] = make_float4(0.f, 0.f, 0.f, 0.f);
            atomicAdd(repaired, 1u);
        } else {
            coef[r.idx] = o;
        }
    }
}

void launch_near_fix64(hipStream_t s, const NsList& ns, const Patch64* patches, const int* qual_ptr,
                       const Pt64* qual, const int* quad_ptr, const Pt64* quad, double kappa,
                       double maxv, unsigned* repaired, float4* coef) {
    if (!ns.rec) return;
    // enough blocks for every CU several times over; each walks the list
    const unsigned blocks = std::max(1u, std::min(2048u, (ns.cap + 255) / 256));
    auto k = kappa != 0.0 ? k_near_fix64<true> : k_near_fix64<false>;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, s, ns, patches, qual_ptr, qual, quad_ptr, quad,
                       kappa, maxv, repaired, coef);
    hip_check(hipGetLastError(), "k_near_fix64 launch");
}

// Guard pass over a stored coefficient set (the sanity check on demand):
// every entry is read once and the bad ones are listed for k_coef_repair.
// Indexing as k_near_coeffs (one thread per entry).
template <typename TC>
__global__ __launch_bounds__(256) void k_coef_scan(
    int n_leaves, const int* __restrict__ blk_ptr, const int* __restrict__ leaf_ptr,
    const int* __restrict__ leaf_idx, const int* __restrict__ nb_ptr,
    const int* __restrict__ nb_idx, const int64_t* __restrict__ coef_off,
    const unsigned* __restrict__ cellkey, int subtract_1pt, const void* __restrict__ coef,
    double maxv, NsList out) {
    const int b = blockIdx.x;
    const int L = find_leaf(blk_ptr, n_leaves, b);
    const int lp = leaf_ptr[L], R = leaf_ptr[L + 1] - lp;
    const int np = nb_ptr[L], Cn = nb_ptr[L + 1] - np;
    const int64_t e = int64_t(b - blk_ptr[L]) * 256 + threadIdx.x;
    if (e >= int64_t(R) * Cn) return;
    const int64_t idx = coef_off[L] + e;
    double v[4];
    if constexpr (sizeof(TC) == 4) {
        const float4 c = reinterpret_cast<const float4*>(coef)[idx];
        v[0] = c.x; v[1] = c.y; v[2] = c.z; v[3] = c.w;
    } else {
        const double2* c2 = reinterpret_cast<const double2*>(coef);
        const double2 a = c2[2 * idx], c = c2[2 * idx + 1];
        v[0] = a.x; v[1] = a.y; v[2] = c.x; v[3] = c.y;
    }
    if (!coef4_bad(v, maxv)) return;
    const int c = int(e / R);
    const int r = int(e - int64_t(c) * R);
    const int i = leaf_idx[lp + r], j = nb_idx[np + c];
    const bool sub = subtract_1pt && !(cellkey && fmm_cells_adjacent(cellkey[i], cellkey[j]));
    const unsigned slot = atomicAdd(out.count, 1u);
    if (slot < out.cap) out.rec[slot] = NsRec{idx, i, 2 * j + (sub ? 1 : 0)};
}

// Recompute the listed entries in FP64 (a self pair, i == j, is zeroed, as
// the reference does for an ill-conditioned patch); an entry still bad is
// zeroed.  Each counts one in *repaired.
template <typename TC, bool KAPPA>
__global__ __launch_bounds__(256) void k_coef_repair(NsList list, const Patch64* __restrict__ patches,
                                                     const int* __restrict__ qual_ptr,
                                                     const Pt64* __restrict__ qual,
                                                     const int* __restrict__ quad_ptr,
                                                     const Pt64* __restrict__ quad, double kappa,
                                                     double maxv, unsigned* __restrict__ repaired,
                                                     void* __restrict__ coef) {
    const unsigned n = min(*list.count, list.cap);
    for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const NsRec r = list.rec[k];
        const int i = r.i, j = r.j2 >> 1;
        double out[4] = {0, 0, 0, 0};
        if (i != j) {
            const Patch64 pi = patches[i], pj = patches[j];
            const int qb = qual_ptr[i], qn = qual_ptr[i + 1] - qb;
            const int pb = quad_ptr[j], pn = quad_ptr[j + 1] - pb;
            pair_coeffs<double, KAPPA>(pi, pj, qb, qn, qual, pb, pn, quad, kappa, r.j2 & 1, out);
            double v[4] = {out[0], out[1], out[2], out[3]};
            if constexpr (sizeof(TC) == 4)
                for (int q = 0; q < 4; ++q) v[q] = double(float(v[q]));
            if (coef4_bad(v, maxv)) out[0] = out[1] = out[2] = out[3] = 0.0;
        }
        if constexpr (sizeof(TC) == 4) {
            reinterpret_cast<float4*>(coef)[r.idx] =
                make_float4(float(out[0]), float(out[1]), float(out[2]), float(out[3]));
        } else {
            double2* c2 = reinterpret_cast<double2*>(coef);
            c2[2 * r.idx] = make_double2(out[0], out[1]);
            c2[2 * r.idx + 1] = make_double2(out[2], out[3]);
        }
        atomicAdd(repaired, 1u);
    }
}

void launch_coef_scan(hipStream_t s, int fp32, int n_blocks, int n_leaves, const int* blk_ptr,
                      const int* leaf_ptr, const int* leaf_idx, const int* nb_ptr, const int* nb_idx,
                      const int64_t* coef_off, const unsigned* cellkey, int subtract_1pt,
                      const void* coef, double maxv, const NsList& out) {
    launch_set_int(s, reinterpret_cast<int*>(out.count), 0);
    if (n_blocks <= 0) return;
    auto k = fp32 ? k_coef_scan<float> : k_coef_scan<double>;
    hipLaunchKernelGGL(k, dim3(n_blocks), dim3(256), 0, s, n_leaves, blk_ptr, leaf_ptr, leaf_idx,
                       nb_ptr, nb_idx, coef_off, cellkey, subtract_1pt, coef, maxv, out);
    hip_check(hipGetLastError(), "k_coef_scan launch");
}

void launch_coef_repair(hipStream_t s, int fp32, const NsList& list, const Patch64* patches,
                        const int* qual_ptr, const Pt64* qual, const int* quad_ptr, const Pt64* quad,
                        double kappa, double maxv, unsigned* repaired, void* coef) {
    const unsigned blocks = std::max(1u, std::min(256u, (list.cap + 255) / 256));
    if (fp32) {
        auto k = kappa != 0.0 ? k_coef_repair<float, true> : k_coef_repair<float, false>;
        hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, s, list, patches, qual_ptr, qual, quad_ptr,
                           quad, kappa, maxv, repaired, coef);
    } else {
        auto k = kappa != 0.0 ? k_coef_repair<double, true> : k_coef_repair<double, false>;
        hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, s, list, patches, qual_ptr, qual, quad_ptr,
                           quad, kappa, maxv, repaired, coef);
    }
    hip_check(hipGetLastError(), "k_coef_repair launch");
}

void launch_near_coeffs_fp64(hipStream_t s, int n_blocks, int n_leaves, const int* blk_ptr,
                             const int* leaf_ptr, const int* leaf_idx, const int* nb_ptr,
                             const int* nb_idx, const int64_t* coef_off, const Patch64* patches,
                             const int* qual_ptr, const Pt64* qual, const int* quad_ptr,
                             const Pt64* quad, double kappa, int subtract_1pt,
                             const unsigned* cellkey, const NsList& guard, double maxv,
                             double2* coef) {
    if (n_blocks <= 0) return;
    launch_set_int(s, reinterpret_cast<int*>(guard.count), 0);
    auto k = kappa != 0.0 ? k_near_coeffs<double, true> : k_near_coeffs<double, false>;
    hipLaunchKernelGGL(k, dim3(n_blocks), dim3(256), 0, s, n_leaves, blk_ptr, leaf_ptr, leaf_idx,
                       nb_ptr, nb_idx, coef_off, patches, qual_ptr, qual, quad_ptr, quad, kappa,
                       subtract_1pt, cellkey, NsList{}, guard, maxv, (void*)coef);
    hip_check(hipGetLastError(), "k_near_coeffs<double> launch");
}

// ---------------------------------------------------------------------------
// Block SpMV over the explicit pairs, plus the diagonal:
//   y_f[i] = sum_j B_ij f_j - A_ij h_j + (B_ii' f_i - A_ii h_i)
//   y_h[i] = sum_j D_ij f_j - C_ij h_j + (D_ii f_i - C_ii' h_i)
// (LocalIntegrations::evaluate_local_contributions, local_integrations.h:123-130).
//
// Pass 1: a block handles one work item = (leaf, group of <= 256 rows, run of
// <= SPMV_CW columns), so blocks carry similar byte counts.  The x values of
// its columns are gathered into LDS; thread (r, s) walks columns s, s+S, ...
// of row r, so each step of the block reads one contiguous run of
// coefficients, U steps in flight per thread.  Partial row sums go to a
// scratch array (no atomics, deterministic).
// Pass 2: a block per row group sums its partials, adds the diagonal and
// writes y.
template <typename TC>
struct Coef4 {
    TC a, b, c, d;
};

typedef float f32x4_nt __attribute__((ext_vector_type(4)));

// Coefficients are read once per mat-vec and exceed every cache.  FP32
// blocks use streaming (non-temporal) loads, measured 3 % faster; for FP64
// blocks they measured 14 % slower, so those use ordinary loads.
template <typename TC>
__device__ __forceinline__ Coef4<TC> load_coef(const void* __restrict__ coef, int64_t idx) {
    if constexpr (sizeof(TC) == 4) {
        const f32x4_nt v = __builtin_nontemporal_load(reinterpret_cast<const f32x4_nt*>(coef) + idx);
        return Coef4<TC>{v.x, v.y, v.z, v.w};
    } else {
        const double2* c2 = reinterpret_cast<const double2*>(coef);
        const double2 v0 = c2[2 * idx], v1 = c2[2 * idx + 1];
        return Coef4<TC>{v0.x, v0.y, v1.x, v1.y};
    }
}

// NC copies of the same near field (instances of one mesh in a pose): copy q
// reads x and writes y at patch offset xoff[q] (f and h halves still n
// apart), so the coefficients are read once for all copies.  NC = 1 is the
// plain mat-vec.  Partials of copy q: part + q * n_work * 512.
struct CopyOffsets {
    int off[2];
};

template <typename TC, typename TA, int U, int NC>
__global__ __launch_bounds__(256) void k_near_spmv_part(
    const int4* __restrict__ work, int n_work, const int* __restrict__ leaf_ptr,
    const int* __restrict__ nb_ptr, const int* __restrict__ nb_idx,
    const int64_t* __restrict__ coef_off, const void* __restrict__ coef,
    const double* __restrict__ x, CopyOffsets co, int n, TA* __restrict__ part) {
    __shared__ TA xs[NC][2][SPMV_CW];
    __shared__ TA red[NC][2][256];
    // with fewer blocks than work items (a residency cap that leaves room on
    // every CU for the far-field kernels of the other stream) a block walks
    // the work items with a grid stride
    for (int wi = blockIdx.x; wi < n_work; wi += gridDim.x) {
    __syncthreads();  // xs / red of the previous item consumed
    const int4 w = work[wi];  // leaf, first row, first column, columns
    const int L = w.x, row0 = w.y, col0 = w.z, nc = w.w;
    const int R = leaf_ptr[L + 1] - leaf_ptr[L];
    const int np = nb_ptr[L];
    const int Rb = min(256, R - row0);
    const int S = 256 / Rb;
    const int t = threadIdx.x;
    const int r = t % Rb, s = t / Rb;
    const bool active = t < Rb * S;
    for (int k = t; k < nc; k += 256) {
        const int j = nb_idx[np + col0 + k];
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            xs[q][0][k] = TA(x[co.off[q] + j]);
            xs[q][1][k] = TA(x[co.off[q] + n + j]);
        }
    }
    __syncthreads();
    TA af[NC], ah[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) af[q] = ah[q] = TA(0);
    if (active) {
        const int64_t base = coef_off[L] + row0 + r + int64_t(col0) * R;
        int c = s;
        for (; c + (U - 1) * S < nc; c += U * S) {
            Coef4<TC> v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = load_coef<TC>(coef, base + int64_t(c + u * S) * R);
            // all U loads are issued before the first use (the scheduler
            // would otherwise interleave them with the sums and keep only a
            // few in flight)
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < U; ++u) {
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    const TA xf = xs[q][0][c + u * S], xh = xs[q][1][c + u * S];
                    af[q] += TA(v[u].b) * xf - TA(v[u].a) * xh;
                    ah[q] += TA(v[u].d) * xf - TA(v[u].c) * xh;
                }
            }
        }
        for (; c < nc; c += S) {
            const Coef4<TC> v = load_coef<TC>(coef, base + int64_t(c) * R);
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                const TA xf = xs[q][0][c], xh = xs[q][1][c];
                af[q] += TA(v.b) * xf - TA(v.a) * xh;
                ah[q] += TA(v.d) * xf - TA(v.c) * xh;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        red[q][0][t] = af[q];
        red[q][1][t] = ah[q];
    }
    __syncthreads();
    if (active && s == 0) {
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            TA f = af[q], h = ah[q];
            for (int k = 1; k < S; ++k) {
                f += red[q][0][k * Rb + r];
                h += red[q][1][k * Rb + r];
            }
            TA* pq = part + size_t(q) * n_work * 512;
            pq[2 * (size_t(wi) * 256 + r)] = f;
            pq[2 * (size_t(wi) * 256 + r) + 1] = h;
        }
    }
    }
}

// grid (row groups, copies); a work item's partials are pstride rows apart
// (256 for the SpMV, FLY_RB for the on-the-fly pass)
template <typename TA>
__global__ __launch_bounds__(256) void k_near_spmv_sum(
    const int4* __restrict__ groups, const int* __restrict__ group_nblk,
    const int* __restrict__ leaf_ptr, const int* __restrict__ leaf_idx,
    const double* __restrict__ diag4, const double* __restrict__ x, CopyOffsets co, int n, int n_work,
    int pstride, const TA* __restrict__ part, double* __restrict__ y) {
    const int4 g = groups[blockIdx.x];  // leaf, first row, rows, first work block
    const int nb = group_nblk[blockIdx.x];
    const int r = threadIdx.x;
    if (r >= g.z) return;
    const int cq = blockIdx.y, off = co.off[cq];
    const TA* pq = part + size_t(cq) * n_work * 2 * pstride;
    TA af = 0, ah = 0;
    for (int k = 0; k < nb; ++k) {
        const size_t q = 2 * (size_t(g.w + k) * pstride + r);
        af += pq[q];
        ah += pq[q + 1];
    }
    const int i = leaf_idx[leaf_ptr[g.x] + g.y + r];
    const double* d = diag4 + 4 * size_t(i);
    const double xf = x[off + i], xh = x[off + n + i];
    y[off + i] = double(af) + (d[1] * xf - d[0] * xh);
    y[off + n + i] = double(ah) + (d[3] * xf - d[2] * xh);
}

// Loads in flight per thread: U = 4 * depth (FP64 coefficients) or 8 * depth
// (FP32); the column order each thread sums in is the same for every depth,
// so y does not depend on it.
template <typename TC, typename TA, int U>
static void spmv_part(hipStream_t s, int grid, const SpmvPlan& p, const int* leaf_ptr,
                      const int* nb_ptr, const int* nb_idx, const int64_t* coef_off,
                      const void* coef, const double* x, int ncopies, CopyOffsets co, int n,
                      TA* part) {
    if (ncopies == 2)
        hipLaunchKernelGGL((k_near_spmv_part<TC, TA, U, 2>), dim3(grid), dim3(256), 0, s, p.work,
                           p.n_work, leaf_ptr, nb_ptr, nb_idx, coef_off, coef, x, co, n, part);
    else
        hipLaunchKernelGGL((k_near_spmv_part<TC, TA, U, 1>), dim3(grid), dim3(256), 0, s, p.work,
                           p.n_work, leaf_ptr, nb_ptr, nb_idx, coef_off, coef, x, co, n, part);
}

void launch_near_spmv(hipStream_t s, int mode, int n, const SpmvPlan& p, const int* leaf_ptr,
                      const int* leaf_idx, const int* nb_ptr, const int* nb_idx,
                      const int64_t* coef_off, const float4* coef32, const double2* coef64,
                      const double* diag4, const double* x, void* part, double* y,
                      int max_blocks, int depth, int ncopies, const int* copy_offset) {
    if (p.n_work <= 0) return;
    if (depth != 1 && depth != 2 && depth != 4) throw std::runtime_error("SpMV depth must be 1, 2 or 4");
    if (ncopies != 1 && ncopies != 2) throw std::runtime_error("the SpMV takes 1 or 2 copies");
    CopyOffsets co{{0, 0}};
    for (int q = 0; q < ncopies; ++q) co.off[q] = copy_offset ? copy_offset[q] : 0;
    const int grid = max_blocks > 0 ? std::min(p.n_work, max_blocks) : p.n_work;
    const int d = depth == 1 ? 0 : (depth == 2 ? 1 : 2);
    const dim3 sgrid(unsigned(p.n_groups), unsigned(ncopies));
    if (mode == 0) {
        using F = void (*)(hipStream_t, int, const SpmvPlan&, const int*, const int*, const int*,
                           const int64_t*, const void*, const double*, int, CopyOffsets, int, double*);
        constexpr F f[3] = {spmv_part<double, double, 4>, spmv_part<double, double, 8>,
                            spmv_part<double, double, 16>};
        f[d](s, grid, p, leaf_ptr, nb_ptr, nb_idx, coef_off, coef64, x, ncopies, co, n, (double*)part);
        hipLaunchKernelGGL((k_near_spmv_sum<double>), sgrid, dim3(256), 0, s, p.groups, p.group_nblk,
                           leaf_ptr, leaf_idx, diag4, x, co, n, p.n_work, 256, (const double*)part, y);
    } else if (mode == 1) {
        using F = void (*)(hipStream_t, int, const SpmvPlan&, const int*, const int*, const int*,
                           const int64_t*, const void*, const double*, int, CopyOffsets, int, float*);
        constexpr F f[3] = {spmv_part<float, float, 8>, spmv_part<float, float, 16>,
                            spmv_part<float, float, 32>};
        f[d](s, grid, p, leaf_ptr, nb_ptr, nb_idx, coef_off, coef32, x, ncopies, co, n, (float*)part);
        hipLaunchKernelGGL((k_near_spmv_sum<float>), sgrid, dim3(256), 0, s, p.groups, p.group_nblk,
                           leaf_ptr, leaf_idx, diag4, x, co, n, p.n_work, 256, (const float*)part, y);
    } else {
        using F = void (*)(hipStream_t, int, const SpmvPlan&, const int*, const int*, const int*,
                           const int64_t*, const void*, const double*, int, CopyOffsets, int, double*);
        constexpr F f[3] = {spmv_part<float, double, 8>, spmv_part<float, double, 16>,
                            spmv_part<float, double, 32>};
        f[d](s, grid, p, leaf_ptr, nb_ptr, nb_idx, coef_off, coef32, x, ncopies, co, n, (double*)part);
        hipLaunchKernelGGL((k_near_spmv_sum<double>), sgrid, dim3(256), 0, s, p.groups, p.group_nblk,
                           leaf_ptr, leaf_idx, diag4, x, co, n, p.n_work, 256, (const double*)part, y);
    }
    hip_check(hipGetLastError(), "k_near_spmv launch");
}

// ---------------------------------------------------------------------------
// Near field on the fly: the semantics of BEM_kernels_instant
// (src/bem/opencl_bem_kernels.cl:237-323, driven per mat-vec by
// BEM_OnDemand_Resources, src/bem/opencl_bem.cpp:346-747): no coefficients are
// stored; every mat-vec integrates each explicit pair and applies it at once.
// The coefficient of a pair is the one the precompute stores for it
// (k_near_coeffs and its FP64 passes): the full rule less the 1-point
// interaction unless the FMM leaf pass covers the pair; in FP32, a
// near-singular pair (when ns64) and a bad entry are integrated in FP64 and
// rounded; an entry still bad is zero.
//
// Work items as the SpMV's (leaf, row group, run of <= SPMV_CW columns), but
// with row groups of <= FLY_RB rows, so that FLY_RB * S of the 256 threads
// are busy.  Thread (r, s) integrates the pairs of row r in columns s, s+S,
// ... and sums their products in that order; the block adds its column lanes
// in order and writes one partial per row (k_near_spmv_sum then adds the
// partials, the diagonal, and writes y): deterministic.
// The FP64 integration of an FP32 pass's near-singular or bad entry, rounded
// to FP32 (zero if still bad), out of line: rare, and inlined it would take
// registers from the FP32 loop.
// It takes indices, not patch records, so the caller keeps nothing on the
// stack for the call.
template <bool KAPPA>
__device__ __attribute__((noinline)) float4 fly_redo64(int i, int j, int sub, const Patch64* __restrict__ patches,
                                                       const int* __restrict__ qual_ptr,
                                                       const Pt64* __restrict__ qual64,
                                                       const int* __restrict__ quad_ptr,
                                                       const Pt64* __restrict__ quad64, double kappa,
                                                       double maxv) {
    const Patch64 pi = patches[i], pj = patches[j];
    const int qb = qual_ptr[i], qn = qual_ptr[i + 1] - qb;
    const int pb = quad_ptr[j], pn = quad_ptr[j + 1] - pb;
    double o[4];
    pair_coeffs<double, KAPPA>(pi, pj, qb, qn, qual64, pb, pn, quad64, kappa, sub, o);
    const float4 v = make_float4(float(o[0]), float(o[1]), float(o[2]), float(o[3]));
    const double d[4] = {v.x, v.y, v.z, v.w};
    return coef4_bad(d, maxv) ? make_float4(0.f, 0.f, 0.f, 0.f) : v;
}

template <typename T, typename TA, bool KAPPA>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 4 : 5) void k_near_fly_part(
    const int4* __restrict__ work, int n_work, const int* __restrict__ leaf_ptr,
    const int* __restrict__ leaf_idx, const int* __restrict__ nb_ptr, const int* __restrict__ nb_idx,
    const Patch64* __restrict__ patches, const int* __restrict__ qual_ptr,
    const Pt<T>* __restrict__ qual, const int* __restrict__ quad_ptr, const Pt<T>* __restrict__ quad,
    const Pt64* __restrict__ qual64, const Pt64* __restrict__ quad64, double kappa,
    const unsigned* __restrict__ cellkey, int ns64, double maxv, const double* __restrict__ x, int n,
    TA* __restrict__ part) {
    // FP32 pass: the near-singular pairs of a thread (up to DEF of them) are
    // set aside (column and leaf-pass bit) and integrated in FP64 by the
    // whole block after the loop, so a wave does not run the FP64 path for a
    // few of its lanes; each thread then adds its own in the order found.
    constexpr int DEF = sizeof(T) == 4 ? 4 : 1;
    __shared__ TA xs[2][SPMV_CW];
    __shared__ TA red[2][256];
    __shared__ int dcode[256 * DEF];
    __shared__ TA dprod[2][256 * DEF];
    const int wi = xcd_remap(blockIdx.x, gridDim.x);
    const int4 w = work[wi];  // leaf, first row, first column, columns
    const int L = w.x, row0 = w.y, col0 = w.z, nc = w.w;
    const int lp = leaf_ptr[L];
    const int R = leaf_ptr[L + 1] - lp;
    const int np = nb_ptr[L];
    const int Rb = min(FLY_RB, R - row0);
    const int S = 256 / Rb;
    const int t = threadIdx.x;
    const int r = t % Rb, s = t / Rb;
    const bool active = t < Rb * S;
    for (int k = t; k < nc; k += 256) {
        const int j = nb_idx[np + col0 + k];
        xs[0][k] = TA(x[j]);
        xs[1][k] = TA(x[n + j]);
    }
    __syncthreads();
    TA af = 0, ah = 0;
    int nd = 0;  // pairs this thread set aside
    if (active) {
        const int i = leaf_idx[lp + row0 + r];
        const Patch64 pi = patches[i];
        const int qb = qual_ptr[i], qn = qual_ptr[i + 1] - qb;
        const unsigned ki = cellkey ? cellkey[i] : 0u;
        for (int c = s; c < nc; c += S) {
            const int j = nb_idx[np + col0 + c];
            if (j == i) continue;  // the self pair is the diagonal's (stored as zero)
            const Patch64 pj = patches[j];
            const int pb = quad_ptr[j], pn = quad_ptr[j + 1] - pb;
            const int sub = !(cellkey && fmm_cells_adjacent(ki, cellkey[j]));
            T v[4];
            bool redo = false;
            if constexpr (sizeof(T) == 4) {
                redo = ns64 && near_singular(pi, pj);
                if (redo && nd < DEF) {
                    dcode[t * DEF + nd] = 2 * c + sub;
                    ++nd;
                    continue;
                }
            }
            if (!redo) {
                pair_coeffs<T, KAPPA>(pi, pj, qb, qn, qual, pb, pn, quad, T(kappa), sub, v);
                const double d[4] = {double(v[0]), double(v[1]), double(v[2]), double(v[3])};
                if (coef4_bad(d, maxv)) {
                    // FP64 recomputes what FP64 already gave: zero at once
                    if constexpr (sizeof(T) == 8) v[0] = v[1] = v[2] = v[3] = T(0);
                    else redo = true;
                }
            }
            if constexpr (sizeof(T) == 4) {
                if (redo) {
                    const float4 o =
                        fly_redo64<KAPPA>(i, j, sub, patches, qual_ptr, qual64, quad_ptr, quad64, kappa, maxv);
                    v[0] = o.x;
                    v[1] = o.y;
                    v[2] = o.z;
                    v[3] = o.w;
                }
            }
            const TA xf = xs[0][c], xh = xs[1][c];
            af += TA(v[1]) * xf - TA(v[0]) * xh;
            ah += TA(v[3]) * xf - TA(v[2]) * xh;
        }
    }
    if constexpr (sizeof(T) == 4) {
        // the set-aside pairs, spread over the block: entry e belongs to
        // thread e / DEF
        for (int k = nd; k < DEF; ++k) dcode[t * DEF + k] = -1;
        __syncthreads();
        for (int e = t; e < 256 * DEF; e += 256) {
            const int code = dcode[e];
            if (code < 0) continue;
            const int owner = e / DEF;
            const int c = code >> 1;
            const int i = leaf_idx[lp + row0 + owner % Rb];
            const int j = nb_idx[np + col0 + c];
            const float4 o =
                fly_redo64<KAPPA>(i, j, code & 1, patches, qual_ptr, qual64, quad_ptr, quad64, kappa, maxv);
            const TA xf = xs[0][c], xh = xs[1][c];
            dprod[0][e] = TA(o.y) * xf - TA(o.x) * xh;
            dprod[1][e] = TA(o.w) * xf - TA(o.z) * xh;
        }
        __syncthreads();
        for (int k = 0; k < nd; ++k) {
            af += dprod[0][t * DEF + k];
            ah += dprod[1][t * DEF + k];
        }
    }
    red[0][t] = af;
    red[1][t] = ah;
    __syncthreads();
    if (active && s == 0) {
        TA f = af, h = ah;
        for (int k = 1; k < S; ++k) {
            f += red[0][k * Rb + r];
            h += red[1][k * Rb + r];
        }
        part[2 * (size_t(wi) * FLY_RB + r)] = f;
        part[2 * (size_t(wi) * FLY_RB + r) + 1] = h;
    }
}

template <typename T, typename TA>
static void fly_launch(hipStream_t s, const SpmvPlan& p, const int* leaf_ptr, const int* leaf_idx,
                       const int* nb_ptr, const int* nb_idx, const Patch64* patches, const int* qual_ptr,
                       const Pt<T>* qual, const int* quad_ptr, const Pt<T>* quad, const Pt64* qual64,
                       const Pt64* quad64, double kappa, const unsigned* cellkey, int ns64, double maxv,
                       const double* diag4, const double* x, int n, TA* part, double* y) {
    auto k = kappa != 0.0 ? k_near_fly_part<T, TA, true> : k_near_fly_part<T, TA, false>;
    hipLaunchKernelGGL(k, dim3(unsigned(p.n_work)), dim3(256), 0, s, p.work, p.n_work, leaf_ptr, leaf_idx,
                       nb_ptr, nb_idx, patches, qual_ptr, qual, quad_ptr, quad, qual64, quad64, kappa,
                       cellkey, ns64, maxv, x, n, part);
    hip_check(hipGetLastError(), "k_near_fly_part launch");
    const CopyOffsets co{{0, 0}};
    hipLaunchKernelGGL((k_near_spmv_sum<TA>), dim3(unsigned(p.n_groups), 1u), dim3(256), 0, s, p.groups,
                       p.group_nblk, leaf_ptr, leaf_idx, diag4, x, co, n, p.n_work, FLY_RB,
                       (const TA*)part, y);
    hip_check(hipGetLastError(), "k_near_spmv_sum launch");
}

void launch_near_fly(hipStream_t s, int mode, int n, const SpmvPlan& p, const int* leaf_ptr,
                     const int* leaf_idx, const int* nb_ptr, const int* nb_idx, const Patch64* patches,
                     const int* qual_ptr, const Pt32* qual32, const Pt64* qual64, const int* quad_ptr,
                     const Pt32* quad32, const Pt64* quad64, double kappa, const unsigned* cellkey,
                     int ns64, double maxv, const double* diag4, const double* x, void* part, double* y) {
    if (p.n_work <= 0) return;
    if (mode == 0)
        fly_launch<double, double>(s, p, leaf_ptr, leaf_idx, nb_ptr, nb_idx, patches, qual_ptr, qual64,
                                   quad_ptr, quad64, qual64, quad64, kappa, cellkey, 0, maxv, diag4, x, n,
                                   (double*)part, y);
    else if (mode == 1)
        fly_launch<float, float>(s, p, leaf_ptr, leaf_idx, nb_ptr, nb_idx, patches, qual_ptr, qual32,
                                 quad_ptr, quad32, qual64, quad64, kappa, cellkey, ns64, maxv, diag4, x,
                                 n, (float*)part, y);
    else
        fly_launch<float, double>(s, p, leaf_ptr, leaf_idx, nb_ptr, nb_idx, patches, qual_ptr, qual32,
                                  quad_ptr, quad32, qual64, quad64, kappa, cellkey, ns64, maxv, diag4, x,
                                  n, (double*)part, y);
}

// ---------------------------------------------------------------------------
// Explicit pairs listed one by one (the near pairs between instances of a
// pose, whose near field otherwise comes from a shared per-mesh context):
// coefficients computed in FP64 (full rule less the 1-point interaction, as
// every explicit pair with the direct far field), stored in FP64 and
// rounded to FP32; bad values zeroed and counted, as the guard does.
template <bool KAPPA>
__global__ __launch_bounds__(256) void k_pair_coeffs(long long n_pairs, const int* __restrict__ pi_,
                                                    const int* __restrict__ pj_,
                                                    const Patch64* __restrict__ patches,
                                                    const int* __restrict__ qual_ptr,
                                                    const Pt64* __restrict__ qual,
                                                    const int* __restrict__ quad_ptr,
                                                    const Pt64* __restrict__ quad, double kappa,
                                                    double maxv, unsigned* __restrict__ repaired,
                                                    float4* __restrict__ c32, double2* __restrict__ c64) {
    const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (k >= n_pairs) return;
    const int i = pi_[k], j = pj_[k];
    double out[4] = {0, 0, 0, 0};
    if (i != j) {
        const Patch64 a = patches[i], b = patches[j];
        const int qb = qual_ptr[i], qn = qual_ptr[i + 1] - qb;
        const int pb = quad_ptr[j], pn = quad_ptr[j + 1] - pb;
        pair_coeffs<double, KAPPA>(a, b, qb, qn, qual, pb, pn, quad, kappa, 1, out);
        if (coef4_bad(out, maxv)) {
            out[0] = out[1] = out[2] = out[3] = 0.0;
            atomicAdd(repaired, 1u);
        }
    }
    c64[2 * k] = make_double2(out[0], out[1]);
    c64[2 * k + 1] = make_double2(out[2], out[3]);
    c32[k] = make_float4(float(out[0]), float(out[1]), float(out[2]), float(out[3]));
}

void launch_pair_coeffs(hipStream_t s, long long n_pairs, const int* pi, const int* pj,
                        const Patch64* patches, const int* qual_ptr, const Pt64* qual, const int* quad_ptr,
                        const Pt64* quad, double kappa, double maxv, unsigned* repaired, float4* c32,
                        double2* c64) {
    if (n_pairs <= 0) return;
    auto k = kappa != 0.0 ? k_pair_coeffs<true> : k_pair_coeffs<false>;
    hipLaunchKernelGGL(k, dim3(unsigned((n_pairs + 255) / 256)), dim3(256), 0, s, n_pairs, pi, pj, patches,
                       qual_ptr, qual, quad_ptr, quad, kappa, maxv, repaired, c32, c64);
    hip_check(hipGetLastError(), "k_pair_coeffs launch");
}

// Pairs between copies closer than rc (the inter-instance pairs of a pose),
// found on the device: one thread per row i, testing every patch of every
// other copy whose bounding sphere (centre, radius) comes within rc of
// patch i, in index order.  The block stages the other copy's centroids
// through LDS, 256 at a time (a chunk), and skips a copy or a chunk none of
// its rows can reach (chunk_sph: the bounding sphere of every chunk of 256
// consecutive patches of each copy; the test is conservative, so the pairs
// found are the brute-force ones).  Pass 1 (cols null) counts per row; pass
// 2 writes each row's columns at ptr[i] and the row of every pair.
__global__ __launch_bounds__(256) void k_cut_pairs(int N, int ncopies, int ns, const int* __restrict__ copy_off,
                                                  const double4* __restrict__ sphere,
                                                  const double4* __restrict__ chunk_sph,
                                                  const Patch64* __restrict__ patches, double rc,
                                                  const int* __restrict__ ptr, int* __restrict__ count,
                                                  int* __restrict__ cols, int* __restrict__ rows_of) {
    __shared__ double4 tile[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool row = i < N;
    const double rc2 = rc * rc;
    double cx = 0, cy = 0, cz = 0;
    int own = -1;
    if (row) {
        cx = patches[i].cx;
        cy = patches[i].cy;
        cz = patches[i].cz;
        for (int q = 0; q < ncopies; ++q)
            if (i >= copy_off[q] && i < copy_off[q] + ns) own = q;
    }
    int n = 0;
    const int o = (row && cols) ? ptr[i] : 0;
    for (int q = 0; q < ncopies; ++q) {
        bool need = false;
        if (row && q != own) {
            const double4 sp = sphere[q];
            const double sx = cx - sp.x, sy = cy - sp.y, sz = cz - sp.z;
            const double reach = sp.w + rc;
            need = sx * sx + sy * sy + sz * sz <= reach * reach;
        }
        if (!__syncthreads_or(need)) continue;  // uniform over the block
        const int j0 = copy_off[q];
        const int nch = (ns + 255) / 256;
        for (int t0 = 0; t0 < ns; t0 += 256) {
            bool need_t = false;
            if (need) {
                const double4 cs = chunk_sph[q * nch + t0 / 256];
                const double sx = cx - cs.x, sy = cy - cs.y, sz = cz - cs.z;
                const double reach = cs.w + rc;
                need_t = sx * sx + sy * sy + sz * sz <= reach * reach * (1.0 + 1e-9);
            }
            // a barrier too: the previous tile is consumed
            if (!__syncthreads_or(need_t)) continue;
            if (t0 + int(threadIdx.x) < ns) {
                const Patch64& p = patches[j0 + t0 + threadIdx.x];
                tile[threadIdx.x] = make_double4(p.cx, p.cy, p.cz, 0.0);
            }
            __syncthreads();
            if (need_t) {
                const int m = min(256, ns - t0);
                for (int k = 0; k < m; ++k) {
                    const double4 pj = tile[k];
                    const double dx = pj.x - cx, dy = pj.y - cy, dz = pj.z - cz;
                    if (dx * dx + dy * dy + dz * dz < rc2) {
                        if (cols) {
                            cols[o + n] = j0 + t0 + k;
                            rows_of[o + n] = i;
                        }
                        ++n;
                    }
                }
            }
        }
    }
    if (row && !cols) count[i] = n;
}

__global__ void k_iota(int n, int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i;
}

void launch_cut_pairs(hipStream_t s, int N, int ncopies, int ns, const int* copy_off, const double4* sphere,
                      const double4* chunk_sph, const Patch64* patches, double rc, const int* ptr, int* count,
                      int* cols, int* rows_of) {
    if (N <= 0) return;
    hipLaunchKernelGGL(k_cut_pairs, dim3((N + 255) / 256), dim3(256), 0, s, N, ncopies, ns, copy_off, sphere,
                       chunk_sph, patches, rc, ptr, count, cols, rows_of);
    hip_check(hipGetLastError(), "k_cut_pairs launch");
}

size_t exclusive_scan_temp_bytes(int n) {
    size_t b = 0;
    hip_check(hipcub::DeviceScan::ExclusiveSum(nullptr, b, static_cast<const int*>(nullptr),
                                               static_cast<int*>(nullptr), n),
              "hipcub scan size");
    return b;
}

void launch_exclusive_scan(hipStream_t s, void* tmp, size_t tmp_bytes, const int* in, int* out, int n) {
    hip_check(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, n, s), "hipcub scan");
}

void launch_iota(hipStream_t s, int n, int* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, n, out);
    hip_check(hipGetLastError(), "k_iota launch");
}

// y += the listed pairs' products: one wave per row that has pairs, lanes
// over its pairs (sorted by column), then a fixed-order shuffle sum.
template <typename TC, typename TA>
__global__ __launch_bounds__(256) void k_pair_spmv(int n, int n_rows, const int* __restrict__ rows,
                                                  const int* __restrict__ ptr, const int* __restrict__ col,
                                                  const void* __restrict__ coef, const double* __restrict__ x,
                                                  double* __restrict__ y) {
    const int wv = (blockIdx.x * 256 + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wv >= n_rows) return;
    const int p0 = ptr[wv], p1 = ptr[wv + 1];
    TA af = 0, ah = 0;
    for (int p = p0 + lane; p < p1; p += 64) {
        const Coef4<TC> v = load_coef<TC>(coef, p);
        const int j = col[p];
        const TA xf = TA(x[j]), xh = TA(x[n + j]);
        af += TA(v.b) * xf - TA(v.a) * xh;
        ah += TA(v.d) * xf - TA(v.c) * xh;
    }
    for (int off = 32; off > 0; off >>= 1) {
        af += __shfl_down(af, off, 64);
        ah += __shfl_down(ah, off, 64);
    }
    if (lane == 0) {
        const int i = rows[wv];
        y[i] += double(af);
        y[n + i] += double(ah);
    }
}

void launch_pair_spmv(hipStream_t s, int mode, int n, int n_rows, const int* rows, const int* ptr,
                      const int* col, const float4* c32, const double2* c64, const double* x, double* y) {
    if (n_rows <= 0) return;
    const dim3 grid(unsigned((n_rows + 3) / 4));
    if (mode == 0)
        hipLaunchKernelGGL((k_pair_spmv<double, double>), grid, dim3(256), 0, s, n, n_rows, rows, ptr, col,
                           (const void*)c64, x, y);
    else if (mode == 1)
        hipLaunchKernelGGL((k_pair_spmv<float, float>), grid, dim3(256), 0, s, n, n_rows, rows, ptr, col,
                           (const void*)c32, x, y);
    else
        hipLaunchKernelGGL((k_pair_spmv<float, double>), grid, dim3(256), 0, s, n, n_rows, rows, ptr, col,
                           (const void*)c32, x, y);
    hip_check(hipGetLastError(), "k_pair_spmv launch");
}

}  // namespace beepdev
