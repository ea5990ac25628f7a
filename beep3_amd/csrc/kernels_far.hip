// Far field of the BEM mat-vec: the 1-point centroid interaction of every
// patch pair (i != j), which the reference evaluates with its 12-component
// Yukawa/Laplace FMM (fmm_bem_hybrid.h:184-642; charges node_patch.h:339-394;
// L2P combination bem_kernels.cpp:12-68).  The explicit-pair coefficients
// carry the difference between the full rules and this interaction, so the
// far field here is summed over all pairs.
//
// The 12 reference components reduce to three potentials of monopole +
// dipole sources (a = area, n = unit normal, ie = 1/eps of the source):
//   Laplace  phi_L : monopole -a h,   dipole  ie a f n
//   Yukawa   phi_1 : monopole  a h,   dipole     -a f n
//   Yukawa   phi_2 : monopole ie a h, dipole  -ie a f n
//   f_i += phi_L(x_i) + phi_1(x_i),   h_i += n_i . grad(phi_L + phi_2)(x_i)
// which per pair, with d = y_j - x_i, s_f = a f, s_h = a h, dd = (s_f n_j).d,
// nd = n_i.d, ndv = n_i.(s_f n_j), gives (times 1/4pi)
//   f += s_h (u - G) + dd (g1 - ie gL1)
//   h += -s_h nd gL1 + ie [ ndv (gL1 - g1) - dd nd (gL2 - g2) + s_h nd g1 ]
// with G = 1/r, u = e/r, gL1 = 1/r^3, g1 = e (1 + k r)/r^3, gL2 = 3/r^5,
// g2 = e (k^2 r^2 + 3 k r + 3)/r^5, e = exp(-k r).
#include <algorithm>
#include <stdexcept>

#include "device.hpp"

namespace beepdev {

__global__ void k_far_pack(int n, const Patch64* __restrict__ patches, const int* __restrict__ order,
                           const double* __restrict__ x, Src64* __restrict__ src) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int j = order[t];
    const Patch64 p = patches[j];
    const double sf = p.area * x[j], sh = p.area * x[n + j];
    src[t] = Src64{p.cx, p.cy, p.cz, p.inv_eps, sf * p.nx, sf * p.ny, sf * p.nz, sh};
}

void launch_far_pack(hipStream_t s, int n, const Patch64* patches, const int* order,
                     const double* x, Src64* src) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_far_pack, dim3((n + 255) / 256), dim3(256), 0, s, n, patches, order, x,
                       src);
    hip_check(hipGetLastError(), "k_far_pack launch");
}

template <typename T>
struct FarMath;
template <>
struct FarMath<float> {
    // hardware reciprocal square root (1 ulp)
    static __device__ __forceinline__ float rsqrt_(float v) { return __builtin_amdgcn_rsqf(v); }
    static __device__ __forceinline__ float exp_(float v) { return __expf(v); }
};
template <>
struct FarMath<double> {
    // hardware estimate refined by two Newton steps y <- y + y (1/2 - (v/2) y^2)
    static __device__ __forceinline__ double rsqrt_(double v) {
        double y = __builtin_amdgcn_rsq(v);
        const double h = 0.5 * v;
        y = fma(y, fma(-h * y, y, 0.5), y);
        return fma(y, fma(-h * y, y, 0.5), y);
    }
    static __device__ __forceinline__ double exp_(double v) { return exp(v); }
};

template <typename T>
struct alignas(16) SrcT {
    T x, y, z, ie;
    T dx, dy, dz, sh;
};

// Targets and sources both in leaf order (spatially compact runs); positions
// are taken relative to the first target of the block, formed in FP64, so
// that nearby pairs keep their relative accuracy in FP32.  A block covers 256
// targets and one of n_split runs of source tiles (grid.y), so small
// problems still fill the chip; with n_split > 1 each run's partial sums go
// to part[run][2n] and k_far_direct_sum adds them in run order.
constexpr int FAR_TILE = 256;

template <typename T>
__global__ __launch_bounds__(256) void k_far_direct(int n, int n_split,
                                                    const int* __restrict__ order,
                                                    const Patch64* __restrict__ patches,
                                                    const Src64* __restrict__ src, T kappa,
                                                    double* __restrict__ y,
                                                    double* __restrict__ part) {
    __shared__ SrcT<T> tile[FAR_TILE];
    __shared__ int tile_idx[FAR_TILE];
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int t0 = blockIdx.x * 256;
    const int sp = blockIdx.y;
    const int n_tiles = (n + FAR_TILE - 1) / FAR_TILE;
    const int s_begin = int((long long)sp * n_tiles / n_split) * FAR_TILE;
    const int s_end = min(n, int((long long)(sp + 1) * n_tiles / n_split) * FAR_TILE);
    const double ox = src[t0].x, oy = src[t0].y, oz = src[t0].z;
    const bool valid = t < n;
    const int i = valid ? order[t] : -1;
    T xi = 0, yi = 0, zi = 0, n0x = 0, n0y = 0, n0z = 0;
    if (valid) {
        const Patch64 p = patches[i];
        xi = T(p.cx - ox); yi = T(p.cy - oy); zi = T(p.cz - oz);
        n0x = T(p.nx); n0y = T(p.ny); n0z = T(p.nz);
    }
    T af = 0, ah = 0;
    const T k = kappa;
    for (int s0 = s_begin; s0 < s_end; s0 += FAR_TILE) {
        __syncthreads();
        const int s = s0 + threadIdx.x;
        if (s < n) {
            const Src64 q = src[s];
            tile[threadIdx.x] = SrcT<T>{T(q.x - ox), T(q.y - oy), T(q.z - oz), T(q.inv_eps),
                                        T(q.dx), T(q.dy), T(q.dz), T(q.sh)};
            tile_idx[threadIdx.x] = order[s];
        }
        __syncthreads();
        const int m = min(FAR_TILE, n - s0);
        if (valid) {
            for (int jj = 0; jj < m; ++jj) {
                const SrcT<T> q = tile[jj];
                if (tile_idx[jj] == i) continue;
                const T dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
                const T r2 = dx * dx + dy * dy + dz * dz;
                const T ir = FarMath<T>::rsqrt_(r2);
                const T r = r2 * ir;
                const T ir2 = ir * ir;
                const T ir3 = ir2 * ir;
                const T kr = k * r;
                const T e = FarMath<T>::exp_(-kr);
                const T u = e * ir;
                const T g1 = e * (T(1) + kr) * ir3;
                const T ir5 = ir3 * ir2;
                const T gL2 = T(3) * ir5;
                const T g2 = e * (kr * kr + T(3) * kr + T(3)) * ir5;
                const T dd = q.dx * dx + q.dy * dy + q.dz * dz;
                const T nd = n0x * dx + n0y * dy + n0z * dz;
                const T ndv = n0x * q.dx + n0y * q.dy + n0z * q.dz;
                af += q.sh * (u - ir) + dd * (g1 - q.ie * ir3);
                ah += -q.sh * nd * ir3 +
                      q.ie * (ndv * (ir3 - g1) - dd * nd * (gL2 - g2) + q.sh * nd * g1);
            }
        }
    }
    if (valid) {
        if (n_split == 1) {
            y[i] += double(af) * FOUR_PI_INV;
            y[n + i] += double(ah) * FOUR_PI_INV;
        } else {
            part[size_t(2 * sp) * n + i] = double(af);
            part[size_t(2 * sp + 1) * n + i] = double(ah);
        }
    }
}

__global__ __launch_bounds__(256) void k_far_direct_sum(int n, int n_split,
                                                        const double* __restrict__ part,
                                                        double* __restrict__ y) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double af = 0, ah = 0;
    for (int sp = 0; sp < n_split; ++sp) {
        af += part[size_t(2 * sp) * n + i];
        ah += part[size_t(2 * sp + 1) * n + i];
    }
    y[i] += af * FOUR_PI_INV;
    y[n + i] += ah * FOUR_PI_INV;
}

int far_direct_splits(int n, int n_cu) {
    if (n <= 0) return 1;
    const int tgt_blocks = (n + 255) / 256;
    const int n_tiles = (n + FAR_TILE - 1) / FAR_TILE;
    const int want = (4 * n_cu + tgt_blocks - 1) / tgt_blocks;  // about 4 blocks per CU
    return std::max(1, std::min(want, n_tiles));
}

void launch_far_direct(hipStream_t s, int fp32, int n, int n_split, const int* order,
                       const Patch64* patches, const Src64* src, double kappa, double* y,
                       double* part) {
    if (n <= 0) return;
    if (n_split > 1 && !part) throw std::runtime_error("direct far field needs its partial sums");
    const dim3 grid(unsigned((n + 255) / 256), unsigned(n_split));
    if (fp32)
        hipLaunchKernelGGL(k_far_direct<float>, grid, dim3(256), 0, s, n, n_split, order, patches,
                           src, float(kappa), y, part);
    else
        hipLaunchKernelGGL(k_far_direct<double>, grid, dim3(256), 0, s, n, n_split, order, patches,
                           src, kappa, y, part);
    hip_check(hipGetLastError(), "k_far_direct launch");
    if (n_split > 1) {
        hipLaunchKernelGGL(k_far_direct_sum, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n,
                           n_split, part, y);
        hip_check(hipGetLastError(), "k_far_direct_sum launch");
    }
}

}  // namespace beepdev
