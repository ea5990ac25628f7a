// The FMM leaf pass folded into the near coefficients (gfx950).
//
// The leaf pass (P2P) adds the 1-point interaction of every pair of patches
// in neighbouring leaf boxes.  Most of those pairs are also explicit near
// pairs, whose coefficients would otherwise carry (full rule - 1 point)
// only for the leaf pass to add the 1 point back on every mat-vec.  With a
// plan, the near coefficients of such pairs keep the full rule
// (fmm_cells_adjacent on the packed leaf cells), and the leaf pass runs over
// the residual: neighbouring-leaf pairs that are not explicit near pairs.
// The operator is the same (every pair: full rule if explicit, else the
// 1-point interaction, beep.cpp:1485-1583).
//
// The residual list is built once per plan and near list, on the device:
//   1. every near-list column as its FMM-order index, sorted per near-list
//      leaf (segmented radix sort);
//   2. per target (FMM order): the sources of its neighbouring leaf boxes
//      that a binary search does not find among its near-list leaf's
//      columns (count, scan on the host, emit).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "device.hpp"
#include "fmm.hpp"

namespace beepdev {

namespace {

__global__ void k_res_keys(long long n, const int* __restrict__ nb_idx,
                           const int* __restrict__ iperm, int* __restrict__ keys) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) keys[e] = iperm[nb_idx[e]];
}

// near-list leaf of every patch
__global__ void k_res_leaf_of(const int* __restrict__ leaf_ptr, const int* __restrict__ leaf_idx,
                              int* __restrict__ leaf_of) {
    const int L = blockIdx.x;
    for (int r = leaf_ptr[L] + int(threadIdx.x); r < leaf_ptr[L + 1]; r += blockDim.x)
        leaf_of[leaf_idx[r]] = L;
}

__device__ __forceinline__ bool sorted_contains(const int* __restrict__ a, int n, int v) {
    int lo = 0, hi = n;  // a[lo..hi) may hold v
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int x = a[mid];
        if (x == v) return true;
        if (x < v) lo = mid + 1; else hi = mid;
    }
    return false;
}

// EMIT = false: count[k] = residual sources of target k; EMIT = true: write
// them at out[res_ptr[k]..] in neighbour-list order, increasing index.  16
// lanes per target walk its neighbour boxes' sources side by side; a ballot
// over the group orders each step's finds.
constexpr int SCAN_LANES = 16;

template <bool EMIT>
__global__ __launch_bounds__(256) void k_res_scan(int n, const FmmDev d,
                                                  const int* __restrict__ tleaf,
                                                  const int* __restrict__ leaf_of,
                                                  const int* __restrict__ nb_ptr,
                                                  const int* __restrict__ keys,
                                                  const int* __restrict__ res_ptr,
                                                  int* __restrict__ out) {
    const int k = blockIdx.x * (256 / SCAN_LANES) + int(threadIdx.x) / SCAN_LANES;
    const int lane = int(threadIdx.x) % SCAN_LANES;
    const int wl = int(threadIdx.x) % 64;  // lane in the wave
    const unsigned long long gmask = 0xFFFFull << (wl & ~(SCAN_LANES - 1));
    if (k >= n) return;  // the whole group: every lane of it has this k
    const int A = tleaf[k];
    const int L = leaf_of[d.perm[k]];
    const int* S = keys + nb_ptr[L];
    const int C = nb_ptr[L + 1] - nb_ptr[L];
    const int o = EMIT ? res_ptr[k] : 0;
    int c = 0;
    for (int q = d.nb_ptr[A]; q < d.nb_ptr[A + 1]; ++q) {
        const int B = d.nb_leaf[q];
        const int s0 = d.leaf_src[B], s1 = d.leaf_src[B + 1];
        for (int sb = s0; sb < s1; sb += SCAN_LANES) {
            const int s = sb + lane;
            const bool keep = s < s1 && s != k && !sorted_contains(S, C, s);
            const unsigned long long m = __ballot(keep) & gmask;
            if (EMIT && keep) out[o + c + __popcll(m & ((1ull << wl) - 1))] = s;
            c += __popcll(m);
        }
    }
    if (!EMIT && lane == 0) out[k] = c;
}

// Residual leaf pass: 16 lanes per target (FMM order), each lane taking
// every 16th source of the target's list, then a fixed-order shuffle
// reduction (lists are short and uneven: a thread per target would leave
// most lanes waiting on the longest list).  Sources come from the packed
// list of the mat-vec (k_far_pack), positions relative to the target formed
// in FP64.
constexpr int RES_LANES = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_p2p_res(int n, const int* __restrict__ perm,
                                                 const int* __restrict__ res_ptr,
                                                 const int* __restrict__ res_src,
                                                 const Src64* __restrict__ src,
                                                 const Patch64* __restrict__ patches, T kappa,
                                                 double* __restrict__ y) {
    const int k = blockIdx.x * (256 / RES_LANES) + threadIdx.x / RES_LANES;
    const int lane = threadIdx.x % RES_LANES;
    const bool valid = k < n;
    const int r0 = valid ? res_ptr[k] : 0, r1 = valid ? res_ptr[k + 1] : 0;
    T af = 0, ah = 0;
    int i = 0;
    if (r0 < r1) {
        i = perm[k];
        const Patch64 p = patches[i];
        const T n0x = T(p.nx), n0y = T(p.ny), n0z = T(p.nz);
        for (int r = r0 + lane; r < r1; r += RES_LANES) {
            const Src64 q = src[res_src[r]];
            p2p_pair<T>(T(q.x - p.cx), T(q.y - p.cy), T(q.z - p.cz), T(q.inv_eps), T(q.dx),
                        T(q.dy), T(q.dz), T(q.sh), n0x, n0y, n0z, kappa, af, ah);
        }
    }
    // every lane of the wave takes part in the shuffles
#pragma unroll
    for (int off = RES_LANES / 2; off > 0; off >>= 1) {
        af += __shfl_down(af, off, RES_LANES);
        ah += __shfl_down(ah, off, RES_LANES);
    }
    if (r0 < r1 && lane == 0) {
        y[i] += double(af) * FOUR_PI_INV;
        y[n + i] += double(ah) * FOUR_PI_INV;
    }
}

}  // namespace

void fmm_build_residual(FmmPlan* pl, int n_leaves, const int* leaf_ptr, const int* leaf_idx,
                        const int* nb_ptr, const int* nb_idx, long long n_nb, hipStream_t s) {
    const int N = pl->N;
    DevArena& ar = *pl->arena;
    // 1. near-list columns in FMM order, sorted per near-list leaf
    int* keys_in = ar.get<int>("res_keys_in", size_t(std::max<long long>(n_nb, 1)));
    int* keys = ar.get<int>("res_keys", size_t(std::max<long long>(n_nb, 1)));
    if (n_nb > 0) {
        hipLaunchKernelGGL(k_res_keys, dim3(unsigned((n_nb + 255) / 256)), dim3(256), 0, s, n_nb,
                           nb_idx, pl->d_iperm, keys_in);
        hip_check(hipGetLastError(), "k_res_keys launch");
        size_t tmp_bytes = 0;
        hip_check(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tmp_bytes, keys_in, keys,
                                                              int(n_nb), n_leaves, nb_ptr,
                                                              nb_ptr + 1, 0, 32, s),
                  "segmented sort size");
        void* tmp = ar.get<unsigned char>("res_sort_tmp", std::max<size_t>(tmp_bytes, 1));
        hip_check(hipcub::DeviceSegmentedRadixSort::SortKeys(tmp, tmp_bytes, keys_in, keys,
                                                              int(n_nb), n_leaves, nb_ptr,
                                                              nb_ptr + 1, 0, 32, s),
                  "segmented sort");
    }
    int* leaf_of = ar.get<int>("res_leaf_of", size_t(N));
    hipLaunchKernelGGL(k_res_leaf_of, dim3(n_leaves), dim3(256), 0, s, leaf_ptr, leaf_idx, leaf_of);
    hip_check(hipGetLastError(), "k_res_leaf_of launch");
    // 2. count, scan, emit
    int* count = ar.get<int>("res_count", size_t(N));
    const unsigned blocks = unsigned((N + 256 / SCAN_LANES - 1) / (256 / SCAN_LANES));
    hipLaunchKernelGGL((k_res_scan<false>), dim3(blocks), dim3(256), 0, s, N, pl->dv,
                       pl->d_tleaf, leaf_of, nb_ptr, keys, nullptr, count);
    hip_check(hipGetLastError(), "k_res_scan launch");
    std::vector<int> h(N);
    hip_check(hipMemcpyAsync(h.data(), count, sizeof(int) * N, hipMemcpyDeviceToHost, s),
              "D2H residual counts");
    hip_check(hipStreamSynchronize(s), "sync residual counts");
    std::vector<int> ptr(size_t(N) + 1, 0);
    long long total = 0;
    for (int k = 0; k < N; ++k) {
        total += h[k];
        if (total > 0x7fffffffLL) throw std::runtime_error("residual leaf-pass list too long");
        ptr[k + 1] = int(total);
    }
    pl->dv.res_ptr = ar.get<int>("res_ptr", size_t(N) + 1);
    pl->dv.res_src = ar.get<int>("res_src", size_t(std::max<long long>(total, 1)));
    hip_check(hipMemcpyAsync(pl->dv.res_ptr, ptr.data(), sizeof(int) * (size_t(N) + 1),
                             hipMemcpyHostToDevice, s), "H2D residual pointers");
    hipLaunchKernelGGL((k_res_scan<true>), dim3(blocks), dim3(256), 0, s, N, pl->dv, pl->d_tleaf,
                       leaf_of, nb_ptr, keys, pl->dv.res_ptr, pl->dv.res_src);
    hip_check(hipGetLastError(), "k_res_scan launch");
    hip_check(hipStreamSynchronize(s), "sync residual list");  // ptr is freed below
    pl->n_res = total;
}

void launch_p2p_residual(FmmPlan* pl, hipStream_t s, int fp32, const Patch64* patches,
                         double kappa, double* y) {
    const int N = pl->N;
    constexpr int per_block = 256 / RES_LANES;
    const unsigned blocks = unsigned((N + per_block - 1) / per_block);
    if (fp32)
        hipLaunchKernelGGL((k_p2p_res<float>), dim3(blocks), dim3(256), 0, s, N, pl->dv.perm,
                           pl->dv.res_ptr, pl->dv.res_src, pl->src, patches, float(kappa), y);
    else
        hipLaunchKernelGGL((k_p2p_res<double>), dim3(blocks), dim3(256), 0, s, N, pl->dv.perm,
                           pl->dv.res_ptr, pl->dv.res_src, pl->src, patches, kappa, y);
    hip_check(hipGetLastError(), "k_p2p_res launch");
}

}  // namespace beepdev
