// Device side of the octree build (beep_hip_build_tree): the descent keys of
// the patch centroids (tree_key.hpp, the same arithmetic as the host build),
// their stable sort, and the gather of the leaves' and neighbourhoods' items
// from per-level sorted item lists.  The tree's shape is decided on the host
// from the sorted keys (host_octree.cpp, Octree::build_sorted); this file does
// the per-item work: one key per patch, a radix sort, and a copy of every
// listed item.
#include <hipcub/hipcub.hpp>

#include "device.hpp"
#include "tree_key.hpp"

namespace beepdev {

namespace {

// full keys (full == nullptr: from the patch centroids) or, from full keys,
// their prefixes at `level`; items[i] = i
__global__ __launch_bounds__(256) void k_tree_keys(int n, const Patch64* __restrict__ patches,
                                                   double cx, double cy, double cz, double edge,
                                                   int level, const uint64_t* __restrict__ full,
                                                   uint64_t* __restrict__ keys,
                                                   int* __restrict__ items) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (full) {
        keys[i] = beeptree::key_prefix(full[i], level);
    } else {
        const Patch64 p = patches[i];
        keys[i] = beeptree::descent_key(p.cx, p.cy, p.cz, cx, cy, cz, edge);
    }
    items[i] = i;
}

// one wave per segment (src offset, length, dst offset): a coalesced copy
__global__ __launch_bounds__(256) void k_gather_segments(int nseg, const int4* __restrict__ seg,
                                                         const int* __restrict__ src,
                                                         int* __restrict__ dst) {
    const int g = blockIdx.x * 4 + int(threadIdx.x >> 6);
    if (g >= nseg) return;
    const int4 q = seg[g];
    for (int k = int(threadIdx.x & 63); k < q.y; k += 64) dst[q.z + k] = src[q.x + k];
}

}  // namespace

void launch_tree_keys(hipStream_t s, int n, const Patch64* patches, const double* centre,
                      double edge, int level, const uint64_t* full, uint64_t* keys, int* items) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_tree_keys, dim3((n + 255) / 256), dim3(256), 0, s, n, patches,
                       centre ? centre[0] : 0.0, centre ? centre[1] : 0.0,
                       centre ? centre[2] : 0.0, edge, level, full, keys, items);
    hip_check(hipGetLastError(), "k_tree_keys launch");
}

size_t tree_sort_bytes(int n, int end_bit) {
    size_t b = 0;
    hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, b, static_cast<const uint64_t*>(nullptr),
                                                  static_cast<uint64_t*>(nullptr),
                                                  static_cast<const int*>(nullptr),
                                                  static_cast<int*>(nullptr), n, 0, end_bit),
              "tree sort size");
    return b;
}

void tree_sort_pairs(hipStream_t s, void* tmp, size_t tmp_bytes, const uint64_t* kin,
                     uint64_t* kout, const int* vin, int* vout, int n, int end_bit) {
    if (n <= 0) return;
    hip_check(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, vout, n, 0,
                                                  end_bit, s),
              "tree sort");
}

void launch_gather_segments(hipStream_t s, int nseg, const int4* seg, const int* src, int* dst) {
    if (nseg <= 0) return;
    hipLaunchKernelGGL(k_gather_segments, dim3((nseg + 3) / 4), dim3(256), 0, s, nseg, seg, src,
                       dst);
    hip_check(hipGetLastError(), "k_gather_segments launch");
}

}  // namespace beepdev
