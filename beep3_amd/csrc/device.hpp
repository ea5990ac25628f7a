// Internal declarations shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace beepdev {

constexpr double FOUR_PI_INV = 0.07957747154594767;  // 1 / (4 pi)

// One point of a node patch, stored relative to the centroid of the patch it
// belongs to (the "local frame": distances between nearby points are then
// formed from small numbers, which keeps FP32 differences accurate).
template <typename T>
struct alignas(16) Pt {
    T x, y, z, w;     // offset from own patch centroid, weight
    T nx, ny, nz, pad;
};
using Pt32 = Pt<float>;
using Pt64 = Pt<double>;

// Patch record: centroid and area, unit normal and 1/eps.
struct alignas(16) Patch64 {
    double cx, cy, cz, area;
    double nx, ny, nz, inv_eps;
};

// Far-field source record (FP64): position + 1/eps, dipole vector a f n and
// monopole a h.
struct alignas(16) Src64 {
    double x, y, z, inv_eps;
    double dx, dy, dz, sh;
};

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  xcd_remap turns block b of n into a logical index such
// that consecutive logical indices run on the same XCD, so blocks that share
// data (same M2L offset, neighbouring leaves) share one L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    constexpr int NX = 8;
    const int q = n / NX, r = n % NX;
    const int x = b % NX, k = b / NX;
    return x * q + min(x, r) + k;
}

// 1/sqrt(v) in FP64: the hardware estimate refined by two Newton steps
// y <- y + y (1/2 - (v/2) y^2), within an ulp or two of the correctly
// rounded value and several times cheaper than 1.0 / sqrt(v)
__device__ __forceinline__ double rsqrt_nr(double v) {
    double y = __builtin_amdgcn_rsq(v);
    const double h = 0.5 * v;
    y = fma(y, fma(-h * y, y, 0.5), y);
    return fma(y, fma(-h * y, y, 0.5), y);
}

// FMM leaf cell of a patch packed as x | y << 10 | z << 20 (leaf level
// <= 10): two patches meet in the FMM leaf pass iff every field differs by
// at most 1.
constexpr int FMM_CELL_BITS = 10;
__host__ __device__ inline bool fmm_cells_adjacent(unsigned a, unsigned b) {
    constexpr unsigned M = (1u << FMM_CELL_BITS) - 1;
    bool ok = true;
    for (int f = 0; f < 3; ++f) {
        const int u = int((a >> (FMM_CELL_BITS * f)) & M), v = int((b >> (FMM_CELL_BITS * f)) & M);
        ok = ok && u - v <= 1 && v - u <= 1;
    }
    return ok;
}

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Named grow-only device buffers.  A name keeps its allocation until a larger
// size is asked for, so a plan rebuilt for a new geometry of no larger size
// allocates and frees nothing (hipFree waits for the whole device, which
// would stall other contexts running on it).
struct DevArena {
    struct Slot {
        void* p = nullptr;
        size_t bytes = 0;
    };
    std::unordered_map<std::string, Slot> slots;
    size_t held = 0;
    DevArena() = default;
    DevArena(const DevArena&) = delete;
    DevArena& operator=(const DevArena&) = delete;
    ~DevArena() { release(); }
    // grow-only, a regrowth with 25 % headroom (plans of varying size reuse
    // their buffers; a free waits for the whole device)
    void* get(const std::string& name, size_t bytes) {
        Slot& sl = slots[name];
        if (sl.p && bytes <= sl.bytes) return sl.p;
        size_t b = bytes ? bytes : 16;
        if (sl.p) {
            b = std::max(b, sl.bytes + sl.bytes / 4);
            (void)hipFree(sl.p);
            held -= sl.bytes;
            sl = Slot{};
        }
        hip_check(hipMalloc(&sl.p, b), "hipMalloc");
        sl.bytes = b;
        held += b;
        return sl.p;
    }
    template <typename T>
    T* get(const std::string& name, size_t count) {
        return static_cast<T*>(get(name, count * sizeof(T)));
    }
    void release() {
        for (auto& kv : slots)
            if (kv.second.p) (void)hipFree(kv.second.p);
        slots.clear();
        held = 0;
    }
};

// ---- launchers (defined in the .hip files) ----
// singular self terms; a bad one (|v| > maxv or not finite) is zeroed and
// counted in *repaired; guard_only: check the stored terms and recompute only
// the bad ones
void launch_singular(hipStream_t s, int n, const Patch64* patches, const double* gc,
                     const int* gal_ptr, const Pt64* gal, double kappa, int guard_only,
                     double maxv, unsigned* repaired, double* diag4);

// Near-singular entries of an FP32 coefficient set, listed by the FP32
// coefficient kernel and computed in FP64 afterwards: global entry index,
// row patch, 2 * column patch + (1 if the entry stores full - 1 point).
struct NsRec {
    int64_t idx;
    int i, j2;
};
struct NsList {
    NsRec* rec = nullptr;      // null: no FP64 pass (every entry in FP32)
    unsigned* count = nullptr; // records listed (may exceed cap: the rest stay FP32)
    unsigned cap = 0;
};

// near-field coefficients; coef points to float4 (FP32) or double2 pairs (FP64)
void launch_near_coeffs_fp32(hipStream_t s, int n_blocks, int n_leaves, const int* blk_ptr,
                             const int* leaf_ptr, const int* leaf_idx, const int* nb_ptr,
                             const int* nb_idx, const int64_t* coef_off, const Patch64* patches,
                             const int* qual_ptr, const Pt32* qual, const int* quad_ptr,
                             const Pt32* quad, float kappa, int subtract_1pt,
                             const unsigned* cellkey, const NsList& ns, const NsList& guard,
                             double maxv, float4* coef);
// the FP64 pass over the near-singular entries the FP32 kernel listed (a bad
// result is zeroed and counted in *repaired)
void launch_near_fix64(hipStream_t s, const NsList& ns, const Patch64* patches, const int* qual_ptr,
                       const Pt64* qual, const int* quad_ptr, const Pt64* quad, double kappa,
                       double maxv, unsigned* repaired, float4* coef);
// coefficient guard: list the entries of a stored set with a value |v| >
// maxv or not finite (the coefficient kernels list theirs as they go), then
// recompute the listed ones in FP64 (zero if still bad, or a self pair)
void launch_coef_scan(hipStream_t s, int fp32, int n_blocks, int n_leaves, const int* blk_ptr,
                      const int* leaf_ptr, const int* leaf_idx, const int* nb_ptr, const int* nb_idx,
                      const int64_t* coef_off, const unsigned* cellkey, int subtract_1pt,
                      const void* coef, double maxv, const NsList& out);
void launch_coef_repair(hipStream_t s, int fp32, const NsList& list, const Patch64* patches,
                        const int* qual_ptr, const Pt64* qual, const int* quad_ptr, const Pt64* quad,
                        double kappa, double maxv, unsigned* repaired, void* coef);
void launch_near_coeffs_fp64(hipStream_t s, int n_blocks, int n_leaves, const int* blk_ptr,
                             const int* leaf_ptr, const int* leaf_idx, const int* nb_ptr,
                             const int* nb_idx, const int64_t* coef_off, const Patch64* patches,
                             const int* qual_ptr, const Pt64* qual, const int* quad_ptr,
                             const Pt64* quad, double kappa, int subtract_1pt,
                             const unsigned* cellkey, const NsList& guard, double maxv,
                             double2* coef);

// Work decomposition of the near-field SpMV (built on the host from the
// tree): work items (leaf, first row, first column, columns) of <= 256 rows x
// <= SPMV_CW columns, and row groups (leaf, first row, rows, first work item)
// with their work-item counts.
constexpr int SPMV_CW = 512;
// rows per work item of the on-the-fly near field (launch_near_fly)
constexpr int FLY_RB = 64;
struct SpmvPlan {
    const int4* work = nullptr;
    int n_work = 0;
    const int4* groups = nullptr;
    const int* group_nblk = nullptr;
    int n_groups = 0;
};
// y = near(x) + diag(x), overwriting y.  mode: 0 FP64 coef/FP64 sums,
// 1 FP32 coef/FP32 sums, 2 FP32 coef/FP64 sums.  part: scratch of
// n_work * 256 * 2 doubles.  max_blocks > 0 caps the grid (blocks then walk
// the work items with a grid stride); depth (1, 2, 4) scales the coefficient
// loads each thread keeps in flight.
// ncopies (1 or 2) copies of the near field at patch offsets copy_offset
// (null: 0) share one pass over the coefficients; part then holds ncopies x
// n_work * 256 * 2 values.
void launch_near_spmv(hipStream_t s, int mode, int n, const SpmvPlan& p, const int* leaf_ptr,
                      const int* leaf_idx, const int* nb_ptr, const int* nb_idx,
                      const int64_t* coef_off, const float4* coef32, const double2* coef64,
                      const double* diag4, const double* x, void* part, double* y,
                      int max_blocks, int depth, int ncopies = 1, const int* copy_offset = nullptr);
// The same product with every explicit pair integrated on the fly (no
// coefficients stored), over a plan of row groups of <= FLY_RB rows; part
// holds 2 * FLY_RB values per work item.  cellkey: the FMM leaf cells (null
// with the direct far field); ns64: FP32 near-singular pairs in FP64.
void launch_near_fly(hipStream_t s, int mode, int n, const SpmvPlan& p, const int* leaf_ptr,
                     const int* leaf_idx, const int* nb_ptr, const int* nb_idx, const Patch64* patches,
                     const int* qual_ptr, const Pt32* qual32, const Pt64* qual64, const int* quad_ptr,
                     const Pt32* quad32, const Pt64* quad64, double kappa, const unsigned* cellkey,
                     int ns64, double maxv, const double* diag4, const double* x, void* part, double* y);

// explicit pairs listed one by one: coefficients (FP64, rounded to FP32 as
// well; bad values zeroed and counted), and y += their products over the
// rows that have pairs (CSR: rows, ptr, col; mode as launch_near_spmv)
void launch_pair_coeffs(hipStream_t s, long long n_pairs, const int* pi, const int* pj,
                        const Patch64* patches, const int* qual_ptr, const Pt64* qual, const int* quad_ptr,
                        const Pt64* quad, double kappa, double maxv, unsigned* repaired, float4* c32,
                        double2* c64);
void launch_pair_spmv(hipStream_t s, int mode, int n, int n_rows, const int* rows, const int* ptr,
                      const int* col, const float4* c32, const double2* c64, const double* x, double* y);
// pairs between copies (contiguous ranges of ns patches at copy_off) closer
// than rc; sphere: per copy (centre xyz, radius); chunk_sph: per copy, the
// spheres of its chunks of 256 consecutive patches.  ptr null: count[i] =
// pairs of row i; else their columns at cols[ptr[i]..] (index order) and
// rows_of[] = i
void launch_cut_pairs(hipStream_t s, int N, int ncopies, int ns, const int* copy_off, const double4* sphere,
                      const double4* chunk_sph, const Patch64* patches, double rc, const int* ptr, int* count,
                      int* cols, int* rows_of);
void launch_iota(hipStream_t s, int n, int* out);
// exclusive prefix sum of n ints (temporary storage of the size the first
// function gives)
size_t exclusive_scan_temp_bytes(int n);
void launch_exclusive_scan(hipStream_t s, void* tmp, size_t tmp_bytes, const int* in, int* out, int n);

// far field: pack sources from x, then y += full 1-point sum over all pairs
void launch_far_pack(hipStream_t s, int n, const Patch64* patches, const int* order,
                     const double* x, Src64* src);
// source-run splits of the direct far field that fill the chip (n_cu CUs)
int far_direct_splits(int n, int n_cu);
// y += direct far field; n_split > 1 needs part: 2 n n_split doubles
void launch_far_direct(hipStream_t s, int fp32, int n, int n_split, const int* order,
                       const Patch64* patches, const Src64* src, double kappa, double* y,
                       double* part);

// RHS and energy coefficients (FP64 direct sums)
void launch_rhs(hipStream_t s, int n, const int* qual_ptr, const Pt64* qual,
                const Patch64* patches, int n_charges, const double* xyzq, double scale,
                double* b);
void launch_energy_coeffs(hipStream_t s, int n, const int* gal_ptr, const Pt64* gal,
                          const Patch64* patches, const int* patch_inst,
                          const int* inst_charge_ptr, const double* xyzq, double* ec_f,
                          double* ec_h);

// vector kernels for the GMRES driver
void launch_multi_dot(hipStream_t s, int n, const double* V, int ncols, const double* w,
                      double* partial, int n_part_blocks);
void launch_reduce_partials(hipStream_t s, int ncols, const double* partial, int n_part_blocks,
                            double* out);
void launch_gemv_sub(hipStream_t s, int n, const double* V, int ncols, const double* h,
                     double* w);
void launch_axpby(hipStream_t s, int n, double a, const double* x, double b, double* y);
void launch_fill(hipStream_t s, int n, double v, double* y);
// Arnoldi step (classical Gram-Schmidt against V[0..j], norm, V[j+1]) on
// the mat-vec output w, with j read from *jp; h receives the j + 2 values.
// h2 non-null: a second Gram-Schmidt pass (its coefficients in h2, added to h)
void launch_arnoldi_step(hipStream_t s, int n, int m, int nb, double* V, const int* jp,
                         double* partial, double* h, double* w, double* h2);
void launch_copy_col(hipStream_t s, int n, const double* V, const int* jp, double* out);
// dst row r = src row idx[r] over 64-byte rows (a scattered read of known size)
void launch_gather_rows64(hipStream_t s, const void* src, const int* idx, size_t rows, void* dst);

// octree build on the device (octree_dev.hip): descent keys of the patch
// centroids in the cube (centre, edge) when full is null, else the prefixes
// at `level` of the full keys; items[i] = i
void launch_tree_keys(hipStream_t s, int n, const Patch64* patches, const double* centre,
                      double edge, int level, const uint64_t* full, uint64_t* keys, int* items);
// stable radix sort of (key, item) pairs over key bits [0, end_bit)
size_t tree_sort_bytes(int n, int end_bit);
void tree_sort_pairs(hipStream_t s, void* tmp, size_t tmp_bytes, const uint64_t* kin,
                     uint64_t* kout, const int* vin, int* vout, int n, int end_bit);
// dst[seg.z + k] = src[seg.x + k] for k < seg.y, per segment
void launch_gather_segments(hipStream_t s, int nseg, const int4* seg, const int* src, int* dst);
void launch_set_int(hipStream_t s, int* p, int v);
void launch_inc_int(hipStream_t s, int* p);
// A triangle rule for point generation: k points per sub-triangle at
// barycentric (u, v) with weights w (k = 0: one centroid point per patch).
struct RuleDev {
    int k;
    double u[16], v[16], w[16];
};
// the point set of every patch from the surface, in the local frame
void launch_gen_points(hipStream_t s, int n, const double* verts, const int* tris,
                       const int* vt_ptr, const int* vt_tri, const int* vt_corner,
                       const int* ctl_idx, const double* ctl, const Patch64* patches,
                       const int* ptr, const RuleDev& rule, int per_area, Pt64* out64,
                       Pt32* out32);
// raw CSR points k0 .. k0+n-1 (7 doubles each) -> local-frame records
void launch_to_local(hipStream_t s, int n, int k0, const double* raw, const int* ptr, int N,
                     const Patch64* patches, Pt64* out64, Pt32* out32);
void launch_mfma_selftest(hipStream_t s, const double* a, const double* b, int fp32, double* out);
void launch_scale_copy(hipStream_t s, int n, const double* x, const double* scale_ptr,
                       double scale, double* y);

}  // namespace beepdev
