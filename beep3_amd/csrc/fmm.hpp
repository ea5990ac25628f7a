// Plan and launch interface of the Chebyshev-interpolation FMM (fmm.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device.hpp"

namespace beepdev {

// Device view of the plan (passed by value to kernels).
struct FmmDev {
    int D = 0;             // leaf level
    int leaf_first = 0;    // global index of the first leaf box
    double4* box_c = nullptr;   // centre xyz, edge
    double* box_h = nullptr;    // edge per level
    int* box_parent = nullptr;
    int* box_oct = nullptr;     // octant of a box within its parent
    int* child_ptr = nullptr;   // per box
    int* child_box = nullptr;
    int* child_oct = nullptr;
    int* leaf_src = nullptr;    // per leaf: first source in FMM order (+ end)
    int* perm = nullptr;        // FMM order -> patch index
    int* nb_ptr = nullptr;      // per leaf: neighbour leaf list
    int* nb_leaf = nullptr;
    int* pair_src = nullptr;    // M2L pairs grouped by offset
    int* pair_sym = nullptr;    // symmetry (0..47) taking each pair to its canonical offset
    int* pair_seg = nullptr;    // target segment of each pair (runs of one target in a chunk)
    int* chunk_seg = nullptr;   // per chunk: its first segment (+ end)
    unsigned short* psi = nullptr;  // [48][P^3] canonical node -> actual node
    unsigned short* phi = nullptr;  // [48][P^3] actual node -> canonical node
    int* tgt_ptr = nullptr;     // per box: its M2L segments, increasing
    int* tgt_seg = nullptr;
    double* tn = nullptr;       // T_j(xi_k), P x P
    double* t_oct = nullptr;    // child-node interpolation factors, 2 x P x P
    // leaf-pass pairs that are not explicit near pairs (the others are folded
    // into the near coefficients): per target in FMM order, sources in FMM
    // order
    int* res_ptr = nullptr;
    int* res_src = nullptr;
};

// 1/sqrt(r2), and 0 for r2 == 0 (the self pair: target and source are the
// same centroid, formed from the same bits, so every term it feeds vanishes).
__device__ __forceinline__ float rsq_or_zero(float r2) {
    const float ir = __builtin_amdgcn_rsqf(r2);
    return r2 > 0.f ? ir : 0.f;
}
__device__ __forceinline__ double rsq_or_zero(double r2) {
    const double ir = rsqrt_nr(r2);
    return r2 > 0.0 ? ir : 0.0;
}
__device__ __forceinline__ float exp_fast(float v) { return __expf(v); }
__device__ __forceinline__ double exp_fast(double v) { return exp(v); }

// 1-point interaction of one packed source (relative position d = source -
// target, 1/eps ie, dipole a f n = (sdx, sdy, sdz), monopole a h = sh) with
// a target of normal n0, added to the target's f (af) and h (ah) sums
// without the factor 1/4pi (the three potentials of kernels_far.hip).
template <typename T>
__device__ __forceinline__ void p2p_pair(T dx, T dy, T dz, T ie, T sdx, T sdy, T sdz, T sh,
                                         T n0x, T n0y, T n0z, T kappa, T& af, T& ah) {
    const T r2 = dx * dx + dy * dy + dz * dz;
    const T ir = rsq_or_zero(r2);
    const T r = r2 * ir;
    const T ir2 = ir * ir;
    const T ir3 = ir2 * ir;
    const T kr = kappa * r;
    const T e = exp_fast(-kr);
    const T u = e * ir;
    const T g1 = e * (T(1) + kr) * ir3;
    const T ir5 = ir3 * ir2;
    const T gL2 = T(3) * ir5;
    const T g2 = e * (kr * kr + T(3) * kr + T(3)) * ir5;
    const T dd = sdx * dx + sdy * dy + sdz * dz;
    const T nd = n0x * dx + n0y * dy + n0z * dz;
    const T ndv = n0x * sdx + n0y * sdy + n0z * sdz;
    af += sh * (u - ir) + dd * (g1 - ie * ir3);
    ah += -sh * nd * ir3 + ie * (ndv * (ir3 - g1) - dd * nd * (gL2 - g2) + sh * nd * g1);
}

struct FmmPlanStage;

// Device tables of the plane-wave engine, per level: [0] Laplace, [1] Yukawa
template <typename T>
struct PwDevLevel {
    T* m2mT[2] = {nullptr, nullptr};  // M2M of this level's children ([8 jp][jp])
    T* l2lT[2] = {nullptr, nullptr};  // L2L into this level's children ([8][jp][jp])
    T* m2x[2] = {nullptr, nullptr};   // [jp][nwp]
    T* x2l[2] = {nullptr, nullptr};   // [nwp][jp]
    T* shift[2] = {nullptr, nullptr}; // [98][nw][2]
    int nw[2] = {0, 0};
};
// One GEMM of a batched plane-wave launch: C[row] (= or +=) alpha A[row] B
// over the rows (box, channel) of a set of boxes.  Row r: box = rows ?
// rows[r / nch] : first + r / nch, channel ch0 + r % nch.  A and C hold
// [box - box0][3 channels][ld].  A row by mode: 0 the box's own row; 1 K = 8
// blocks of kblk, block o the row of child o (zero when absent); 2 the row of
// the box's parent.  B: K rows of N used columns, row stride ldb.  tiles =
// row tiles x column tiles (64 x 64).
template <typename T>
struct PwProb {
    int mode, first, nrows, nch, ch0, K, N, kblk, lda, ldb, ldc, accumulate, abox0, cbox0, tiles, ctiles;
    const int* rows;
    const T* A;
    const T* B;
    T* C;
    T alpha;
};
template <typename T>
struct PwDevTables {
    std::vector<PwDevLevel<T>> lev;
    // the problems of every batched launch, in device memory, and the
    // launches: [first problem, problems, blocks]
    PwProb<T>* probs = nullptr;
    std::vector<int3> launch_m2m;  // one per level D-1 .. 2
    int3 launch_m2x{0, 0, 0}, launch_x2l{0, 0, 0};
    std::vector<int3> launch_l2l;  // one per level 3 .. D
    // X2X over every level >= 2 in one launch: per level the Laplace and
    // Yukawa wave counts and shift tables
    int* x2x_nw = nullptr;         // [2 (D + 1)]
    T** x2x_sh = nullptr;          // [2 (D + 1)]
};

struct FmmPlan {
    static constexpr int CHUNK = 64;  // M2L pairs per block
    static constexpr int NCANON = 16; // canonical M2L offsets
    unsigned long long uid = 0;       // unique per plan built in the process
    int N = 0, P = 4, D = 0, nbox = 0, nleaf = 0, n_pairs = 0, n_segs = 0;
    long long n_p2p = 0;              // target-source pairs of the leaf pass (self included)
    unsigned* cellkey = nullptr;      // per patch (original index); null when D > 10
    int* d_iperm = nullptr;           // patch -> FMM order
    int* d_tleaf = nullptr;           // FMM order -> leaf box
    long long n_res = -1;             // residual leaf-pass pairs (-1: not built)
    FmmPlanStage* stage = nullptr;    // host data of a plan begun and not finished
    double edge = 0;
    size_t n_exp = 0;                 // values per expansion set (nbox * 3 * P^3)
    std::vector<int> lvl_first;       // D + 2
    std::vector<int> lvl_chunk;       // D + 2
    std::vector<double> box_h;
    FmmDev dv;
    int4* d_chunks = nullptr;
    int4* d_canon = nullptr;          // canonical offsets (x, y, z, 0)
    double* d_xi = nullptr;
    std::vector<double*> ktab64;      // [l] Yukawa at level l, [D+1] Laplace
    std::vector<float*> ktab32;
    double* exp64 = nullptr;          // M then L
    float* exp32 = nullptr;
    double** ytab64 = nullptr;        // device copies of the table pointers
    float** ytab32 = nullptr;
    double* slots64 = nullptr;        // M2L results per target segment (n_segs x 3 x P^3)
    float* slots32 = nullptr;
    Src64* src = nullptr;
    bool have64 = false, have32 = false;
    double tables_kappa = -1;
    DevArena* arena = nullptr;        // owner of every device buffer above
    // plane-wave engine (engine 1): expansion order, quadrature nodes, padded
    // coefficient and wave widths, the children of every box (8 per box, -1
    // absent), the X2X lists (per box at levels >= 2 and direction: source
    // boxes and local offset codes), the L2L rows of every level grouped by
    // child octant (9 offsets per level into pw_l2l_rows)
    int engine = 0;                   // 0: Chebyshev interpolation, 1: plane waves
    int pw_p = 0, pw_nlam = 0, pw_jp = 0, pw_nwp = 0;
    long long pw_n_x2x = 0;           // X2X translations per mat-vec (all levels)
    int* pw_child8 = nullptr;
    int* pw_xl_ptr = nullptr;
    int* pw_xl_src = nullptr;
    int* pw_xl_code = nullptr;
    int* pw_l2l_rows = nullptr;
    int* pw_l2l_gptr = nullptr;
    std::vector<int> pw_l2l_gptr_h;   // host copy (9 per level)
    int* pw_box_level = nullptr;      // level of every box
    PwDevTables<double> pw64;
    PwDevTables<float> pw32;
    double *pw_M64 = nullptr, *pw_L64 = nullptr, *pw_W64 = nullptr, *pw_Wacc64 = nullptr, *pw_T64 = nullptr;
    float *pw_M32 = nullptr, *pw_L32 = nullptr, *pw_W32 = nullptr, *pw_Wacc32 = nullptr, *pw_T32 = nullptr;
};

// The plan's device buffers come from the arena (owned by the caller, reused
// by the next plan built on it); one live plan per arena.  engine 0:
// Chebyshev interpolation of the given order (3..6); engine 1: plane waves
// of expansion order 9, 12 or 18 (9 quadrature nodes for order 9, else 18).
FmmPlan* fmm_plan_create(int N, const double* centroid, int order, int leaf_target,
                         DevArena& arena, hipStream_t s, int engine = 0);
// The same in two halves: begin chooses the leaf level and uploads the cell
// keys (stream-ordered on s; uid, D, cellkey valid), finish builds the rest
// and synchronises s.
FmmPlan* fmm_plan_begin(int N, const double* centroid, int order, int leaf_target,
                        DevArena& arena, hipStream_t s, int engine = 0);
void fmm_plan_finish(FmmPlan* p, hipStream_t s);
// plane-wave engine (fmm_pw.hip): tables of the plan's levels for kappa and
// the expansion buffers; the far field without the leaf pass into y
bool pw_order_supported(int p);
void pw_prepare(FmmPlan* p, double kappa, bool fp64, bool fp32, hipStream_t s);
void pw_evaluate(FmmPlan* p, hipStream_t s, int fp32, const Patch64* patches, double kappa, double* y,
                 hipEvent_t* m2l_ev);
void fmm_plan_destroy(FmmPlan* p);
// algorithmic HBM bytes of one far-field evaluation: every far-field array
// (sources, expansions, plane waves, the leaf-pass list, the far vector)
// read once and written once
double fmm_algorithmic_bytes(const FmmPlan* p, int fp32);
void fmm_prepare(FmmPlan* p, double kappa, bool fp64, bool fp32, hipStream_t s);
inline bool fmm_tables_ready(const FmmPlan* p, double kappa, bool fp64, bool fp32) {
    return p->tables_kappa == kappa && (!fp64 || p->have64) && (!fp32 || p->have32);
}
// canonical offset of the M2L offset t (index returned, components in
// canon_xyz) and the map canonical node -> actual node (order^3 entries)
int fmm_canonical(int tx, int ty, int tz, int order, int* canon_xyz, int* node_map);
// y += far field (1-point interaction of every pair) of x
// y += the far field of x without its P2P part (pack, upward pass, M2L,
// downward pass, L2P); m2l_ev: optional pair of events recorded around the
// M2L launches; pack_ev: optional event recorded once the sources are packed
void fmm_evaluate(FmmPlan* p, hipStream_t s, int fp32, const Patch64* patches, const double* x,
                  double kappa, double* y, hipEvent_t* m2l_ev = nullptr,
                  hipEvent_t pack_ev = nullptr);
// y += the P2P part (neighbour leaf boxes) of the sources packed by the last
// fmm_evaluate (stream-ordered after its pack): with a residual list, only
// the pairs the near coefficients do not hold in full
void fmm_p2p(FmmPlan* p, hipStream_t s, int fp32, const Patch64* patches, double kappa, double* y);
// Build the residual leaf-pass list against the explicit near pairs (per
// near-list leaf L: rows leaf_idx[leaf_ptr[L]..], columns nb_idx[nb_ptr[L]..],
// device arrays); needs cellkey.  Synchronises the stream.
void fmm_build_residual(FmmPlan* p, int n_leaves, const int* leaf_ptr, const int* leaf_idx,
                        const int* nb_ptr, const int* nb_idx, long long n_nb, hipStream_t s);
// y += the residual leaf pass (fmm_p2p with a residual list)
void launch_p2p_residual(FmmPlan* p, hipStream_t s, int fp32, const Patch64* patches,
                         double kappa, double* y);

}  // namespace beepdev
