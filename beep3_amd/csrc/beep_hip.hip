// Device context and the C ABI of include/beep_hip.h: geometry and tree
// upload, precompute, mat-vec (do_bem_fmm_iteration), RHS, energy
// coefficients and the device GMRES / iterative-refinement driver.
#include <algorithm>
#include <mutex>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/beep_hip.h"
#include "device.hpp"
#include "fmm.hpp"
#include "pw.hpp"
#include "host.hpp"
#include "tree_key.hpp"

using namespace beepdev;

namespace {

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // grow-only: keeps the buffer when it is already large enough, and a
    // regrowth takes 25 % headroom, so that a context serving problems of
    // varying size rarely frees (hipFree waits for the whole device and
    // blocks every other thread's launches meanwhile)
    void alloc(size_t count) {
        if (p && count <= n) return;
        const size_t want = n ? std::max(count, n + n / 4) : count;
        reset();
        if (want == 0) return;
        hip_check(hipMalloc(&p, want * sizeof(T)), "hipMalloc");
        n = want;
    }
    void upload(const T* src, size_t count, hipStream_t s) {
        alloc(count);
        if (count) hip_check(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s),
                             "hipMemcpyAsync H2D");
    }
};

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Scratch of the allocation-free mat-vec (beep_hip_matvec_device_noalloc):
// one array per device, held from the library's load to the end of the
// process.  Host code never passes `scratch` itself to a kernel (in host code
// the name is a stand-in, not a device address): it uses the address
// hipGetSymbolAddress returns for the current device, looked up once per
// device when the first context is created on it.
constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// the device address of `scratch` on `device` (current), looked up once
void* scratch_address(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

}  // namespace

struct beep_hip_ctx {
    int device = 0;
    void* scratch = nullptr;  // device address of this device's scratch array
    hipStream_t stream = nullptr;
    // far field runs on its own stream, concurrently with the near SpMV, into
    // yfar; the context stream adds it to y (fork / join through events)
    hipStream_t far_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_pack = nullptr;
    bool overlap = true;
    int n_cu = 256;               // compute units of the device
    // near-SpMV settings per coefficient mode (0: FP64, 1: FP32, 2: FP32
    // coefficients with FP64 sums): residency cap in blocks per CU (0: one
    // block per work item) and loads in flight per thread (x 4 FP64, x 8 FP32)
    int spmv_cap[3] = {1, 1, 1};
    int spmv_depth[3] = {1, 1, 1};
    std::string err;

    // geometry
    int N = 0;
    DevBuf<Patch64> patches;
    DevBuf<double> gc;
    DevBuf<int> qual_ptr, quad_ptr, gal_ptr;
    DevBuf<Pt64> qual64, quad64, gal64;
    DevBuf<Pt32> qual32, quad32;
    DevBuf<double> stage;  // raw point upload staging (7 doubles per point)
    DevBuf<double> sverts; // surface for point generation
    DevBuf<int> stris, svt_ptr, svt_tri, svt_corner;
    DevBuf<int> sctl_idx;  // curved triangles: index and control points, while points are made
    DevBuf<double> sctl;
    // octree build on the device: keys, sort buffers, per-level sorted
    // patches and gather segments (grow-only, kept across builds)
    DevBuf<uint64_t> tk_full, tk_a, tk_b;
    DevBuf<int> ti_a, ti_b, t_sorted;
    DevBuf<unsigned char> t_tmp;
    DevBuf<int4> t_segl, t_segn;
    bool have_geom = false;

    // tree
    int L = 0;
    std::vector<int> h_leaf_ptr, h_nb_ptr;
    DevBuf<int> leaf_ptr, leaf_idx, nb_ptr, nb_idx;
    DevBuf<int64_t> coef_off;
    DevBuf<int> coef_blk;
    int n_coef_blocks = 0;
    DevBuf<int4> spmv_work, spmv_groups;
    DevBuf<int> spmv_group_nblk;
    DevBuf<double> spmv_part;
    SpmvPlan spmv;
    // the on-the-fly near field: its plan (row groups of <= FLY_RB rows),
    // the mode, and whether its FP32 pass integrates near-singular pairs in
    // FP64 (as of the last FP32 precompute)
    DevBuf<int4> fly_work, fly_groups;
    DevBuf<int> fly_group_nblk;
    SpmvPlan fly;
    int near_mode = BEEP_NEAR_PRECOMPUTE;
    bool fly_ns64 = true;
    int64_t nnz = 0;
    long long n_nb = 0;  // neighbourhood entries (nb_ptr[L])
    bool have_tree = false;

    // precomputed operator
    DevBuf<double> diag;
    DevBuf<float4> coef32;
    DevBuf<double2> coef64;
    bool have32 = false, have64 = false, have_diag = false;
    // near-singular entries of the FP32 set, computed in FP64 (on by
    // default): the list the FP32 kernel fills and its device count
    bool ns_fp64 = true;
    DevBuf<NsRec> ns_rec;
    DevBuf<unsigned> ns_count;
    bool ns_listed = false;  // the current FP32 set was built with the FP64 pass
    unsigned ns_cap = 0;     // the list capacity it was built with
    // near field shared with another context (the instances of a pose): its
    // coefficient sets, diagonal and tree applied to copies of its patches
    // at patch offsets share_off, plus this context's own explicit pairs
    // (CSR over the rows that have any, columns sorted; pair_i / pair_j per
    // pair for the coefficient kernel)
    beep_hip_ctx* share = nullptr;
    std::vector<int> share_off;
    long long n_pairs = 0;
    int n_pair_rows = 0;
    DevBuf<int> pair_rows, pair_ptr, pair_col, pair_i;
    DevBuf<int> cut_off_d, cut_count;  // device pair search: copy offsets, per-row counts
    DevBuf<double4> cut_sph_d;         // and the copies' bounding spheres
    DevBuf<double4> cut_chunk_d;       // and those of each copy's chunks of 256 patches
    DevBuf<char> cut_tmp;              // scan storage
    DevBuf<float4> pcoef32;
    DevBuf<double2> pcoef64;
    // coefficient guard (|v| > guard_max or not finite): the list of bad
    // entries and counters [0] list length, entries repaired by the last
    // precompute of [1] the FP64 set, [2] the FP32 set, [3] the diagonal,
    // [4] by the last sanity check
    static constexpr unsigned GUARD_CAP = 1u << 16;
    double guard_max = 1e6;
    DevBuf<NsRec> guard_rec;
    DevBuf<unsigned> guard_cnt;
    // far-field plan each coefficient set was built for (0: none, the direct
    // far field): pairs of the FMM leaf pass keep the full rule
    unsigned long long uid32 = 0, uid64 = 0;
    double kappa = 0;

    // far field
    int far_mode = BEEP_FAR_AUTO, fmm_order = 4, fmm_leaf = 48;
    double plan_ms = 0;  // host time of the current plan and its residual list
    std::vector<double> h_centroid;
    DevArena fmm_arena;  // device buffers of the FMM plan, reused across plans
    FmmPlan* fmm = nullptr;
    bool use_fmm = false;

    // per mat-vec scratch
    DevBuf<Src64> src;
    DevBuf<double> xbuf, ybuf, yfar;
    DevBuf<double> far_part;  // partial sums of the split direct far field
    // charges and outputs of the RHS / energy-coefficient calls (kept, so a
    // solve frees nothing: hipFree waits for the whole device)
    DevBuf<double> qbuf, rhsbuf, ecf, ech;
    DevBuf<int> pinst_d, icp_d;

    // GMRES workspace
    DevBuf<double> V, hvec, partial, wtmp, bvec, rvec, dvec;
    DevBuf<double> xin, wvec;   // fixed mat-vec input / output of an Arnoldi step
    DevBuf<int> jstep;          // Arnoldi step index on the device
    int gs_passes = 1;          // Gram-Schmidt passes per Arnoldi step (1: CGS, 2: CGS2)
    DevBuf<double> hvec2;       // the second pass's coefficients
    double* hh_pinned = nullptr;  // step results read by the host (pinned)
    size_t hh_cap = 0;
    int part_blocks = 0;
    // one Arnoldi step captured as a HIP graph, replayed for every step
    bool use_graphs = true;
    bool capturing = false;
    // captured Arnoldi steps, one per distinct key (precision, buffers, ...),
    // most recently used first; mixed solves alternate two precisions
    struct StepGraph {
        std::vector<const void*> key;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<StepGraph> step_graphs;
    unsigned long long gen = 0;  // bumped whenever buffers or plans may change

    // ring of event sets (start, near done, end, far done, M2L start, M2L
    // done), one per mat-vec
    static constexpr int RING = 64;
    hipEvent_t ev[RING][6] = {};
    int ring_head = 0, ring_pending = 0;  // next slot, slots not yet read
    bool slot_fmm[RING] = {};             // the slot's mat-vec recorded M2L events
    beep_matvec_timing timing{};
};

namespace {

template <typename F>
int guarded(beep_hip_ctx* ctx, F&& f) {
    if (!ctx) return -2;
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return -1;
    } catch (...) {
        ctx->err = "unknown error";
        return -1;
    }
}

void require(bool cond, const char* msg) {
    if (!cond) throw std::runtime_error(msg);
}

void set_device(beep_hip_ctx* c) { hip_check(hipSetDevice(c->device), "hipSetDevice"); }

// wait for both streams (before buffers they may use are freed or replaced)
void sync_streams(beep_hip_ctx* c) {
    if (c->far_stream) hip_check(hipStreamSynchronize(c->far_stream), "hipStreamSynchronize");
    if (c->stream) hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

// (re)build the far-field plan for the current geometry and options
void far_setup(beep_hip_ctx* c) {
    sync_streams(c);
    c->gen++;
    if (c->fmm) {
        fmm_plan_destroy(c->fmm);
        c->fmm = nullptr;
    }
    c->use_fmm = c->far_mode == BEEP_FAR_FMM || c->far_mode == BEEP_FAR_PW ||
                 (c->far_mode == BEEP_FAR_AUTO && c->N > 20000);
    // the plan itself is built on first use (ensure_far), so its host work
    // can overlap the coefficient precompute already running on the device
}

// Build the far-field plan if it is missing, on the far stream (which holds
// no other work then), and with the near list set, its residual leaf-pass
// list.  Host-synchronous.
void begin_plan(beep_hip_ctx* c) {
    if (!c->use_fmm || c->fmm) return;
    hip_check(hipStreamSynchronize(c->far_stream), "hipStreamSynchronize");
    const auto t0 = std::chrono::steady_clock::now();
    c->fmm = fmm_plan_begin(c->N, c->h_centroid.data(), c->fmm_order, c->fmm_leaf, c->fmm_arena,
                            c->far_stream, c->far_mode == BEEP_FAR_PW ? 1 : 0);
    c->plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c->gen++;
}

void ensure_plan(beep_hip_ctx* c) {
    if (!c->use_fmm) return;
    begin_plan(c);
    if (c->fmm->stage) {
        const auto t0 = std::chrono::steady_clock::now();
        fmm_plan_finish(c->fmm, c->far_stream);
        c->plan_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        c->gen++;
    }
    if (c->have_tree && c->fmm->cellkey && c->fmm->n_res < 0) {
        const auto t0 = std::chrono::steady_clock::now();
        fmm_build_residual(c->fmm, c->L, c->leaf_ptr.p, c->leaf_idx.p, c->nb_ptr.p, c->nb_idx.p,
                           c->n_nb, c->far_stream);
        c->plan_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        c->gen++;
    }
}

// The plan (ensure_plan) and the kernel tables of the given precision for
// the current kappa.  Tables change only through a new kappa (precompute,
// which drains both streams first) or a precision not built yet (buffers of
// their own), so no pending far-field work reads what is written here.
void ensure_far(beep_hip_ctx* c, int precision = -1) {
    if (!c->use_fmm) return;
    ensure_plan(c);
    if (precision < 0 || !c->have_diag) return;
    const bool f64 = precision != BEEP_FP32, f32 = precision != BEEP_FP64;
    if (!fmm_tables_ready(c->fmm, c->kappa, f64, f32)) {
        fmm_prepare(c->fmm, c->kappa, f64, f32, c->far_stream);
        c->gen++;
    }
}

// state common to both ways of giving the geometry
void geometry_done(beep_hip_ctx* c, int N, const double* centroid) {
    c->N = N;
    c->h_centroid.assign(centroid, centroid + 3 * size_t(N));
    far_setup(c);
    c->part_blocks = std::max(1, std::min(256, (2 * N + 1023) / 1024));
    c->have_geom = true;
    c->have_tree = false;
    c->have32 = c->have64 = c->have_diag = false;
    c->share = nullptr;
    c->share_off.clear();
    c->n_pairs = 0;
    c->n_pair_rows = 0;
}

void check_ptr(int N, const int* ptr, const char* what) {
    require(ptr != nullptr, what);
    require(ptr[0] == 0, what);
    for (int i = 0; i < N; ++i) require(ptr[i + 1] >= ptr[i], what);
}

// read the oldest pending event triple into the timing sums
void timing_pop(beep_hip_ctx* c) {
    const int slot = (c->ring_head - c->ring_pending + beep_hip_ctx::RING) % beep_hip_ctx::RING;
    hipEvent_t* e = c->ev[slot];
    float a = 0, b = 0, t = 0;
    hip_check(hipEventSynchronize(e[2]), "hipEventSynchronize");
    // e0 start, e1 near done, e3 far done (far stream), e2 end; serially
    // (one stream) the near part runs first and the far part after e1
    hip_check(hipEventElapsedTime(&a, e[0], e[1]), "hipEventElapsedTime");
    hip_check(hipEventElapsedTime(&b, c->overlap ? e[0] : e[1], e[3]), "hipEventElapsedTime");
    hip_check(hipEventElapsedTime(&t, e[0], e[2]), "hipEventElapsedTime");
    float m2l = 0;
    if (c->slot_fmm[slot]) hip_check(hipEventElapsedTime(&m2l, e[4], e[5]), "hipEventElapsedTime");
    c->timing.m2l_ms = m2l;
    c->timing.m2l_ms_sum += m2l;
    c->timing.near_ms = a;
    c->timing.far_ms = b;
    c->timing.total_ms = t;
    c->timing.count += 1;
    c->timing.near_ms_sum += a;
    c->timing.far_ms_sum += b;
    c->timing.total_ms_sum += t;
    c->ring_pending -= 1;
}

void timing_collect(beep_hip_ctx* c) {
    while (c->ring_pending > 0) timing_pop(c);
}

// near-SpMV coefficient mode of a mat-vec precision
int spmv_mode(int precision) {
    return precision == BEEP_FP64 ? 0 : (precision == BEEP_FP32 ? 1 : 2);
}

void matvec_dev(beep_hip_ctx* c, const double* dx, double* dy, int precision) {
    require(c->have_geom && (c->have_tree || c->share), "geometry and tree must be set before a mat-vec");
    require(c->have_diag, "precompute before the mat-vec");
    require(precision == BEEP_FP64 || precision == BEEP_FP32 || precision == BEEP_FP32_RESID,
            "unknown precision");
    if (precision == BEEP_FP64) require(c->have64, "FP64 coefficients not precomputed");
    else require(c->have32, "FP32 coefficients not precomputed");
    if (c->share) {
        require(precision == BEEP_FP64 ? c->share->have64 : c->share->have32,
                "the shared near field lacks this precision's coefficients");
    } else {
        const unsigned long long built = precision == BEEP_FP64 ? c->uid64 : c->uid32;
        require(c->use_fmm ? (c->fmm && c->fmm->uid == built) : built == 0,
                "near coefficients were precomputed for another far field: call "
                "beep_hip_precompute again");
    }
    hipStream_t s = c->stream;
    // plan and kernel tables are built on first use for the current kappa
    // (before any capture: a captured step finds them ready)
    if (!c->capturing) ensure_far(c, precision == BEEP_FP32 ? BEEP_FP32 : BEEP_FP64);
    const bool timed = !c->capturing;
    if (timed && c->ring_pending == beep_hip_ctx::RING) timing_pop(c);
    hipEvent_t* ev = c->ev[c->ring_head];
    const int n2 = 2 * c->N;
    c->yfar.alloc(size_t(n2));
    if (!c->use_fmm) c->src.alloc(size_t(c->N));
    if (timed) hip_check(hipEventRecord(ev[0], s), "hipEventRecord");
    hipStream_t sf = s;
    if (c->overlap) {
        sf = c->far_stream;
        hip_check(hipEventRecord(c->ev_fork, s), "hipEventRecord");
        hip_check(hipStreamWaitEvent(sf, c->ev_fork, 0), "hipStreamWaitEvent");
    }
    // near field (explicit pairs) + diagonal into y, launched first: the
    // SpMV is the longer of the two, and the far field's dozen short kernels
    // take the host a while to launch
    const int mode = spmv_mode(precision);
    if (c->share) {
        // every copy of the shared patches (rows and columns at its offset,
        // f and h halves N apart), then this context's own pairs
        const beep_hip_ctx* h = c->share;
        const int nco = int(c->share_off.size());
        for (int k = 0; k < nco; k += 2)
            launch_near_spmv(s, mode, c->N, h->spmv, h->leaf_ptr.p, h->leaf_idx.p, h->nb_ptr.p,
                             h->nb_idx.p, h->coef_off.p, h->coef32.p, h->coef64.p, h->diag.p, dx,
                             c->spmv_part.p, dy, c->spmv_cap[mode] * c->n_cu, c->spmv_depth[mode],
                             std::min(2, nco - k), c->share_off.data() + k);
        launch_pair_spmv(s, mode, c->N, c->n_pair_rows, c->pair_rows.p, c->pair_ptr.p, c->pair_col.p,
                         c->pcoef32.p, c->pcoef64.p, dx, dy);
    } else if (c->near_mode == BEEP_NEAR_ON_THE_FLY) {
        launch_near_fly(s, mode, c->N, c->fly, c->leaf_ptr.p, c->leaf_idx.p, c->nb_ptr.p, c->nb_idx.p,
                        c->patches.p, c->qual_ptr.p, c->qual32.p, c->qual64.p, c->quad_ptr.p, c->quad32.p,
                        c->quad64.p, c->kappa, c->use_fmm ? c->fmm->cellkey : nullptr, c->fly_ns64 ? 1 : 0,
                        c->guard_max, c->diag.p, dx, c->spmv_part.p, dy);
    } else {
        launch_near_spmv(s, mode, c->N, c->spmv, c->leaf_ptr.p, c->leaf_idx.p, c->nb_ptr.p,
                         c->nb_idx.p, c->coef_off.p, c->coef32.p, c->coef64.p, c->diag.p, dx,
                         c->spmv_part.p, dy, c->spmv_cap[mode] * c->n_cu, c->spmv_depth[mode]);
    }
    if (timed) hip_check(hipEventRecord(ev[1], s), "hipEventRecord");
    // far field (1-point interaction of all pairs) into yfar
    launch_fill(sf, n2, 0.0, c->yfar.p);
    if (c->use_fmm) {
        fmm_evaluate(c->fmm, sf, precision == BEEP_FP32 ? 1 : 0, c->patches.p, dx, c->kappa,
                     c->yfar.p, timed ? &ev[4] : nullptr, c->ev_pack);
    } else {
        launch_far_pack(sf, c->N, c->patches.p, c->leaf_idx.p, dx, c->src.p);
        const int ns = far_direct_splits(c->N, c->n_cu);
        if (ns > 1 && !c->capturing) c->far_part.alloc(size_t(2) * c->N * ns);
        launch_far_direct(sf, precision == BEEP_FP32 ? 1 : 0, c->N, ns, c->leaf_idx.p,
                          c->patches.p, c->src.p, c->kappa, c->yfar.p, c->far_part.p);
    }
    if (timed) hip_check(hipEventRecord(ev[3], sf), "hipEventRecord");
    if (c->overlap) hip_check(hipEventRecord(c->ev_join, sf), "hipEventRecord");
    if (c->use_fmm) {
        // the P2P part (VALU) follows the SpMV on this stream, beside the
        // MFMA-bound M2L of the far stream; it needs the packed sources
        if (c->overlap) hip_check(hipStreamWaitEvent(s, c->ev_pack, 0), "hipStreamWaitEvent");
        fmm_p2p(c->fmm, s, precision == BEEP_FP32 ? 1 : 0, c->patches.p, c->kappa, dy);
    }
    if (c->overlap) hip_check(hipStreamWaitEvent(s, c->ev_join, 0), "hipStreamWaitEvent");
    launch_axpby(s, n2, 1.0, c->yfar.p, 1.0, dy);
    if (!timed) return;
    hip_check(hipEventRecord(ev[2], s), "hipEventRecord");
    c->slot_fmm[c->ring_head] = c->use_fmm;
    c->ring_head = (c->ring_head + 1) % beep_hip_ctx::RING;
    c->ring_pending += 1;
}

double dev_dot(beep_hip_ctx* c, const double* a, const double* b, int n) {
    c->partial.alloc(size_t(c->part_blocks));
    c->hvec.alloc(std::max<size_t>(c->hvec.n, 1));
    launch_multi_dot(c->stream, n, a, 1, b, c->partial.p, c->part_blocks);
    launch_reduce_partials(c->stream, 1, c->partial.p, c->part_blocks, c->hvec.p);
    double r = 0;
    hip_check(hipMemcpyAsync(&r, c->hvec.p, sizeof(double), hipMemcpyDeviceToHost, c->stream),
              "D2H dot");
    hip_check(hipStreamSynchronize(c->stream), "sync dot");
    return r;
}

// The launches of one Arnoldi step: w = K V[j] through the fixed buffers xin
// and wvec, Gram-Schmidt against V[0..j], V[j+1] = w / |w|, the j + 2
// coefficients to pinned host memory, j += 1 (all indices read from the
// device, so the same launches serve every step).
void arnoldi_launches(beep_hip_ctx* c, int precision, int m) {
    const int n = 2 * c->N;
    hipStream_t s = c->stream;
    launch_copy_col(s, n, c->V.p, c->jstep.p, c->xin.p);
    matvec_dev(c, c->xin.p, c->wvec.p, precision);
    launch_arnoldi_step(s, n, m, c->part_blocks, c->V.p, c->jstep.p, c->partial.p, c->hvec.p,
                        c->wvec.p, c->gs_passes == 2 ? c->hvec2.p : nullptr);
    hip_check(hipMemcpyAsync(c->hh_pinned, c->hvec.p, sizeof(double) * (m + 2),
                             hipMemcpyDeviceToHost, s), "D2H arnoldi");
    launch_inc_int(s, c->jstep.p);
}

constexpr size_t STEP_GRAPHS = 4;  // captured steps kept per context

void drop_step_graph(beep_hip_ctx* c) {
    for (auto& g : c->step_graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    c->step_graphs.clear();
}

void arnoldi_step(beep_hip_ctx* c, int precision, int m) {
    if (!c->use_graphs) {
        arnoldi_launches(c, precision, m);
        return;
    }
    // everything the captured launches bake in: buffers, plan, sizes
    ensure_far(c, precision == BEEP_FP32 ? BEEP_FP32 : BEEP_FP64);
    c->yfar.alloc(size_t(2) * c->N);
    if (!c->use_fmm) {
        c->src.alloc(size_t(c->N));
        const int ns = far_direct_splits(c->N, c->n_cu);
        if (ns > 1) c->far_part.alloc(size_t(2) * c->N * ns);
    }
    const std::vector<const void*> key = {
        reinterpret_cast<const void*>(uintptr_t(c->gen)), reinterpret_cast<const void*>(intptr_t(precision)),
        reinterpret_cast<const void*>(intptr_t(m)), reinterpret_cast<const void*>(intptr_t(c->N)),
        c->V.p, c->xin.p, c->wvec.p, c->hvec.p, c->partial.p, c->jstep.p, c->hh_pinned,
        c->yfar.p, c->src.p, c->far_part.p, c->coef32.p, c->coef64.p, c->diag.p,
        c->spmv_part.p, c->fmm, c->hvec2.p, c->share, c->pair_ptr.p, c->pcoef32.p, c->pcoef64.p,
        reinterpret_cast<const void*>(intptr_t(c->gs_passes)),
        reinterpret_cast<const void*>(intptr_t(c->overlap)),
        reinterpret_cast<const void*>(intptr_t(c->near_mode * 2 + (c->fly_ns64 ? 1 : 0))),
        reinterpret_cast<const void*>(intptr_t(c->spmv_cap[spmv_mode(precision)])),
        reinterpret_cast<const void*>(intptr_t(c->spmv_depth[spmv_mode(precision)]))};
    // graphs of an older generation bake in buffers or plans that changed
    for (size_t k = c->step_graphs.size(); k-- > 0;)
        if (c->step_graphs[k].key[0] != key[0]) {
            (void)hipGraphExecDestroy(c->step_graphs[k].exec);
            c->step_graphs.erase(c->step_graphs.begin() + k);
        }
    size_t hit = 0;
    while (hit < c->step_graphs.size() && c->step_graphs[hit].key != key) ++hit;
    if (hit < c->step_graphs.size()) {
        // most recently used first
        std::rotate(c->step_graphs.begin(), c->step_graphs.begin() + hit,
                    c->step_graphs.begin() + hit + 1);
    } else {
        if (c->step_graphs.size() >= STEP_GRAPHS) {
            (void)hipGraphExecDestroy(c->step_graphs.back().exec);
            c->step_graphs.pop_back();
        }
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal),
                  "hipStreamBeginCapture");
        c->capturing = true;
        try {
            arnoldi_launches(c, precision, m);
        } catch (...) {
            c->capturing = false;
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        c->capturing = false;
        hip_check(hipStreamEndCapture(c->stream, &g), "hipStreamEndCapture");
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "hipGraphInstantiate");
        c->step_graphs.insert(c->step_graphs.begin(), beep_hip_ctx::StepGraph{key, ex});
    }
    hip_check(hipGraphLaunch(c->step_graphs.front().exec, c->stream), "hipGraphLaunch");
}

// GMRES without restart (BEEP::gmres, beep.cpp:1343-1432) on device vectors:
// r = K x - b, classical Gram-Schmidt, Givens on the host, stop when
// |y_j| <= eps |b|, x -= V y.  Returns inner iterations; adds mat-vecs.
// bnorm: |b| when the caller has it (the same dot product), else < 0.
int gmres(beep_hip_ctx* c, int precision, const double* d_b, double* d_x, double eps, int m,
          int* matvecs, bool* converged, double* est_resid, bool x_is_zero, double bnorm) {
    const int n = 2 * c->N;
    hipStream_t s = c->stream;
    c->V.alloc(size_t(n) * (m + 1));
    c->hvec.alloc(size_t(m + 2));
    c->hvec2.alloc(size_t(m + 2));
    c->partial.alloc(size_t(m + 2) * c->part_blocks);
    c->wtmp.alloc(size_t(m + 1));
    c->xin.alloc(size_t(n));
    c->wvec.alloc(size_t(n));
    c->jstep.alloc(1);
    if (c->hh_cap < size_t(m + 2)) {
        if (c->hh_pinned) hip_check(hipHostFree(c->hh_pinned), "hipHostFree");
        c->hh_pinned = nullptr;
        c->hh_cap = 0;
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->hh_pinned), sizeof(double) * (m + 2)),
                  "hipHostMalloc");
        c->hh_cap = size_t(m + 2);
    }
    auto NEAR0 = [](double v) { return v == 0 ? 2.2250738585072014e-308 : v; };
    const double nrm2b = bnorm >= 0 ? bnorm : std::sqrt(dev_dot(c, d_b, d_b, n));
    double* v0 = c->V.p;
    if (x_is_zero) {
        // K 0 - b = -b without a mat-vec
        launch_scale_copy(s, n, d_b, nullptr, -1.0, v0);
    } else {
        matvec_dev(c, d_x, v0, precision);
        ++*matvecs;
        launch_axpby(s, n, -1.0, d_b, 1.0, v0);
    }
    // |-b| is |b| to the bit (the same squares summed in the same order)
    const double beta = x_is_zero ? nrm2b : std::sqrt(dev_dot(c, v0, v0, n));
    launch_scale_copy(s, n, v0, nullptr, 1.0 / NEAR0(beta), v0);
    std::vector<double> U(size_t(m) * (m + 1) / 2 + 1, 0.0), y(m + 2, 0.0), cs(m + 1), sn(m + 1),
        hh(m + 2);
    y[0] = beta;
    int j = 0, uij = 0;
    if (beta == 0) {
        *converged = true;
        *est_resid = 0;
        return 0;
    }
    launch_set_int(s, c->jstep.p, 0);
    do {
        const int u0j = uij;
        // V[j+1] from V[j]: mat-vec + Gram-Schmidt + normalisation on the
        // device (one graph launch when graphs are on), then h to the host
        arnoldi_step(c, precision, m);
        ++*matvecs;
        hip_check(hipStreamSynchronize(s), "sync arnoldi");
        for (int q = 0; q <= j + 1; ++q) hh[q] = c->hh_pinned[q];
        const double h = std::sqrt(std::max(0.0, hh[j + 1]));
        for (int q = 0; q <= j; ++q) U[u0j + q] = hh[q];
        for (int i = 0; i < j; ++i) {
            const double tmp = cs[i] * U[uij] - sn[i] * U[uij + 1];
            U[uij + 1] = sn[i] * U[uij] + cs[i] * U[uij + 1];
            U[uij] = tmp;
            ++uij;
        }
        const double rd = U[uij];
        const double dd = std::sqrt(rd * rd + h * h);
        cs[j] = rd / NEAR0(dd);
        sn[j] = -h / NEAR0(dd);
        U[uij] = dd;
        ++uij;
        y[j + 1] = sn[j] * y[j];
        y[j] = cs[j] * y[j];
        ++j;
    } while (j < m && std::fabs(y[j]) > eps * nrm2b);
    *est_resid = std::fabs(y[j]);
    *converged = std::fabs(y[j]) <= eps * nrm2b;
    double su = 0, sy = 0;
    for (int k = 0; k < uij; ++k) su += std::fabs(U[k]);
    for (int k = 0; k < j; ++k) sy += std::fabs(y[k]);
    if (!(su == 0 && sy == 0)) {
        // packed column-major upper triangle: (row i, col q) at q(q+1)/2 + i
        for (int i = j - 1; i >= 0; --i) {
            double acc = y[i];
            for (int q = i + 1; q < j; ++q) acc -= U[size_t(q) * (q + 1) / 2 + i] * y[q];
            y[i] = acc / U[size_t(i) * (i + 1) / 2 + i];
        }
    }
    hip_check(hipMemcpyAsync(c->wtmp.p, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, s),
              "H2D y");
    launch_gemv_sub(s, n, c->V.p, j, c->wtmp.p, d_x);
    hip_check(hipStreamSynchronize(s), "sync gmres");
    return j;
}

}  // namespace

extern "C" {

static thread_local std::string g_create_err = "no error";

const char* beep_hip_last_error(const beep_hip_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int beep_hip_create(int device, beep_hip_ctx** out) {
    if (!out) return -2;
    *out = nullptr;
    auto c = std::make_unique<beep_hip_ctx>();
    c->device = device;
    int rc = guarded(c.get(), [&] {
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        require(device >= 0 && device < count, "no such HIP device");
        set_device(c.get());
        hip_check(hipStreamCreate(&c->stream), "hipStreamCreate");
        hip_check(hipStreamCreate(&c->far_stream), "hipStreamCreate");
        hip_check(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device),
                  "hipDeviceGetAttribute");
        c->scratch = scratch_address(device);
        hip_check(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&c->ev_pack, hipEventDisableTiming), "hipEventCreate");
        for (auto& set : c->ev)
            for (auto& e : set) hip_check(hipEventCreate(&e), "hipEventCreate");
    });
    if (rc != 0) {
        // no context survives: the message is kept for beep_hip_last_error(NULL)
        g_create_err = c->err;
        for (auto& set : c->ev)
            for (auto& e : set)
                if (e) (void)hipEventDestroy(e);
        if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
        if (c->ev_join) (void)hipEventDestroy(c->ev_join);
        if (c->ev_pack) (void)hipEventDestroy(c->ev_pack);
        if (c->far_stream) (void)hipStreamDestroy(c->far_stream);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        return rc;
    }
    *out = c.release();
    return 0;
}

int beep_hip_destroy(beep_hip_ctx* ctx) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    if (ctx->far_stream) (void)hipStreamSynchronize(ctx->far_stream);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    drop_step_graph(ctx);
    if (ctx->hh_pinned) (void)hipHostFree(ctx->hh_pinned);
    if (ctx->fmm) fmm_plan_destroy(ctx->fmm);
    for (auto& set : ctx->ev)
        for (auto& e : set)
            if (e) (void)hipEventDestroy(e);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->ev_pack) (void)hipEventDestroy(ctx->ev_pack);
    if (ctx->far_stream) (void)hipStreamDestroy(ctx->far_stream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

int beep_hip_set_geometry(beep_hip_ctx* ctx, const beep_hip_geometry* g) {
    return guarded(ctx, [&] {
        require(g && g->n_patches > 0, "empty geometry");
        set_device(ctx);
        sync_streams(ctx);
        const int N = g->n_patches;
        require(g->centroid && g->normal && g->area && g->eps_ratio && g->gc, "null patch array");
        check_ptr(N, g->qual_ptr, "bad qualocation pointer array");
        check_ptr(N, g->quad_ptr, "bad quadrature pointer array");
        check_ptr(N, g->gal_ptr, "bad galerkin pointer array");
        require(g->qual_pts && g->quad_pts && g->gal_pts, "null point array");
        for (int i = 0; i < N; ++i) {
            require(g->qual_ptr[i + 1] > g->qual_ptr[i], "patch without qualocation points");
            require(g->quad_ptr[i + 1] > g->quad_ptr[i], "patch without quadrature points");
            require(g->gal_ptr[i + 1] > g->gal_ptr[i], "patch without galerkin points");
            require(g->eps_ratio[i] > 0 && g->area[i] > 0, "non-positive area or eps ratio");
        }
        std::vector<Patch64> pr(N);
        for (int i = 0; i < N; ++i)
            pr[i] = Patch64{g->centroid[3 * i], g->centroid[3 * i + 1], g->centroid[3 * i + 2],
                            g->area[i], g->normal[3 * i], g->normal[3 * i + 1],
                            g->normal[3 * i + 2], 1.0 / g->eps_ratio[i]};
        hipStream_t s = ctx->stream;
        ctx->patches.upload(pr.data(), N, s);
        ctx->gc.upload(g->gc, N, s);
        ctx->qual_ptr.upload(g->qual_ptr, N + 1, s);
        ctx->quad_ptr.upload(g->quad_ptr, N + 1, s);
        ctx->gal_ptr.upload(g->gal_ptr, N + 1, s);
        // raw points go up in chunks through a staging buffer and are put
        // in the local frame on the device (FP64 and FP32 copies)
        constexpr int CHUNK = 1 << 20;
        ctx->stage.alloc(size_t(7) * CHUNK);
        auto points = [&](const int* hptr, const int* dptr, const double* raw, DevBuf<Pt64>& o64,
                          DevBuf<Pt32>* o32) {
            const int total = hptr[N];
            o64.alloc(size_t(total));
            if (o32) o32->alloc(size_t(total));
            for (int k0 = 0; k0 < total; k0 += CHUNK) {
                const int n = std::min(CHUNK, total - k0);
                hip_check(hipMemcpyAsync(ctx->stage.p, raw + 7 * size_t(k0), sizeof(double) * 7 * n,
                                         hipMemcpyHostToDevice, s), "H2D points");
                launch_to_local(s, n, k0, ctx->stage.p, dptr, N, ctx->patches.p, o64.p,
                                o32 ? o32->p : nullptr);
            }
        };
        points(g->qual_ptr, ctx->qual_ptr.p, g->qual_pts, ctx->qual64, &ctx->qual32);
        points(g->quad_ptr, ctx->quad_ptr.p, g->quad_pts, ctx->quad64, &ctx->quad32);
        points(g->gal_ptr, ctx->gal_ptr.p, g->gal_pts, ctx->gal64, nullptr);
        hip_check(hipStreamSynchronize(s), "sync geometry upload");
        geometry_done(ctx, N, g->centroid);
    });
}

int beep_hip_set_surface(beep_hip_ctx* ctx, const beep_hip_surface* g) {
    return guarded(ctx, [&] {
        require(g && g->n_patches > 0 && g->n_tris > 0, "empty surface");
        set_device(ctx);
        sync_streams(ctx);
        const int N = g->n_patches, T = g->n_tris;
        require(g->centroid && g->normal && g->area && g->eps_ratio && g->gc, "null patch array");
        require(g->verts && g->tris && g->vt_ptr && g->vt_tri && g->vt_corner, "null surface array");
        auto rule_ok = [](int r) { return r == 0 || r == 1 || r == 4 || r == 6 || r == 7 || r == 16; };
        require(rule_ok(g->qual_rule) && rule_ok(g->quad_rule), "rule must be 0, 1, 4, 6, 7 or 16");
        check_ptr(N, g->vt_ptr, "bad vertex-triangle pointer array");
        for (int t = 0; t < 3 * T; ++t) require(g->tris[t] >= 0 && g->tris[t] < N, "triangle vertex out of range");
        for (int p = 0; p < N; ++p) {
            require(g->vt_ptr[p + 1] > g->vt_ptr[p], "vertex without triangles");
            require(g->eps_ratio[p] > 0 && g->area[p] > 0, "non-positive area or eps ratio");
            for (int k = g->vt_ptr[p]; k < g->vt_ptr[p + 1]; ++k) {
                const int t = g->vt_tri[k], c = g->vt_corner[k];
                require(t >= 0 && t < T && c >= 0 && c < 3 && g->tris[3 * t + c] == p,
                        "vertex-triangle list does not match the triangles");
            }
        }
        const int NC = g->n_curved;
        require(NC >= 0 && (NC == 0 || (g->ctl_idx && g->ctl)), "null curved-triangle array");
        if (g->ctl_idx)
            for (int t = 0; t < T; ++t)
                require(g->ctl_idx[t] >= -1 && g->ctl_idx[t] < NC, "curved-triangle index out of range");
        std::vector<Patch64> pr(N);
        for (int i = 0; i < N; ++i)
            pr[i] = Patch64{g->centroid[3 * i], g->centroid[3 * i + 1], g->centroid[3 * i + 2],
                            g->area[i], g->normal[3 * i], g->normal[3 * i + 1],
                            g->normal[3 * i + 2], 1.0 / g->eps_ratio[i]};
        hipStream_t s = ctx->stream;
        ctx->patches.upload(pr.data(), N, s);
        ctx->gc.upload(g->gc, N, s);
        ctx->sverts.upload(g->verts, size_t(3) * N, s);
        ctx->stris.upload(g->tris, size_t(3) * T, s);
        const int nvt = g->vt_ptr[N];
        ctx->svt_ptr.upload(g->vt_ptr, N + 1, s);
        ctx->svt_tri.upload(g->vt_tri, nvt, s);
        ctx->svt_corner.upload(g->vt_corner, nvt, s);
        if (g->ctl_idx) ctx->sctl_idx.upload(g->ctl_idx, T, s);
        if (NC > 0) ctx->sctl.upload(g->ctl, size_t(63) * NC, s);
        // one point set per rule: counts (1, or 2 sub-triangles x k points per
        // triangle around the vertex), then the points on the device
        std::vector<int> ptr(size_t(N) + 1);
        auto make = [&](int order, DevBuf<int>& dptr, DevBuf<Pt64>& o64, DevBuf<Pt32>* o32,
                        int per_area) {
            RuleDev rd{};
            if (order != 0) {
                const beep::Rule& r = beep::rule_for_order(order);
                rd.k = int(r.w.size());
                for (int j = 0; j < rd.k; ++j) {
                    rd.u[j] = r.u[j];
                    rd.v[j] = r.v[j];
                    rd.w[j] = r.w[j];
                }
            }
            ptr[0] = 0;
            for (int p = 0; p < N; ++p) {
                const int64_t c = rd.k == 0 ? 1 : int64_t(2) * (g->vt_ptr[p + 1] - g->vt_ptr[p]) * rd.k;
                require(ptr[p] + c < int64_t(0x7fffffff), "too many points for 32-bit offsets");
                ptr[p + 1] = ptr[p] + int(c);
            }
            dptr.upload(ptr.data(), N + 1, s);
            o64.alloc(size_t(ptr[N]));
            if (o32) o32->alloc(size_t(ptr[N]));
            launch_gen_points(s, N, ctx->sverts.p, ctx->stris.p, ctx->svt_ptr.p, ctx->svt_tri.p,
                              ctx->svt_corner.p, g->ctl_idx ? ctx->sctl_idx.p : nullptr,
                              NC > 0 ? ctx->sctl.p : nullptr, ctx->patches.p, dptr.p, rd, per_area,
                              o64.p, o32 ? o32->p : nullptr);
            hip_check(hipStreamSynchronize(s), "sync point generation");  // ptr is reused
        };
        make(g->qual_rule, ctx->qual_ptr, ctx->qual64, &ctx->qual32, 1);
        make(g->quad_rule, ctx->quad_ptr, ctx->quad64, &ctx->quad32, 0);
        make(4, ctx->gal_ptr, ctx->gal64, nullptr, 0);
        geometry_done(ctx, N, g->centroid);
    });
}

// Forget the previous near-field structure before a new tree is set.
static void begin_tree(beep_hip_ctx* ctx) {
    ctx->share = nullptr;
    ctx->share_off.clear();
    ctx->n_pairs = 0;
    ctx->n_pair_rows = 0;
    set_device(ctx);
    sync_streams(ctx);
    ctx->gen++;
}

// The SpMV and on-the-fly plans of a tree whose leaf_idx and nb_idx are on
// the device already: per leaf the offset of its coefficient block, and the
// work items (row group x column slab) of both plans.
static void install_tree(beep_hip_ctx* ctx, int L, const int* leaf_ptr, const int* nb_ptr) {
    std::vector<int64_t> off(L + 1, 0);
    std::vector<int> cblk(L + 1, 0);
    std::vector<int4> work, groups, fwork, fgroups;
    std::vector<int> gnb, fgnb;
    for (int l = 0; l < L; ++l) {
        const int64_t R = leaf_ptr[l + 1] - leaf_ptr[l];
        const int64_t C = nb_ptr[l + 1] - nb_ptr[l];
        require(R > 0 && C > 0, "empty leaf or neighbourhood");
        off[l + 1] = off[l] + R * C;
        const int64_t nb = (R * C + 255) / 256;
        require(int64_t(cblk[l]) + nb < (int64_t(1) << 31), "too many coefficient blocks");
        cblk[l + 1] = cblk[l] + int(nb);
        for (int r0 = 0; r0 < R; r0 += 256) {
            const int first = int(work.size());
            for (int c0 = 0; c0 < C; c0 += SPMV_CW)
                work.push_back(make_int4(l, r0, c0, int(std::min<int64_t>(SPMV_CW, C - c0))));
            groups.push_back(make_int4(l, r0, int(std::min<int64_t>(256, R - r0)), first));
            gnb.push_back(int(work.size()) - first);
        }
        for (int r0 = 0; r0 < R; r0 += FLY_RB) {
            const int first = int(fwork.size());
            for (int c0 = 0; c0 < C; c0 += SPMV_CW)
                fwork.push_back(make_int4(l, r0, c0, int(std::min<int64_t>(SPMV_CW, C - c0))));
            fgroups.push_back(make_int4(l, r0, int(std::min<int64_t>(FLY_RB, R - r0)), first));
            fgnb.push_back(int(fwork.size()) - first);
        }
    }
    hipStream_t s = ctx->stream;
    ctx->leaf_ptr.upload(leaf_ptr, L + 1, s);
    ctx->nb_ptr.upload(nb_ptr, L + 1, s);
    ctx->coef_off.upload(off.data(), L + 1, s);
    ctx->coef_blk.upload(cblk.data(), L + 1, s);
    ctx->spmv_work.upload(work.data(), work.size(), s);
    ctx->spmv_groups.upload(groups.data(), groups.size(), s);
    ctx->spmv_group_nblk.upload(gnb.data(), gnb.size(), s);
    ctx->fly_work.upload(fwork.data(), fwork.size(), s);
    ctx->fly_groups.upload(fgroups.data(), fgroups.size(), s);
    ctx->fly_group_nblk.upload(fgnb.data(), fgnb.size(), s);
    ctx->spmv_part.alloc(std::max(work.size() * 256, fwork.size() * FLY_RB) * 2);
    hip_check(hipStreamSynchronize(s), "sync tree upload");
    ctx->L = L;
    ctx->nnz = off[L];
    ctx->n_coef_blocks = cblk[L];
    ctx->spmv.work = ctx->spmv_work.p;
    ctx->spmv.n_work = int(work.size());
    ctx->spmv.groups = ctx->spmv_groups.p;
    ctx->spmv.group_nblk = ctx->spmv_group_nblk.p;
    ctx->spmv.n_groups = int(groups.size());
    ctx->fly = SpmvPlan{};
    ctx->fly.work = ctx->fly_work.p;
    ctx->fly.n_work = int(fwork.size());
    ctx->fly.groups = ctx->fly_groups.p;
    ctx->fly.group_nblk = ctx->fly_group_nblk.p;
    ctx->fly.n_groups = int(fgroups.size());
    ctx->h_leaf_ptr.assign(leaf_ptr, leaf_ptr + L + 1);
    ctx->h_nb_ptr.assign(nb_ptr, nb_ptr + L + 1);
    ctx->n_nb = nb_ptr[L];
    ctx->have_tree = true;
    ctx->have32 = ctx->have64 = ctx->have_diag = false;
    if (ctx->fmm) ctx->fmm->n_res = -1;  // the residual leaf-pass list follows the near list
    // the coefficient buffers stay allocated for the next precompute
}

int beep_hip_set_tree(beep_hip_ctx* ctx, const beep_hip_tree* t) {
    return guarded(ctx, [&] {
        require(ctx->have_geom, "set the geometry before the tree");
        begin_tree(ctx);
        require(t && t->n_leaves > 0, "empty tree");
        const int N = ctx->N, L = t->n_leaves;
        check_ptr(L, t->leaf_ptr, "bad leaf pointer array");
        check_ptr(L, t->nb_ptr, "bad neighbourhood pointer array");
        require(t->leaf_ptr[L] == N, "leaves must cover every patch exactly once");
        std::vector<char> seen(N, 0);
        for (int k = 0; k < N; ++k) {
            const int i = t->leaf_idx[k];
            require(i >= 0 && i < N && !seen[i], "leaf contents are not a permutation of the patches");
            seen[i] = 1;
        }
        const int nnb = t->nb_ptr[L];
        for (int k = 0; k < nnb; ++k)
            require(t->nb_idx[k] >= 0 && t->nb_idx[k] < N, "neighbourhood index out of range");
        hipStream_t s = ctx->stream;
        ctx->leaf_idx.upload(t->leaf_idx, N, s);
        ctx->nb_idx.upload(t->nb_idx, nnb, s);
        install_tree(ctx, L, t->leaf_ptr, t->nb_ptr);
    });
}

int beep_hip_build_tree(beep_hip_ctx* ctx, const double* centre, double edge, int nbsize,
                        int* n_leaves, long long* n_nb) {
    return guarded(ctx, [&] {
        require(ctx->have_geom, "set the geometry before the tree");
        require(centre && edge > 0 && nbsize > 0, "bad bounding cube or neighbourhood size");
        begin_tree(ctx);
        const int N = ctx->N;
        hipStream_t s = ctx->stream;
        // descent keys of the centroids, sorted with their patches
        DevBuf<uint64_t>& kfull = ctx->tk_full;
        DevBuf<uint64_t>& ka = ctx->tk_a;
        DevBuf<uint64_t>& kb = ctx->tk_b;
        DevBuf<int>& ia = ctx->ti_a;
        DevBuf<int>& ib = ctx->ti_b;
        DevBuf<unsigned char>& tmp = ctx->t_tmp;
        kfull.alloc(N);
        ka.alloc(N);
        kb.alloc(N);
        ia.alloc(N);
        ib.alloc(N);
        const int key_bits = 3 * beeptree::KEY_LEVELS;
        size_t tb = tree_sort_bytes(N, key_bits);
        for (int l = 1; l < beeptree::KEY_LEVELS; ++l) tb = std::max(tb, tree_sort_bytes(N, 3 * l));
        tmp.alloc(std::max<size_t>(tb, 1));
        launch_tree_keys(s, N, ctx->patches.p, centre, edge, 0, nullptr, kfull.p, ia.p);
        tree_sort_pairs(s, tmp.p, tb, kfull.p, ka.p, ia.p, ib.p, N, key_bits);
        std::vector<uint64_t> hk(N);
        hip_check(hipMemcpyAsync(hk.data(), ka.p, sizeof(uint64_t) * N, hipMemcpyDeviceToHost, s),
                  "D2H tree keys");
        hip_check(hipStreamSynchronize(s), "sync tree keys");
        // the tree's shape from the counts the sorted keys give
        beep::Octree tree;
        tree.build_sorted(hk.data(), N, beep::Vec3(centre[0], centre[1], centre[2]), edge, nbsize);
        std::vector<int> lev, sptr, slo, shi;
        tree.leaf_segments(lev, sptr, slo, shi);
        const int L = int(lev.size());
        require(L > 0, "empty tree");
        // per level that has leaves, the patches sorted by their box at that
        // level, each box's in ascending order (stable sort from patch order)
        std::vector<int> slot(size_t(beeptree::KEY_LEVELS) + 1, -1);
        int nlev = 0;
        for (int l : lev)
            if (slot[l] < 0) slot[l] = nlev++;
        require(int64_t(nlev) * N < int64_t(0x7fffffff), "too many patches for 32-bit offsets");
        DevBuf<int>& S = ctx->t_sorted;
        S.alloc(size_t(nlev) * N);
        for (int l = 0; l <= beeptree::KEY_LEVELS; ++l) {
            if (slot[l] < 0) continue;
            int* dst = S.p + size_t(slot[l]) * N;
            if (l == 0) {
                launch_tree_keys(s, N, nullptr, nullptr, 0.0, 0, kfull.p, ka.p, dst);  // all in one box
                continue;
            }
            launch_tree_keys(s, N, nullptr, nullptr, 0.0, l, kfull.p, ka.p, ia.p);
            tree_sort_pairs(s, tmp.p, tb, ka.p, kb.p, ia.p, dst, N, 3 * l);
        }
        // the leaves' items and their neighbourhoods' items, gathered
        std::vector<int> leaf_ptr(size_t(L) + 1, 0), nb_ptr(size_t(L) + 1, 0);
        std::vector<int4> segl, segn;
        segl.reserve(L);
        segn.reserve(sptr[L]);
        for (int f = 0; f < L; ++f) {
            const size_t base = size_t(slot[lev[f]]) * N;
            const int g0 = sptr[f];
            leaf_ptr[f + 1] = leaf_ptr[f] + (shi[g0] - slo[g0]);
            segl.push_back(make_int4(int(base) + slo[g0], shi[g0] - slo[g0], leaf_ptr[f], 0));
            int64_t at = nb_ptr[f];
            for (int g = g0 + 1; g < sptr[f + 1]; ++g) {
                segn.push_back(make_int4(int(base) + slo[g], shi[g] - slo[g], int(at), 0));
                at += shi[g] - slo[g];
            }
            require(at < int64_t(0x7fffffff), "neighbourhood lists exceed 32-bit offsets");
            nb_ptr[f + 1] = int(at);
        }
        require(leaf_ptr[L] == N, "octree: leaves do not cover the patches");
        DevBuf<int4>& dsegl = ctx->t_segl;
        DevBuf<int4>& dsegn = ctx->t_segn;
        dsegl.upload(segl.data(), segl.size(), s);
        dsegn.upload(segn.data(), segn.size(), s);
        ctx->leaf_idx.alloc(N);
        ctx->nb_idx.alloc(std::max(nb_ptr[L], 1));
        launch_gather_segments(s, int(segl.size()), dsegl.p, S.p, ctx->leaf_idx.p);
        launch_gather_segments(s, int(segn.size()), dsegn.p, S.p, ctx->nb_idx.p);
        install_tree(ctx, L, leaf_ptr.data(), nb_ptr.data());  // synchronises the stream
        if (n_leaves) *n_leaves = L;
        if (n_nb) *n_nb = nb_ptr[L];
    });
}

int beep_hip_get_tree(beep_hip_ctx* ctx, int* leaf_ptr, int* leaf_idx, int* nb_ptr, int* nb_idx) {
    return guarded(ctx, [&] {
        require(ctx->have_tree, "no tree set");
        set_device(ctx);
        sync_streams(ctx);
        const int L = ctx->L;
        if (leaf_ptr) std::copy(ctx->h_leaf_ptr.begin(), ctx->h_leaf_ptr.end(), leaf_ptr);
        if (nb_ptr) std::copy(ctx->h_nb_ptr.begin(), ctx->h_nb_ptr.end(), nb_ptr);
        if (leaf_idx)
            hip_check(hipMemcpy(leaf_idx, ctx->leaf_idx.p, sizeof(int) * size_t(ctx->h_leaf_ptr[L]),
                                hipMemcpyDeviceToHost), "D2H leaf contents");
        if (nb_idx && ctx->n_nb > 0)
            hip_check(hipMemcpy(nb_idx, ctx->nb_idx.p, sizeof(int) * size_t(ctx->n_nb),
                                hipMemcpyDeviceToHost), "D2H neighbourhoods");
    });
}

// Singular terms and the near coefficients of one precision.  With the FMM
// the plan is built first: explicit pairs its leaf pass covers keep the full
// rule (their packed leaf cells are adjacent) and the leaf pass runs over
// the residual pairs only.
// Precompute of a context whose near field is shared (beep_hip_share_near):
// the listed pairs' coefficients (both precisions) for the shared sets'
// kappa; the diagonal and the other explicit pairs are the shared context's.
void precompute_shared(beep_hip_ctx* ctx, double kappa, int precision) {
    const beep_hip_ctx* h = ctx->share;
    require(h->have_diag && h->kappa == kappa, "the shared near field has another kappa");
    require(precision == BEEP_FP64 ? h->have64 : h->have32,
            "the shared near field lacks this precision's coefficients");
    sync_streams(ctx);
    ctx->gen++;
    hipStream_t s = ctx->stream;
    ctx->kappa = kappa;
    if (!ctx->guard_cnt.p) {
        ctx->guard_cnt.alloc(5);
        for (int k = 0; k < 5; ++k) launch_set_int(s, reinterpret_cast<int*>(ctx->guard_cnt.p + k), 0);
    }
    unsigned* gcnt = ctx->guard_cnt.p;
    launch_set_int(s, reinterpret_cast<int*>(gcnt + 1), 0);
    launch_set_int(s, reinterpret_cast<int*>(gcnt + 2), 0);
    launch_pair_coeffs(s, ctx->n_pairs, ctx->pair_i.p, ctx->pair_col.p, ctx->patches.p, ctx->qual_ptr.p,
                       ctx->qual64.p, ctx->quad_ptr.p, ctx->quad64.p, kappa, ctx->guard_max, gcnt + 1,
                       ctx->pcoef32.p, ctx->pcoef64.p);
    ctx->have_diag = true;
    ctx->have64 = h->have64;
    ctx->have32 = h->have32;
    ctx->uid32 = ctx->uid64 = 0;
    ensure_far(ctx, precision);
}

void precompute_impl(beep_hip_ctx* ctx, double kappa, int precision) {
    if (ctx->share) {
        require(precision == BEEP_FP64 || precision == BEEP_FP32 || precision == BEEP_FP32_RESID,
                "unknown precision");
        require(kappa >= 0, "kappa must be >= 0");
        precompute_shared(ctx, kappa, precision);
        return;
    }
    require(ctx->have_geom && ctx->have_tree, "geometry and tree must be set");
    require(kappa >= 0, "kappa must be >= 0");
    require(precision == BEEP_FP64 || precision == BEEP_FP32 || precision == BEEP_FP32_RESID,
            "unknown precision");
    sync_streams(ctx);
    ctx->gen++;
    if (ctx->have_diag && kappa != ctx->kappa) {
        ctx->have32 = ctx->have64 = false;
        ctx->coef32.reset();
        ctx->coef64.reset();
    }
    hipStream_t s = ctx->stream;
    ctx->kappa = kappa;
    ctx->diag.alloc(size_t(4) * ctx->N);
    ctx->guard_rec.alloc(beep_hip_ctx::GUARD_CAP);
    if (!ctx->guard_cnt.p) {
        ctx->guard_cnt.alloc(5);
        for (int k = 0; k < 5; ++k) launch_set_int(s, reinterpret_cast<int*>(ctx->guard_cnt.p + k), 0);
    }
    unsigned* gcnt = ctx->guard_cnt.p;
    const NsList guard{ctx->guard_rec.p, gcnt, beep_hip_ctx::GUARD_CAP};
    launch_set_int(s, reinterpret_cast<int*>(gcnt + 3), 0);
    launch_singular(s, ctx->N, ctx->patches.p, ctx->gc.p, ctx->gal_ptr.p, ctx->gal64.p, kappa, 0,
                    ctx->guard_max, gcnt + 3, ctx->diag.p);
    ctx->have_diag = true;
    const unsigned* cellkey = nullptr;
    unsigned long long uid = 0;
    if (ctx->use_fmm) {
        // the leaf level and cell keys first (the rest of the plan is built on
        // the host while the coefficient kernel runs, below)
        begin_plan(ctx);
        cellkey = ctx->fmm->cellkey;  // null above leaf level 10: no folding
        uid = ctx->fmm->uid;
        hip_check(hipEventRecord(ctx->ev_pack, ctx->far_stream), "hipEventRecord");
        hip_check(hipStreamWaitEvent(s, ctx->ev_pack, 0), "hipStreamWaitEvent");
    }
    if (ctx->near_mode == BEEP_NEAR_ON_THE_FLY) {
        // nothing stored: every mat-vec integrates the explicit pairs for
        // this kappa and far-field plan (the guard and the near-singular
        // FP64 pass run inside it)
        if (precision == BEEP_FP64) {
            ctx->have64 = true;
            ctx->uid64 = uid;
        } else {
            ctx->have32 = true;
            ctx->uid32 = uid;
            ctx->fly_ns64 = ctx->ns_fp64;
        }
        ensure_far(ctx, precision);
        if (ctx->use_fmm)
            require(!cellkey || ctx->fmm->n_res >= 0, "residual leaf-pass list missing");
        return;
    }
    size_t freeb = 0, totalb = 0;
    hip_check(hipMemGetInfo(&freeb, &totalb), "hipMemGetInfo");
    if (precision == BEEP_FP64) {
        const size_t need = size_t(ctx->nnz) * 2 * sizeof(double2);
        const size_t have = ctx->coef64.n * sizeof(double2);
        require(need <= have || need + (size_t(1) << 30) < freeb + have,
                "not enough device memory for FP64 coefficients");
        ctx->coef64.alloc(size_t(ctx->nnz) * 2);
        launch_set_int(s, reinterpret_cast<int*>(gcnt + 1), 0);
        launch_near_coeffs_fp64(s, ctx->n_coef_blocks, ctx->L, ctx->coef_blk.p, ctx->leaf_ptr.p,
                                ctx->leaf_idx.p, ctx->nb_ptr.p, ctx->nb_idx.p, ctx->coef_off.p,
                                ctx->patches.p, ctx->qual_ptr.p, ctx->qual64.p, ctx->quad_ptr.p,
                                ctx->quad64.p, kappa, 1, cellkey, guard, ctx->guard_max,
                                ctx->coef64.p);
        launch_coef_repair(s, 0, guard, ctx->patches.p, ctx->qual_ptr.p, ctx->qual64.p,
                           ctx->quad_ptr.p, ctx->quad64.p, kappa, ctx->guard_max, gcnt + 1,
                           ctx->coef64.p);
        ctx->have64 = true;
        ctx->uid64 = uid;
    } else {
        const size_t need = size_t(ctx->nnz) * sizeof(float4);
        const size_t have = ctx->coef32.n * sizeof(float4);
        require(need <= have || need + (size_t(1) << 30) < freeb + have,
                "not enough device memory for FP32 coefficients");
        ctx->coef32.alloc(size_t(ctx->nnz));
        NsList ns{};
        if (ctx->ns_fp64) {
            // room for 24 near-singular partners per row (about 12 on a
            // regular surface mesh); entries past it stay FP32
            const size_t cap = std::min<size_t>(size_t(24) * ctx->N, size_t(ctx->nnz) + 1);
            ctx->ns_rec.alloc(cap);
            ctx->ns_count.alloc(1);
            ns = NsList{ctx->ns_rec.p, ctx->ns_count.p, unsigned(cap)};
            ctx->ns_cap = unsigned(cap);
        }
        launch_set_int(s, reinterpret_cast<int*>(gcnt + 2), 0);
        launch_near_coeffs_fp32(s, ctx->n_coef_blocks, ctx->L, ctx->coef_blk.p, ctx->leaf_ptr.p,
                                ctx->leaf_idx.p, ctx->nb_ptr.p, ctx->nb_idx.p, ctx->coef_off.p,
                                ctx->patches.p, ctx->qual_ptr.p, ctx->qual32.p, ctx->quad_ptr.p,
                                ctx->quad32.p, float(kappa), 1, cellkey, ns, guard,
                                ctx->guard_max, ctx->coef32.p);
        launch_near_fix64(s, ns, ctx->patches.p, ctx->qual_ptr.p, ctx->qual64.p, ctx->quad_ptr.p,
                          ctx->quad64.p, kappa, ctx->guard_max, gcnt + 2, ctx->coef32.p);
        launch_coef_repair(s, 1, guard, ctx->patches.p, ctx->qual_ptr.p, ctx->qual64.p,
                           ctx->quad_ptr.p, ctx->quad64.p, kappa, ctx->guard_max, gcnt + 2,
                           ctx->coef32.p);
        ctx->ns_listed = ctx->ns_fp64;
        ctx->have32 = true;
        ctx->uid32 = uid;
    }
    // the rest of the plan, its residual leaf-pass list (with cell keys,
    // the list the coefficients assume) and the kernel tables are built on
    // the host and the far stream while the coefficient kernels run (no
    // wait here: the precompute is stream-ordered); FP32 coefficients serve
    // the FP32 mat-vec (FP32 far field) and the refinement residual (FP64
    // far field)
    ensure_far(ctx, precision);
    if (ctx->use_fmm)
        require(!cellkey || ctx->fmm->n_res >= 0, "residual leaf-pass list missing");
}

int beep_hip_precompute(beep_hip_ctx* ctx, double kappa, int precision) {
    return guarded(ctx, [&] {
        set_device(ctx);
        precompute_impl(ctx, kappa, precision);
    });
}

int beep_hip_set_far_field(beep_hip_ctx* ctx, int mode, int order, int leaf_target) {
    return guarded(ctx, [&] {
        require(mode == BEEP_FAR_AUTO || mode == BEEP_FAR_DIRECT || mode == BEEP_FAR_FMM ||
                    mode == BEEP_FAR_PW,
                "unknown far-field mode");
        if (mode == BEEP_FAR_PW)
            require(pw_order_supported(order), "plane-wave FMM order must be 9, 12 or 18");
        else
            require(order >= 3 && order <= 6, "FMM order must be 3..6");
        require(leaf_target >= 1, "leaf_target must be positive");
        set_device(ctx);
        ctx->far_mode = mode;
        ctx->fmm_order = order;
        ctx->fmm_leaf = leaf_target;
        if (ctx->have_geom) far_setup(ctx);  // plan and tables follow on first use
    });
}

int beep_hip_set_overlap(beep_hip_ctx* ctx, int on) {
    return guarded(ctx, [&] {
        set_device(ctx);
        sync_streams(ctx);
        timing_collect(ctx);  // event sets recorded in the old mode
        ctx->overlap = on != 0;
        ctx->gen++;
    });
}

namespace {

// Checks and state common to both ways of sharing a near field: src, the
// copies, the patch order of the direct far field and the SpMV partials.
void share_common(beep_hip_ctx* ctx, beep_hip_ctx* src, int n_copies, const int* copy_offset) {
    require(ctx->have_geom, "set the geometry first");
    require(src && src != ctx && !src->share, "the shared near field must come from another plain context");
    require(ctx->near_mode == BEEP_NEAR_PRECOMPUTE && src->near_mode == BEEP_NEAR_PRECOMPUTE,
            "a shared near field needs stored coefficients (near mode BEEP_NEAR_PRECOMPUTE)");
    require(src->have_geom && src->have_tree && src->have_diag && (src->have32 || src->have64),
            "the shared context must have its geometry, tree and coefficients");
    require(src->uid32 == 0 && src->uid64 == 0,
            "the shared coefficients must be precomputed with the direct far field");
    require(n_copies >= 1 && copy_offset, "no copies");
    const int N = ctx->N, Ns = src->N;
    require(size_t(n_copies) * Ns == size_t(N), "the copies must cover every patch");
    std::vector<char> cover(N, 0);
    for (int k = 0; k < n_copies; ++k) {
        require(copy_offset[k] >= 0 && copy_offset[k] + Ns <= N, "copy offset out of range");
        for (int i = copy_offset[k]; i < copy_offset[k] + Ns; ++i) {
            require(!cover[i], "copies overlap");
            cover[i] = 1;
        }
    }
    set_device(ctx);
    sync_streams(ctx);
    hipStream_t s = ctx->stream;
    ctx->gen++;
    ctx->have_tree = false;
    ctx->have32 = ctx->have64 = ctx->have_diag = false;
    ctx->share = src;
    ctx->ns_listed = false;
    ctx->share_off.assign(copy_offset, copy_offset + n_copies);
    // the patch order of the direct far field's source tiles (leaf_idx of a
    // plain context): each copy in the shared context's leaf order (copied on
    // this context's stream: a copy on the legacy stream would break the
    // graph capture of any other context's stream meanwhile)
    std::vector<int> src_order(Ns), order;
    order.reserve(N);
    hip_check(hipMemcpyAsync(src_order.data(), src->leaf_idx.p, sizeof(int) * size_t(Ns),
                             hipMemcpyDeviceToHost, s), "D2H shared leaf order");
    hip_check(hipStreamSynchronize(s), "sync shared leaf order");
    for (int k = 0; k < n_copies; ++k)
        for (int q = 0; q < Ns; ++q) order.push_back(src_order[q] + copy_offset[k]);
    ctx->leaf_idx.upload(order.data(), order.size(), s);
    ctx->spmv_part.alloc(size_t(2) * src->spmv.n_work * 256 * 2);  // two copies per pass
    ctx->guard_rec.alloc(beep_hip_ctx::GUARD_CAP);
    hip_check(hipStreamSynchronize(s), "sync share");  // host arrays die here
}

}  // namespace

int beep_hip_share_near(beep_hip_ctx* ctx, beep_hip_ctx* src, int n_copies, const int* copy_offset,
                        long long n_pairs, const int* pair_i, const int* pair_j) {
    return guarded(ctx, [&] {
        require(n_pairs >= 0 && (n_pairs == 0 || (pair_i && pair_j)), "bad pair list");
        const int N = ctx->N;
        // pairs sorted by row, then column; CSR over the rows that have any
        std::vector<std::pair<int, int>> pr(static_cast<size_t>(n_pairs));
        for (long long q = 0; q < n_pairs; ++q) {
            require(pair_i[q] >= 0 && pair_i[q] < N && pair_j[q] >= 0 && pair_j[q] < N &&
                        pair_i[q] != pair_j[q], "pair index out of range");
            pr[size_t(q)] = {pair_i[q], pair_j[q]};
        }
        std::sort(pr.begin(), pr.end());
        require(std::adjacent_find(pr.begin(), pr.end()) == pr.end(), "repeated pair");
        std::vector<int> rows, ptr{0}, col(pr.size()), pi(pr.size());
        for (size_t q = 0; q < pr.size(); ++q) {
            if (q == 0 || pr[q].first != pr[q - 1].first) {
                if (q) ptr.push_back(int(q));
                rows.push_back(pr[q].first);
            }
            pi[q] = pr[q].first;
            col[q] = pr[q].second;
        }
        if (!pr.empty()) ptr.push_back(int(pr.size()));
        share_common(ctx, src, n_copies, copy_offset);
        hipStream_t s = ctx->stream;
        ctx->n_pairs = n_pairs;
        ctx->n_pair_rows = int(rows.size());
        const size_t np = std::max<size_t>(pr.size(), 1);
        ctx->pair_rows.upload(rows.data(), rows.size(), s);
        ctx->pair_ptr.upload(ptr.data(), ptr.size(), s);
        ctx->pair_col.upload(col.data(), col.size(), s);
        ctx->pair_i.upload(pi.data(), pi.size(), s);
        ctx->pcoef32.alloc(np);
        ctx->pcoef64.alloc(2 * np);
        hip_check(hipStreamSynchronize(s), "sync pairs");  // host arrays die here
        // no residual leaf pass: the explicit pairs store full - 1 point
        far_setup(ctx);
    });
}

int beep_hip_share_near_cutoff(beep_hip_ctx* ctx, beep_hip_ctx* src, int n_copies, const int* copy_offset,
                               double cutoff, long long* n_pairs) {
    return guarded(ctx, [&] {
        require(cutoff > 0, "the cutoff must be positive");
        share_common(ctx, src, n_copies, copy_offset);
        const int N = ctx->N, Ns = src->N;
        hipStream_t s = ctx->stream;
        // bounding spheres (host centroids) of each copy and of each chunk of
        // 256 consecutive patches of a copy
        auto bound = [&](const double* c, int n) {
            double cx = 0, cy = 0, cz = 0;
            for (int i = 0; i < n; ++i) {
                cx += c[3 * i];
                cy += c[3 * i + 1];
                cz += c[3 * i + 2];
            }
            cx /= n;
            cy /= n;
            cz /= n;
            double r2 = 0;
            for (int i = 0; i < n; ++i) {
                const double dx = c[3 * i] - cx, dy = c[3 * i + 1] - cy, dz = c[3 * i + 2] - cz;
                r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
            }
            return make_double4(cx, cy, cz, std::sqrt(r2) * (1.0 + 1e-12));
        };
        const int nch = (Ns + 255) / 256;
        std::vector<double4> sph(n_copies), csph(size_t(n_copies) * nch);
        for (int k = 0; k < n_copies; ++k) {
            const double* c = ctx->h_centroid.data() + 3 * size_t(copy_offset[k]);
            sph[k] = bound(c, Ns);
            for (int q = 0; q < nch; ++q)
                csph[size_t(k) * nch + q] = bound(c + 3 * size_t(q) * 256, std::min(256, Ns - 256 * q));
        }
        ctx->cut_off_d.upload(copy_offset, size_t(n_copies), s);
        ctx->cut_sph_d.upload(sph.data(), sph.size(), s);
        ctx->cut_chunk_d.upload(csph.data(), csph.size(), s);
        ctx->cut_count.alloc(size_t(N) + 1);
        ctx->pair_ptr.alloc(size_t(N) + 1);
        launch_set_int(s, ctx->cut_count.p + N, 0);
        launch_cut_pairs(s, N, n_copies, Ns, ctx->cut_off_d.p, ctx->cut_sph_d.p, ctx->cut_chunk_d.p, ctx->patches.p,
                         cutoff, nullptr, ctx->cut_count.p, nullptr, nullptr);
        const size_t tb = exclusive_scan_temp_bytes(N + 1);
        ctx->cut_tmp.alloc(std::max<size_t>(tb, 1));
        launch_exclusive_scan(s, ctx->cut_tmp.p, tb, ctx->cut_count.p, ctx->pair_ptr.p, N + 1);
        int total = 0;
        hip_check(hipMemcpyAsync(&total, ctx->pair_ptr.p + N, sizeof(int), hipMemcpyDeviceToHost, s), "D2H pairs");
        hip_check(hipStreamSynchronize(s), "sync pair count");
        require(total >= 0, "too many pairs for 32-bit offsets");
        const size_t np = std::max<size_t>(size_t(total), 1);
        ctx->pair_col.alloc(np);
        ctx->pair_i.alloc(np);
        ctx->pcoef32.alloc(np);
        ctx->pcoef64.alloc(2 * np);
        ctx->pair_rows.alloc(size_t(N));
        if (total > 0)
            launch_cut_pairs(s, N, n_copies, Ns, ctx->cut_off_d.p, ctx->cut_sph_d.p, ctx->cut_chunk_d.p,
                             ctx->patches.p, cutoff, ctx->pair_ptr.p, nullptr, ctx->pair_col.p, ctx->pair_i.p);
        launch_iota(s, N, ctx->pair_rows.p);
        ctx->n_pairs = total;
        ctx->n_pair_rows = total > 0 ? N : 0;
        if (n_pairs) *n_pairs = total;
        far_setup(ctx);
    });
}

int beep_hip_set_near_singular_fp64(beep_hip_ctx* ctx, int on) {
    return guarded(ctx, [&] { ctx->ns_fp64 = on != 0; });
}

int beep_hip_set_near_mode(beep_hip_ctx* ctx, int mode) {
    return guarded(ctx, [&] {
        require(mode == BEEP_NEAR_PRECOMPUTE || mode == BEEP_NEAR_ON_THE_FLY, "unknown near-field mode");
        require(!ctx->share || mode == BEEP_NEAR_PRECOMPUTE,
                "a shared near field needs stored coefficients (near mode BEEP_NEAR_PRECOMPUTE)");
        if (mode == ctx->near_mode) return;
        set_device(ctx);
        sync_streams(ctx);
        ctx->gen++;
        ctx->near_mode = mode;
        // the precision sets of the other mode are gone: precompute again
        ctx->have32 = ctx->have64 = false;
        ctx->ns_listed = false;
        if (mode == BEEP_NEAR_ON_THE_FLY) {
            ctx->coef32.reset();
            ctx->coef64.reset();
        }
    });
}

int beep_hip_near_singular_count(beep_hip_ctx* ctx, long long* listed, long long* overflow) {
    return guarded(ctx, [&] {
        require(listed && overflow, "null output");
        set_device(ctx);
        *listed = *overflow = 0;
        if (!ctx->have32 || !ctx->ns_listed || ctx->near_mode != BEEP_NEAR_PRECOMPUTE) return;
        unsigned cnt = 0;
        hip_check(hipMemcpyAsync(&cnt, ctx->ns_count.p, sizeof(unsigned), hipMemcpyDeviceToHost,
                                 ctx->stream), "D2H near-singular count");
        hip_check(hipStreamSynchronize(ctx->stream), "sync near-singular count");
        const unsigned cap = ctx->ns_cap;
        *listed = std::min(cnt, cap);
        *overflow = cnt > cap ? cnt - cap : 0;
    });
}

int beep_hip_set_gram_schmidt(beep_hip_ctx* ctx, int passes) {
    return guarded(ctx, [&] {
        require(passes == 1 || passes == 2, "Gram-Schmidt passes must be 1 or 2");
        set_device(ctx);
        sync_streams(ctx);
        ctx->gs_passes = passes;
    });
}

int beep_hip_set_guard(beep_hip_ctx* ctx, double max_value) {
    return guarded(ctx, [&] {
        require(max_value > 0, "the guard bound must be positive");
        ctx->guard_max = max_value;
    });
}

int beep_hip_guard_stats(beep_hip_ctx* ctx, long long* repaired4) {
    return guarded(ctx, [&] {
        require(repaired4 != nullptr, "null output");
        set_device(ctx);
        for (int k = 0; k < 4; ++k) repaired4[k] = 0;
        if (!ctx->guard_cnt.p) return;
        unsigned h[5] = {0, 0, 0, 0, 0};
        hip_check(hipMemcpyAsync(h, ctx->guard_cnt.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream),
                  "D2H guard counters");
        hip_check(hipStreamSynchronize(ctx->stream), "sync guard counters");
        for (int k = 0; k < 4; ++k) repaired4[k] = h[k + 1];
    });
}

int beep_hip_sanity_check(beep_hip_ctx* ctx, long long* repaired) {
    return guarded(ctx, [&] {
        require(repaired != nullptr, "null output");
        require(ctx->have_diag, "precompute before the sanity check");
        require(!ctx->share, "the coefficients are shared: check the context that owns them");
        set_device(ctx);
        sync_streams(ctx);
        hipStream_t s = ctx->stream;
        unsigned* gcnt = ctx->guard_cnt.p;
        const NsList guard{ctx->guard_rec.p, gcnt, beep_hip_ctx::GUARD_CAP};
        launch_set_int(s, reinterpret_cast<int*>(gcnt + 4), 0);
        launch_singular(s, ctx->N, ctx->patches.p, ctx->gc.p, ctx->gal_ptr.p, ctx->gal64.p,
                        ctx->kappa, 1, ctx->guard_max, gcnt + 4, ctx->diag.p);
        for (int fp32 = 0; fp32 < 2; ++fp32) {
            // on the fly, the mat-vec guards every entry it integrates
            if (ctx->near_mode != BEEP_NEAR_PRECOMPUTE) continue;
            if (!(fp32 ? ctx->have32 : ctx->have64)) continue;
            const unsigned long long built = fp32 ? ctx->uid32 : ctx->uid64;
            require(ctx->use_fmm ? (ctx->fmm && ctx->fmm->uid == built) : built == 0,
                    "near coefficients were precomputed for another far field: call "
                    "beep_hip_precompute again");
            const unsigned* cellkey = ctx->use_fmm ? ctx->fmm->cellkey : nullptr;
            void* coef = fp32 ? static_cast<void*>(ctx->coef32.p) : static_cast<void*>(ctx->coef64.p);
            launch_coef_scan(s, fp32, ctx->n_coef_blocks, ctx->L, ctx->coef_blk.p, ctx->leaf_ptr.p,
                             ctx->leaf_idx.p, ctx->nb_ptr.p, ctx->nb_idx.p, ctx->coef_off.p, cellkey,
                             1, coef, ctx->guard_max, guard);
            launch_coef_repair(s, fp32, guard, ctx->patches.p, ctx->qual_ptr.p, ctx->qual64.p,
                               ctx->quad_ptr.p, ctx->quad64.p, ctx->kappa, ctx->guard_max, gcnt + 4,
                               coef);
        }
        unsigned h = 0;
        hip_check(hipMemcpyAsync(&h, gcnt + 4, sizeof(h), hipMemcpyDeviceToHost, s), "D2H guard count");
        hip_check(hipStreamSynchronize(s), "sync sanity check");
        *repaired = h;
        ctx->gen++;
    });
}

int beep_hip_debug_poke(beep_hip_ctx* ctx, int which, long long entry, const double* v4) {
    return guarded(ctx, [&] {
        require(v4 != nullptr, "null values");
        require(!ctx->share, "the coefficients are shared: poke the context that owns them");
        require(which == 2 || ctx->near_mode == BEEP_NEAR_PRECOMPUTE,
                "no stored coefficients on the fly (near mode BEEP_NEAR_ON_THE_FLY)");
        set_device(ctx);
        sync_streams(ctx);
        hipStream_t s = ctx->stream;
        if (which == 0) {
            require(ctx->have64 && entry >= 0 && entry < ctx->nnz, "no such FP64 coefficient");
            const double2 h[2] = {make_double2(v4[0], v4[1]), make_double2(v4[2], v4[3])};
            hip_check(hipMemcpyAsync(ctx->coef64.p + 2 * entry, h, sizeof(h), hipMemcpyHostToDevice, s),
                      "H2D poke");
        } else if (which == 1) {
            require(ctx->have32 && entry >= 0 && entry < ctx->nnz, "no such FP32 coefficient");
            const float4 h = make_float4(float(v4[0]), float(v4[1]), float(v4[2]), float(v4[3]));
            hip_check(hipMemcpyAsync(ctx->coef32.p + entry, &h, sizeof(h), hipMemcpyHostToDevice, s),
                      "H2D poke");
        } else if (which == 2) {
            require(ctx->have_diag && entry >= 0 && entry < ctx->N, "no such patch");
            hip_check(hipMemcpyAsync(ctx->diag.p + 4 * entry, v4, 4 * sizeof(double), hipMemcpyHostToDevice, s),
                      "H2D poke");
        } else {
            throw std::runtime_error("which must be 0 (FP64 set), 1 (FP32 set) or 2 (diagonal)");
        }
        hip_check(hipStreamSynchronize(s), "sync poke");
        ctx->gen++;
    });
}

int beep_hip_set_graphs(beep_hip_ctx* ctx, int on) {
    return guarded(ctx, [&] {
        set_device(ctx);
        sync_streams(ctx);
        ctx->use_graphs = on != 0;
        drop_step_graph(ctx);
    });
}

int beep_hip_far_info(beep_hip_ctx* ctx, beep_far_info* out) {
    return guarded(ctx, [&] {
        require(out != nullptr, "null output");
        *out = beep_far_info{};
        out->mode = ctx->use_fmm ? (ctx->far_mode == BEEP_FAR_PW ? BEEP_FAR_PW : BEEP_FAR_FMM)
                                 : BEEP_FAR_DIRECT;
        ensure_far(ctx);
        if (ctx->use_fmm && ctx->fmm) {
            out->order = ctx->fmm->engine == 1 ? ctx->fmm->pw_p : ctx->fmm->P;
            out->leaf_level = ctx->fmm->D;
            out->n_leaf_boxes = ctx->fmm->nleaf;
            out->n_boxes = ctx->fmm->nbox;
            out->m2l_pairs = ctx->fmm->n_pairs;
            out->p2p_pairs = ctx->fmm->n_p2p;
            out->p2p_residual_pairs = ctx->fmm->n_res;
            out->plan_ms = ctx->plan_ms;
            out->bytes_fp64 = fmm_algorithmic_bytes(ctx->fmm, 0);
            out->bytes_fp32 = fmm_algorithmic_bytes(ctx->fmm, 1);
        } else if (!ctx->use_fmm) {
            // direct sum: x and patches in, sources out, then every target
            // reads the source records once per tile pass (counted once)
            // and writes f and h
            const double N = ctx->N;
            out->bytes_fp64 = out->bytes_fp32 = N * (16.0 + 64.0 + 64.0) + N * 64.0 + 4.0 * N * 8.0;
        }
    });
}

int beep_hip_matvec_device(beep_hip_ctx* ctx, const double* dx, double* dy, int precision) {
    return guarded(ctx, [&] {
        require(dx && dy && dx != dy, "mat-vec needs distinct device vectors");
        set_device(ctx);
        matvec_dev(ctx, dx, dy, precision);
    });
}

int beep_hip_matvec_device_noalloc(beep_hip_ctx* ctx, const double* dx, double* dy, int precision) {
    return guarded(ctx, [&] {
        require(dx && dy && dx != dy, "mat-vec needs distinct device vectors");
        require(ctx->have_geom && ctx->have_tree && ctx->have_diag,
                "geometry, tree and precompute must come first");
        require(precision == BEEP_FP64 || precision == BEEP_FP32 || precision == BEEP_FP32_RESID,
                "unknown precision");
        require(precision == BEEP_FP64 ? ctx->have64 : ctx->have32, "coefficients not precomputed");
        require(!ctx->use_fmm, "the allocation-free mat-vec runs with the direct far field only");
        require((precision == BEEP_FP64 ? ctx->uid64 : ctx->uid32) == 0,
                "near coefficients were precomputed for another far field: call "
                "beep_hip_precompute again");
        require(ctx->scratch != nullptr, "no scratch address for this device");
        // carve the scratch array: sources, far vector, far partial sums,
        // SpMV row partials (256-byte aligned pieces)
        const int N = ctx->N, n2 = 2 * N;
        const int ns = far_direct_splits(N, ctx->n_cu);
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t b_src = up(sizeof(Src64) * size_t(N));
        const size_t b_yfar = up(sizeof(double) * size_t(n2));
        const size_t b_part = ns > 1 ? up(sizeof(double) * size_t(n2) * ns) : 0;
        const bool fly = ctx->near_mode == BEEP_NEAR_ON_THE_FLY;
        const size_t b_spmv = up(sizeof(double) * (fly ? size_t(ctx->fly.n_work) * FLY_RB
                                                        : size_t(ctx->spmv.n_work) * 256) * 2);
        const size_t need = b_src + b_yfar + b_part + b_spmv;
        if (need > SCRATCH_BYTES)
            throw std::runtime_error("the allocation-free mat-vec needs " + std::to_string(need) +
                                     " bytes of scratch for " + std::to_string(N) +
                                     " patches; its limit is " + std::to_string(SCRATCH_BYTES));
        set_device(ctx);
        unsigned char* base = static_cast<unsigned char*>(ctx->scratch);
        Src64* src = reinterpret_cast<Src64*>(base);
        double* yfar = reinterpret_cast<double*>(base + b_src);
        double* part = ns > 1 ? reinterpret_cast<double*>(base + b_src + b_yfar) : nullptr;
        double* spart = reinterpret_cast<double*>(base + b_src + b_yfar + b_part);
        hipStream_t s = ctx->stream;
        // the kernels of matvec_dev with the direct far field, in its order,
        // all on the context stream
        const int mode = spmv_mode(precision);
        if (fly)
            launch_near_fly(s, mode, N, ctx->fly, ctx->leaf_ptr.p, ctx->leaf_idx.p, ctx->nb_ptr.p,
                            ctx->nb_idx.p, ctx->patches.p, ctx->qual_ptr.p, ctx->qual32.p, ctx->qual64.p,
                            ctx->quad_ptr.p, ctx->quad32.p, ctx->quad64.p, ctx->kappa, nullptr,
                            ctx->fly_ns64 ? 1 : 0, ctx->guard_max, ctx->diag.p, dx, spart, dy);
        else
            launch_near_spmv(s, mode, N, ctx->spmv, ctx->leaf_ptr.p, ctx->leaf_idx.p, ctx->nb_ptr.p,
                             ctx->nb_idx.p, ctx->coef_off.p, ctx->coef32.p, ctx->coef64.p, ctx->diag.p,
                             dx, spart, dy, ctx->spmv_cap[mode] * ctx->n_cu, ctx->spmv_depth[mode]);
        launch_fill(s, n2, 0.0, yfar);
        launch_far_pack(s, N, ctx->patches.p, ctx->leaf_idx.p, dx, src);
        launch_far_direct(s, precision == BEEP_FP32 ? 1 : 0, N, ns, ctx->leaf_idx.p, ctx->patches.p,
                          src, ctx->kappa, yfar, part);
        launch_axpby(s, n2, 1.0, yfar, 1.0, dy);
    });
}

int beep_hip_autotune_leaf(beep_hip_ctx* ctx, int precision, const int* leaf_targets, int n,
                           int reps, double* ms_out, int* best) {
    return guarded(ctx, [&] {
        require(leaf_targets && n > 0 && reps > 0, "bad autotune arguments");
        require(ctx->have_geom && ctx->have_tree && ctx->have_diag, "precompute before tuning");
        set_device(ctx);
        const int saved_mode = ctx->far_mode;
        // an FMM engine: the plane-wave one when selected, else Chebyshev
        ctx->far_mode = saved_mode == BEEP_FAR_PW ? BEEP_FAR_PW : BEEP_FAR_FMM;
        const size_t n2 = size_t(2) * ctx->N;
        ctx->xbuf.alloc(n2);
        ctx->ybuf.alloc(n2);
        launch_fill(ctx->stream, int(n2), 1.0, ctx->xbuf.p);
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        double best_ms = 1e300;
        int best_leaf = ctx->fmm_leaf;
        try {
            for (int k = 0; k < n; ++k) {
                require(leaf_targets[k] >= 1, "leaf targets must be positive");
                ctx->fmm_leaf = leaf_targets[k];
                far_setup(ctx);
                precompute_impl(ctx, ctx->kappa, precision);  // coefficients follow the plan
                matvec_dev(ctx, ctx->xbuf.p, ctx->ybuf.p, precision);  // tables + warm-up
                hip_check(hipEventRecord(e0, ctx->stream), "hipEventRecord");
                for (int r = 0; r < reps; ++r) matvec_dev(ctx, ctx->xbuf.p, ctx->ybuf.p, precision);
                hip_check(hipEventRecord(e1, ctx->stream), "hipEventRecord");
                hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                float ms = 0;
                hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                const double per = double(ms) / reps;
                if (ms_out) ms_out[k] = per;
                if (per < best_ms) {
                    best_ms = per;
                    best_leaf = leaf_targets[k];
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            ctx->far_mode = saved_mode;
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        timing_collect(ctx);
        ctx->timing = beep_matvec_timing{};
        ctx->fmm_leaf = best_leaf;
        far_setup(ctx);
        // every coefficient set the context held, for the chosen plan
        const bool f64 = ctx->have64, f32 = ctx->have32;
        if (f64) precompute_impl(ctx, ctx->kappa, BEEP_FP64);
        if (f32) precompute_impl(ctx, ctx->kappa, BEEP_FP32);
        if (best) *best = best_leaf;
    });
}

int beep_hip_set_spmv_residency(beep_hip_ctx* ctx, int blocks_per_cu) {
    return guarded(ctx, [&] {
        require(blocks_per_cu >= 0 && blocks_per_cu <= 64, "blocks per CU must be 0..64");
        set_device(ctx);
        sync_streams(ctx);
        for (int& v : ctx->spmv_cap) v = blocks_per_cu;
        ctx->gen++;
    });
}

int beep_hip_set_spmv_depth(beep_hip_ctx* ctx, int depth) {
    return guarded(ctx, [&] {
        require(depth == 1 || depth == 2 || depth == 4, "SpMV depth must be 1, 2 or 4");
        set_device(ctx);
        sync_streams(ctx);
        for (int& v : ctx->spmv_depth) v = depth;
        ctx->gen++;
    });
}

// Times `reps` whole mat-vecs (after one warm-up) for each candidate value of
// one SpMV setting, keeps the fastest.  `table` is the per-mode setting; the
// entry of the precision's mode is tuned.
static void tune_spmv_setting(beep_hip_ctx* ctx, int precision, int (beep_hip_ctx::*table)[3],
                              const int* values, int n, int reps, bool (*valid)(int),
                              const char* what, double* ms_out, int* best) {
    require(precision == BEEP_FP64 || precision == BEEP_FP32 || precision == BEEP_FP32_RESID,
            "unknown precision");
    int* field = &(ctx->*table)[spmv_mode(precision)];
    require(values && n > 0 && reps > 0, "bad autotune arguments");
    require(ctx->have_geom && ctx->have_tree && ctx->have_diag, "precompute before tuning");
    require(ctx->near_mode == BEEP_NEAR_PRECOMPUTE, "the SpMV settings apply to stored coefficients only");
    for (int k = 0; k < n; ++k) require(valid(values[k]), what);
    set_device(ctx);
    const size_t n2 = size_t(2) * ctx->N;
    ctx->xbuf.alloc(n2);
    ctx->ybuf.alloc(n2);
    launch_fill(ctx->stream, int(n2), 1.0, ctx->xbuf.p);
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    const int saved = *field;
    double best_ms = 1e300;
    int best_v = saved;
    try {
        for (int k = 0; k < n; ++k) {
            *field = values[k];
            matvec_dev(ctx, ctx->xbuf.p, ctx->ybuf.p, precision);  // warm-up
            hip_check(hipEventRecord(e0, ctx->stream), "hipEventRecord");
            for (int r = 0; r < reps; ++r) matvec_dev(ctx, ctx->xbuf.p, ctx->ybuf.p, precision);
            hip_check(hipEventRecord(e1, ctx->stream), "hipEventRecord");
            hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            const double per = double(ms) / reps;
            if (ms_out) ms_out[k] = per;
            if (per < best_ms) {
                best_ms = per;
                best_v = values[k];
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *field = saved;
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    timing_collect(ctx);
    ctx->timing = beep_matvec_timing{};
    *field = best_v;
    ctx->gen++;
    if (best) *best = best_v;
}

int beep_hip_autotune_spmv(beep_hip_ctx* ctx, int precision, const int* blocks_per_cu, int n,
                           int reps, double* ms_out, int* best) {
    return guarded(ctx, [&] {
        tune_spmv_setting(ctx, precision, &beep_hip_ctx::spmv_cap, blocks_per_cu, n, reps,
                          [](int v) { return v >= 0 && v <= 64; }, "blocks per CU must be 0..64",
                          ms_out, best);
    });
}

int beep_hip_autotune_spmv_depth(beep_hip_ctx* ctx, int precision, const int* depths, int n,
                                 int reps, double* ms_out, int* best) {
    return guarded(ctx, [&] {
        tune_spmv_setting(ctx, precision, &beep_hip_ctx::spmv_depth, depths, n, reps,
                          [](int v) { return v == 1 || v == 2 || v == 4; },
                          "SpMV depth must be 1, 2 or 4", ms_out, best);
    });
}

int beep_hip_matvec(beep_hip_ctx* ctx, const double* x, double* y, int precision) {
    return guarded(ctx, [&] {
        require(x && y, "null vector");
        set_device(ctx);
        const size_t n = size_t(2) * ctx->N;
        ctx->xbuf.alloc(n);
        ctx->ybuf.alloc(n);
        hip_check(hipMemcpyAsync(ctx->xbuf.p, x, n * sizeof(double), hipMemcpyHostToDevice,
                                 ctx->stream), "H2D x");
        matvec_dev(ctx, ctx->xbuf.p, ctx->ybuf.p, precision);
        hip_check(hipMemcpyAsync(y, ctx->ybuf.p, n * sizeof(double), hipMemcpyDeviceToHost,
                                 ctx->stream), "D2H y");
        hip_check(hipStreamSynchronize(ctx->stream), "sync matvec");
    });
}

int beep_hip_matvec_timing(beep_hip_ctx* ctx, beep_matvec_timing* out) {
    return guarded(ctx, [&] {
        require(out != nullptr, "null output");
        set_device(ctx);
        timing_collect(ctx);
        *out = ctx->timing;
    });
}

int beep_hip_timing_reset(beep_hip_ctx* ctx) {
    return guarded(ctx, [&] {
        set_device(ctx);
        timing_collect(ctx);
        ctx->timing = beep_matvec_timing{};
    });
}

int beep_hip_rhs(beep_hip_ctx* ctx, int n_charges, const double* xyzq, double Dsolvent,
                 double* b) {
    return guarded(ctx, [&] {
        require(ctx->have_geom, "geometry not set");
        require(n_charges >= 0 && (n_charges == 0 || xyzq) && b && Dsolvent > 0, "bad rhs args");
        set_device(ctx);
        const int N = ctx->N;
        DevBuf<double>& q = ctx->qbuf;
        DevBuf<double>& out = ctx->rhsbuf;
        if (n_charges > 0) q.upload(xyzq, size_t(4) * n_charges, ctx->stream);
        out.alloc(size_t(2) * N);
        launch_rhs(ctx->stream, N, ctx->qual_ptr.p, ctx->qual64.p, ctx->patches.p, n_charges,
                   n_charges > 0 ? q.p : nullptr, FOUR_PI_INV / Dsolvent, out.p);
        hip_check(hipMemcpyAsync(b, out.p, sizeof(double) * 2 * N, hipMemcpyDeviceToHost,
                                 ctx->stream), "D2H rhs");
        hip_check(hipStreamSynchronize(ctx->stream), "sync rhs");
    });
}

int beep_hip_rhs_device(beep_hip_ctx* ctx, int n_charges, const double* xyzq, double Dsolvent,
                        double* d_b) {
    return guarded(ctx, [&] {
        require(ctx->have_geom, "geometry not set");
        require(n_charges >= 0 && (n_charges == 0 || xyzq) && d_b && Dsolvent > 0,
                "bad rhs args");
        set_device(ctx);
        DevBuf<double>& q = ctx->qbuf;
        if (n_charges > 0) q.upload(xyzq, size_t(4) * n_charges, ctx->stream);
        launch_rhs(ctx->stream, ctx->N, ctx->qual_ptr.p, ctx->qual64.p, ctx->patches.p, n_charges,
                   n_charges > 0 ? q.p : nullptr, FOUR_PI_INV / Dsolvent, d_b);
        hip_check(hipStreamSynchronize(ctx->stream), "sync rhs");  // the charges buffer is reused
    });
}

int beep_hip_energy_coefficients(beep_hip_ctx* ctx, int n_inst, const int* ipp, const int* icp,
                                 const double* xyzq, double* ec_f, double* ec_h) {
    return guarded(ctx, [&] {
        require(ctx->have_geom, "geometry not set");
        require(n_inst > 0 && ipp && icp && ec_f && ec_h, "bad energy args");
        require(ipp[0] == 0 && ipp[n_inst] == ctx->N && icp[0] == 0, "bad instance offsets");
        set_device(ctx);
        const int N = ctx->N;
        std::vector<int> pinst(N);
        for (int k = 0; k < n_inst; ++k) {
            require(ipp[k + 1] >= ipp[k] && icp[k + 1] >= icp[k], "bad instance offsets");
            for (int i = ipp[k]; i < ipp[k + 1]; ++i) pinst[i] = k;
        }
        const int nch = icp[n_inst];
        DevBuf<double>& q = ctx->qbuf;
        DevBuf<double>& ef = ctx->ecf;
        DevBuf<double>& eh = ctx->ech;
        DevBuf<int>& pi = ctx->pinst_d;
        DevBuf<int>& cp = ctx->icp_d;
        if (nch > 0) q.upload(xyzq, size_t(4) * nch, ctx->stream);
        pi.upload(pinst.data(), N, ctx->stream);
        cp.upload(icp, n_inst + 1, ctx->stream);
        ef.alloc(N);
        eh.alloc(N);
        launch_energy_coeffs(ctx->stream, N, ctx->gal_ptr.p, ctx->gal64.p, ctx->patches.p, pi.p,
                             cp.p, q.p, ef.p, eh.p);
        hip_check(hipMemcpyAsync(ec_f, ef.p, sizeof(double) * N, hipMemcpyDeviceToHost,
                                 ctx->stream), "D2H ec_f");
        hip_check(hipMemcpyAsync(ec_h, eh.p, sizeof(double) * N, hipMemcpyDeviceToHost,
                                 ctx->stream), "D2H ec_h");
        hip_check(hipStreamSynchronize(ctx->stream), "sync energy");
    });
}

// The solve on device vectors: b read-only, x the initial guess on entry and
// the solution on exit (x_zero: the guess is zero, a cold start).
void solve_impl(beep_hip_ctx* ctx, const double* d_b, double* d_x, bool x_zero,
                const beep_solve_opts* o, beep_solve_stats* st) {
    require(o->max_iter > 0 && o->tol > 0, "bad solve options");
    const int n = 2 * ctx->N;
    hipStream_t s = ctx->stream;
    beep_solve_stats stats{};
    const double t0 = now_ms();
    ctx->dvec.alloc(n);
    ctx->rvec.alloc(n);
    const double bnorm = std::sqrt(dev_dot(ctx, d_b, d_b, n));
    int matvecs = 0;
    if (o->mode == BEEP_SOLVE_FP64 || o->mode == BEEP_SOLVE_FP32) {
        const int prec = o->mode == BEEP_SOLVE_FP64 ? BEEP_FP64 : BEEP_FP32;
        bool conv = false;
        double est = 0;
        stats.iterations = gmres(ctx, prec, d_b, d_x, o->tol, o->max_iter, &matvecs, &conv, &est,
                                 x_zero, bnorm);
        stats.converged = conv;
        stats.residual = est;
        stats.rel_residual = bnorm > 0 ? est / bnorm : 0;
    } else if (o->mode == BEEP_SOLVE_MIXED) {
        require(o->max_refine > 0, "bad refinement options");
        const int resid_prec = o->resid_fp64 ? BEEP_FP64 : BEEP_FP32_RESID;
        require(!o->resid_fp64 || ctx->have64, "an FP64 residual needs the FP64 precompute");
        // inner FP32 GMRES tolerance; <= 0: half the outer one (one
        // refinement step then usually suffices)
        const double inner_tol = o->inner_tol > 0 ? o->inner_tol : 0.5 * o->tol;
        bool done = false;
        double rn = 0;
        for (int k = 0; k <= o->max_refine; ++k) {
            // r = b - K x with the FP64 residual operator
            if (k == 0 && x_zero) {
                hip_check(hipMemcpyAsync(ctx->rvec.p, d_b, sizeof(double) * n,
                                         hipMemcpyDeviceToDevice, s), "D2D r");
            } else {
                matvec_dev(ctx, d_x, ctx->rvec.p, resid_prec);
                ++matvecs;
                launch_axpby(s, n, 1.0, d_b, -1.0, ctx->rvec.p);
            }
            // r = b on a cold start: its norm is |b|
            rn = (k == 0 && x_zero) ? bnorm : std::sqrt(dev_dot(ctx, ctx->rvec.p, ctx->rvec.p, n));
            if (o->verbose)
                fprintf(stderr, "refine %d: |r|/|b| = %.3e\n", k, bnorm > 0 ? rn / bnorm : 0.0);
            if (rn <= o->tol * bnorm) {
                done = true;
                break;
            }
            if (k == o->max_refine) break;
            launch_fill(s, n, 0.0, ctx->dvec.p);
            bool conv = false;
            double est = 0;
            const int it = gmres(ctx, BEEP_FP32, ctx->rvec.p, ctx->dvec.p, inner_tol,
                                 o->max_iter, &matvecs, &conv, &est, true, rn);
            if (o->verbose)
                fprintf(stderr, "inner %d: %d iterations, estimate |r - K d|/|r| = %.3e%s\n", k,
                        it, rn > 0 ? est / rn : 0.0, conv ? "" : " (not converged)");
            stats.iterations += it;
            launch_axpby(s, n, 1.0, ctx->dvec.p, 1.0, d_x);
            stats.refinements = k + 1;
        }
        stats.converged = done;
        stats.residual = rn;
        stats.rel_residual = bnorm > 0 ? rn / bnorm : 0;
    } else {
        throw std::runtime_error("unknown solve mode");
    }
    hip_check(hipStreamSynchronize(s), "sync solve");
    stats.matvecs = matvecs;
    stats.time_ms = now_ms() - t0;
    if (st) *st = stats;
}

int beep_hip_solve(beep_hip_ctx* ctx, const double* b, double* x, const beep_solve_opts* o,
                   beep_solve_stats* st) {
    return guarded(ctx, [&] {
        require(b && x && o, "bad solve args");
        set_device(ctx);
        const int n = 2 * ctx->N;
        hipStream_t s = ctx->stream;
        ctx->bvec.alloc(n);
        ctx->xbuf.alloc(n);
        hip_check(hipMemcpyAsync(ctx->bvec.p, b, sizeof(double) * n, hipMemcpyHostToDevice, s), "H2D b");
        hip_check(hipMemcpyAsync(ctx->xbuf.p, x, sizeof(double) * n, hipMemcpyHostToDevice, s), "H2D x");
        bool x_zero = true;
        for (int k = 0; k < n && x_zero; ++k) x_zero = (x[k] == 0.0);
        solve_impl(ctx, ctx->bvec.p, ctx->xbuf.p, x_zero, o, st);
        hip_check(hipMemcpyAsync(x, ctx->xbuf.p, sizeof(double) * n, hipMemcpyDeviceToHost, s), "D2H x");
        hip_check(hipStreamSynchronize(s), "sync solve");
    });
}

int beep_hip_solve_from_charges(beep_hip_ctx* ctx, int n_charges, const double* xyzq,
                                double Dsolvent, const beep_solve_opts* o, double* x,
                                beep_solve_stats* st) {
    return guarded(ctx, [&] {
        require(ctx->have_geom, "geometry not set");
        require(n_charges >= 0 && (n_charges == 0 || xyzq) && x && o && Dsolvent > 0,
                "bad solve args");
        set_device(ctx);
        const int n = 2 * ctx->N;
        hipStream_t s = ctx->stream;
        ctx->bvec.alloc(n);
        ctx->xbuf.alloc(n);
        DevBuf<double>& q = ctx->qbuf;
        if (n_charges > 0) q.upload(xyzq, size_t(4) * n_charges, s);
        launch_rhs(s, ctx->N, ctx->qual_ptr.p, ctx->qual64.p, ctx->patches.p, n_charges,
                   n_charges > 0 ? q.p : nullptr, FOUR_PI_INV / Dsolvent, ctx->bvec.p);
        launch_fill(s, n, 0.0, ctx->xbuf.p);
        solve_impl(ctx, ctx->bvec.p, ctx->xbuf.p, true, o, st);
        hip_check(hipMemcpyAsync(x, ctx->xbuf.p, sizeof(double) * n, hipMemcpyDeviceToHost, s), "D2H x");
        hip_check(hipStreamSynchronize(s), "sync solve");
    });
}

int beep_hip_solve_device(beep_hip_ctx* ctx, const double* d_b, double* d_x,
                          const beep_solve_opts* o, beep_solve_stats* st) {
    return guarded(ctx, [&] {
        require(d_b && d_x && o && d_b != d_x, "bad solve args");
        set_device(ctx);
        const int n = 2 * ctx->N;
        const bool x_zero = dev_dot(ctx, d_x, d_x, n) == 0.0;
        solve_impl(ctx, d_b, d_x, x_zero, o, st);
    });
}

int beep_hip_pw_selftest(int order, int nodes, double beta, double* err5) {
    try {
        if (!err5) throw std::runtime_error("null output");
        if (order < 1 || order > PW_MAXP) throw std::runtime_error("plane-wave order out of range");
        if (nodes != 9 && nodes != 18) throw std::runtime_error("quadrature nodes must be 9 or 18");
        if (!(beta >= 0)) throw std::runtime_error("beta must be >= 0");
        pw_selftest(order, nodes, beta, err5);
        return 0;
    } catch (const std::exception& e) {
        g_create_err = e.what();
        return -1;
    }
}

int beep_hip_fmm_symmetry(int tx, int ty, int tz, int order, int* canon_xyz, int* node_map) {
    try {
        if (!canon_xyz || !node_map) throw std::runtime_error("null output");
        if (std::max(std::abs(tx), std::max(std::abs(ty), std::abs(tz))) < 2 ||
            std::max(std::abs(tx), std::max(std::abs(ty), std::abs(tz))) > 3)
            throw std::runtime_error("offset is not an M2L offset");
        return fmm_canonical(tx, ty, tz, order, canon_xyz, node_map);
    } catch (const std::exception& e) {
        g_create_err = e.what();
        return -1;
    }
}

int beep_hip_selftest_mfma(beep_hip_ctx* ctx, int fp32, const double* a, const double* b,
                           double* d) {
    return guarded(ctx, [&] {
        require(a && b && d, "null array");
        set_device(ctx);
        DevBuf<double> da, db, dd;
        da.upload(a, 64, ctx->stream);
        db.upload(b, 64, ctx->stream);
        dd.alloc(256);
        launch_mfma_selftest(ctx->stream, da.p, db.p, fp32, dd.p);
        hip_check(hipMemcpyAsync(d, dd.p, 256 * sizeof(double), hipMemcpyDeviceToHost,
                                 ctx->stream), "D2H selftest");
        hip_check(hipStreamSynchronize(ctx->stream), "sync selftest");
    });
}

int beep_hip_device_memory(int device, size_t* free_bytes, size_t* total_bytes) {
    try {
        if (!free_bytes || !total_bytes) throw std::runtime_error("null output");
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
        return 0;
    } catch (const std::exception& e) {
        g_create_err = e.what();
        return -1;
    }
}

int beep_hip_device_alloc(beep_hip_ctx* ctx, size_t bytes, void** out) {
    return guarded(ctx, [&] {
        require(out != nullptr, "null output");
        set_device(ctx);
        hip_check(hipMalloc(out, bytes), "hipMalloc");
    });
}

int beep_hip_device_free(beep_hip_ctx* ctx, void* p) {
    return guarded(ctx, [&] {
        set_device(ctx);
        if (p) hip_check(hipFree(p), "hipFree");
    });
}

int beep_hip_memcpy(beep_hip_ctx* ctx, void* dst, const void* src, size_t bytes, int kind) {
    return guarded(ctx, [&] {
        set_device(ctx);
        hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                          : kind == 2 ? hipMemcpyDeviceToHost
                          : kind == 3 ? hipMemcpyDeviceToDevice
                                      : hipMemcpyDefault;
        require(kind >= 1 && kind <= 3, "memcpy kind must be 1, 2 or 3");
        hip_check(hipMemcpyAsync(dst, src, bytes, k, ctx->stream), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(ctx->stream), "sync memcpy");
    });
}

int beep_hip_gather_rows64(beep_hip_ctx* ctx, const void* src, const int* idx, size_t rows, void* dst) {
    return guarded(ctx, [&] {
        require(src && idx && dst && rows > 0, "bad gather arguments");
        require(rows <= (size_t(1) << 33), "too many rows");
        set_device(ctx);
        launch_gather_rows64(ctx->stream, src, idx, rows, dst);
    });
}

int beep_hip_synchronize(beep_hip_ctx* ctx) {
    return guarded(ctx, [&] {
        set_device(ctx);
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}

int beep_hip_memory_held(beep_hip_ctx* ctx, size_t* bytes) {
    return guarded(ctx, [&] {
        require(bytes != nullptr, "null output");
        size_t b = 0;
        auto add = [&](const auto& buf) { b += buf.n * sizeof(*buf.p); };
        add(ctx->patches); add(ctx->gc); add(ctx->qual_ptr); add(ctx->quad_ptr); add(ctx->gal_ptr);
        add(ctx->qual64); add(ctx->quad64); add(ctx->gal64); add(ctx->qual32); add(ctx->quad32);
        add(ctx->stage); add(ctx->sverts); add(ctx->stris); add(ctx->svt_ptr); add(ctx->svt_tri);
        add(ctx->svt_corner);
        add(ctx->leaf_ptr); add(ctx->leaf_idx); add(ctx->nb_ptr); add(ctx->nb_idx);
        add(ctx->coef_off); add(ctx->coef_blk); add(ctx->spmv_work); add(ctx->spmv_groups);
        add(ctx->spmv_group_nblk); add(ctx->spmv_part); add(ctx->diag); add(ctx->coef32);
        add(ctx->coef64); add(ctx->src); add(ctx->xbuf); add(ctx->ybuf); add(ctx->yfar);
        add(ctx->far_part);
        add(ctx->qbuf); add(ctx->rhsbuf); add(ctx->ecf); add(ctx->ech); add(ctx->pinst_d);
        add(ctx->icp_d); add(ctx->V); add(ctx->hvec); add(ctx->partial); add(ctx->wtmp);
        add(ctx->xin); add(ctx->wvec); add(ctx->jstep);
        add(ctx->bvec); add(ctx->rvec); add(ctx->dvec);
        add(ctx->pair_rows); add(ctx->pair_ptr); add(ctx->pair_col); add(ctx->pair_i);
        add(ctx->pcoef32); add(ctx->pcoef64); add(ctx->ns_rec); add(ctx->guard_rec); add(ctx->hvec2);
        add(ctx->cut_off_d); add(ctx->cut_count); add(ctx->cut_sph_d); add(ctx->cut_chunk_d); add(ctx->cut_tmp);
        add(ctx->sctl_idx); add(ctx->sctl); add(ctx->tk_full); add(ctx->tk_a); add(ctx->tk_b);
        add(ctx->ti_a); add(ctx->ti_b); add(ctx->t_sorted); add(ctx->t_tmp); add(ctx->t_segl);
        add(ctx->t_segn);
        *bytes = b + ctx->fmm_arena.held;
    });
}

int beep_hip_memory_info(beep_hip_ctx* ctx, size_t* free_bytes, size_t* total_bytes) {
    return guarded(ctx, [&] {
        set_device(ctx);
        size_t f = 0, t = 0;
        hip_check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
    });
}

}  // extern "C"
