// Solve server (include/beep_serve.h): worker slots with resident device
// contexts, admission by a device-memory budget, continuous batching of
// requests as they arrive.  Host C++ only; the device work goes through the
// C ABI of include/beep_hip.h.
#include "../../include/beep_serve.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/beep_hip.h"
#include "host.hpp"

namespace {

// src/bem/constants.h:14-51; energy in kJ/mol (mesh.cpp:1373-1403)
constexpr double AVOGADRO = 6.02214179e23;
constexpr double EPSILON0 = 8.85418782e-12;
constexpr double ELEMENTARY_CHARGE = 1.60217646e-19;
constexpr double ANGSTROM = 1.0e-10;
constexpr double KJ_PER_MOL =
    AVOGADRO * ELEMENTARY_CHARGE * ELEMENTARY_CHARGE / (EPSILON0 * ANGSTROM * 1000.0);

struct Job {
    long long id;
    beep_scene* scene;
    double t_submit;
};

struct Done {
    beep_result r{};
    std::vector<double> f, h;
};

void check_hip(beep_hip_ctx* c, int rc, const char* what) {
    if (rc != 0)
        throw std::runtime_error(std::string(what) + ": " + beep_hip_last_error(c));
}

// device bytes a scene will need in a context (upper estimate, reconciled
// with beep_hip_memory_held after the request): point sets in both
// precisions, coefficients of the solve's precisions, GMRES basis and
// per-patch FMM storage (tables, expansions, M2L slots: about 12 KB/patch)
double estimate_bytes(const beep_scene* s, const beep_server_opts& o, bool stored) {
    const double N = double(s->centroid.size());
    // point counts from the offsets: the point sets themselves are empty when
    // the device generates them
    auto count = [&](int k) { return s->ptr[k].empty() ? 0.0 : double(s->ptr[k].back()); };
    const double nq = count(0), nd = count(1), ng = count(2);
    double nnz = 0;
    if (s->with_tree) {
        for (size_t L = 0; L + 1 < s->leaf_ptr.size(); ++L)
            nnz += double(s->leaf_ptr[L + 1] - s->leaf_ptr[L]) * double(s->nb_ptr[L + 1] - s->nb_ptr[L]);
    } else {
        // no tree yet (the device builds it after admission): the octree is
        // refined until its patches interact with fewer than nbsize + 1
        // patches each on average (optimize_neighbourhoods, octree.h:867-986)
        nnz = N * std::min(N, double(s->nbsize) + 1.0);
    }
    double b = 96.0 * (nq + nd) + 64.0 * ng + 64.0 * N;
    if (stored && o.solve_mode != BEEP_SOLVE_FP64) b += 16.0 * nnz;
    if (stored && o.solve_mode == BEEP_SOLVE_FP64) b += 32.0 * nnz;
    b += 16.0 * N * double(o.max_iter + 1) + 16.0 * N * 8.0;
    b += 12288.0 * N;
    return b;
}

// The built scene's surface to the context (placed vertices, triangles,
// triangles around each vertex, the curved triangles' control points, patch
// arrays); the device generates the point sets.
void upload_surface(beep_hip_ctx* ctx, beep_scene* sc) {
    const int N = int(sc->centroid.size());
    beep_scene_sizes sz{};
    if (beep_scene_get_sizes(sc, &sz) != 0)
        throw std::runtime_error(std::string("scene sizes: ") + beep_host_last_error());
    // every element is written by the export: no zero fill (about 20 MB at
    // 163,842 patches)
    std::unique_ptr<double[]> verts(new double[size_t(3) * N]);
    std::unique_ptr<int[]> tris(new int[size_t(3) * sz.n_triangles]), vt_ptr(new int[size_t(N) + 1]),
        vt_tri(new int[size_t(sz.n_vertex_triangles)]), vt_corner(new int[size_t(sz.n_vertex_triangles)]);
    if (beep_scene_surface(sc, verts.get(), tris.get(), vt_ptr.get(), vt_tri.get(), vt_corner.get()) != 0)
        throw std::runtime_error(std::string("scene surface: ") + beep_host_last_error());
    int nc = 0;
    std::unique_ptr<int[]> ctl_idx(new int[size_t(sz.n_triangles)]);
    std::vector<double> ctl;
    if (beep_scene_surface_curved(sc, &nc, nullptr, nullptr) != 0)
        throw std::runtime_error(std::string("scene surface: ") + beep_host_last_error());
    ctl.resize(size_t(63) * nc);
    if (nc > 0 && beep_scene_surface_curved(sc, nullptr, ctl_idx.get(), ctl.data()) != 0)
        throw std::runtime_error(std::string("scene surface: ") + beep_host_last_error());
    beep_hip_surface g{};
    g.n_patches = N;
    g.centroid = &sc->centroid[0].x;
    g.normal = &sc->normal[0].x;
    g.area = sc->area.data();
    g.eps_ratio = sc->eps.data();
    g.gc = sc->gc.data();
    g.verts = verts.get();
    g.n_tris = sz.n_triangles;
    g.tris = tris.get();
    g.vt_ptr = vt_ptr.get();
    g.vt_tri = vt_tri.get();
    g.vt_corner = vt_corner.get();
    g.qual_rule = sc->qual;
    g.quad_rule = sc->quad;
    if (nc > 0) {
        g.n_curved = nc;
        g.ctl_idx = ctl_idx.get();
        g.ctl = ctl.data();
    }
    check_hip(ctx, beep_hip_set_surface(ctx, &g), "geometry");
}

// An instance at the canonical placement of its mesh: no rotation and the
// centre left where it is (beep_scene_add_instance with no pose).
bool canonical_placement(const Instance& in) {
    return in.rot.w == 1.0 && in.rot.x == 0.0 && in.rot.y == 0.0 && in.rot.z == 0.0 &&
           in.offset.x == in.mesh->mesh.centre.x && in.offset.y == in.mesh->mesh.centre.y &&
           in.offset.z == in.mesh->mesh.centre.z;
}

}  // namespace

struct beep_server {
    beep_server_opts o{};
    std::chrono::steady_clock::time_point t0;
    mutable std::mutex mu;
    std::condition_variable cv_job, cv_done, cv_mem;
    std::deque<Job> queue;
    std::deque<long long> done_order;
    std::unordered_map<long long, Done> done;
    std::vector<std::thread> workers;
    std::vector<double> held;  // bytes charged to each slot
    double used = 0;
    // device bytes each slot's context really holds (measured after its last
    // request; 0 when released): with the device's free bytes, what this
    // server's slots could hold together right now
    std::vector<double> actual;
    static constexpr double RESERVE = 1073741824.0;  // left free on the device
    double device_total = 0;
    // the bytes the slots may hold together: the measured free device memory
    // plus what the slots hold, less a reserve, capped by the configured
    // budget (0: no cap); *free_out receives the free bytes read
    double budget_now(double* free_out) const;
    // raw energy coefficients of a mesh at its canonical placement, per mesh
    // stamp (surface and charges): Mesh::init_energy_precalcs runs once per
    // mesh, in the mesh's own frame (mesh.cpp:730-887).  An instance placed
    // otherwise uses them, so its energy does not depend on the requests
    // that ran before it.
    struct EnergyCoeffs {
        std::vector<double> ef, eh;
    };
    std::mutex ecache_mu;
    std::unordered_map<uint64_t, EnergyCoeffs> ecache;
    static constexpr size_t ECACHE_MAX = 64;  // meshes kept
    int waiting_for_memory = 0;
    long long next_id = 1;
    int pending = 0;  // submitted, not yet polled
    bool stop = false;
    std::string err = "no error";

    double clock() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void worker(int slot);
    void solve_one(int slot, beep_hip_ctx*& ctx, Job& j, Done& d);
    void energy_coefficients(beep_hip_ctx* ctx, beep_scene* sc, std::vector<double>& ef,
                             std::vector<double>& eh);
    void release_slot(int slot, beep_hip_ctx*& ctx);  // caller holds mu
};

void beep_server::release_slot(int slot, beep_hip_ctx*& ctx) {
    if (ctx) beep_hip_destroy(ctx);
    ctx = nullptr;
    used -= held[slot];
    held[slot] = 0;
    actual[slot] = 0;
    cv_mem.notify_all();
}

double beep_server::budget_now(double* free_out) const {
    size_t fr = 0, tot = 0;
    if (beep_hip_device_memory(o.device, &fr, &tot) != 0)
        throw std::runtime_error(std::string("device context: ") + beep_hip_last_error(nullptr));
    double ours = 0;
    for (double a : actual) ours += a;
    if (free_out) *free_out = double(fr);
    const double real = double(fr) + ours - RESERVE;
    return o.hbm_budget > 0 ? std::min(o.hbm_budget, real) : real;
}

// Raw energy coefficients of every instance of a built scene whose geometry
// is on the context.  Instances at the canonical placement of their mesh are
// computed from the context; the others take their mesh's coefficients at
// the canonical placement from the cache, computed on a miss from a
// one-instance scene (which replaces the context's geometry).
void beep_server::energy_coefficients(beep_hip_ctx* ctx, beep_scene* sc, std::vector<double>& ef,
                                      std::vector<double>& eh) {
    const int ni = int(sc->inst.size());
    const int nch = int(sc->charge_inst.size());
    bool any_canonical = false;
    for (int k = 0; k < ni; ++k) any_canonical |= canonical_placement(sc->inst[k]);
    if (any_canonical)
        check_hip(ctx, beep_hip_energy_coefficients(ctx, ni, sc->inst_patch_ptr.data(),
                                                    sc->inst_charge_ptr.data(),
                                                    nch ? sc->charges.data() : nullptr, ef.data(),
                                                    eh.data()), "energy coefficients");
    for (int k = 0; k < ni; ++k) {
        const Instance& in = sc->inst[k];
        if (canonical_placement(in)) continue;
        const int p0 = sc->inst_patch_ptr[k], n_k = sc->inst_patch_ptr[k + 1] - p0;
        const uint64_t key = (in.stamp << 10) | (uint64_t(sc->quad) << 5) | uint64_t(sc->qual);
        bool hit = false;
        {
            std::lock_guard<std::mutex> lk(ecache_mu);
            const auto it = ecache.find(key);
            if (it != ecache.end() && it->second.ef.size() == size_t(n_k)) {
                std::copy(it->second.ef.begin(), it->second.ef.end(), ef.begin() + p0);
                std::copy(it->second.eh.begin(), it->second.eh.end(), eh.begin() + p0);
                hit = true;
            }
        }
        if (hit) continue;
        beep_scene* one = nullptr;
        if (beep_scene_create(sc->Dsolvent, sc->kappa, sc->quad, sc->qual, sc->nbsize, &one) != 0)
            throw std::runtime_error(std::string("scene: ") + beep_host_last_error());
        try {
            if (beep_scene_add_instance(one, in.mesh, nullptr, nullptr, in.Dprotein) != 0)
                throw std::runtime_error(std::string("scene: ") + beep_host_last_error());
            one->host_points = false;  // the device generates the point sets
            if (beep_scene_build(one) != 0)
                throw std::runtime_error(std::string("scene build: ") + beep_host_last_error());
            if (int(one->centroid.size()) != n_k)
                throw std::runtime_error("canonical instance differs in size");
            upload_surface(ctx, one);
            EnergyCoeffs c{std::vector<double>(size_t(n_k)), std::vector<double>(size_t(n_k))};
            const int nc1 = int(one->charge_inst.size());
            const int ipp[2] = {0, n_k}, icp[2] = {0, nc1};
            check_hip(ctx, beep_hip_energy_coefficients(ctx, 1, ipp, icp,
                                                        nc1 ? one->charges.data() : nullptr,
                                                        c.ef.data(), c.eh.data()),
                      "energy coefficients");
            std::copy(c.ef.begin(), c.ef.end(), ef.begin() + p0);
            std::copy(c.eh.begin(), c.eh.end(), eh.begin() + p0);
            std::lock_guard<std::mutex> lk(ecache_mu);
            if (ecache.size() >= ECACHE_MAX) ecache.clear();
            ecache.emplace(key, std::move(c));
        } catch (...) {
            beep_scene_destroy(one);
            throw;
        }
        beep_scene_destroy(one);
    }
}

namespace {
void set_tree_from(beep_hip_ctx* ctx, const beep_scene* sc);
}  // namespace

void beep_server::solve_one(int slot, beep_hip_ctx*& ctx, Job& j, Done& d) {
    beep_result& r = d.r;
    beep_scene* sc = j.scene;
    const double th = clock();
    sc->host_points = false;  // the device generates the point sets
    size_t n_items = 0;
    for (const Instance& in : sc->inst) n_items += in.mesh->mesh.verts.size();
    const bool device_tree = o.device_tree == BEEP_SERVE_TREE_AUTO
                                 ? n_items >= size_t(BEEP_SERVE_TREE_DEVICE_MIN)
                                 : o.device_tree != 0;
    sc->with_tree = !device_tree;  // else the device builds the octree
    sc->built = false;
    if (beep_scene_build(sc) != 0)
        throw std::runtime_error(std::string("scene build: ") + beep_host_last_error());
    r.host_ms = (clock() - th) * 1e3;
    const int N = int(sc->centroid.size());
    r.n_patches = N;
    // admission: a free slot (this thread) and room in the memory budget
    // admission against the device's measured free memory (OpenCL_Handler
    // compares a blob's device_mem_reqd with its own count of the device's
    // memory, opencl_handler.cpp:451-463)
    bool fly = o.near_mode == BEEP_NEAR_ON_THE_FLY;
    double need_total = estimate_bytes(sc, o, !fly);
    {
        std::unique_lock<std::mutex> lk(mu);
        double cap = o.hbm_budget > 0 ? o.hbm_budget : 1e300;
        if (device_total > 0) cap = std::min(cap, device_total - RESERVE);
        if (need_total > cap && o.near_mode == BEEP_SERVE_NEAR_AUTO) {
            // the stored coefficients do not fit: integrate them on the fly
            const double need_fly = estimate_bytes(sc, o, false);
            if (need_fly <= cap) {
                fly = true;
                need_total = need_fly;
            }
        }
        if (need_total > cap)
            throw std::runtime_error("request needs more device memory than the budget and the "
                                     "device allow");
        double budget = budget_now(&r.device_free);
        double grow = std::max(0.0, need_total - held[slot]);
        if (used + grow > budget && ctx) {
            // give back what this slot holds before waiting, so that waiting
            // slots never hold memory the others need
            release_slot(slot, ctx);
            grow = need_total;
        }
        ++waiting_for_memory;
        cv_job.notify_all();  // idle slots release their contexts under pressure
        // woken by every release and reconciliation; a timed wait also
        // notices memory freed by other processes
        while (!(used + grow <= budget || used == 0)) {
            cv_mem.wait_for(lk, std::chrono::milliseconds(20));
            budget = budget_now(&r.device_free);
        }
        --waiting_for_memory;
        used += grow;
        held[slot] += grow;
    }
    r.t_admit = clock();
    const double td = clock();
    if (!ctx) {
        if (beep_hip_create(o.device, &ctx) != 0)
            throw std::runtime_error(std::string("device context: ") + beep_hip_last_error(nullptr));
    }
    check_hip(ctx, beep_hip_set_far_field(ctx, o.far_mode, o.fmm_order, o.fmm_leaf), "far field");
    upload_surface(ctx, sc);
    if (device_tree) {
        const double cc[3] = {sc->cube_centre.x, sc->cube_centre.y, sc->cube_centre.z};
        int nl = 0;
        long long nn = 0;
        check_hip(ctx, beep_hip_build_tree(ctx, cc, sc->cube_edge, sc->nbsize, &nl, &nn), "tree");
    } else {
        set_tree_from(ctx, sc);
    }
    const double tp = clock();
    r.upload_ms = (tp - td) * 1e3;
    check_hip(ctx, beep_hip_set_near_mode(ctx, fly ? BEEP_NEAR_ON_THE_FLY : BEEP_NEAR_PRECOMPUTE), "near mode");
    r.near_on_the_fly = fly ? 1 : 0;
    if (o.solve_mode != BEEP_SOLVE_FP64)
        check_hip(ctx, beep_hip_precompute(ctx, sc->kappa, BEEP_FP32), "precompute");
    else
        check_hip(ctx, beep_hip_precompute(ctx, sc->kappa, BEEP_FP64), "precompute");
    check_hip(ctx, beep_hip_synchronize(ctx), "precompute");
    const double ts = clock();
    r.precompute_ms = (ts - tp) * 1e3;
    std::vector<double> x(size_t(2) * N, 0.0);
    const int nch = int(sc->charge_inst.size());
    beep_solve_opts so{o.tol, o.max_iter, o.solve_mode, o.max_refine, o.inner_tol, 0, 0};
    beep_solve_stats st{};
    check_hip(ctx, beep_hip_solve_from_charges(ctx, nch, nch ? sc->charges.data() : nullptr,
                                               sc->Dsolvent, &so, x.data(), &st), "solve");
    // energy: Gauss-law renormalised coefficients (mesh.cpp:891-925), then
    // 1/2 sum (h eps ec_h - f ec_f) per instance (mesh.cpp:1373-1403)
    const int ni = int(sc->inst.size());
    std::vector<double> ef(N), eh(N);
    energy_coefficients(ctx, sc, ef, eh);
    double E = 0;
    for (int k = 0; k < ni; ++k) {
        const int p0 = sc->inst_patch_ptr[k], p1 = sc->inst_patch_ptr[k + 1];
        double net = 0, total_f = 0, area = 0;
        for (int c = sc->inst_charge_ptr[k]; c < sc->inst_charge_ptr[k + 1]; ++c)
            net += sc->charges[4 * size_t(c) + 3];
        for (int i = p0; i < p1; ++i) {
            total_f -= ef[i];
            area += sc->area[i];
        }
        const double e = net - total_f;
        double s = 0;
        for (int i = p0; i < p1; ++i) {
            const double efi = ef[i] - sc->area[i] / area * e;
            s += x[size_t(N) + i] * sc->eps[i] * eh[i] - x[i] * efi;
        }
        E += 0.5 * s * KJ_PER_MOL;
    }
    r.energy = E;
    r.iterations = st.iterations;
    r.matvecs = st.matvecs;
    r.refinements = st.refinements;
    r.converged = st.converged;
    r.rel_residual = st.rel_residual;
    if (o.keep_fh) {
        d.f.assign(x.begin(), x.begin() + N);
        d.h.assign(x.begin() + N, x.end());
    }
    size_t hb = 0;
    check_hip(ctx, beep_hip_memory_held(ctx, &hb), "memory");
    r.bytes = double(hb);
    r.device_ms = (clock() - td) * 1e3;
    r.solve_ms = (clock() - ts) * 1e3;
    {
        // charge the slot what its context really holds
        std::lock_guard<std::mutex> lk(mu);
        used += double(hb) - held[slot];
        held[slot] = double(hb);
        actual[slot] = double(hb);
        cv_mem.notify_all();
    }
}

void beep_server::worker(int slot) {
    beep_hip_ctx* ctx = nullptr;
    for (;;) {
        Job j{};
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_job.wait(lk, [&] {
                if (!queue.empty() || stop) return true;
                // an idle slot gives its memory back when another slot waits
                if (waiting_for_memory > 0 && held[slot] > 0) release_slot(slot, ctx);
                return false;
            });
            if (queue.empty()) break;  // stop requested and nothing left
            j = queue.front();
            queue.pop_front();
        }
        Done d;
        d.r.id = j.id;
        d.r.slot = slot;
        d.r.t_submit = j.t_submit;
        try {
            solve_one(slot, ctx, j, d);
            d.r.status = 0;
            std::snprintf(d.r.message, sizeof(d.r.message), "ok");
        } catch (const std::exception& e) {
            d.r.status = -1;
            std::snprintf(d.r.message, sizeof(d.r.message), "%s", e.what());
            std::lock_guard<std::mutex> lk(mu);
            release_slot(slot, ctx);  // a failed context is not reused
        }
        beep_scene_destroy(j.scene);
        d.r.t_done = clock();
        {
            std::lock_guard<std::mutex> lk(mu);
            done_order.push_back(j.id);
            done.emplace(j.id, std::move(d));
            cv_done.notify_all();
        }
    }
    std::lock_guard<std::mutex> lk(mu);
    release_slot(slot, ctx);
}

extern "C" {

void beep_server_default_opts(beep_server_opts* o) {
    if (!o) return;
    *o = beep_server_opts{};
    o->device = 0;
    o->slots = 4;
    o->hbm_budget = 0.0;  // the device's free memory
    o->solve_mode = BEEP_SOLVE_MIXED;
    o->tol = 1e-6;
    o->max_iter = 100;
    o->max_refine = 4;
    o->inner_tol = 0.0;  // half the outer tolerance
    o->far_mode = BEEP_FAR_AUTO;
    o->fmm_order = 5;
    o->fmm_leaf = 64;
    o->near_mode = BEEP_SERVE_NEAR_AUTO;
    o->device_tree = BEEP_SERVE_TREE_AUTO;
}

int beep_server_estimate(const beep_server_opts* o, beep_scene* sc, double* stored_bytes, double* fly_bytes) {
    try {
        if (!o || !sc || !stored_bytes || !fly_bytes) throw std::runtime_error("null argument");
        if (!sc->built) {
            if (beep_scene_build(sc) != 0)
                throw std::runtime_error(std::string("scene build: ") + beep_host_last_error());
        }
        *stored_bytes = estimate_bytes(sc, *o, true);
        *fly_bytes = estimate_bytes(sc, *o, false);
        return 0;
    } catch (const std::exception& e) {
        beep_host_set_error(e.what());
        return -1;
    }
}

int beep_server_create(const beep_server_opts* opts, beep_server** out) {
    if (!out) return -2;
    *out = nullptr;
    try {
        auto s = std::make_unique<beep_server>();
        if (opts) s->o = *opts;
        else beep_server_default_opts(&s->o);
        if (s->o.slots < 1 || s->o.slots > 64) throw std::runtime_error("slots must be 1..64");
        if (s->o.near_mode != BEEP_SERVE_NEAR_AUTO && s->o.near_mode != BEEP_NEAR_PRECOMPUTE &&
            s->o.near_mode != BEEP_NEAR_ON_THE_FLY)
            throw std::runtime_error("unknown near-field mode");
        if (!(s->o.hbm_budget >= 0)) throw std::runtime_error("budget must be >= 0");
        // the device's size (0 when it cannot be read: requests then fail
        // at admission with the error)
        size_t fr = 0, tot = 0;
        if (beep_hip_device_memory(s->o.device, &fr, &tot) == 0) s->device_total = double(tot);
        s->t0 = std::chrono::steady_clock::now();
        s->held.assign(size_t(s->o.slots), 0.0);
        s->actual.assign(size_t(s->o.slots), 0.0);
        for (int k = 0; k < s->o.slots; ++k) s->workers.emplace_back([p = s.get(), k] { p->worker(k); });
        *out = s.release();
        return 0;
    } catch (const std::exception& e) {
        beep_host_set_error(e.what());
        return -1;
    }
}

int beep_server_submit(beep_server* s, beep_scene* scene, long long* id) {
    if (!s || !scene) return -2;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
        s->err = "server is stopping";
        return -1;
    }
    const long long k = s->next_id++;
    s->queue.push_back(Job{k, scene, s->clock()});
    ++s->pending;
    if (id) *id = k;
    s->cv_job.notify_one();
    return 0;
}

int beep_server_poll(beep_server* s, double timeout_s, beep_result* out) {
    if (!s || !out) return -2;
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return !s->done_order.empty(); };
    if (!ready()) {
        if (timeout_s <= 0) return 1;
        if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return 1;
    }
    const long long k = s->done_order.front();
    s->done_order.pop_front();
    *out = s->done.at(k).r;
    if (!s->o.keep_fh || s->done.at(k).f.empty()) s->done.erase(k);
    --s->pending;
    return 0;
}

int beep_server_take_fh(beep_server* s, long long id, double* f, double* h) {
    if (!s) return -2;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->done.find(id);
    if (it == s->done.end() || it->second.f.empty()) {
        s->err = "no surface values for this request";
        return -1;
    }
    if (f) std::copy(it->second.f.begin(), it->second.f.end(), f);
    if (h) std::copy(it->second.h.begin(), it->second.h.end(), h);
    s->done.erase(it);
    return 0;
}

double beep_server_clock(const beep_server* s) { return s ? s->clock() : 0.0; }

int beep_server_pending(const beep_server* s) {
    if (!s) return -2;
    std::lock_guard<std::mutex> lk(s->mu);
    return s->pending;
}

const char* beep_server_last_error(const beep_server* s) {
    return s ? s->err.c_str() : beep_host_last_error();
}

int beep_server_destroy(beep_server* s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_job.notify_all();
    }
    for (auto& t : s->workers) t.join();
    delete s;  // results not taken are dropped with it
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Pose batches (include/beep_serve.h): one shared near field per mesh, many
// two-instance poses on worker slots.

struct beep_poses {
    beep_pose_opts o{};
    const beep_mesh* mesh = nullptr;
    double Dsolvent = 80, kappa = 0, Dprotein = 2;
    int quad = 0, qual = 0, nbsize = 1500;
    beep_hip_ctx* src = nullptr;     // the mesh's own near field (canonical placement)
    int N1 = 0;                      // patches of one instance
    long long intra_pairs = 0;
    double cutoff = 0, setup_ms = 0;
    std::vector<double> efr, eh;     // energy coefficients of one instance (f: Gauss-renormalised)
    std::vector<beep_hip_ctx*> slot_ctx;
    std::vector<std::vector<double>> fh;  // f then h of every pose of the last solve (keep_fh)
};

namespace {

// Median over (a sample of) the rows of the largest centroid distance to any
// of the row's explicit columns: how far the mesh's own near lists reach.
double median_reach(const beep_scene* sc) {
    std::vector<double> reach;
    const auto& c = sc->centroid;
    const int L = int(sc->leaf_ptr.size()) - 1;
    for (int l = 0; l < L; ++l)
        for (int k = sc->leaf_ptr[l]; k < sc->leaf_ptr[l + 1]; k += 7) {
            const int i = sc->leaf_idx[k];
            double m = 0;
            for (int q = sc->nb_ptr[l]; q < sc->nb_ptr[l + 1]; ++q) {
                const int j = sc->nb_idx[q];
                const double ex = c[j].x - c[i].x, ey = c[j].y - c[i].y, ez = c[j].z - c[i].z;
                m = std::max(m, ex * ex + ey * ey + ez * ez);
            }
            reach.push_back(std::sqrt(m));
        }
    if (reach.empty()) throw std::runtime_error("the mesh has no explicit pairs");
    std::nth_element(reach.begin(), reach.begin() + reach.size() / 2, reach.end());
    return reach[reach.size() / 2];
}

void set_tree_from(beep_hip_ctx* ctx, const beep_scene* sc) {
    beep_hip_tree t{};
    t.n_leaves = int(sc->leaf_ptr.size()) - 1;
    t.leaf_ptr = sc->leaf_ptr.data();
    t.leaf_idx = sc->leaf_idx.data();
    t.nb_ptr = sc->nb_ptr.data();
    t.nb_idx = sc->nb_idx.data();
    check_hip(ctx, beep_hip_set_tree(ctx, &t), "tree");
}

void pose_one(beep_poses* p, beep_hip_ctx*& ctx, int slot, const double* quat, const double* off,
              beep_pose_result& r, std::vector<double>* fh) {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    beep_scene* sc = nullptr;
    if (beep_scene_create(p->Dsolvent, p->kappa, p->quad, p->qual, p->nbsize, &sc) != 0)
        throw std::runtime_error(std::string("scene: ") + beep_host_last_error());
    std::unique_ptr<beep_scene, void (*)(beep_scene*)> guard(sc, beep_scene_destroy);
    const double origin[3] = {0.0, 0.0, 0.0};
    if (beep_scene_add_instance(sc, p->mesh, nullptr, origin, p->Dprotein) != 0 ||
        beep_scene_add_instance(sc, p->mesh, quat, off, p->Dprotein) != 0 ||
        beep_scene_set_host_points(sc, 0) != 0 || beep_scene_set_tree(sc, 0) != 0 ||
        beep_scene_build(sc) != 0)
        throw std::runtime_error(std::string("pose scene: ") + beep_host_last_error());
    const int n1 = p->N1, N = 2 * n1;
    if (int(sc->centroid.size()) != N) throw std::runtime_error("pose scene size");
    const auto t1 = clk::now();
    r.host_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (!ctx && beep_hip_create(p->o.device, &ctx) != 0)
        throw std::runtime_error(std::string("device context: ") + beep_hip_last_error(nullptr));
    check_hip(ctx, beep_hip_set_far_field(ctx, p->o.far_mode, p->o.fmm_order, p->o.fmm_leaf), "far field");
    upload_surface(ctx, sc);
    const int offs[2] = {0, n1};
    check_hip(ctx, beep_hip_share_near_cutoff(ctx, p->src, 2, offs, p->cutoff, &r.n_pairs), "shared near field");
    check_hip(ctx, beep_hip_set_gram_schmidt(ctx, p->o.gram_schmidt), "Gram-Schmidt");
    check_hip(ctx, beep_hip_precompute(ctx, p->kappa, p->o.solve_mode == BEEP_SOLVE_FP64 ? BEEP_FP64 : BEEP_FP32),
              "pair coefficients");
    std::vector<double> x(size_t(2) * N, 0.0);
    const int nch = int(sc->charge_inst.size());
    beep_solve_opts so{p->o.tol, p->o.max_iter, p->o.solve_mode, p->o.max_refine, p->o.inner_tol, 0, 0};
    beep_solve_stats st{};
    check_hip(ctx, beep_hip_solve_from_charges(ctx, nch, nch ? sc->charges.data() : nullptr, p->Dsolvent,
                                               &so, x.data(), &st), "solve");
    // energy of each instance with the mesh's coefficients (pose-invariant):
    // 1/2 sum (h eps ec_h - f ec_f) (mesh.cpp:1373-1403)
    double E = 0;
    const double eps = p->Dsolvent / p->Dprotein;
    for (int k = 0; k < 2; ++k) {
        double s = 0;
        for (int i = 0; i < n1; ++i)
            s += x[size_t(N) + k * n1 + i] * eps * p->eh[i] - x[size_t(k) * n1 + i] * p->efr[i];
        E += 0.5 * s * KJ_PER_MOL;
    }
    r.energy = E;
    r.iterations = st.iterations;
    r.matvecs = st.matvecs;
    r.refinements = st.refinements;
    r.converged = st.converged;
    r.rel_residual = st.rel_residual;
    r.slot = slot;
    r.device_ms = std::chrono::duration<double, std::milli>(clk::now() - t1).count();
    if (fh) *fh = std::move(x);
}

}  // namespace

extern "C" {

void beep_pose_default_opts(beep_pose_opts* o) {
    if (!o) return;
    *o = beep_pose_opts{};
    o->device = 0;
    o->slots = 8;
    o->solve_mode = BEEP_SOLVE_MIXED;
    o->tol = 1e-6;
    o->max_iter = 100;
    o->max_refine = 4;
    o->inner_tol = 0.0;
    o->far_mode = BEEP_FAR_FMM;
    o->fmm_order = 5;
    o->fmm_leaf = 64;
    o->pair_cutoff = 0.0;
    o->gram_schmidt = 1;
    o->keep_fh = 0;
}

int beep_poses_create(const beep_pose_opts* opts, const beep_mesh* mesh, double Dsolvent, double kappa,
                      int quad_rule, int qual_rule, int nbsize, double Dprotein, beep_poses** out) {
    if (!out) return -2;
    *out = nullptr;
    auto p = std::make_unique<beep_poses>();
    beep_scene* sc = nullptr;
    try {
        const auto t0 = std::chrono::steady_clock::now();
        if (!mesh) throw std::runtime_error("null mesh");
        if (opts) p->o = *opts;
        else beep_pose_default_opts(&p->o);
        if (p->o.slots < 1 || p->o.slots > 64) throw std::runtime_error("slots must be 1..64");
        if (p->o.gram_schmidt != 1 && p->o.gram_schmidt != 2) throw std::runtime_error("gram_schmidt must be 1 or 2");
        p->mesh = mesh;
        p->Dsolvent = Dsolvent;
        p->kappa = kappa;
        p->quad = quad_rule;
        p->qual = qual_rule;
        p->nbsize = nbsize;
        p->Dprotein = Dprotein;
        // the mesh at its canonical placement, with its own octree
        if (beep_scene_create(Dsolvent, kappa, quad_rule, qual_rule, nbsize, &sc) != 0 ||
            beep_scene_add_instance(sc, mesh, nullptr, nullptr, Dprotein) != 0 ||
            beep_scene_set_host_points(sc, 0) != 0 || beep_scene_build(sc) != 0)
            throw std::runtime_error(std::string("mesh scene: ") + beep_host_last_error());
        p->N1 = int(sc->centroid.size());
        for (size_t l = 0; l + 1 < sc->leaf_ptr.size(); ++l)
            p->intra_pairs += (long long)(sc->leaf_ptr[l + 1] - sc->leaf_ptr[l]) * (sc->nb_ptr[l + 1] - sc->nb_ptr[l]);
        p->cutoff = p->o.pair_cutoff > 0 ? p->o.pair_cutoff : median_reach(sc);
        if (beep_hip_create(p->o.device, &p->src) != 0)
            throw std::runtime_error(std::string("device context: ") + beep_hip_last_error(nullptr));
        // direct far field: every explicit pair stores full - 1 point, which
        // holds whatever far field a pose then uses
        check_hip(p->src, beep_hip_set_far_field(p->src, BEEP_FAR_DIRECT, 5, 40), "far field");
        upload_surface(p->src, sc);
        set_tree_from(p->src, sc);
        check_hip(p->src, beep_hip_precompute(p->src, kappa,
                                              p->o.solve_mode == BEEP_SOLVE_FP64 ? BEEP_FP64 : BEEP_FP32),
                  "mesh coefficients");
        // energy coefficients of one instance, Gauss-renormalised
        // (mesh.cpp:891-925)
        const int n1 = p->N1, nch = int(sc->charge_inst.size());
        std::vector<double> ef(n1);
        p->eh.assign(n1, 0.0);
        const int ipp[2] = {0, n1}, icp[2] = {0, nch};
        check_hip(p->src, beep_hip_energy_coefficients(p->src, 1, ipp, icp, nch ? sc->charges.data() : nullptr,
                                                       ef.data(), p->eh.data()),
                  "energy coefficients");
        double net = 0, total_f = 0, area = 0;
        for (int c = 0; c < nch; ++c) net += sc->charges[4 * size_t(c) + 3];
        for (int i = 0; i < n1; ++i) {
            total_f -= ef[i];
            area += sc->area[i];
        }
        const double e = net - total_f;
        p->efr.resize(n1);
        for (int i = 0; i < n1; ++i) p->efr[i] = ef[i] - sc->area[i] / area * e;
        check_hip(p->src, beep_hip_synchronize(p->src), "set-up");
        p->slot_ctx.assign(size_t(p->o.slots), nullptr);
        p->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        beep_scene_destroy(sc);
        *out = p.release();
        return 0;
    } catch (const std::exception& e) {
        if (sc) beep_scene_destroy(sc);
        if (p->src) beep_hip_destroy(p->src);
        beep_host_set_error(e.what());
        return -1;
    }
}

int beep_poses_solve(beep_poses* p, int n, const double* quat4, const double* offset3,
                     beep_pose_result* results) {
    if (!p || n < 0 || (n > 0 && (!quat4 || !offset3 || !results))) return -2;
    p->fh.assign(p->o.keep_fh ? size_t(n) : 0, std::vector<double>());
    std::atomic<int> next{0};
    auto work = [&](int slot) {
        for (;;) {
            const int k = next.fetch_add(1);
            if (k >= n) break;
            beep_pose_result& r = results[k];
            r = beep_pose_result{};
            try {
                pose_one(p, p->slot_ctx[slot], slot, quat4 + 4 * size_t(k), offset3 + 3 * size_t(k), r,
                         p->o.keep_fh ? &p->fh[k] : nullptr);
                r.status = 0;
                std::snprintf(r.message, sizeof(r.message), "ok");
            } catch (const std::exception& e) {
                r.status = -1;
                r.slot = slot;
                std::snprintf(r.message, sizeof(r.message), "%s", e.what());
                if (p->slot_ctx[slot]) beep_hip_destroy(p->slot_ctx[slot]);  // not reused
                p->slot_ctx[slot] = nullptr;
            }
        }
    };
    std::vector<std::thread> th;
    for (int s = 0; s < p->o.slots; ++s) th.emplace_back(work, s);
    for (auto& t : th) t.join();
    return 0;
}

int beep_poses_take_fh(beep_poses* p, int k, double* f, double* h) {
    if (!p || k < 0 || size_t(k) >= p->fh.size() || p->fh[k].empty()) {
        beep_host_set_error("no f and h kept for that pose");
        return -1;
    }
    const size_t N = p->fh[k].size() / 2;
    if (f) std::copy(p->fh[k].begin(), p->fh[k].begin() + N, f);
    if (h) std::copy(p->fh[k].begin() + N, p->fh[k].end(), h);
    return 0;
}

int beep_poses_info(const beep_poses* p, int* n_patches, long long* intra_pairs, double* cutoff,
                    double* setup_ms) {
    if (!p) return -2;
    if (n_patches) *n_patches = p->N1;
    if (intra_pairs) *intra_pairs = p->intra_pairs;
    if (cutoff) *cutoff = p->cutoff;
    if (setup_ms) *setup_ms = p->setup_ms;
    return 0;
}

void beep_poses_destroy(beep_poses* p) {
    if (!p) return;
    for (auto* c : p->slot_ctx)
        if (c) beep_hip_destroy(c);
    if (p->src) beep_hip_destroy(p->src);
    delete p;
}

}  // extern "C"
