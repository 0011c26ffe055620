// fm_batch.hip -- many independent domains (replicas of uneven size) advanced together.
//
// Each step is two packed launches over every member: VV pass 1 with the displacement check
// (one block per 256 molecules of a member), and the forces with the fused VV pass 2 (one block
// per brick of a member).  Per-member arguments live in device descriptor arrays, written per
// step from pinned host staging; block -> (member, offset) tables change only when the
// membership or a member's brick count changes.
//
// Steps are queued in windows of up to BATCH_WINDOW steps with one host synchronisation per
// window.  A member's steps in a window end before its next rebuild that the host knows in
// advance (the record step's forced rebuild, or the fixed interval); such a rebuild opens the
// member's next window, queued without host waits.  A displacement check that fires inside the
// window records its step in the member's flag words, and the member's later launches of the
// window do nothing (the packed kernels read each member's flags); at the window's
// synchronisation the host reads every member's flags, and the member's next window starts
// with that step's rebuild and force.  A queued list build beyond the capacities halts the
// member the same way and is repeated, with grown capacities, at the synchronisation.  Each
// completed step has the arithmetic of the member's own step-by-step loop.  A window's
// descriptors go to the device in three copies: pass 1 of its first step, then (after the
// rebuilds that open members' windows, which change their state pointers) everything else.
// Optionally the launches after the rebuilds are replayed as one HIP graph per window length
// (kernel nodes only), re-instantiated when a table, the launch sizes or the LDS size change.
#include "fm_internal.h"

#include <string.h>

#include <algorithm>
#include <vector>

#define BATCH_WINDOW 16

struct Member {
    fm_ctx* c;
    int iter;                 // iteration number of the member's next step
    hipStream_t own_stream;   // restored when the member leaves; the member's rebuilds run on it
    int fired;                // step whose displacement check fired (its pass 1 has run), or -1
    hipEvent_t done = nullptr;  // the member's rebuild, for the batch stream to wait on
};

struct WindowGraph {          // the launches of a window of W steps after its rebuilds
    hipGraphExec_t g = nullptr;
    int nblk1 = -1, nblk2 = -1;
    size_t lds = 0;
    uint32_t plain = 0;       // bit j: step j's force launch has no record or virial step
};

struct fm_batch {
    std::vector<Member> m;
    hipStream_t stream = nullptr;
    int kind = -1;
    // device: BATCH_WINDOW slots of cap_m descriptors each
    P1Desc* d_p1 = nullptr;
    char* d_fargs = nullptr;
    int2* d_t1 = nullptr;
    int2* d_t2 = nullptr;
    size_t cap_m = 0, cap_t1 = 0, cap_t2 = 0;
    // pinned host staging, the same slots
    P1Desc* h_p1 = nullptr;
    char* h_fargs = nullptr;
    std::vector<int2> t1, t2;
    bool dirty = true;
    int nblk1 = 0, nblk2 = 0;
    // graphs
    bool use_graph = true;
    WindowGraph wg[BATCH_WINDOW + 1];
    uint32_t plain = 0;       // the current window's plain force launches (bit j)
    int64_t rebuilds = 0;
    int64_t windows = 0;
    hipEvent_t p1done = nullptr;   // pass 1 of a window's first step, for the members' rebuilds
};


static void free_graphs(fm_batch* b) {
    for (WindowGraph& w : b->wg) {
        if (w.g) hipGraphExecDestroy(w.g);
        w = WindowGraph();
    }
}

static void free_buffers(fm_batch* b) {
    free_graphs(b);
    void* d[] = {b->d_p1, b->d_fargs, b->d_t1, b->d_t2};
    for (void* p : d)
        if (p) hipFree(p);
    void* h[] = {b->h_p1, b->h_fargs};
    for (void* p : h)
        if (p) hipHostFree(p);
    b->d_p1 = nullptr;
    b->d_fargs = nullptr;
    b->d_t1 = b->d_t2 = nullptr;
    b->h_p1 = nullptr;
    b->h_fargs = nullptr;
    b->cap_m = b->cap_t1 = b->cap_t2 = 0;
}

static int ensure_member_buffers(fm_batch* b, size_t n) {
    if (n <= b->cap_m) return 0;
    size_t cap = std::max<size_t>(n, 2 * b->cap_m + 8);
    free_graphs(b);
    if (b->d_p1) hipFree(b->d_p1);
    if (b->d_fargs) hipFree(b->d_fargs);
    if (b->h_p1) hipHostFree(b->h_p1);
    if (b->h_fargs) hipHostFree(b->h_fargs);
    b->d_p1 = nullptr;
    b->d_fargs = nullptr;
    b->h_p1 = nullptr;
    b->h_fargs = nullptr;
    b->cap_m = 0;
    const size_t W = BATCH_WINDOW;
    FM_HIP(hipMalloc((void**)&b->d_p1, W * cap * sizeof(P1Desc)));
    FM_HIP(hipMalloc((void**)&b->d_fargs, W * cap * fargs_bytes()));
    FM_HIP(hipHostMalloc((void**)&b->h_p1, W * cap * sizeof(P1Desc), hipHostMallocDefault));
    FM_HIP(hipHostMalloc((void**)&b->h_fargs, W * cap * fargs_bytes(), hipHostMallocDefault));
    b->cap_m = cap;
    return 0;
}

static int upload_tables(fm_batch* b) {
    b->t1.clear();
    b->t2.clear();
    for (size_t k = 0; k < b->m.size(); ++k) {
        const fm_ctx* c = b->m[k].c;
        const int nb1 = (c->np + FM_BLOCK - 1) / FM_BLOCK;
        for (int q = 0; q < nb1; ++q) b->t1.push_back(make_int2((int)k, q));
        const int nb2 = brick_count(c);
        for (int q = 0; q < nb2; ++q) b->t2.push_back(make_int2((int)k, q));
    }
    b->nblk1 = (int)b->t1.size();
    b->nblk2 = (int)b->t2.size();
    if (b->t1.size() > b->cap_t1) {
        free_graphs(b);
        if (b->d_t1) hipFree(b->d_t1);
        b->cap_t1 = b->t1.size() + b->t1.size() / 2 + 64;
        FM_HIP(hipMalloc((void**)&b->d_t1, b->cap_t1 * sizeof(int2)));
    }
    if (b->t2.size() > b->cap_t2) {
        free_graphs(b);
        if (b->d_t2) hipFree(b->d_t2);
        b->cap_t2 = b->t2.size() + b->t2.size() / 2 + 64;
        FM_HIP(hipMalloc((void**)&b->d_t2, b->cap_t2 * sizeof(int2)));
    }
    if (!b->t1.empty())
        FM_HIP(hipMemcpyAsync(b->d_t1, b->t1.data(), b->t1.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream));
    if (!b->t2.empty())
        FM_HIP(hipMemcpyAsync(b->d_t2, b->t2.data(), b->t2.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream));
    FM_HIP(hipStreamSynchronize(b->stream));
    b->dirty = false;
    return 0;
}

static P1Desc* slot_p1(fm_batch* b, P1Desc* base, int j) { return base + (size_t)j * b->cap_m; }
static char* slot_fargs(fm_batch* b, char* base, int j) { return base + (size_t)j * b->cap_m * fargs_bytes(); }

// the launches of window steps j0.. of a W-step window: pass 1 (except at step 0, launched
// before the rebuilds) and the forces of each step, from the descriptors already on the device
static int enqueue_window(fm_batch* b, int W, size_t lds) {
    for (int j = 0; j < W; ++j) {
        if (j > 0) TRY(launch_pass1_multi(b->stream, slot_p1(b, b->d_p1, j), b->d_t1, b->nblk1));
        TRY(launch_forces_multi(b->stream, b->kind, slot_fargs(b, b->d_fargs, j), b->d_t2, b->nblk2, lds,
                                (b->plain >> j) & 1u));
    }
    return 0;
}

static int launch_window(fm_batch* b, int W, size_t lds) {
    if (!b->use_graph) return enqueue_window(b, W, lds);
    WindowGraph& w = b->wg[W];
    // a graph's LDS size is fixed: keep it until a member needs more, with headroom
    if (!w.g || w.nblk1 != b->nblk1 || w.nblk2 != b->nblk2 || lds > w.lds || w.plain != b->plain) {
        if (w.g) hipGraphExecDestroy(w.g);
        w.g = nullptr;
        size_t glds = std::max(lds, lds + lds / 8);
        size_t limit = b->m[0].c->lds_limit;
        if (glds > limit) glds = std::max(lds, limit);
        hipGraph_t g = nullptr;
        FM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_window(b, W, glds);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fm_set_error(FM_EHIP, "graph capture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&w.g, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) return fm_set_error(FM_EHIP, "graph instantiate: %s", hipGetErrorString(e));
        w.nblk1 = b->nblk1;
        w.nblk2 = b->nblk2;
        w.lds = glds;
        w.plain = b->plain;
    }
    FM_HIP(hipGraphLaunch(w.g, b->stream));
    return 0;
}

extern "C" int fm_batch_create(fm_batch** out) {
    if (!out) return fm_set_error(FM_EINVAL, "null argument");
    fm_batch* b = new fm_batch();
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->p1done, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (b->stream) hipStreamDestroy(b->stream);
        delete b;
        return fm_set_error(FM_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = b;
    return 0;
}

extern "C" void fm_batch_destroy(fm_batch* b) {
    if (!b) return;
    hipStreamSynchronize(b->stream);
    for (auto& mm : b->m) {
        mm.c->stream = mm.own_stream;
        if (mm.done) hipEventDestroy(mm.done);
    }
    free_buffers(b);
    if (b->p1done) hipEventDestroy(b->p1done);
    hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int fm_batch_use_graphs(fm_batch* b, int32_t enable) {
    if (!b) return fm_set_error(FM_EINVAL, "null batch");
    b->use_graph = enable != 0;
    free_graphs(b);
    return 0;
}

extern "C" int fm_batch_add(fm_batch* b, fm_ctx* c, int32_t iter0) {
    if (!b || !c) return fm_set_error(FM_EINVAL, "null argument");
    for (auto& mm : b->m)
        if (mm.c == c) return fm_set_error(FM_EINVAL, "context already in the batch");
    if (!c->nlist_valid) return fm_set_error(FM_ESTATE, "a batch member needs its neighbour list and initial forces");
    if (c->p.rebuild_criteria == FM_REBUILD_VMAX)
        return fm_set_error(FM_EINVAL, "batch members use the displacement or fixed rebuild criterion");
    if (c->p.integration_algorithm != FM_INTEGRATION_VV)
        return fm_set_error(FM_EINVAL, "batch members use velocity Verlet");
    char tmp[1024];
    size_t lds;
    int kind;
    if (fargs_bytes() > sizeof tmp) return fm_set_error(FM_EINVAL, "descriptor too large");
    if (int rc = make_fargs(c, iter0, true, tmp, &lds, &kind)) return rc;
    if (b->kind >= 0 && kind != b->kind)
        return fm_set_error(FM_EINVAL, "batch members must share precision, pair law and bond/tether kind");
    if (int rc = ensure_member_buffers(b, b->m.size() + 1)) return rc;
    // no rebuild pending from the member's own loops (the batch's halt rule reads these words)
    static const int reset[] = {FLAG_REBUILD, FLAG_HALT, FLAG_NEED};
    if (int rc = launch_reset_flags(c, reset, 3)) return rc;
    FM_HIP(hipStreamSynchronize(c->stream));
    b->kind = kind;
    hipEvent_t ev = nullptr;
    FM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    b->m.push_back({c, iter0, c->stream, -1, ev});
    c->stream = b->stream;
    b->dirty = true;
    return 0;
}

extern "C" int fm_batch_remove(fm_batch* b, fm_ctx* c) {
    if (!b || !c) return fm_set_error(FM_EINVAL, "null argument");
    for (size_t k = 0; k < b->m.size(); ++k)
        if (b->m[k].c == c) {
            FM_HIP(hipStreamSynchronize(b->stream));
            c->stream = b->m[k].own_stream;
            if (b->m[k].done) hipEventDestroy(b->m[k].done);
            b->m.erase(b->m.begin() + k);
            b->dirty = true;
            if (b->m.empty()) b->kind = -1;
            return 0;
        }
    return fm_set_error(FM_EINVAL, "context not in the batch");
}

extern "C" int fm_batch_size(const fm_batch* b, int32_t* n) {
    if (!b || !n) return fm_set_error(FM_EINVAL, "null argument");
    *n = (int32_t)b->m.size();
    return 0;
}

extern "C" int fm_batch_counters(const fm_batch* b, int64_t* rebuilds, int64_t* host_syncs) {
    if (!b) return fm_set_error(FM_EINVAL, "null batch");
    if (rebuilds) *rebuilds = b->rebuilds;
    if (host_syncs) *host_syncs = b->windows;
    return 0;
}

extern "C" int fm_batch_get_iter(const fm_batch* b, const fm_ctx* c, int32_t* iter) {
    if (!b || !c || !iter) return fm_set_error(FM_EINVAL, "null argument");
    for (auto& mm : b->m)
        if (mm.c == c) {
            *iter = mm.iter;
            return 0;
        }
    return fm_set_error(FM_EINVAL, "context not in the batch");
}

// Does member c rebuild at step it whatever its displacement check finds (the record step's
// forced rebuild, or the fixed interval)?
static bool host_rebuild(const fm_ctx* c, int it) {
    return ((it + 1) % c->p.tplot) == 0 || (c->p.rebuild_criteria == FM_REBUILD_FIXED && it % c->p.fixed_rebuild == 0);
}

// md_advance_vv (src/md.f90:143-163) nsteps times for every member, in windows of queued steps
extern "C" int fm_batch_run(fm_batch* b, int32_t nsteps, int32_t* n_rebuilds) {
    if (!b) return fm_set_error(FM_EINVAL, "null batch");
    const size_t n = b->m.size();
    int nrb = 0;
    if (n == 0 || nsteps <= 0) {
        if (n_rebuilds) *n_rebuilds = 0;
        return 0;
    }
    if (b->dirty) TRY(upload_tables(b));
    const size_t fsz = fargs_bytes();
    std::vector<int> rem(n, nsteps), len(n), it0(n);
    std::vector<char> rb0(n);
    for (;;) {
        // ---- the window: each member's steps up to its next host-known rebuild
        int W = 0;
        for (size_t k = 0; k < n; ++k) {
            Member& mm = b->m[k];
            it0[k] = mm.iter;
            len[k] = 0;
            rb0[k] = 0;
            if (rem[k] <= 0) continue;
            rb0[k] = mm.fired == mm.iter || host_rebuild(mm.c, mm.iter);
            int l = 1;
            while (l < BATCH_WINDOW && l < rem[k] && !host_rebuild(mm.c, mm.iter + l)) ++l;
            len[k] = l;
            W = std::max(W, l);
        }
        if (W == 0) break;
        const size_t n_ = n;
        size_t lds = 0;
        b->plain = 0;
        for (int j = 0; j < W; ++j) {
            // ---- VV pass 1 and the displacement check of every member's step j
            for (size_t k = 0; k < n_; ++k) {
                Member& mm = b->m[k];
                fm_ctx* c = mm.c;
                const int it = it0[k] + j;
                P1Desc& d = slot_p1(b, b->h_p1, j)[k];
                d.active = j < len[k] && !(j == 0 && mm.fired == it0[k]);
                d.st = c->x();
                d.rd = c->rdisp;
                d.flags = c->d_flags;
                d.np = c->np;
                d.cap = c->cap;
                d.step = it;
                d.check = (c->p.rebuild_criteria == FM_REBUILD_DISPLACEMENT && !host_rebuild(c, it)) ? 1 : 0;
                d.fresh = c->rdisp_valid ? 0 : 1;
                d.dt = c->p.delta_t;
                d.hdt2 = 0.5 * (c->p.delta_t * c->p.delta_t);
                d.hdt = 0.5 * c->p.delta_t;
                d.thresh = 0.5 * c->dmin;
                if (d.active) {
                    c->moved_since_bin = true;
                    if (d.check) c->rdisp_valid = true;
                }
            }
            if (j == 0) {
                // pass 1 of the first step, then the rebuilds that open members' windows, queued
                // without host waits (they move members' state to their other buffer)
                FM_HIP(hipMemcpyAsync(b->d_p1, b->h_p1, n_ * sizeof(P1Desc), hipMemcpyHostToDevice, b->stream));
                TRY(launch_pass1_multi(b->stream, b->d_p1, b->d_t1, b->nblk1));
                // the rebuilds run side by side, each on its member's own stream after the pass 1
                // launch; the batch stream waits for all of them before the forces
                bool any = false;
                for (size_t k = 0; k < n_; ++k) any = any || (len[k] > 0 && rb0[k]);
                if (any) FM_HIP(hipEventRecord(b->p1done, b->stream));
                for (size_t k = 0; k < n_; ++k) {
                    if (len[k] == 0 || !rb0[k]) continue;
                    Member& mm = b->m[k];
                    fm_ctx* c = mm.c;
                    c->rdisp_valid = false;
                    const int old_bricks = brick_count(c);
                    FM_HIP(hipStreamWaitEvent(mm.own_stream, b->p1done, 0));
                    c->stream = mm.own_stream;
                    const int rc = rebuild_lists_queued_internal(c);
                    c->stream = b->stream;
                    TRY(rc);
                    FM_HIP(hipEventRecord(mm.done, mm.own_stream));
                    FM_HIP(hipStreamWaitEvent(b->stream, mm.done, 0));
                    if (brick_count(c) != old_bricks) b->dirty = true;
                    ++nrb;
                }
                if (b->dirty) TRY(upload_tables(b));
            }
            // ---- forces with the fused VV pass 2 of every member's step j
            bool plain = true;
            for (size_t k = 0; k < n_; ++k) {
                size_t l = 0;
                int kind = 0;
                char* fa = slot_fargs(b, b->h_fargs, j) + k * fsz;
                TRY(make_fargs(b->m[k].c, it0[k] + j, true, fa, &l, &kind));
                fargs_set_queued(fa, b->m[k].c->d_flags, j < len[k]);
                if (j < len[k] && !fargs_plain(fa)) plain = false;
                lds = std::max(lds, l);
            }
            if (plain) b->plain |= 1u << j;
        }
        // the window's other descriptors in two copies, then its launches
        if (W > 1)
            FM_HIP(hipMemcpyAsync(slot_p1(b, b->d_p1, 1), slot_p1(b, b->h_p1, 1), (size_t)(W - 1) * b->cap_m * sizeof(P1Desc),
                                  hipMemcpyHostToDevice, b->stream));
        FM_HIP(hipMemcpyAsync(b->d_fargs, b->h_fargs, (size_t)W * b->cap_m * fsz, hipMemcpyHostToDevice, b->stream));
        TRY(launch_window(b, W, lds));
        // ---- one synchronisation: every member's flag words
        for (size_t k = 0; k < n; ++k) {
            fm_ctx* c = b->m[k].c;
            FM_HIP(hipMemcpyAsync(c->h_flags, c->d_flags, FLAG_COUNT * sizeof(int), hipMemcpyDeviceToHost, b->stream));
        }
        FM_HIP(hipStreamSynchronize(b->stream));
        ++b->windows;
        for (size_t k = 0; k < n; ++k) {
            if (len[k] == 0) continue;
            Member& mm = b->m[k];
            fm_ctx* c = mm.c;
            TRY(check_flags_internal(c));
            int done = len[k];
            mm.fired = -1;
            if (c->h_flags[FLAG_NEED]) {
                // the window's list build met a region or list beyond the capacities: every launch
                // of the member after it did nothing; build again, then the step's force
                if (!rb0[k]) return fm_set_error(FM_ESTATE, "capacity request without a queued list build");
                TRY(redo_lists_internal(c));
                TRY(launch_forces(c, it0[k], true));
                done = 1;
            } else if (c->h_flags[FLAG_REBUILD]) {
                const int s = c->h_flags[FLAG_RB_STEP];
                if (s < it0[k] || s >= it0[k] + len[k] || (s == it0[k] && rb0[k]))
                    return fm_set_error(FM_ESTATE, "rebuild requested at step %d outside %d..%d", s, it0[k],
                                        it0[k] + len[k] - 1);
                // steps before s are complete; s has had its pass 1 and opens the next window
                done = s - it0[k];
                mm.fired = s;
            }
            for (int q = 0; q < done; ++q) {
                const int it = it0[k] + q;
                const bool rec = (it % c->p.tplot) == 0;
                const bool ev = c->p.teval > 0 && (it % c->p.teval) == 0;
                c->pe_valid = rec;
                c->vir_valid = rec || ev;
                c->simtime = c->simtime + c->p.delta_t;
            }
            mm.iter = it0[k] + done;
            rem[k] -= done;
        }
    }
    b->rebuilds += nrb;
    if (n_rebuilds) *n_rebuilds = nrb;
    return 0;
}
