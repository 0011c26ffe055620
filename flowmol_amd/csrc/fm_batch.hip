// fm_batch.hip -- many independent domains (replicas of uneven size) advanced together.
//
// Each step is two packed launches over every member: VV pass 1 with the displacement check
// (one block per 256 molecules of a member), and the forces with the fused VV pass 2 (one block
// per brick of a member).  Per-member arguments live in device descriptor arrays, rebuilt per
// step from pinned host staging; block -> (member, offset) tables change only when the
// membership or a member's brick count changes.  One host synchronisation per step reads all
// members' rebuild flags; only flagged members rebuild their lists.  Both packed launches are
// captured as HIP graphs (copy nodes and kernel nodes only) and re-instantiated when a table,
// the launch size or the LDS size changes.
#include "fm_internal.h"

#include <string.h>

#include <algorithm>
#include <vector>

struct Member {
    fm_ctx* c;
    int iter;                 // iteration number of the member's next step
    hipStream_t own_stream;   // restored when the member leaves
};

struct fm_batch {
    std::vector<Member> m;
    hipStream_t stream = nullptr;
    int kind = -1;
    // device
    P1Desc* d_p1 = nullptr;
    int* d_bflags = nullptr;
    char* d_fargs = nullptr;
    int2* d_t1 = nullptr;
    int2* d_t2 = nullptr;
    size_t cap_m = 0, cap_t1 = 0, cap_t2 = 0;
    // pinned host staging
    P1Desc* h_p1 = nullptr;
    int* h_bflags = nullptr;
    char* h_fargs = nullptr;
    std::vector<int2> t1, t2;
    bool dirty = true;
    int nblk1 = 0, nblk2 = 0;
    // graphs
    bool use_graph = true;
    hipGraphExec_t g1 = nullptr, g2 = nullptr;
    int g1_nblk = -1, g2_nblk = -1, g2_kind = -1;
    size_t g1_n = 0, g2_n = 0;          // member counts baked into the graphs' copy nodes
    size_t g2_lds = 0;
    int64_t rebuilds = 0;
};


static void free_graphs(fm_batch* b) {
    if (b->g1) hipGraphExecDestroy(b->g1);
    if (b->g2) hipGraphExecDestroy(b->g2);
    b->g1 = b->g2 = nullptr;
    b->g1_nblk = b->g2_nblk = -1;
}

static void free_buffers(fm_batch* b) {
    free_graphs(b);
    void* d[] = {b->d_p1, b->d_bflags, b->d_fargs, b->d_t1, b->d_t2};
    for (void* p : d)
        if (p) hipFree(p);
    void* h[] = {b->h_p1, b->h_bflags, b->h_fargs};
    for (void* p : h)
        if (p) hipHostFree(p);
    b->d_p1 = nullptr;
    b->d_bflags = nullptr;
    b->d_fargs = nullptr;
    b->d_t1 = b->d_t2 = nullptr;
    b->h_p1 = nullptr;
    b->h_bflags = nullptr;
    b->h_fargs = nullptr;
    b->cap_m = b->cap_t1 = b->cap_t2 = 0;
}

static int ensure_member_buffers(fm_batch* b, size_t n) {
    if (n <= b->cap_m) return 0;
    size_t cap = std::max<size_t>(n, 2 * b->cap_m + 8);
    free_graphs(b);
    if (b->d_p1) hipFree(b->d_p1);
    if (b->d_bflags) hipFree(b->d_bflags);
    if (b->d_fargs) hipFree(b->d_fargs);
    if (b->h_p1) hipHostFree(b->h_p1);
    if (b->h_bflags) hipHostFree(b->h_bflags);
    if (b->h_fargs) hipHostFree(b->h_fargs);
    FM_HIP(hipMalloc((void**)&b->d_p1, cap * sizeof(P1Desc)));
    FM_HIP(hipMalloc((void**)&b->d_bflags, cap * sizeof(int)));
    FM_HIP(hipMalloc((void**)&b->d_fargs, cap * fargs_bytes()));
    FM_HIP(hipHostMalloc((void**)&b->h_p1, cap * sizeof(P1Desc), hipHostMallocDefault));
    FM_HIP(hipHostMalloc((void**)&b->h_bflags, cap * sizeof(int), hipHostMallocDefault));
    FM_HIP(hipHostMalloc((void**)&b->h_fargs, cap * fargs_bytes(), hipHostMallocDefault));
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

// the work of one packed launch, enqueued on the batch stream (also used under capture)
static int enqueue_pass1(fm_batch* b) {
    const size_t n = b->m.size();
    FM_HIP(hipMemcpyAsync(b->d_p1, b->h_p1, n * sizeof(P1Desc), hipMemcpyHostToDevice, b->stream));
    FM_HIP(hipMemcpyAsync(b->d_bflags, b->h_bflags, n * sizeof(int), hipMemcpyHostToDevice, b->stream));
    if (int rc = launch_pass1_multi(b->stream, b->d_p1, b->d_t1, b->nblk1)) return rc;
    FM_HIP(hipMemcpyAsync(b->h_bflags, b->d_bflags, n * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    return 0;
}

static int enqueue_forces(fm_batch* b, size_t lds) {
    const size_t n = b->m.size();
    FM_HIP(hipMemcpyAsync(b->d_fargs, b->h_fargs, n * fargs_bytes(), hipMemcpyHostToDevice, b->stream));
    return launch_forces_multi(b->stream, b->kind, b->d_fargs, b->d_t2, b->nblk2, lds);
}

static int capture(fm_batch* b, hipGraphExec_t* exec, int which, size_t lds) {
    hipGraph_t g = nullptr;
    FM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = which == 1 ? enqueue_pass1(b) : enqueue_forces(b, lds);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fm_set_error(FM_EHIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return fm_set_error(FM_EHIP, "graph instantiate: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int fm_batch_create(fm_batch** out) {
    if (!out) return fm_set_error(FM_EINVAL, "null argument");
    fm_batch* b = new fm_batch();
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return fm_set_error(FM_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = b;
    return 0;
}

extern "C" void fm_batch_destroy(fm_batch* b) {
    if (!b) return;
    hipStreamSynchronize(b->stream);
    for (auto& mm : b->m) mm.c->stream = mm.own_stream;
    free_buffers(b);
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
    FM_HIP(hipStreamSynchronize(c->stream));
    b->kind = kind;
    b->m.push_back({c, iter0, c->stream});
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

extern "C" int fm_batch_get_iter(const fm_batch* b, const fm_ctx* c, int32_t* iter) {
    if (!b || !c || !iter) return fm_set_error(FM_EINVAL, "null argument");
    for (auto& mm : b->m)
        if (mm.c == c) {
            *iter = mm.iter;
            return 0;
        }
    return fm_set_error(FM_EINVAL, "context not in the batch");
}

// md_advance_vv (src/md.f90:143-163) nsteps times for every member
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
    for (int step = 0; step < nsteps; ++step) {
        // ---- VV pass 1 and the displacement check, all members in one launch
        for (size_t k = 0; k < n; ++k) {
            Member& mm = b->m[k];
            fm_ctx* c = mm.c;
            const bool forced = ((mm.iter + 1) % c->p.tplot) == 0;
            P1Desc& d = b->h_p1[k];
            d.st = c->x();
            d.rd = c->rdisp;
            d.flag_out = b->d_bflags + k;
            d.np = c->np;
            d.cap = c->cap;
            d.check = (c->p.rebuild_criteria == FM_REBUILD_DISPLACEMENT && !forced) ? 1 : 0;
            d.fresh = c->rdisp_valid ? 0 : 1;
            d.dt = c->p.delta_t;
            d.hdt2 = 0.5 * (c->p.delta_t * c->p.delta_t);
            d.hdt = 0.5 * c->p.delta_t;
            d.thresh = 0.5 * c->dmin;
            b->h_bflags[k] = 0;
            c->moved_since_bin = true;
        }
        if (b->use_graph) {
            if (!b->g1 || b->g1_nblk != b->nblk1 || b->g1_n != n) {
                if (b->g1) hipGraphExecDestroy(b->g1);
                b->g1 = nullptr;
                TRY(capture(b, &b->g1, 1, 0));
                b->g1_nblk = b->nblk1;
                b->g1_n = n;
            }
            FM_HIP(hipGraphLaunch(b->g1, b->stream));
        } else {
            TRY(enqueue_pass1(b));
        }
        FM_HIP(hipStreamSynchronize(b->stream));
        // ---- per-member rebuild decisions and list rebuilds
        for (size_t k = 0; k < n; ++k) {
            Member& mm = b->m[k];
            fm_ctx* c = mm.c;
            const P1Desc& d = b->h_p1[k];
            if (d.check) c->rdisp_valid = true;
            int rb = ((mm.iter + 1) % c->p.tplot) == 0;
            if (d.check && b->h_bflags[k]) rb = 1;
            if (c->p.rebuild_criteria == FM_REBUILD_FIXED && mm.iter % c->p.fixed_rebuild == 0) rb = 1;
            if (rb) {
                c->rdisp_valid = false;
                const int old_bricks = brick_count(c);
                TRY(rebuild_lists_internal(c));
                if (brick_count(c) != old_bricks) b->dirty = true;
                ++nrb;
            }
        }
        if (b->dirty) TRY(upload_tables(b));
        // ---- forces with the fused VV pass 2, all members in one launch
        size_t lds = 0;
        for (size_t k = 0; k < n; ++k) {
            size_t l = 0;
            int kind = 0;
            TRY(make_fargs(b->m[k].c, b->m[k].iter, true, b->h_fargs + k * fsz, &l, &kind));
            lds = std::max(lds, l);
        }
        if (b->use_graph) {
            // a graph's LDS size is fixed: keep it until a member needs more, with headroom
            if (!b->g2 || b->g2_nblk != b->nblk2 || b->g2_n != n || lds > b->g2_lds) {
                if (b->g2) hipGraphExecDestroy(b->g2);
                b->g2 = nullptr;
                size_t glds = std::max(lds, lds + lds / 8);
                size_t limit = b->m[0].c->lds_limit;
                if (glds > limit) glds = std::max(lds, limit);
                TRY(capture(b, &b->g2, 2, glds));
                b->g2_nblk = b->nblk2;
                b->g2_lds = glds;
                b->g2_n = n;
            }
            FM_HIP(hipGraphLaunch(b->g2, b->stream));
        } else {
            TRY(enqueue_forces(b, lds));
        }
        for (size_t k = 0; k < n; ++k) {
            Member& mm = b->m[k];
            fm_ctx* c = mm.c;
            const bool rec = (mm.iter % c->p.tplot) == 0;
            const bool ev = c->p.teval > 0 && (mm.iter % c->p.teval) == 0;
            c->pe_valid = rec;
            c->vir_valid = rec || ev;
            c->simtime = c->simtime + c->p.delta_t;
            mm.iter += 1;
        }
    }
    FM_HIP(hipStreamSynchronize(b->stream));
    for (auto& mm : b->m) TRY(ctx_sync_flags(mm.c));
    b->rebuilds += nrb;
    if (n_rebuilds) *n_rebuilds = nrb;
    return 0;
}
