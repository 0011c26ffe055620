// fm_api.hip -- C ABI of libflowmol_hip: context lifecycle, host<->device state and the
// per-routine entry points that replace the reference's force / neighbour / integration bodies.
#include "fm_internal.h"

#include <climits>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

static thread_local char g_err[1024];

int fm_set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char* fm_last_error(void) { return g_err; }
extern "C" int fm_abi_version(void) { return FM_ABI_VERSION; }

// ---- kernel timing ----------------------------------------------------------------------------
const char* const kt_names[KT_NCLASS] = {"vv_pass1", "checkrebuild", "sendmols", "assign_to_cell",
                                         "neighbourlist", "forces", "vv_pass2", "macros", "forces_pass1"};

struct KTimer {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    struct Rec {
        int cls;
        size_t idx;  // index of the start event
        int step;    // step the launch belongs to (fm_run_vv), for dropping launches that did nothing
    };
    std::vector<Rec> rec;
    int open_cls = -1;
    size_t open_idx = 0;
    double total_ms[KT_NCLASS] = {0};
    long long count[KT_NCLASS] = {0};
};

static hipEvent_t kt_event(KTimer* t) {
    if (t->used == t->pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        t->pool.push_back(e);
    }
    return t->pool[t->used++];
}

static void kt_collect(fm_ctx* c) {
    KTimer* t = c->timer;
    if (!t || t->rec.empty()) return;
    hipStreamSynchronize(c->stream);
    for (auto& r : t->rec) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t->pool[r.idx], t->pool[r.idx + 1]) == hipSuccess) {
            t->total_ms[r.cls] += ms;
            t->count[r.cls] += 1;
        }
    }
    t->rec.clear();
    t->used = 0;
}

void ktimer_begin(fm_ctx* c, int cls) {
    KTimer* t = c->timer;
    if (!t || !t->on) return;
    if (t->used + 2 > 4096) kt_collect(c);
    hipEvent_t a = kt_event(t);
    hipEvent_t b = kt_event(t);
    if (!a || !b) return;
    t->open_cls = cls;
    t->open_idx = t->used - 2;
    hipEventRecord(a, c->stream);
}

void ktimer_end(fm_ctx* c, int cls) {
    KTimer* t = c->timer;
    if (!t || !t->on || t->open_cls != cls) return;
    hipEventRecord(t->pool[t->open_idx + 1], c->stream);
    t->rec.push_back({cls, t->open_idx, c->kt_step});
    t->open_cls = -1;
}

// forget the pending records of class cls from step `from` on (launches that found a rebuild
// pending and did nothing)
static void ktimer_drop(fm_ctx* c, int cls, int from) {
    KTimer* t = c->timer;
    if (!t || !t->on) return;
    size_t k = 0;
    for (auto& r : t->rec)
        if (!(r.cls == cls && r.step >= from)) t->rec[k++] = r;
    t->rec.resize(k);
}

// collect before a queue of launches whose records may be dropped afterwards
static void ktimer_make_room(fm_ctx* c, size_t events) {
    KTimer* t = c->timer;
    if (t && t->on && t->used + events > 4096) kt_collect(c);
}

// ---- helpers ----------------------------------------------------------------------------------
#define CHECK_CTX(c) \
    if (!(c)) return fm_set_error(FM_EINVAL, "null context")

template <typename T>
static int dalloc(T** p, size_t n) {
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return fm_set_error(FM_ENOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
    return 0;
}

static int check_flags(fm_ctx* c);

// read the device flag words; report the first device-side error
static int sync_flags(fm_ctx* c) {
    FM_HIP(hipMemcpyAsync(c->h_flags, c->d_flags, FLAG_COUNT * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipStreamSynchronize(c->stream));
    return check_flags(c);
}

// the first device-side error in the flag words read last (h_flags)
static int check_flags(fm_ctx* c) {
    int code = c->h_flags[FLAG_ERROR];
    if (code) {
        int mol = c->h_flags[FLAG_ERROR_MOL];
        // clear so that the context can be inspected afterwards
        launch_fill_i32(c, c->d_flags + FLAG_ERROR, 0, 2);
        switch (code) {
        case FM_EESCAPE:
            return fm_set_error(code, "Aborting due to escaped mol. (molecule %d outside the domain and its halo)", mol + 1);
        case FM_EOVERFLOW:
            return fm_set_error(code, "capacity exceeded (neighbour list or cell) at molecule/cell %d", mol);
        case FM_EBOND:
            return fm_set_error(code, "polymer bond broken or too many bonds for molno %d", mol + 1);
        default:
            return fm_set_error(code, "device error %d at %d", code, mol);
        }
    }
    return 0;
}

int ctx_sync_flags(fm_ctx* c) { return sync_flags(c); }
int check_flags_internal(fm_ctx* c) { return check_flags(c); }

static void free_ctx(fm_ctx* c) {
    if (!c) return;
    if (c->dev >= 0) hipSetDevice(c->dev);
    void* ptrs[] = {c->st[0], c->st[1], c->rdisp, c->ext[0], c->ext[1], c->slot_of, c->moltype[0], c->moltype[1],
                    c->tag[0], c->tag[1], c->teth[0], c->teth[1], c->mono[0], c->mono[1], c->cell_of,
                    c->cell_count, c->cell_start, c->bin_pos, c->scan_part, c->perm, c->nbl, c->nbn, c->bond, c->bondn,
                    c->pe_lj, c->pe_poly, c->vir, c->rf, c->pxy, c->cvface, c->planes, c->out_flag, c->out_vals,
                    c->out_slot, c->out_count, c->red, c->d_macros, c->d_ring, c->d_mie, c->d_vmax, c->d_rmax,
                    c->d_flags, c->stage, c->brick_order, c->btab, c->hcurve, c->sort_key,
                    c->sort_val, c->sort_key2, c->sort_val2, c->sort_cnt, c->sort_tmp};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (c->h_flags) hipHostFree(c->h_flags);
    if (c->h_ring) hipHostFree(c->h_ring);
    if (c->timer) {
        for (auto e : c->timer->pool) hipEventDestroy(e);
        delete c->timer;
    }
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

static int alloc_nlist(fm_ctx* c, int maxnb) {
    if (c->nbl) hipFree(c->nbl);
    c->nbl = nullptr;
    c->maxnb = (maxnb + 3) & ~3;
    return dalloc(&c->nbl, (size_t)(c->maxnb / 4) * c->cap);
}

// a build met a list of `need` entries: room for a quarter more
static int grow_nlist(fm_ctx* c, int need) { return alloc_nlist(c, need + need / 4 + 8); }

// Brick shape b and the brick order of the brick kernels: tiles of 4x4x4 bricks, x fastest, bricks
// x fastest inside a tile.  Workgroups are dealt to the 8 XCDs in turn and each XCD takes a
// contiguous run of this order, so the bricks running together on one XCD form compact blocks
// that share their staged halo cells in that XCD's L2.
static int set_brick(fm_ctx* c, const int b[3]) {
    int nb[3];
    for (int d = 0; d < 3; ++d) {
        c->brick[d] = std::max(1, std::min(b[d], c->nc[d]));
        nb[d] = (c->nc[d] + c->brick[d] - 1) / c->brick[d];
    }
    const int T = 4;
    std::vector<int> order;
    order.reserve((size_t)nb[0] * nb[1] * nb[2]);
    for (int tz = 0; tz < nb[2]; tz += T)
        for (int ty = 0; ty < nb[1]; ty += T)
            for (int tx = 0; tx < nb[0]; tx += T)
                for (int z = tz; z < std::min(tz + T, nb[2]); ++z)
                    for (int y = ty; y < std::min(ty + T, nb[1]); ++y)
                        for (int x = tx; x < std::min(tx + T, nb[0]); ++x) order.push_back(x + nb[0] * (y + nb[1] * z));
    if (c->brick_order) hipFree(c->brick_order);
    c->brick_order = nullptr;
    int rc = dalloc(&c->brick_order, order.size());
    if (rc) return rc;
    FM_HIP(hipMemcpy(c->brick_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
    return 0;
}

// Brick shape: the largest number of cells whose expected molecule count fits one pass of the
// workgroup (15/16 of its threads), preferring the shape with the least staged halo per cell.
static void choose_brick(const int nc[3], double occ, int out[3]) {
    const double target = FB_BLOCK * 15.0 / 16.0;
    double best = -1.0;
    out[0] = out[1] = out[2] = 1;
    for (int bx = 1; bx <= std::min(nc[0], 12); ++bx)
        for (int by = 1; by <= std::min(nc[1], 12); ++by)
            for (int bz = 1; bz <= std::min(nc[2], 12); ++bz) {
                const int P = bx * by * bz;
                if (P > 1 && P * occ > target) continue;
                const double eff = (double)P / ((bx + 2) * (by + 2) * (bz + 2));
                if (eff > best + 1e-12) {
                    best = eff;
                    out[0] = bx;
                    out[1] = by;
                    out[2] = bz;
                }
            }
}

// Block curve of sort_mols: SORT_FLAG 1 numbers the blocks x fastest
// (src/setup_set_parameters.f90:1141-1153); SORT_FLAG 2 walks them along the 3D Hilbert curve of
// build_hilbert (src/functions_lib.f90:2555-2636): the curve of order ceiling(log2(max nblocks))
// in the unit cube, scaled to the block grid, keeping the points that fall inside it.
static int build_block_curve(int flag, const int nb[3], std::vector<int>& out) {
    const size_t nblk = (size_t)nb[0] * nb[1] * nb[2];
    out.assign(nblk, -666);
    auto at = [&](int i, int j, int k) -> int& { return out[(size_t)(i - 1) + nb[0] * ((size_t)(j - 1) + (size_t)nb[1] * (k - 1))]; };
    if (flag == 1) {
        for (int i = 1; i <= nb[0]; ++i)
            for (int j = 1; j <= nb[1]; ++j)
                for (int k = 1; k <= nb[2]; ++k) at(i, j, k) = i + nb[0] * (j - 1) + nb[0] * nb[1] * (k - 1);
        return 0;
    }
    int maxcells = std::max(nb[0], std::max(nb[1], nb[2]));
    const int order = (int)ceil(log((double)maxcells) / log(2.0));
    // the curve of order n: eight transformed copies of the curve of order n-1
    std::vector<double> x(1, 0.0), y(1, 0.0), z(1, 0.0);
    for (int lev = 1; lev <= order; ++lev) {
        const size_t m = x.size();
        std::vector<double> X(8 * m), Y(8 * m), Z(8 * m);
        for (size_t k = 0; k < m; ++k) {
            const double a = x[k], b = y[k], c = z[k];
            const double px[8] = {0.5 + c, 0.5 + b, -0.5 + b, -0.5 - a, -0.5 - a, -0.5 - b, 0.5 - b, 0.5 + c};
            const double py[8] = {0.5 + a, 0.5 + c, 0.5 + c, 0.5 + b, -0.5 + b, -0.5 - c, -0.5 - c, -0.5 - a};
            const double pz[8] = {0.5 + b, -0.5 + a, -0.5 + a, 0.5 - c, 0.5 - c, -0.5 + a, -0.5 + a, 0.5 - b};
            for (int q = 0; q < 8; ++q) {
                X[q * m + k] = 0.5 * px[q];
                Y[q * m + k] = 0.5 * py[q];
                Z[q * m + k] = 0.5 * pz[q];
            }
        }
        x.swap(X);
        y.swap(Y);
        z.swap(Z);
    }
    if (order - log((double)maxcells) / log(2.0) > 0.0001) maxcells = (int)ceil(pow(8.0, (double)order / 3.0));
    double mn = x[0];
    for (size_t k = 0; k < x.size(); ++k) mn = std::min(mn, std::min(x[k], std::min(y[k], z[k])));
    const double shift = 0.00001 - mn;
    int m = 0;
    for (size_t k = 0; k < x.size(); ++k) {
        const int i = (int)ceil((x[k] + shift) * maxcells), j = (int)ceil((y[k] + shift) * maxcells),
                  kk = (int)ceil((z[k] + shift) * maxcells);
        if (i <= nb[0] && j <= nb[1] && kk <= nb[2]) at(i, j, kk) = ++m;
    }
    for (int v : out)
        if (v == -666) return fm_set_error(FM_EINVAL, "Error in generating Hilbert curve - change sort flag");
    return 0;
}

// ---- lifecycle --------------------------------------------------------------------------------
extern "C" int fm_init(const fm_params* p, int32_t np, int32_t capacity, fm_ctx** out) {
    if (!p || !out) return fm_set_error(FM_EINVAL, "null argument");
    *out = nullptr;
    if (np < 0) return fm_set_error(FM_EINVAL, "np < 0");
    for (int d = 0; d < 3; ++d)
        if (!(p->domain[d] > 0.0)) return fm_set_error(FM_EINVAL, "domain must be positive");
    if (!(p->rcutoff > 0.0) || !(p->delta_t > 0.0)) return fm_set_error(FM_EINVAL, "rcutoff and delta_t must be positive");
    if (p->tplot < 1) return fm_set_error(FM_EINVAL, "tplot must be >= 1");
    if (p->rebuild_criteria < 0 || p->rebuild_criteria > 2) return fm_set_error(FM_EINVAL, "rebuild_criteria must be 0, 1 or 2");
    if (p->rebuild_criteria == FM_REBUILD_FIXED && p->fixed_rebuild < 1) return fm_set_error(FM_EINVAL, "fixed_rebuild must be >= 1");
    if (p->ensemble != FM_ENSEMBLE_NVE && p->ensemble != FM_ENSEMBLE_TAG_MOVE)
        return fm_set_error(FM_EINVAL, "ensemble must be NVE (0) or tag_move (6)");
    if (p->sort_flag < 0 || p->sort_flag > 2) return fm_set_error(FM_EINVAL, "Incorrect value of sort_flag");
    if (p->sort_flag != 0 && (p->sort_freq < 1 || p->sortblocksize < 1))
        return fm_set_error(FM_EINVAL, "sort_freq and sortblocksize must be >= 1");
    if (p->sort_flag != 0 && p->potential_flag == 1)
        return fm_set_error(FM_EINVAL, "Sort should be turned off - Not developed for polymers");
    if (p->sort_flag != 0 && p->mie_potential == 1)
        return fm_set_error(FM_EINVAL, "Sort should be turned off - Not developed for Mie potential");
    if (p->integration_algorithm != FM_INTEGRATION_VV && p->integration_algorithm != FM_INTEGRATION_LFV)
        return fm_set_error(FM_EINVAL, "integration_algorithm must be leap-frog (0) or velocity Verlet (1)");
    if (p->precision != FM_PREC_FP64 && p->precision != FM_PREC_MIXED) return fm_set_error(FM_EINVAL, "bad precision");
    if (p->precision == FM_PREC_MIXED && (p->mie_potential || p->potential_flag))
        return fm_set_error(FM_EINVAL, "mixed precision supports the LJ pair law without bonds");
    if (p->vflux_outflag < 0 || p->vflux_outflag > 4) return fm_set_error(FM_EINVAL, "vflux_outflag must be 0..4");
    if (p->vflux_outflag == 4 && p->cv_conserve != 0)
        return fm_set_error(FM_EINVAL, "CV_conserve 1 (CV stresses every step) is not on this path");
    if (p->vflux_outflag >= 1 && p->vflux_outflag <= 3 &&
        (p->nplanes < 1 || p->nplanes > FM_MAX_PLANES || !(p->planespacing > 0.0)))
        return fm_set_error(FM_EINVAL, "MOP needs 1 <= nplanes <= %d and planespacing > 0", FM_MAX_PLANES);
    if (p->mie_potential == 1 && (p->ntypes < 1 || p->ntypes > FM_MAX_TYPES))
        return fm_set_error(FM_EINVAL, "ntypes must be 1..%d", FM_MAX_TYPES);

    fm_ctx* c = new fm_ctx();
    c->p = *p;
    c->dev = -1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        delete c;
        return fm_set_error(FM_EHIP, "no HIP device");
    }
    c->dev = dev;
    c->np = np;
    c->cap = std::max(std::max(np, capacity), 1);
    if (c->cap >= (1 << 27)) {
        free_ctx(c);
        return fm_set_error(FM_EINVAL, "capacity %d exceeds 2^27 molecules per context", c->cap);
    }

    // set_parameters_cells (src/setup_set_parameters.f90:1848-1903)
    for (int d = 0; d < 3; ++d) c->half[d] = 0.5 * p->domain[d];
    double rneighbr = p->rcutoff + (p->delta_rneighbr[0] + p->delta_rneighbr[1] + p->delta_rneighbr[2]) / 3.0;
    c->rn2 = rneighbr * rneighbr;
    if (p->potential_flag == 1 && rneighbr < p->R_0) {
        rneighbr = p->R_0;
        c->rn2 = p->R_0 * p->R_0;
    }
    for (int d = 0; d < 3; ++d) c->nc[d] = (int)floor(p->domain[d] / (p->rcutoff + p->delta_rneighbr[d]));
    if (c->nc[0] < 3 || c->nc[1] < 3 || c->nc[2] < 3) {
        free_ctx(c);
        return fm_set_error(FM_EINVAL, "ERROR - DOMAIN SHOULD HAVE AT LEAST 3 CELLS, IN X, Y AND Z - INCREASE NUMBER OF UNITS IN INPUT");
    }
    for (int d = 0; d < 3; ++d) c->cs[d] = p->domain[d] / c->nc[d];
    c->ncell = c->nc[0] * c->nc[1] * c->nc[2];
    c->rc2 = p->rcutoff * p->rcutoff;
    {
        double rc = p->rcutoff, x3 = rc * (rc * rc), x6 = x3 * x3, x12 = x6 * x6;
        c->potshift = 4.0 * (1.0 / x12 - 1.0 / x6);
    }
    c->dmin = std::min(p->delta_rneighbr[0], std::min(p->delta_rneighbr[1], p->delta_rneighbr[2]));
    if (p->brick[0] > 0 && p->brick[1] > 0 && p->brick[2] > 0) {
        for (int d = 0; d < 3; ++d) c->brick[d] = std::min(p->brick[d], c->nc[d]);
    } else {
        choose_brick(c->nc, std::max(1.0, (double)np / c->ncell), c->brick);
    }
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
            c->lds_limit = (size_t)prop.sharedMemPerBlock;
            c->lds_cu = std::max((size_t)prop.maxSharedMemoryPerMultiProcessor, c->lds_limit);
        }
        const char* tight = getenv("FLOWMOL_TIGHT_REGIONS");
        c->tight_regions = tight && tight[0] == '1';
    }

    // neighbour capacity: 1.5x the expected count at the mean density, plus margin
    int maxnb = p->max_neighbours;
    if (maxnb <= 0) {
        double vol = p->domain[0] * p->domain[1] * p->domain[2];
        double dens = std::max(p->density, np / vol);
        double expect = 4.0 / 3.0 * M_PI * rneighbr * rneighbr * rneighbr * dens;
        maxnb = (int)(1.5 * expect) + 32;
        maxnb = (maxnb + 3) & ~3;
    }

    int rc = 0;
    hipError_t he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        free_ctx(c);
        return fm_set_error(FM_EHIP, "hipStreamCreate: %s", hipGetErrorString(he));
    }
    const size_t cap = c->cap;
    const int T2 = FM_MAX_TYPES * FM_MAX_TYPES;
    c->nred_blocks = (int)((cap + FM_BLOCK - 1) / FM_BLOCK) + 1;
    if ((rc = dalloc(&c->st[0], 9 * cap)) || (rc = dalloc(&c->st[1], 9 * cap)) || (rc = dalloc(&c->rdisp, 3 * cap)) ||
        (rc = dalloc(&c->ext[0], cap)) || (rc = dalloc(&c->ext[1], cap)) || (rc = dalloc(&c->slot_of, cap)) ||
        (rc = dalloc(&c->cell_of, cap)) || (rc = dalloc(&c->cell_count, (size_t)c->ncell + 1)) ||
        (rc = dalloc(&c->cell_start, (size_t)c->ncell + 1)) || (rc = dalloc(&c->bin_pos, cap)) ||
        (rc = dalloc(&c->scan_part, (size_t)c->ncell / 1024 + 2)) ||
        (rc = dalloc(&c->perm, cap)) || (rc = dalloc(&c->nbn, cap)) || (rc = dalloc(&c->pe_lj, cap)) ||
        (rc = dalloc(&c->pe_poly, cap)) || (rc = dalloc(&c->vir, cap)) || (rc = dalloc(&c->out_flag, cap)) ||
        (rc = dalloc(&c->out_vals, cap)) || (rc = dalloc(&c->out_slot, cap)) || (rc = dalloc(&c->out_count, (cap + FM_BLOCK - 1) / FM_BLOCK + 4)) ||
        (rc = dalloc(&c->red, (size_t)c->nred_blocks * 8)) || (rc = dalloc(&c->d_macros, 16)) ||
        (rc = dalloc(&c->d_ring, 8 * (size_t)FM_RING_SLOTS)) ||
        (rc = dalloc(&c->d_mie, 7 * (size_t)T2 + FM_MAX_TYPES)) || (rc = dalloc(&c->d_vmax, 1)) ||
        (rc = dalloc(&c->d_rmax, 1)) || (rc = dalloc(&c->d_flags, FLAG_COUNT)) || (rc = dalloc(&c->stage, 9 * cap)) ||
        (rc = alloc_nlist(c, maxnb))) {
        free_ctx(c);
        return rc;
    }
    if (p->potential_flag == 1) {
        if ((rc = dalloc(&c->bond, FM_MAX_FUNCY * cap)) || (rc = dalloc(&c->bondn, cap)) ||
            (rc = dalloc(&c->mono[0], 8 * cap)) || (rc = dalloc(&c->mono[1], 8 * cap))) {
            free_ctx(c);
            return rc;
        }
    }
    if (p->mie_potential == 1) {
        if ((rc = dalloc(&c->moltype[0], cap)) || (rc = dalloc(&c->moltype[1], cap))) {
            free_ctx(c);
            return rc;
        }
    }
    if (p->ensemble == FM_ENSEMBLE_TAG_MOVE) {
        if ((rc = dalloc(&c->tag[0], cap)) || (rc = dalloc(&c->tag[1], cap)) || (rc = dalloc(&c->teth[0], 6 * cap)) ||
            (rc = dalloc(&c->teth[1], 6 * cap))) {
            free_ctx(c);
            return rc;
        }
    }
    if (p->pressure_outflag == 1 && (rc = dalloc(&c->rf, 9 * cap))) {
        free_ctx(c);
        return rc;
    }
    if (p->vflux_outflag >= 1 && p->vflux_outflag <= 3) {
        if ((rc = dalloc(&c->pxy, 3 * (size_t)p->nplanes)) || (rc = dalloc(&c->planes, (size_t)p->nplanes))) {
            free_ctx(c);
            return rc;
        }
        // MOP planes, src/setup_set_parameters.f90:2298-2340
        int pn = p->vflux_outflag - 1;
        double shift = 0.5 * (p->domain[pn] - p->planespacing * (p->nplanes - 1));
        c->mop_shift = shift;
        std::vector<double> pl(p->nplanes);
        for (int n = 1; n <= p->nplanes; ++n) pl[n - 1] = p->planespacing * (n - 1) + shift - c->half[pn];
        if (hipMemcpy(c->planes, pl.data(), pl.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            free_ctx(c);
            return fm_set_error(FM_EHIP, "planes upload failed");
        }
    }
    if (p->vflux_outflag == 4) {
        // averaging bins for one process, src/setup_set_parameters.f90:1990-2011
        size_t nbo = 1;
        for (int d = 0; d < 3; ++d) {
            c->cv_nbins[d] = (int)(p->binspercell[d] * c->nc[d]);  // real -> integer assignment
            if (c->cv_nbins[d] < 1) {
                free_ctx(c);
                return fm_set_error(FM_EINVAL, "Error -- bin size less than one, decrease bin to cell ratio or use bigger domain");
            }
            c->cv_nhb[d] = (int)ceil((double)c->cv_nbins[d] / (double)c->nc[d]);
            c->cv_nbo[d] = c->cv_nbins[d] + 2 * c->cv_nhb[d];
            c->cv_bs[d] = p->domain[d] / c->cv_nbins[d];
            nbo *= (size_t)c->cv_nbo[d];
        }
        if ((rc = dalloc(&c->cvface, 9 * nbo)) || (rc = launch_fill_f64(c, c->cvface, 0.0, 9 * nbo))) {
            free_ctx(c);
            return rc;
        }
    }
    if ((rc = set_brick(c, c->brick))) {
        free_ctx(c);
        return rc;
    }
    if (p->sort_flag != 0) {
        for (int d = 0; d < 3; ++d) {
            c->sort_bsl[d] = p->sortblocksize * c->cs[d];
            c->sort_nb[d] = (int)ceil(p->domain[d] / c->sort_bsl[d]);
        }
        std::vector<int> hc;
        if ((rc = build_block_curve(p->sort_flag, c->sort_nb, hc))) {
            free_ctx(c);
            return rc;
        }
        c->sort_tmp_bytes = radix_sort_bytes((int)cap);
        if ((rc = dalloc(&c->hcurve, hc.size())) || (rc = dalloc(&c->sort_key, cap)) ||
            (rc = dalloc(&c->sort_val, cap)) || (rc = dalloc(&c->sort_key2, cap)) || (rc = dalloc(&c->sort_val2, cap)) ||
            (rc = dalloc(&c->sort_cnt, hc.size() + 1)) || (rc = dalloc((char**)&c->sort_tmp, c->sort_tmp_bytes))) {
            free_ctx(c);
            return rc;
        }
        if (hipMemcpy(c->hcurve, hc.data(), hc.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
            free_ctx(c);
            return fm_set_error(FM_EHIP, "block curve upload failed");
        }
    }
    {
        std::vector<double> mie(7 * T2 + FM_MAX_TYPES, 0.0);
        for (int t = 0; t < T2; ++t) {
            mie[t] = p->epsilon_lookup[t];
            mie[T2 + t] = p->lambdar_lookup[t];
            mie[2 * T2 + t] = p->lambdaa_lookup[t];
            mie[3 * T2 + t] = p->C_lookup[t];
            mie[4 * T2 + t] = p->alpha_lookup[t];
            mie[5 * T2 + t] = p->potshift_lookup[t];
            mie[6 * T2 + t] = p->rcutoff2_lookup[t];
        }
        for (int t = 0; t < FM_MAX_TYPES; ++t) mie[7 * T2 + t] = p->mie_potential == 1 ? p->mass_lookup[t] : 1.0;
        if (hipMemcpy(c->d_mie, mie.data(), mie.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            free_ctx(c);
            return fm_set_error(FM_EHIP, "table upload failed");
        }
    }
    he = hipHostMalloc((void**)&c->h_flags, FLAG_COUNT * sizeof(int), hipHostMallocDefault);
    if (he == hipSuccess) he = hipHostMalloc((void**)&c->h_ring, 8 * FM_RING_SLOTS * sizeof(double), hipHostMallocDefault);
    if (he != hipSuccess) {
        free_ctx(c);
        return fm_set_error(FM_EHIP, "hipHostMalloc: %s", hipGetErrorString(he));
    }
    c->timer = new KTimer();
    if ((rc = launch_fill_i32(c, c->d_flags, 0, FLAG_COUNT)) ||
        (rc = launch_fill_i32(c, c->cell_count, 0, (size_t)c->ncell + 1)) || (rc = launch_fill_f64(c, c->d_rmax, 0.0, 1)) ||
        (rc = launch_fill_f64(c, c->st[0], 0.0, 9 * cap)) || (rc = launch_fill_f64(c, c->pe_lj, 0.0, cap)) ||
        (rc = launch_fill_f64(c, c->pe_poly, 0.0, cap)) || (rc = launch_fill_f64(c, c->vir, 0.0, cap))) {
        free_ctx(c);
        return rc;
    }
    if (c->pxy && (rc = launch_fill_f64(c, c->pxy, 0.0, 3 * (size_t)p->nplanes))) {
        free_ctx(c);
        return rc;
    }
    if (launch_fill_f64(c, (double*)c->d_vmax, 0.0, 1) || hipStreamSynchronize(c->stream) != hipSuccess) {
        free_ctx(c);
        return fm_set_error(FM_EHIP, "initialisation failed");
    }
    *out = c;
    return 0;
}

// Device bytes fm_init allocates for these parameters (admission control by HBM footprint).
extern "C" int fm_memory_footprint(const fm_params* p, int32_t np, int64_t* bytes) {
    if (!p || !bytes || np < 0) return fm_set_error(FM_EINVAL, "bad argument");
    int nc[3];
    for (int d = 0; d < 3; ++d) {
        nc[d] = (int)floor(p->domain[d] / (p->rcutoff + p->delta_rneighbr[d]));
        if (nc[d] < 3) return fm_set_error(FM_EINVAL, "fewer than 3 cells");
    }
    const size_t ncell = (size_t)nc[0] * nc[1] * nc[2];
    const size_t cap = std::max(np, 1);
    double rneighbr = p->rcutoff + (p->delta_rneighbr[0] + p->delta_rneighbr[1] + p->delta_rneighbr[2]) / 3.0;
    if (p->potential_flag == 1 && rneighbr < p->R_0) rneighbr = p->R_0;
    int maxnb = p->max_neighbours;
    if (maxnb <= 0) {
        const double vol = p->domain[0] * p->domain[1] * p->domain[2];
        const double dens = std::max(p->density, np / vol);
        maxnb = (int)(1.5 * (4.0 / 3.0 * M_PI * rneighbr * rneighbr * rneighbr * dens)) + 32;
    }
    maxnb = (maxnb + 3) & ~3;
    size_t b = 0;
    b += 2 * 9 * cap * 8 + 3 * cap * 8 + 9 * cap * 8;          // state copies, rdisp, staging
    b += (2 + 1 + 1 + 1 + 1 + 1 + 3) * cap * 4;                 // ext, slot_of, cell_of, bin_pos, perm, nbn, sendmols scratch
    b += 3 * cap * 8;                                           // pe_lj, pe_poly, vir
    b += (size_t)(maxnb / 4) * cap * 8;                         // neighbour words
    b += 3 * (ncell + 1) * 4;                                   // cell counts, starts, cursors
    b += cap * 4 + 4 * 1024 * 1024;                             // scan/select temporary, reductions
    b += 8 * 1024 * 1024;                                       // the allocator's rounding of ~45 buffers
    {
        // staging tables: 10 B per molecule of region capacity per brick (1.125 of the capacity);
        // the regions hold each molecule prod((B+2)/B) times over, and the capacity is up to
        // 1.6 times the mean region
        int br[3];
        if (p->brick[0] > 0 && p->brick[1] > 0 && p->brick[2] > 0) {
            for (int d = 0; d < 3; ++d) br[d] = std::min(p->brick[d], nc[d]);
        } else {
            choose_brick(nc, std::max(1.0, (double)np / ncell), br);
        }
        double halo = 1.0;
        for (int d = 0; d < 3; ++d) halo *= (br[d] + 2.0) / br[d];
        b += (size_t)(10.0 * 1.125 * 1.6 * halo * (double)cap) + 4096;
    }
    if (p->potential_flag == 1) b += FM_MAX_FUNCY * cap * 4 + cap * 4 + 2 * 7 * cap * 4;
    if (p->mie_potential == 1) b += 2 * cap * 4;
    if (p->ensemble == FM_ENSEMBLE_TAG_MOVE) b += 2 * cap * 4 + 2 * 6 * cap * 8;
    if (p->pressure_outflag == 1) b += 9 * cap * 8;
    if (p->vflux_outflag >= 1 && p->vflux_outflag <= 3) b += 4 * (size_t)std::max(p->nplanes, 0) * 8;
    if (p->sort_flag != 0) b += 4 * cap * 4 + 2 * (ncell + 1) * 4 + cap * 8 + 8192;  // keys, values, curve, counts, sort temporary
    if (p->vflux_outflag == 4) {
        size_t nbo = 1;
        for (int d = 0; d < 3; ++d) {
            const int nb = std::max((int)(p->binspercell[d] * nc[d]), 1);
            nbo *= (size_t)(nb + 2 * (int)ceil((double)nb / nc[d]));
        }
        b += 9 * nbo * 8;
    }
    *bytes = (int64_t)b;
    return 0;
}

extern "C" void fm_finalize(fm_ctx* c) {
    if (c) hipStreamSynchronize(c->stream);
    free_ctx(c);
}

extern "C" int fm_get_geometry(const fm_ctx* c, int32_t ncells[3], double cellsidelength[3], double* rneighbr2,
                               double* rcutoff2, double* potshift) {
    CHECK_CTX(c);
    for (int d = 0; d < 3; ++d) {
        if (ncells) ncells[d] = c->nc[d];
        if (cellsidelength) cellsidelength[d] = c->cs[d];
    }
    if (rneighbr2) *rneighbr2 = c->rn2;
    if (rcutoff2) *rcutoff2 = c->rc2;
    if (potshift) *potshift = c->potshift;
    return 0;
}

extern "C" int fm_get_capacity(const fm_ctx* c, int32_t* list_entries, int32_t* region_molecules) {
    CHECK_CTX(c);
    if (list_entries) *list_entries = c->maxnb;
    if (region_molecules) *region_molecules = c->maxreg;
    return 0;
}

extern "C" int fm_get_np(const fm_ctx* c, int32_t* np) {
    CHECK_CTX(c);
    *np = c->np;
    return 0;
}

extern "C" double fm_simtime(const fm_ctx* c) { return c ? c->simtime : 0.0; }

// simtime of a context resumed from a checkpoint (src/modules.f90:263; final_state header)
extern "C" int fm_set_simtime(fm_ctx* c, double simtime) {
    CHECK_CTX(c);
    c->simtime = simtime;
    return 0;
}

// ---- state transfer ---------------------------------------------------------------------------
extern "C" int fm_upload_state(fm_ctx* c, int32_t np, const double* r, const double* v, const double* a) {
    CHECK_CTX(c);
    if (np < 0 || np > c->cap) return fm_set_error(FM_EINVAL, "np %d exceeds capacity %d", np, c->cap);
    if (!r || !v) return fm_set_error(FM_EINVAL, "r and v are required");
    c->np = np;
    const size_t n3 = 3 * (size_t)np;
    FM_HIP(hipMemcpyAsync(c->stage, r, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    FM_HIP(hipMemcpyAsync(c->stage + n3, v, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (a) FM_HIP(hipMemcpyAsync(c->stage + 2 * n3, a, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    else TRY(launch_fill_f64(c, c->stage + 2 * n3, 0.0, n3));
    TRY(launch_from_host(c, c->stage));
    FM_HIP(hipStreamSynchronize(c->stream));
    c->cells_valid = c->nlist_valid = c->rdisp_valid = false;
    c->pe_valid = c->vir_valid = false;
    c->moved_since_bin = true;
    // per-molecule properties are in the previous slot order: they must be uploaded again
    c->has_moltype = c->has_tags = c->has_mono = false;
    return 0;
}

static int upload_i32(fm_ctx* c, int* dst, const int32_t* src, int nf) {
    const size_t n = (size_t)nf * c->np;
    int* tmp = (int*)c->stage;  // staging holds 9*cap doubles >= 7*cap ints
    FM_HIP(hipMemcpyAsync(tmp, src, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    TRY(launch_ext_to_slot_i32(c, dst, tmp, nf));
    FM_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int fm_upload_moltype(fm_ctx* c, const int32_t* moltype) {
    CHECK_CTX(c);
    if (!c->moltype[0]) return fm_set_error(FM_EINVAL, "moltype needs mie_potential = 1");
    for (int n = 0; n < c->np; ++n)
        if (moltype[n] < 0 || moltype[n] >= c->p.ntypes) return fm_set_error(FM_EINVAL, "moltype out of range at %d", n);
    TRY(upload_i32(c, c->moltype[c->cur], moltype, 1));
    c->has_moltype = true;
    return 0;
}

extern "C" int fm_upload_tags(fm_ctx* c, const int32_t* tag, const double* rtether) {
    CHECK_CTX(c);
    if (!c->tag[0]) return fm_set_error(FM_EINVAL, "tags need ensemble = tag_move");
    for (int n = 0; n < c->np; ++n)
        if (tag[n] != FM_TAG_FREE && tag[n] != FM_TAG_FIXED && tag[n] != FM_TAG_FIXED_SLIDE && tag[n] != FM_TAG_TETH &&
            tag[n] != FM_TAG_TETH_SLIDE)
            return fm_set_error(FM_EINVAL, "Invalid molecular Tag %d for molecule %d (thermostatted tags are not on this path)",
                                tag[n], n + 1);
    TRY(upload_i32(c, c->tag[c->cur], tag, 1));
    // rtether (AoS) and slidev from read_tag (src/molecular_properties.f90:764-830)
    std::vector<double> h(6 * (size_t)c->np, 0.0), r(3 * (size_t)c->np);
    FM_HIP(hipStreamSynchronize(c->stream));
    if (fm_download_state(c, r.data(), nullptr, nullptr)) return FM_EHIP;
    for (int n = 0; n < c->np; ++n) {
        for (int d = 0; d < 3; ++d) h[6 * (size_t)n + d] = rtether ? rtether[3 * (size_t)n + d] : r[3 * (size_t)n + d];
        double sg = tag[n] == FM_TAG_TETH_SLIDE ? copysign(1.0, r[3 * (size_t)n + 1]) : 0.0;
        for (int d = 0; d < 3; ++d)
            h[6 * (size_t)n + 3 + d] = tag[n] == FM_TAG_TETH_SLIDE    ? c->p.wallslidev[d] * sg
                                       : tag[n] == FM_TAG_FIXED_SLIDE ? c->p.wallslidev[d]
                                                                      : 0.0;
    }
    FM_HIP(hipMemcpyAsync(c->stage, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TRY(launch_ext_to_slot_f64(c, c->teth[c->cur], c->stage, 6));
    FM_HIP(hipStreamSynchronize(c->stream));
    c->has_tags = true;
    return 0;
}

// glob_no may be null: then molecule n carries glob_no n+1, as setup numbers them on one process
extern "C" int fm_upload_monomers(fm_ctx* c, const int32_t* chainID, const int32_t* subchainID, const int32_t* funcy,
                                  const int32_t* bin_bflag, const int32_t* glob_no) {
    CHECK_CTX(c);
    if (!c->mono[0]) return fm_set_error(FM_EINVAL, "monomers need potential_flag = 1");
    const size_t np = c->np;
    std::vector<int32_t> h(8 * np);
    for (size_t n = 0; n < np; ++n) {
        if (funcy[n] < 0 || funcy[n] > FM_MAX_FUNCY) return fm_set_error(FM_EINVAL, "funcy out of range at %zu", n);
        h[8 * n] = chainID[n];
        h[8 * n + 1] = subchainID[n];
        h[8 * n + 2] = funcy[n];
        for (int g = 0; g < 4; ++g) h[8 * n + 3 + g] = bin_bflag[4 * n + g];
        h[8 * n + 7] = glob_no ? glob_no[n] : (int32_t)(n + 1);
    }
    TRY(upload_i32(c, c->mono[c->cur], h.data(), 8));
    c->has_mono = true;
    return 0;
}

extern "C" int fm_download_state(fm_ctx* c, double* r, double* v, double* a) {
    CHECK_CTX(c);
    const size_t n3 = 3 * (size_t)c->np;
    TRY(launch_scatter_state_to_host_order(c, c->stage, nullptr));
    if (r) FM_HIP(hipMemcpyAsync(r, c->stage, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (v) FM_HIP(hipMemcpyAsync(v, c->stage + n3, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (a) FM_HIP(hipMemcpyAsync(a, c->stage + 2 * n3, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int fm_download_tags(fm_ctx* c, double* rtether) {
    CHECK_CTX(c);
    if (!c->has_tags) return fm_set_error(FM_EINVAL, "no tags uploaded");
    TRY(launch_slot_to_ext_f64(c, c->stage, c->teth[c->cur], 3));
    FM_HIP(hipMemcpyAsync(rtether, c->stage, 3 * (size_t)c->np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// Per-molecule labels in the current molecule numbering: sendmols and sort_mols renumber the
// molecules, and tag, moltype and monomer move with them (src/messenger.MPI.f90:1460-1904,
// src/linklist.f90:2820-2860), so a host mirror reads them back after a rebuild.
extern "C" int fm_download_labels(fm_ctx* c, int32_t* tag, int32_t* moltype, int32_t* chainID, int32_t* subchainID,
                                  int32_t* funcy, int32_t* bin_bflag, int32_t* glob_no) {
    CHECK_CTX(c);
    const size_t np = c->np;
    int* st = (int*)c->stage;
    if (tag) {
        if (!c->has_tags) return fm_set_error(FM_EINVAL, "no tags uploaded");
        TRY(launch_slot_to_ext_i32(c, st, c->tag[c->cur], 1));
        FM_HIP(hipMemcpyAsync(tag, st, np * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        FM_HIP(hipStreamSynchronize(c->stream));
    }
    if (moltype) {
        if (!c->has_moltype) return fm_set_error(FM_EINVAL, "no moltype uploaded");
        TRY(launch_slot_to_ext_i32(c, st, c->moltype[c->cur], 1));
        FM_HIP(hipMemcpyAsync(moltype, st, np * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        FM_HIP(hipStreamSynchronize(c->stream));
    }
    if (chainID || subchainID || funcy || bin_bflag || glob_no) {
        if (!c->has_mono) return fm_set_error(FM_EINVAL, "no monomer data uploaded");
        std::vector<int32_t> h(8 * np);
        TRY(launch_slot_to_ext_i32(c, st, c->mono[c->cur], 8));
        FM_HIP(hipMemcpyAsync(h.data(), st, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        FM_HIP(hipStreamSynchronize(c->stream));
        for (size_t n = 0; n < np; ++n) {
            if (chainID) chainID[n] = h[8 * n];
            if (subchainID) subchainID[n] = h[8 * n + 1];
            if (funcy) funcy[n] = h[8 * n + 2];
            if (bin_bflag)
                for (int g = 0; g < 4; ++g) bin_bflag[4 * n + g] = h[8 * n + 3 + g];
            if (glob_no) glob_no[n] = h[8 * n + 7];
        }
    }
    return 0;
}

// potenergymol_LJ and potenergymol_POLY of the last record step (src/modules.f90:371-375)
extern "C" int fm_download_potenergy_parts(fm_ctx* c, double* pe_lj, double* pe_poly) {
    CHECK_CTX(c);
    const size_t np = c->np;
    const bool poly = c->p.potential_flag == 1 && c->has_mono;
    for (int k = 0; k < 2; ++k) {
        double* out = k == 0 ? pe_lj : pe_poly;
        if (!out) continue;
        if (!c->pe_valid || (k == 1 && !poly)) {
            for (size_t n = 0; n < np; ++n) out[n] = 0.0;
            continue;
        }
        TRY(launch_slot_to_ext_f64(c, c->stage, k == 0 ? c->pe_lj : c->pe_poly, 1));
        FM_HIP(hipMemcpyAsync(out, c->stage, np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        FM_HIP(hipStreamSynchronize(c->stream));
    }
    return 0;
}

extern "C" int fm_download_molecular(fm_ctx* c, double* potenergymol, double* virialmol) {
    CHECK_CTX(c);
    const size_t np = c->np;
    std::vector<double> pl(np), pp(np, 0.0);
    if (potenergymol) {
        if (c->pe_valid) {
            TRY(launch_slot_to_ext_f64(c, c->stage, c->pe_lj, 1));
            FM_HIP(hipMemcpyAsync(pl.data(), c->stage, np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            FM_HIP(hipStreamSynchronize(c->stream));
            if (c->p.potential_flag == 1 && c->has_mono) {
                TRY(launch_slot_to_ext_f64(c, c->stage, c->pe_poly, 1));
                FM_HIP(hipMemcpyAsync(pp.data(), c->stage, np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
                FM_HIP(hipStreamSynchronize(c->stream));
            }
            for (size_t n = 0; n < np; ++n) potenergymol[n] = pl[n] + pp[n];
        } else {
            for (size_t n = 0; n < np; ++n) potenergymol[n] = 0.0;
        }
    }
    if (virialmol) {
        if (c->vir_valid) {
            TRY(launch_slot_to_ext_f64(c, c->stage, c->vir, 1));
            FM_HIP(hipMemcpyAsync(virialmol, c->stage, np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            FM_HIP(hipStreamSynchronize(c->stream));
        } else {
            for (size_t n = 0; n < np; ++n) virialmol[n] = 0.0;
        }
    }
    return 0;
}

extern "C" int fm_download_rfmol(fm_ctx* c, double* rfmol) {
    CHECK_CTX(c);
    if (!c->rf) return fm_set_error(FM_EINVAL, "rfmol needs pressure_outflag = 1");
    TRY(launch_slot_to_ext_f64(c, c->stage, c->rf, 9));
    FM_HIP(hipMemcpyAsync(rfmol, c->stage, 9 * (size_t)c->np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// ---- the routines -----------------------------------------------------------------------------
extern "C" int fm_move_particles_vv(fm_ctx* c, int32_t pass_num) {
    CHECK_CTX(c);
    if (c->p.integration_algorithm != FM_INTEGRATION_VV)
        return fm_set_error(FM_ESTATE, "integration_algorithm is leap-frog: use fm_move_particles_lfv");
    if (pass_num == 1) {
        c->moved_since_bin = true;
        TRY(launch_pass1(c));
    } else if (pass_num == 2) {
        TRY(launch_pass2(c));
        c->simtime = c->simtime + c->p.delta_t;
    } else {
        return fm_set_error(FM_EINVAL, "pass_num must be 1 or 2");
    }
    return 0;
}

extern "C" int fm_move_particles_lfv(fm_ctx* c) {
    CHECK_CTX(c);
    if (c->p.integration_algorithm != FM_INTEGRATION_LFV)
        return fm_set_error(FM_ESTATE, "integration_algorithm is velocity Verlet: use fm_move_particles_vv");
    if (c->p.ensemble == FM_ENSEMBLE_TAG_MOVE && !c->has_tags) return fm_set_error(FM_ESTATE, "tag_move needs tags");
    c->moved_since_bin = true;
    TRY(launch_move_lfv(c, false));
    c->simtime = c->simtime + c->p.delta_t;
    if (c->p.ensemble == FM_ENSEMBLE_TAG_MOVE) TRY(sync_flags(c));
    return 0;
}

// queue the rebuild check of step iter on the device; the decision lands in d_flags[FLAG_REBUILD]
// except for the host-known cases returned in *host_rb
static int queue_checkrebuild(fm_ctx* c, int iter, int* host_rb) {
    const int forced = ((iter + 1) % c->p.tplot) == 0;
    *host_rb = forced;
    switch (c->p.rebuild_criteria) {
    case FM_REBUILD_VMAX:
        TRY(launch_check_vmax(c));
        if (forced) {  // rmax is reset on any rebuild
            TRY(launch_fill_f64(c, c->d_rmax, 0.0, 1));
        }
        break;
    case FM_REBUILD_FIXED:
        if (iter % c->p.fixed_rebuild == 0) *host_rb = 1;
        break;
    case FM_REBUILD_DISPLACEMENT:
        // on a forced rebuild the displacement sums are discarded, so the update is skipped
        if (!forced) TRY(launch_check_disp(c));
        break;
    }
    return 0;
}

extern "C" int fm_checkrebuild(fm_ctx* c, int32_t iter, int32_t* rebuild) {
    CHECK_CTX(c);
    if (!rebuild) return fm_set_error(FM_EINVAL, "null rebuild");
    int host_rb = 0;
    TRY(queue_checkrebuild(c, iter, &host_rb));
    TRY(sync_flags(c));
    int rb = host_rb || c->h_flags[FLAG_REBUILD];
    if (c->h_flags[FLAG_REBUILD]) TRY(launch_fill_i32(c, c->d_flags + FLAG_REBUILD, 0, 1));
    if (rb) {
        c->rdisp_valid = false;
        if (c->p.rebuild_criteria == FM_REBUILD_VMAX) TRY(launch_fill_f64(c, c->d_rmax, 0.0, 1));
    }
    *rebuild = rb;
    return 0;
}

extern "C" int fm_linklist_deallocateall(fm_ctx* c) {
    CHECK_CTX(c);
    c->cells_valid = false;
    c->nlist_valid = false;
    return 0;
}

extern "C" int fm_sendmols(fm_ctx* c) {
    CHECK_CTX(c);
    TRY(launch_sendmols(c));
    c->cells_valid = c->nlist_valid = false;
    return 0;
}

// sort_mols, src/linklist.f90:2744-2865.  The state stays in slot (cell) order; only the
// molecule numbers (ext / slot_of) change, to those the reference's reordering gives.
extern "C" int fm_sort_mols(fm_ctx* c) {
    CHECK_CTX(c);
    if (c->sort_nrebuilds != c->p.sort_freq) {
        c->sort_nrebuilds = c->sort_nrebuilds + 1;
        return 0;
    }
    c->sort_nrebuilds = 1;
    if (c->p.sort_flag == 0) return 0;
    int skipped = 0;
    return launch_sort_mols(c, &skipped);
}

// The bricks' staging tables (fm_ctx::btab) for the current region capacity: reallocated when the
// capacity or the brick count grew (at a synchronisation point: no launch uses the old ones), and
// invalid until the next list build either way.
static int size_brick_tables(fm_ctx* c) {
    c->btab_valid = false;
    const int nbrick = brick_count(c);
    if (c->btab && c->maxreg <= c->btab_reg && nbrick <= c->btab_nbrick) {
        c->btab_stride = (2 + 2 * c->maxreg + (c->maxreg + 1) / 2 + 3) & ~3;
        return 0;
    }
    if (c->btab) hipFree(c->btab);
    c->btab = nullptr;
    c->btab_reg = c->btab_nbrick = 0;
    const int reg = c->maxreg + c->maxreg / 8;  // headroom for later growth
    const int stride = (2 + 2 * reg + (reg + 1) / 2 + 3) & ~3;
    TRY(dalloc(&c->btab, (size_t)stride * nbrick));
    c->btab_reg = reg;
    c->btab_nbrick = nbrick;
    c->btab_stride = (2 + 2 * c->maxreg + (c->maxreg + 1) / 2 + 3) & ~3;
    return 0;
}

// LDS of a brick kernel staging m molecules: the largest over the kernel kinds
static size_t region_bytes(const fm_ctx* c, int m) {
    return std::max(brick_lds_bytes(c, m, 0), brick_lds_bytes(c, m, 2));
}

// Region capacity from the largest region of a binning (molecules): the LDS those need sets how
// many workgroups share a CU; the capacity is raised to what that share of the CU's LDS holds
// (occupancy unchanged), capped by the 16-bit list indices.
static int size_regions(fm_ctx* c, int measured) {
    if (measured > 65536)
        return fm_set_error(FM_EOVERFLOW, "brick region holds %d molecules (16-bit list indices); use a smaller brick", measured);
    const int m0 = std::max(measured, 1);
    const size_t need = region_bytes(c, m0);
    if (need > c->lds_limit)
        return fm_set_error(FM_EOVERFLOW, "brick region needs %zu bytes of LDS (limit %zu); use a smaller brick", need,
                            c->lds_limit);
    if (c->tight_regions) {
        c->maxreg = m0;
        return size_brick_tables(c);
    }
    const size_t wgs = std::max<size_t>(c->lds_cu / need, 1);
    const size_t budget = std::min(c->lds_limit, (c->lds_cu / wgs) & ~(size_t)1023);
    int lo = m0, hi = 65536;  // largest m in [m0, 65536] with region_bytes(m) <= budget
    while (lo < hi) {
        const int mid = lo + (hi - lo + 1) / 2;
        if (region_bytes(c, mid) <= budget) lo = mid;
        else hi = mid - 1;
    }
    // the branch-free Mie launch stages fewer bytes per molecule: its workgroups per CU (two at
    // most, as its registers allow) stay what its LDS at the measured size gives
    const size_t mb = brick_lds_bytes(c, m0, 3);
    if (mb > 0) {
        const size_t w = std::min<size_t>(2, std::max<size_t>(c->lds_cu / mb, 1));
        const size_t bm = std::min(c->lds_limit, (c->lds_cu / w) & ~(size_t)1023);
        int l2 = m0, h2 = lo;
        while (l2 < h2) {
            const int mid = l2 + (h2 - l2 + 1) / 2;
            if (brick_lds_bytes(c, mid, 3) <= bm) l2 = mid;
            else h2 = mid - 1;
        }
        lo = l2;
    }
    c->maxreg = lo;
    return size_brick_tables(c);
}

extern "C" int fm_assign_to_cell(fm_ctx* c) {
    CHECK_CTX(c);
    static const int reset[] = {FLAG_MAX_REGION, FLAG_NEED, FLAG_HALT};
    TRY(launch_reset_flags(c, reset, 3));
    TRY(launch_bin(c));
    TRY(launch_region_max(c));
    TRY(sync_flags(c));
    TRY(size_regions(c, c->h_flags[FLAG_MAX_REGION]));
    c->moved_since_bin = false;
    c->cells_valid = true;
    c->nlist_valid = false;
    // the binning permutes the slots and the displacement sums are kept per slot: they start
    // again from zero (a rebuild that simulation_checkrebuild asked for has discarded them
    // already; the rebuild before a final_state write also lands here)
    c->rdisp_valid = false;
    return 0;
}

extern "C" int fm_updateborders(fm_ctx* c, int32_t) {
    CHECK_CTX(c);
    return 0;
}

extern "C" int fm_assign_to_neighbourlist(fm_ctx* c) {
    CHECK_CTX(c);
    if (!c->cells_valid) return fm_set_error(FM_ESTATE, "assign_to_neighbourlist before assign_to_cell");
    if (c->p.potential_flag == 1 && !c->has_mono) return fm_set_error(FM_ESTATE, "potential_flag 1 needs monomer data (fm_upload_monomers)");
    for (int attempt = 0;; ++attempt) {
        static const int reset[] = {FLAG_MAX_NB, FLAG_NEED, FLAG_HALT};
        TRY(launch_reset_flags(c, reset, 3));
        TRY(launch_nlist(c));
        TRY(sync_flags(c));
        const int need = c->h_flags[FLAG_NEED];
        if (need == 0) break;
        if (attempt == 2 || (need & NEED_REGION))
            return fm_set_error(FM_EOVERFLOW, "neighbour list build needs %d entries, region %d molecules",
                                c->h_flags[FLAG_MAX_NB], c->h_flags[FLAG_MAX_REGION]);
        // grow the list and build again
        TRY(grow_nlist(c, c->h_flags[FLAG_MAX_NB]));
    }
    c->nlist_valid = true;
    return 0;
}

extern "C" int fm_compute_forces(fm_ctx* c, int32_t iter) {
    CHECK_CTX(c);
    if (!c->nlist_valid) return fm_set_error(FM_ESTATE, "simulation_compute_forces before assign_to_neighbourlist");
    if (c->p.mie_potential == 1 && !c->has_moltype) return fm_set_error(FM_ESTATE, "Mie potential needs moltype (fm_upload_moltype)");
    if (c->p.potential_flag == 1 && !c->has_mono) return fm_set_error(FM_ESTATE, "potential_flag 1 needs monomer data (fm_upload_monomers)");
    if (c->p.ensemble == FM_ENSEMBLE_TAG_MOVE && !c->has_tags) return fm_set_error(FM_ESTATE, "tag_move needs tags (fm_upload_tags)");
    TRY(launch_forces(c, iter));
    const bool rec = (iter % c->p.tplot) == 0;
    const bool ev = c->p.teval > 0 && (iter % c->p.teval) == 0;
    c->pe_valid = rec;
    c->vir_valid = rec || ev;
    if (c->p.ensemble == FM_ENSEMBLE_TAG_MOVE || c->p.potential_flag == 1) TRY(sync_flags(c));
    return 0;
}

static void derive_macros_np(int32_t np_, double density, int32_t potential_flag, const double* s, fm_macros* o) {
    memset(o, 0, sizeof *o);
    const double np = (double)np_;
    o->potenergysum_LJ = s[0];
    o->potenergysum_POLY = s[1];
    o->potenergysum = (potential_flag == 1) ? s[2] : s[0];
    o->virial = s[3];
    o->mv2sum = s[4];
    o->vsum = s[5];
    o->msum = s[6];
    o->max_potenergymol = s[7];
    o->kinenergy = (0.5 * o->mv2sum) / np;
    o->potenergy = (0.5 * o->potenergysum) / np;
    o->totenergy = o->kinenergy + o->potenergy;
    o->temperature = o->mv2sum / (double)(3 * np_);
    o->pressure = (density / (double)(np_ * 3)) * (o->mv2sum + o->virial / 2.0);
}

// The macroscopic properties from the eight sums of fm_macros_sums_noalloc (host arithmetic only).
extern "C" int fm_derive_macros(const double sums[8], int32_t np, double density, int32_t potential_flag,
                                fm_macros* out) {
    if (!sums || !out || np < 1) return fm_set_error(FM_EINVAL, "fm_derive_macros: bad arguments");
    derive_macros_np(np, density, potential_flag, sums, out);
    return 0;
}

static void derive_macros(const fm_ctx* c, const double* s, fm_macros* o) {
    memset(o, 0, sizeof *o);
    const double np = (double)c->np;
    o->potenergysum_LJ = s[0];
    o->potenergysum_POLY = s[1];
    o->potenergysum = (c->p.potential_flag == 1) ? s[2] : s[0];
    o->virial = s[3];
    o->mv2sum = s[4];
    o->vsum = s[5];
    o->msum = s[6];
    o->max_potenergymol = s[7];
    o->kinenergy = (0.5 * o->mv2sum) / np;
    o->potenergy = (0.5 * o->potenergysum) / np;
    o->totenergy = o->kinenergy + o->potenergy;
    o->temperature = o->mv2sum / (double)(3 * c->np);
    o->pressure = (c->p.density / (double)(c->np * 3)) * (o->mv2sum + o->virial / 2.0);
}

// The reference aborts when a molecule's potential energy exceeds 1e4 at a record step
// (src/simulation_record.f90:1271-1282); the sums are still returned.
static int check_max_pe(const fm_macros* m) {
    if (m->max_potenergymol > 10000.0)
        return fm_set_error(FM_EABORT,
                            "Simulation aborted because max PE has reached an unreasonably high value "
                            "(max(potenergymol) = %.6g)",
                            m->max_potenergymol);
    return 0;
}

extern "C" int fm_evaluate_macroscopic_properties(fm_ctx* c, fm_macros* out) {
    CHECK_CTX(c);
    if (!out) return fm_set_error(FM_EINVAL, "null output");
    TRY(launch_macros(c));
    double s[8];
    FM_HIP(hipMemcpyAsync(s, c->d_macros, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipStreamSynchronize(c->stream));
    derive_macros(c, s, out);
    return check_max_pe(out);
}

extern "C" int fm_get_mop(fm_ctx* c, double* pxy, int32_t nplanes) {
    CHECK_CTX(c);
    if (!c->pxy) return fm_set_error(FM_EINVAL, "MOP needs vflux_outflag 1..3");
    if (nplanes != c->p.nplanes) return fm_set_error(FM_EINVAL, "nplanes mismatch");
    FM_HIP(hipMemcpyAsync(pxy, c->pxy, 3 * (size_t)nplanes * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int fm_reset_mop(fm_ctx* c) {
    CHECK_CTX(c);
    if (!c->pxy) return 0;
    TRY(launch_fill_f64(c, c->pxy, 0.0, 3 * (size_t)c->p.nplanes));
    return 0;
}

extern "C" int fm_get_cv_bins(const fm_ctx* c, int32_t nbins[3], int32_t nhb[3], double binsize[3]) {
    CHECK_CTX(c);
    if (!c->cvface) return fm_set_error(FM_EINVAL, "CV stresses need vflux_outflag 4");
    for (int d = 0; d < 3; ++d) {
        nbins[d] = c->cv_nbins[d];
        nhb[d] = c->cv_nhb[d];
        binsize[d] = c->cv_bs[d];
    }
    return 0;
}

// Faces 4..6 are not tallied on the device: cumulative_flux_opt only ever assigns them the running
// total of face n of the bin one up along n (src/simulation_record.f90:5053-5057), and each bin's
// face n changes only together with that copy, so the copy is taken here.
extern "C" int fm_get_cv_stress(fm_ctx* c, double* out, int64_t n) {
    CHECK_CTX(c);
    if (!c->cvface) return fm_set_error(FM_EINVAL, "CV stresses need vflux_outflag 4");
    const int64_t nb0 = c->cv_nbo[0], nb1 = c->cv_nbo[1], nb2 = c->cv_nbo[2];
    const int64_t nbin = nb0 * nb1 * nb2;
    if (!out || n != 18 * nbin) return fm_set_error(FM_EINVAL, "Pxyface has %lld elements", (long long)(18 * nbin));
    FM_HIP(hipMemcpyAsync(out, c->cvface, 9 * (size_t)nbin * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipStreamSynchronize(c->stream));
    std::fill(out + 9 * nbin, out + 18 * nbin, 0.0);
    for (int f = 0; f < 3; ++f)
        for (int comp = 0; comp < 3; ++comp) {
            const double* src = out + (comp + 3 * (int64_t)f) * nbin;
            double* dst = out + (comp + 3 * (int64_t)(f + 3)) * nbin;
            for (int64_t k = 0; k < nb2; ++k)
                for (int64_t j = 0; j < nb1; ++j)
                    for (int64_t i = 0; i < nb0; ++i) {
                        int64_t u[3] = {i, j, k};
                        u[f] += 1;
                        if (u[f] >= c->cv_nbo[f]) continue;
                        dst[i + nb0 * (j + nb1 * k)] = src[u[0] + nb0 * (u[1] + nb1 * u[2])];
                    }
        }
    return 0;
}

extern "C" int fm_reset_cv_stress(fm_ctx* c) {
    CHECK_CTX(c);
    if (!c->cvface) return 0;
    TRY(launch_fill_f64(c, c->cvface, 0.0, 9 * (size_t)c->cv_nbo[0] * c->cv_nbo[1] * c->cv_nbo[2]));
    return 0;
}

// ---- fused md_advance_vv loop (src/md.f90:143-163) ---------------------------------------------
// Same operations as the call-per-routine sequence; VV pass 1 and the displacement update share
// one kernel, and VV pass 2 is applied by the force kernel as it writes a.
static int rebuild_lists(fm_ctx* c) {
    c->cells_valid = c->nlist_valid = false;
    TRY(launch_sendmols(c));
    TRY(fm_sort_mols(c));
    TRY(fm_assign_to_cell(c));
    return fm_assign_to_neighbourlist(c);
}

int rebuild_lists_internal(fm_ctx* c) { return rebuild_lists(c); }

static int rebuild_lists_queued(fm_ctx* c, bool resets, bool keep_acc);

// The rebuild of a fused loop whose list one force call reads, at the positions it is built from
// (the next step rebuilds before its force, and that step belongs to this loop): the pairs that
// call counts are those inside rc, so the list keeps only those (r^2 below rc^2, with a relative
// 1e-12 margin over the force kernels' own test on the same values).  The entries keep the order
// of the full list (the cell windows and the runs are those of the list radius), and an entry
// beyond rc adds exactly zero to every force, PE and virial sum, so each sum is the full list's,
// bit for bit.  Only for pairs that all end at rc: no bonds, no Mie tables.
static int rebuild_lists_for(fm_ctx* c, bool read_once, bool queued) {
    const bool skinless = read_once && c->p.potential_flag == 0 && c->p.mie_potential == 0;
    c->list_ra2 = skinless ? c->rc2 * (1.0 + 1e-12) : 0.0;
    const int rc = queued ? rebuild_lists_queued(c, true, false) : rebuild_lists(c);
    c->list_ra2 = 0.0;
    return rc;
}

// Whether step `it` + 1 rebuilds its lists whatever the displacement or velocity checks find (its
// record step is next, or the fixed interval)
static bool next_rebuild_certain(const fm_ctx* c, int it) {
    const bool fixed = c->p.rebuild_criteria != FM_REBUILD_DISPLACEMENT && c->p.rebuild_criteria != FM_REBUILD_VMAX;
    return ((it + 2) % c->p.tplot) == 0 || (fixed && (it + 1) % c->p.fixed_rebuild == 0);
}

// The same rebuild queued without waiting on the device (fused loops): the list build uses the
// region and list capacities of the last measured build, and a region or list beyond them raises
// FLAG_NEED and FLAG_HALT, so the conditional launches queued after it do nothing until
// redo_lists() has grown the capacities at the next synchronisation.
// resets: clear FLAG_REBUILD, FLAG_HALT, FLAG_NEED and the measured maxima first (a window with
// one rebuild at its start); without, a halt raised by an earlier build of the window stays.
// keep_acc: permute the accelerations too (no force call follows before the state is read).
static int rebuild_lists_queued(fm_ctx* c, bool resets = true, bool keep_acc = false) {
    static const int reset[] = {FLAG_REBUILD, FLAG_HALT, FLAG_NEED, FLAG_MAX_REGION, FLAG_MAX_NB};
    if (c->maxreg <= 0) {  // no measured capacity yet: the synchronous build
        TRY(launch_reset_flags(c, reset, 2));
        return rebuild_lists(c);
    }
    c->cells_valid = c->nlist_valid = false;
    TRY(launch_sendmols(c, reset, resets ? 5 : 0));
    TRY(fm_sort_mols(c));
    // every launch queued after this rebuild (the force launch, or nothing until the lists are
    // rebuilt) recomputes all accelerations before reading them
    TRY(launch_bin(c, keep_acc));
    c->moved_since_bin = false;
    c->cells_valid = true;
    TRY(launch_nlist(c));
    c->nlist_valid = true;
    return 0;
}

int rebuild_lists_queued_internal(fm_ctx* c) { return rebuild_lists_queued(c, true, false); }

// A synchronisation found FLAG_NEED set by the queued list build: grow the region and list
// capacities to what it met and build the lists again (the cells are unchanged, and every launch
// queued after the build did nothing).
static int redo_lists(fm_ctx* c) {
    for (int attempt = 0; attempt < 3; ++attempt) {
        const int need = c->h_flags[FLAG_NEED];
        if (need == 0) return 0;
        if (need & NEED_REGION) TRY(size_regions(c, std::max(c->h_flags[FLAG_MAX_REGION], c->maxreg + 1)));
        if (need & NEED_LIST) TRY(grow_nlist(c, std::max(c->h_flags[FLAG_MAX_NB], c->maxnb + 1)));
        static const int reset[] = {FLAG_HALT, FLAG_NEED, FLAG_MAX_REGION, FLAG_MAX_NB};
        TRY(launch_reset_flags(c, reset, 4));
        TRY(launch_nlist(c));
        TRY(sync_flags(c));
    }
    return c->h_flags[FLAG_NEED] ? fm_set_error(FM_EOVERFLOW, "list build still beyond its capacities") : 0;
}

int redo_lists_internal(fm_ctx* c) { return redo_lists(c); }

// bookkeeping of a completed step of the fused loops; ring: the step's record sums, taken on the
// device when the step was queued (null: evaluate them now, the state being that of this step)
static int finish_step(fm_ctx* c, int iter, fm_macros* macros, int n_macros, int* nm, const double* ring = nullptr) {
    const bool rec = (iter % c->p.tplot) == 0;
    const bool ev = c->p.teval > 0 && (iter % c->p.teval) == 0;
    c->pe_valid = rec;
    c->vir_valid = rec || ev;
    c->simtime = c->simtime + c->p.delta_t;
    if (rec && macros && *nm < n_macros) {
        if (ring) {
            derive_macros(c, ring, &macros[*nm]);
            ++*nm;
            return check_max_pe(&macros[*nm - 1]);
        }
        TRY(fm_evaluate_macroscopic_properties(c, &macros[*nm]));
        ++*nm;
    }
    return 0;
}

// md_advance_vv without a host synchronisation per step (rebuild criteria 1 and 2; bonds and
// tethers too, whose force launches are then not fused with the next pass 1).  Steps are queued in windows that end before the next step whose rebuild the host
// knows in advance (forced at mod(iter+1,tplot)==0, or the fixed interval), or after
// FM_RING_SLOTS steps; the host synchronises once per window.  The sums of a record step are
// reduced on the device right after its force launch, from the buffer that holds the step's
// final velocities, into a ring slot that the host reads at the window's synchronisation.  With the
// displacement criterion the queued launches are conditional: the first step whose check fires
// records its number and the launches after it do nothing; the next window then starts with that
// step's rebuild and force call.  A window that starts with a rebuild queues it too (the list
// capacities are checked at the window's synchronisation, and a build beyond them is repeated
// after growing them).  Inside a window the force launch of a plain step also applies VV pass 1
// of the next step and its displacement check, writing the state into the other buffer; the host
// keeps which buffer holds the state after each step's pass 1 for the rollback to a fired check.
// The arithmetic of every completed step is that of the step-by-step loop.
static int run_vv_queued(fm_ctx* c, int iter0, int nsteps, fm_macros* macros, int n_macros, int* nrb, int* nm) {
    const int end = iter0 + nsteps, W = FM_RING_SLOTS;
    const bool disp = c->p.rebuild_criteria == FM_REBUILD_DISPLACEMENT;
    auto host_rebuild = [&](int it) {
        return ((it + 1) % c->p.tplot) == 0 || (!disp && it % c->p.fixed_rebuild == 0);
    };
    static const int reset[] = {FLAG_REBUILD, FLAG_HALT};
    TRY(launch_reset_flags(c, reset, 2));
    int iter = iter0;
    int rb = -1;  // step whose displacement check fired (its pass 1 has run)
    // pass 1 of `iter` ran inside the last force launch of the previous window
    bool p1done = false;
    int scur_p1[W];
    while (iter < end) {
        ktimer_make_room(c, 4 * W + 64);
        const bool rebuild_first = rb == iter || host_rebuild(iter);
        if (rebuild_first) {
            if (rb == iter && c->rb_last >= 0) c->rb_gap = iter - c->rb_last;
            c->rb_last = iter;
        }
        // With the displacement criterion the window ends one step after the step at which the
        // check is expected to fire (the last rebuild plus the last interval between a rebuild and
        // a fired check): launches queued after a fired check only cost their launch, and a
        // window that ends before the check fires costs one more synchronisation.
        int wmax = W;
        if (disp && c->rb_gap > 0 && c->rb_last >= 0) wmax = std::min(W, std::max(3, c->rb_last + c->rb_gap + 2 - iter));
        int wend = iter + 1;
        while (wend < end && wend - iter < wmax && !host_rebuild(wend)) ++wend;
        int lists = -1, scur_lists = 0;
        bool fused_prev = p1done;
        bool ringed[W];
        int want = macros ? n_macros - *nm : 0;  // record sums still wanted
        for (int k = iter; k < wend; ++k) {
            c->kt_step = k;
            if (k == iter && rebuild_first) {
                if (rb != iter && !p1done) {
                    c->moved_since_bin = true;
                    TRY(launch_pass1(c));
                }
                scur_p1[0] = c->scur;
                c->rdisp_valid = false;
                TRY(rebuild_lists_for(c, iter + 1 < end && host_rebuild(iter + 1), true));
                ++*nrb;
                lists = iter;
                scur_lists = c->scur;
            } else if (!fused_prev) {
                c->moved_since_bin = true;
                if (disp) TRY(launch_pass1_check_disp(c, true, k));
                else TRY(launch_pass1(c, lists >= 0));
                scur_p1[k - iter] = c->scur;
            } else {
                scur_p1[k - iter] = c->scur;  // pass 1 of k ran inside the force launch of k-1
            }
            // (the window's last step too, unless it ends the run: its pass 1 opens the next window)
            const int next = (disp && k + 1 < end) ? k + 1 : -1;
            TRY(launch_forces(c, k, true, disp || lists >= 0, next, &fused_prev));
            if (fused_prev) c->moved_since_bin = true;
            ringed[k - iter] = want > 0 && (k % c->p.tplot) == 0;
            if (ringed[k - iter]) {
                // a fused launch left this step's final velocities in the buffer it read
                TRY(launch_macros(c, fused_prev ? 1 - c->scur : c->scur, c->d_ring + 8 * (k - iter)));
                --want;
            }
        }
        FM_HIP(hipMemcpyAsync(c->h_ring, c->d_ring, 8 * (size_t)(wend - iter) * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        TRY(sync_flags(c));
        if (c->h_flags[FLAG_NEED]) {
            if (lists < 0) return fm_set_error(FM_ESTATE, "capacity request without a queued list build");
            ktimer_drop(c, KT_NLIST, lists);
            ktimer_drop(c, KT_FORCES, lists);
            ktimer_drop(c, KT_FORCES_P1, lists);
            ktimer_drop(c, KT_PASS1, lists + 1);
            // the launches after the build did nothing: the state is that of the binning, and the
            // displacement sums start again after the rebuild
            c->kt_step = lists;
            c->scur = scur_lists;
            c->moved_since_bin = false;
            c->rdisp_valid = false;
            TRY(redo_lists(c));
            TRY(launch_forces(c, lists, true));
            TRY(finish_step(c, lists, macros, n_macros, nm));
            iter = lists + 1;
            p1done = false;
            continue;
        }
        if (disp && c->h_flags[FLAG_REBUILD] && c->h_flags[FLAG_RB_STEP] == wend && fused_prev) {
            // the check fired in the pass 1 of step wend, inside the window's last force launch:
            // every step of the window completed, and the state after that pass 1 is current
            for (int k = iter; k < wend; ++k)
                TRY(finish_step(c, k, macros, n_macros, nm, ringed[k - iter] ? c->h_ring + 8 * (k - iter) : nullptr));
            rb = wend;
            iter = wend;
            p1done = true;
            continue;
        }
        p1done = false;
        if (disp && c->h_flags[FLAG_REBUILD]) {
            const int s = c->h_flags[FLAG_RB_STEP];
            if (s < iter || s >= wend || s == lists)
                return fm_set_error(FM_ESTATE, "rebuild requested at step %d outside %d..%d", s, iter, wend - 1);
            // (the rebuild's first launch clears FLAG_REBUILD and FLAG_HALT)
            ktimer_drop(c, KT_FORCES, s);
            ktimer_drop(c, KT_FORCES_P1, s);
            ktimer_drop(c, KT_PASS1, s + 1);
            c->scur = scur_p1[s - iter];
            for (int k = iter; k < s; ++k)
                TRY(finish_step(c, k, macros, n_macros, nm, ringed[k - iter] ? c->h_ring + 8 * (k - iter) : nullptr));
            rb = s;
            iter = s;
        } else {
            for (int k = iter; k < wend; ++k)
                TRY(finish_step(c, k, macros, n_macros, nm, ringed[k - iter] ? c->h_ring + 8 * (k - iter) : nullptr));
            iter = wend;
            p1done = fused_prev;  // the last launch carried pass 1 of step wend
        }
    }
    return 0;
}

extern "C" int fm_run_vv(fm_ctx* c, int32_t iter0, int32_t nsteps, fm_macros* macros, int32_t n_macros,
                         int32_t* n_rebuilds) {
    CHECK_CTX(c);
    if (c->p.integration_algorithm != FM_INTEGRATION_VV)
        return fm_set_error(FM_ESTATE, "integration_algorithm is leap-frog: use fm_run_lfv");
    if (!c->nlist_valid) return fm_set_error(FM_ESTATE, "run before the initial neighbour list");
    if (c->p.mie_potential == 1 && !c->has_moltype) return fm_set_error(FM_ESTATE, "Mie potential needs moltype");
    if (c->p.potential_flag == 1 && !c->has_mono) return fm_set_error(FM_ESTATE, "potential_flag 1 needs monomer data");
    if (c->p.ensemble == FM_ENSEMBLE_TAG_MOVE && !c->has_tags) return fm_set_error(FM_ESTATE, "tag_move needs tags");
    int nrb = 0, nm = 0;
    if (c->p.rebuild_criteria == FM_REBUILD_DISPLACEMENT || c->p.rebuild_criteria == FM_REBUILD_FIXED) {
        TRY(run_vv_queued(c, iter0, nsteps, macros, n_macros, &nrb, &nm));
        TRY(sync_flags(c));
        if (n_rebuilds) *n_rebuilds = nrb;
        return 0;
    }
    const bool sync_each_force = c->p.ensemble == FM_ENSEMBLE_TAG_MOVE || c->p.potential_flag == 1;
    for (int iter = iter0; iter < iter0 + nsteps; ++iter) {
        const int forced = ((iter + 1) % c->p.tplot) == 0;
        int rb = forced;
        bool need_flag = false;
        switch (c->p.rebuild_criteria) {
        case FM_REBUILD_DISPLACEMENT:
            c->moved_since_bin = true;
            if (forced) {
                TRY(launch_pass1(c));
            } else {
                TRY(launch_pass1_check_disp(c));
                need_flag = true;
            }
            break;
        case FM_REBUILD_VMAX:
            c->moved_since_bin = true;
            TRY(launch_pass1(c));
            TRY(launch_check_vmax(c));
            need_flag = !forced;
            break;
        default:
            c->moved_since_bin = true;
            TRY(launch_pass1(c));
            if (iter % c->p.fixed_rebuild == 0) rb = 1;
            break;
        }
        if (need_flag) {
            TRY(sync_flags(c));
            if (c->h_flags[FLAG_REBUILD]) {
                rb = 1;
                TRY(launch_fill_i32(c, c->d_flags + FLAG_REBUILD, 0, 1));
            }
        }
        if (rb) {
            c->rdisp_valid = false;
            if (c->p.rebuild_criteria == FM_REBUILD_VMAX) {
                TRY(launch_fill_f64(c, c->d_rmax, 0.0, 1));
                TRY(launch_fill_i32(c, c->d_flags + FLAG_REBUILD, 0, 1));
            }
            TRY(rebuild_lists_for(c, iter + 1 < iter0 + nsteps && next_rebuild_certain(c, iter), false));
            ++nrb;
        }
        TRY(launch_forces(c, iter, true));
        const bool rec = (iter % c->p.tplot) == 0;
        const bool ev = c->p.teval > 0 && (iter % c->p.teval) == 0;
        c->pe_valid = rec;
        c->vir_valid = rec || ev;
        c->simtime = c->simtime + c->p.delta_t;
        if (sync_each_force) TRY(sync_flags(c));
        if (rec && macros && nm < n_macros) {
            TRY(fm_evaluate_macroscopic_properties(c, &macros[nm]));
            ++nm;
        }
    }
    TRY(sync_flags(c));
    if (n_rebuilds) *n_rebuilds = nrb;
    return 0;
}

// ---- fused md_advance_lfv loop (src/md.f90:109-139): forces, record, move, rebuild check ------
// md_advance_lfv without a host synchronisation per step, for domains whose rebuilds are all
// known in advance (TPLOT 1, where every step rebuilds for its record step, or the fixed interval)
// and have no bonds, sort_mols or velocity criterion.  Steps are queued in windows of up to
// FM_RING_SLOTS steps: force (conditional), the record sums into a ring slot, the move
// (conditional) and the rebuild, queued without host waits.  A list build beyond the capacities
// raises FLAG_NEED and FLAG_HALT; the force and move launches after it do nothing (the first
// halted force launch stores its step in FLAG_HALT_STEP), while the later rebuilds only reorder
// the unchanged state.  At the window's synchronisation the steps before the halted one are
// complete; the list is built again with grown capacities and the next window starts at the
// halted step.  Every completed step has the arithmetic of the step-by-step loop.
static bool lfv_rebuild_known(const fm_ctx* c, int it) {
    return ((it + 1) % c->p.tplot) == 0 || (c->p.rebuild_criteria == FM_REBUILD_FIXED && it % c->p.fixed_rebuild == 0);
}

static int run_lfv_queued(fm_ctx* c, int iter0, int nsteps, fm_macros* macros, int n_macros, int* nrb, int* nm) {
    const int end = iter0 + nsteps, W = FM_RING_SLOTS;
    static const int reset[] = {FLAG_REBUILD, FLAG_HALT, FLAG_NEED};
    TRY(launch_reset_flags(c, reset, 3));
    int iter = iter0;
    while (iter < end) {
        const int wend = std::min(end, iter + W);
        bool ringed[W], rebuilt[W];
        int want = macros ? n_macros - *nm : 0;
        for (int k = iter; k < wend; ++k) {
            c->kt_step = k;
            TRY(launch_forces(c, k, false, true));
            ringed[k - iter] = want > 0 && (k % c->p.tplot) == 0;
            if (ringed[k - iter]) {
                TRY(launch_macros(c, -1, c->d_ring + 8 * (k - iter)));
                --want;
            }
            c->moved_since_bin = true;
            TRY(launch_move_lfv(c, false, true));
            rebuilt[k - iter] = lfv_rebuild_known(c, k);
            if (rebuilt[k - iter]) {
                c->rdisp_valid = false;
                // the loop's last list serves whatever runs next: full radius, accelerations kept
                const bool last = k + 1 >= end;
                const bool skinless = !last && lfv_rebuild_known(c, k + 1) && c->p.potential_flag == 0 &&
                                      c->p.mie_potential == 0;
                c->list_ra2 = skinless ? c->rc2 * (1.0 + 1e-12) : 0.0;
                const int rc = rebuild_lists_queued(c, false, last);
                c->list_ra2 = 0.0;
                TRY(rc);
            }
        }
        FM_HIP(hipMemcpyAsync(c->h_ring, c->d_ring, 8 * (size_t)(wend - iter) * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        TRY(sync_flags(c));
        int done = wend;
        if (c->h_flags[FLAG_HALT]) {
            // a list build met a region or list beyond the capacities: the first force launch
            // after it halted (none did when it was the window's last rebuild)
            const int s = c->h_flags[FLAG_HALT_STEP];
            if (!c->h_flags[FLAG_NEED] || (s != INT_MAX && (s <= iter || s >= wend)))
                return fm_set_error(FM_ESTATE, "queued leap-frog halted at step %d outside %d..%d", s, iter + 1, wend - 1);
            // steps iter .. s-1 are complete; the cells are those of the last rebuild, the state
            // unchanged since step s-1
            TRY(redo_lists(c));
            if (s != INT_MAX) done = s;
        }
        for (int k = iter; k < done; ++k) {
            if (rebuilt[k - iter]) ++*nrb;
            TRY(finish_step(c, k, macros, n_macros, nm, ringed[k - iter] ? c->h_ring + 8 * (k - iter) : nullptr));
        }
        iter = done;
    }
    return 0;
}

extern "C" int fm_run_lfv(fm_ctx* c, int32_t iter0, int32_t nsteps, fm_macros* macros, int32_t n_macros,
                          int32_t* n_rebuilds) {
    CHECK_CTX(c);
    if (c->p.integration_algorithm != FM_INTEGRATION_LFV)
        return fm_set_error(FM_ESTATE, "integration_algorithm is velocity Verlet: use fm_run_vv");
    if (!c->nlist_valid) return fm_set_error(FM_ESTATE, "run before the initial neighbour list");
    if (c->p.mie_potential == 1 && !c->has_moltype) return fm_set_error(FM_ESTATE, "Mie potential needs moltype");
    if (c->p.potential_flag == 1 && !c->has_mono) return fm_set_error(FM_ESTATE, "potential_flag 1 needs monomer data");
    if (c->p.ensemble == FM_ENSEMBLE_TAG_MOVE && !c->has_tags) return fm_set_error(FM_ESTATE, "tag_move needs tags");
    const bool sync_each_force = c->p.potential_flag == 1;
    int nrb = 0, nm = 0;
    const bool known = c->p.tplot == 1 || c->p.rebuild_criteria == FM_REBUILD_FIXED;
    if (known && !sync_each_force && c->p.sort_flag == 0 && c->p.rebuild_criteria != FM_REBUILD_VMAX &&
        c->maxreg > 0 && nsteps > 0) {
        TRY(run_lfv_queued(c, iter0, nsteps, macros, n_macros, &nrb, &nm));
        TRY(sync_flags(c));
        if (n_rebuilds) *n_rebuilds = nrb;
        return 0;
    }
    for (int iter = iter0; iter < iter0 + nsteps; ++iter) {
        TRY(launch_forces(c, iter, false));
        const bool rec = (iter % c->p.tplot) == 0;
        const bool ev = c->p.teval > 0 && (iter % c->p.teval) == 0;
        c->pe_valid = rec;
        c->vir_valid = rec || ev;
        if (sync_each_force) TRY(sync_flags(c));
        if (rec && macros && nm < n_macros) {
            TRY(fm_evaluate_macroscopic_properties(c, &macros[nm]));
            ++nm;
        }
        const int forced = ((iter + 1) % c->p.tplot) == 0;
        int rb = forced;
        bool need_flag = false;
        c->moved_since_bin = true;
        switch (c->p.rebuild_criteria) {
        case FM_REBUILD_DISPLACEMENT:
            TRY(launch_move_lfv(c, !forced));
            need_flag = !forced;
            break;
        case FM_REBUILD_VMAX:
            TRY(launch_move_lfv(c, false));
            TRY(launch_check_vmax(c));
            need_flag = !forced;
            break;
        default:
            TRY(launch_move_lfv(c, false));
            if (iter % c->p.fixed_rebuild == 0) rb = 1;
            break;
        }
        c->simtime = c->simtime + c->p.delta_t;
        if (need_flag || c->p.ensemble == FM_ENSEMBLE_TAG_MOVE) {
            TRY(sync_flags(c));
            if (c->h_flags[FLAG_REBUILD]) {
                if (need_flag) rb = 1;
                TRY(launch_fill_i32(c, c->d_flags + FLAG_REBUILD, 0, 1));
            }
        }
        if (rb) {
            c->rdisp_valid = false;
            if (c->p.rebuild_criteria == FM_REBUILD_VMAX) {
                TRY(launch_fill_f64(c, c->d_rmax, 0.0, 1));
                TRY(launch_fill_i32(c, c->d_flags + FLAG_REBUILD, 0, 1));
            }
            TRY(rebuild_lists_for(c, iter + 1 < iter0 + nsteps && next_rebuild_certain(c, iter), false));
            ++nrb;
        }
    }
    TRY(sync_flags(c));
    if (n_rebuilds) *n_rebuilds = nrb;
    return 0;
}

// ---- inspection -------------------------------------------------------------------------------
extern "C" int fm_download_cells(fm_ctx* c, int32_t* cell_of_mol) {
    CHECK_CTX(c);
    if (!c->cells_valid) return fm_set_error(FM_ESTATE, "no cell lists");
    std::vector<int> cs(c->np);
    TRY(launch_slot_to_ext_i32(c, (int*)c->stage, c->cell_of, 1));
    FM_HIP(hipMemcpyAsync(cs.data(), c->stage, c->np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipStreamSynchronize(c->stream));
    const int nx = c->nc[0], ny = c->nc[1];
    for (int n = 0; n < c->np; ++n) {
        int ci = cs[n];
        int cx = ci % nx, cy = (ci / nx) % ny, cz = ci / (nx * ny);
        cell_of_mol[n] = ((cz + 1) * (ny + 2) + (cy + 1)) * (nx + 2) + (cx + 1);
    }
    return 0;
}

extern "C" int fm_download_neighbourlist(fm_ctx* c, int32_t* mol_i, int32_t* mol_j, int32_t* shift, int64_t capacity,
                                         int64_t* n_entries) {
    CHECK_CTX(c);
    if (!c->nlist_valid) return fm_set_error(FM_ESTATE, "no neighbour list");
    const size_t cap = c->cap, np = c->np;
    std::vector<int> nbn(np), ext(np);
    FM_HIP(hipMemcpyAsync(nbn.data(), c->nbn, np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipMemcpyAsync(ext.data(), c->ext[c->cur], np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipStreamSynchronize(c->stream));
    int64_t total = 0;
    for (size_t i = 0; i < np; ++i) total += nbn[i];
    if (n_entries) *n_entries = total;
    if (!mol_i) return 0;
    if (capacity < total) return fm_set_error(FM_EINVAL, "capacity %lld < %lld entries", (long long)capacity, (long long)total);
    std::vector<uint32_t> row(np);
    std::vector<int64_t> off(np + 1, 0);
    for (size_t i = 0; i < np; ++i) off[i + 1] = off[i] + nbn[i];
    uint32_t* dec = nullptr;
    FM_HIP(hipMalloc((void**)&dec, (size_t)c->maxnb * cap * sizeof(uint32_t)));
    struct Free {
        uint32_t* p;
        ~Free() { hipFree(p); }
    } guard{dec};
    TRY(launch_decode_nlist(c, dec));
    FM_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->maxnb; ++k) {
        FM_HIP(hipMemcpy(row.data(), dec + (size_t)k * cap, np * sizeof(uint32_t), hipMemcpyDeviceToHost));
        bool any = false;
        for (size_t i = 0; i < np; ++i) {
            if (k >= nbn[i]) continue;
            any = true;
            int64_t q = off[i] + k;
            mol_i[q] = ext[i];
            mol_j[q] = ext[row[i] >> 5];
            shift[q] = (int)(row[i] & 31u);
        }
        if (!any) break;
    }
    return 0;
}

extern "C" int fm_download_bonds(fm_ctx* c, int32_t* bond, int32_t* bond_shift) {
    CHECK_CTX(c);
    if (!c->bond || !c->nlist_valid) return fm_set_error(FM_ESTATE, "no bond lists");
    const size_t cap = c->cap, np = c->np;
    std::vector<int> bn(np), ext(np);
    std::vector<uint32_t> rows(FM_MAX_FUNCY * cap);
    FM_HIP(hipMemcpyAsync(bn.data(), c->bondn, np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipMemcpyAsync(ext.data(), c->ext[c->cur], np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipMemcpyAsync(rows.data(), c->bond, FM_MAX_FUNCY * cap * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    FM_HIP(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < np; ++i) {
        size_t e = ext[i];
        for (int b = 0; b < FM_MAX_FUNCY; ++b) {
            if (b < bn[i]) {
                uint32_t w = rows[(size_t)b * cap + i];
                bond[FM_MAX_FUNCY * e + b] = ext[w >> 5];
                bond_shift[FM_MAX_FUNCY * e + b] = (int)(w & 31u);
            } else {
                bond[FM_MAX_FUNCY * e + b] = -1;
                bond_shift[FM_MAX_FUNCY * e + b] = 13;
            }
        }
    }
    return 0;
}

extern "C" int fm_synchronize(fm_ctx* c) {
    CHECK_CTX(c);
    FM_HIP(hipStreamSynchronize(c->stream));
    return sync_flags(c);
}

extern "C" int fm_kernel_timing(fm_ctx* c, int32_t enable) {
    CHECK_CTX(c);
    kt_collect(c);
    c->timer->on = enable != 0;
    for (int k = 0; k < KT_NCLASS; ++k) {
        c->timer->total_ms[k] = 0.0;
        c->timer->count[k] = 0;
    }
    return 0;
}

extern "C" int fm_get_kernel_timing(fm_ctx* c, char* names, int32_t names_cap, double* avg_ms, int64_t* counts,
                                    int32_t max_classes, int32_t* n_classes) {
    CHECK_CTX(c);
    kt_collect(c);
    std::string s;
    int n = std::min<int>(KT_NCLASS, max_classes);
    for (int k = 0; k < n; ++k) {
        s += kt_names[k];
        s += '\n';
        if (avg_ms) avg_ms[k] = c->timer->count[k] ? c->timer->total_ms[k] / c->timer->count[k] : 0.0;
        if (counts) counts[k] = c->timer->count[k];
    }
    if (names && names_cap > 0) {
        strncpy(names, s.c_str(), names_cap - 1);
        names[names_cap - 1] = 0;
    }
    if (n_classes) *n_classes = n;
    return 0;
}
