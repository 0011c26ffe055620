// fm_internal.h -- device context and kernel launchers of libflowmol_hip (gfx950).
//
// Data layout in HBM (one context = one periodic domain):
//   * molecule state in "slot" order, structure of arrays: x,y,z, vx,vy,vz, ax,ay,az (FP64).
//     Slots are re-ordered by cell at every assign_to_cell (counting sort), so a cell's molecules
//     are contiguous; inside a cell they are ordered by decreasing molecule number, the order of
//     the reference's head-inserted cell lists.  ext[slot] is the host-visible molecule number
//     (0-based), slot_of[molno] its inverse.
//   * cells: interior cells only (ncx*ncy*ncz), x fastest; cell_start[c]..cell_start[c+1].
//   * neighbour list: ELL, entry k of slot i at nbl[k*cap + i] (coalesced across a wavefront),
//     word = (j_slot << 5) | shift_code, shift_code = (sx+1)*9+(sy+1)*3+(sz+1): the image of j is
//     r_j + s*domain, formed component by component exactly as the reference's halo copy.
//   * bonds: bond[b*cap + i] same encoding, FM_MAX_FUNCY per slot.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/flowmol_hip.h"

#define FM_BLOCK 256
// workgroup size of the brick kernels (neighbour list, forces): one workgroup per brick, so a
// larger workgroup shares one staged region among more waves (less LDS per wave, less halo per
// molecule).  The brick shape is chosen so that a brick holds about this many molecules.
#define FB_BLOCK 512
// record steps of one queued window of the fused loop whose sums stay on the device until the
// window's synchronisation (the window is at most this many steps)
#define FM_RING_SLOTS 16

// device flag words (d_flags)
enum {
    FLAG_REBUILD = 0,     // displacement / vmax criterion fired
    FLAG_ERROR = 1,       // first FM_E* error code
    FLAG_ERROR_MOL = 2,   // molecule number of the first error
    FLAG_MAX_CELL = 3,    // (not tracked)
    FLAG_MAX_NB = 4,      // largest neighbour count at the last build
    FLAG_NOUT = 5,        // molecules outside the box at sendmols
    FLAG_EDGE = 6,        // molecules exactly on the lower domain face (clamped into cell 1)
    FLAG_MAX_REGION = 7,  // largest LDS region (brick + one-cell halo) in molecules
    FLAG_HALT = 8,        // a rebuild is pending: conditional launches of later steps do nothing
    FLAG_RB_STEP = 9,     // step whose displacement check set FLAG_REBUILD
    FLAG_SORT_MAX = 10,   // largest sort block (sort_mols)
    FLAG_NEED = 11,       // capacity requests of queued list builds: NEED_REGION, NEED_LIST bits
    FLAG_HALT_STEP = 12,  // the first step whose conditional force launch found a halt (INT_MAX: none)
    FLAG_COUNT = 13
};
enum { NEED_REGION = 1, NEED_LIST = 2 };

struct KTimer;

struct fm_ctx {
    fm_params p;
    int np = 0, cap = 0;
    int dev = 0;
    hipStream_t stream = nullptr;

    // geometry (set_parameters_cells)
    int nc[3] = {0, 0, 0};
    int ncell = 0;
    double half[3], cs[3], rn2, rc2, potshift, dmin;
    // > 0: the next list build keeps only the pairs with r^2 below this (a list that one force
    // call reads at the build's positions needs no skin); 0: the list radius
    double list_ra2 = 0.0;

    // state, slot order (two copies for the permutation at rebuild)
    double* st[2] = {nullptr, nullptr};   // each: 9*cap doubles: x y z vx vy vz ax ay az
    int cur = 0;
    double* rdisp = nullptr;              // 3*cap
    int* ext[2] = {nullptr, nullptr};
    int* slot_of = nullptr;
    int* moltype[2] = {nullptr, nullptr};
    int* tag[2] = {nullptr, nullptr};
    double* teth[2] = {nullptr, nullptr}; // 6*cap: rtether xyz, slidev xyz
    int* mono[2] = {nullptr, nullptr};    // 8*cap: chainID subchainID funcy bflag[4] glob_no
    bool has_moltype = false, has_tags = false, has_mono = false;

    // cells
    int* cell_of = nullptr;               // per slot: interior cell index
    int* cell_count = nullptr;            // ncell
    int* cell_start = nullptr;            // ncell+1
    int* bin_pos = nullptr;               // cap: each molecule's place among its cell's (binning)
    int* scan_part = nullptr;             // cell scan: count sums of the 1024-cell chunks
    int* perm = nullptr;                  // cap

    // LDS bricks
    int brick[3] = {1, 1, 1};
    int* brick_order = nullptr;           // bricks in 4x4x4-tile order (XCD-aware workgroup mapping)
    int kt_step = 0;                      // step number attached to kernel timing records

    // sort_mols: block curve, the routine's saved counter and safety factors
    int* hcurve = nullptr;                // nblocks^3, 1-based curve positions, x fastest
    int sort_nb[3] = {0, 0, 0};
    double sort_bsl[3] = {0.0, 0.0, 0.0};
    int sort_nrebuilds = 1;
    double sort_sf1 = 1.5;
    int sort_sf2 = 5;
    int *sort_key = nullptr, *sort_val = nullptr, *sort_key2 = nullptr, *sort_val2 = nullptr, *sort_cnt = nullptr;
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    // LDS capacity of a staged region (molecules): at least the largest region of the last
    // measured binning, raised to what the LDS of as many workgroups per CU still holds, so that
    // queued rebuilds run without a measurement; a kernel meeting a larger region raises
    // NEED_REGION and FLAG_HALT instead of staging it
    int maxreg = 0;
    size_t lds_limit = 65536;             // dynamic LDS per workgroup
    size_t lds_cu = 65536;                // LDS per compute unit
    bool tight_regions = false;           // capacity = measured largest region (no headroom)
    // per brick, written by the quantised list build for the plain force launches: region total
    // and interior count, each staged molecule's slot << 5 | image shift (maxreg), each interior
    // molecule's slot (maxreg) and LDS index (maxreg 16-bit); valid until the next binning
    int* btab = nullptr;
    int btab_stride = 0, btab_reg = 0, btab_nbrick = 0;
    bool btab_valid = false;

    // neighbour list: 16-bit indices into the brick's LDS region layout, four per 64-bit word:
    // entries 4c..4c+3 of slot i in nbl[c*cap + i]
    int maxnb = 0;
    uint64_t* nbl = nullptr;
    int* nbn = nullptr;
    uint32_t* bond = nullptr;             // FM_MAX_FUNCY*cap
    int* bondn = nullptr;

    // per-molecule tallies
    double* pe_lj = nullptr;
    double* pe_poly = nullptr;
    double* vir = nullptr;
    double* rf = nullptr;                 // 9*cap
    double* pxy = nullptr;                // 3*nplanes
    double* cvface = nullptr;             // faces 1..3 of Pxyface: nbinso^3 x 3 components x 3 faces
    int cv_nbins[3] = {0, 0, 0}, cv_nhb[3] = {0, 0, 0}, cv_nbo[3] = {0, 0, 0};
    double cv_bs[3] = {0.0, 0.0, 0.0};
    double* planes = nullptr;
    double mop_shift = 0.0;

    // sendmols scratch
    int* out_flag = nullptr;              // cap
    int* out_vals = nullptr;              // cap
    int* out_slot = nullptr;              // cap
    int* out_count = nullptr;             // per-workgroup leaver counts

    // reductions
    double* red = nullptr;                // partial sums
    int nred_blocks = 0;
    double* d_macros = nullptr;           // 16 doubles
    double* d_ring = nullptr;             // FM_RING_SLOTS x 8 record-step sums of a queued window
    double* h_ring = nullptr;             // pinned copy of d_ring, read at the window's synchronisation
    double* d_mie = nullptr;              // eps, lr, la, C, alpha, potshift, rc2 tables + mass[4]
    unsigned long long* d_vmax = nullptr; // bits of max |v|^2
    double* d_rmax = nullptr;

    int* d_flags = nullptr;
    int* h_flags = nullptr;               // pinned
    double* stage = nullptr;              // 9*cap doubles: host-order staging
    bool pe_valid = false, vir_valid = false;


    bool cells_valid = false, nlist_valid = false, rdisp_valid = false;
    bool forces_valid = false;
    bool moved_since_bin = true;          // cells are x-sorted only right after binning
    double simtime = 0.0;
    int last_record_iter = -1;

    KTimer* timer = nullptr;

    int scur = 0;                         // state buffer in use (st[scur]); ext & co. follow cur
    int rb_last = -1, rb_gap = 0;         // fused VV loop: last rebuild step, last rebuild-to-check interval
    double* x() const { return st[scur]; }
};

// ---- error handling -----------------------------------------------------------------------------
int fm_set_error(int code, const char* fmt, ...);
#define FM_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess)                                                          \
            return fm_set_error(FM_EHIP, "%s failed: %s (%s:%d)", #call,               \
                                hipGetErrorString(_e), __FILE__, __LINE__);            \
    } while (0)

#define TRY(x)               \
    do {                     \
        int _rc = (x);       \
        if (_rc) return _rc; \
    } while (0)

// ---- kernel timing (HIP events on the context stream) -------------------------------------------
void ktimer_begin(fm_ctx* c, int cls);
void ktimer_end(fm_ctx* c, int cls);
enum { KT_PASS1 = 0, KT_CHECK, KT_SENDMOLS, KT_BIN, KT_NLIST, KT_FORCES, KT_PASS2, KT_MACROS, KT_FORCES_P1, KT_NCLASS };
extern const char* const kt_names[KT_NCLASS];

// ---- launchers (fm_kernels.hip, fm_brick.hip) ----------------------------------------------------
int launch_fill_i32(fm_ctx* c, int* p, int v, size_t n);
int launch_fill_f64(fm_ctx* c, double* p, double v, size_t n);
int launch_pass1(fm_ctx* c, bool cond = false);  // cond: do nothing while FLAG_HALT is set
int launch_pass2(fm_ctx* c);
int launch_check_disp(fm_ctx* c);
int launch_check_vmax(fm_ctx* c);
int launch_sendmols(fm_ctx* c, const int* reset = nullptr, int nreset = 0);  // also zeroes the listed flags
// binning (counting sort into cells, x order per cell, permutation of the state); keep_acc false
// leaves the accelerations out of the permutation (the next force launch overwrites them all)
int launch_bin(fm_ctx* c, bool keep_acc = true);
int launch_reset_flags(fm_ctx* c, const int* which, int n);  // zero the listed flag words
int launch_nlist(fm_ctx* c);
// cond: do nothing when a rebuild is pending (FLAG_REBUILD or FLAG_HALT), and raise FLAG_HALT.
// next >= 0 with fused != null: on a plain LJ step with the kick, also apply VV pass 1 of step
// `next` and its displacement check in the same launch, into the other state buffer (c->scur
// flips); *fused reports whether that happened.
int launch_forces(fm_ctx* c, int iter, bool kick = false, bool cond = false, int next = -1, bool* fused = nullptr);
int launch_pass1_check_disp(fm_ctx* c, bool cond = false, int step = 0);
int launch_move_lfv(fm_ctx* c, bool check_disp, bool cond = false);  // cond: nothing while FLAG_HALT is set
// record-step sums over state buffer st[buf] (buf < 0: the current one) and the per-molecule
// tallies into out[8] (null: d_macros)
int launch_macros(fm_ctx* c, int buf = -1, double* out = nullptr);
int launch_scatter_state_to_host_order(fm_ctx* c, double* d_out9, double* d_tally);
int launch_from_host(fm_ctx* c, const double* d_in9);
int launch_ext_to_slot_i32(fm_ctx* c, int* dst, const int* src, int nf);
int launch_ext_to_slot_f64(fm_ctx* c, double* dst, const double* src, int nf);
int launch_slot_to_ext_f64(fm_ctx* c, double* dst, const double* src, int nf);
int launch_slot_to_ext_i32(fm_ctx* c, int* dst, const int* src, int nf);

// packed launches over many domains (replica batches)
struct P1Desc {               // VV pass 1 (+ displacement check) of one domain
    double* st;               // state block (9*cap doubles)
    double* rd;               // rdisp (3*cap)
    int* flags;               // the domain's flag words (halt rule, FLAG_REBUILD / FLAG_RB_STEP)
    int np, cap;
    int check, fresh;         // apply the displacement update / start it from zero
    int active, step;         // the domain has a step in this launch; its number
    double dt, hdt2, hdt, thresh;
};
int launch_pass1_multi(hipStream_t st, const P1Desc* d_descs, const int2* d_table, int nblk);
size_t fargs_bytes();
void fargs_set_queued(void* fargs, int* flags, bool active);
// force-kernel arguments of one domain for step iter; kind: bit0 mixed, bit1 Mie, bit2 bonds/tethers
int make_fargs(fm_ctx* c, int iter, bool kick, void* out, size_t* lds, int* kind);
// plain: no domain of the launch is on a record or virial step
int launch_forces_multi(hipStream_t st, int kind, const void* d_descs, const int2* d_table, int nblk, size_t lds,
                        bool plain);
bool fargs_plain(const void* fargs);
int rebuild_lists_internal(fm_ctx* c);
int rebuild_lists_queued_internal(fm_ctx* c);  // the rebuild without host waits (capacity requests in FLAG_NEED)
int redo_lists_internal(fm_ctx* c);            // after FLAG_NEED: grow the capacities, build the list again
int check_flags_internal(fm_ctx* c);           // report a device-side error in the flag words read last

// LDS-brick kernels (fm_brick.hip)
int brick_count(const fm_ctx* c);
size_t brick_lds_bytes(const fm_ctx* c, int maxreg, int kind);  // kind 0 list build, 1 forces, 2 MOP, 3 Mie plain
int launch_region_max(fm_ctx* c);
int launch_decode_nlist(fm_ctx* c, uint32_t* out);

// stable radix sort of (key, value) pairs (fm_kernels.hip)
size_t radix_sort_bytes(int n_items);
int radix_sort_pairs(fm_ctx* c, int* k0, int* v0, int* k1, int* v1, int n, int end_bit, int** vals);
int launch_sort_mols(fm_ctx* c, int* overflow_out);
int ctx_sync_flags(fm_ctx* c);  // read the flag words, report a device-side error
