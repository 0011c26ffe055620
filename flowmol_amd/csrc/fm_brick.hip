// fm_brick.hip -- LDS-staged cell-brick kernels: neighbour-list build, pair/bond forces with the
// energy/virial/MOP tallies, and list decoding for inspection.
//
// A brick is Bx*By*Bz interior cells; its region is the brick plus one cell on every side
// ((Bx+2)*(By+2)*(Bz+2) cells).  One workgroup stages the region's molecule positions into LDS,
// cell by cell in region order, with each cell's periodic image shift applied while staging
// (x + s*L, one add per shifted component, as the reference's halo copy forms its images).
// Molecules are contiguous per cell in slot order, so staging reads are coalesced.
//
// The neighbour list of molecule i holds 16-bit indices into its brick's LDS region layout.
// Between rebuilds the cell contents and the slot order do not change, so the force kernel
// re-stages the identical layout and the indices stay valid.
#include "fm_internal.h"

#include <limits.h>
#include <math.h>
#include <string.h>

#include <type_traits>

#pragma clang fp contract(off)

// doubles in the Mie tables: 7 fields per type pair, then the mass of each type
#define FM_MIE_WORDS (7 * FM_MAX_TYPES * FM_MAX_TYPES + FM_MAX_TYPES)

namespace {

struct St {
    double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az;
};

__host__ __device__ inline St state_of(double* base, int cap) {
    St s;
    s.x = base;
    s.y = base + (size_t)cap;
    s.z = base + 2 * (size_t)cap;
    s.vx = base + 3 * (size_t)cap;
    s.vy = base + 4 * (size_t)cap;
    s.vz = base + 5 * (size_t)cap;
    s.ax = base + 6 * (size_t)cap;
    s.ay = base + 7 * (size_t)cap;
    s.az = base + 8 * (size_t)cap;
    return s;
}

__device__ inline void set_err(int* flags, int code, int mol) {
    if (atomicCAS(&flags[FLAG_ERROR], 0, code) == 0) flags[FLAG_ERROR_MOL] = mol;
}

// 1/x from v_rcp_f64 refined by two Newton steps (within an ulp of the IEEE quotient; the
// pair law needs 1e-12 relative, not the reference's last bit)
__device__ inline double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ inline double fmaf64(double a, double b, double c) { return __builtin_fma(a, b, c); }
// one Newton step: relative error below 3e-15 (v_rcp_f64 alone is good to about 5e-8)
__device__ inline double recip1(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}

// x^e for the Mie exponents (lambda/2 and lambda/2 + 1): a whole e in 0..31 by binary powers,
// branch-free (within a few ulp of pow, far inside the pair law's 1e-12), any other e by pow
__device__ inline double mie_pow(double x, double e) {
    const int n = (int)e;
    if ((double)n == e && n >= 0 && n < 32) {
        double r = 1.0, b = x;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            r = ((n >> k) & 1) ? r * b : r;
            b = b * b;
        }
        return r;
    }
    return pow(x, e);
}

__device__ inline double p3(double x) { return x * (x * x); }
__device__ inline double p4(double x) { double x2 = x * x; return x2 * x2; }
__device__ inline double p6(double x) { double x3 = x * (x * x); return x3 * x3; }
__device__ inline double p7(double x) { double x2 = x * x; return (x * x2) * (x2 * x2); }

__device__ inline int shift_code(int sx, int sy, int sz) { return (sx + 1) * 9 + (sy + 1) * 3 + (sz + 1); }
__device__ inline void decode_shift(int code, int& sx, int& sy, int& sz) {
    sx = code / 9 - 1;
    sy = (code / 3) % 3 - 1;
    sz = code % 3 - 1;
}
__device__ inline double img(double r, int s, double L) { return s == 0 ? r : r + (double)s * L; }

// ---- region metadata and staging ---------------------------------------------------------------
struct Region {
    // LDS arrays
    int* off;      // nreg+1: prefix of molecule counts (region layout)
    int* gc;       // nreg: global cell
    int* sh;       // nreg: shift code of the region cell
    int* ic_off;   // nic+1: prefix over interior (i-) cells
    int* ic_rc;    // nic: region cell of each i-cell
    int* cst;      // nreg: first slot of each region cell's global cell
    uint16_t* rq;  // staged molecule -> region cell
    double* pos;   // staged positions, x y z interleaved (one address, three reads)
    int* lslot;    // optional: slot of each staged molecule (the force kernels: its molecule number)
    int* ltype;    // optional: molecule type (Mie tables) of each staged molecule
    uint8_t* ltype8 = nullptr;  // optional: the same in one byte (the branch-free Mie launch)
    int nic;
    int total;     // staged molecules
    int ni;        // i-molecules
};

struct BG {  // brick geometry
    int nc[3], B[3], nb[3], R[3], nreg;
    double L[3];
};

// Carve the LDS buffer (8-byte aligned pieces first).
__device__ inline void carve(Region& r, char* smem, const BG& g, int maxreg, bool with_slot,
                           bool with_type = false) {
    double* d = (double*)smem;
    r.pos = d;
    int* ip = (int*)(d + 3 * maxreg);
    r.lslot = nullptr;
    if (with_slot) {
        r.lslot = ip;
        ip += maxreg;
    }
    r.ltype = nullptr;
    if (with_type) {
        r.ltype = ip;
        ip += maxreg;
    }
    r.rq = (uint16_t*)ip;
    ip += (maxreg + 1) / 2;
    r.off = ip;
    ip += g.nreg + 1;
    r.gc = ip;
    ip += g.nreg;
    r.sh = ip;
    ip += g.nreg;
    r.cst = ip;
    ip += g.nreg;
    const int nbmax = g.B[0] * g.B[1] * g.B[2];
    r.ic_off = ip;
    ip += nbmax + 1;
    r.ic_rc = ip;
}

// Bytes of the region tables, rounded up to 16 so that an array of doubles placed after them
// (the MOP plane tallies, updated with 64-bit LDS atomics) is naturally aligned.
__host__ __device__ inline size_t region_lds_bytes(const BG& g, int maxreg, bool with_slot, bool with_type = false) {
    const int nbmax = g.B[0] * g.B[1] * g.B[2];
    const size_t b = (size_t)maxreg * (3 * sizeof(double) + (with_slot ? sizeof(int) : 0) + (with_type ? sizeof(int) : 0)) +
                     (size_t)((maxreg + 1) / 2 + 4 * g.nreg + 1 + 2 * nbmax + 1) * sizeof(int) + 16;
    return (b + 15) & ~(size_t)15;
}

// exclusive scan of cnt[0..n) into off[0..n] by the first wavefront
__device__ inline void wave0_scan(int* off, int n) {
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int carry = 0;
        for (int base = 0; base < n; base += 64) {
            int q = base + lane;
            int v = q < n ? off[q] : 0;
            int x = v;
            for (int d = 1; d < 64; d <<= 1) {
                int y = __shfl_up(x, d, 64);
                if (lane >= d) x += y;
            }
            if (q < n) off[q] = carry + x - v;
            carry += __shfl(x, 63, 64);
        }
        if (lane == 0) off[n] = carry;
    }
}

// The region tables of a brick: global cell, image shift and molecule offset of each region cell,
// and the interior-cell list.  A region of more than capm molecules is not staged: the workgroup
// raises NEED_REGION (and the size in FLAG_MAX_REGION) and FLAG_HALT, and the call returns false
// on every thread.
__device__ bool stage_tables(Region& r, const BG& g, int brick, const int* start, int capm, int* flags) {
    const int bx = brick % g.nb[0], by = (brick / g.nb[0]) % g.nb[1], bz = brick / (g.nb[0] * g.nb[1]);
    const int org[3] = {bx * g.B[0], by * g.B[1], bz * g.B[2]};
    for (int t = threadIdx.x; t < g.nreg; t += blockDim.x) {
        const int rr[3] = {t % g.R[0], (t / g.R[0]) % g.R[1], t / (g.R[0] * g.R[1])};
        int gg[3], ss[3];
        for (int d = 0; d < 3; ++d) {
            int v = org[d] + rr[d] - 1, sv = 0;
            if (v < 0) {
                v += g.nc[d];
                sv = -1;
            } else if (v >= g.nc[d]) {
                v -= g.nc[d];
                sv = 1;
            }
            gg[d] = v;
            ss[d] = sv;
        }
        const int c = gg[0] + g.nc[0] * (gg[1] + g.nc[1] * gg[2]);
        r.gc[t] = c;
        r.sh[t] = shift_code(ss[0], ss[1], ss[2]);
        const int s0 = start[c];
        r.cst[t] = s0;
        r.off[t] = start[c + 1] - s0;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        wave0_scan(r.off, g.nreg);
        // interior cells of this brick (partial bricks at the far faces), x fastest, and the
        // prefix of their molecule counts, by the same wavefront (its LDS accesses stay in order)
        const int lane = threadIdx.x;
        const int ex = min(g.B[0], g.nc[0] - org[0]), ey = min(g.B[1], g.nc[1] - org[1]),
                  ez = min(g.B[2], g.nc[2] - org[2]);
        const int nic = ex * ey * ez;
        int carry = 0;
        for (int base = 0; base < nic; base += 64) {
            const int k = base + lane;
            int cnt = 0;
            if (k < nic) {
                const int x = k % ex + 1, y = (k / ex) % ey + 1, z = k / (ex * ey) + 1;
                const int rc = x + g.R[0] * (y + g.R[1] * z);
                r.ic_rc[k] = rc;
                cnt = r.off[rc + 1] - r.off[rc];
            }
            int xs = cnt;
            for (int d = 1; d < 64; d <<= 1) {
                const int y2 = __shfl_up(xs, d, 64);
                if (lane >= d) xs += y2;
            }
            if (k < nic) r.ic_off[k] = carry + xs - cnt;
            carry += __shfl(xs, 63, 64);
        }
        if (lane == 0) r.ic_off[nic] = carry;
    }
    __syncthreads();
    r.total = r.off[g.nreg];
    r.nic = min(g.B[0], g.nc[0] - org[0]) * min(g.B[1], g.nc[1] - org[1]) * min(g.B[2], g.nc[2] - org[2]);
    r.ni = r.ic_off[r.nic];
    if (r.total > capm) {
        if (threadIdx.x == 0 && flags) {
            atomicOr(&flags[FLAG_NEED], NEED_REGION);
            atomicMax(&flags[FLAG_MAX_REGION], r.total);
            flags[FLAG_HALT] = 1;
        }
        return false;
    }
    // region cell of every staged molecule, cell by cell
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int c = threadIdx.x >> 6; c < g.nreg; c += nw) {
        const int b = r.off[c], n = r.off[c + 1] - b;
        for (int l = lane; l < n; l += 64) r.rq[b + l] = (uint16_t)c;
    }
    __syncthreads();
    return true;
}

// Region tables + staged positions (and slots / chain ids when asked).  Ends with __syncthreads().
__device__ bool stage_region(Region& r, const BG& g, int brick, const int* start, const St& s, bool stage_pos,
                             const int* mono_src, int cap, int* lchain, int* lsub, int capm, int* flags,
                             const int* type_src = nullptr, const int* slot_map = nullptr) {
    if (!stage_tables(r, g, brick, start, capm, flags)) return false;
    if (stage_pos || r.lslot || r.ltype8 || lchain) {
        // up to SU molecules per thread and pass, all their loads issued before the first store
        // (out-of-range indices are clamped to the last molecule and not stored)
        constexpr int SU = 6;
        for (int q0 = threadIdx.x; q0 < r.total; q0 += SU * blockDim.x) {
            int sl[SU], sc[SU];
            double px[SU], py[SU], pz[SU];
#pragma unroll
            for (int k = 0; k < SU; ++k) {
                const int qc = min(q0 + k * (int)blockDim.x, r.total - 1);
                const int rc = r.rq[qc];
                sl[k] = r.cst[rc] + (qc - r.off[rc]);
                sc[k] = r.sh[rc];
                if (stage_pos) {
                    px[k] = s.x[sl[k]];
                    py[k] = s.y[sl[k]];
                    pz[k] = s.z[sl[k]];
                }
            }
            // stores without branches, so that the loads stay ahead of them: an index past the
            // region writes the last molecule again, with the same values
#pragma unroll
            for (int k = 0; k < SU; ++k) {
                const int q = min(q0 + k * (int)blockDim.x, r.total - 1);
                if (stage_pos) {
                    int sx, sy, sz;
                    decode_shift(sc[k], sx, sy, sz);
                    r.pos[3 * q] = img(px[k], sx, g.L[0]);
                    r.pos[3 * q + 1] = img(py[k], sy, g.L[1]);
                    r.pos[3 * q + 2] = img(pz[k], sz, g.L[2]);
                }
                if (r.lslot) r.lslot[q] = slot_map ? slot_map[sl[k]] : sl[k];
                if (r.ltype) r.ltype[q] = type_src[sl[k]];
                if (r.ltype8) r.ltype8[q] = (uint8_t)type_src[sl[k]];
                if (lchain) {
                    lchain[q] = mono_src[sl[k]];
                    lsub[q] = mono_src[sl[k] + (size_t)cap];
                }
            }
        }
    }
    __syncthreads();
    return true;
}

// Staged positions from the brick's table, which the list build wrote since the last binning: the
// slot and image shift of each staged molecule in the region layout (total and interior counts
// first), then the slot and LDS index of each interior molecule.  No region tables are formed.
// Ends with __syncthreads().
// (type_src: the one-byte types of the staged molecules into r.ltype8 too; key_src: the owner keys
// of the MOP tally, key_src[slot] << 1 | (image of a periodic shift), into r.lslot)
__device__ bool stage_from_table(Region& r, const BG& g, const St& s, const int* T, int capm,
                                 const int* type_src = nullptr, const int* key_src = nullptr) {
    r.total = T[0];
    r.ni = T[1];
    if (r.total > capm) return false;  // (a build that met a larger region halted its loop)
    const int* stg = T + 2;
    constexpr int SU = 6;
    for (int q0 = threadIdx.x; q0 < r.total; q0 += SU * blockDim.x) {
        int w[SU];
        double px[SU], py[SU], pz[SU];
#pragma unroll
        for (int k = 0; k < SU; ++k) w[k] = stg[min(q0 + k * (int)blockDim.x, r.total - 1)];
#pragma unroll
        for (int k = 0; k < SU; ++k) {
            const int sl = w[k] >> 5;
            px[k] = s.x[sl];
            py[k] = s.y[sl];
            pz[k] = s.z[sl];
        }
#pragma unroll
        for (int k = 0; k < SU; ++k) {
            const int q = min(q0 + k * (int)blockDim.x, r.total - 1);
            int sx, sy, sz;
            decode_shift(w[k] & 31, sx, sy, sz);
            r.pos[3 * q] = img(px[k], sx, g.L[0]);
            r.pos[3 * q + 1] = img(py[k], sy, g.L[1]);
            r.pos[3 * q + 2] = img(pz[k], sz, g.L[2]);
            if (type_src) r.ltype8[q] = (uint8_t)type_src[w[k] >> 5];
            if (key_src) r.lslot[q] = (key_src[w[k] >> 5] << 1) | ((w[k] & 31) != 13 ? 1 : 0);
        }
    }
    __syncthreads();
    return true;
}

// i-molecule q of the brick -> (region cell, LDS index, slot)
__device__ inline void imol(const Region& r, const int* start, int q, int& rc, int& li, int& slot);
// the same (without the region cell) from the brick's staging table when the region was staged
// from it (capm: the region capacity the table was laid out with)
__device__ inline void imol_tab(const Region& r, const int* start, const int* tab, int capm, int q, int& li, int& i) {
    if (tab) {
        i = tab[2 + capm + q];
        li = ((const uint16_t*)(tab + 2 + 2 * capm))[q];
    } else {
        int rc;
        imol(r, start, q, rc, li, i);
    }
}
__device__ inline void imol(const Region& r, const int* start, int q, int& rc, int& li, int& slot) {
    int lo = 0, hi = r.nic - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (r.ic_off[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    rc = r.ic_rc[lo];
    const int k = q - r.ic_off[lo];
    li = r.off[rc] + k;
    slot = r.cst[rc] + k;
}

__device__ inline int region_cell_of(const Region& r, int, int q) { return r.rq[q]; }

// Brick of workgroup b of a one-domain launch (gridDim.x == number of bricks).  Workgroups are
// dealt to the 8 XCDs in turn; XCD x takes the contiguous run x of the tiled brick order.
__device__ inline int brick_of_block(const int* order, int b, int nb) {
    if (!order) return b;
    const int x = b & 7, k = b >> 3, base = nb >> 3, rem = nb & 7;
    return order[x * base + min(x, rem) + k];
}

// ---- neighbour list: src/linklist.f90:660-804 (set semantics), bonds 2542-2584 -----------------
struct NlbArgs {
    St s;
    BG g;
    const int* start;
    const int* ext;
    const int* mono;  // slot order: chainID, subchainID, funcy, bflag[4] (stride cap) or null
    uint64_t* nbl;   // entries 4c..4c+3 of slot i in word nbl[c*cap + i] (16 bits each)
    int* nbn;
    uint32_t* bond;
    int* bondn;
    int* flags;
    int cap, maxnb, maxreg;
    double ra2;      // pairs kept: r^2 < ra2 (the list radius; rc^2 for a list read once, at the
                     // positions it is built from)
    int xsorted;     // cells are sorted by x (positions unchanged since binning)
    double xwin;     // window half-width: above sqrt(ra2), so no pair inside ra2 is skipped
    double cs[3], hd[3];  // cell side and half domain (cell faces in image coordinates)
    const int* order;     // brick order (brick_of_block)
};

__device__ inline int get_bondflag(const int* mono, int cap, int i, int scID) {
    // get_bondflag, src/modules.f90:597-610 (intbits = 31)
    int group = (int)ceilf((float)scID / 31.0f);
    int expo = scID % 31 - 1;
    if (expo == -1) expo = 30;
    if (group < 1 || group > 4) return 0;
    int w = mono[i + (size_t)(2 + group) * cap];
    return (w >> expo) & 1;
}
// the same from the molecule's four bin_bflag words held in registers
__device__ inline int bondflag_of(const int (&bf)[4], int scID) {
    const int group = (int)ceilf((float)scID / 31.0f);
    int expo = scID % 31 - 1;
    if (expo == -1) expo = 30;
    const int w = group == 1 ? bf[0] : group == 2 ? bf[1] : group == 3 ? bf[2] : bf[3];
    return (group >= 1 && group <= 4) ? (w >> expo) & 1 : 0;
}

template <bool POLY>
__global__ __launch_bounds__(FB_BLOCK) void k_nlist_brick(NlbArgs a) {
    extern __shared__ __align__(16) char smem[];
    Region r;
    carve(r, smem, a.g, a.maxreg, POLY);
    int* lchain = nullptr;
    int* lsub = nullptr;
    if (POLY) {  // chain ids of the staged molecules after the brick tables
        lchain = r.ic_rc + a.g.B[0] * a.g.B[1] * a.g.B[2];
        lsub = lchain + a.maxreg;
    }
    const int brick = brick_of_block(a.order, blockIdx.x, gridDim.x);
    if (!stage_region(r, a.g, brick, a.start, a.s, true, a.mono, a.cap, lchain, lsub, a.maxreg, a.flags)) return;
    int wmax = 0;  // longest list of this wavefront
    const size_t cap = a.cap;
    const int R0 = a.g.R[0], R1 = a.g.R[1];
    const int nrounds = (r.ni + blockDim.x - 1) / blockDim.x;
    for (int round = 0; round < nrounds; ++round) {
        const int q = round * blockDim.x + threadIdx.x;
        const bool active = q < r.ni;
        int cnt = 0;
        if (active) {
            int rc, li, slot;
            imol(r, a.start, q, rc, li, slot);
            const int rx = rc % R0, ry = (rc / R0) % R1, rz = rc / (R0 * R1);
            const double xi = r.pos[3 * li], yi = r.pos[3 * li + 1], zi = r.pos[3 * li + 2];
            int nb = 0, chain = 0, funcy = 0;
            if (POLY) {
                chain = lchain[li];
                funcy = a.mono[slot + 2 * cap];
            }
            uint64_t word = 0;
            // Distances from the molecule to its cell's faces in y and z (cell faces in image
            // coordinates; at a rebuild every molecule lies in its cell).  A run of cells one row
            // up or down is at least that far away, which narrows its x window, or removes the
            // run when the faces alone are beyond rn.  The margin keeps the window conservative.
            const int by = (brick / a.g.nb[0]) % a.g.nb[1], bz = brick / (a.g.nb[0] * a.g.nb[1]);
            const int vy = by * a.g.B[1] + ry - 1, vz = bz * a.g.B[2] + rz - 1;
            const double ylo = -a.hd[1] + vy * a.cs[1], zlo = -a.hd[2] + vz * a.cs[2];
            const double gyl = fmax(yi - ylo - 1e-9, 0.0), gyh = fmax(ylo + a.cs[1] - yi - 1e-9, 0.0);
            const double gzl = fmax(zi - zlo - 1e-9, 0.0), gzh = fmax(zlo + a.cs[2] - zi - 1e-9, 0.0);
            // the 27 neighbour cells as 9 runs of three x-consecutive region cells, which are
            // contiguous in the LDS layout; run k covers dz = k/3 - 1, dy = k%3 - 1
            auto run_window = [&](int k, int& jb, int& je) {
                const int dz = k / 3 - 1, dy = k % 3 - 1;
                const int c0 = (rx - 1) + R0 * ((ry + dy) + R1 * (rz + dz));
                jb = r.off[c0];
                je = r.off[c0 + 3];
                // (with bonds too: a bonded partner is closer than R0 <= r_n, so inside the window,
                // and the runs keep their order, so do the list and bond entries)
                if (!a.xsorted) return;
                const double gy = dy < 0 ? gyl : (dy > 0 ? gyh : 0.0);
                const double gz = dz < 0 ? gzl : (dz > 0 ? gzh : 0.0);
                const double w2 = a.ra2 - (gy * gy + gz * gz);
                if (!(w2 > 0.0)) {
                    je = jb;
                    return;
                }
                const double xw = (dy == 0 && dz == 0) ? a.xwin : sqrt(w2) * (1.0 + 1e-9) + 1e-12;
                // the run is sorted by x: keep only xi - xw <= x <= xi + xw
                const double xlo = xi - xw, xhi = xi + xw;
                int lo = jb, hi = je;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (r.pos[3 * mid] < xlo) lo = mid + 1;
                    else hi = mid;
                }
                jb = lo;
                lo = jb;
                hi = je;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (r.pos[3 * mid] <= xhi) lo = mid + 1;
                    else hi = mid;
                }
                je = lo;
            };
            if (!POLY) {
                // All nine windows first, then the runs nearest first: the lanes of a wavefront
                // then walk runs of similar length together (the order of a list's entries is
                // free; the set is the reference's).  The order is that of the distance g^2 from
                // the molecule to the run's y and z faces (its top 28 bits), then of the run
                // number, which does not depend on the radius: a list of a smaller radius is the
                // larger one's with the farther entries left out.  Packed as g2:28 | k:4 | jb:16 |
                // je:16 (a region's LDS holds far fewer than 65,536 molecules).
                uint64_t w[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    int jb, je;
                    run_window(k, jb, je);
                    const int dz = k / 3 - 1, dy = k % 3 - 1;
                    const double gy = dy < 0 ? gyl : (dy > 0 ? gyh : 0.0);
                    const double gz = dz < 0 ? gzl : (dz > 0 ? gzh : 0.0);
                    w[k] = ((uint64_t)(__float_as_uint((float)(gy * gy + gz * gz)) >> 4) << 36) | ((uint64_t)k << 32) |
                           ((uint64_t)jb << 16) | (uint64_t)je;
                }
#pragma unroll
                for (int round = 0; round < 9; ++round)  // odd-even transposition sort, ascending
#pragma unroll
                    for (int k = round & 1; k + 1 < 9; k += 2) {
                        const uint64_t x = w[k], y = w[k + 1];
                        w[k] = x < y ? x : y;
                        w[k + 1] = x < y ? y : x;
                    }
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    const int jb = (int)((w[k] >> 16) & 0xffffu), je = (int)(w[k] & 0xffffu);
                    // 32 candidates at a time: a hit mask first (distance tests only), then the
                    // hits in candidate order into the list words
                    for (int j0 = jb; j0 < je; j0 += 32) {
                        const int nn = min(32, je - j0);
                        uint32_t m = 0;
                        for (int u = 0; u < nn; ++u) {
                            const int j = j0 + u;
                            const double ddx = xi - r.pos[3 * j], ddy = yi - r.pos[3 * j + 1],
                                         ddz = zi - r.pos[3 * j + 2];
                            const double r2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                            m |= (r2 < a.ra2) ? (1u << u) : 0u;
                        }
                        if (li >= j0 && li < j0 + nn) m &= ~(1u << (li - j0));
                        // entries enter the word from the top: after four, the first is in
                        // bits 0-15
                        while (m) {
                            const int j = j0 + __builtin_ctz(m);
                            m &= m - 1;
                            word = (word >> 16) | ((uint64_t)(uint32_t)j << 48);
                            if ((cnt & 3) == 3 && cnt < a.maxnb) a.nbl[(size_t)(cnt >> 2) * cap + slot] = word;
                            ++cnt;
                        }
                    }
                }
            } else {
                for (int k = 0; k < 9; ++k) {
                    int jb, je;
                    run_window(k, jb, je);
                    for (int j = jb; j < je; ++j) {
                        const double ddx = xi - r.pos[3 * j], ddy = yi - r.pos[3 * j + 1], ddz = zi - r.pos[3 * j + 2];
                        const double r2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                        const bool hit = (r2 < a.ra2) && (j != li);
                        const uint32_t sh = (uint32_t)(cnt & 3) << 4;
                        word |= hit ? ((uint64_t)(uint32_t)j << sh) : 0ull;
                        if (hit && (cnt & 3) == 3) {
                            if (cnt < a.maxnb) a.nbl[(size_t)(cnt >> 2) * cap + slot] = word;
                            word = 0;
                        }
                        cnt += hit ? 1 : 0;
                        if (POLY && j != li && lchain[j] == chain && get_bondflag(a.mono, a.cap, slot, lsub[j])) {
                            const int c2 = region_cell_of(r, a.g.nreg, j);
                            if (nb < funcy && nb < FM_MAX_FUNCY)
                                a.bond[(size_t)nb * cap + slot] = ((uint32_t)r.lslot[j] << 5) | (uint32_t)r.sh[c2];
                            ++nb;
                        }
                    }
                }
            }
            if ((cnt & 3) != 0 && cnt < a.maxnb) {
                // the last, partial word: the non-polymer path fills words from the top
                const uint64_t last = POLY ? word : word >> (16 * (4 - (cnt & 3)));
                a.nbl[(size_t)(cnt >> 2) * cap + slot] = last;
            }
            a.nbn[slot] = cnt < a.maxnb ? cnt : a.maxnb;
            if (POLY) {
                a.bondn[slot] = nb < funcy ? nb : funcy;
                if (nb > funcy) set_err(a.flags, FM_EBOND, a.ext[slot]);
            }
        }
        int m = active ? cnt : 0;
        for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
        wmax = max(wmax, m);
    }
    // a list longer than the capacity: the workgroup's longest into FLAG_MAX_NB, NEED_LIST and
    // FLAG_HALT (the region tables are no longer read, so their first words hold the per-wave values)
    __syncthreads();
    if ((threadIdx.x & 63) == 0) r.off[threadIdx.x >> 6] = wmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = max(m, r.off[w]);
        if (m > a.maxnb) {
            atomicMax(&a.flags[FLAG_MAX_NB], m);
            atomicOr(&a.flags[FLAG_NEED], NEED_LIST);
            a.flags[FLAG_HALT] = 1;
        }
    }
}

// ---- neighbour list with a quantised prefilter: cells freshly binned (sorted by x), no bonds ----
// The same set as k_nlist_brick, from fewer and cheaper operations per candidate.  Positions are
// staged as 15-bit integers q = round((x - o)/h) on a grid of step h from o, two steps below the
// region's low corner, so a distance test is two packed 16-bit subtractions and two 16-bit dot
// products (8 bytes of LDS per candidate instead of 24).  Each component of a quantised
// difference is within eps = 1 + 1e-6 grid steps of the exact difference over h, so with
// s = |rij|/h and t = rn/h (Cauchy-Schwarz on the three components):
//   r2q <  lo = t^2 - 2 sqrt(3) eps t, less margins              =>  s < t: a pair,
//   r2q >= hi = t^2 + 2 sqrt(3) eps t + 3 eps^2, plus margins    =>  s > t: no pair,
// and a candidate between the two is decided by the reference's FP64 test
// (rij2 < rneighbr2, src/linklist.f90:717-720,750-754) on its positions in HBM.  The margins (1e-9 t^2)
// cover the rounding of the FP64 test itself, so the set is the FP64 set exactly.
typedef short s16x2 __attribute__((ext_vector_type(2)));

struct NlqArgs {
    St s;
    BG g;
    const int* start;
    uint64_t* nbl;
    int* nbn;
    int* flags;
    int cap, maxnb, maxreg;
    double ra2, xwin, h, inv_h;  // ra2, xwin, lo, hi: the radius of the pairs kept (as NlbArgs)
    int lo, hi;
    double cs[3], hd[3];
    const int* order;
    // bond discovery (POLY): monomer fields in slot order, the bond lists, molecule numbers
    const int* mono;
    uint32_t* bond;
    int* bondn;
    const int* ext;
    // optional: the bricks' staging tables for the plain force launches (stage_from_table)
    int* btab;
    int btab_stride;
};

__device__ inline int quantise(double x, double o, double inv_h) {
    const int q = (int)floor((x - o) * inv_h + 0.5);
    return min(max(q, 0), 32767);
}

// registers held to what lets three workgroups share a CU (its LDS region is a third of the
// force kernels').  POLY: the bonds of the chains are found among the list entries (a bonded
// partner lies within R0 <= r_n, so it is an entry), as check_update_adjacentbeadinfo finds them
// during the reference's scan (src/linklist.f90:2542-2584): the low 17 bits of each chain id
// ride in the spare bits of the quantised z, so the candidate test also marks the candidates of
// the molecule's own chain, and only hits among those read the full chain and subchain ids (from
// HBM, by the staged slot) and test the molecule's bond flags (held in registers).  The region's
// LDS is then that of the plain build, and so are its three workgroups per CU.
template <bool POLY>
__global__ __launch_bounds__(FB_BLOCK) __attribute__((amdgpu_waves_per_eu(3 * FB_BLOCK / 256))) void
k_nlist_quant(NlqArgs a) {
    extern __shared__ __align__(16) char smem[];
    const BG& g = a.g;
    uint2* qv = (uint2*)smem;               // maxreg: x | y << 16, z (POLY: | chain id << 15)
    int* ls = (int*)(qv + a.maxreg);        // maxreg: slot << 5 | image shift code
    Region r;
    r.pos = nullptr;
    r.lslot = nullptr;
    int* ip = ls + a.maxreg;
    r.rq = (uint16_t*)ip;
    ip += (a.maxreg + 1) / 2;
    r.off = ip;
    ip += g.nreg + 1;
    r.gc = ip;
    ip += g.nreg;
    r.sh = ip;
    ip += g.nreg;
    r.cst = ip;
    ip += g.nreg;
    r.ic_off = ip;
    ip += g.B[0] * g.B[1] * g.B[2] + 1;
    r.ic_rc = ip;
    const int brick = brick_of_block(a.order, blockIdx.x, gridDim.x);
    if (!stage_tables(r, g, brick, a.start, a.maxreg, a.flags)) return;
    const int bx = brick % g.nb[0], by = (brick / g.nb[0]) % g.nb[1], bz = brick / (g.nb[0] * g.nb[1]);
    const double ox = -a.hd[0] + (bx * g.B[0] - 1) * a.cs[0] - 2.0 * a.h;
    const double oy = -a.hd[1] + (by * g.B[1] - 1) * a.cs[1] - 2.0 * a.h;
    const double oz = -a.hd[2] + (bz * g.B[2] - 1) * a.cs[2] - 2.0 * a.h;
    constexpr int SU = 6;  // molecules per thread and pass, loads first (as stage_region)
    for (int q0 = threadIdx.x; q0 < r.total; q0 += SU * blockDim.x) {
        int sl[SU], sc[SU];
        double px[SU], py[SU], pz[SU];
#pragma unroll
        for (int k = 0; k < SU; ++k) {
            const int qc = min(q0 + k * (int)blockDim.x, r.total - 1);
            const int rc = r.rq[qc];
            sl[k] = r.cst[rc] + (qc - r.off[rc]);
            sc[k] = r.sh[rc];
            px[k] = a.s.x[sl[k]];
            py[k] = a.s.y[sl[k]];
            pz[k] = a.s.z[sl[k]];
        }
#pragma unroll
        for (int k = 0; k < SU; ++k) {
            const int q = q0 + k * (int)blockDim.x;
            if (q >= r.total) break;
            int sx, sy, sz;
            decode_shift(sc[k], sx, sy, sz);
            const int xq = quantise(img(px[k], sx, g.L[0]), ox, a.inv_h);
            const int yq = quantise(img(py[k], sy, g.L[1]), oy, a.inv_h);
            const int zq = quantise(img(pz[k], sz, g.L[2]), oz, a.inv_h);
            // (POLY: z | low 17 bits of the chain id << 15)
            qv[q] = make_uint2((uint32_t)xq | ((uint32_t)yq << 16),
                               (uint32_t)zq | (POLY ? ((uint32_t)a.mono[sl[k]] & 0x1ffffu) << 15 : 0u));
            ls[q] = (sl[k] << 5) | sc[k];
            if (a.btab) a.btab[(size_t)brick * a.btab_stride + 2 + q] = (sl[k] << 5) | sc[k];
        }
    }
    if (a.btab && threadIdx.x == 0) {
        a.btab[(size_t)brick * a.btab_stride] = r.total;
        a.btab[(size_t)brick * a.btab_stride + 1] = r.ni;
    }
    __syncthreads();
    int wmax = 0;
    const size_t cap = a.cap;
    const int R0 = g.R[0], R1 = g.R[1];
    const int nrounds = (r.ni + blockDim.x - 1) / blockDim.x;
    for (int round = 0; round < nrounds; ++round) {
        const int qi = round * blockDim.x + threadIdx.x;
        const bool active = qi < r.ni;
        int cnt = 0;
        if (active) {
            int rc, li, slot;
            imol(r, a.start, qi, rc, li, slot);
            if (a.btab) {
                int* const T = a.btab + (size_t)brick * a.btab_stride + 2 + a.maxreg;
                T[qi] = slot;
                ((uint16_t*)(T + a.maxreg))[qi] = (uint16_t)li;
            }
            const int rx = rc % R0, ry = (rc / R0) % R1, rz = rc / (R0 * R1);
            // interior cells carry no image shift: the staged position is the stored one
            const double xi = a.s.x[slot], yi = a.s.y[slot], zi = a.s.z[slot];
            const uint2 wi = qv[li];
            const s16x2 ixy = __builtin_bit_cast(s16x2, wi.x);
            const int izq = (int)(wi.y & 0x7fffu);
            const uint32_t ich = wi.y >> 15;
            int nb = 0, chain = 0, funcy = 0;
            int bf[4] = {0, 0, 0, 0};
            if (POLY) {
                chain = a.mono[slot];
                funcy = a.mono[slot + 2 * cap];
#pragma unroll
                for (int g4 = 0; g4 < 4; ++g4) bf[g4] = a.mono[slot + (size_t)(3 + g4) * cap];
            }
            // (POLY) region indices of the waiting hits of the molecule's own chain, two per word
            uint32_t pend[4] = {0u, 0u, 0u, 0u};
            int npend = 0;
            // the bond test of the waiting hits, four at a time: the chain and subchain ids of
            // their slots loaded together, then the tests and bond entries in hit order
            auto flush_bonds = [&]() {
                for (int k0 = 0; k0 < npend; k0 += 4) {
                    int sj[4], cj[4], sc[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int jj = (int)((pend[(k0 + k) >> 1] >> (16 * ((k0 + k) & 1))) & 0xffffu);
                        sj[k] = k0 + k < npend ? ls[jj] : 0;
                    }
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int sl = sj[k] >> 5;
                        cj[k] = k0 + k < npend ? a.mono[sl] : ~chain;
                        sc[k] = k0 + k < npend ? a.mono[sl + cap] : 0;
                    }
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        if (cj[k] == chain && bondflag_of(bf, sc[k])) {
                            if (nb < funcy && nb < FM_MAX_FUNCY) a.bond[(size_t)nb * cap + slot] = (uint32_t)sj[k];
                            ++nb;
                        }
                }
                npend = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k) pend[k] = 0u;
            };
            // faces of the molecule's cell in y and z, as in k_nlist_brick
            const int vy = by * g.B[1] + ry - 1, vz = bz * g.B[2] + rz - 1;
            const double ylo = -a.hd[1] + vy * a.cs[1], zlo = -a.hd[2] + vz * a.cs[2];
            const double gyl = fmax(yi - ylo - 1e-9, 0.0), gyh = fmax(ylo + a.cs[1] - yi - 1e-9, 0.0);
            const double gzl = fmax(zi - zlo - 1e-9, 0.0), gzh = fmax(zlo + a.cs[2] - zi - 1e-9, 0.0);
            // nine runs of three x-consecutive region cells; each x window (of the radius of the
            // pairs kept) widened by a grid step on both sides of its quantised bounds.  The runs
            // go nearest first: by the distance g^2 from the molecule to the run's y and z faces
            // (its top 28 bits), then by run number.  That order does not depend on the radius,
            // so a list of a smaller radius is the larger one's with the farther entries left out.
            // (the order keys only, g2:28 | k:4; each run's window is searched when its turn comes)
            uint32_t w[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const int dz = k / 3 - 1, dy = k % 3 - 1;
                const double gy = dy < 0 ? gyl : (dy > 0 ? gyh : 0.0);
                const double gz = dz < 0 ? gzl : (dz > 0 ? gzh : 0.0);
                w[k] = ((__float_as_uint((float)(gy * gy + gz * gz)) >> 4) << 4) | (uint32_t)k;
            }
#pragma unroll
            for (int sr = 0; sr < 9; ++sr)  // odd-even transposition sort, ascending
#pragma unroll
                for (int k = sr & 1; k + 1 < 9; k += 2) {
                    const uint32_t x = w[k], y = w[k + 1];
                    w[k] = x < y ? x : y;
                    w[k + 1] = x < y ? y : x;
                }
            uint64_t word = 0;
            for (int kk = 0; kk < 9; ++kk) {
                const int k = (int)(w[kk] & 15u);
                const int dz = k / 3 - 1, dy = k % 3 - 1;
                const int c0 = (rx - 1) + R0 * ((ry + dy) + R1 * (rz + dz));
                int jb = r.off[c0], je = r.off[c0 + 3];
                const double gy = dy < 0 ? gyl : (dy > 0 ? gyh : 0.0);
                const double gz = dz < 0 ? gzl : (dz > 0 ? gzh : 0.0);
                const double w2 = a.ra2 - (gy * gy + gz * gz);
                if (!(w2 > 0.0)) continue;
                {
                    const double xw = (dy == 0 && dz == 0) ? a.xwin : sqrt(w2) * (1.0 + 1e-9) + 1e-12;
                    const int xl = (int)floor((xi - xw - ox) * a.inv_h) - 1;
                    const int xh = (int)ceil((xi + xw - ox) * a.inv_h) + 1;
                    int lo = jb, hi = je;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if ((int)(qv[mid].x & 0xffffu) < xl) lo = mid + 1;
                        else hi = mid;
                    }
                    jb = lo;
                    hi = je;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if ((int)(qv[mid].x & 0xffffu) <= xh) lo = mid + 1;
                        else hi = mid;
                    }
                    je = lo;
                }
                for (int j0 = jb; j0 < je; j0 += 32) {
                    const int nn = min(32, je - j0);
                    // each test shifts its two bits in from the bottom (the first candidate ends
                    // highest); one bit reversal afterwards puts candidate u at bit u
                    uint32_t m = 0, b = 0, sc = 0;
                    auto test = [&](int j) {
                        const uint2 wj = qv[j];
                        const s16x2 d1 = ixy - __builtin_bit_cast(s16x2, wj.x);
                        const int dz = izq - (int)(POLY ? wj.y & 0x7fffu : wj.y);
                        const int r2q = __builtin_amdgcn_sdot2(d1, d1, __mul24(dz, dz), false);
                        m = m + m + (uint32_t)(r2q < a.lo);
                        b = b + b + (uint32_t)(r2q < a.hi);
                        if (POLY) sc = sc + sc + (uint32_t)((wj.y >> 15) == ich);
                    };
                    int u = 0;
                    for (; u + 4 <= nn; u += 4) {
                        test(j0 + u);
                        test(j0 + u + 1);
                        test(j0 + u + 2);
                        test(j0 + u + 3);
                    }
                    for (; u < nn; ++u) test(j0 + u);
                    m = __builtin_bitreverse32(m) >> (32 - nn);
                    b = __builtin_bitreverse32(b) >> (32 - nn);
                    if (POLY) sc = __builtin_bitreverse32(sc) >> (32 - nn);
                    if (li >= j0 && li < j0 + nn) {
                        m &= ~(1u << (li - j0));
                        b &= ~(1u << (li - j0));
                    }
                    // the undecided candidates: the FP64 test on the positions in HBM
                    for (uint32_t u8 = b & ~m; u8; u8 &= u8 - 1) {
                        const int u = __builtin_ctz(u8);
                        const int pk = ls[j0 + u];
                        const int sl = pk >> 5;
                        int sx, sy, sz;
                        decode_shift(pk & 31, sx, sy, sz);
                        const double ddx = xi - img(a.s.x[sl], sx, g.L[0]), ddy = yi - img(a.s.y[sl], sy, g.L[1]),
                                     ddz = zi - img(a.s.z[sl], sz, g.L[2]);
                        const double r2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                        if (r2 < a.ra2) m |= 1u << u;
                    }
                    // entries enter the word from the top: after four, the first is in bits 0-15
                    // (POLY: the hits of the molecule's own chain, for the bond test below)
                    const uint32_t same = POLY ? m & sc : 0u;
                    while (m) {
                        const int j = j0 + __builtin_ctz(m);
                        m &= m - 1;
                        word = (word >> 16) | ((uint64_t)(uint32_t)j << 48);
                        if ((cnt & 3) == 3 && cnt < a.maxnb) a.nbl[(size_t)(cnt >> 2) * cap + slot] = word;
                        ++cnt;
                    }
                    if (POLY) {
                        // hits of the molecule's own chain wait in registers (up to eight) for
                        // the bond test after the runs, where their id loads go out together; a
                        // ninth first has the eight tested (hit order is kept)
                        for (uint32_t h = same; h; h &= h - 1) {
                            if (npend == 8) flush_bonds();
                            const int j = j0 + __builtin_ctz(h);
                            pend[npend >> 1] |= (uint32_t)j << (16 * (npend & 1));
                            ++npend;
                        }
                    }
                }
            }
            if ((cnt & 3) != 0 && cnt < a.maxnb) a.nbl[(size_t)(cnt >> 2) * cap + slot] = word >> (16 * (4 - (cnt & 3)));
            a.nbn[slot] = cnt < a.maxnb ? cnt : a.maxnb;
            if (POLY) {
                flush_bonds();
                a.bondn[slot] = nb < funcy ? nb : funcy;
                if (nb > funcy) set_err(a.flags, FM_EBOND, a.ext[slot]);
            }
        }
        int m = active ? cnt : 0;
        for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
        wmax = max(wmax, m);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) r.off[threadIdx.x >> 6] = wmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = max(m, r.off[w]);
        if (m > a.maxnb) {
            atomicMax(&a.flags[FLAG_MAX_NB], m);
            atomicOr(&a.flags[FLAG_NEED], NEED_LIST);
            a.flags[FLAG_HALT] = 1;
        }
    }
}

// ---- forces: src/simulation_compute_forces.f90:628-726 (pairs), 731-789 (bonds) ----------------
struct FbArgs {
    St s;
    BG g;
    const int* start;
    const uint64_t* nbl;
    const int* nbn;
    const uint32_t* bond;
    const int* bondn;
    const int* ext;
    const int* moltype;    // null unless Mie
    const int* tag;        // null unless tag_move
    double* teth;          // rtether xyz, slidev xyz
    double* pe_lj;
    double* pe_poly;
    double* vir;
    double* rf;            // 9*cap or null
    double* pxy;           // 3*nplanes or null
    const double* planes;
    int* flags;
    int cap, maxreg;
    int rec, ev;
    int kick;              // fused simulation_move_particles_vv(2): v += (0.5*dt)*a
    double hdt;
    double rc2, potshift;
    double k_c, R0, R02, dt;
    double teth_k2, teth_k4, teth_k6;
    double hx, hz;
    int nplanes;
    double hp, spacing, shift;
    const double* mie;     // [7][FM_MAX_TYPES^2] + mass[FM_MAX_TYPES]
    const int* order;      // brick order of one-domain launches (brick_of_block)
    int* halt;             // flag words of a conditional launch (null: unconditional)
    int step;              // step whose forces this launch computes (the halt rule compares it)
    int active;            // packed launches: 0 = this domain has no step in this launch
    float band32;          // mixed, FP32 staging: |r^2 - rc^2| below which FP64 decides the cutoff
    double* cv;            // faces 1..3 of the control-volume tally, or null
    int cvnbo[3], cvnhb[3];
    double cvbs[3], cvhd[3];
    // pass 1 of the next step fused into the force launch (k_forces_brick_plain<true>): the state
    // after it goes to s2 (the other state buffer), with the displacement update of rd
    St s2;
    double* rd;
    double hdt2, thresh;
    int fresh, p1step;
    // the bricks' staging tables written by the list build since the last binning, or null
    // (stage_from_table; the LJ plain launches)
    const int* btab;
    int btab_stride;
    // the branch-free Mie launch's partner types and exponent class (mie_plain_types), or 0
    int mie_nt, mie_nb;
};

// forward cell offsets of the reference's half-list scan (src/linklist.f90:686-688)
// (dz = 1: all nine; dz = 0: (1,0), (1,1), (0,1), (-1,1))
__device__ inline bool is_forward(int dx, int dy, int dz) {
    if (dz != 0) return dz == 1;
    if (dy != 0) return dy == 1;
    return dx == 1;
}

__device__ inline double dp_heaviside(double x) { return ceil(copysign(0.5, x)); }

// pressure_tensor_forces_MOP (src/simulation_record.f90:7068-7112) for one list entry
__device__ void mop_entry(const FbArgs& a, const double ri[3], const double rj[3], const double rij[3],
                          double acc, double* spxy) {
    int pi = (int)llround((ri[1] + a.hp - a.shift) / a.spacing) + 1;
    pi = pi < 1 ? 1 : (pi > a.nplanes ? a.nplanes : pi);
    int pj = (int)llround((rj[1] + a.hp - a.shift) / a.spacing) + 1;
    pj = pj < 1 ? 1 : (pj > a.nplanes ? a.nplanes : pj);
    const int step = (pj - pi) >= 0 ? 1 : -1;
    for (int n = pi;; n += step) {
        const double plane = a.planes[n - 1];
        const double pyb0 = ri[0] + (rij[0] / rij[1]) * (plane - ri[1]);
        const double pyb2 = ri[2] + (rij[2] / rij[1]) * (plane - ri[1]);
        const double fx = copysign(1.0, ri[1] - plane) - copysign(1.0, rj[1] - plane);
        const double hxx = dp_heaviside(a.hx - pyb0) - dp_heaviside(-a.hx - pyb0);
        const double hzz = dp_heaviside(a.hz - pyb2) - dp_heaviside(-a.hz - pyb2);
        if (fx != 0.0 && hxx != 0.0 && hzz != 0.0)
            for (int d = 0; d < 3; ++d) atomicAdd(&spxy[3 * (n - 1) + d], acc * rij[d] * fx * hxx * hzz);
        if (n == pj) break;
    }
}

// control_volume_stresses_opt -> cumulative_flux_opt -> get_crossings for one list entry
// (src/simulation_record.f90:5281-5317, 4878-5064, 308-353) with fij = f*rij: the segment ri->rj
// crosses the bin planes between the two molecules' bins along each axis n; every crossing adds
// crossdir*2*fij to face n of the bin just above the plane.  (Face n+3 is a copy taken on
// download.)  Bins, crossing points and the floor of tiny separations follow the reference's
// FP64 operations exactly, so every contribution lands in the reference's bin.
__device__ void cv_entry(const FbArgs& a, const double ri[3], const double rj[3], double f, int ext_i) {
    const double eps = (double)1e-12f;  // a default-real literal in the reference
    double q[3], r12[3];
    int b1[3], b2[3];
    for (int d = 0; d < 3; ++d) {
        const double rd = ri[d] - rj[d];
        q[d] = 2.0 * (f * rd);
        r12[d] = fabs(rd) < 0.000001 ? 0.000001 : rd;
        b1[d] = (int)ceil((ri[d] + a.cvhd[d]) / a.cvbs[d]) + a.cvnhb[d];
        b2[d] = (int)ceil((rj[d] + a.cvhd[d]) / a.cvbs[d]) + a.cvnhb[d];
    }
    const size_t nbin = (size_t)a.cvnbo[0] * a.cvnbo[1] * a.cvnbo[2];
    for (int n = 0; n < 3; ++n) {
        if (b1[n] == b2[n]) continue;
        const int t1 = n == 2 ? 0 : n + 1, t2 = n == 0 ? 2 : n - 1;
        const double crossdir = copysign(1.0, r12[n]);
        const int lo = min(b1[n], b2[n]), hi = max(b1[n], b2[n]);
        for (int i = lo; i < hi; ++i) {
            const double pt = (double)(i - a.cvnhb[n]) * a.cvbs[n] - a.cvhd[n];
            double rc[3];
            rc[t1] = ri[t1] + (r12[t1] / r12[n]) * (pt - ri[n]);
            rc[t2] = ri[t2] + (r12[t2] / r12[n]) * (pt - ri[n]);
            rc[n] = pt + eps;
            int cb[3];
            bool ok = true;
            for (int d = 0; d < 3; ++d) {
                cb[d] = (int)ceil((rc[d] + a.cvhd[d]) / a.cvbs[d]) + a.cvnhb[d];
                ok = ok && cb[d] >= 1 && cb[d] <= a.cvnbo[d];
            }
            if (!ok) {
                set_err(a.flags, FM_EOVERFLOW, ext_i);
                continue;
            }
            double* F = a.cv + (size_t)(cb[0] - 1) + a.cvnbo[0] * ((size_t)(cb[1] - 1) + a.cvnbo[1] * (size_t)(cb[2] - 1));
            for (int c = 0; c < 3; ++c) atomicAdd(F + (size_t)(c + 3 * n) * nbin, crossdir * q[c]);
        }
    }
}

// Does the reference visit this entry in its MOP tally?  Shifted entries: always (one per end).
// Unshifted pairs: once, from the molecule whose cell scans the other's as a forward cell, or in
// the same cell from the higher molecule number (head-inserted list order).
__device__ inline bool mop_owner(const Region& r, const BG& g, int rci, int rcj, int exti, int extj) {
    const int shj = r.sh[rcj];
    if (shj != 13 || r.sh[rci] != 13) return true;
    if (rcj == rci) return exti > extj;
    const int dx = rcj % g.R[0] - rci % g.R[0];
    const int dy = (rcj / g.R[0]) % g.R[1] - (rci / g.R[0]) % g.R[1];
    const int dz = rcj / (g.R[0] * g.R[1]) - rci / (g.R[0] * g.R[1]);
    return is_forward(dx, dy, dz);
}

// LJ list loop of a step without tallies.  Four entries per word, evaluated without branches: an
// entry beyond rc, or past the end of the list (index 0 pads a last, partial word), gets a zero
// magnitude.  Two words are in flight: the next word's positions are read from LDS while this
// word is evaluated, and the word after it from HBM.  The pair law 48 x^4 (x^3 - 1/2) is summed
// as x^4 (2 x^3 - 1) r_ij and scaled by 24 at the end, with x = 1/r^2 from one Newton step;
// within 2e-14 relative of the reference's 48 (x^7 - x^4/2) per pair.
struct P4 {
    double x[4], y[4], z[4];
};
__device__ __forceinline__ void lj_stage(P4& p, const double* pos, uint64_t w) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int j = (int)((w >> (16 * q)) & 0xffffu);
        p.x[q] = pos[3 * j];
        p.y[q] = pos[3 * j + 1];
        p.z[q] = pos[3 * j + 2];
    }
}
__device__ __forceinline__ void lj_word(const P4& p, int left, double xi, double yi, double zi, double rc2,
                                        double& fx, double& fy, double& fz) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double rx = xi - p.x[q], ry = yi - p.y[q], rz = zi - p.z[q];
        const double r2 = (rx * rx + ry * ry) + rz * rz;
        double x = recip1(r2);
        // an entry that does not count keeps only the low word of x: at most a denormal, whose
        // square underflows to zero, so its magnitude below is exactly -0
        x = (r2 < rc2 && q < left) ? x : __hiloint2double(0, __double2loint(x));
        const double x2 = x * x;
        const double x4 = x2 * x2;
        const double acc = x4 * fmaf64(2.0, x2 * x, -1.0);
        fx = fmaf64(acc, rx, fx);
        fy = fmaf64(acc, ry, fy);
        fz = fmaf64(acc, rz, fz);
    }
}
// The same loop on a record step with only PE and virial tallied: 1/r^2 from two Newton steps (the
// IEEE quotient), the PE term in the reference's expansion 4(x^6 - x^3) - potshift (an entry that
// does not count adds exactly zero), and the virial as x^4 (2x^3 - 1) r^2, scaled by 24 at the end.
__device__ __forceinline__ void lj_word_rec(const P4& p, int left, double xi, double yi, double zi, double rc2,
                                            double potshift, double& fx, double& fy, double& fz, double& pe,
                                            double& vt) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double rx = xi - p.x[q], ry = yi - p.y[q], rz = zi - p.z[q];
        const double r2 = (rx * rx + ry * ry) + rz * rz;
        const bool in = r2 < rc2 && q < left;
        double x = recip(r2);
        x = in ? x : __hiloint2double(0, __double2loint(x));
        const double x2 = x * x;
        const double x3 = x * x2;
        const double x4 = x2 * x2;
        const double t = x4 * fmaf64(2.0, x3, -1.0);
        fx = fmaf64(t, rx, fx);
        fy = fmaf64(t, ry, fy);
        fz = fmaf64(t, rz, fz);
        vt = fmaf64(t, r2, vt);
        const double e = 4.0 * (x3 * x3 - x3) - potshift;
        pe = pe + (in ? e : 0.0);
    }
}

// REC: also the molecule's PE and virial (lj_word_rec).
template <bool REC = false>
__device__ __forceinline__ void lj_plain(const double* pos, const uint64_t* row, int n, size_t cap, double xi,
                                         double yi, double zi, double rc2, double& fx, double& fy, double& fz,
                                         double potshift = 0.0, double* pe_out = nullptr, double* vir_out = nullptr) {
    double sx = 0, sy = 0, sz = 0;
    double pe = 0.0, vt = 0.0;
    P4 pa, pb;
    // At the top of each pass pa holds the positions of word k = kw/4, w1 is word k+1 and w2
    // word k+2 (in flight).  Word loads past the end re-read the last word (the row always has
    // at least one word), so no load sits behind a branch and each wait is for the oldest load;
    // a pass always evaluates both words (the second one masked off past the end) so that the
    // loop body is one block and the loads stay ahead of their uses.
    const int lastw = n > 0 ? (n - 1) >> 2 : 0;
    uint64_t w0 = row[0];
    lj_stage(pa, pos, n > 0 ? w0 : 0);
    uint64_t w1 = row[(size_t)min(1, lastw) * cap];
    uint64_t w2 = row[(size_t)min(2, lastw) * cap];
    for (int kw = 0; kw < n; kw += 8) {
        const int k = kw >> 2;
        lj_stage(pb, pos, kw + 4 < n ? w1 : 0);
        // both words two ahead are loaded here, at the top of the pass: the wait the compiler
        // places at the loop head is then for loads a whole pass old
        const uint64_t wk2 = w2;
        w1 = row[(size_t)min(k + 3, lastw) * cap];
        w2 = row[(size_t)min(k + 4, lastw) * cap];
        __builtin_amdgcn_sched_barrier(0);  // keep the reads above the evaluation
        if constexpr (REC) lj_word_rec(pa, n - kw, xi, yi, zi, rc2, potshift, sx, sy, sz, pe, vt);
        else lj_word(pa, n - kw, xi, yi, zi, rc2, sx, sy, sz);
        lj_stage(pa, pos, kw + 8 < n ? wk2 : 0);
        __builtin_amdgcn_sched_barrier(0);
        // the second word only when a lane of the wavefront has entries in it
        if (__builtin_amdgcn_ballot_w64(kw + 4 < n) != 0ull) {
            if constexpr (REC) lj_word_rec(pb, n - kw - 4, xi, yi, zi, rc2, potshift, sx, sy, sz, pe, vt);
            else lj_word(pb, n - kw - 4, xi, yi, zi, rc2, sx, sy, sz);
        }
    }
    fx = 24.0 * sx;
    fy = 24.0 * sy;
    fz = 24.0 * sz;
    if constexpr (REC) {
        *pe_out = pe;
        *vir_out = 24.0 * vt;
    }
}

// The launch's argument block read again from the kernel's argument segment (the launch's only
// argument, at its start): an empty asm hides that the pointer is the same on each call, so the
// fields read through it after a pair loop are loaded there and not held in scalar registers
// through the loop.  Only for kernels whose one argument is an FbArgs.
__device__ __forceinline__ const FbArgs& fargs_again() {
    const __attribute__((address_space(4))) char* kp =
        (const __attribute__((address_space(4))) char*)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(kp));
    return *(const FbArgs*)kp;
}

// simulation_move_particles_vv(2) of this step, then (1) of the next step and the rdisp update of
// rebuild criterion 2 (the arithmetic of k_pass2 and k_pass1_disp), the state after them written
// to s2; the first lane of the wavefront that sees a displacement beyond the threshold raises the
// rebuild flag with the next step's number.  KEEP_V (record steps): v at the end of this step stays
// in s for the record-step sums (no other thread of the launch reads this molecule's velocity).
template <bool KEEP_V>
__device__ __forceinline__ void vv_pass2_pass1(const FbArgs& a, int i, double fx, double fy, double fz) {
    const double ux = a.s.vx[i] + a.hdt * fx, uy = a.s.vy[i] + a.hdt * fy, uz = a.s.vz[i] + a.hdt * fz;
    a.s2.x[i] = a.s.x[i] + a.dt * ux + a.hdt2 * fx;
    a.s2.y[i] = a.s.y[i] + a.dt * uy + a.hdt2 * fy;
    a.s2.z[i] = a.s.z[i] + a.dt * uz + a.hdt2 * fz;
    const double wx = ux + a.hdt * fx, wy = uy + a.hdt * fy, wz = uz + a.hdt * fz;
    a.s2.vx[i] = wx;
    a.s2.vy[i] = wy;
    a.s2.vz[i] = wz;
    a.s2.ax[i] = fx;
    a.s2.ay[i] = fy;
    a.s2.az[i] = fz;
    const size_t cp = a.cap;
    double d0 = a.fresh ? 0.0 : a.rd[i], d1 = a.fresh ? 0.0 : a.rd[i + cp], d2 = a.fresh ? 0.0 : a.rd[i + 2 * cp];
    d0 = d0 + wx * a.dt;
    d1 = d1 + wy * a.dt;
    d2 = d2 + wz * a.dt;
    a.rd[i] = d0;
    a.rd[i + cp] = d1;
    a.rd[i + 2 * cp] = d2;
    const bool hit = fabs(d0) > a.thresh || fabs(d1) > a.thresh || fabs(d2) > a.thresh;
    const unsigned long long m = __ballot(hit);
    if (m != 0ull && (int)(threadIdx.x & 63) == __ffsll((long long)m) - 1 &&
        __atomic_load_n(&a.flags[FLAG_REBUILD], __ATOMIC_RELAXED) == 0) {
        atomicOr(&a.flags[FLAG_REBUILD], 1);
        atomicExch(&a.flags[FLAG_RB_STEP], a.p1step);
    }
    if constexpr (KEEP_V) {
        a.s.vx[i] = ux;
        a.s.vy[i] = uy;
        a.s.vz[i] = uz;
    }
}

// FENE bonds (simulation_compute_forces_poly) and tethers (tether_force) of molecule i, added to
// its force (fx, fy, fz already per unit mass); on record steps the bond energy and tallies too.
template <bool RECORD, bool VIR>
__device__ __forceinline__ void bond_tether_forces(const FbArgs& a, int i, double xi, double yi, double zi,
                                                   double inv_mi, bool do_rf, double* rf, double& fx, double& fy,
                                                   double& fz, double& vir) {
    const size_t cap = a.cap;
    if (a.bondn) {  // FENE, simulation_compute_forces_poly
        const int nb = a.bondn[i];
        double pp = 0.0;
        for (int b = 0; b < nb; ++b) {
            const uint32_t w = a.bond[(size_t)b * cap + i];
            const int j = (int)(w >> 5);
            int sx, sy, sz;
            decode_shift((int)(w & 31u), sx, sy, sz);
            const double xj = img(a.s.x[j], sx, a.g.L[0]), yj = img(a.s.y[j], sy, a.g.L[1]),
                         zj = img(a.s.z[j], sz, a.g.L[2]);
            const double rx = xi - xj, ry = yi - yj, rz = zi - zj;
            const double r2 = (rx * rx + ry * ry) + rz * rz;
            if (r2 >= a.R02) {
                set_err(a.flags, FM_EBOND, a.ext[i]);
                continue;
            }
            const double acc = -a.k_c / (1.0 - (r2 / a.R02));
            fx = fx + acc * rx * inv_mi;
            fy = fy + acc * ry * inv_mi;
            fz = fz + acc * rz * inv_mi;
            if (VIR) vir = vir + acc * r2;
            if (RECORD) {
                if (a.cv) {
                    const double ri[3] = {xi, yi, zi}, rj[3] = {xj, yj, zj};
                    cv_entry(a, ri, rj, acc, a.ext[i]);
                }
                pp = pp + 0.5 * a.k_c * a.R0 * a.R0 * log(1.0 - (r2 / a.R02));
                if (do_rf) {
                    const double rr[3] = {rx, ry, rz};
                    for (int p = 0; p < 3; ++p)
                        for (int qq = 0; qq < 3; ++qq) rf[3 * p + qq] = rf[3 * p + qq] + acc * rr[p] * rr[qq];
                }
            }
        }
        if (RECORD) a.pe_poly[i] = pp;
    }
    if (a.tag) {  // tether_force, src/molecular_properties.f90:833-915
        const int tg = a.tag[i];
        if (tg == FM_TAG_TETH || tg == FM_TAG_TETH_SLIDE) {
            double* T = a.teth;
            const double rio[3] = {xi - T[i], yi - T[i + cap], zi - T[i + 2 * cap]};
            const double mag = sqrt((rio[0] * rio[0] + rio[1] * rio[1]) + rio[2] * rio[2]);
            const double am = -2.0 * a.teth_k2 * mag - 4.0 * a.teth_k4 * (mag * mag) -
                              6.0 * a.teth_k6 * pow(mag, 4.0);
            fx = fx + am * rio[0];
            fy = fy + am * rio[1];
            fz = fz + am * rio[2];
            for (int d = 0; d < 3; ++d) T[i + d * cap] = T[i + d * cap] + T[i + (3 + d) * cap] * a.dt;
        }
    }
}

// VIR: virial on a non-record step (teval); record steps always tally it.  P1: VV pass 2 of this
// step, pass 1 of the next one and its displacement update follow the force (same arithmetic as
// the separate kernels), written to the other state buffer.
template <bool MIE, bool RECORD, bool EXTRA, bool VIR = RECORD, bool P1 = false, bool SIMPLE = false>
__device__ __forceinline__ void forces_body(const FbArgs& a, int brick, char* smem) {
    Region r;
    const bool mop = !SIMPLE && RECORD && a.pxy != nullptr;
    // the stress tallies need each staged molecule's slot (molecule numbers for the owner rule)
    const bool cvrec = !SIMPLE && RECORD && a.cv != nullptr;
    carve(r, smem, a.g, a.maxreg, mop || MIE || cvrec, MIE);
    double* spxy = nullptr;
    if (mop) spxy = (double*)(smem + region_lds_bytes(a.g, a.maxreg, true, MIE));
    // the Mie tables (7 fields per type pair, then the mass per type) in LDS after the region and
    // the MOP planes
    const int T2 = FM_MAX_TYPES * FM_MAX_TYPES;
    const double* mtab = a.mie;
    if constexpr (MIE) {
        double* t = (double*)(smem + region_lds_bytes(a.g, a.maxreg, mop || MIE || cvrec, MIE) +
                              (mop ? 3 * (size_t)a.nplanes * sizeof(double) : 0));
        for (int k = threadIdx.x; k < FM_MIE_WORDS; k += blockDim.x) t[k] = a.mie[k];
        mtab = t;  // visible after the staging's barriers below
    }
    // the staged slot array carries molecule numbers here (the stress tallies' owner rule is its
    // only reader); the LJ plain launches stage from the list build's table when there is one
    const int* tab = SIMPLE && a.btab ? a.btab + (size_t)brick * a.btab_stride : nullptr;
    if (tab) {
        if (!stage_from_table(r, a.g, a.s, tab, a.maxreg)) return;
    } else if (!stage_region(r, a.g, brick, a.start, a.s, true, nullptr, a.cap, nullptr, nullptr, a.maxreg,
                             a.flags, MIE ? a.moltype : nullptr, a.ext))
        return;
    if (mop) {
        for (int t = threadIdx.x; t < 3 * a.nplanes; t += blockDim.x) spxy[t] = 0.0;
        __syncthreads();
    }
    const size_t cap = a.cap;
    // a record step that tallies PE and virial only (no MOP, CV or rfmol)
    const bool simple_rec = RECORD && (SIMPLE || (!mop && !cvrec && a.rf == nullptr));
    for (int q = threadIdx.x; q < r.ni; q += blockDim.x) {
        int rci = 0, li, i;
        if (tab) {
            i = tab[2 + a.maxreg + q];
            li = ((const uint16_t*)(tab + 2 + 2 * a.maxreg))[q];
        } else {
            imol(r, a.start, q, rci, li, i);
        }
        const double xi = r.pos[3 * li], yi = r.pos[3 * li + 1], zi = r.pos[3 * li + 2];
        const int n = a.nbn[i];
        double fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0, vir = 0.0;
        double rf[9];
        const bool do_rf = !SIMPLE && RECORD && a.rf != nullptr;
        if (do_rf)
            for (int k = 0; k < 9; ++k) rf[k] = 0.0;
        int ti = 0;
        // 1/mass of molecule i, once: the per-pair a += f rij / m becomes a multiplication (exact
        // for unit masses, within an ulp otherwise)
        double inv_mi = 1.0;
        if (MIE) {
            ti = a.moltype[i];
            inv_mi = 1.0 / mtab[7 * T2 + ti];
        }
        const int exti = (mop || (RECORD && a.cv)) ? a.ext[i] : 0;
        const uint64_t* row = a.nbl + i;
        bool looped = false;  // the pairs went through the branch-free loop
        if constexpr (!MIE && !RECORD && !VIR) {
            lj_plain(r.pos, row, n, cap, xi, yi, zi, a.rc2, fx, fy, fz);
            looped = true;
        } else if constexpr (!MIE && RECORD && VIR && !EXTRA) {
            if (simple_rec) {  // PE and virial only
                lj_plain<true>(r.pos, row, n, cap, xi, yi, zi, a.rc2, fx, fy, fz, a.potshift, &pe, &vir);
                looped = true;
            }
        }
        // Four 16-bit entries per word.  The next word is loaded while this one is worked on,
        // and the four entries' positions are read from LDS together (the unused entries of a
        // last, partial word are 0, a valid region index).
        uint64_t wnext = n > 0 ? row[0] : 0;
        for (int kw = 0; !looped && kw < n; kw += 4) {
            const uint64_t word = wnext;
            if (kw + 4 < n) wnext = row[(size_t)((kw >> 2) + 1) * cap];
            int jq[4];
            double rxq[4], ryq[4], rzq[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                jq[q] = (int)((word >> (16 * q)) & 0xffffu);
                rxq[q] = xi - r.pos[3 * (jq[q])];
                ryq[q] = yi - r.pos[3 * (jq[q]) + 1];
                rzq[q] = zi - r.pos[3 * (jq[q]) + 2];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
            if (kw + q >= n) break;
            const int j = jq[q];
            const double rx = rxq[q], ry = ryq[q], rz = rzq[q];
            const double r2 = (rx * rx + ry * ry) + rz * rz;
            double rc2 = a.rc2;
            int t = 0;
            if (MIE) {
                t = ti * FM_MAX_TYPES + r.ltype[j];
                const double c = mtab[6 * T2 + t];
                if (c > 0.0) rc2 = c;
            }
            if (!(r2 < rc2)) continue;
            const double inv = recip(r2);
            double acc, e = 0.0;
            if (!MIE) {
                acc = 48.0 * (p7(inv) - 0.5 * p4(inv));
                if (RECORD) e = 4.0 * (p6(inv) - p3(inv)) - a.potshift;
            } else {
                const double eps = mtab[t], lr = mtab[T2 + t], la = mtab[2 * T2 + t], C = mtab[3 * T2 + t],
                             al = mtab[4 * T2 + t];
                acc = C * eps * (lr * mie_pow(inv, 0.5 * lr + 1) - al * la * mie_pow(inv, 0.5 * la + 1));
                if (RECORD) e = C * eps * (mie_pow(inv, 0.5 * lr) - al * mie_pow(inv, 0.5 * la)) - mtab[5 * T2 + t];
            }
            if (MIE) {
                fx = fx + acc * rx * inv_mi;
                fy = fy + acc * ry * inv_mi;
                fz = fz + acc * rz * inv_mi;
            } else {  // LJ mass is 1
                fx = fmaf64(acc, rx, fx);
                fy = fmaf64(acc, ry, fy);
                fz = fmaf64(acc, rz, fz);
            }
            if (VIR) vir = vir + acc * r2;
            if (RECORD) {
                pe = pe + e;
                if (do_rf) {
                    const double rr[3] = {rx, ry, rz};
                    for (int p = 0; p < 3; ++p)
                        for (int qq = 0; qq < 3; ++qq) rf[3 * p + qq] = rf[3 * p + qq] + acc * rr[p] * rr[qq];
                }
                if (mop || a.cv) {
                    const int rcj = region_cell_of(r, a.g.nreg, j);
                    if (mop_owner(r, a.g, rci, rcj, exti, r.lslot[j])) {
                        const double ri[3] = {xi, yi, zi}, rj[3] = {r.pos[3 * j], r.pos[3 * j + 1], r.pos[3 * j + 2]}, rij[3] = {rx, ry, rz};
                        if (mop) mop_entry(a, ri, rj, rij, acc, spxy);
                        // interior pairs carry 2 acc rij, entries against a periodic image acc rij
                        // (each such pair is met once from either end)
                        const bool img_entry = r.sh[rcj] != 13 || r.sh[rci] != 13;
                        if (a.cv) cv_entry(a, ri, rj, img_entry ? acc : 2.0 * acc, exti);
                    }
                }
            }
        }
        }
        if constexpr (EXTRA) bond_tether_forces<RECORD, VIR>(a, i, xi, yi, zi, inv_mi, do_rf, rf, fx, fy, fz, vir);
        if constexpr (P1) {  // (k_forces_brick_plain only: its one argument is the FbArgs)
            const FbArgs& b = fargs_again();
            vv_pass2_pass1<RECORD>(b, i, fx, fy, fz);
            if (VIR) b.vir[i] = vir;
            if (RECORD) b.pe_lj[i] = pe;
            continue;
        }
        a.s.ax[i] = fx;
        a.s.ay[i] = fy;
        a.s.az[i] = fz;
        if (a.kick) {
            a.s.vx[i] = a.s.vx[i] + a.hdt * fx;
            a.s.vy[i] = a.s.vy[i] + a.hdt * fy;
            a.s.vz[i] = a.s.vz[i] + a.hdt * fz;
        }
        if (VIR) a.vir[i] = vir;
        if (RECORD) {
            a.pe_lj[i] = pe;
            if (do_rf)
                for (int k = 0; k < 9; ++k) a.rf[i + k * cap] = rf[k];
        }
    }
    if (mop) {
        __syncthreads();
        for (int t = threadIdx.x; t < 3 * a.nplanes; t += blockDim.x)
            if (spxy[t] != 0.0) atomicAdd(&a.pxy[t], spxy[t]);
    }
}

// ---- Mie steps without tallies, and record steps of PE and virial -------------------------------
// The Mie pair law of forces_body's general loop over every list entry without branches, when each
// type pair's exponents lr/2 + 1 and la/2 + 1 are whole numbers below 2^NB.  Molecule i's row of
// the type tables (NT partner types) sits in registers: the pair cutoff, A = C eps lr and
// B = C eps alpha la, and the exponents; an entry picks them by its partner's type.  The powers
// x^n of x = 1/r^2 come from the binary expansion of n, with the multiplications of mie_pow; the
// magnitude is A x^nr - B x^na, and an entry beyond its pair's cutoff or past the end of the list
// gets zero.  The pair forces are summed per unit mass and divided by the mass at the end.  REC
// (record steps tallying PE and virial only): x from two Newton steps (the IEEE quotient, as the
// general loop), the PE term C eps (x^(lr/2) - alpha x^(la/2)) - potshift of the pair and the
// virial magnitude times r^2, both zero for an entry that does not count.  Plain steps take x
// from one Newton step.  NB = 0: the tables' usual class, la = 6 for every pair (x^4) and
// lr = 12..18 (x^7..x^10), where the powers are x^4 and x^4 x^3 x^(nr-7) (and x^3, x^6 x^(nr-7)
// for the PE): a few products instead of the bit expansions (within a few ulp of them).
template <int NT, int REC>
struct MieRow {
    double rc2[NT], A[NT], B[NT];
    int e[NT];  // lr/2 + 1 | (la/2 + 1) << 8 | (lr/2) << 16 | (la/2) << 24 (NB = 0: lr/2 - 6)
    double ce[REC ? NT : 1], al[REC ? NT : 1], sh[REC ? NT : 1];  // C eps, alpha, potshift
};

template <int NT, int NB, int REC>
__device__ __forceinline__ void mie_row(const FbArgs& a, int ti, MieRow<NT, REC>& m) {
    const int T2 = FM_MAX_TYPES * FM_MAX_TYPES;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int k = ti * FM_MAX_TYPES + t;
        const double eps = a.mie[k], lr = a.mie[T2 + k], la = a.mie[2 * T2 + k], C = a.mie[3 * T2 + k],
                     al = a.mie[4 * T2 + k], c = a.mie[6 * T2 + k];
        m.rc2[t] = c > 0.0 ? c : a.rc2;
        m.A[t] = (C * eps) * lr;
        m.B[t] = (C * eps) * (al * la);
        const int nr = (int)(0.5 * lr + 1.0), na = (int)(0.5 * la + 1.0);
        if constexpr (NB == 0) m.e[t] = nr - 7;  // (the host checks 7 <= nr <= 10 and na = 4)
        else m.e[t] = nr | na << 8 | (nr - 1) << 16 | (na - 1) << 24;
        if constexpr (REC) {
            m.ce[t] = C * eps;
            m.al[t] = al;
            m.sh[t] = a.mie[5 * T2 + k];
        }
    }
}

// v[tj] of a row array (types are below NT)
template <int NT, typename T>
__device__ __forceinline__ T mie_sel(const T (&v)[NT], int tj) {
    static_assert(NT == 2 || NT == 4, "two or four partner types");
    if constexpr (NT == 2) return tj != 0 ? v[1] : v[0];
    const bool s0 = (tj & 1) != 0, s1 = (tj & 2) != 0;
    // the values pass an empty asm, so that the selects are not turned into an indexed load (which
    // would keep the row in scratch memory)
    T v0 = v[0], v1 = v[1], v2 = v[2], v3 = v[3];
    asm("" : "+v"(v0), "+v"(v1), "+v"(v2), "+v"(v3));
    return s1 ? (s0 ? v3 : v2) : (s0 ? v1 : v0);
}

// x^n from pw[k] = x^(2^k) and the bits of n (mie_pow's products, in its order)
template <int NB>
__device__ __forceinline__ double mie_bits_pow(const double (&pw)[NB], int n) {
    double r = (n & 1) ? pw[0] : 1.0;
#pragma unroll
    for (int k = 1; k < NB; ++k) r = ((n >> k) & 1) ? r * pw[k] : r;
    return r;
}

struct P4M {
    double x[4], y[4], z[4];
    int t[4];
};
__device__ __forceinline__ void mie_stage(P4M& p, const double* pos, const uint8_t* lt, uint64_t w) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int j = (int)((w >> (16 * q)) & 0xffffu);
        p.x[q] = pos[3 * j];
        p.y[q] = pos[3 * j + 1];
        p.z[q] = pos[3 * j + 2];
        p.t[q] = lt[j];
    }
}
template <int NT, int NB, bool REC>
__device__ __forceinline__ void mie_word(const P4M& p, int left, double xi, double yi, double zi,
                                         const MieRow<NT, REC>& m, double& fx, double& fy, double& fz, double& pe,
                                         double& vt) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double rx = xi - p.x[q], ry = yi - p.y[q], rz = zi - p.z[q];
        const double r2 = (rx * rx + ry * ry) + rz * rz;
        const int tj = p.t[q];
        const int e = mie_sel<NT>(m.e, tj);
        const double x = REC ? recip(r2) : recip1(r2);
        // x^nr, x^na, and on record steps x^(nr-1), x^(na-1)
        double pr, pa, ur = 0.0, ua = 0.0;
        if constexpr (NB == 0) {
            const double x2 = x * x, x3 = x2 * x, x4 = x2 * x2;
            const double xd = ((e & 1) ? x : 1.0) * ((e & 2) ? x2 : 1.0);  // x^(nr-7)
            pr = (x4 * x3) * xd;
            pa = x4;
            if constexpr (REC) {
                ur = (x3 * x3) * xd;
                ua = x3;
            }
        } else {
            double pw[NB > 0 ? NB : 1];
            pw[0] = x;
#pragma unroll
            for (int k = 1; k < NB; ++k) pw[k] = pw[k - 1] * pw[k - 1];
            pr = mie_bits_pow<NB>(pw, e);
            pa = mie_bits_pow<NB>(pw, e >> 8);
            if constexpr (REC) {
                ur = mie_bits_pow<NB>(pw, e >> 16);
                ua = mie_bits_pow<NB>(pw, e >> 24);
            }
        }
        const double acc = fmaf64(mie_sel<NT>(m.A, tj), pr, -(mie_sel<NT>(m.B, tj) * pa));
        // a select, not a product: a padding entry may sit at r = 0 (an infinite or NaN acc)
        const bool in = r2 < mie_sel<NT>(m.rc2, tj) && q < left;
        const double g = in ? acc : 0.0;
        fx = fmaf64(g, rx, fx);
        fy = fmaf64(g, ry, fy);
        fz = fmaf64(g, rz, fz);
        if constexpr (REC) {
            vt = fmaf64(g, r2, vt);
            const double u = mie_sel<NT>(m.ce, tj) * fmaf64(-mie_sel<NT>(m.al, tj), ua, ur) - mie_sel<NT>(m.sh, tj);
            pe = pe + (in ? u : 0.0);
        }
    }
}

// The list loop of lj_plain (two words in flight, the word after them loading) over mie_word.
template <int NT, int NB, bool REC>
__device__ __forceinline__ void mie_plain(const double* pos, const uint8_t* lt, const uint64_t* row, int n,
                                          size_t cap, double xi, double yi, double zi, const MieRow<NT, REC>& m,
                                          double& fx, double& fy, double& fz, double& pe, double& vt) {
    double sx = 0, sy = 0, sz = 0, spe = 0, svt = 0;
    P4M pa, pb;
    const int lastw = n > 0 ? (n - 1) >> 2 : 0;
    const uint64_t w0 = row[0];
    mie_stage(pa, pos, lt, n > 0 ? w0 : 0);
    uint64_t w1 = row[(size_t)min(1, lastw) * cap];
    uint64_t w2 = row[(size_t)min(2, lastw) * cap];
    for (int kw = 0; kw < n; kw += 8) {
        const int k = kw >> 2;
        mie_stage(pb, pos, lt, kw + 4 < n ? w1 : 0);
        const uint64_t wk2 = w2;
        w1 = row[(size_t)min(k + 3, lastw) * cap];
        w2 = row[(size_t)min(k + 4, lastw) * cap];
        __builtin_amdgcn_sched_barrier(0);
        mie_word<NT, NB, REC>(pa, n - kw, xi, yi, zi, m, sx, sy, sz, spe, svt);
        mie_stage(pa, pos, lt, kw + 8 < n ? wk2 : 0);
        __builtin_amdgcn_sched_barrier(0);
        if (__builtin_amdgcn_ballot_w64(kw + 4 < n) != 0ull)
            mie_word<NT, NB, REC>(pb, n - kw - 4, xi, yi, zi, m, sx, sy, sz, spe, svt);
    }
    fx = sx;
    fy = sy;
    fz = sz;
    pe = spe;
    vt = svt;
}

// LDS of the branch-free Mie launch: positions, region cell and one-byte type of each staged
// molecule, then the region tables (no slots; the type rows are read from HBM once per molecule)
__device__ inline void carve_mie(Region& r, char* smem, const BG& g, int maxreg) {
    r.pos = (double*)smem;
    int* ip = (int*)(r.pos + 3 * (size_t)maxreg);
    r.lslot = nullptr;
    r.ltype = nullptr;
    r.rq = (uint16_t*)ip;
    ip += (maxreg + 1) / 2;
    r.ltype8 = (uint8_t*)ip;
    ip += (maxreg + 3) / 4;
    r.off = ip;
    ip += g.nreg + 1;
    r.gc = ip;
    ip += g.nreg;
    r.sh = ip;
    ip += g.nreg;
    r.cst = ip;
    ip += g.nreg;
    const int nbmax = g.B[0] * g.B[1] * g.B[2];
    r.ic_off = ip;
    ip += nbmax + 1;
    r.ic_rc = ip;
}
__host__ __device__ inline size_t mie_lds_bytes(const BG& g, int maxreg) {
    const int nbmax = g.B[0] * g.B[1] * g.B[2];
    const size_t b = (size_t)maxreg * 3 * sizeof(double) +
                     (size_t)((maxreg + 1) / 2 + (maxreg + 3) / 4 + 4 * g.nreg + 1 + 2 * nbmax + 1) * sizeof(int);
    return (b + 15) & ~(size_t)15;
}

// Forces of a Mie step without tallies (REC: a record step tallying PE and virial only), bonds and
// tethers added to the pair sum after it; the kick of a fused simulation_move_particles_vv(2) as
// forces_body does it, or with P1 pass 2 of this step and pass 1 of the next (vv_pass2_pass1).
// (KA: a is the launch's argument block, read again after the pair loop; packed launches pass
// false: a is a domain's descriptor in device memory)
template <int NT, int NB, bool EXTRA, bool REC, bool P1, bool KA = true>
__device__ __forceinline__ void forces_body_mie(const FbArgs& a, int brick, char* smem) {
    Region r;
    carve_mie(r, smem, a.g, a.maxreg);
    // staged from the list build's table when there is one
    const int* tab = a.btab ? a.btab + (size_t)brick * a.btab_stride : nullptr;
    if (tab) {
        if (!stage_from_table(r, a.g, a.s, tab, a.maxreg, a.moltype)) return;
    } else if (!stage_region(r, a.g, brick, a.start, a.s, true, nullptr, a.cap, nullptr, nullptr, a.maxreg, a.flags,
                             a.moltype, nullptr))
        return;
    const int T2 = FM_MAX_TYPES * FM_MAX_TYPES;
    if constexpr (EXTRA) {
        // bonds and tethers in a pass of their own (each thread its own molecules), their force
        // (and bond virial) parked in this launch's output: what they read is then not held
        // through the pair loops
        for (int q = threadIdx.x; q < r.ni; q += blockDim.x) {
            int li, i;
            imol_tab(r, a.start, tab, a.maxreg, q, li, i);
            const double inv_mi = 1.0 / a.mie[7 * T2 + a.moltype[i]];
            double bx = 0.0, by = 0.0, bz = 0.0, vir = 0.0;
            bond_tether_forces<REC, REC>(a, i, r.pos[3 * li], r.pos[3 * li + 1], r.pos[3 * li + 2], inv_mi, false,
                                         nullptr, bx, by, bz, vir);
            (P1 ? a.s2.ax : a.s.ax)[i] = bx;  // where this launch leaves the force
            (P1 ? a.s2.ay : a.s.ay)[i] = by;
            (P1 ? a.s2.az : a.s.az)[i] = bz;
            if constexpr (REC) a.vir[i] = vir;
        }
    }
    for (int q = threadIdx.x; q < r.ni; q += blockDim.x) {
        int li, i;
        imol_tab(r, a.start, tab, a.maxreg, q, li, i);
        const double xi = r.pos[3 * li], yi = r.pos[3 * li + 1], zi = r.pos[3 * li + 2];
        const int ti = a.moltype[i];
        MieRow<NT, REC> m;
        mie_row<NT, NB, REC>(a, ti, m);
        const double inv_mi = 1.0 / a.mie[7 * T2 + ti];
        double fx, fy, fz, pe, vir;
        mie_plain<NT, NB, REC>(r.pos, r.ltype8, a.nbl + i, a.nbn[i], a.cap, xi, yi, zi, m, fx, fy, fz, pe, vir);
        const FbArgs& b = KA ? fargs_again() : a;  // the arguments below: not held through the pair loop
        fx = fx * inv_mi;
        fy = fy * inv_mi;
        fz = fz * inv_mi;
        if constexpr (EXTRA) {
            fx = fx + (P1 ? b.s2.ax : b.s.ax)[i];
            fy = fy + (P1 ? b.s2.ay : b.s.ay)[i];
            fz = fz + (P1 ? b.s2.az : b.s.az)[i];
            if constexpr (REC) vir = vir + b.vir[i];
        }
        if constexpr (REC) {
            b.vir[i] = vir;
            b.pe_lj[i] = pe;
        }
        if constexpr (P1) {
            vv_pass2_pass1<REC>(b, i, fx, fy, fz);
            continue;
        }
        b.s.ax[i] = fx;
        b.s.ay[i] = fy;
        b.s.az[i] = fz;
        if (b.kick) {
            b.s.vx[i] = b.s.vx[i] + b.hdt * fx;
            b.s.vy[i] = b.s.vy[i] + b.hdt * fy;
            b.s.vz[i] = b.s.vz[i] + b.hdt * fz;
        }
    }
}

// pressure_tensor_forces_MOP for one list entry of the record pass below, with the same
// contributions as mop_entry.  The reference loops over the planes between the ones nearest to
// ri(2) and rj(2); only planes strictly between the two molecules add (the sign factor is zero
// elsewhere).  With t = (y + halfdomain - shift)/planespacing, plane n sits at t = n - 1, so the
// candidates here are k0 <= n - 1 <= k1, the whole numbers between t_i and t_j widened by 1e-9
// (the t's come from a multiplication, within 1e-14 of the quotient), and the sign factor decides
// exactly as the reference's.  The intercept factors are 1 whenever both molecules lie strictly
// inside the box in x and z (the intercept lies between them); otherwise they are computed with
// the reference's divisions.
__device__ inline void mop_entry_fast(const FbArgs& a, const double ri[3], const double rj[3], const double rij[3],
                                      double acc, double* spxy, const double* planes, int k0, int k1, bool inside) {
    const int n0 = k0 < 0 ? 1 : k0 + 1, n1 = k1 >= a.nplanes ? a.nplanes : k1 + 1;
    for (int n = n0; n <= n1; ++n) {
        const double plane = planes[n - 1];
        const double fx = copysign(1.0, ri[1] - plane) - copysign(1.0, rj[1] - plane);
        if (fx == 0.0) continue;
        double hxx = 1.0, hzz = 1.0;
        if (!inside) {
            const double pyb0 = ri[0] + (rij[0] / rij[1]) * (plane - ri[1]);
            const double pyb2 = ri[2] + (rij[2] / rij[1]) * (plane - ri[1]);
            hxx = dp_heaviside(a.hx - pyb0) - dp_heaviside(-a.hx - pyb0);
            hzz = dp_heaviside(a.hz - pyb2) - dp_heaviside(-a.hz - pyb2);
            if (hxx == 0.0 || hzz == 0.0) continue;
        }
        for (int d = 0; d < 3; ++d) atomicAdd(&spxy[3 * (n - 1) + d], acc * rij[d] * fx * hxx * hzz);
    }
}

// Record step of an LJ domain with the MOP tally and nothing else (no Mie tables, bonds, tethers
// in the forces, CV faces or rfmol), in FP64 for both precisions (every pair term the tallies need
// is an FP64 term, so the FP32 pair law would only add work).  One pass over the list: the forces,
// PE and virial as the branch-free record loop computes them, and per entry the MOP test; the
// entries inside rc that cross a plane and that this end tallies add their contribution, in FP64
// as the reference's loop computes it (src/simulation_compute_forces.f90:683-701).
//
// Which end tallies: the reference visits every pair with both molecules in the domain once (its
// half list) and every pair with a halo image from the molecule's side (so twice, once per end).
// A pair's contribution is symmetric in its two ends (r_ij and the sign factor both change sign,
// the crossed planes and the intercept are the same), so any rule that picks one end of each
// unshifted pair tallies the same sum: the end with the higher molecule number.  Each staged
// molecule's key is its molecule number << 1 | (its region cell is an image).
struct P4K {
    double x[4], y[4], z[4];
    int k[4];
};
__device__ __forceinline__ void lj_stage_k(P4K& p, const double* pos, const int* key, uint64_t w) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int j = (int)((w >> (16 * q)) & 0xffffu);
        p.x[q] = pos[3 * j];
        p.y[q] = pos[3 * j + 1];
        p.z[q] = pos[3 * j + 2];
        p.k[q] = key[j];
    }
}
struct MopI {             // the i-molecule's side of the MOP test
    double ti;            // (y_i + halfdomain - shift) / planespacing
    double hps, inv_sp;   // halfdomain - shift, 1 / planespacing
    double bx, bz;        // the box's half widths in x and z, less 1e-12 relative
    int keyi, nplanes;
    bool in_i;            // molecule i strictly inside the box in x and z
    double* spxy;
    const double* planes;
    // SLOTS: the (at most three) planes a pair of molecule i can cross, kb .. kb+2 (0-based plane
    // numbers; a slot outside 0 .. nplanes-1 holds 1e300, where no pair crosses), their y, the
    // sign of y_i against each, and the thread's sums of the contributions to each
    int kb, sneg;          // sneg bit s: y_i lies below slot s's plane
    double pl[3];
    double c[3][3];
};
// SLOTS (when 2 rc < 3 planespacing, so a pair inside rc crosses at most three planes near
// molecule i): the MOP tally without branches.  For each slot the reference's sign factor
// sign(y_i - plane) - sign(y_j - plane) is zero unless the plane lies between the two molecules,
// so every entry adds (acc r_ij) x (sign factor) to each slot's sums, with acc zero for an entry
// beyond rc or not tallied from this end; the entries with a molecule near the box's x or z
// faces (whose intercept factors need the reference's divisions) go through mop_entry_fast.
template <bool SLOTS, int NQ = 4>
__device__ __forceinline__ void lj_word_mop(const FbArgs& a, const P4K& p, int left, double xi, double yi,
                                            double zi, MopI& m, double& fx, double& fy, double& fz,
                                            double& pe, double& vt) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double rx = xi - p.x[q], ry = yi - p.y[q], rz = zi - p.z[q];
        const double r2 = (rx * rx + ry * ry) + rz * rz;
        const bool in = r2 < a.rc2 && q < left;
        const double xr = recip(r2);
        const double x = in ? xr : __hiloint2double(0, __double2loint(xr));
        const double x2 = x * x;
        const double x3 = x * x2;
        const double x4 = x2 * x2;
        const double t = x4 * fmaf64(2.0, x3, -1.0);
        fx = fmaf64(t, rx, fx);
        fy = fmaf64(t, ry, fy);
        fz = fmaf64(t, rz, fz);
        vt = fmaf64(t, r2, vt);
        const double e = 4.0 * (x3 * x3 - x3) - a.potshift;
        pe = pe + (in ? e : 0.0);
        const int kj = p.k[q];
        const bool own = ((m.keyi | kj) & 1) != 0 || (m.keyi >> 1) > (kj >> 1);
        if constexpr (SLOTS) {
            const bool ins = m.in_i && fabs(p.x[q]) < m.bx && fabs(p.z[q]) < m.bz;
            // (t is zero for an entry beyond rc)
            const double af = (own && ins) ? 24.0 * t : 0.0;
            const double v[3] = {af * rx, af * ry, af * rz};
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const double f = (((m.sneg >> s) & 1) ? -1.0 : 1.0) - copysign(1.0, p.y[q] - m.pl[s]);
#pragma unroll
                for (int d = 0; d < 3; ++d) m.c[s][d] = fmaf64(v[d], f, m.c[s][d]);
            }
            if (!(in && own && !ins)) continue;
        }
        // the whole numbers strictly between t_i and t_j, widened (see mop_entry_fast)
        const double tj = (p.y[q] + m.hps) * m.inv_sp;
        const int k0 = (int)ceil(fmin(m.ti, tj) - 1e-9), k1 = (int)floor(fmax(m.ti, tj) + 1e-9);
        if (in && own && k0 <= k1 && k1 >= 0 && k0 < m.nplanes) {
            // the pair law's magnitude 48 (x^7 - x^4/2) as the force sum takes it (within 1e-15)
            const double acc = 24.0 * t;
            const double ri[3] = {xi, yi, zi}, rj[3] = {p.x[q], p.y[q], p.z[q]}, rij[3] = {rx, ry, rz};
            mop_entry_fast(a, ri, rj, rij, acc, m.spxy, m.planes, k0, k1,
                           m.in_i && fabs(p.x[q]) < m.bx && fabs(p.z[q]) < m.bz);
        }
    }
}

// The list loop with the MOP test per entry: one word at a time, the next word's index in flight.
// With SLOTS the word's entries go one at a time (the nine plane sums hold registers that four
// staged entries would need).
template <bool SLOTS>
__device__ __forceinline__ void lj_plain_mop(const FbArgs& a, const double* pos, const int* key,
                                             const uint64_t* row, int n, size_t cap, double xi, double yi,
                                             double zi, MopI& m, double& fx, double& fy, double& fz,
                                             double& pe_out, double& vir_out) {
    double sx = 0, sy = 0, sz = 0, pe = 0.0, vt = 0.0;
    uint64_t wn = row[0];
    for (int kw = 0; kw < n; kw += 4) {
        const uint64_t w = wn;
        if (kw + 4 < n) wn = row[(size_t)((kw >> 2) + 1) * cap];
        if constexpr (SLOTS) {
#pragma unroll 1
            for (int q = 0; q < 4; ++q) {
                P4K p;
                const int j = (int)((w >> (16 * q)) & 0xffffu);
                p.x[0] = pos[3 * j];
                p.y[0] = pos[3 * j + 1];
                p.z[0] = pos[3 * j + 2];
                p.k[0] = key[j];
                lj_word_mop<true, 1>(a, p, n - kw - q, xi, yi, zi, m, sx, sy, sz, pe, vt);
            }
        } else {
            P4K p;
            lj_stage_k(p, pos, key, w);
            lj_word_mop<SLOTS>(a, p, n - kw, xi, yi, zi, m, sx, sy, sz, pe, vt);
        }
    }
    fx = 24.0 * sx;
    fy = 24.0 * sy;
    fz = 24.0 * sz;
    pe_out = pe;
    vir_out = 24.0 * vt;
}

template <bool SLOTS>
__device__ __forceinline__ void forces_body_mop(const FbArgs& a, int brick, char* smem) {
    Region r;
    carve(r, smem, a.g, a.maxreg, true);
    // after the region: the plane tallies and the plane positions
    double* spxy = (double*)(smem + region_lds_bytes(a.g, a.maxreg, true));
    double* planes = spxy + 3 * a.nplanes;
    for (int t = threadIdx.x; t < a.nplanes; t += blockDim.x) planes[t] = a.planes[t];
    for (int t = threadIdx.x; t < 3 * a.nplanes; t += blockDim.x) spxy[t] = 0.0;
    // staged from the list build's table when there is one, with the owner keys
    const int* tab = a.btab ? a.btab + (size_t)brick * a.btab_stride : nullptr;
    if (tab) {
        if (!stage_from_table(r, a.g, a.s, tab, a.maxreg, nullptr, a.ext)) return;
    } else {
        if (!stage_region(r, a.g, brick, a.start, a.s, true, nullptr, a.cap, nullptr, nullptr, a.maxreg, a.flags,
                          nullptr, a.ext))
            return;
        // the staged molecule numbers become the owner keys
        for (int t = threadIdx.x; t < r.total; t += blockDim.x) r.lslot[t] = (r.lslot[t] << 1) | (r.sh[r.rq[t]] != 13);
        __syncthreads();
    }
    const size_t cap = a.cap;
    MopI m;
    m.hps = a.hp - a.shift;
    m.inv_sp = 1.0 / a.spacing;
    m.bx = a.hx * (1.0 - 1e-12);
    m.bz = a.hz * (1.0 - 1e-12);
    m.nplanes = a.nplanes;
    m.spxy = spxy;
    m.planes = planes;
    const double rct = sqrt(a.rc2) * m.inv_sp;  // rc in plane spacings
    for (int q = threadIdx.x; q < r.ni; q += blockDim.x) {
        int li, i;
        imol_tab(r, a.start, tab, a.maxreg, q, li, i);
        const double xi = r.pos[3 * li], yi = r.pos[3 * li + 1], zi = r.pos[3 * li + 2];
        m.ti = (yi + m.hps) * m.inv_sp;
        m.keyi = r.lslot[li];
        m.in_i = fabs(xi) < m.bx && fabs(zi) < m.bz;
        if constexpr (SLOTS) {
            // the planes strictly within rc of y_i: t in (t_i - rc/spacing, t_i + rc/spacing)
            m.kb = (int)floor(m.ti - rct - 1e-9) + 1;
            m.sneg = 0;
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const int k = m.kb + s;
                m.pl[s] = (k >= 0 && k < a.nplanes) ? planes[k] : 1e300;
                m.sneg |= (copysign(1.0, yi - m.pl[s]) < 0.0 ? 1 : 0) << s;
                m.c[s][0] = m.c[s][1] = m.c[s][2] = 0.0;
            }
        }
        double fx, fy, fz, pe, vir;
        lj_plain_mop<SLOTS>(a, r.pos, r.lslot, a.nbl + i, a.nbn[i], cap, xi, yi, zi, m, fx, fy, fz, pe, vir);
        if constexpr (SLOTS) {
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const int k = m.kb + s;
                if (k >= 0 && k < a.nplanes)
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        if (m.c[s][d] != 0.0) atomicAdd(&spxy[3 * k + d], m.c[s][d]);
            }
        }
        a.s.ax[i] = fx;
        a.s.ay[i] = fy;
        a.s.az[i] = fz;
        if (a.kick) {
            a.s.vx[i] = a.s.vx[i] + a.hdt * fx;
            a.s.vy[i] = a.s.vy[i] + a.hdt * fy;
            a.s.vz[i] = a.s.vz[i] + a.hdt * fz;
        }
        a.vir[i] = vir;
        a.pe_lj[i] = pe;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * a.nplanes; t += blockDim.x)
        if (spxy[t] != 0.0) atomicAdd(&a.pxy[t], spxy[t]);
}

// ---- mixed precision staged in FP32 --------------------------------------------------------------
// The mixed mode's plain steps (no tallies, bonds or tethers) stage each molecule as a float4:
// x y z relative to the region's lower corner in FP32, and the molecule's slot in w.  That is 16
// B instead of 24 B (one LDS read per entry instead of two), so three workgroups share a CU.  The
// pair law and the per-molecule sum are FP32 as in the FP64-staged mixed loop.  Whether a pair
// counts is decided on the FP32 r^2, except within band32 of rc^2, where the FP64 r^2 of the two
// molecules' positions in HBM decides, in the FP64 path's operation order: band32 bounds the
// difference between the FP32 and FP64 r^2 (the rounding of the staged offsets, of their
// differences and of the sum), so every pair counts exactly when the FP64 path counts it.
__device__ inline void carve32(Region& r, float4*& p4, char* smem, const BG& g, int maxreg) {
    p4 = (float4*)smem;
    int* ip = (int*)(p4 + maxreg);
    r.pos = nullptr;
    r.lslot = nullptr;
    r.ltype = nullptr;
    r.rq = (uint16_t*)ip;
    ip += (maxreg + 1) / 2;
    r.off = ip;
    ip += g.nreg + 1;
    r.gc = ip;
    ip += g.nreg;
    r.sh = ip;
    ip += g.nreg;
    r.cst = ip;
    ip += g.nreg;
    const int nbmax = g.B[0] * g.B[1] * g.B[2];
    r.ic_off = ip;
    ip += nbmax + 1;
    r.ic_rc = ip;
}

// staged offsets and slots of the region from the brick's table (stage_from_table); o: the
// region's lower corner
__device__ bool stage_table32(Region& r, float4* p4, const BG& g, const St& s, const int* T, int capm,
                              const double o[3]) {
    r.total = T[0];
    r.ni = T[1];
    if (r.total > capm) return false;  // (a build that met a larger region halted its loop)
    const int* stg = T + 2;
    constexpr int SU = 6;
    for (int q0 = threadIdx.x; q0 < r.total; q0 += SU * blockDim.x) {
        int w[SU];
        double px[SU], py[SU], pz[SU];
#pragma unroll
        for (int k = 0; k < SU; ++k) w[k] = stg[min(q0 + k * (int)blockDim.x, r.total - 1)];
#pragma unroll
        for (int k = 0; k < SU; ++k) {
            const int sl = w[k] >> 5;
            px[k] = s.x[sl];
            py[k] = s.y[sl];
            pz[k] = s.z[sl];
        }
#pragma unroll
        for (int k = 0; k < SU; ++k) {
            const int q = min(q0 + k * (int)blockDim.x, r.total - 1);
            int sx, sy, sz;
            decode_shift(w[k] & 31, sx, sy, sz);
            p4[q] = make_float4((float)(img(px[k], sx, g.L[0]) - o[0]), (float)(img(py[k], sy, g.L[1]) - o[1]),
                                (float)(img(pz[k], sz, g.L[2]) - o[2]), __int_as_float(w[k] >> 5));
        }
    }
    __syncthreads();
    return true;
}

// staged offsets and slots of the region of `brick`; o: the region's lower corner
__device__ bool stage_region32(Region& r, float4* p4, const BG& g, int brick, const int* start, const St& s,
                               int capm, int* flags, const double o[3]) {
    if (!stage_tables(r, g, brick, start, capm, flags)) return false;
    constexpr int SU = 6;
    for (int q0 = threadIdx.x; q0 < r.total; q0 += SU * blockDim.x) {
        int sl[SU], sc[SU];
        double px[SU], py[SU], pz[SU];
#pragma unroll
        for (int k = 0; k < SU; ++k) {
            const int qc = min(q0 + k * (int)blockDim.x, r.total - 1);
            const int rc = r.rq[qc];
            sl[k] = r.cst[rc] + (qc - r.off[rc]);
            sc[k] = r.sh[rc];
            px[k] = s.x[sl[k]];
            py[k] = s.y[sl[k]];
            pz[k] = s.z[sl[k]];
        }
#pragma unroll
        for (int k = 0; k < SU; ++k) {
            const int q = min(q0 + k * (int)blockDim.x, r.total - 1);
            int sx, sy, sz;
            decode_shift(sc[k], sx, sy, sz);
            p4[q] = make_float4((float)(img(px[k], sx, g.L[0]) - o[0]), (float)(img(py[k], sy, g.L[1]) - o[1]),
                                (float)(img(pz[k], sz, g.L[2]) - o[2]), __int_as_float(sl[k]));
        }
    }
    __syncthreads();
    return true;
}

struct P4F {
    float x[4], y[4], z[4];
    int s[4];
};
__device__ __forceinline__ void lj_stage32(P4F& p, const float4* p4, uint64_t w) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 v = p4[(int)((w >> (16 * q)) & 0xffffu)];
        p.x[q] = v.x;
        p.y[q] = v.y;
        p.z[q] = v.z;
        p.s[q] = __float_as_int(v.w);
    }
}

// the FP64 path's cutoff test for molecule i (interior: its stored position is its staged one)
// and region entry j (slot sj), from their positions in HBM
__device__ __forceinline__ bool in_rc_fp64(const FbArgs& a, const Region& r, int i, int j, int sj, const int* tab) {
    int sx, sy, sz;
    decode_shift(tab ? tab[2 + j] & 31 : r.sh[r.rq[j]], sx, sy, sz);
    const double dx = a.s.x[i] - img(a.s.x[sj], sx, a.g.L[0]), dy = a.s.y[i] - img(a.s.y[sj], sy, a.g.L[1]),
                 dz = a.s.z[i] - img(a.s.z[sj], sz, a.g.L[2]);
    return ((dx * dx + dy * dy) + dz * dz) < a.rc2;
}

// One entry's FP32 pair term (the same arithmetic in the list loop and for a band entry).  REC:
// also the PE term 4 (x^6 - x^3) - potshift in FP64 from the FP32 x, and the virial term
// x^4 (2 x^3 - 1) r^2 (scaled by 24 with the force), summed in FP64; an entry that does not
// count adds exactly zero.
template <bool REC = false>
__device__ __forceinline__ void pair32(float dx, float dy, float dz, bool in, float& fx, float& fy, float& fz,
                                       double potshift = 0.0, double* pe = nullptr, double* vt = nullptr) {
    // the FP32 operations spelled out (explicit fused multiply-adds), so that the one-domain and
    // the packed launches, and the loop and the band entries, round alike
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    float x = __builtin_amdgcn_rcpf(r2);
    x = in ? x : 0.0f;
    const float x2 = x * x;
    const float acc = (x2 * x2) * fmaf(2.0f, x2 * x, -1.0f);
    fx = fmaf(acc, dx, fx);
    fy = fmaf(acc, dy, fy);
    fz = fmaf(acc, dz, fz);
    if constexpr (REC) {
        const double xd = (double)x;
        *pe = *pe + (in ? 4.0 * (p6(xd) - p3(xd)) - potshift : 0.0);
        *vt = *vt + (double)acc * (double)r2;
    }
}
__device__ __forceinline__ float r2_32(float dx, float dy, float dz) { return fmaf(dz, dz, fmaf(dy, dy, dx * dx)); }

// A band entry is left out of the loop's sum and its list position (12 bits) kept: up to four
// per molecule in band (count in nb); the loop itself has no branch.
template <bool REC>
__device__ __forceinline__ void lj_word32(const P4F& p, int kw, int left, float xi, float yi, float zi, float rc2f,
                                          float band, float& fx, float& fy, float& fz, uint32_t& bl0,
                                          uint32_t& bl1, int& nb, double potshift, double* pe, double* vt) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float dx = xi - p.x[q], dy = yi - p.y[q], dz = zi - p.z[q];
        const float r2 = r2_32(dx, dy, dz);
        const bool inb = fabsf(r2 - rc2f) < band && q < left;
        const bool in = r2 < rc2f && q < left && !inb;
        pair32<REC>(dx, dy, dz, in, fx, fy, fz, potshift, pe, vt);
        const uint32_t e = (uint32_t)(kw + q);
        bl1 = inb ? (bl1 << 12) | (bl0 >> 20) : bl1;
        bl0 = inb ? (bl0 << 12) | e : bl0;
        nb += inb ? 1 : 0;
    }
}

// the list loop of lj_plain (two words in flight) over the FP32 staging, then the band entries
// (REC: with the PE and virial sums)
// (tab: the brick's staging table when the region was staged from it, for the image shifts)
template <bool REC>
__device__ __forceinline__ void lj_plain32(const FbArgs& a, const Region& r, const int* tab, const float4* p4, int i,
                                           const uint64_t* row, int n, size_t cap, float xi, float yi, float zi,
                                           double& fx, double& fy, double& fz, double& pe, double& vir) {
    float sx = 0.f, sy = 0.f, sz = 0.f;
    double vt = 0.0;
    pe = 0.0;
    const float rc2f = (float)a.rc2, band = a.band32;
    uint32_t bl0 = 0, bl1 = 0;
    int nb = 0;
    P4F pa, pb;
    const int lastw = n > 0 ? (n - 1) >> 2 : 0;
    uint64_t w0 = row[0];
    uint64_t wa = n > 0 ? w0 : 0;
    lj_stage32(pa, p4, wa);
    uint64_t w1 = row[(size_t)min(1, lastw) * cap];
    uint64_t w2 = row[(size_t)min(2, lastw) * cap];
    for (int kw = 0; kw < n; kw += 8) {
        const int k = kw >> 2;
        const uint64_t wb = kw + 4 < n ? w1 : 0;
        lj_stage32(pb, p4, wb);
        const uint64_t wk2 = w2;
        w1 = row[(size_t)min(k + 3, lastw) * cap];
        w2 = row[(size_t)min(k + 4, lastw) * cap];
        __builtin_amdgcn_sched_barrier(0);
        lj_word32<REC>(pa, kw, n - kw, xi, yi, zi, rc2f, band, sx, sy, sz, bl0, bl1, nb, a.potshift, &pe, &vt);
        wa = kw + 8 < n ? wk2 : 0;
        lj_stage32(pa, p4, wa);
        __builtin_amdgcn_sched_barrier(0);
        if (__builtin_amdgcn_ballot_w64(kw + 4 < n) != 0ull)
            lj_word32<REC>(pb, kw + 4, n - kw - 4, xi, yi, zi, rc2f, band, sx, sy, sz, bl0, bl1, nb, a.potshift, &pe,
                           &vt);
    }
    if (nb > 0) {
        // the band entries, each decided on the FP64 r^2 (all of the list again in the improbable
        // case of more than four: every entry in band re-decided, in list order)
        const bool all = nb > 4 || n > 4096;  // (positions are kept in 12 bits)
        const int cnt = all ? n : nb;
        for (int b = 0; b < cnt; ++b) {
            int e;
            if (all) {
                e = b;
            } else {
                const uint64_t packed = ((uint64_t)bl1 << 32) | bl0;
                e = (int)((packed >> (12 * b)) & 0xfffu);
            }
            const int j = (int)((row[(size_t)(e >> 2) * cap] >> (16 * (e & 3))) & 0xffffu);
            const float4 v = p4[j];
            const float dx = xi - v.x, dy = yi - v.y, dz = zi - v.z;
            if (!(fabsf(r2_32(dx, dy, dz) - rc2f) < band)) continue;
            pair32<REC>(dx, dy, dz, in_rc_fp64(a, r, i, j, __float_as_int(v.w), tab), sx, sy, sz, a.potshift, &pe, &vt);
        }
    }
    fx = 24.0 * (double)sx;
    fy = 24.0 * (double)sy;
    fz = 24.0 * (double)sz;
    vir = 24.0 * vt;
}

// P1: VV pass 2 of this step, pass 1 of the next one and its displacement update follow the force,
// written to the other state buffer (the arithmetic of forces_body's P1 branch).  REC: a record
// step tallying PE and virial only.
template <bool P1, bool REC = false, bool TAB = false>
__device__ __forceinline__ void forces_body32(const FbArgs& a, int brick, char* smem) {
    Region r;
    float4* p4;
    carve32(r, p4, smem, a.g, a.maxreg);
    const int bxi = brick % a.g.nb[0], byi = (brick / a.g.nb[0]) % a.g.nb[1], bzi = brick / (a.g.nb[0] * a.g.nb[1]);
    const double o[3] = {-0.5 * a.g.L[0] + (bxi * a.g.B[0] - 1) * (a.g.L[0] / a.g.nc[0]),
                         -0.5 * a.g.L[1] + (byi * a.g.B[1] - 1) * (a.g.L[1] / a.g.nc[1]),
                         -0.5 * a.g.L[2] + (bzi * a.g.B[2] - 1) * (a.g.L[2] / a.g.nc[2])};
    // TAB (the one-domain launch): stage from the list build's table when there is one
    const int* tab = TAB && a.btab ? a.btab + (size_t)brick * a.btab_stride : nullptr;
    if (tab) {
        if (!stage_table32(r, p4, a.g, a.s, tab, a.maxreg, o)) return;
    } else if (!stage_region32(r, p4, a.g, brick, a.start, a.s, a.maxreg, a.flags, o))
        return;
    const size_t cap = a.cap;
    for (int q = threadIdx.x; q < r.ni; q += blockDim.x) {
        int rci = 0, li, i;
        if (tab) {
            i = tab[2 + a.maxreg + q];
            li = ((const uint16_t*)(tab + 2 + 2 * a.maxreg))[q];
        } else {
            imol(r, a.start, q, rci, li, i);
        }
        const float4 me = p4[li];
        double fx, fy, fz, pe, vir;
        lj_plain32<REC>(a, r, tab, p4, i, a.nbl + i, a.nbn[i], cap, me.x, me.y, me.z, fx, fy, fz, pe, vir);
        if constexpr (REC) {
            a.vir[i] = vir;
            a.pe_lj[i] = pe;
        }
        if constexpr (P1) {  // (k_forces_brick_plain32 only: its one argument is the FbArgs)
            vv_pass2_pass1<REC>(fargs_again(), i, fx, fy, fz);
            continue;
        }
        a.s.ax[i] = fx;
        a.s.ay[i] = fy;
        a.s.az[i] = fz;
        if (a.kick) {
            a.s.vx[i] = a.s.vx[i] + a.hdt * fx;
            a.s.vy[i] = a.s.vy[i] + a.hdt * fy;
            a.s.vz[i] = a.s.vz[i] + a.hdt * fz;
        }
    }
}

// A conditional launch does nothing when a rebuild is pending for its own step or an earlier one,
// and passes the halt on to the launches queued after it.  A force launch that carries the next
// step's pass 1 may itself raise FLAG_REBUILD for that next step while its later workgroups are
// still being dispatched: those must not halt.  FLAG_RB_STEP holds INT_MAX from every reset of
// FLAG_REBUILD until a check stores its step, so a workgroup of the launch of step k reads either
// INT_MAX or k+1 for a request raised by its own launch (both above k), and a step <= k only for a
// request raised by an earlier launch (ordered before it on the stream, so every workgroup reads
// the same values).
__device__ inline bool halted(int* f, int step) {
    if (!f) return false;
    const bool h = f[FLAG_HALT] != 0 ||
                   (__atomic_load_n(&f[FLAG_REBUILD], __ATOMIC_RELAXED) != 0 &&
                    __atomic_load_n(&f[FLAG_RB_STEP], __ATOMIC_RELAXED) <= step);
    if (h && blockIdx.x == 0 && threadIdx.x == 0) {
        f[FLAG_HALT] = 1;
        atomicMin(&f[FLAG_HALT_STEP], step);
    }
    return h;
}

// (the LJ steps without tallies, with bonds or tethers: registers held to what lets two
// workgroups share a CU)
template <bool MIE, bool RECORD, bool EXTRA, bool VIR>
__global__ __launch_bounds__(FB_BLOCK) __attribute__((amdgpu_waves_per_eu(!MIE && !RECORD ? FB_BLOCK / 128 : 1))) void
k_forces_brick(FbArgs a) {
    extern __shared__ __align__(16) char smem[];
    if (halted(a.halt, a.step)) return;
    forces_body<MIE, RECORD, EXTRA, VIR>(a, brick_of_block(a.order, blockIdx.x, gridDim.x), smem);
}

// LJ record and virial steps: registers held to what lets two workgroups share a CU (a few
// spill on record steps, which beats one workgroup per CU)
template <bool RECORD, bool VIR>
__global__ __launch_bounds__(FB_BLOCK) __attribute__((amdgpu_waves_per_eu(FB_BLOCK / 128))) void
k_forces_brick_lj(FbArgs a) {
    extern __shared__ __align__(16) char smem[];
    if (halted(a.halt, a.step)) return;
    forces_body<false, RECORD, false, VIR>(a, brick_of_block(a.order, blockIdx.x, gridDim.x), smem);
}

// record steps with the MOP tally only (forces_body_mop): two workgroups per CU, as the LDS of
// two regions allows
template <bool SLOTS>
__global__ __launch_bounds__(FB_BLOCK) __attribute__((amdgpu_waves_per_eu(FB_BLOCK / 128))) void
k_forces_brick_mop(FbArgs a) {
    extern __shared__ __align__(16) char smem[];
    if (halted(a.halt, a.step)) return;
    forces_body_mop<SLOTS>(a, brick_of_block(a.order, blockIdx.x, gridDim.x), smem);
}

// LJ steps without tallies, bonds or tethers (the common step): registers held to what lets two
// workgroups share a CU (the LDS of two regions fits its 160 KB)
// (P1: with pass 1 of the next step; REC: a record step tallying PE and virial only)
template <bool P1, bool REC = false>
__global__ __launch_bounds__(FB_BLOCK) __attribute__((amdgpu_waves_per_eu(FB_BLOCK / 128, FB_BLOCK / 128))) void
k_forces_brick_plain(FbArgs a) {
    extern __shared__ __align__(16) char smem[];
    if (halted(a.halt, a.step)) return;
    forces_body<false, REC, false, REC, P1, true>(a, brick_of_block(a.order, blockIdx.x, gridDim.x), smem);
}

// Mie steps without tallies and record steps of PE and virial (forces_body_mie): the steps without
// tallies of up to two types held to the registers that let two workgroups share a CU (the LDS of
// two regions with one-byte types fits its 160 KB); record steps (one in TPLOT) and four types
// (the row of four partner types) without a cap
template <int NT, int NB, bool EXTRA, bool REC, bool P1>
__global__ __launch_bounds__(FB_BLOCK) __attribute__((amdgpu_waves_per_eu(REC || NT > 2 ? 1 : FB_BLOCK / 128,
                                                                        REC || NT > 2 ? 8 : FB_BLOCK / 128))) void
k_forces_brick_mie(FbArgs a) {
    extern __shared__ __align__(16) char smem[];
    if (halted(a.halt, a.step)) return;
    forces_body_mie<NT, NB, EXTRA, REC, P1>(a, brick_of_block(a.order, blockIdx.x, gridDim.x), smem);
}

// A domain's step in a packed launch does nothing when the domain has no step in this launch, or
// when a rebuild is pending for this step or an earlier one (the rule of halted(), read only:
// a packed launch's first workgroup belongs to one domain, so none raises FLAG_HALT).
__device__ inline bool skip_domain(const FbArgs& a) {
    if (!a.active) return true;
    const int* f = a.halt;
    return f && (f[FLAG_HALT] != 0 || (__atomic_load_n(&f[FLAG_REBUILD], __ATOMIC_RELAXED) != 0 &&
                                       __atomic_load_n(&f[FLAG_RB_STEP], __ATOMIC_RELAXED) <= a.step));
}

// the mixed mode's plain steps over the FP32 staging: registers held to what lets three workgroups
// share a CU (the LDS of three regions fits)
template <bool P1, bool REC = false>
__global__ __launch_bounds__(FB_BLOCK) __attribute__((amdgpu_waves_per_eu(3 * FB_BLOCK / 256, 3 * FB_BLOCK / 256))) void
k_forces_brick_plain32(FbArgs a) {
    extern __shared__ __align__(16) char smem[];
    if (halted(a.halt, a.step)) return;
    forces_body32<P1, REC, true>(a, brick_of_block(a.order, blockIdx.x, gridDim.x), smem);
}

// the branch-free Mie body of a packed launch's domain, by its partner types and exponent class
template <bool EXTRA, bool REC>
__device__ __forceinline__ void mie_dispatch(const FbArgs& a, int brick, char* smem) {
    if (a.mie_nt == 2) {
        if (a.mie_nb == 0) forces_body_mie<2, 0, EXTRA, REC, false, false>(a, brick, smem);
        else if (a.mie_nb == 4) forces_body_mie<2, 4, EXTRA, REC, false, false>(a, brick, smem);
        else forces_body_mie<2, 5, EXTRA, REC, false, false>(a, brick, smem);
    } else {
        if (a.mie_nb == 0) forces_body_mie<4, 0, EXTRA, REC, false, false>(a, brick, smem);
        else if (a.mie_nb == 4) forces_body_mie<4, 4, EXTRA, REC, false, false>(a, brick, smem);
        else forces_body_mie<4, 5, EXTRA, REC, false, false>(a, brick, smem);
    }
}

// packed launch over many domains: block b works on brick table[b].y of domain table[b].x.
// PLAIN: no domain of the launch is on a record or virial step (only the plain path is compiled,
// and its registers are held to what lets two workgroups share a CU, as for the one-domain launch)
template <bool MIE, bool EXTRA, bool PLAIN>
__global__ __launch_bounds__(FB_BLOCK) __attribute__((amdgpu_waves_per_eu(PLAIN && !MIE ? FB_BLOCK / 128 : 1))) void
k_forces_brick_multi(const FbArgs* descs, const int2* table) {
    extern __shared__ __align__(16) char smem[];
    const int2 t = table[blockIdx.x];
    const FbArgs& a = descs[t.x];
    if (skip_domain(a)) return;
    if constexpr (MIE) {
        // a Mie domain takes the branch-free Mie body on the steps where its one-domain launch
        // does (steps without tallies, record steps of PE and virial only)
        const bool simple_rec = a.rec && a.pxy == nullptr && a.cv == nullptr && a.rf == nullptr;
        if (a.mie_nt > 0 && ((!a.rec && !a.ev) || simple_rec)) {
            if (a.rec) mie_dispatch<EXTRA, true>(a, t.y, smem);
            else mie_dispatch<EXTRA, false>(a, t.y, smem);
            return;
        }
    }
    if constexpr (PLAIN) {
        // (LJ: staged from the member's table when its last list build wrote one; a member's
        // rebuilds open its windows, before the window's descriptors are made)
        forces_body<MIE, false, EXTRA, false, false, !MIE>(a, t.y, smem);
        return;
    }
    if (a.rec) forces_body<MIE, true, EXTRA>(a, t.y, smem);
    else if (a.ev) forces_body<MIE, false, EXTRA, true>(a, t.y, smem);
    else forces_body<MIE, false, EXTRA>(a, t.y, smem);
}

// Mixed precision: image coordinates and differences in FP64, the pair law and the per-molecule
// force sum in FP32; energy, virial and the tallies accumulated in FP64.
template <bool RECORD, bool EXTRA, bool VIR = RECORD>
__device__ __forceinline__ void forces_body_mixed(const FbArgs& a, int brick, char* smem) {
    Region r;
    const bool mop = RECORD && a.pxy != nullptr;
    const bool cvrec = RECORD && a.cv != nullptr;
    carve(r, smem, a.g, a.maxreg, mop || cvrec);
    double* spxy = nullptr;
    if (mop) spxy = (double*)(smem + region_lds_bytes(a.g, a.maxreg, true));
    // staged molecule numbers in the slot array (the owner rule is its only reader)
    if (!stage_region(r, a.g, brick, a.start, a.s, true, nullptr, a.cap, nullptr, nullptr, a.maxreg, a.flags,
                      nullptr, a.ext))
        return;
    if (mop) {
        for (int t = threadIdx.x; t < 3 * a.nplanes; t += blockDim.x) spxy[t] = 0.0;
        __syncthreads();
    }
    const size_t cap = a.cap;
    for (int q = threadIdx.x; q < r.ni; q += blockDim.x) {
        int rci, li, i;
        imol(r, a.start, q, rci, li, i);
        const double xi = r.pos[3 * li], yi = r.pos[3 * li + 1], zi = r.pos[3 * li + 2];
        const int n = a.nbn[i];
        float fx = 0.f, fy = 0.f, fz = 0.f;
        double pe = 0.0, vir = 0.0;
        const int exti = (mop || (RECORD && a.cv)) ? a.ext[i] : 0;
        const uint64_t* row = a.nbl + i;
        // next word prefetched; the four entries' positions read together (see forces_body)
        uint64_t wnext = n > 0 ? row[0] : 0;
        for (int kw = 0; kw < n; kw += 4) {
            const uint64_t word = wnext;
            if (kw + 4 < n) wnext = row[(size_t)((kw >> 2) + 1) * cap];
            int jq[4];
            double rxq[4], ryq[4], rzq[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                jq[q] = (int)((word >> (16 * q)) & 0xffffu);
                rxq[q] = xi - r.pos[3 * (jq[q])];
                ryq[q] = yi - r.pos[3 * (jq[q]) + 1];
                rzq[q] = zi - r.pos[3 * (jq[q]) + 2];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
            if (kw + q >= n) break;
            const int j = jq[q];
            const double rxd = rxq[q], ryd = ryq[q], rzd = rzq[q];
            // the cutoff decided on the FP64 r^2, as the FP64 path decides it
            if (!(((rxd * rxd + ryd * ryd) + rzd * rzd) < a.rc2)) continue;
            const float rx = (float)rxd, ry = (float)ryd, rz = (float)rzd;
            float acc, inv;
            {
                // explicit fused multiply-adds: the one-domain and packed launches round alike
                const float r2 = fmaf(rz, rz, fmaf(ry, ry, rx * rx));
                inv = __builtin_amdgcn_rcpf(r2);
                const float i2 = inv * inv, i4 = i2 * i2, i3 = i2 * inv;
                acc = 48.f * fmaf(i4, i3, -0.5f * i4);
                fx = fmaf(acc, rx, fx);
                fy = fmaf(acc, ry, fy);
                fz = fmaf(acc, rz, fz);
            }
            if (VIR) vir = vir + (double)acc * ((rxd * rxd + ryd * ryd) + rzd * rzd);
            if (RECORD) {
                const double id = (double)inv;
                pe = pe + (4.0 * (p6(id) - p3(id)) - a.potshift);
                if (mop || a.cv) {
                    const int rcj = region_cell_of(r, a.g.nreg, j);
                    if (mop_owner(r, a.g, rci, rcj, exti, r.lslot[j])) {
                        const double ri[3] = {xi, yi, zi}, rj[3] = {r.pos[3 * j], r.pos[3 * j + 1], r.pos[3 * j + 2]},
                                     rij[3] = {rxd, ryd, rzd};
                        if (mop) mop_entry(a, ri, rj, rij, (double)acc, spxy);
                        const bool img_entry = r.sh[rcj] != 13 || r.sh[rci] != 13;
                        if (a.cv) cv_entry(a, ri, rj, img_entry ? (double)acc : 2.0 * (double)acc, exti);
                    }
                }
            }
        }
        }
        double dfx = fx, dfy = fy, dfz = fz;
        if (EXTRA && a.tag) {
            const int tg = a.tag[i];
            if (tg == FM_TAG_TETH || tg == FM_TAG_TETH_SLIDE) {
                double* T = a.teth;
                const double rio[3] = {xi - T[i], yi - T[i + cap], zi - T[i + 2 * cap]};
                const double mag = sqrt((rio[0] * rio[0] + rio[1] * rio[1]) + rio[2] * rio[2]);
                const double am = -2.0 * a.teth_k2 * mag - 4.0 * a.teth_k4 * (mag * mag) -
                                  6.0 * a.teth_k6 * pow(mag, 4.0);
                dfx = dfx + am * rio[0];
                dfy = dfy + am * rio[1];
                dfz = dfz + am * rio[2];
                for (int d = 0; d < 3; ++d) T[i + d * cap] = T[i + d * cap] + T[i + (3 + d) * cap] * a.dt;
            }
        }
        a.s.ax[i] = dfx;
        a.s.ay[i] = dfy;
        a.s.az[i] = dfz;
        if (a.kick) {
            a.s.vx[i] = a.s.vx[i] + a.hdt * dfx;
            a.s.vy[i] = a.s.vy[i] + a.hdt * dfy;
            a.s.vz[i] = a.s.vz[i] + a.hdt * dfz;
        }
        if (VIR) a.vir[i] = vir;
        if (RECORD) a.pe_lj[i] = pe;
    }
    if (mop) {
        __syncthreads();
        for (int t = threadIdx.x; t < 3 * a.nplanes; t += blockDim.x)
            if (spxy[t] != 0.0) atomicAdd(&a.pxy[t], spxy[t]);
    }
}

// registers held to what lets two workgroups share a CU, as for k_forces_brick_lj
template <bool RECORD, bool EXTRA, bool VIR>
__global__ __launch_bounds__(FB_BLOCK) __attribute__((amdgpu_waves_per_eu(FB_BLOCK / 128))) void
k_forces_brick_mixed(FbArgs a) {
    extern __shared__ __align__(16) char smem[];
    if (halted(a.halt, a.step)) return;
    forces_body_mixed<RECORD, EXTRA, VIR>(a, brick_of_block(a.order, blockIdx.x, gridDim.x), smem);
}

template <bool EXTRA, bool PLAIN>
__global__ __launch_bounds__(FB_BLOCK) __attribute__((amdgpu_waves_per_eu(PLAIN ? FB_BLOCK / 128 : 1))) void
k_forces_brick_mixed_multi(const FbArgs* descs, const int2* table) {
    extern __shared__ __align__(16) char smem[];
    const int2 t = table[blockIdx.x];
    const FbArgs& a = descs[t.x];
    if (skip_domain(a)) return;
    if constexpr (PLAIN) {
        if constexpr (!EXTRA) forces_body32<false>(a, t.y, smem);
        else forces_body_mixed<false, EXTRA>(a, t.y, smem);
        return;
    }
    const bool simple_rec = a.rec && a.pxy == nullptr && a.cv == nullptr && a.rf == nullptr;
    if constexpr (!EXTRA) {
        if (simple_rec) {
            forces_body32<false, true>(a, t.y, smem);
            return;
        }
    }
    if (a.rec) forces_body_mixed<true, EXTRA>(a, t.y, smem);
    else if (a.ev) forces_body_mixed<false, EXTRA, true>(a, t.y, smem);
    else if constexpr (!EXTRA) forces_body32<false>(a, t.y, smem);
    else forces_body_mixed<false, EXTRA>(a, t.y, smem);
}

// ---- decode the LDS-indexed list into (slot << 5 | shift code) for inspection ------------------
__global__ __launch_bounds__(FM_BLOCK) void k_decode_brick(St s, BG g, const int* start, const uint64_t* nbl,
                                                           const int* nbn, int cap, int maxreg, uint32_t* out) {
    extern __shared__ __align__(16) char smem[];
    Region r;
    carve(r, smem, g, maxreg, true);
    if (!stage_region(r, g, blockIdx.x, start, s, false, nullptr, cap, nullptr, nullptr, maxreg, nullptr)) return;
    for (int q = threadIdx.x; q < r.ni; q += blockDim.x) {
        int rci, li, i;
        imol(r, start, q, rci, li, i);
        const int n = nbn[i];
        for (int k = 0; k < n; ++k) {
            const int j = (int)((nbl[(size_t)(k >> 2) * cap + i] >> (16 * (k & 3))) & 0xffffu);
            const int rcj = region_cell_of(r, g.nreg, j);
            out[(size_t)k * cap + i] = ((uint32_t)r.lslot[j] << 5) | (uint32_t)r.sh[rcj];
        }
    }
}

// staged region sizes (LDS sizing): one wavefront per brick sums its region cells; the largest
// of each workgroup goes into flags[FLAG_MAX_REGION]
__global__ __launch_bounds__(FM_BLOCK) void k_region_max(BG g, const int* start, int nbrick, int* flags) {
    __shared__ int wm[FM_BLOCK / 64];
    const int b = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    int tot = 0;
    if (b < nbrick) {
        const int bx = b % g.nb[0], by = (b / g.nb[0]) % g.nb[1], bz = b / (g.nb[0] * g.nb[1]);
        const int org[3] = {bx * g.B[0], by * g.B[1], bz * g.B[2]};
        for (int t = lane; t < g.nreg; t += 64) {
            const int rr[3] = {t % g.R[0], (t / g.R[0]) % g.R[1], t / (g.R[0] * g.R[1])};
            int v[3];
            for (int d = 0; d < 3; ++d) {
                v[d] = org[d] + rr[d] - 1;
                if (v[d] < 0) v[d] += g.nc[d];
                else if (v[d] >= g.nc[d]) v[d] -= g.nc[d];
            }
            const int c = v[0] + g.nc[0] * (v[1] + g.nc[1] * v[2]);
            tot += start[c + 1] - start[c];
        }
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off, 64);
    }
    if (lane == 0) wm[threadIdx.x >> 6] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = wm[0];
        for (int w = 1; w < FM_BLOCK / 64; ++w) m = max(m, wm[w]);
        atomicMax(&flags[FLAG_MAX_REGION], m);
    }
}

inline int nblk(size_t n) { return (int)((n + FM_BLOCK - 1) / FM_BLOCK); }

BG make_bg(const fm_ctx* c) {
    BG g;
    for (int d = 0; d < 3; ++d) {
        g.nc[d] = c->nc[d];
        g.B[d] = c->brick[d];
        g.nb[d] = (c->nc[d] + c->brick[d] - 1) / c->brick[d];
        g.R[d] = c->brick[d] + 2;
        g.L[d] = c->p.domain[d];
    }
    g.nreg = g.R[0] * g.R[1] * g.R[2];
    return g;
}

}  // namespace

#define LAUNCH_CHECK()                                                                       \
    do {                                                                                     \
        hipError_t _e = hipGetLastError();                                                   \
        if (_e != hipSuccess) return fm_set_error(FM_EHIP, "kernel launch: %s (%s:%d)",       \
                                                  hipGetErrorString(_e), __FILE__, __LINE__); \
    } while (0)

int brick_count(const fm_ctx* c) {
    BG g = make_bg(c);
    return g.nb[0] * g.nb[1] * g.nb[2];
}

// Partner types of the branch-free Mie launch (2 or 4; bits of the exponents in *nb, 0 for the
// class la = 6, lr = 12..18), or 0 when some type pair's exponent lr/2 + 1 or la/2 + 1 is not a
// whole number below 32 (the general loop then runs)
static int mie_plain_types(const fm_ctx* c, int* nb) {
    const fm_params& p = c->p;
    if (p.mie_potential != 1) return 0;
    int emax = 0;
    bool la6 = true;  // la = 6 for every pair and lr = 12, 14, 16 or 18 (nb = 0)
    for (int ti = 0; ti < p.ntypes; ++ti)
        for (int tj = 0; tj < p.ntypes; ++tj) {
            const int k = ti * FM_MAX_TYPES + tj;
            for (double lam : {p.lambdar_lookup[k], p.lambdaa_lookup[k]}) {
                const double e = 0.5 * lam + 1.0;
                if (!(e >= 0.0 && e < 32.0) || e != floor(e)) return 0;
                emax = std::max(emax, (int)e);
            }
            const double lr = p.lambdar_lookup[k];
            la6 = la6 && p.lambdaa_lookup[k] == 6.0 && (lr == 12.0 || lr == 14.0 || lr == 16.0 || lr == 18.0);
        }
    *nb = la6 ? 0 : (emax < 16 ? 4 : 5);
    return p.ntypes <= 2 ? 2 : 4;
}

size_t brick_lds_bytes(const fm_ctx* c, int maxreg, int kind) {
    BG g = make_bg(c);
    // kind 0: list build (+ slots and chain ids with bonds); 1: forces; 2: forces with MOP;
    // 3: the branch-free Mie launch (0 when it does not run)
    if (kind == 0) {
        bool poly = c->p.potential_flag == 1;
        return region_lds_bytes(g, maxreg, poly) + (poly ? 2 * (size_t)maxreg * sizeof(int) : 0);
    }
    const bool mie = c->p.mie_potential == 1;
    if (kind == 3) {  // the branch-free Mie launch, or 0 when it does not run
        int nb;
        return mie_plain_types(c, &nb) > 0 ? mie_lds_bytes(g, maxreg) : 0;
    }
    const size_t tab = mie ? FM_MIE_WORDS * sizeof(double) : 0;  // Mie tables
    if (kind == 1) return region_lds_bytes(g, maxreg, mie || c->p.vflux_outflag == 4, mie) + tab;
    // kind 2: the plane tallies, and for the MOP record kernel the plane positions
    return region_lds_bytes(g, maxreg, true, mie) + 4 * (size_t)c->p.nplanes * sizeof(double) + tab;
}

// measure the largest region (after binning); result in d_flags[FLAG_MAX_REGION]
int launch_region_max(fm_ctx* c) {
    BG g = make_bg(c);
    int nbrick = g.nb[0] * g.nb[1] * g.nb[2];
    k_region_max<<<nblk((size_t)nbrick * 64), FM_BLOCK, 0, c->stream>>>(g, c->cell_start, nbrick, c->d_flags);
    LAUNCH_CHECK();
    return 0;
}

int launch_nlist(fm_ctx* c) {
    if (c->np == 0) return 0;
    BG g = make_bg(c);
    NlbArgs a;
    a.s = state_of(c->x(), c->cap);
    a.g = g;
    a.start = c->cell_start;
    a.ext = c->ext[c->cur];
    a.mono = (c->p.potential_flag == 1 && c->has_mono) ? c->mono[c->cur] : nullptr;
    a.nbl = c->nbl;
    a.nbn = c->nbn;
    a.bond = c->bond;
    a.bondn = c->bondn;
    a.flags = c->d_flags;
    a.cap = c->cap;
    a.maxnb = c->maxnb;
    a.maxreg = c->maxreg;
    a.ra2 = c->list_ra2 > 0.0 ? c->list_ra2 : c->rn2;
    a.xsorted = c->moved_since_bin ? 0 : 1;
    a.xwin = sqrt(a.ra2) * (1.0 + 1e-9) + 1e-12;
    a.order = c->brick_order;
    for (int d = 0; d < 3; ++d) {
        a.cs[d] = c->cs[d];
        a.hd[d] = c->half[d];
    }
    if (a.xsorted && c->cap < (1 << 26)) {
        // quantised prefilter (slots packed with the shift code in 31 bits)
        NlqArgs q;
        q.mono = a.mono;
        q.bond = a.bond;
        q.bondn = a.bondn;
        q.ext = a.ext;
        q.s = a.s;
        q.g = g;
        q.start = a.start;
        q.nbl = a.nbl;
        q.nbn = a.nbn;
        q.flags = a.flags;
        q.cap = a.cap;
        q.maxnb = a.maxnb;
        q.maxreg = a.maxreg;
        q.ra2 = a.ra2;
        q.xwin = a.xwin;
        q.order = a.order;
        // the staging tables of the plain force launches, while the cells stay as binned
        q.btab = c->btab;
        q.btab_stride = c->btab_stride;
        double ext = 0.0;
        for (int d = 0; d < 3; ++d) {
            q.cs[d] = a.cs[d];
            q.hd[d] = a.hd[d];
            ext = fmax(ext, (g.B[d] + 2) * c->cs[d]);
        }
        // grid step: the region and two steps on each side within 0..32767; the quantised
        // squared distances of neighbouring-cell candidates then stay below 2^31
        q.h = ext / 32700.0;
        q.inv_h = 1.0 / q.h;
        const double t = sqrt(a.ra2) / q.h, eps = 1.0 + 1e-6, r3 = sqrt(3.0);
        q.lo = (int)floor(t * t - 2.0 * r3 * eps * t - 1e-9 * t * t) - 2;
        q.hi = (int)ceil(t * t * (1.0 + 3e-9) + 2.0 * r3 * eps * t * (1.0 + 1e-9) + 3.0 * eps * eps) + 2;
        const int nbmax = g.B[0] * g.B[1] * g.B[2];
        const size_t lds = (size_t)c->maxreg * (sizeof(uint2) + sizeof(int)) +
                           (size_t)((c->maxreg + 1) / 2 + 4 * g.nreg + 1 + 2 * nbmax + 1) * sizeof(int);
        ktimer_begin(c, KT_NLIST);
        if (a.mono) k_nlist_quant<true><<<brick_count(c), FB_BLOCK, lds, c->stream>>>(q);
        else k_nlist_quant<false><<<brick_count(c), FB_BLOCK, lds, c->stream>>>(q);
        LAUNCH_CHECK();
        ktimer_end(c, KT_NLIST);
        c->btab_valid = c->btab != nullptr;
        return 0;
    }
    c->btab_valid = false;
    const size_t lds = brick_lds_bytes(c, c->maxreg, 0);
    ktimer_begin(c, KT_NLIST);
    if (a.mono) k_nlist_brick<true><<<brick_count(c), FB_BLOCK, lds, c->stream>>>(a);
    else k_nlist_brick<false><<<brick_count(c), FB_BLOCK, lds, c->stream>>>(a);
    LAUNCH_CHECK();
    ktimer_end(c, KT_NLIST);
    return 0;
}

size_t fargs_bytes() { return sizeof(FbArgs); }

// LDS of the FP32-staged mixed launch
static size_t lds32(const fm_ctx* c) {
    const BG g = make_bg(c);
    const int nbmax = g.B[0] * g.B[1] * g.B[2];
    const size_t b = (size_t)c->maxreg * sizeof(float4) +
                     (size_t)((c->maxreg + 1) / 2 + 4 * g.nreg + 1 + 2 * nbmax + 1) * sizeof(int);
    return (b + 15) & ~(size_t)15;
}

// The band of the FP32-staged cutoff test: with u = 2^-24, a staged offset is within u*ext of the
// exact one (ext bounds the offsets: the region's extent plus a cell of drift), a difference
// within e = 2 u ext + u rc, and near the cutoff (|d| <= rc per component) the FP32 r^2 within
// e (2 sqrt(3) rc + 3 e) + 6 u rc^2 of the FP64 one; the band is four times that.
static float band32_of(const fm_ctx* c) {
    const BG g = make_bg(c);
    double ext = 0.0;
    for (int d = 0; d < 3; ++d) ext = fmax(ext, (g.B[d] + 3) * c->cs[d]);
    const double u = 1.0 / 16777216.0, rc = sqrt(c->rc2);
    const double e = 2.0 * u * ext + u * rc;
    return (float)(4.0 * (e * (2.0 * sqrt(3.0) * rc + 3.0 * e) + 6.0 * u * c->rc2));
}

// a step without record or virial tallies
bool fargs_plain(const void* fargs) {
    const FbArgs* a = (const FbArgs*)fargs;
    return !a->rec && !a->ev;
}

// a domain's force arguments in a queued packed launch: halted by its own flag words, or idle
void fargs_set_queued(void* fargs, int* flags, bool active) {
    FbArgs* a = (FbArgs*)fargs;
    a->halt = flags;
    a->active = active ? 1 : 0;
}

// Arguments of the force kernel for one domain at step iter.  kind: bit0 mixed precision,
// bit1 Mie tables, bit2 bonds or tethers present.
int make_fargs(fm_ctx* c, int iter, bool kick, void* out, size_t* lds_out, int* kind_out) {
    BG g = make_bg(c);
    FbArgs a;
    const int cur = c->cur;
    a.s = state_of(c->x(), c->cap);
    a.g = g;
    a.start = c->cell_start;
    a.nbl = c->nbl;
    a.nbn = c->nbn;
    const bool poly = c->p.potential_flag == 1 && c->has_mono;
    a.bond = poly ? c->bond : nullptr;
    a.bondn = poly ? c->bondn : nullptr;
    a.ext = c->ext[cur];
    const bool mie = c->p.mie_potential == 1;
    a.moltype = mie ? c->moltype[cur] : nullptr;
    // velocity Verlet: tethers act with the forces; leap-frog applies them in the move
    const bool tm = c->p.ensemble == FM_ENSEMBLE_TAG_MOVE && c->has_tags &&
                    c->p.integration_algorithm == FM_INTEGRATION_VV;
    a.tag = tm ? c->tag[cur] : nullptr;
    a.teth = tm ? c->teth[cur] : nullptr;
    a.pe_lj = c->pe_lj;
    a.pe_poly = c->pe_poly;
    a.vir = c->vir;
    a.flags = c->d_flags;
    a.cap = c->cap;
    a.maxreg = c->maxreg;
    const int tplot = c->p.tplot, teval = c->p.teval;
    a.step = iter;
    a.rec = (iter % tplot) == 0;
    a.ev = teval > 0 && (iter % teval) == 0;
    a.rf = (a.rec && c->p.pressure_outflag == 1) ? c->rf : nullptr;
    const bool mop = a.rec && c->p.vflux_outflag >= 1 && c->p.vflux_outflag <= 3 && c->pxy;
    a.pxy = mop ? c->pxy : nullptr;
    a.planes = c->planes;
    a.rc2 = c->rc2;
    a.potshift = c->potshift;
    a.k_c = c->p.k_c;
    a.R0 = c->p.R_0;
    a.R02 = c->p.R_0 * c->p.R_0;
    a.dt = c->p.delta_t;
    a.teth_k2 = c->p.teth_k2;
    a.teth_k4 = c->p.teth_k4;
    a.teth_k6 = c->p.teth_k6;
    a.hx = c->half[0];
    a.hz = c->half[2];
    a.nplanes = c->p.nplanes;
    const int pn = c->p.vflux_outflag - 1;
    a.hp = (pn >= 0 && pn < 3) ? c->half[pn] : 0.0;
    a.spacing = c->p.planespacing;
    a.shift = c->mop_shift;
    a.mie = mie ? c->d_mie : nullptr;
    a.order = c->brick_order;
    a.halt = nullptr;
    a.active = 1;
    a.band32 = band32_of(c);
    a.btab = c->btab_valid ? c->btab : nullptr;
    a.btab_stride = c->btab_stride;
    a.mie_nb = 0;
    // (Mie tables compute in FP64 in either precision)
    a.mie_nt = mie ? mie_plain_types(c, &a.mie_nb) : 0;
    a.cv = (a.rec && c->cvface) ? c->cvface : nullptr;
    for (int d = 0; d < 3; ++d) {
        a.cvnbo[d] = c->cv_nbo[d];
        a.cvnhb[d] = c->cv_nhb[d];
        a.cvbs[d] = c->cv_bs[d];
        a.cvhd[d] = c->half[d];
    }
    a.kick = kick ? 1 : 0;
    a.hdt = 0.5 * c->p.delta_t;
    // LDS: the MOP layout whenever MOP is enabled, so that one launch can mix record steps
    const bool mop_enabled = c->p.vflux_outflag >= 1 && c->p.vflux_outflag <= 3 && c->pxy;
    const size_t lds = brick_lds_bytes(c, c->maxreg, mop_enabled ? 2 : 1);
    if (mop && (region_lds_bytes(g, c->maxreg, true, mie) % sizeof(double)) != 0)
        return fm_set_error(FM_EINVAL, "MOP tallies would be misaligned in LDS");
    memcpy(out, &a, sizeof a);
    if (lds_out) *lds_out = lds;
    if (kind_out) *kind_out = (c->p.precision == FM_PREC_MIXED && !mie && !poly ? 1 : 0) | (mie ? 2 : 0) |
                              (poly || tm ? 4 : 0);
    return 0;
}

// the branch-free Mie launch with nt partner types and nb exponent bits (mie_plain_types)
static void launch_forces_mie(const fm_ctx* c, const FbArgs& a, int nt, int nb, bool rec, bool p1) {
    const size_t lm = brick_lds_bytes(c, c->maxreg, 3);
    const dim3 grid(brick_count(c)), blk(FB_BLOCK);
    const bool extra = a.bondn != nullptr || a.tag != nullptr;
#define FMIE_E(NT, NB, E)                                                                                \
    (rec ? (p1 ? k_forces_brick_mie<NT, NB, E, true, true><<<grid, blk, lm, c->stream>>>(a)             \
               : k_forces_brick_mie<NT, NB, E, true, false><<<grid, blk, lm, c->stream>>>(a))           \
         : (p1 ? k_forces_brick_mie<NT, NB, E, false, true><<<grid, blk, lm, c->stream>>>(a)            \
               : k_forces_brick_mie<NT, NB, E, false, false><<<grid, blk, lm, c->stream>>>(a)))
#define FMIE(NT, NB) (extra ? FMIE_E(NT, NB, true) : FMIE_E(NT, NB, false))
    if (nt == 2) {
        if (nb == 0) FMIE(2, 0);
        else if (nb == 4) FMIE(2, 4);
        else FMIE(2, 5);
    } else {
        if (nb == 0) FMIE(4, 0);
        else if (nb == 4) FMIE(4, 4);
        else FMIE(4, 5);
    }
#undef FMIE
#undef FMIE_E
}

int launch_forces(fm_ctx* c, int iter, bool kick, bool cond, int next, bool* fused) {
    if (fused) *fused = false;
    if (c->np == 0) return 0;
    FbArgs a;
    size_t lds = 0;
    int kind = 0;
    if (int rc = make_fargs(c, iter, kick, &a, &lds, &kind)) return rc;
    if (cond) a.halt = c->d_flags;
    const bool mixed = kind & 1, mie = kind & 2, extra = kind & 4;
    // a record step tallying PE and virial only takes the branch-free loop too (FP64, and the
    // FP32-staged one in mixed precision)
    const bool simple_rec = (kind == 0 || kind == 1) && a.rec && a.pxy == nullptr && a.cv == nullptr && a.rf == nullptr;
    // the branch-free Mie launch (FP64): steps without tallies, and record steps of PE and virial
    const int mnb = a.mie_nb, mnt = mixed ? 0 : a.mie_nt;
    const bool mie_rec = mnt > 0 && a.rec && a.pxy == nullptr && a.cv == nullptr && a.rf == nullptr;
    const bool mie_fast = mnt > 0 && ((!a.rec && !a.ev) || mie_rec);
    if (next >= 0 && ((((kind == 0 || kind == 1) && !a.rec && !a.ev) || simple_rec) || mie_fast) && kick && fused) {
        // this step's force and kick, then pass 1 of step `next` into the other state buffer
        a.s2 = state_of(c->st[1 - c->scur], c->cap);
        a.rd = c->rdisp;
        a.hdt2 = 0.5 * (c->p.delta_t * c->p.delta_t);
        a.thresh = 0.5 * c->dmin;
        a.fresh = c->rdisp_valid ? 0 : 1;
        a.p1step = next;
        ktimer_begin(c, KT_FORCES_P1);
        if (mie_fast) launch_forces_mie(c, a, mnt, mnb, mie_rec, true);
        else if (simple_rec && mixed) k_forces_brick_plain32<true, true><<<brick_count(c), FB_BLOCK, lds32(c), c->stream>>>(a);
        else if (simple_rec) k_forces_brick_plain<true, true><<<brick_count(c), FB_BLOCK, lds, c->stream>>>(a);
        else if (mixed) k_forces_brick_plain32<true><<<brick_count(c), FB_BLOCK, lds32(c), c->stream>>>(a);
        else k_forces_brick_plain<true><<<brick_count(c), FB_BLOCK, lds, c->stream>>>(a);
        ktimer_end(c, KT_FORCES_P1);
        LAUNCH_CHECK();
        c->scur = 1 - c->scur;
        c->rdisp_valid = true;
        *fused = true;
        return 0;
    }
    ktimer_begin(c, KT_FORCES);
    dim3 grid(brick_count(c)), blk(FB_BLOCK);
    // a record step whose only tally beyond PE and virial is MOP: the one-pass FP64 kernel (for
    // both precisions)
    if ((kind == 0 || kind == 1) && a.rec && a.pxy != nullptr && a.cv == nullptr && a.rf == nullptr) {
        // a pair inside rc crosses at most three planes near molecule i when 2 rc < 3 spacings
        if (2.0 * sqrt(a.rc2) / a.spacing < 3.0 - 1e-6) k_forces_brick_mop<true><<<grid, blk, lds, c->stream>>>(a);
        else k_forces_brick_mop<false><<<grid, blk, lds, c->stream>>>(a);
        ktimer_end(c, KT_FORCES);
        LAUNCH_CHECK();
        return 0;
    }
#define FB(M, R, E) \
    (R ? k_forces_brick<M, true, E, true><<<grid, blk, lds, c->stream>>>(a) \
       : (a.ev ? k_forces_brick<M, false, E, true><<<grid, blk, lds, c->stream>>>(a) \
               : k_forces_brick<M, false, E, false><<<grid, blk, lds, c->stream>>>(a)))
#define FM(R, E) \
    (R ? k_forces_brick_mixed<true, E, true><<<grid, blk, lds, c->stream>>>(a) \
       : (a.ev ? k_forces_brick_mixed<false, E, true><<<grid, blk, lds, c->stream>>>(a) \
               : k_forces_brick_mixed<false, E, false><<<grid, blk, lds, c->stream>>>(a)))
    if (mixed) {
        if (extra) { if (a.rec) FM(true, true); else FM(false, true); }
        else if (simple_rec) k_forces_brick_plain32<false, true><<<grid, blk, lds32(c), c->stream>>>(a);
        else if (a.rec) FM(true, false);
        else if (a.ev) FM(false, false);
        else k_forces_brick_plain32<false><<<grid, blk, lds32(c), c->stream>>>(a);
    } else if (mie) {
        if (mie_fast) launch_forces_mie(c, a, mnt, mnb, mie_rec, false);
        else if (extra) { if (a.rec) FB(true, true, true); else FB(true, false, true); }
        else { if (a.rec) FB(true, true, false); else FB(true, false, false); }
    } else {
        if (extra) { if (a.rec) FB(false, true, true); else FB(false, false, true); }
        else if (simple_rec) k_forces_brick_plain<false, true><<<grid, blk, lds, c->stream>>>(a);
        else if (a.rec) k_forces_brick_lj<true, true><<<grid, blk, lds, c->stream>>>(a);
        else if (a.ev) k_forces_brick_lj<false, true><<<grid, blk, lds, c->stream>>>(a);
        else k_forces_brick_plain<false><<<grid, blk, lds, c->stream>>>(a);
    }
#undef FB
#undef FM
    ktimer_end(c, KT_FORCES);
    LAUNCH_CHECK();
    return 0;
}

int launch_forces_multi(hipStream_t st, int kind, const void* d_descs, const int2* d_table, int nblk, size_t lds,
                        bool plain) {
    if (nblk == 0) return 0;
    const FbArgs* d = (const FbArgs*)d_descs;
    const bool mixed = kind & 1, mie = kind & 2, extra = kind & 4;
#define MULTI(K, ...) \
    (plain ? K<__VA_ARGS__, true><<<nblk, FB_BLOCK, lds, st>>>(d, d_table) \
           : K<__VA_ARGS__, false><<<nblk, FB_BLOCK, lds, st>>>(d, d_table))
    if (mixed) {
        if (extra) MULTI(k_forces_brick_mixed_multi, true);
        else MULTI(k_forces_brick_mixed_multi, false);
    } else if (mie) {
        if (extra) MULTI(k_forces_brick_multi, true, true);
        else MULTI(k_forces_brick_multi, true, false);
    } else {
        if (extra) MULTI(k_forces_brick_multi, false, true);
        else MULTI(k_forces_brick_multi, false, false);
    }
#undef MULTI
    LAUNCH_CHECK();
    return 0;
}

int launch_decode_nlist(fm_ctx* c, uint32_t* out) {
    if (c->np == 0) return 0;
    BG g = make_bg(c);
    const size_t lds = brick_lds_bytes(c, c->maxreg, 2) ;
    k_decode_brick<<<brick_count(c), FM_BLOCK, lds, c->stream>>>(state_of(c->x(), c->cap), g, c->cell_start, c->nbl,
                                                                 c->nbn, c->cap, c->maxreg, out);
    LAUNCH_CHECK();
    return 0;
}
