// fm_kernels.hip -- gfx950 kernels of the force / neighbour / integration path.
//
// Parity-critical FP64 arithmetic (cell index, image coordinates, squared distances, velocity
// Verlet) is written in the reference's operation order and compiled without fused multiply-add
// (file-wide fp contract off), so it is bit-identical to the reference's FP64 operations.
// The mixed-precision pair law re-enables contraction locally.
#include "fm_internal.h"

#include <math.h>

#pragma clang fp contract(off)

namespace {

struct St {  // molecule state in slot order
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

inline int nblocks(size_t n) { return (int)((n + FM_BLOCK - 1) / FM_BLOCK); }

__device__ inline void set_err(int* flags, int code, int mol) {
    if (atomicCAS(&flags[FLAG_ERROR], 0, code) == 0) flags[FLAG_ERROR_MOL] = mol;
}

// gcc __builtin_powi expansion (what gfortran emits for x**n): x7 = (x*x2)*(x2*x2) ...
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
// image coordinate r + s*L, one add per shifted component (the reference's halo copy)
__device__ inline double img(double r, int s, double L) { return s == 0 ? r : r + (double)s * L; }

// ---------------------------------------------------------------------------------------------
__global__ void k_fill_i32(int* p, int v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double* p, double v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- velocity Verlet, src/simulation_move_particles_vv.f90:64-69 (pass 1), 132-135 (pass 2) ---
__global__ __launch_bounds__(FM_BLOCK) void k_pass1(St s, int np, double dt, double hdt2, double hdt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double ax = s.ax[i], ay = s.ay[i], az = s.az[i];
    double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
    s.x[i] = s.x[i] + dt * vx + hdt2 * ax;
    s.y[i] = s.y[i] + dt * vy + hdt2 * ay;
    s.z[i] = s.z[i] + dt * vz + hdt2 * az;
    s.vx[i] = vx + hdt * ax;
    s.vy[i] = vy + hdt * ay;
    s.vz[i] = vz + hdt * az;
}

__global__ __launch_bounds__(FM_BLOCK) void k_pass2(St s, int np, double hdt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    s.vx[i] = s.vx[i] + hdt * s.ax[i];
    s.vy[i] = s.vy[i] + hdt * s.ay[i];
    s.vz[i] = s.vz[i] + hdt * s.az[i];
}

// ---- rebuild criterion 2, src/simulation_checkrebuild.f90:105-121 ----------------------------
__global__ __launch_bounds__(FM_BLOCK) void k_check_disp(St s, double* rd, int np, int cap, double dt,
                                                          double thresh, int fresh, int* flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double d0 = fresh ? 0.0 : rd[i];
    double d1 = fresh ? 0.0 : rd[i + cap];
    double d2 = fresh ? 0.0 : rd[i + 2 * (size_t)cap];
    d0 = d0 + s.vx[i] * dt;
    d1 = d1 + s.vy[i] * dt;
    d2 = d2 + s.vz[i] * dt;
    rd[i] = d0;
    rd[i + cap] = d1;
    rd[i + 2 * (size_t)cap] = d2;
    bool hit = fabs(d0) > thresh || fabs(d1) > thresh || fabs(d2) > thresh;
    unsigned long long m = __ballot(hit);
    if (m != 0ull && (int)(threadIdx.x & 63) == __ffsll((long long)m) - 1) atomicOr(&flags[FLAG_REBUILD], 1);
}

// ---- rebuild criterion 0 (Rapaport vmax), src/simulation_checkrebuild.f90:84-99 --------------
__global__ __launch_bounds__(FM_BLOCK) void k_vmax(St s, int np, unsigned long long* vmax_bits) {
    __shared__ double red[FM_BLOCK / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (i < np) {
        double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
        m = (vx * vx + vy * vy) + vz * vz;
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = red[0];
        for (int w = 1; w < FM_BLOCK / 64; ++w) b = fmax(b, red[w]);
        // non-negative doubles order like their bit patterns
        atomicMax(vmax_bits, (unsigned long long)__double_as_longlong(b));
    }
}

__global__ void k_vmax_finish(unsigned long long* vmax_bits, double* rmax, double dt, double thresh,
                              int forced, int* flags) {
    double vmax = sqrt(__longlong_as_double((long long)*vmax_bits));
    double r = *rmax + vmax * dt;
    int rb = forced || r > thresh;
    if (rb) {
        flags[FLAG_REBUILD] = 1;
        r = 0.0;
    }
    *rmax = r;
    *vmax_bits = 0ull;
}

// ---- sendmols on one periodic rank, src/messenger.MPI.f90:1383-1904 ---------------------------
// 1) mark molecules (in molecule-number order) that lie outside the box on any axis
__global__ __launch_bounds__(FM_BLOCK) void k_out_flags(St s, const int* slot_of, int np, double hx,
                                                        double hy, double hz, int* flag, int* flags) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= np) return;
    int sl = slot_of[e];
    double x = s.x[sl], y = s.y[sl], z = s.z[sl];
    int out = (x < -hx) | (x >= hx) | (y < -hy) | (y >= hy) | (z < -hz) | (z >= hz);
    flag[e] = out;
}

// 2) one workgroup replays the 3 x (3 axes x 2 directions) passes on the compacted set.
//    vals[q] ascending molecule numbers of the set (fixed); sl[q] slot currently holding vals[q].
//    Each pass: leavers in q order (= ascending molecule number) are wrapped by -dir*domain and
//    their molecule numbers reversed, as checksendbuild + sendrecvface + reorderdata do.
__global__ __launch_bounds__(1024) void k_sendmols_passes(St s, double* teth, const int* tag, int* ext,
                                                          int* slot_of, const int* vals, int* sl,
                                                          int* lst, int* tmp, const int* count,
                                                          int cap, double Lx, double Ly, double Lz,
                                                          int tag_move, double wsx, double wsy,
                                                          double wsz) {
    __shared__ int wsum[16];
    __shared__ int chunk_base;
    const int n = *count;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (int q = tid; q < n; q += blockDim.x) sl[q] = slot_of[vals[q]];
    __syncthreads();
    const double L[3] = {Lx, Ly, Lz};
    double* pos[3] = {s.x, s.y, s.z};
    for (int it = 0; it < 3; ++it)
        for (int ax = 0; ax < 3; ++ax)
            for (int dir = -1; dir <= 1; dir += 2) {
                const double h = 0.5 * L[ax];
                // ordered compaction of leavers into lst[0..k)
                if (tid == 0) chunk_base = 0;
                __syncthreads();
                for (int base = 0; base < n; base += blockDim.x) {
                    int q = base + tid;
                    int lv = 0;
                    if (q < n) {
                        double r = pos[ax][sl[q]];
                        lv = dir == -1 ? (r < -h) : (r >= h);
                    }
                    unsigned long long m = __ballot(lv);
                    int rank = __popcll(m & ((1ull << lane) - 1ull));
                    if (lane == 0) wsum[wid] = __popcll(m);
                    __syncthreads();
                    int off = 0, tot = 0;
                    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
                        if (w < wid) off += wsum[w];
                        tot += wsum[w];
                    }
                    if (lv) lst[chunk_base + off + rank] = q;
                    __syncthreads();
                    if (tid == 0) chunk_base += tot;
                    __syncthreads();
                }
                const int k = chunk_base;
                if (k == 0) continue;
                // wrap the leavers: r(ixyz) - dir*domain(ixyz)
                for (int m = tid; m < k; m += blockDim.x) {
                    int slot = sl[lst[m]];
                    pos[ax][slot] = pos[ax][slot] - (double)dir * L[ax];
                    if (tag_move && (tag[slot] == FM_TAG_TETH || tag[slot] == FM_TAG_TETH_SLIDE))
                        teth[slot + (size_t)ax * cap] = teth[slot + (size_t)ax * cap] - (double)dir * L[ax];
                }
                __syncthreads();
                for (int m = tid; m < k; m += blockDim.x) tmp[m] = sl[lst[k - 1 - m]];
                __syncthreads();
                for (int m = tid; m < k; m += blockDim.x) {
                    int slot = tmp[m];
                    sl[lst[m]] = slot;
                    if (tag_move) {  // read_tag on the renumbered molecule
                        double sg = tag[slot] == FM_TAG_TETH_SLIDE ? copysign(1.0, s.y[slot]) : 0.0;
                        teth[slot + 3 * (size_t)cap] = tag[slot] == FM_TAG_TETH_SLIDE ? wsx * sg : 0.0;
                        teth[slot + 4 * (size_t)cap] = tag[slot] == FM_TAG_TETH_SLIDE ? wsy * sg : 0.0;
                        teth[slot + 5 * (size_t)cap] = tag[slot] == FM_TAG_TETH_SLIDE ? wsz * sg : 0.0;
                    }
                }
                __syncthreads();
            }
    for (int q = tid; q < n; q += blockDim.x) {
        ext[sl[q]] = vals[q];
        slot_of[vals[q]] = sl[q];
    }
}

// ---- assign_to_cell as a counting sort, src/linklist.f90:2645-2676 -----------------------------
__global__ __launch_bounds__(FM_BLOCK) void k_bin_count(St s, const int* ext, int np, double hx, double hy,
                                                        double hz, double csx, double csy, double csz,
                                                        int nx, int ny, int nz, int* cell_of,
                                                        int* count, int* flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double r[3] = {s.x[i], s.y[i], s.z[i]};
    const double h[3] = {hx, hy, hz}, cs[3] = {csx, csy, csz};
    const int n[3] = {nx, ny, nz};
    int c[3];
    for (int d = 0; d < 3; ++d) {
        double t = r[d] + h[d];
        if (!(t > -cs[d] && t < (double)(n[d] + 1) * cs[d])) {  // also rejects NaN
            set_err(flags, FM_EESCAPE, ext[i]);
            return;
        }
        int ic = (int)ceil(t / cs[d]) + 1;  // 1-based with one halo layer
        if (ic < 1 || ic > n[d] + 2) {
            set_err(flags, FM_EESCAPE, ext[i]);
            return;
        }
        if (ic == 1) {
            if (t == 0.0) {  // exactly on the lower face: the reference's ceiling() gives the halo
                atomicAdd(&flags[FLAG_EDGE], 1);
                ic = 2;
            } else {
                set_err(flags, FM_EESCAPE, ext[i]);
                return;
            }
        }
        if (ic == n[d] + 2) {
            set_err(flags, FM_EESCAPE, ext[i]);
            return;
        }
        c[d] = ic - 2;
    }
    int cell = c[0] + nx * (c[1] + ny * c[2]);
    cell_of[i] = cell;
    atomicAdd(&count[cell], 1);
}

__global__ __launch_bounds__(FM_BLOCK) void k_bin_scatter(const int* cell_of, const int* start, int* cursor,
                                                          int* perm, int np) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    int c = cell_of[i];
    int pos = start[c] + atomicAdd(&cursor[c], 1);
    perm[pos] = i;
}

// order each cell by decreasing molecule number (the reference's head-inserted cell list order)
__global__ __launch_bounds__(FM_BLOCK) void k_bin_sort(const int* start, int* perm, const int* ext, int ncell,
                                                       int maxcell, int* flags) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int b = start[c], e = start[c + 1];
    if (e - b > maxcell) set_err(flags, FM_EOVERFLOW, c);
    atomicMax(&flags[FLAG_MAX_CELL], e - b);
    for (int k = b + 1; k < e; ++k) {
        int v = perm[k], key = ext[v];
        int m = k - 1;
        while (m >= b && ext[perm[m]] < key) {
            perm[m + 1] = perm[m];
            --m;
        }
        perm[m + 1] = v;
    }
}

struct PermArgs {
    const double* src;
    double* dst;
    const int *ext_src, *cell_src, *mt_src, *tag_src, *mono_src;
    int *ext_dst, *cell_dst, *mt_dst, *tag_dst, *mono_dst;
    const double* teth_src;
    double* teth_dst;
    int* slot_of;
    const int* perm;
    int np, cap;
};

__global__ __launch_bounds__(FM_BLOCK) void k_permute(PermArgs a) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.np) return;
    int o = a.perm[k];
    const size_t cap = a.cap;
#pragma unroll
    for (int f = 0; f < 9; ++f) a.dst[k + f * cap] = a.src[o + f * cap];
    int e = a.ext_src[o];
    a.ext_dst[k] = e;
    a.slot_of[e] = k;
    a.cell_dst[k] = a.cell_src[o];
    if (a.mt_dst) a.mt_dst[k] = a.mt_src[o];
    if (a.tag_dst) {
        a.tag_dst[k] = a.tag_src[o];
        for (int f = 0; f < 6; ++f) a.teth_dst[k + f * cap] = a.teth_src[o + f * cap];
    }
    if (a.mono_dst)
        for (int f = 0; f < 7; ++f) a.mono_dst[k + f * cap] = a.mono_src[o + f * cap];
}

// ---- neighbour list, src/linklist.f90:660-804 (set semantics), bonds 2542-2584 -----------------
struct NlArgs {
    St s;
    const int* cell_of;
    const int* start;
    const int* ext;
    const int* mono;      // chainID, subchainID, funcy, bflag[4] (slot order) or null
    uint32_t* nbl;
    int* nbn;
    uint32_t* bond;
    int* bondn;
    int* flags;
    int np, cap, maxnb;
    int nx, ny, nz;
    double Lx, Ly, Lz, rn2;
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

__global__ __launch_bounds__(FM_BLOCK) void k_nlist(NlArgs a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.np) return;
    const int ci = a.cell_of[i];
    const int cx = ci % a.nx, cy = (ci / a.nx) % a.ny, cz = ci / (a.nx * a.ny);
    const double xi = a.s.x[i], yi = a.s.y[i], zi = a.s.z[i];
    const size_t cap = a.cap;
    int cnt = 0, nb = 0;
    int chain = 0, funcy = 0;
    if (a.mono) {
        chain = a.mono[i];
        funcy = a.mono[i + 2 * cap];
    }
    for (int dz = -1; dz <= 1; ++dz) {
        int zz = cz + dz, sz = 0;
        if (zz < 0) { zz += a.nz; sz = -1; } else if (zz >= a.nz) { zz -= a.nz; sz = 1; }
        for (int dy = -1; dy <= 1; ++dy) {
            int yy = cy + dy, sy = 0;
            if (yy < 0) { yy += a.ny; sy = -1; } else if (yy >= a.ny) { yy -= a.ny; sy = 1; }
            for (int dx = -1; dx <= 1; ++dx) {
                int xx = cx + dx, sx = 0;
                if (xx < 0) { xx += a.nx; sx = -1; } else if (xx >= a.nx) { xx -= a.nx; sx = 1; }
                const int c = xx + a.nx * (yy + a.ny * zz);
                const uint32_t code = (uint32_t)shift_code(sx, sy, sz);
                const int jb = a.start[c], je = a.start[c + 1];
                for (int j = jb; j < je; ++j) {
                    if (j == i) continue;
                    double rx = xi - img(a.s.x[j], sx, a.Lx);
                    double ry = yi - img(a.s.y[j], sy, a.Ly);
                    double rz = zi - img(a.s.z[j], sz, a.Lz);
                    double r2 = (rx * rx + ry * ry) + rz * rz;
                    if (r2 < a.rn2) {
                        if (cnt < a.maxnb) a.nbl[(size_t)cnt * cap + i] = ((uint32_t)j << 5) | code;
                        ++cnt;
                    }
                    if (a.mono && a.mono[j] == chain) {
                        if (get_bondflag(a.mono, a.cap, i, a.mono[j + cap])) {
                            if (nb < funcy && nb < FM_MAX_FUNCY)
                                a.bond[(size_t)nb * cap + i] = ((uint32_t)j << 5) | code;
                            ++nb;
                        }
                    }
                }
            }
        }
    }
    a.nbn[i] = cnt < a.maxnb ? cnt : a.maxnb;
    atomicMax(&a.flags[FLAG_MAX_NB], cnt);
    if (cnt > a.maxnb) set_err(a.flags, FM_EOVERFLOW, a.ext[i]);
    if (a.mono) {
        a.bondn[i] = nb < funcy ? nb : funcy;
        if (nb > funcy) set_err(a.flags, FM_EBOND, a.ext[i]);
    }
}

// ---- forces, src/simulation_compute_forces.f90:628-726 (pairs), 731-789 (bonds) ---------------
struct FArgs {
    St s;
    const uint32_t* nbl;
    const int* nbn;
    const uint32_t* bond;
    const int* bondn;
    const int* ext;
    const int* cell_of;
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
    int np, cap, nx, ny;
    int rec, ev;           // record step (PE, virial, tallies) / virial-only step
    double Lx, Ly, Lz, hx, hz, rc2, potshift;
    double k_c, R0, R02, dt;
    double teth_k2, teth_k4, teth_k6;
    // MOP
    int nplanes, pn;
    double hp, spacing, shift;
    // Mie tables
    const double* mie;     // [7][FM_MAX_TYPES^2]: eps, lr, la, C, alpha, potshift, rc2 ; + mass[4]
};

// forward cell offsets of the reference's half-list scan (src/linklist.f90:686-688)
__device__ inline bool is_forward(int dx, int dy, int dz) {
    if (dz == 1) return true;
    if (dz == -1) return false;
    if (dy == 1) return true;
    if (dy == -1) return dx == 1;
    return dx == 1;
}

__device__ inline double dp_heaviside(double x) { return ceil(copysign(0.5, x)); }

// pressure_tensor_forces_MOP (src/simulation_record.f90:7068-7112) for one list entry
__device__ void mop_entry(const FArgs& a, const double ri[3], const double rj[3], const double rij[3],
                          double acc, double* spxy) {
    int pi = (int)llround((ri[1] + a.hp - a.shift) / a.spacing) + 1;
    pi = pi < 1 ? 1 : (pi > a.nplanes ? a.nplanes : pi);
    int pj = (int)llround((rj[1] + a.hp - a.shift) / a.spacing) + 1;
    pj = pj < 1 ? 1 : (pj > a.nplanes ? a.nplanes : pj);
    int step = (pj - pi) >= 0 ? 1 : -1;
    for (int n = pi;; n += step) {
        double plane = a.planes[n - 1];
        double pyb0 = ri[0] + (rij[0] / rij[1]) * (plane - ri[1]);
        double pyb2 = ri[2] + (rij[2] / rij[1]) * (plane - ri[1]);
        double fx = copysign(1.0, ri[1] - plane) - copysign(1.0, rj[1] - plane);
        double hxx = dp_heaviside(a.hx - pyb0) - dp_heaviside(-a.hx - pyb0);
        double hzz = dp_heaviside(a.hz - pyb2) - dp_heaviside(-a.hz - pyb2);
        if (fx != 0.0 && hxx != 0.0 && hzz != 0.0) {
            for (int d = 0; d < 3; ++d) atomicAdd(&spxy[3 * (n - 1) + d], acc * rij[d] * fx * hxx * hzz);
        }
        if (n == pj) break;
    }
}

template <bool MIE, bool RECORD>
__global__ __launch_bounds__(FM_BLOCK) void k_forces_fp64(FArgs a) {
    extern __shared__ double spxy[];
    const bool mop = RECORD && a.pxy != nullptr;
    if (mop) {
        for (int t = threadIdx.x; t < 3 * a.nplanes; t += blockDim.x) spxy[t] = 0.0;
        __syncthreads();
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.np) {
        const size_t cap = a.cap;
        const double xi = a.s.x[i], yi = a.s.y[i], zi = a.s.z[i];
        const int n = a.nbn[i];
        double fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0, vir = 0.0;
        double rf[9];
        const bool do_rf = RECORD && a.rf != nullptr;
        if (do_rf)
            for (int q = 0; q < 9; ++q) rf[q] = 0.0;
        int ti = 0;
        double mi = 1.0;
        if (MIE) {
            ti = a.moltype[i];
            mi = a.mie[7 * FM_MAX_TYPES * FM_MAX_TYPES + ti];
        }
        const int exti = mop ? a.ext[i] : 0;
        const int ci = mop ? a.cell_of[i] : 0;
        for (int k = 0; k < n; ++k) {
            const uint32_t w = a.nbl[(size_t)k * cap + i];
            const int j = (int)(w >> 5);
            int sx, sy, sz;
            decode_shift((int)(w & 31u), sx, sy, sz);
            const double rj[3] = {img(a.s.x[j], sx, a.Lx), img(a.s.y[j], sy, a.Ly), img(a.s.z[j], sz, a.Lz)};
            const double rx = xi - rj[0], ry = yi - rj[1], rz = zi - rj[2];
            const double r2 = (rx * rx + ry * ry) + rz * rz;
            double rc2 = a.rc2;
            int t = 0;
            if (MIE) {
                t = ti * FM_MAX_TYPES + a.moltype[j];
                double c = a.mie[6 * FM_MAX_TYPES * FM_MAX_TYPES + t];
                if (c > 0.0) rc2 = c;
            }
            if (!(r2 < rc2)) continue;
            const double inv = 1.0 / r2;
            double acc, e = 0.0;
            if (!MIE) {
                acc = 48.0 * (p7(inv) - 0.5 * p4(inv));
                if (RECORD) e = 4.0 * (p6(inv) - p3(inv)) - a.potshift;
            } else {
                const int T2 = FM_MAX_TYPES * FM_MAX_TYPES;
                double eps = a.mie[t], lr = a.mie[T2 + t], la = a.mie[2 * T2 + t], C = a.mie[3 * T2 + t],
                       al = a.mie[4 * T2 + t];
                acc = C * eps * (lr * pow(inv, 0.5 * lr + 1) - al * la * pow(inv, 0.5 * la + 1));
                if (RECORD) e = C * eps * (pow(inv, 0.5 * lr) - al * pow(inv, 0.5 * la)) - a.mie[5 * T2 + t];
            }
            if (MIE) {
                fx = fx + acc * rx / mi;
                fy = fy + acc * ry / mi;
                fz = fz + acc * rz / mi;
            } else {  // mass 1: x/1.0 == x
                fx = fx + acc * rx;
                fy = fy + acc * ry;
                fz = fz + acc * rz;
            }
            if (RECORD || a.ev) vir = vir + acc * r2;
            if (RECORD) {
                pe = pe + e;
                if (do_rf) {
                    const double rr[3] = {rx, ry, rz};
                    for (int p = 0; p < 3; ++p)
                        for (int q = 0; q < 3; ++q) rf[3 * p + q] = rf[3 * p + q] + acc * rr[p] * rr[q];
                }
                if (mop) {
                    // the reference visits each unshifted pair once, from the list of the molecule
                    // whose cell scans the other's as a forward cell (same cell: the higher molno)
                    bool own = true;
                    if (sx == 0 && sy == 0 && sz == 0) {
                        const int cj = a.cell_of[j];
                        if (cj == ci) own = exti > a.ext[j];
                        else {
                            int ddx = cj % a.nx - ci % a.nx, ddy = (cj / a.nx) % a.ny - (ci / a.nx) % a.ny,
                                ddz = cj / (a.nx * a.ny) - ci / (a.nx * a.ny);
                            own = is_forward(ddx, ddy, ddz);
                        }
                    }
                    if (own) {
                        const double ri[3] = {xi, yi, zi}, rij[3] = {rx, ry, rz};
                        mop_entry(a, ri, rj, rij, acc, spxy);
                    }
                }
            }
        }
        if (a.bondn) {  // FENE, simulation_compute_forces_poly
            const int nb = a.bondn[i];
            double pp = 0.0;
            for (int b = 0; b < nb; ++b) {
                const uint32_t w = a.bond[(size_t)b * cap + i];
                const int j = (int)(w >> 5);
                int sx, sy, sz;
                decode_shift((int)(w & 31u), sx, sy, sz);
                const double rx = xi - img(a.s.x[j], sx, a.Lx), ry = yi - img(a.s.y[j], sy, a.Ly),
                             rz = zi - img(a.s.z[j], sz, a.Lz);
                const double r2 = (rx * rx + ry * ry) + rz * rz;
                if (r2 >= a.R02) {
                    set_err(a.flags, FM_EBOND, a.ext[i]);
                    continue;
                }
                const double acc = -a.k_c / (1.0 - (r2 / a.R02));
                fx = fx + acc * rx / mi;
                fy = fy + acc * ry / mi;
                fz = fz + acc * rz / mi;
                if (RECORD || a.ev) vir = vir + acc * r2;
                if (RECORD) {
                    pp = pp + 0.5 * a.k_c * a.R0 * a.R0 * log(1.0 - (r2 / a.R02));
                    if (do_rf) {
                        const double rr[3] = {rx, ry, rz};
                        for (int p = 0; p < 3; ++p)
                            for (int q = 0; q < 3; ++q) rf[3 * p + q] = rf[3 * p + q] + acc * rr[p] * rr[q];
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
        a.s.ax[i] = fx;
        a.s.ay[i] = fy;
        a.s.az[i] = fz;
        if (RECORD || a.ev) a.vir[i] = vir;
        if (RECORD) {
            a.pe_lj[i] = pe;
            if (do_rf)
                for (int q = 0; q < 9; ++q) a.rf[i + q * cap] = rf[q];
        }
    }
    if (mop) {
        __syncthreads();
        for (int t = threadIdx.x; t < 3 * a.nplanes; t += blockDim.x)
            if (spxy[t] != 0.0) atomicAdd(&a.pxy[t], spxy[t]);
    }
}

// Mixed precision: image coordinates and differences in FP64, the pair law in FP32, per-molecule
// force summed in FP32; energy, virial and tallies accumulated in FP64.
template <bool RECORD>
__global__ __launch_bounds__(FM_BLOCK) void k_forces_mixed(FArgs a) {
    extern __shared__ double spxy[];
    const bool mop = RECORD && a.pxy != nullptr;
    if (mop) {
        for (int t = threadIdx.x; t < 3 * a.nplanes; t += blockDim.x) spxy[t] = 0.0;
        __syncthreads();
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.np) {
        const size_t cap = a.cap;
        const double xi = a.s.x[i], yi = a.s.y[i], zi = a.s.z[i];
        const int n = a.nbn[i];
        float fx = 0.f, fy = 0.f, fz = 0.f;
        double pe = 0.0, vir = 0.0;
        const float rc2 = (float)a.rc2;
        const int exti = mop ? a.ext[i] : 0;
        const int ci = mop ? a.cell_of[i] : 0;
        for (int k = 0; k < n; ++k) {
            const uint32_t w = a.nbl[(size_t)k * cap + i];
            const int j = (int)(w >> 5);
            int sx, sy, sz;
            decode_shift((int)(w & 31u), sx, sy, sz);
            const double rj[3] = {img(a.s.x[j], sx, a.Lx), img(a.s.y[j], sy, a.Ly), img(a.s.z[j], sz, a.Lz)};
            const double rxd = xi - rj[0], ryd = yi - rj[1], rzd = zi - rj[2];
            const float rx = (float)rxd, ry = (float)ryd, rz = (float)rzd;
            float acc, inv;
            {
#pragma clang fp contract(fast)
                const float r2 = rx * rx + ry * ry + rz * rz;
                if (!(r2 < rc2)) continue;
                inv = __builtin_amdgcn_rcpf(r2);
                const float i2 = inv * inv, i4 = i2 * i2, i3 = i2 * inv;
                acc = 48.f * (i4 * i3 - 0.5f * i4);
                fx += acc * rx;
                fy += acc * ry;
                fz += acc * rz;
            }
            if (RECORD || a.ev) {
                const double r2d = (rxd * rxd + ryd * ryd) + rzd * rzd;
                vir = vir + (double)acc * r2d;
            }
            if (RECORD) {
                const double id = (double)inv;
                pe = pe + (4.0 * (p6(id) - p3(id)) - a.potshift);
                if (mop) {
                    bool own = true;
                    if (sx == 0 && sy == 0 && sz == 0) {
                        const int cj = a.cell_of[j];
                        if (cj == ci) own = exti > a.ext[j];
                        else {
                            int ddx = cj % a.nx - ci % a.nx, ddy = (cj / a.nx) % a.ny - (ci / a.nx) % a.ny,
                                ddz = cj / (a.nx * a.ny) - ci / (a.nx * a.ny);
                            own = is_forward(ddx, ddy, ddz);
                        }
                    }
                    if (own) {
                        const double ri[3] = {xi, yi, zi}, rij[3] = {rxd, ryd, rzd};
                        mop_entry(a, ri, rj, rij, (double)acc, spxy);
                    }
                }
            }
        }
        double dfx = fx, dfy = fy, dfz = fz;
        if (a.tag) {
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
        if (RECORD || a.ev) a.vir[i] = vir;
        if (RECORD) a.pe_lj[i] = pe;
    }
    if (mop) {
        __syncthreads();
        for (int t = threadIdx.x; t < 3 * a.nplanes; t += blockDim.x)
            if (spxy[t] != 0.0) atomicAdd(&a.pxy[t], spxy[t]);
    }
}

// ---- macroscopic sums, src/simulation_record.f90:1205-1314 ------------------------------------
// 7 sums per block: pe_lj, pe_poly, pe_lj+pe_poly, vir, mv2, vsum, msum ; + max(pe)
#define NRED 8
__global__ __launch_bounds__(FM_BLOCK) void k_macros_partial(St s, const double* pe_lj, const double* pe_poly,
                                                             const double* vir, const int* moltype,
                                                             const double* mass, int np, double* part) {
    __shared__ double red[NRED][FM_BLOCK / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[NRED] = {0, 0, 0, 0, 0, 0, 0, -1.0e300};
    if (i < np) {
        double m = moltype ? mass[moltype[i]] : 1.0;
        double pl = pe_lj[i], pp = pe_poly ? pe_poly[i] : 0.0;
        v[0] = pl;
        v[1] = pp;
        v[2] = pl + pp;
        v[3] = vir[i];
        double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
        v[4] = (m * (vx * vx) + m * (vy * vy)) + m * (vz * vz);
        v[5] = (m * vx + m * vy) + m * vz;
        v[6] = m;
        v[7] = pl + pp;
    }
    for (int q = 0; q < NRED; ++q) {
        double x = v[q];
        for (int off = 32; off > 0; off >>= 1) {
            double o = __shfl_down(x, off, 64);
            x = q == 7 ? fmax(x, o) : x + o;
        }
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < NRED) {
        int q = threadIdx.x;
        double x = red[q][0];
        for (int w = 1; w < FM_BLOCK / 64; ++w) x = q == 7 ? fmax(x, red[q][w]) : x + red[q][w];
        part[(size_t)blockIdx.x * NRED + q] = x;
    }
}

__global__ __launch_bounds__(64) void k_macros_final(const double* part, int nb, double* out) {
    int q = threadIdx.x;
    if (q >= NRED) return;
    double x = q == 7 ? -1.0e300 : 0.0;
    for (int b = 0; b < nb; ++b) x = q == 7 ? fmax(x, part[(size_t)b * NRED + q]) : x + part[(size_t)b * NRED + q];
    out[q] = x;
}

// ---- host-order staging ---------------------------------------------------------------------
__global__ __launch_bounds__(FM_BLOCK) void k_to_host(St s, const int* ext, int np, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    size_t e = ext[i];
    const size_t n3 = 3 * (size_t)np;
    out[3 * e] = s.x[i];
    out[3 * e + 1] = s.y[i];
    out[3 * e + 2] = s.z[i];
    out[n3 + 3 * e] = s.vx[i];
    out[n3 + 3 * e + 1] = s.vy[i];
    out[n3 + 3 * e + 2] = s.vz[i];
    out[2 * n3 + 3 * e] = s.ax[i];
    out[2 * n3 + 3 * e + 1] = s.ay[i];
    out[2 * n3 + 3 * e + 2] = s.az[i];
}

// host (molecule-number) order <-> slot order
__global__ __launch_bounds__(FM_BLOCK) void k_from_host(St s, int* ext, int* slot_of, int np, const double* in) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= np) return;
    const size_t n3 = 3 * (size_t)np;
    s.x[e] = in[3 * (size_t)e];
    s.y[e] = in[3 * (size_t)e + 1];
    s.z[e] = in[3 * (size_t)e + 2];
    s.vx[e] = in[n3 + 3 * (size_t)e];
    s.vy[e] = in[n3 + 3 * (size_t)e + 1];
    s.vz[e] = in[n3 + 3 * (size_t)e + 2];
    s.ax[e] = in[2 * n3 + 3 * (size_t)e];
    s.ay[e] = in[2 * n3 + 3 * (size_t)e + 1];
    s.az[e] = in[2 * n3 + 3 * (size_t)e + 2];
    ext[e] = e;
    slot_of[e] = e;
}

// dst[s + f*cap] = src[ext[s]*nf + f]
template <typename T>
__global__ __launch_bounds__(FM_BLOCK) void k_ext_to_slot(T* dst, int cap, const T* src, int nf, const int* ext, int np) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    size_t e = ext[s];
    for (int f = 0; f < nf; ++f) dst[s + (size_t)f * cap] = src[e * nf + f];
}

// dst[ext[s]*nf + f] = src[s + f*cap]
template <typename T>
__global__ __launch_bounds__(FM_BLOCK) void k_slot_to_ext(T* dst, const T* src, int cap, int nf, const int* ext, int np) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= np) return;
    size_t e = ext[s];
    for (int f = 0; f < nf; ++f) dst[e * nf + f] = src[s + (size_t)f * cap];
}

}  // namespace

// =============================================================================================
// launchers
#define LAUNCH_CHECK()                                                                       \
    do {                                                                                     \
        hipError_t _e = hipGetLastError();                                                   \
        if (_e != hipSuccess) return fm_set_error(FM_EHIP, "kernel launch: %s (%s:%d)",       \
                                                  hipGetErrorString(_e), __FILE__, __LINE__); \
    } while (0)

int launch_fill_i32(fm_ctx* c, int* p, int v, size_t n) {
    if (n == 0) return 0;
    k_fill_i32<<<nblocks(n), FM_BLOCK, 0, c->stream>>>(p, v, n);
    LAUNCH_CHECK();
    return 0;
}

int launch_fill_f64(fm_ctx* c, double* p, double v, size_t n) {
    if (n == 0) return 0;
    k_fill_f64<<<nblocks(n), FM_BLOCK, 0, c->stream>>>(p, v, n);
    LAUNCH_CHECK();
    return 0;
}

int launch_pass1(fm_ctx* c) {
    if (c->np == 0) return 0;
    double dt = c->p.delta_t;
    ktimer_begin(c, KT_PASS1);
    k_pass1<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(state_of(c->x(), c->cap), c->np, dt, 0.5 * (dt * dt),
                                                         0.5 * dt);
    ktimer_end(c, KT_PASS1);
    LAUNCH_CHECK();
    return 0;
}

int launch_pass2(fm_ctx* c) {
    if (c->np == 0) return 0;
    ktimer_begin(c, KT_PASS2);
    k_pass2<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(state_of(c->x(), c->cap), c->np, 0.5 * c->p.delta_t);
    ktimer_end(c, KT_PASS2);
    LAUNCH_CHECK();
    return 0;
}

int launch_check_disp(fm_ctx* c) {
    if (c->np == 0) return 0;
    ktimer_begin(c, KT_CHECK);
    k_check_disp<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(state_of(c->x(), c->cap), c->rdisp, c->np, c->cap,
                                                              c->p.delta_t, 0.5 * c->dmin,
                                                              c->rdisp_valid ? 0 : 1, c->d_flags);
    ktimer_end(c, KT_CHECK);
    LAUNCH_CHECK();
    c->rdisp_valid = true;
    return 0;
}

int launch_check_vmax(fm_ctx* c) {
    ktimer_begin(c, KT_CHECK);
    if (c->np > 0) k_vmax<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(state_of(c->x(), c->cap), c->np, c->d_vmax);
    k_vmax_finish<<<1, 1, 0, c->stream>>>(c->d_vmax, c->d_rmax, c->p.delta_t, 0.5 * c->dmin, 0, c->d_flags);
    ktimer_end(c, KT_CHECK);
    LAUNCH_CHECK();
    return 0;
}

int launch_sendmols(fm_ctx* c) {
    if (c->np == 0) return 0;
    ktimer_begin(c, KT_SENDMOLS);
    St s = state_of(c->x(), c->cap);
    k_out_flags<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(s, c->slot_of, c->np, c->half[0], c->half[1],
                                                             c->half[2], c->out_flag, c->d_flags);
    LAUNCH_CHECK();
    if (int rc = cub_select_flagged_index(c, c->out_flag, c->out_vals, c->out_count, c->np)) return rc;
    const bool tm = c->p.ensemble == FM_ENSEMBLE_TAG_MOVE && c->has_tags;
    // lst and tmp reuse the cell permutation buffer and the flag buffer
    k_sendmols_passes<<<1, 1024, 0, c->stream>>>(
        s, tm ? c->teth[c->cur] : nullptr, tm ? c->tag[c->cur] : nullptr, c->ext[c->cur], c->slot_of,
        c->out_vals, c->out_slot, c->perm, c->out_flag, c->out_count, c->cap, c->p.domain[0], c->p.domain[1],
        c->p.domain[2], tm ? 1 : 0, c->p.wallslidev[0], c->p.wallslidev[1], c->p.wallslidev[2]);
    ktimer_end(c, KT_SENDMOLS);
    LAUNCH_CHECK();
    return 0;
}

int launch_bin(fm_ctx* c) {
    if (c->np == 0) return 0;
    ktimer_begin(c, KT_BIN);
    St s = state_of(c->x(), c->cap);
    const int nxt = 1 - c->cur;
    int* cell_tmp = c->out_flag;  // cell index in old slot order
    if (int rc = launch_fill_i32(c, c->cell_count, 0, (size_t)c->ncell + 1)) return rc;
    if (int rc = launch_fill_i32(c, c->cell_cursor, 0, (size_t)c->ncell)) return rc;
    k_bin_count<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(s, c->ext[c->cur], c->np, c->half[0], c->half[1],
                                                             c->half[2], c->cs[0], c->cs[1], c->cs[2], c->nc[0],
                                                             c->nc[1], c->nc[2], cell_tmp, c->cell_count,
                                                             c->d_flags);
    LAUNCH_CHECK();
    if (int rc = cub_exclusive_scan(c, c->cell_count, c->cell_start, c->ncell + 1)) return rc;
    k_bin_scatter<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(cell_tmp, c->cell_start, c->cell_cursor, c->perm,
                                                               c->np);
    LAUNCH_CHECK();
    int maxcell = c->p.max_cell_np > 0 ? c->p.max_cell_np : (1 << 26);
    k_bin_sort<<<nblocks(c->ncell), FM_BLOCK, 0, c->stream>>>(c->cell_start, c->perm, c->ext[c->cur], c->ncell,
                                                               maxcell, c->d_flags);
    LAUNCH_CHECK();
    PermArgs a;
    a.src = c->st[c->cur];
    a.dst = c->st[nxt];
    a.ext_src = c->ext[c->cur];
    a.ext_dst = c->ext[nxt];
    a.cell_src = cell_tmp;
    a.cell_dst = c->cell_of;
    a.mt_src = c->has_moltype ? c->moltype[c->cur] : nullptr;
    a.mt_dst = c->has_moltype ? c->moltype[nxt] : nullptr;
    a.tag_src = c->has_tags ? c->tag[c->cur] : nullptr;
    a.tag_dst = c->has_tags ? c->tag[nxt] : nullptr;
    a.teth_src = c->has_tags ? c->teth[c->cur] : nullptr;
    a.teth_dst = c->has_tags ? c->teth[nxt] : nullptr;
    a.mono_src = c->has_mono ? c->mono[c->cur] : nullptr;
    a.mono_dst = c->has_mono ? c->mono[nxt] : nullptr;
    a.slot_of = c->slot_of;
    a.perm = c->perm;
    a.np = c->np;
    a.cap = c->cap;
    k_permute<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(a);
    ktimer_end(c, KT_BIN);
    LAUNCH_CHECK();
    c->cur = nxt;
    return 0;
}

int launch_nlist(fm_ctx* c) {
    if (c->np == 0) return 0;
    ktimer_begin(c, KT_NLIST);
    NlArgs a;
    a.s = state_of(c->x(), c->cap);
    a.cell_of = c->cell_of;
    a.start = c->cell_start;
    a.ext = c->ext[c->cur];
    a.mono = (c->p.potential_flag == 1 && c->has_mono) ? c->mono[c->cur] : nullptr;
    a.nbl = c->nbl;
    a.nbn = c->nbn;
    a.bond = c->bond;
    a.bondn = c->bondn;
    a.flags = c->d_flags;
    a.np = c->np;
    a.cap = c->cap;
    a.maxnb = c->maxnb;
    a.nx = c->nc[0];
    a.ny = c->nc[1];
    a.nz = c->nc[2];
    a.Lx = c->p.domain[0];
    a.Ly = c->p.domain[1];
    a.Lz = c->p.domain[2];
    a.rn2 = c->rn2;
    k_nlist<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(a);
    ktimer_end(c, KT_NLIST);
    LAUNCH_CHECK();
    return 0;
}

int launch_forces(fm_ctx* c, int iter) {
    if (c->np == 0) return 0;
    FArgs a;
    const int cur = c->cur;
    a.s = state_of(c->x(), c->cap);
    a.nbl = c->nbl;
    a.nbn = c->nbn;
    const bool poly = c->p.potential_flag == 1 && c->has_mono;
    a.bond = poly ? c->bond : nullptr;
    a.bondn = poly ? c->bondn : nullptr;
    a.ext = c->ext[cur];
    a.cell_of = c->cell_of;
    const bool mie = c->p.mie_potential == 1;
    a.moltype = mie ? c->moltype[cur] : nullptr;
    const bool tm = c->p.ensemble == FM_ENSEMBLE_TAG_MOVE && c->has_tags;
    a.tag = tm ? c->tag[cur] : nullptr;
    a.teth = tm ? c->teth[cur] : nullptr;
    a.pe_lj = c->pe_lj;
    a.pe_poly = c->pe_poly;
    a.vir = c->vir;
    a.flags = c->d_flags;
    a.np = c->np;
    a.cap = c->cap;
    a.nx = c->nc[0];
    a.ny = c->nc[1];
    const int tplot = c->p.tplot, teval = c->p.teval;
    a.rec = (iter % tplot) == 0;
    a.ev = teval > 0 && (iter % teval) == 0;
    a.rf = (a.rec && c->p.pressure_outflag == 1) ? c->rf : nullptr;
    const bool mop = a.rec && c->p.vflux_outflag >= 1 && c->p.vflux_outflag <= 3 && c->pxy;
    a.pxy = mop ? c->pxy : nullptr;
    a.planes = c->planes;
    a.Lx = c->p.domain[0];
    a.Ly = c->p.domain[1];
    a.Lz = c->p.domain[2];
    a.hx = c->half[0];
    a.hz = c->half[2];
    a.rc2 = c->rc2;
    a.potshift = c->potshift;
    a.k_c = c->p.k_c;
    a.R0 = c->p.R_0;
    a.R02 = c->p.R_0 * c->p.R_0;
    a.dt = c->p.delta_t;
    a.teth_k2 = c->p.teth_k2;
    a.teth_k4 = c->p.teth_k4;
    a.teth_k6 = c->p.teth_k6;
    a.nplanes = c->p.nplanes;
    a.pn = c->p.vflux_outflag - 1;
    a.hp = (a.pn >= 0 && a.pn < 3) ? c->half[a.pn] : 0.0;
    a.spacing = c->p.planespacing;
    a.shift = c->mop_shift;
    a.mie = mie ? c->d_mie : nullptr;
    size_t shm = mop ? 3 * (size_t)c->p.nplanes * sizeof(double) : 0;
    ktimer_begin(c, KT_FORCES);
    dim3 g(nblocks(c->np)), b(FM_BLOCK);
    if (c->p.precision == FM_PREC_MIXED && !mie && !poly) {
        if (a.rec) k_forces_mixed<true><<<g, b, shm, c->stream>>>(a);
        else k_forces_mixed<false><<<g, b, shm, c->stream>>>(a);
    } else if (mie) {
        if (a.rec) k_forces_fp64<true, true><<<g, b, shm, c->stream>>>(a);
        else k_forces_fp64<true, false><<<g, b, shm, c->stream>>>(a);
    } else {
        if (a.rec) k_forces_fp64<false, true><<<g, b, shm, c->stream>>>(a);
        else k_forces_fp64<false, false><<<g, b, shm, c->stream>>>(a);
    }
    ktimer_end(c, KT_FORCES);
    LAUNCH_CHECK();
    return 0;
}

int launch_macros(fm_ctx* c) {
    int nb = nblocks(c->np);
    if (nb == 0) nb = 1;
    ktimer_begin(c, KT_MACROS);
    const bool mie = c->p.mie_potential == 1;
    const bool poly = c->p.potential_flag == 1 && c->has_mono;
    k_macros_partial<<<nb, FM_BLOCK, 0, c->stream>>>(state_of(c->x(), c->cap), c->pe_lj, poly ? c->pe_poly : nullptr,
                                                      c->vir, mie ? c->moltype[c->cur] : nullptr,
                                                      c->d_mie + 7 * FM_MAX_TYPES * FM_MAX_TYPES,
                                                      c->np, c->red);
    k_macros_final<<<1, 64, 0, c->stream>>>(c->red, nb, c->d_macros);
    ktimer_end(c, KT_MACROS);
    LAUNCH_CHECK();
    return 0;
}

int launch_scatter_state_to_host_order(fm_ctx* c, double* d_out9, double*) {
    if (c->np == 0) return 0;
    k_to_host<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(state_of(c->x(), c->cap), c->ext[c->cur], c->np, d_out9);
    LAUNCH_CHECK();
    return 0;
}

int launch_from_host(fm_ctx* c, const double* d_in9) {
    if (c->np == 0) return 0;
    k_from_host<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(state_of(c->x(), c->cap), c->ext[c->cur], c->slot_of,
                                                             c->np, d_in9);
    LAUNCH_CHECK();
    return 0;
}

int launch_ext_to_slot_i32(fm_ctx* c, int* dst, const int* src, int nf) {
    if (c->np == 0) return 0;
    k_ext_to_slot<int><<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(dst, c->cap, src, nf, c->ext[c->cur], c->np);
    LAUNCH_CHECK();
    return 0;
}

int launch_ext_to_slot_f64(fm_ctx* c, double* dst, const double* src, int nf) {
    if (c->np == 0) return 0;
    k_ext_to_slot<double><<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(dst, c->cap, src, nf, c->ext[c->cur], c->np);
    LAUNCH_CHECK();
    return 0;
}

int launch_slot_to_ext_f64(fm_ctx* c, double* dst, const double* src, int nf) {
    if (c->np == 0) return 0;
    k_slot_to_ext<double><<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(dst, src, c->cap, nf, c->ext[c->cur], c->np);
    LAUNCH_CHECK();
    return 0;
}

int launch_slot_to_ext_i32(fm_ctx* c, int* dst, const int* src, int nf) {
    if (c->np == 0) return 0;
    k_slot_to_ext<int><<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(dst, src, c->cap, nf, c->ext[c->cur], c->np);
    LAUNCH_CHECK();
    return 0;
}
