// fm_kernels.hip -- gfx950 kernels of the force / neighbour / integration path.
//
// Parity-critical FP64 arithmetic (cell index, image coordinates, squared distances, velocity
// Verlet) is written in the reference's operation order and compiled without fused multiply-add
// (file-wide fp contract off), so it is bit-identical to the reference's FP64 operations.
// The mixed-precision pair law re-enables contraction locally.
#include "fm_internal.h"

#include <limits.h>
#include <math.h>

#include <mutex>

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
// halt: flag words of a conditional launch (does nothing while FLAG_HALT is set), or null
__global__ __launch_bounds__(FM_BLOCK) void k_pass1(St s, int np, double dt, double hdt2, double hdt, const int* halt) {
    if (halt && halt[FLAG_HALT] != 0) return;  // whole workgroup
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

// ---- simulation_move_particles_lfv, src/simulation_move_particles_lfv.f90:38-177 ---------------
// NVE, or tag_move through simulation_move_particles_lfv_tag (308-490) for the tags without a
// thermostat; tethered molecules get tether_force (src/molecular_properties.f90:833-915) here,
// inside the move.  With disp != 0 the displacement update of simulation_checkrebuild (rebuild
// criterion 2) follows in the same pass.
__global__ __launch_bounds__(FM_BLOCK) void k_move_lfv(St s, const int* tag, double* teth, const int* ext, int np,
                                                        int cap, double dt, double k2, double k4, double k6,
                                                        double* rd, double thresh, int disp, int fresh,
                                                        int* flags, int cond) {
    if (cond && flags[FLAG_HALT] != 0) return;  // whole launch: a queued rebuild is pending
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool hit = false;
    if (i < np) {
        const int tg = tag ? tag[i] : FM_TAG_FREE;
        double r[3] = {s.x[i], s.y[i], s.z[i]};
        double v[3] = {s.vx[i], s.vy[i], s.vz[i]};
        double a[3] = {s.ax[i], s.ay[i], s.az[i]};
        const size_t c = (size_t)cap;
        if (tg == FM_TAG_TETH || tg == FM_TAG_TETH_SLIDE) {
            double rio[3];
            for (int d = 0; d < 3; ++d) rio[d] = r[d] - teth[i + d * c];
            const double mag = sqrt((rio[0] * rio[0] + rio[1] * rio[1]) + rio[2] * rio[2]);
            const double am = -2.0 * k2 * mag - 4.0 * k4 * (mag * mag) - 6.0 * k6 * pow(mag, 4.0);
            for (int d = 0; d < 3; ++d) a[d] = a[d] + am * rio[d];
            for (int d = 0; d < 3; ++d) teth[i + d * c] = teth[i + d * c] + teth[i + (3 + d) * c] * dt;
            s.ax[i] = a[0];
            s.ay[i] = a[1];
            s.az[i] = a[2];
        }
        switch (tg) {
        case FM_TAG_FREE:
        case FM_TAG_TETH:
            for (int d = 0; d < 3; ++d) v[d] = v[d] + dt * a[d];
            for (int d = 0; d < 3; ++d) r[d] = r[d] + dt * v[d];
            break;
        case FM_TAG_TETH_SLIDE:
            for (int d = 0; d < 3; ++d) v[d] = v[d] + dt * a[d];
            for (int d = 0; d < 3; ++d) r[d] = r[d] + dt * v[d] + dt * teth[i + (3 + d) * c];
            break;
        case FM_TAG_FIXED_SLIDE:
            for (int d = 0; d < 3; ++d) r[d] = r[d] + dt * teth[i + (3 + d) * c];
            break;
        case FM_TAG_FIXED:
            break;
        default:
            set_err(flags, FM_EINVAL, ext[i]);
        }
        s.x[i] = r[0];
        s.y[i] = r[1];
        s.z[i] = r[2];
        s.vx[i] = v[0];
        s.vy[i] = v[1];
        s.vz[i] = v[2];
        if (disp) {
            double d0 = fresh ? 0.0 : rd[i];
            double d1 = fresh ? 0.0 : rd[i + c];
            double d2 = fresh ? 0.0 : rd[i + 2 * c];
            d0 = d0 + v[0] * dt;
            d1 = d1 + v[1] * dt;
            d2 = d2 + v[2] * dt;
            rd[i] = d0;
            rd[i + c] = d1;
            rd[i + 2 * c] = d2;
            hit = fabs(d0) > thresh || fabs(d1) > thresh || fabs(d2) > thresh;
        }
    }
    unsigned long long m = __ballot(hit);
    if (m != 0ull && (int)(threadIdx.x & 63) == __ffsll((long long)m) - 1 &&
        __atomic_load_n(&flags[FLAG_REBUILD], __ATOMIC_RELAXED) == 0)
        atomicOr(&flags[FLAG_REBUILD], 1);
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
    if (m != 0ull && (int)(threadIdx.x & 63) == __ffsll((long long)m) - 1 &&
        __atomic_load_n(&flags[FLAG_REBUILD], __ATOMIC_RELAXED) == 0)
        atomicOr(&flags[FLAG_REBUILD], 1);
}

// simulation_move_particles_vv(1) followed by the displacement update of simulation_checkrebuild,
// in one pass over the state (same operations as k_pass1 then k_check_disp)
// cond: a launch queued ahead of the host's check; it does nothing when an earlier step of the
// queue asked for a rebuild (FLAG_HALT), and records its step when it asks for one itself.
__global__ __launch_bounds__(FM_BLOCK) void k_pass1_disp(St s, double* rd, int np, int cap, double dt, double hdt2,
                                                          double hdt, double thresh, int fresh, int* flags, int cond,
                                                          int step) {
    if (cond && flags[FLAG_HALT] != 0) return;  // whole workgroup
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool hit = false;
    if (i < np) {
        const double ax = s.ax[i], ay = s.ay[i], az = s.az[i];
        double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
        s.x[i] = s.x[i] + dt * vx + hdt2 * ax;
        s.y[i] = s.y[i] + dt * vy + hdt2 * ay;
        s.z[i] = s.z[i] + dt * vz + hdt2 * az;
        vx = vx + hdt * ax;
        vy = vy + hdt * ay;
        vz = vz + hdt * az;
        s.vx[i] = vx;
        s.vy[i] = vy;
        s.vz[i] = vz;
        double d0 = fresh ? 0.0 : rd[i];
        double d1 = fresh ? 0.0 : rd[i + cap];
        double d2 = fresh ? 0.0 : rd[i + 2 * (size_t)cap];
        d0 = d0 + vx * dt;
        d1 = d1 + vy * dt;
        d2 = d2 + vz * dt;
        rd[i] = d0;
        rd[i + cap] = d1;
        rd[i + 2 * (size_t)cap] = d2;
        hit = fabs(d0) > thresh || fabs(d1) > thresh || fabs(d2) > thresh;
    }
    unsigned long long m = __ballot(hit);
    if (m != 0ull && (int)(threadIdx.x & 63) == __ffsll((long long)m) - 1 &&
        __atomic_load_n(&flags[FLAG_REBUILD], __ATOMIC_RELAXED) == 0) {
        atomicOr(&flags[FLAG_REBUILD], 1);
        if (cond) atomicExch(&flags[FLAG_RB_STEP], step);
    }
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
        const unsigned long long bits = (unsigned long long)__double_as_longlong(b);
        if (bits > __atomic_load_n(vmax_bits, __ATOMIC_RELAXED)) atomicMax(vmax_bits, bits);
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
// zero the listed flag words
struct FlagList {
    int n;
    int w[8];
};

// zero one flag word; a reset of FLAG_REBUILD also returns FLAG_RB_STEP to "no step" (INT_MAX),
// which the halt rule of conditional force launches relies on
__device__ inline void reset_flag(int* flags, int w) {
    flags[w] = 0;
    if (w == FLAG_REBUILD) flags[FLAG_RB_STEP] = 0x7fffffff;
    if (w == FLAG_HALT) flags[FLAG_HALT_STEP] = 0x7fffffff;
}

// 1) molecules outside the box, in molecule-number order: workgroup b writes its leavers e
//    (ascending) to seg[b*FM_BLOCK ..] and their count to bcnt[b].  Workgroup 0 also zeroes the
//    flag words of `reset` (the resets of a queued rebuild, folded into its first launch).
__global__ __launch_bounds__(FM_BLOCK) void k_out_select(St s, const int* slot_of, int np, double hx, double hy,
                                                         double hz, int* seg, int* bcnt, int* flags, FlagList reset) {
    __shared__ int wc[FM_BLOCK / 64];
    if (blockIdx.x == 0 && (int)threadIdx.x < reset.n) reset_flag(flags, reset.w[threadIdx.x]);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int out = 0;
    if (e < np) {
        const int sl = slot_of[e];
        const double x = s.x[sl], y = s.y[sl], z = s.z[sl];
        out = (x < -hx) | (x >= hx) | (y < -hy) | (y >= hy) | (z < -hz) | (z >= hz);
    }
    const unsigned long long m = __ballot(out);
    const int rank = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wc[wid] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    for (int w = 0; w < FM_BLOCK / 64; ++w) {
        if (w < wid) off += wc[w];
        tot += wc[w];
    }
    if (out) seg[(size_t)blockIdx.x * FM_BLOCK + off + rank] = e;
    if (threadIdx.x == 0) bcnt[blockIdx.x] = tot;
}

// 2) one workgroup replays the 3 x (3 axes x 2 directions) passes on the compacted set.
//    vals[q] ascending molecule numbers of the set (fixed); sl[q] slot currently holding vals[q].
//    Each pass: leavers in q order (= ascending molecule number) are wrapped by -dir*domain and
//    their molecule numbers reversed, as checksendbuild + sendrecvface + reorderdata do.
//    The set is first gathered from the segments of k_out_select in workgroup order.
__global__ __launch_bounds__(1024) void k_sendmols_passes(St s, double* teth, const int* tag, int* ext,
                                                          int* slot_of, const int* seg, const int* bcnt, int nblk,
                                                          int* vals, int* sl, int* lst, int* tmp,
                                                          int cap, double Lx, double Ly, double Lz,
                                                          int tag_move, double wsx, double wsy,
                                                          double wsz) {
    __shared__ int wsum[16];
    __shared__ int chunk_base;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    {  // thread t gathers the segments of a contiguous range of workgroups
        const int per = (nblk + (int)blockDim.x - 1) / (int)blockDim.x;
        const int b0 = min(nblk, tid * per), b1 = min(nblk, b0 + per);
        int mine = 0;
        for (int b = b0; b < b1; ++b) mine += bcnt[b];
        int x = mine;
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) wsum[wid] = x;
        __syncthreads();
        int pos = x - mine, tot = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            if (w < wid) pos += wsum[w];
            tot += wsum[w];
        }
        for (int b = b0; b < b1; ++b) {
            const int cb = bcnt[b];
            for (int k = 0; k < cb; ++k) vals[pos++] = seg[(size_t)b * FM_BLOCK + k];
        }
        if (tid == 0) chunk_base = tot;
        __syncthreads();
    }
    const int n = chunk_base;
    __syncthreads();
    const double L[3] = {Lx, Ly, Lz};
    double* pos[3] = {s.x, s.y, s.z};
    if (n <= (int)blockDim.x) {
        // One leaver per thread, held in LDS by list position: its slot and position.  The same
        // passes as below, without a global read in them; a pass without leavers ends at its
        // count.  The renumbering moves slots and positions between list positions together.
        __shared__ int cs[1024], cl[1024], ct[1024];
        __shared__ double cp[3][1024], cq[3][1024];
        if (tid < n) {
            const int slot = slot_of[vals[tid]];
            cs[tid] = slot;
            cp[0][tid] = pos[0][slot];
            cp[1][tid] = pos[1][slot];
            cp[2][tid] = pos[2][slot];
        }
        __syncthreads();
        for (int it = 0; it < 3; ++it)
            for (int ax = 0; ax < 3; ++ax)
                for (int dir = -1; dir <= 1; dir += 2) {
                    const double h = 0.5 * L[ax];
                    const bool lv = tid < n && (dir == -1 ? (cp[ax][tid] < -h) : (cp[ax][tid] >= h));
                    const unsigned long long m = __ballot(lv);
                    if (lane == 0) wsum[wid] = __popcll(m);
                    __syncthreads();
                    int off = 0, k = 0;
                    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
                        if (w < wid) off += wsum[w];
                        k += wsum[w];
                    }
                    __syncthreads();  // (wsum is written again by the next pass)
                    if (k == 0) continue;
                    if (lv) cl[off + __popcll(m & ((1ull << lane) - 1ull))] = tid;
                    __syncthreads();
                    // wrap the leavers: r(ixyz) - dir*domain(ixyz)
                    if (tid < k) {
                        const int q = cl[tid], slot = cs[q];
                        cp[ax][q] = cp[ax][q] - (double)dir * L[ax];
                        pos[ax][slot] = cp[ax][q];
                        if (tag_move && (tag[slot] == FM_TAG_TETH || tag[slot] == FM_TAG_TETH_SLIDE))
                            teth[slot + (size_t)ax * cap] = teth[slot + (size_t)ax * cap] - (double)dir * L[ax];
                    }
                    __syncthreads();
                    if (tid < k) {
                        const int q = cl[k - 1 - tid];
                        ct[tid] = cs[q];
                        cq[0][tid] = cp[0][q];
                        cq[1][tid] = cp[1][q];
                        cq[2][tid] = cp[2][q];
                    }
                    __syncthreads();
                    if (tid < k) {
                        const int q = cl[tid], slot = ct[tid];
                        cs[q] = slot;
                        cp[0][q] = cq[0][tid];
                        cp[1][q] = cq[1][tid];
                        cp[2][q] = cq[2][tid];
                        if (tag_move) {  // read_tag on the renumbered molecule
                            const int tg = tag[slot];
                            const double sg = tg == FM_TAG_TETH_SLIDE   ? copysign(1.0, cq[1][tid])
                                              : tg == FM_TAG_FIXED_SLIDE ? 1.0
                                                                         : 0.0;
                            teth[slot + 3 * (size_t)cap] = sg != 0.0 ? wsx * sg : 0.0;
                            teth[slot + 4 * (size_t)cap] = sg != 0.0 ? wsy * sg : 0.0;
                            teth[slot + 5 * (size_t)cap] = sg != 0.0 ? wsz * sg : 0.0;
                        }
                    }
                    __syncthreads();
                }
        if (tid < n) {
            ext[cs[tid]] = vals[tid];
            slot_of[vals[tid]] = cs[tid];
        }
        return;
    }
    for (int q = tid; q < n; q += blockDim.x) sl[q] = slot_of[vals[q]];
    __syncthreads();
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
                        const int tg = tag[slot];
                        const double sg = tg == FM_TAG_TETH_SLIDE   ? copysign(1.0, s.y[slot])
                                          : tg == FM_TAG_FIXED_SLIDE ? 1.0
                                                                     : 0.0;
                        teth[slot + 3 * (size_t)cap] = sg != 0.0 ? wsx * sg : 0.0;
                        teth[slot + 4 * (size_t)cap] = sg != 0.0 ? wsy * sg : 0.0;
                        teth[slot + 5 * (size_t)cap] = sg != 0.0 ? wsz * sg : 0.0;
                    }
                }
                __syncthreads();
            }
    for (int q = tid; q < n; q += blockDim.x) {
        ext[sl[q]] = vals[q];
        slot_of[vals[q]] = sl[q];
    }
}

// Lanes of a wavefront that add to the same counter are served by one atomic: returns each
// active lane's position among the additions to its counter (consecutive slots mostly share a
// cell, so a wavefront issues a handful of atomics instead of 64).
__device__ inline int wave_aggregated_add(int* counters, int key, bool active) {
    const int lane = threadIdx.x & 63;
    unsigned long long mask = __ballot(active);
    int pos = -1;
    while (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        const int lk = __shfl(key, leader, 64);
        const unsigned long long same = __ballot(active && key == lk) & mask;
        int base = 0;
        if (lane == leader) base = atomicAdd(&counters[lk], __popcll(same));
        base = __shfl(base, leader, 64);
        if ((same >> lane) & 1ull) pos = base + __popcll(same & ((1ull << lane) - 1ull));
        mask &= ~same;
    }
    return pos;
}

// ---- assign_to_cell as a counting sort, src/linklist.f90:2645-2676 -----------------------------
__global__ __launch_bounds__(FM_BLOCK) void k_bin_count(St s, const int* ext, int np, double hx, double hy,
                                                        double hz, double csx, double csy, double csz,
                                                        int nx, int ny, int nz, int* cell_of,
                                                        int* count, int* pos_out, int* flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int cell = -1;
    if (i < np) {
        double r[3] = {s.x[i], s.y[i], s.z[i]};
        const double h[3] = {hx, hy, hz}, cs[3] = {csx, csy, csz};
        const int n[3] = {nx, ny, nz};
        int c[3];
        bool ok = true;
        for (int d = 0; d < 3 && ok; ++d) {
            double t = r[d] + h[d];
            if (!(t > -cs[d] && t < (double)(n[d] + 1) * cs[d])) {  // also rejects NaN
                ok = false;
                break;
            }
            int ic = (int)ceil(t / cs[d]) + 1;  // 1-based with one halo layer
            if (ic == 1 && t == 0.0) {  // exactly on the lower face: the reference's ceiling() gives the halo
                atomicAdd(&flags[FLAG_EDGE], 1);
                ic = 2;
            }
            if (ic < 2 || ic > n[d] + 1) {
                ok = false;
                break;
            }
            c[d] = ic - 2;
        }
        if (ok) {
            cell = c[0] + nx * (c[1] + ny * c[2]);
            cell_of[i] = cell;
        } else {
            // counted in cell 0 so that the permutation stays complete; the error stops the run
            set_err(flags, FM_EESCAPE, ext[i]);
            cell = 0;
            cell_of[i] = 0;
        }
    }
    // the molecule's place among its cell's molecules (the order within a cell is arbitrary
    // here: k_bin_sort orders each cell by x, ties by molecule number)
    const int pos = wave_aggregated_add(count, cell, cell >= 0);
    if (i < np) pos_out[i] = pos;
}

// start[0..n) = exclusive prefix of count[0..n), in chunks of SCAN_CH counts (FM_BLOCK threads,
// four consecutive counts each): k_scan_part sums each chunk; k_scan_fin adds up the sums of the
// chunks before its own, scans its chunk (by thread, wavefront and workgroup) and writes the
// prefixes.  The counts are zeroed as they are read (ready for the next binning) and so is
// cursor[0..ncur).
#define SCAN_CH (4 * FM_BLOCK)
__device__ inline int block_sum(int v, int* wsum) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    if (lane == 0) wsum[wid] = v;
    __syncthreads();
    int t = 0;
    for (int w = 0; w < FM_BLOCK / 64; ++w) t += wsum[w];
    return t;
}

__global__ __launch_bounds__(FM_BLOCK) void k_scan_part(const int* count, int* part, int n) {
    __shared__ int wsum[FM_BLOCK / 64];
    const int i0 = blockIdx.x * SCAN_CH + 4 * threadIdx.x;
    int v = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) v += i0 + k < n ? count[i0 + k] : 0;
    const int t = block_sum(v, wsum);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(FM_BLOCK) void k_scan_fin(int* count, int* start, const int* part, int n, int* cursor,
                                                       int ncur) {
    __shared__ int wsum[FM_BLOCK / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    // the sums of the chunks before this one
    int off = 0;
    for (int b = tid; b < (int)blockIdx.x; b += FM_BLOCK) off += part[b];
    off = block_sum(off, wsum);
    __syncthreads();  // (wsum is used again below)
    const int i0 = blockIdx.x * SCAN_CH + 4 * tid;
    int v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = i0 + k < n ? count[i0 + k] : 0;
        s += v[k];
    }
    int x = s;
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    int pre = off + x - s;
    for (int w = 0; w < wid; ++w) pre += wsum[w];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (i0 + k < n) {
            start[i0 + k] = pre;
            count[i0 + k] = 0;
            pre += v[k];
        }
    for (int i = blockIdx.x * FM_BLOCK + tid; i < ncur; i += gridDim.x * FM_BLOCK) cursor[i] = 0;
}

__global__ __launch_bounds__(FM_BLOCK) void k_bin_scatter(const int* cell_of, const int* start, const int* pos,
                                                          int* perm, int np) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) perm[start[cell_of[i]] + pos[i]] = i;
}

// order each cell by increasing x (ties: molecule number), which lets the list build window
// its candidate runs by x.  A wavefront takes two cells: when both hold at most 32 molecules,
// each half ranks its molecules (each lane counts the keys of its cell below its own, the other
// lanes' keys coming through shuffles); otherwise the whole wavefront takes each cell in turn,
// ranking up to 64 molecules with the keys through readlane, and cells above 64 molecules fall
// back to an insertion sort by lane 0.  Keys are distinct (ties in x are broken by the molecule
// number), so the ranks are a permutation.
__device__ inline bool xkey_less(double xa, int ea, double xb, int eb) { return xa < xb || (xa == xb && ea < eb); }

__device__ inline void bin_sort_cell(int b, int e, int* perm, const int* ext, const double* x) {
    const int lane = threadIdx.x & 63, n = e - b;
    if (n <= 1) return;
    if (n <= 64) {
        const int v = lane < n ? perm[b + lane] : -1;
        const double kx = lane < n ? x[v] : 0.0;
        const int ke = lane < n ? ext[v] : 0;
        const long long kb = __double_as_longlong(kx);
        const int klo = (int)kb, khi = (int)(kb >> 32);
        int rank = 0;
        for (int k = 0; k < n; ++k) {
            const int olo = __builtin_amdgcn_readlane(klo, k), ohi = __builtin_amdgcn_readlane(khi, k);
            const int oe = __builtin_amdgcn_readlane(ke, k);
            const double ox = __longlong_as_double(((long long)ohi << 32) | (long long)(unsigned)olo);
            rank += xkey_less(ox, oe, kx, ke) ? 1 : 0;
        }
        if (lane < n) perm[b + rank] = v;
    } else if (lane == 0) {
        for (int k = b + 1; k < e; ++k) {
            const int v = perm[k];
            const double kx = x[v];
            const int ke = ext[v];
            int m = k - 1;
            while (m >= b && xkey_less(kx, ke, x[perm[m]], ext[perm[m]])) {
                perm[m + 1] = perm[m];
                --m;
            }
            perm[m + 1] = v;
        }
    }
}

// cells 2w and 2w + 1 by wavefront w
__global__ __launch_bounds__(FM_BLOCK) void k_bin_sort(const int* start, int* perm, const int* ext, const double* x,
                                                       int ncell, int maxcell, int* flags) {
    const int lane = threadIdx.x & 63;
    const int c0 = 2 * (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    if (c0 >= ncell) return;  // whole wavefront
    const int hl = lane & 31, c = c0 + (lane >> 5);
    const int b = c < ncell ? start[c] : 0, n = c < ncell ? start[c + 1] - b : 0;
    if (n > maxcell && hl == 0) set_err(flags, FM_EOVERFLOW, c);
    if (__builtin_amdgcn_ballot_w64(n > 32) == 0ull) {
        const bool own = hl < n;
        const int v = own ? perm[b + hl] : -1;
        const double kx = own ? x[v] : 0.0;
        const int ke = own ? ext[v] : 0;
        const long long kb = __double_as_longlong(kx);
        const int klo = (int)kb, khi = (int)(kb >> 32);
        const int nmax = max(n, __shfl_xor(n, 32, 64));
        int rank = 0;
        for (int k = 0; k < nmax; ++k) {
            const int src = (lane & 32) + k;
            const int olo = __shfl(klo, src, 64), ohi = __shfl(khi, src, 64), oe = __shfl(ke, src, 64);
            const double ox = __longlong_as_double(((long long)ohi << 32) | (long long)(unsigned)olo);
            rank += (k < n && xkey_less(ox, oe, kx, ke)) ? 1 : 0;
        }
        if (own) perm[b + rank] = v;
        return;
    }
    bin_sort_cell(start[c0], start[c0 + 1], perm, ext, x);
    if (c0 + 1 < ncell) bin_sort_cell(start[c0 + 1], start[c0 + 2], perm, ext, x);
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
    int nf;  // state fields moved: 9, or 6 when the accelerations are about to be recomputed
};

__global__ __launch_bounds__(FM_BLOCK) void k_permute(PermArgs a) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.np) return;
    int o = a.perm[k];
    const size_t cap = a.cap;
#pragma unroll
    for (int f = 0; f < 9; ++f)
        if (f < a.nf) a.dst[k + f * cap] = a.src[o + f * cap];
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
        for (int f = 0; f < 8; ++f) a.mono_dst[k + f * cap] = a.mono_src[o + f * cap];
}

// ---- macroscopic sums, src/simulation_record.f90:1205-1314 ------------------------------------
// 7 sums per block: pe_lj, pe_poly, pe_lj+pe_poly, vir, mv2, vsum, msum ; + max(pe).  Each thread
// sums MACRO_MPT molecules a grid apart (in that order), then the wavefront and the workgroup.
#define NRED 8
#define MACRO_MPT 4
__global__ __launch_bounds__(FM_BLOCK) void k_macros_partial(St s, const double* pe_lj, const double* pe_poly,
                                                             const double* vir, const int* moltype,
                                                             const double* mass, int np, int lfv, double dt,
                                                             double* part) {
    __shared__ double red[NRED][FM_BLOCK / 64];
    double v[NRED] = {0, 0, 0, 0, 0, 0, 0, -1.0e300};
    const int stride = gridDim.x * blockDim.x;
#pragma unroll
    for (int k = 0; k < MACRO_MPT; ++k) {
        const int i = blockIdx.x * blockDim.x + threadIdx.x + k * stride;
        if (i >= np) break;
        double m = moltype ? mass[moltype[i]] : 1.0;
        double pl = pe_lj[i], pp = pe_poly ? pe_poly[i] : 0.0;
        v[0] = v[0] + pl;
        v[1] = v[1] + pp;
        v[2] = v[2] + (pl + pp);
        v[3] = v[3] + vir[i];
        double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
        if (lfv) {  // leap-frog: velocity shifted half a step, src/simulation_record.f90:1238-1247
            vx = vx + 0.5 * s.ax[i] * dt;
            vy = vy + 0.5 * s.ay[i] * dt;
            vz = vz + 0.5 * s.az[i] * dt;
        }
        v[4] = v[4] + ((m * (vx * vx) + m * (vy * vy)) + m * (vz * vz));
        v[5] = v[5] + ((m * vx + m * vy) + m * vz);
        v[6] = v[6] + m;
        v[7] = fmax(v[7], pl + pp);
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

// one workgroup: thread t sums partials t, t+256, ... in that order, then a fixed tree over the
// threads (deterministic; the partial count is the number of k_macros_partial workgroups)
__global__ __launch_bounds__(256) void k_macros_final(const double* part, int nb, double* out) {
    __shared__ double red[NRED][256];
    const int t = threadIdx.x;
    double x[NRED];
#pragma unroll
    for (int q = 0; q < NRED; ++q) x[q] = q == 7 ? -1.0e300 : 0.0;
    for (int b = t; b < nb; b += 256) {
#pragma unroll
        for (int q = 0; q < NRED; ++q) {
            const double o = part[(size_t)b * NRED + q];
            x[q] = q == 7 ? fmax(x[q], o) : x[q] + o;
        }
    }
#pragma unroll
    for (int q = 0; q < NRED; ++q) red[q][t] = x[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
#pragma unroll
            for (int q = 0; q < NRED; ++q) red[q][t] = q == 7 ? fmax(red[q][t], red[q][t + h]) : red[q][t] + red[q][t + h];
        __syncthreads();
    }
    if (t < NRED) out[t] = red[t][0];
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

// packed VV pass 1 (+ displacement check) over many domains: block b handles molecules
// table[b].y*FM_BLOCK .. of domain table[b].x (same operations as k_pass1 / k_pass1_disp).  A
// domain's pass does nothing when it has no step in this launch, or when a rebuild is pending
// for an earlier step (FLAG_HALT, or FLAG_REBUILD with FLAG_RB_STEP below this step: a check of
// this same launch stores this step, so its other workgroups still run); a check that fires
// records this step.
__global__ __launch_bounds__(FM_BLOCK) void k_pass1_multi(const P1Desc* descs, const int2* table) {
    const int2 t = table[blockIdx.x];
    const P1Desc& d = descs[t.x];
    if (!d.active) return;
    int* f = d.flags;
    if (f[FLAG_HALT] != 0 || (__atomic_load_n(&f[FLAG_REBUILD], __ATOMIC_RELAXED) != 0 &&
                              __atomic_load_n(&f[FLAG_RB_STEP], __ATOMIC_RELAXED) < d.step))
        return;
    const int i = t.y * FM_BLOCK + threadIdx.x;
    bool hit = false;
    if (i < d.np) {
        St s = state_of(d.st, d.cap);
        const double ax = s.ax[i], ay = s.ay[i], az = s.az[i];
        double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
        s.x[i] = s.x[i] + d.dt * vx + d.hdt2 * ax;
        s.y[i] = s.y[i] + d.dt * vy + d.hdt2 * ay;
        s.z[i] = s.z[i] + d.dt * vz + d.hdt2 * az;
        vx = vx + d.hdt * ax;
        vy = vy + d.hdt * ay;
        vz = vz + d.hdt * az;
        s.vx[i] = vx;
        s.vy[i] = vy;
        s.vz[i] = vz;
        if (d.check) {
            const size_t cap = d.cap;
            double d0 = d.fresh ? 0.0 : d.rd[i];
            double d1 = d.fresh ? 0.0 : d.rd[i + cap];
            double d2 = d.fresh ? 0.0 : d.rd[i + 2 * cap];
            d0 = d0 + vx * d.dt;
            d1 = d1 + vy * d.dt;
            d2 = d2 + vz * d.dt;
            d.rd[i] = d0;
            d.rd[i + cap] = d1;
            d.rd[i + 2 * cap] = d2;
            hit = fabs(d0) > d.thresh || fabs(d1) > d.thresh || fabs(d2) > d.thresh;
        }
    }
    unsigned long long m = __ballot(hit);
    if (m != 0ull && (int)(threadIdx.x & 63) == __ffsll((long long)m) - 1 &&
        __atomic_load_n(&f[FLAG_REBUILD], __ATOMIC_RELAXED) == 0) {
        atomicOr(&f[FLAG_REBUILD], 1);
        atomicExch(&f[FLAG_RB_STEP], d.step);
    }
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

int launch_pass1(fm_ctx* c, bool cond) {
    if (c->np == 0) return 0;
    double dt = c->p.delta_t;
    ktimer_begin(c, KT_PASS1);
    k_pass1<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(state_of(c->x(), c->cap), c->np, dt, 0.5 * (dt * dt),
                                                         0.5 * dt, cond ? c->d_flags : nullptr);
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

// zero the listed flag words (one launch for the resets before a queued stage)
__global__ void k_reset_flags(int* flags, FlagList l) {
    if ((int)threadIdx.x < l.n) reset_flag(flags, l.w[threadIdx.x]);
}

int launch_reset_flags(fm_ctx* c, const int* which, int n) {
    if (n < 1 || n > 8) return fm_set_error(FM_EINVAL, "reset of %d flag words", n);
    FlagList l;
    l.n = n;
    for (int k = 0; k < n; ++k) l.w[k] = which[k];
    k_reset_flags<<<1, 64, 0, c->stream>>>(c->d_flags, l);
    LAUNCH_CHECK();
    return 0;
}

// ---- sort_mols, src/linklist.f90:2744-2865 ------------------------------------------------------
// key[e] = curve position of molecule e's block, val[e] = e; cnt[key] counts the block
__global__ __launch_bounds__(FM_BLOCK) void k_sort_keys(St s, const int* slot_of, int np, const int* hcurve,
                                                         int nb0, int nb1, int nb2, double b0, double b1, double b2,
                                                         double h0, double h1, double h2, int* key, int* val,
                                                         int* cnt, int* flags) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= np) return;
    const int slot = slot_of[e];
    const int i = (int)ceil((s.x[slot] + h0) / b0), j = (int)ceil((s.y[slot] + h1) / b1),
              k = (int)ceil((s.z[slot] + h2) / b2);
    if (i < 1 || i > nb0 || j < 1 || j > nb1 || k < 1 || k > nb2) {
        set_err(flags, FM_EESCAPE, e);
        key[e] = 0;
        val[e] = e;
        return;
    }
    const int c = hcurve[(i - 1) + nb0 * ((j - 1) + nb1 * (k - 1))];
    key[e] = c;
    val[e] = e;
    atomicAdd(&cnt[c], 1);
}

// order[k] = old number of the molecule that becomes number k -> relabel the slots
__global__ __launch_bounds__(FM_BLOCK) void k_sort_newof(const int* order, int np, int* newof) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < np) newof[order[k]] = k;
}

__global__ __launch_bounds__(FM_BLOCK) void k_sort_relabel(St s, int* ext, int* slot_of, const int* newof, int np,
                                                            const int* tag, double* teth, int cap, double w0,
                                                            double w1, double w2) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= np) return;
    const int e = newof[ext[slot]];
    ext[slot] = e;
    slot_of[e] = slot;
    if (tag) {  // read_tag on every renumbered molecule
        const int tg = tag[slot];
        const double sg = tg == FM_TAG_TETH_SLIDE ? copysign(1.0, s.y[slot]) : tg == FM_TAG_FIXED_SLIDE ? 1.0 : 0.0;
        teth[slot + 3 * (size_t)cap] = sg != 0.0 ? w0 * sg : 0.0;
        teth[slot + 4 * (size_t)cap] = sg != 0.0 ? w1 * sg : 0.0;
        teth[slot + 5 * (size_t)cap] = sg != 0.0 ? w2 * sg : 0.0;
    }
}

// flags[flag] = max(flags[flag], v[0..n)) for an arbitrary array
__global__ __launch_bounds__(FM_BLOCK) void k_max_of(const int* v, int n, int* flag) {
    __shared__ int sm[FM_BLOCK / 64];
    int m = INT_MIN;
    for (int k = threadIdx.x; k < n; k += blockDim.x) m = max(m, v[k]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < FM_BLOCK / 64; ++w) m = max(m, sm[w]);
        m = max(m, sm[0]);
        *flag = m;
    }
}

// ---- stable LSD radix sort of (key, value) pairs, 8 bits a pass ---------------------------------
// Each workgroup takes a tile of FM_BLOCK consecutive pairs.  k_radix_hist counts the tile's digits
// into hist[digit * ntiles + tile]: digit-major, so the exclusive prefix of the whole array (the
// chunked scan above) is, for each tile and digit, the first output position of that tile's pairs
// with that digit.  k_radix_scatter places each pair after the pairs of its digit that come before
// it in its tile: the lanes of its wavefront below it whose digit matches (eight ballots, one per
// digit bit) and the matching counts of the wavefronts before its own.  Tiles, wavefronts and lanes
// keep the input order, so every pass is stable.
static_assert(FM_BLOCK == 256, "one 8-bit digit bucket per thread of the tile");

__global__ __launch_bounds__(FM_BLOCK) void k_radix_hist(const int* key, int n, int shift, int* hist, int ntiles) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * FM_BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&h[(key[i] >> shift) & 255], 1);
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(FM_BLOCK) void k_radix_scatter(const int* kin, const int* vin, int* kout, int* vout,
                                                             int n, int shift, const int* start, int ntiles) {
    __shared__ int wc[FM_BLOCK / 64][256];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (int w = 0; w < FM_BLOCK / 64; ++w) wc[w][tid] = 0;
    __syncthreads();
    const int i = blockIdx.x * FM_BLOCK + tid;
    const bool active = i < n;
    const int k = active ? kin[i] : 0;
    const int d = (k >> shift) & 255;
    uint64_t m = __ballot(active);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const uint64_t on = __ballot((d >> b) & 1);
        m &= ((d >> b) & 1) ? on : ~on;
    }
    const int rank = __popcll(m & ((1ull << lane) - 1ull));
    if (active && rank == 0) wc[wid][d] = __popcll(m);
    __syncthreads();
    if (!active) return;
    int pos = start[(size_t)d * ntiles + blockIdx.x] + rank;
    for (int w = 0; w < wid; ++w) pos += wc[w][d];
    kout[pos] = k;
    vout[pos] = vin[i];
}

size_t radix_sort_bytes(int n) {
    const size_t nt = ((size_t)std::max(n, 1) + FM_BLOCK - 1) / FM_BLOCK, nh = 256 * nt;
    return (2 * nh + nh / SCAN_CH + 2) * sizeof(int);
}

// Sorts the pairs (k0, v0)[0..n) by key bits [0, end_bit), stably, with (k1, v1) as the other
// buffer of each pass; *vals is the buffer that ends with the sorted values.
int radix_sort_pairs(fm_ctx* c, int* k0, int* v0, int* k1, int* v1, int n, int end_bit, int** vals) {
    const int nt = (n + FM_BLOCK - 1) / FM_BLOCK, nh = 256 * nt, nscan = (nh + SCAN_CH - 1) / SCAN_CH;
    if ((size_t)(2 * (size_t)nh + nh / SCAN_CH + 2) * sizeof(int) > c->sort_tmp_bytes)
        return fm_set_error(FM_EOVERFLOW, "radix sort of %d pairs exceeds its temporary", n);
    int* hist = (int*)c->sort_tmp;
    int* start = hist + nh;
    int* part = start + nh;
    for (int shift = 0; shift < end_bit; shift += 8) {
        k_radix_hist<<<nt, FM_BLOCK, 0, c->stream>>>(k0, n, shift, hist, nt);
        k_scan_part<<<nscan, FM_BLOCK, 0, c->stream>>>(hist, part, nh);
        k_scan_fin<<<nscan, FM_BLOCK, 0, c->stream>>>(hist, start, part, nh, nullptr, 0);
        k_radix_scatter<<<nt, FM_BLOCK, 0, c->stream>>>(k0, v0, k1, v1, n, shift, start, nt);
        LAUNCH_CHECK();
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    *vals = v0;
    return 0;
}

// One sort_mols renumbering (the caller handles the call counter).  *overflow_out = 1 when a
// block exceeds the estimate and the round is skipped, as the reference does.
int launch_sort_mols(fm_ctx* c, int* overflow_out) {
    *overflow_out = 0;
    if (c->np == 0) return 0;
    const int blocks = c->sort_nb[0] * c->sort_nb[1] * c->sort_nb[2];
    int ave = c->np / blocks;
    ave = (int)ceil(ave * c->sort_sf1 + c->sort_sf2);
    if (int rc = launch_fill_i32(c, c->sort_cnt, 0, (size_t)blocks + 1)) return rc;
    St s = state_of(c->x(), c->cap);
    k_sort_keys<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(
        s, c->slot_of, c->np, c->hcurve, c->sort_nb[0], c->sort_nb[1], c->sort_nb[2], c->sort_bsl[0], c->sort_bsl[1],
        c->sort_bsl[2], c->half[0], c->half[1], c->half[2], c->sort_key, c->sort_val, c->sort_cnt, c->d_flags);
    LAUNCH_CHECK();
    k_max_of<<<1, FM_BLOCK, 0, c->stream>>>(c->sort_cnt, blocks + 1, c->d_flags + FLAG_SORT_MAX);
    LAUNCH_CHECK();
    if (int rc = ctx_sync_flags(c)) return rc;
    if (c->h_flags[FLAG_SORT_MAX] > ave) {  // "Miss this round of sorting"
        c->sort_sf1 = c->sort_sf1 + 0.1;
        c->sort_sf2 = c->sort_sf2 + 5;
        *overflow_out = 1;
        return 0;
    }
    int end_bit = 1;
    while ((1 << end_bit) <= blocks) ++end_bit;
    int* order = nullptr;
    if (int rc = radix_sort_pairs(c, c->sort_key, c->sort_val, c->sort_key2, c->sort_val2, c->np, end_bit, &order))
        return rc;
    // the keys are spent: sort_key takes the new numbers (order is a value buffer)
    k_sort_newof<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(order, c->np, c->sort_key);
    LAUNCH_CHECK();
    const bool tm = c->p.ensemble == FM_ENSEMBLE_TAG_MOVE && c->has_tags;
    k_sort_relabel<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(
        s, c->ext[c->cur], c->slot_of, c->sort_key, c->np, tm ? c->tag[c->cur] : nullptr,
        tm ? c->teth[c->cur] : nullptr, c->cap, c->p.wallslidev[0], c->p.wallslidev[1], c->p.wallslidev[2]);
    LAUNCH_CHECK();
    return 0;
}

// simulation_move_particles_lfv; check_disp: fused displacement update of rebuild criterion 2
int launch_move_lfv(fm_ctx* c, bool check_disp, bool cond) {
    if (c->np == 0) return 0;
    const bool tm = c->p.ensemble == FM_ENSEMBLE_TAG_MOVE;
    ktimer_begin(c, KT_PASS1);
    k_move_lfv<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(
        state_of(c->x(), c->cap), tm ? c->tag[c->cur] : nullptr, tm ? c->teth[c->cur] : nullptr, c->ext[c->cur],
        c->np, c->cap, c->p.delta_t, c->p.teth_k2, c->p.teth_k4, c->p.teth_k6, c->rdisp, 0.5 * c->dmin,
        check_disp ? 1 : 0, c->rdisp_valid ? 0 : 1, c->d_flags, cond ? 1 : 0);
    ktimer_end(c, KT_PASS1);
    LAUNCH_CHECK();
    if (check_disp) c->rdisp_valid = true;
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

int launch_sendmols(fm_ctx* c, const int* reset, int nreset) {
    if (nreset > 8) return fm_set_error(FM_EINVAL, "reset of %d flag words", nreset);
    FlagList l;
    l.n = nreset;
    for (int k = 0; k < nreset; ++k) l.w[k] = reset[k];
    if (c->np == 0) return nreset ? launch_reset_flags(c, reset, nreset) : 0;
    ktimer_begin(c, KT_SENDMOLS);
    St s = state_of(c->x(), c->cap);
    const int nb = nblocks(c->np);
    // segments in the flag buffer, counts in out_count; lst and tmp of the passes reuse the cell
    // permutation buffer and the flag buffer (after the gather)
    k_out_select<<<nb, FM_BLOCK, 0, c->stream>>>(s, c->slot_of, c->np, c->half[0], c->half[1], c->half[2],
                                                   c->out_flag, c->out_count, c->d_flags, l);
    LAUNCH_CHECK();
    const bool tm = c->p.ensemble == FM_ENSEMBLE_TAG_MOVE && c->has_tags;
    k_sendmols_passes<<<1, 1024, 0, c->stream>>>(
        s, tm ? c->teth[c->cur] : nullptr, tm ? c->tag[c->cur] : nullptr, c->ext[c->cur], c->slot_of,
        c->out_flag, c->out_count, nb, c->out_vals, c->out_slot, c->perm, c->out_flag, c->cap, c->p.domain[0],
        c->p.domain[1], c->p.domain[2], tm ? 1 : 0, c->p.wallslidev[0], c->p.wallslidev[1], c->p.wallslidev[2]);
    ktimer_end(c, KT_SENDMOLS);
    LAUNCH_CHECK();
    return 0;
}

int launch_bin(fm_ctx* c, bool keep_acc) {
    c->btab_valid = false;  // new cells: the staging tables wait for the next list build
    if (c->np == 0) return 0;
    ktimer_begin(c, KT_BIN);
    St s = state_of(c->x(), c->cap);
    const int nxt = 1 - c->cur;
    int* cell_tmp = c->out_flag;  // cell index in old slot order
    k_bin_count<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(s, c->ext[c->cur], c->np, c->half[0], c->half[1],
                                                             c->half[2], c->cs[0], c->cs[1], c->cs[2], c->nc[0],
                                                             c->nc[1], c->nc[2], cell_tmp, c->cell_count,
                                                             c->bin_pos, c->d_flags);
    LAUNCH_CHECK();
    const int nscan = (c->ncell + 1 + SCAN_CH - 1) / SCAN_CH;
    k_scan_part<<<nscan, FM_BLOCK, 0, c->stream>>>(c->cell_count, c->scan_part, c->ncell + 1);
    LAUNCH_CHECK();
    k_scan_fin<<<nscan, FM_BLOCK, 0, c->stream>>>(c->cell_count, c->cell_start, c->scan_part, c->ncell + 1,
                                                  nullptr, 0);
    LAUNCH_CHECK();
    k_bin_scatter<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(cell_tmp, c->cell_start, c->bin_pos, c->perm,
                                                               c->np);
    LAUNCH_CHECK();
    int maxcell = c->p.max_cell_np > 0 ? c->p.max_cell_np : (1 << 26);
    const int nsort = nblocks((size_t)((c->ncell + 1) / 2) * 64);  // two cells per wavefront
    k_bin_sort<<<nsort, FM_BLOCK, 0, c->stream>>>(c->cell_start, c->perm, c->ext[c->cur], s.x, c->ncell, maxcell,
                                                  c->d_flags);
    LAUNCH_CHECK();
    PermArgs a;
    a.src = c->st[c->scur];
    a.dst = c->st[1 - c->scur];
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
    a.nf = keep_acc ? 9 : 6;
    k_permute<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(a);
    ktimer_end(c, KT_BIN);
    LAUNCH_CHECK();
    c->cur = nxt;
    c->scur = 1 - c->scur;
    return 0;
}

int launch_macros(fm_ctx* c, int buf, double* out) {
    int nb = nblocks((c->np + MACRO_MPT - 1) / MACRO_MPT);
    if (nb == 0) nb = 1;
    if (buf < 0) buf = c->scur;
    ktimer_begin(c, KT_MACROS);
    const bool mie = c->p.mie_potential == 1;
    const bool poly = c->p.potential_flag == 1 && c->has_mono;
    k_macros_partial<<<nb, FM_BLOCK, 0, c->stream>>>(state_of(c->st[buf], c->cap), c->pe_lj, poly ? c->pe_poly : nullptr,
                                                      c->vir, mie ? c->moltype[c->cur] : nullptr,
                                                      c->d_mie + 7 * FM_MAX_TYPES * FM_MAX_TYPES,
                                                      c->np, c->p.integration_algorithm == FM_INTEGRATION_LFV ? 1 : 0,
                                                      c->p.delta_t, c->red);
    k_macros_final<<<1, 256, 0, c->stream>>>(c->red, nb, out ? out : c->d_macros);
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

int launch_pass1_check_disp(fm_ctx* c, bool cond, int step) {
    if (c->np == 0) return 0;
    const double dt = c->p.delta_t;
    ktimer_begin(c, KT_PASS1);
    k_pass1_disp<<<nblocks(c->np), FM_BLOCK, 0, c->stream>>>(state_of(c->x(), c->cap), c->rdisp, c->np, c->cap, dt,
                                                              0.5 * (dt * dt), 0.5 * dt, 0.5 * c->dmin,
                                                              c->rdisp_valid ? 0 : 1, c->d_flags, cond ? 1 : 0, step);
    ktimer_end(c, KT_PASS1);
    LAUNCH_CHECK();
    c->rdisp_valid = true;
    return 0;
}

int launch_pass1_multi(hipStream_t st, const P1Desc* d_descs, const int2* d_table, int nblk) {
    if (nblk == 0) return 0;
    k_pass1_multi<<<nblk, FM_BLOCK, 0, st>>>(d_descs, d_table);
    LAUNCH_CHECK();
    return 0;
}

// =============================================================================================
// Record-step sums of caller-held device arrays without any allocation (fm_macros_sums_noalloc).
// The partial sums live in a file-scope device array; each device's copy is found once with
// hipGetSymbolAddress and its address is what the kernels receive.
#define FM_NOALLOC_MAX_NP (1 << 24)
#define FM_NOALLOC_BLOCKS ((FM_NOALLOC_MAX_NP + FM_BLOCK - 1) / FM_BLOCK)
static __device__ __align__(256) double scratch[FM_NOALLOC_BLOCKS * NRED];

namespace {
// the sums of k_macros_partial for (3,np) arrays in molecule order; a (leap-frog) shifts v by
// half a step, m = per-molecule masses (null: 1)
__global__ __launch_bounds__(FM_BLOCK) void k_macros_partial_aos(const double* v, const double* a, double dt,
                                                                 const double* pe_lj, const double* pe_poly,
                                                                 const double* vir, const double* mass, int np,
                                                                 double* part) {
    __shared__ double red[NRED][FM_BLOCK / 64];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double s[NRED] = {0, 0, 0, 0, 0, 0, 0, -1.0e300};
    if (i < np) {
        const double m = mass ? mass[i] : 1.0;
        const double pl = pe_lj[i], pp = pe_poly ? pe_poly[i] : 0.0;
        s[0] = pl;
        s[1] = pp;
        s[2] = pl + pp;
        s[3] = vir[i];
        double vx = v[3 * (size_t)i], vy = v[3 * (size_t)i + 1], vz = v[3 * (size_t)i + 2];
        if (a) {
            vx = vx + 0.5 * a[3 * (size_t)i] * dt;
            vy = vy + 0.5 * a[3 * (size_t)i + 1] * dt;
            vz = vz + 0.5 * a[3 * (size_t)i + 2] * dt;
        }
        s[4] = (m * (vx * vx) + m * (vy * vy)) + m * (vz * vz);
        s[5] = (m * vx + m * vy) + m * vz;
        s[6] = m;
        s[7] = pl + pp;
    }
    for (int q = 0; q < NRED; ++q) {
        double x = s[q];
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_down(x, off, 64);
            x = q == 7 ? fmax(x, o) : x + o;
        }
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x < NRED) {
        const int q = threadIdx.x;
        double x = red[q][0];
        for (int w = 1; w < FM_BLOCK / 64; ++w) x = q == 7 ? fmax(x, red[q][w]) : x + red[q][w];
        part[(size_t)blockIdx.x * NRED + q] = x;
    }
}
}  // namespace

// device address of `scratch` per device, looked up at each device's first call
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[64] = {nullptr};

extern "C" int fm_macros_sums_noalloc(int32_t np, const double* v, const double* a, double delta_t,
                                      const double* potenergymol_LJ, const double* potenergymol_POLY,
                                      const double* virialmol, const double* mass, double* sums, void* stream) {
    if (np < 1 || np > FM_NOALLOC_MAX_NP)
        return fm_set_error(FM_EINVAL, "fm_macros_sums_noalloc: np %d outside 1..%d (the size of its static scratch)",
                            np, FM_NOALLOC_MAX_NP);
    if (!v || !potenergymol_LJ || !virialmol || !sums) return fm_set_error(FM_EINVAL, "fm_macros_sums_noalloc: null array");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return fm_set_error(FM_EHIP, "fm_macros_sums_noalloc: no current device");
    double* ptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return fm_set_error(FM_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            g_scratch_ptr[dev] = (double*)p;
        }
        ptr = g_scratch_ptr[dev];
    }
    const hipStream_t st = (hipStream_t)stream;
    const int nb = (np + FM_BLOCK - 1) / FM_BLOCK;
    k_macros_partial_aos<<<nb, FM_BLOCK, 0, st>>>(v, a, delta_t, potenergymol_LJ, potenergymol_POLY, virialmol, mass,
                                                 np, ptr);
    k_macros_final<<<1, 256, 0, st>>>(ptr, nb, sums);
    LAUNCH_CHECK();
    return 0;
}
