"""Host-side mirror of the reference's module interface for the force / neighbour / integration path.

Each method forwards to one C-ABI entry point of libflowmol_hip and keeps the reference routine's
name and argument meaning (paths relative to the reference's root):

  assign_to_cell               src/linklist.f90:2645-2676
  assign_to_neighbourlist      src/linklist.f90:2710-2737 (FORCE_LIST 3)
  linklist_deallocateall       src/linklist.f90:2872-2963
  sendmols                     src/messenger.MPI.f90:1383-1421 (one periodic rank)
  sort_mols                    src/linklist.f90:2744-2865 (SORT_FLAG 0, 1, 2)
  messenger_updateborders      src/messenger.MPI.f90:594-628
  simulation_checkrebuild      src/simulation_checkrebuild.f90:17-174
  simulation_compute_forces    src/simulation_compute_forces.f90:182-275
  simulation_move_particles_vv src/simulation_move_particles_vv.f90:44-193
  evaluate_macroscopic_properties src/simulation_record.f90:1205-1314

Errors come back as FlowmolHipError carrying the reference's abort message (error_abort).
"""
import ctypes as C

import numpy as np

from ._abi import FM_ENSEMBLE_TAG_MOVE, FM_MAX_FUNCY, FmMacros, dptr, iptr
from ._lib import check, load


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class MDContext:
    """One periodic domain on the current HIP device."""

    def __init__(self, params, r, v, a=None, capacity=0):
        L = load()
        r = _f64(r)
        v = _f64(v)
        np_ = r.shape[0]
        self.params = params
        h = C.c_void_p()
        check(L.fm_init(C.byref(params), np_, int(capacity), C.byref(h)))
        self._h = h
        self._np = np_
        check(L.fm_upload_state(self._h, np_, dptr(r), dptr(v), dptr(None if a is None else _f64(a))))

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            load().fm_finalize(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def np(self):
        return self._np

    def geometry(self):
        nc = np.zeros(3, np.int32)
        cs = np.zeros(3)
        rn2, rc2, ps = C.c_double(), C.c_double(), C.c_double()
        check(load().fm_get_geometry(self._h, iptr(nc), dptr(cs), C.byref(rn2), C.byref(rc2), C.byref(ps)))
        return dict(ncells=nc, cellsidelength=cs, rneighbr2=rn2.value, rcutoff2=rc2.value, potshift=ps.value)

    def autotune_brick(self, iter_=0, max_candidates=4):
        """In-place retuning of a live context is switched off: its first GPU run ended in a host
        segmentation fault whose cause was not found.  tune_brick() tunes on fresh contexts."""
        raise RuntimeError("in-place brick retuning is switched off; use flowmol_amd.md.tune_brick")

    def capacity(self):
        """(neighbour-list entries per molecule, molecules per staged brick region)."""
        n, m = C.c_int32(), C.c_int32()
        check(load().fm_get_capacity(self._h, C.byref(n), C.byref(m)))
        return n.value, m.value

    # ---- state ---------------------------------------------------------------------------------
    def upload_state(self, r, v, a=None):
        r, v = _f64(r), _f64(v)
        self._np = r.shape[0]
        check(load().fm_upload_state(self._h, self._np, dptr(r), dptr(v), dptr(None if a is None else _f64(a))))

    def set_moltype(self, moltype):
        check(load().fm_upload_moltype(self._h, iptr(np.ascontiguousarray(moltype, np.int32))))

    def set_tags(self, tag, rtether=None):
        rt = None if rtether is None else _f64(rtether)
        check(load().fm_upload_tags(self._h, iptr(np.ascontiguousarray(tag, np.int32)), dptr(rt)))

    def set_monomers(self, chainID, subchainID, funcy, bin_bflag, glob_no=None):
        c = lambda x: None if x is None else np.ascontiguousarray(x, np.int32)
        check(load().fm_upload_monomers(self._h, iptr(c(chainID)), iptr(c(subchainID)), iptr(c(funcy)),
                                        iptr(c(bin_bflag)), iptr(c(glob_no))))

    def state(self):
        n = self._np
        r, v, a = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        check(load().fm_download_state(self._h, dptr(r), dptr(v), dptr(a)))
        return r, v, a

    def rtether(self):
        out = np.zeros((self._np, 3))
        check(load().fm_download_tags(self._h, dptr(out)))
        return out

    def labels(self):
        """(tag, moltype, monomer) per molecule in the current numbering; entries the context does
        not hold are None.  monomer = (chainID, subchainID, funcy, bin_bflag (np, 4), glob_no)."""
        n = self._np
        p = self.params
        tag = np.zeros(n, np.int32) if p.ensemble == FM_ENSEMBLE_TAG_MOVE else None
        mt = np.zeros(n, np.int32) if p.mie_potential == 1 else None
        mono = None
        if p.potential_flag == 1:
            mono = (np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros((n, 4), np.int32),
                    np.zeros(n, np.int32))
        m = mono or (None,) * 5
        check(load().fm_download_labels(self._h, iptr(tag), iptr(mt), iptr(m[0]), iptr(m[1]), iptr(m[2]), iptr(m[3]),
                                        iptr(m[4])))
        return tag, mt, mono

    def potenergy_parts(self):
        lj, poly = np.zeros(self._np), np.zeros(self._np)
        check(load().fm_download_potenergy_parts(self._h, dptr(lj), dptr(poly)))
        return lj, poly

    def molecular(self):
        pe, vir = np.zeros(self._np), np.zeros(self._np)
        check(load().fm_download_molecular(self._h, dptr(pe), dptr(vir)))
        return pe, vir

    def rfmol(self):
        out = np.zeros((self._np, 3, 3))
        check(load().fm_download_rfmol(self._h, dptr(out)))
        return out

    @property
    def simtime(self):
        return load().fm_simtime(self._h)

    @simtime.setter
    def simtime(self, t):
        check(load().fm_set_simtime(self._h, float(t)))

    # ---- the reference routines ------------------------------------------------------------------
    def simulation_move_particles_vv(self, pass_num):
        check(load().fm_move_particles_vv(self._h, int(pass_num)))

    def simulation_move_particles_lfv(self):
        check(load().fm_move_particles_lfv(self._h))

    def simulation_checkrebuild(self, iter_):
        rb = C.c_int32()
        check(load().fm_checkrebuild(self._h, int(iter_), C.byref(rb)))
        return rb.value

    def linklist_deallocateall(self):
        check(load().fm_linklist_deallocateall(self._h))

    def sendmols(self):
        check(load().fm_sendmols(self._h))

    def sort_mols(self):
        check(load().fm_sort_mols(self._h))

    def assign_to_cell(self):
        check(load().fm_assign_to_cell(self._h))

    def messenger_updateborders(self, rebuild):
        check(load().fm_updateborders(self._h, int(rebuild)))

    def assign_to_neighbourlist(self):
        check(load().fm_assign_to_neighbourlist(self._h))

    def simulation_compute_forces(self, iter_):
        check(load().fm_compute_forces(self._h, int(iter_)))

    def evaluate_macroscopic_properties(self):
        m = FmMacros()
        check(load().fm_evaluate_macroscopic_properties(self._h, C.byref(m)))
        return m.as_dict()

    def mop(self):
        n = self.params.nplanes
        out = np.zeros((n, 3))
        check(load().fm_get_mop(self._h, dptr(out), n))
        return out

    def reset_mop(self):
        check(load().fm_reset_mop(self._h))

    def cv_bins(self):
        """(nbins, nhb, binsize) of the control-volume tally (vflux_outflag 4)."""
        nb = np.zeros(3, np.int32)
        nh = np.zeros(3, np.int32)
        bs = np.zeros(3)
        check(load().fm_get_cv_bins(self._h, iptr(nb), iptr(nh), dptr(bs)))
        return nb, nh, bs

    def cv_stress(self):
        """Pxyface indexed [bx, by, bz, component, face] (0-based, faces 0..5, halo bins included)."""
        nb, nh, _ = self.cv_bins()
        nbo = nb + 2 * nh
        out = np.zeros(int(nbo.prod()) * 18)
        check(load().fm_get_cv_stress(self._h, dptr(out), out.size))
        return out.reshape((6, 3, nbo[2], nbo[1], nbo[0])).transpose(4, 3, 2, 1, 0)

    def reset_cv_stress(self):
        check(load().fm_reset_cv_stress(self._h))

    # ---- call sequences of src/md.f90 -----------------------------------------------------------
    def setup(self, iter0=0):
        """setup_MD list build and the initial force call (src/md.f90:60-64,
        src/setup_initial_record.f90:88)."""
        self.assign_to_cell()
        self.messenger_updateborders(1)
        self.assign_to_neighbourlist()
        self.simulation_compute_forces(iter0)

    def rebuild(self):
        """The rebuild branch of md_advance_vv (src/md.f90:150-157)."""
        self.linklist_deallocateall()
        self.sendmols()
        self.sort_mols()
        self.assign_to_cell()
        self.messenger_updateborders(1)
        self.assign_to_neighbourlist()

    def md_advance_vv(self, iter_):
        """One velocity Verlet step, routine by routine (src/md.f90:143-163)."""
        self.simulation_move_particles_vv(1)
        self.messenger_updateborders(0)
        rb = self.simulation_checkrebuild(iter_)
        if rb:
            self.rebuild()
        self.simulation_compute_forces(iter_)
        self.simulation_move_particles_vv(2)
        return rb

    def run_vv(self, iter0, nsteps, n_macros=0):
        """Fused time loop (same kernels and arithmetic as md_advance_vv per step)."""
        ms = (FmMacros * max(n_macros, 1))()
        nrb = C.c_int32()
        check(load().fm_run_vv(self._h, int(iter0), int(nsteps), ms, int(n_macros), C.byref(nrb)))
        return [ms[k].as_dict() for k in range(n_macros)], nrb.value

    def md_advance_lfv(self, iter_, macros=False):
        """One leap-frog step, routine by routine (src/md.f90:109-139): forces, record, move,
        rebuild check.  Returns (rebuilt, macros of the record step or None)."""
        self.simulation_compute_forces(iter_)
        m = None
        if macros and iter_ % self.params.tplot == 0:
            m = self.evaluate_macroscopic_properties()
        self.simulation_move_particles_lfv()
        self.messenger_updateborders(0)
        rb = self.simulation_checkrebuild(iter_)
        if rb:
            self.rebuild()
        return rb, m

    def run_lfv(self, iter0, nsteps, n_macros=0):
        """Fused leap-frog loop (same kernels and arithmetic as md_advance_lfv per step)."""
        ms = (FmMacros * max(n_macros, 1))()
        nrb = C.c_int32()
        check(load().fm_run_lfv(self._h, int(iter0), int(nsteps), ms, int(n_macros), C.byref(nrb)))
        return [ms[k].as_dict() for k in range(n_macros)], nrb.value

    # ---- inspection ----------------------------------------------------------------------------
    def cells(self):
        out = np.zeros(self._np, np.int32)
        check(load().fm_download_cells(self._h, iptr(out)))
        return out

    def neighbourlist(self):
        n = C.c_int64()
        check(load().fm_download_neighbourlist(self._h, None, None, None, 0, C.byref(n)))
        i, j, s = (np.zeros(n.value, np.int32) for _ in range(3))
        check(load().fm_download_neighbourlist(self._h, iptr(i), iptr(j), iptr(s), n.value, C.byref(n)))
        return i, j, s

    def bonds(self):
        b = np.zeros((self._np, FM_MAX_FUNCY), np.int32)
        s = np.zeros((self._np, FM_MAX_FUNCY), np.int32)
        check(load().fm_download_bonds(self._h, iptr(b), iptr(s)))
        return b, s

    def synchronize(self):
        check(load().fm_synchronize(self._h))

    def kernel_timing(self, enable=True):
        check(load().fm_kernel_timing(self._h, 1 if enable else 0))

    def get_kernel_timing(self):
        buf = C.create_string_buffer(1024)
        avg = np.zeros(16)
        cnt = np.zeros(16, np.int64)
        n = C.c_int32()
        check(load().fm_get_kernel_timing(self._h, buf, 1024, dptr(avg), cnt.ctypes.data_as(C.POINTER(C.c_int64)),
                                          16, C.byref(n)))
        names = buf.value.decode().split("\n")[: n.value]
        return {names[k]: (avg[k], int(cnt[k])) for k in range(n.value)}


# ---- brick-shape autotuner ---------------------------------------------------------------------
# No counterpart in the reference, whose cells are not grouped into bricks.
FB_BLOCK = 512  # threads of one brick workgroup (FB_BLOCK of the brick kernels)


def brick_candidates(ncells, nmol, max_candidates=4):
    """Brick shapes (cells along x, y, z) whose expected molecule count fits one pass of a
    workgroup (15/16 of FB_BLOCK threads), least staged halo per cell first.  The first one is the
    library's default shape for the same cells and occupancy."""
    nc = [int(x) for x in ncells]
    occ = max(1.0, nmol / float(nc[0] * nc[1] * nc[2]))
    target = FB_BLOCK * 15.0 / 16.0
    out = []
    for bx in range(1, min(nc[0], 12) + 1):
        for by in range(1, min(nc[1], 12) + 1):
            for bz in range(1, min(nc[2], 12) + 1):
                P = bx * by * bz
                if P > 1 and P * occ > target:
                    continue
                out.append((P / ((bx + 2) * (by + 2) * (bz + 2)), (bx, by, bz)))
    # a stable sort: equal halo efficiency keeps the enumeration order, as the library's choice
    return [b for _, b in sorted(out, key=lambda t: -t[0])][:max(1, int(max_candidates))]


def tune_brick(params, r, v, a=None, steps=20, warmup=5, max_candidates=4, prepare=None):
    """Time `steps` fused velocity Verlet steps (after `warmup` more, list rebuilds included) of the
    same state on a fresh context per candidate brick shape; returns (fastest shape, {shape:
    seconds per step}).  `params` is not changed: set params.brick to the shape to use it.
    prepare(ctx), when given, loads what the state needs beyond r, v, a (molecule types, monomer
    data, tags) into each fresh context before its setup."""
    import time

    probe = MDContext(params, r, v, a)
    nc = probe.geometry()["ncells"]
    probe.close()
    times = {}
    for b in brick_candidates(nc, np.asarray(r).reshape(-1, 3).shape[0], max_candidates):
        p = type(params).from_buffer_copy(params)
        p.brick[0], p.brick[1], p.brick[2] = b
        g = MDContext(p, r, v, a)
        try:
            if prepare is not None:
                prepare(g)
            g.setup(0)
            g.run_vv(1, warmup)
            g.synchronize()
            t0 = time.perf_counter()
            g.run_vv(1 + warmup, steps)
            g.synchronize()
            times[b] = (time.perf_counter() - t0) / steps
        finally:
            g.close()
    best = min(times, key=times.get)
    return best, times
