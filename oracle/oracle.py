"""ctypes wrapper of the CPU restatement (libflowmol_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os

import numpy as np

from flowmol_amd._abi import FmMacros, FmParams, dptr, iptr, FM_MAX_FUNCY

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libflowmol_oracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        vp, i32, i64, dp, ip = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_int32)
        sig = {
            "orc_create": (vp, [C.POINTER(FmParams), i32, i32]),
            "orc_destroy": (None, [vp]),
            "orc_last_error": (C.c_char_p, []),
            "orc_geometry": (None, [vp, ip, dp, dp, dp, dp]),
            "orc_set_state": (C.c_int, [vp, i32, dp, dp, dp]),
            "orc_set_moltype": (C.c_int, [vp, ip]),
            "orc_set_tags": (C.c_int, [vp, ip, dp]),
            "orc_set_monomers": (C.c_int, [vp, ip, ip, ip, ip]),
            "orc_get_state": (C.c_int, [vp, dp, dp, dp]),
            "orc_get_tags": (C.c_int, [vp, dp]),
            "orc_get_molecular": (C.c_int, [vp, dp, dp]),
            "orc_get_rfmol": (C.c_int, [vp, dp]),
            "orc_np": (i32, [vp]),
            "orc_halo_np": (i32, [vp]),
            "orc_linklist_deallocateall": (C.c_int, [vp]),
            "orc_sendmols": (C.c_int, [vp]),
            "orc_sort_mols": (C.c_int, [vp]),
            "orc_get_hcurve": (C.c_int, [vp, ip]),
            "orc_assign_to_cell": (C.c_int, [vp]),
            "orc_updateborders": (C.c_int, [vp, i32]),
            "orc_assign_to_neighbourlist": (C.c_int, [vp]),
            "orc_compute_forces": (C.c_int, [vp, i32]),
            "orc_compute_forces_allpairs": (C.c_int, [vp, i32]),
            "orc_move_particles_vv": (C.c_int, [vp, i32]),
            "orc_checkrebuild": (C.c_int, [vp, i32, ip]),
            "orc_macros": (C.c_int, [vp, C.POINTER(FmMacros)]),
            "orc_get_mop": (C.c_int, [vp, dp]),
            "orc_reset_mop": (C.c_int, [vp]),
            "orc_get_planes": (C.c_int, [vp, dp]),
            "orc_get_cv_bins": (C.c_int, [vp, ip, ip, dp]),
            "orc_get_cv_stress": (C.c_int, [vp, dp]),
            "orc_reset_cv_stress": (C.c_int, [vp]),
            "orc_run_vv": (C.c_int, [vp, i32, i32, C.POINTER(FmMacros), i32, ip]),
            "orc_move_particles_lfv": (C.c_int, [vp]),
            "orc_run_lfv": (C.c_int, [vp, i32, i32, C.POINTER(FmMacros), i32, ip]),
            "orc_get_cells": (C.c_int, [vp, ip]),
            "orc_neighbour_count": (i64, [vp]),
            "orc_get_neighbourlist": (C.c_int, [vp, ip, ip, ip]),
            "orc_get_bonds": (C.c_int, [vp, ip, ip]),
            "orc_get_force_scale": (C.c_int, [vp, dp]),
            "orc_bruteforce_pairs": (i64, [vp, ip, ip, ip, i64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


class Oracle:
    """One periodic domain on the CPU, driven routine by routine like the reference's md.f90."""

    def __init__(self, params, r, v, a=None, capacity=None):
        L = lib()
        r = np.ascontiguousarray(r, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        np_ = r.shape[0]
        self.params = params
        if capacity is None:
            capacity = 3 * np_ + 20000
        self._h = L.orc_create(C.byref(params), np_, capacity)
        if not self._h:
            raise OracleError(L.orc_last_error().decode())
        _check(L.orc_set_state(self._h, np_, dptr(r), dptr(v),
                               dptr(None if a is None else np.ascontiguousarray(a, dtype=np.float64))))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    @property
    def np(self):
        return lib().orc_np(self._h)

    def geometry(self):
        nc = np.zeros(3, np.int32)
        cs = np.zeros(3)
        rn2, rc2, ps = C.c_double(), C.c_double(), C.c_double()
        lib().orc_geometry(self._h, iptr(nc), dptr(cs), C.byref(rn2), C.byref(rc2), C.byref(ps))
        return dict(ncells=nc, cellsidelength=cs, rneighbr2=rn2.value, rcutoff2=rc2.value,
                    potshift=ps.value)

    def set_moltype(self, moltype):
        _check(lib().orc_set_moltype(self._h, iptr(np.ascontiguousarray(moltype, np.int32))))

    def set_tags(self, tag, rtether=None):
        rt = None if rtether is None else np.ascontiguousarray(rtether, np.float64)
        _check(lib().orc_set_tags(self._h, iptr(np.ascontiguousarray(tag, np.int32)), dptr(rt)))

    def set_monomers(self, chainID, subchainID, funcy, bin_bflag):
        c = lambda x: np.ascontiguousarray(x, np.int32)
        _check(lib().orc_set_monomers(self._h, iptr(c(chainID)), iptr(c(subchainID)), iptr(c(funcy)),
                                      iptr(c(bin_bflag))))

    def state(self):
        n = self.np
        r, v, a = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        _check(lib().orc_get_state(self._h, dptr(r), dptr(v), dptr(a)))
        return r, v, a

    def rtether(self):
        out = np.zeros((self.np, 3))
        _check(lib().orc_get_tags(self._h, dptr(out)))
        return out

    def molecular(self):
        pe, vir = np.zeros(self.np), np.zeros(self.np)
        _check(lib().orc_get_molecular(self._h, dptr(pe), dptr(vir)))
        return pe, vir

    def rfmol(self):
        out = np.zeros((self.np, 3, 3))
        _check(lib().orc_get_rfmol(self._h, dptr(out)))
        return out

    def force_scale(self):
        out = np.zeros(self.np)
        _check(lib().orc_get_force_scale(self._h, dptr(out)))
        return out

    # reference routines -------------------------------------------------------------------
    def linklist_deallocateall(self):
        _check(lib().orc_linklist_deallocateall(self._h))

    def sendmols(self):
        _check(lib().orc_sendmols(self._h))

    def sort_mols(self):
        _check(lib().orc_sort_mols(self._h))

    def hcurve(self, nblocks):
        """Hcurve indexed [i, j, k] (0-based block indices; 1-based curve positions)."""
        out = np.zeros(int(np.prod(nblocks)), np.int32)
        _check(lib().orc_get_hcurve(self._h, iptr(out)))
        return out.reshape((nblocks[2], nblocks[1], nblocks[0])).transpose(2, 1, 0)

    def assign_to_cell(self):
        _check(lib().orc_assign_to_cell(self._h))

    def updateborders(self, rebuild):
        _check(lib().orc_updateborders(self._h, int(rebuild)))

    def assign_to_neighbourlist(self):
        _check(lib().orc_assign_to_neighbourlist(self._h))

    def compute_forces(self, iter_):
        _check(lib().orc_compute_forces(self._h, int(iter_)))

    def compute_forces_allpairs(self, iter_):
        _check(lib().orc_compute_forces_allpairs(self._h, int(iter_)))

    def move_particles_vv(self, pass_num):
        _check(lib().orc_move_particles_vv(self._h, int(pass_num)))

    def move_particles_lfv(self):
        _check(lib().orc_move_particles_lfv(self._h))

    def checkrebuild(self, iter_):
        rb = C.c_int32()
        _check(lib().orc_checkrebuild(self._h, int(iter_), C.byref(rb)))
        return rb.value

    def macros(self):
        m = FmMacros()
        _check(lib().orc_macros(self._h, C.byref(m)))
        return m.as_dict()

    def mop(self):
        out = np.zeros((self.params.nplanes, 3))
        _check(lib().orc_get_mop(self._h, dptr(out)))
        return out

    def reset_mop(self):
        _check(lib().orc_reset_mop(self._h))

    def cv_bins(self):
        """(nbins, nhb, binsize) of the control-volume tally."""
        nb = np.zeros(3, np.int32)
        nh = np.zeros(3, np.int32)
        bs = np.zeros(3)
        _check(lib().orc_get_cv_bins(self._h, nb.ctypes.data_as(C.POINTER(C.c_int32)),
                                     nh.ctypes.data_as(C.POINTER(C.c_int32)), dptr(bs)))
        return nb, nh, bs

    def cv_stress(self):
        """Pxyface as an array indexed [bx, by, bz, component, face] (0-based; faces 0..5)."""
        nb, nh, _ = self.cv_bins()
        nbo = nb + 2 * nh
        out = np.zeros(int(nbo.prod()) * 18)
        _check(lib().orc_get_cv_stress(self._h, dptr(out)))
        return out.reshape((6, 3, nbo[2], nbo[1], nbo[0])).transpose(4, 3, 2, 1, 0)

    def reset_cv_stress(self):
        _check(lib().orc_reset_cv_stress(self._h))

    def planes(self):
        out = np.zeros(self.params.nplanes)
        _check(lib().orc_get_planes(self._h, dptr(out)))
        return out

    def setup(self):
        """setup_MD's list build: sendmols is not called at setup (src/md.f90:60-63)."""
        self.assign_to_cell()
        self.updateborders(1)
        self.assign_to_neighbourlist()

    def rebuild(self):
        self.linklist_deallocateall()
        self.sendmols()
        self.sort_mols()
        self.assign_to_cell()
        self.updateborders(1)
        self.assign_to_neighbourlist()

    def run_vv(self, iter0, nsteps, n_macros=0):
        ms = (FmMacros * max(n_macros, 1))()
        nrb = C.c_int32()
        _check(lib().orc_run_vv(self._h, int(iter0), int(nsteps), ms, int(n_macros), C.byref(nrb)))
        return [ms[k].as_dict() for k in range(n_macros)], nrb.value

    def run_lfv(self, iter0, nsteps, n_macros=0):
        """md_advance_lfv (src/md.f90:109-139) nsteps times."""
        ms = (FmMacros * max(n_macros, 1))()
        nrb = C.c_int32()
        _check(lib().orc_run_lfv(self._h, int(iter0), int(nsteps), ms, int(n_macros), C.byref(nrb)))
        return [ms[k].as_dict() for k in range(n_macros)], nrb.value

    # inspection -----------------------------------------------------------------------------
    def cells(self):
        out = np.zeros(self.np, np.int32)
        _check(lib().orc_get_cells(self._h, iptr(out)))
        return out

    def halo_np(self):
        return lib().orc_halo_np(self._h)

    def neighbourlist(self):
        n = lib().orc_neighbour_count(self._h)
        if n < 0:
            raise OracleError("no neighbour list")
        i, j, s = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        _check(lib().orc_get_neighbourlist(self._h, iptr(i), iptr(j), iptr(s)))
        return i, j, s

    def bonds(self):
        b = np.zeros((self.np, FM_MAX_FUNCY), np.int32)
        s = np.zeros((self.np, FM_MAX_FUNCY), np.int32)
        _check(lib().orc_get_bonds(self._h, iptr(b), iptr(s)))
        return b, s

    def bruteforce_pairs(self):
        n = lib().orc_bruteforce_pairs(self._h, None, None, None, 0)
        i, j, s = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        lib().orc_bruteforce_pairs(self._h, iptr(i), iptr(j), iptr(s), n)
        return i, j, s
