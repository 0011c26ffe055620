"""fm_macros_sums_noalloc (the record-step sums without any allocation) against
fm_evaluate_macroscopic_properties on the same states, through the C ABI: two calls in a row, the
second with fewer molecules, both within 1e-12 relative of the context's sums (the two reduce in
different orders: slot order in the context, molecule order here), the maximum exactly.  The
caller's device buffers are made here with the HIP runtime; the entry point itself allocates
nothing (the device's free memory is the same before and after a call)."""
import ctypes as C
import sys

import numpy as np
import pytest

from flowmol_amd import config
from flowmol_amd._abi import FmMacros
from flowmol_amd._lib import check, load
from flowmol_amd.md import MDContext

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device
        self.bufs.append(p)
        return p

    def zeros(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        assert self.L.hipMemset(p, 0, C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p

    def download(self, p, n):
        out = np.zeros(n)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free_bytes(self):
        f, t = C.c_size_t(), C.c_size_t()
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    def close(self):
        for p in self.bufs:
            self.L.hipFree(p)


def _ctx_and_arrays(nunits, seed):
    p, r, v = config.config_fcc(nunits, 0.8, 2.5, seed=seed, tplot=5)
    r = r + np.random.default_rng(seed).uniform(-0.03, 0.03, r.shape)
    g = MDContext(p, r, v)
    g.setup(0)
    g.run_vv(1, 10)          # step 10 is a record step: PE and virial valid
    m = g.evaluate_macroscopic_properties()
    _, v10, _ = g.state()
    lj, poly = g.potenergy_parts()
    _, vir = g.molecular()
    return p, g, m, v10, lj, vir


def test_noalloc_sums_match_the_context_entry_point():
    L = load()
    hip = Hip()
    try:
        for nunits, seed in (((8, 8, 8), 91), ((6, 6, 6), 92)):   # the second call is smaller
            p, g, m, v, lj, vir = _ctx_and_arrays(nunits, seed)
            n = v.shape[0]
            say(f"case: fm_macros_sums_noalloc, {n} molecules")
            dv, dlj, dvir = hip.upload(v), hip.upload(lj), hip.upload(vir)
            dsums = hip.zeros(8 * 8)
            free0 = hip.free_bytes()
            check(L.fm_macros_sums_noalloc(n, dv, None, p.delta_t, dlj, None, dvir, None, dsums, None))
            sums = hip.download(dsums, 8)
            assert hip.free_bytes() == free0
            out = FmMacros()
            check(L.fm_derive_macros(sums.ctypes.data_as(C.POINTER(C.c_double)), n, p.density, p.potential_flag,
                                     C.byref(out)))
            got = out.as_dict()
            for k in ("potenergysum", "virial", "mv2sum", "kinenergy", "potenergy", "totenergy", "temperature",
                      "pressure", "msum"):
                assert got[k] == pytest.approx(m[k], rel=1e-12, abs=1e-12), k
            assert abs(got["vsum"] - m["vsum"]) <= 1e-10
            assert got["max_potenergymol"] == m["max_potenergymol"]
            g.close()
        # a size beyond the static scratch is refused before any launch
        rc = L.fm_macros_sums_noalloc((1 << 24) + 1, dv, None, 0.005, dlj, None, dvir, None, dsums, None)
        assert rc != 0 and b"static scratch" in L.fm_last_error()
    finally:
        hip.close()
