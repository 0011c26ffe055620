"""The C-ABI library loads and exports every symbol include/flowmol_hip.h declares; the ctypes
struct mirrors match the C layout.  No compute calls (runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from flowmol_amd import _abi, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "flowmol_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fm_[a-z0-9_]+)\s*\(", txt)))


def test_header_parses():
    names = declared_functions()
    assert "fm_init" in names and "fm_compute_forces" in names and len(names) >= 30


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libflowmol_hip.so not built (python -m flowmol_amd.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (fm_[a-z0-9_]+)$", out, flags=re.M))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    # the ctypes table covers exactly the declared functions
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_library_loads_and_reports_version():
    L = _lib.load()
    txt = open(HEADER).read()
    assert L.fm_abi_version() == int(re.search(r"#define FM_ABI_VERSION (\d+)", txt).group(1))


def _c_layout():
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "flowmol_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(fm_params), sizeof(fm_macros),
         offsetof(fm_params, k_c), offsetof(fm_params, planespacing),
         offsetof(fm_params, max_cell_np), offsetof(fm_macros, max_potenergymol),
         offsetof(fm_params, binspercell), offsetof(fm_params, cv_conserve), offsetof(fm_params, sortblocksize));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        exe = os.path.join(d, "l")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        return [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]


def test_struct_layout_matches_header():
    sp, sm, okc, ops, omc, omax, obin, ocv, osb = _c_layout()
    assert C.sizeof(_abi.FmParams) == sp
    assert C.sizeof(_abi.FmMacros) == sm
    assert _abi.FmParams.k_c.offset == okc
    assert _abi.FmParams.planespacing.offset == ops
    assert _abi.FmParams.max_cell_np.offset == omc
    assert _abi.FmMacros.max_potenergymol.offset == omax
    assert _abi.FmParams.binspercell.offset == obin
    assert _abi.FmParams.cv_conserve.offset == ocv
    assert _abi.FmParams.sortblocksize.offset == osb


def test_memory_footprint_terms():
    """fm_memory_footprint (host arithmetic, no device): grows with the molecule count; the
    staging tables scale with the brick's halo factor (1-cell bricks hold each molecule 27 times
    over, 3-cell bricks about 4.6 times); for a large system the estimate covers the state
    copies, the list words and the tables, plus fixed terms of at most 16 MiB."""
    from flowmol_amd import config
    from flowmol_amd.serve import memory_footprint

    p, r, _ = config.config_fcc((16, 16, 16), 0.8, 2.5, seed=1, tplot=10)
    n = r.shape[0]
    small, big = memory_footprint(p, n // 2), memory_footprint(p, n)
    assert 0 < small < big
    p1 = type(p).from_buffer_copy(p)
    p3 = type(p).from_buffer_copy(p)
    for d in range(3):
        p1.brick[d], p3.brick[d] = 1, 3
    f1, f3 = memory_footprint(p1, n), memory_footprint(p3, n)
    tables = lambda halo: 10.0 * 1.125 * 1.6 * halo * n
    assert abs((f1 - f3) - (tables(27.0) - tables(125.0 / 27.0))) < 8192
    state = (2 * 9 + 3 + 9) * 8 * n           # state copies, rdisp, staging
    assert f3 >= state + tables(125.0 / 27.0)
    assert f3 <= 2 * state + 16 * 2**20 + tables(125.0 / 27.0) + 400 * n
