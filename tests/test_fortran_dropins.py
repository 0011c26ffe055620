"""The reference-side drop-in bodies (flowmol_amd/fortran/flowmol_hip_routines.f90) compile
against the reference's own modules.

The modules the drop-ins use are compiled with amdflang from the reference's sources in a
temporary directory: modules.f90 (physical_constants_MD, computational_constants_MD, arrays_MD,
polymer_info_MD, calculated_properties_MD), interfaces.f90 (error_abort), linklist.f90 and the
module part of setup_set_parameters.f90 (module_set_parameters, whose Mie tables the drop-ins
read; the rest of that file needs the MPI messenger, which this image cannot build).  Nothing is
run and nothing is written into the repository.  Skipped where the reference or amdflang is
absent (the GPU box has no reference tree).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
FC = shutil.which("amdflang")

pytestmark = pytest.mark.skipif(not (os.path.isdir(REF) and FC), reason="reference sources or amdflang absent")

REF_FILES = ["minpack.f90", "interfaces.f90", "bilinear_intersect.f90", "functions_lib.f90",
             "intrinsic_interface.f90", "modules.f90", "linklist.f90"]


def _fc(args, cwd):
    r = subprocess.run([FC, "-cpp", "-DUSE_COUPLER=0", "-O0", "-c"] + args, cwd=cwd, capture_output=True, text=True,
                       timeout=600)
    return r.returncode, r.stderr


@pytest.fixture(scope="module")
def refmods(tmp_path_factory):
    d = tmp_path_factory.mktemp("refmods")
    for f in REF_FILES:
        rc, err = _fc([os.path.join(REF, f), "-J", str(d), "-o", str(d / (f + ".o"))], str(d))
        assert rc == 0, (f, err[-2000:])
    # the module part of setup_set_parameters.f90
    lines = open(os.path.join(REF, "setup_set_parameters.f90")).read().splitlines(keepends=True)
    end = next(k for k, l in enumerate(lines) if l.strip().lower().startswith("end module module_set_parameters"))
    (d / "module_set_parameters_part.f90").write_text("".join(lines[:end + 1]))
    rc, err = _fc([str(d / "module_set_parameters_part.f90"), "-J", str(d), "-o", str(d / "msp.o")], str(d))
    assert rc == 0, err[-2000:]
    return d


def test_dropins_compile_against_reference_modules(refmods):
    d = refmods
    fdir = os.path.join(ROOT, "flowmol_amd", "fortran")
    for f in ("flowmol_hip_c.f90", "flowmol_hip_routines.f90"):
        rc, err = _fc([os.path.join(fdir, f), "-I", str(d), "-J", str(d), "-o", str(d / (f + ".o"))], str(d))
        assert rc == 0, (f, err[-3000:])
    # every replaced routine is defined with the reference's name
    out = subprocess.run(["nm", str(d / "flowmol_hip_routines.f90.o")], capture_output=True, text=True).stdout
    for name in ("assign_to_cell", "assign_to_neighbourlist", "linklist_deallocateall", "sort_mols", "sendmols",
                 "messenger_updateborders", "simulation_checkrebuild", "simulation_move_particles_vv",
                 "simulation_move_particles_lfv", "simulation_compute_forces"):
        assert any(l.split()[-1].rstrip("_") == name and " T " in l for l in out.splitlines()), name
