"""ctypes mirror of the C structs declared in include/flowmol_hip.h.

The field order and types must match the header exactly; tests/test_abi.py checks the sizes
against the compiled library.
"""
import ctypes as C

FM_MAX_TYPES = 4
FM_MAX_FUNCY = 4
FM_MAX_PLANES = 1024

FM_OK = 0
FM_EABORT = 8
FM_ENSEMBLE_NVE = 0
FM_ENSEMBLE_TAG_MOVE = 6
FM_REBUILD_VMAX = 0
FM_REBUILD_FIXED = 1
FM_REBUILD_DISPLACEMENT = 2
FM_PREC_FP64 = 0
FM_PREC_MIXED = 1
FM_TAG_FREE = 0
FM_TAG_FIXED = 1
FM_TAG_FIXED_SLIDE = 2
FM_TAG_TETH = 3
FM_TAG_TETH_SLIDE = 6
FM_INTEGRATION_LFV = 0
FM_INTEGRATION_VV = 1

_T2 = FM_MAX_TYPES * FM_MAX_TYPES


class FmParams(C.Structure):
    _fields_ = [
        ("domain", C.c_double * 3),
        ("density", C.c_double),
        ("rcutoff", C.c_double),
        ("delta_rneighbr", C.c_double * 3),
        ("delta_t", C.c_double),
        ("tplot", C.c_int32),
        ("teval", C.c_int32),
        ("rebuild_criteria", C.c_int32),
        ("fixed_rebuild", C.c_int32),
        ("potential_flag", C.c_int32),
        ("mie_potential", C.c_int32),
        ("ntypes", C.c_int32),
        ("precision", C.c_int32),
        ("mass_lookup", C.c_double * FM_MAX_TYPES),
        ("epsilon_lookup", C.c_double * _T2),
        ("lambdar_lookup", C.c_double * _T2),
        ("lambdaa_lookup", C.c_double * _T2),
        ("C_lookup", C.c_double * _T2),
        ("alpha_lookup", C.c_double * _T2),
        ("potshift_lookup", C.c_double * _T2),
        ("rcutoff2_lookup", C.c_double * _T2),
        ("k_c", C.c_double),
        ("R_0", C.c_double),
        ("ensemble", C.c_int32),
        ("integration_algorithm", C.c_int32),
        ("teth_k2", C.c_double),
        ("teth_k4", C.c_double),
        ("teth_k6", C.c_double),
        ("wallslidev", C.c_double * 3),
        ("vflux_outflag", C.c_int32),
        ("pressure_outflag", C.c_int32),
        ("nplanes", C.c_int32),
        ("_pad1", C.c_int32),
        ("planespacing", C.c_double),
        ("max_neighbours", C.c_int32),
        ("max_cell_np", C.c_int32),
        ("brick", C.c_int32 * 3),
        ("_pad2", C.c_int32),
        ("binspercell", C.c_double * 3),
        ("cv_conserve", C.c_int32),
        ("_pad3", C.c_int32),
        ("sort_flag", C.c_int32),
        ("sort_freq", C.c_int32),
        ("sortblocksize", C.c_int32),
        ("_pad4", C.c_int32),
    ]


class FmMacros(C.Structure):
    _fields_ = [
        ("potenergysum", C.c_double),
        ("potenergysum_LJ", C.c_double),
        ("potenergysum_POLY", C.c_double),
        ("virial", C.c_double),
        ("mv2sum", C.c_double),
        ("vsum", C.c_double),
        ("msum", C.c_double),
        ("kinenergy", C.c_double),
        ("potenergy", C.c_double),
        ("totenergy", C.c_double),
        ("temperature", C.c_double),
        ("pressure", C.c_double),
        ("max_potenergymol", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def dptr(a):
    """double* of a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype.name == "float64" and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    """int32* of a C-contiguous int32 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype.name == "int32" and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(C.POINTER(C.c_int32))
