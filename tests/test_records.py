"""Record-step output formats (results/macroscopic_properties, results/pplane) and the kinetic MOP
term, pinned by hand-written lines and hand-computed crossings.  CPU only."""
import numpy as np
import pytest

from flowmol_amd import records as R


def test_fortran_fixed_point_descriptor():
    # Fw.d: right-justified; a leading zero only when there is room; w asterisks on overflow
    assert R.fortran_f(0.5, 10, 4) == "    0.5000"
    assert R.fortran_f(-0.5, 10, 4) == "   -0.5000"
    assert R.fortran_f(1234.56789, 10, 4) == " 1234.5679"
    assert R.fortran_f(123456.0, 10, 4) == "**********"
    assert R.fortran_f(-0.123, 6, 4) == "-.1230"
    assert R.fortran_f(-5.25, 19, 15) == " -5.250000000000000"
    assert R.fortran_f(-0.00001, 10, 4) == "   -0.0000"
    assert R.fortran_i(12, 8) == "      12"
    assert R.fortran_i(123456789, 8) == "********"


def test_macroscopic_properties_line_lj():
    m = dict(vsum=1e-12, mv2sum=6144.0, temperature=1.0, kinenergy=1.5, potenergy=-2.25, totenergy=-0.75,
             pressure=0.8)
    want = ("       10;         0.0500;         0.0000;      6144.0000;    1.0000;  1.500000000000000;"
            " -2.250000000000000; -0.750000000000000;    0.8000")
    assert R.macroscopic_properties_line(10, 0.05, m) == want  # 1x then i8


def test_macroscopic_properties_line_polymer_keeps_the_shifted_descriptors():
    # with potential_flag 1 the reference's format gives the temperature f15.4 and KE f10.4
    m = dict(vsum=0.0, mv2sum=3.0, temperature=1.25, kinenergy=0.625, potenergysum_LJ=-8.0,
             potenergysum_POLY=4.0, potenergy=-1.0, totenergy=-0.375, pressure=2.5, np=2)
    line = R.macroscopic_properties_line(20, 0.1, m, potential_flag=1)
    f = line.split(";")
    assert f[0] == "       20" and len(f) == 13
    assert f[4] == "         1.2500" and f[5] == "    0.6250"
    assert f[6] == " -2.000000000000000" and f[7] == "  1.000000000000000"
    assert f[10] == "  2.500000000000000" and f[11] == "    0.0000"


def test_macroscopic_properties_file_round_trip(tmp_path):
    w = R.MacroscopicProperties(str(tmp_path), 0)
    m = dict(vsum=0.0, mv2sum=6.0, temperature=1.0, kinenergy=1.5, potenergy=-2.0, totenergy=-0.5, pressure=0.7)
    w.record(5, 0.025, m)
    w.record(10, 0.05, dict(m, potenergy=-2.5))
    lines = open(w.path).read().splitlines()
    assert lines[0] == " iter; simtime; VSum; V^2Sum; Temp; KE; PE; TE; Pressure"
    cols = R.read_macroscopic_properties(w.path)
    assert list(cols["iter"]) == [5, 10]
    assert list(cols["PE"]) == [-2.0, -2.5]


def test_mop_kinetic_hand_crossings():
    # planes of a 10-unit box, 10 planes 1 apart: shift 0.5, planes at -4.5 ... 4.5; plane 5 at -0.5
    planes, shift = R.mop_planes((10.0, 10.0, 10.0), 10, 1.0)
    assert shift == 0.5 and planes[4] == -0.5 and planes[0] == -4.5
    dt = 0.005
    r = np.array([[0.0, -0.4, 0.0],    # came from -0.6: up through plane 5
                  [0.0, -0.6, 0.0],    # came from -0.4: down through plane 5
                  [0.0, 0.1, 0.0]])    # from -0.05: no plane between
    v = np.array([[1.0, 40.0, 2.0], [3.0, -40.0, 4.0], [5.0, 30.0, 6.0]])
    out = R.mop_kinetic(r, v, 1.0, (10.0, 10.0, 10.0), 10, 1.0, dt)
    want = np.zeros((3, 10))
    want[:, 4] = v[0] + v[1]
    assert np.array_equal(out, want)
    # mass scales the term
    out2 = R.mop_kinetic(r[:1], v[:1], 2.5, (10.0, 10.0, 10.0), 10, 1.0, dt)
    assert np.array_equal(out2[:, 4], 2.5 * v[0])


def test_pplane_normalisation_and_record_layout(tmp_path):
    P = np.arange(12, dtype=np.float64).reshape(3, 4) + 1.0
    n = R.pplane_normalise(P, 2, (3.0, 7.0, 5.0))
    assert np.allclose(n, P / 2 / (4 * 3.0 * 5.0), rtol=1e-15)
    # record m = iter/(tplot*Nmass_ave) at byte (m-1)*3*nplanes*8, first nplanes doubles of the
    # column-major array: P(1,1), P(2,1), P(3,1), P(1,2)
    R.write_pplane(str(tmp_path), 10, 10, 1, P)
    R.write_pplane(str(tmp_path), 20, 10, 1, 2 * P)
    raw = np.fromfile(str(tmp_path / "results" / "pplane"), dtype="<f8")
    assert raw.size == 12 + 4
    assert list(raw[:4]) == [1.0, 5.0, 9.0, 2.0]
    assert np.all(raw[4:12] == 0.0)
    assert list(raw[12:16]) == [2.0, 10.0, 18.0, 4.0]
