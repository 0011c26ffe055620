"""Pins of the CPU restatement (oracle) -- each check is independent of the restated formula:
closed forms and known answers, numerical derivatives of the oracle's own energies, brute force
over all pairs and images, the reference's all-pairs routine as a second algorithm, hand-built
trajectories, invariants, and the reference's own restart fixture.  CPU only.
"""
import math

import numpy as np
import pytest

from flowmol_amd import config
from flowmol_amd._abi import FM_REBUILD_DISPLACEMENT, FM_REBUILD_FIXED, FM_REBUILD_VMAX
from flowmol_amd.restart import read_final_state, write_final_state
from oracle.oracle import Oracle, OracleError
from tests.helpers import canonical_pairs

GOLDEN = "tests/golden/intrinsic_interface_initial_state"


def box_params(L=20.0, rc=2.5, **kw):
    return config.make_params((L, L, L), 0.5, rc, **kw)


def two_body(r, rc=2.5, L=20.0, **kw):
    p = box_params(L, rc, **kw)
    pos = np.array([[0.1, 0.2, 0.3], [0.1 + r, 0.2, 0.3]])
    o = Oracle(p, pos, np.zeros((2, 3)))
    o.setup()
    o.compute_forces(0)
    return o


# ---- geometry (set_parameters_cells) ------------------------------------------------------------
def test_geometry_matches_reference_derivation():
    # values derived from the reference's formulas for the two headline lattices
    p, r, v = config.config_c1()
    g = Oracle(p, r, v).geometry()
    assert list(g["ncells"]) == [9, 9, 9]
    assert abs(p.domain[0] - 13.680) < 1e-3 and abs(g["cellsidelength"][0] - 1.520) < 1e-3
    assert abs(g["rneighbr2"] - (2 ** (1 / 6) + 0.3) ** 2) < 1e-14
    assert g["potshift"] == pytest.approx(-1.0, abs=1e-14)  # 4(rc^-12 - rc^-6) at rc = 2^(1/6)
    p, r, v = config.config_c2(nunits=8)
    p.domain[0] = p.domain[1] = p.domain[2] = 109.44
    g = Oracle(p, r, v).geometry()
    assert list(g["ncells"]) == [39, 39, 39] and abs(g["cellsidelength"][0] - 2.806) < 1e-3


# ---- pair law ------------------------------------------------------------------------------------
def test_lj_force_zero_at_minimum_and_energy_shift():
    o = two_body(2 ** (1 / 6))
    _, _, a = o.state()
    assert np.abs(a).max() < 1e-12
    pe, _ = o.molecular()
    # U(2^(1/6)) = -1 unshifted; shifted by -potshift
    ps = o.geometry()["potshift"]
    assert pe[0] == pytest.approx(-1.0 - ps, abs=1e-13)


def test_lj_energy_vanishes_at_cutoff():
    o = two_body(2.5 - 1e-12)
    pe, _ = o.molecular()
    assert abs(pe[0]) < 1e-11


@pytest.mark.parametrize("r", [0.95, 1.05, 1.3, 1.8, 2.3])
def test_lj_force_is_minus_energy_gradient(r):
    h = 1e-6
    e = lambda d: two_body(d).molecular()[0][0]
    dedr = (e(r + h) - e(r - h)) / (2 * h)
    _, _, a = two_body(r).state()
    # molecule 0 sits at -x of molecule 1: its force is -(-dU/dr) along +x ... a0_x = dU/dr
    assert a[0, 0] == pytest.approx(dedr, rel=1e-7)
    assert a[1, 0] == pytest.approx(-dedr, rel=1e-7)
    assert np.abs(a[:, 1:]).max() == 0.0


def test_three_body_superposition_and_newton():
    p = box_params()
    pos = np.array([[0.0, 0.0, 0.0], [1.1, 0.0, 0.0], [0.3, 1.2, -0.4]])
    o = Oracle(p, pos, np.zeros((3, 3)))
    o.setup()
    o.compute_forces(0)
    _, _, a = o.state()
    total = np.zeros(3)
    for (i, j) in [(0, 1), (0, 2)]:
        q = Oracle(p, pos[[i, j]], np.zeros((2, 3)))
        q.setup()
        q.compute_forces(0)
        total += q.state()[2][0]
    assert np.allclose(a[0], total, rtol=1e-14, atol=1e-14)
    assert np.abs(a.sum(axis=0)).max() < 1e-13


def test_virial_two_body():
    r = 1.2
    o = two_body(r)
    _, vir = o.molecular()
    _, _, a = o.state()
    # per molecule: accijmag * r^2 with accijmag = a_x / r_x
    acc = -a[0, 0] / r
    assert vir[0] == pytest.approx(acc * r * r, rel=1e-14)
    assert vir[1] == pytest.approx(acc * r * r, rel=1e-14)


def test_mie_12_6_reduces_to_lj():
    assert config.mie_C(12.0, 6.0) == pytest.approx(4.0, abs=1e-15)
    p, r, v = config.config_fcc((5, 5, 5), 0.8, 2.5, seed=3)
    rng = np.random.default_rng(1)
    r = r + rng.uniform(-0.05, 0.05, r.shape)
    o1 = Oracle(p, r, v)
    o1.setup()
    o1.compute_forces(0)
    p2, _, _ = config.config_fcc((5, 5, 5), 0.8, 2.5, seed=3)
    config.set_mie_types(p2, {(0, 0): dict(lr=12.0, la=6.0, eps=1.0)})
    o2 = Oracle(p2, r, v)
    o2.set_moltype(np.zeros(r.shape[0], np.int32))
    o2.setup()
    o2.compute_forces(0)
    a1, a2 = o1.state()[2], o2.state()[2]
    assert np.abs(a1 - a2).max() <= 1e-12 * np.abs(a1).max()
    assert np.allclose(o1.molecular()[0], o2.molecular()[0], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("r", [1.0, 1.2, 1.7])
def test_mie_14_6_force_is_minus_energy_gradient(r):
    def run(d):
        p = box_params()
        config.set_mie_types(p, {(0, 0): dict(lr=14.0, la=6.0, eps=1.0, rc=2.5)})
        pos = np.array([[0.1, 0.2, 0.3], [0.1 + d, 0.2, 0.3]])
        o = Oracle(p, pos, np.zeros((2, 3)))
        o.set_moltype(np.zeros(2, np.int32))
        o.setup()
        o.compute_forces(0)
        return o
    h = 1e-6
    dedr = (run(r + h).molecular()[0][0] - run(r - h).molecular()[0][0]) / (2 * h)
    assert run(r).state()[2][0, 0] == pytest.approx(dedr, rel=1e-7)


def _mie_pair(d, eps, lr=14.0, la=6.0, rc=2.5):
    p = box_params()
    config.set_mie_types(p, {(0, 0): dict(lr=lr, la=la, eps=eps, rc=rc)})
    pos = np.array([[0.1, 0.2, 0.3], [0.1 + d, 0.2, 0.3]])
    o = Oracle(p, pos, np.zeros((2, 3)))
    o.set_moltype(np.zeros(2, np.int32))
    o.setup()
    o.compute_forces(0)
    return p, o


@pytest.mark.parametrize("eps", [1.0, 1.7])
def test_mie_14_6_energy_vanishes_at_cutoff_and_minimum_is_minus_eps(eps):
    """Closed forms of the 14-6 law, written out by hand: C = (14/8)(14/6)^(6/8); the minimum of
    C eps (r^-14 - r^-6) lies at r_min = (14/6)^(1/8) with depth -eps; after the shift by the
    pair's own energy at rc the energy is zero at rc.  A dropped eps, a wrong C, a swapped
    exponent or a shift taken at another radius fails one of them."""
    C = 1.75 * (7.0 / 3.0) ** 0.75
    assert config.mie_C(14.0, 6.0) == pytest.approx(C, rel=1e-15)
    U_rc = C * eps * (2.5 ** -14 - 2.5 ** -6)
    p, o = _mie_pair(2.5 - 1e-12, eps)
    assert p.potshift_lookup[0] == pytest.approx(U_rc, rel=1e-14)
    assert abs(o.molecular()[0][0]) < 1e-11
    r_min = (14.0 / 6.0) ** 0.125
    _, o = _mie_pair(r_min, eps)
    pe, _ = o.molecular()
    assert pe[0] == pytest.approx(-eps - U_rc, rel=1e-13)
    assert pe[1] == pe[0]
    assert np.abs(o.state()[2]).max() < 1e-12 * eps


# ---- pressure: pressure = density/(3N) (sum m v^2 + virial/2), src/simulation_record.f90:1286 ----
def test_pressure_without_pairs_is_density_times_temperature():
    """Eight molecules 10 apart (no pair within rc + delta): virial 0 and P = rho T exactly."""
    L = 20.0
    pos = np.array([[x, y, z] for x in (-5.0, 5.0) for y in (-5.0, 5.0) for z in (-5.0, 5.0)])
    v = np.random.default_rng(3).standard_normal((8, 3))
    p = config.make_params((L, L, L), 8 / L ** 3, 2.5)
    o = Oracle(p, pos, v)
    o.setup()
    o.compute_forces(0)
    m = o.macros()
    assert m["virial"] == 0.0
    T = float((v * v).sum()) / 24.0
    assert m["temperature"] == pytest.approx(T, rel=1e-14)
    assert m["pressure"] == pytest.approx((8 / L ** 3) * T, rel=1e-14)


@pytest.mark.parametrize("r", [1.05, 1.2, 1.6])
def test_pressure_two_body_closed_form(r):
    """Two molecules at rest at separation r in a box of volume V (density 2/V): the pressure is
    the pair's virial r.f(r)/(3V) with f(r) = 48 (r^-13 - r^-7/2), counted once although both
    molecules tally it (the virial/2 of the reference)."""
    L = 20.0
    V = L ** 3
    p = config.make_params((L, L, L), 2 / V, 2.5)
    pos = np.array([[0.1, 0.2, 0.3], [0.1 + r, 0.2, 0.3]])
    o = Oracle(p, pos, np.zeros((2, 3)))
    o.setup()
    o.compute_forces(0)
    m = o.macros()
    rf = r * 48.0 * (r ** -13 - 0.5 * r ** -7)
    assert m["virial"] == pytest.approx(2 * rf, rel=1e-13)
    assert m["pressure"] == pytest.approx(rf / (3 * V), rel=1e-13)
    assert m["temperature"] == 0.0


# ---- neighbour list ------------------------------------------------------------------------------
def _disordered(nunits, seed, density=0.8, rc=2.5, steps=150, **kw):
    p, r, v = config.config_fcc(nunits, density, rc, seed=seed, **kw)
    r = r + np.random.default_rng(seed).uniform(-1e-3, 1e-3, r.shape)
    o = Oracle(p, r, v)
    o.setup()
    o.compute_forces(0)
    o.run_vv(1, steps)
    o.rebuild()
    return p, o


@pytest.mark.parametrize("nunits,rc", [((5, 5, 5), 2.5), ((8, 8, 8), 2 ** (1 / 6)), ((6, 5, 7), 2.5)])
def test_half_list_equals_brute_force_set(nunits, rc):
    p, o = _disordered(nunits, seed=4, rc=rc)
    i, j, s = o.neighbourlist()
    bi, bj, bs = o.bruteforce_pairs()
    assert np.array_equal(canonical_pairs(i, j, s), canonical_pairs(bi, bj, bs))
    keys = (i.astype(np.int64) << 32) | (j.astype(np.int64) << 5) | s
    assert np.unique(keys).size == keys.size  # no duplicate entries
    # every shifted pair appears once per end, every unshifted pair once
    shifted = s != 13
    assert np.sum(shifted) == np.sum(bs != 13)
    assert np.sum(~shifted) * 2 == np.sum(bs == 13)


def test_list_forces_equal_all_pairs_routine():
    p, o = _disordered((6, 6, 6), seed=5)
    o.compute_forces(0)
    _, _, a1 = o.state()
    pe1, vir1 = o.molecular()
    scale = o.force_scale()
    o.compute_forces_allpairs(0)
    _, _, a2 = o.state()
    pe2, vir2 = o.molecular()
    assert (np.abs(a1 - a2).max(axis=1) / scale).max() < 1e-13
    assert np.allclose(pe1, pe2, rtol=1e-12, atol=1e-14)
    assert np.allclose(vir1, vir2, rtol=1e-12, atol=1e-12)
    assert np.abs(a1.sum(axis=0)).max() < 1e-10 * scale.sum()


# ---- integrator and rebuild check ----------------------------------------------------------------
def test_vv_free_flight_and_kick():
    p = box_params(dt=0.01)
    pos = np.array([[0.0, 0.0, 0.0], [5.0, 5.0, 5.0]])
    vel = np.array([[0.3, -0.2, 0.1], [0.0, 0.0, 0.0]])
    o = Oracle(p, pos, vel, a=np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]]))
    o.move_particles_vv(1)
    r, v, _ = o.state()
    # r + dt v + dt^2/2 a ; v + dt/2 a
    assert r[0, 0] == pytest.approx(0.003 + 0.00005, abs=1e-16)
    assert v[0, 0] == pytest.approx(0.305, abs=1e-16)
    o.move_particles_vv(2)
    assert o.state()[1][0, 0] == pytest.approx(0.31, abs=1e-16)


def test_vv_energy_drift_is_second_order():
    def drift(dt, steps):
        p, r, v = config.config_fcc((5, 5, 5), 0.8, 2.5, seed=6, dt=dt, tplot=1)
        r = r + np.random.default_rng(6).uniform(-0.02, 0.02, r.shape)
        o = Oracle(p, r, v)
        o.setup()
        o.compute_forces(0)
        ms, _ = o.run_vv(1, steps, n_macros=steps)
        e = np.array([m["totenergy"] for m in ms])
        return np.abs(e - e[0]).max()
    d1 = drift(0.004, 200)
    d2 = drift(0.002, 400)
    assert 2.5 < d1 / d2 < 6.0   # energy error of velocity Verlet scales as dt^2


@pytest.mark.parametrize("crit", [FM_REBUILD_DISPLACEMENT, FM_REBUILD_VMAX])
def test_rebuild_fires_when_displacement_exceeds_half_skin(crit):
    # one molecule moving at 1.1 along y: 0.0055 per step crosses 0.15 at the 28th check
    p = box_params(dt=0.005, tplot=10 ** 6, rebuild_criteria=crit)
    pos = np.array([[0.0, 0.0, 0.0], [6.0, 6.0, 6.0]])
    vel = np.array([[0.0, 1.1, 0.0], [0.0, 0.0, 0.0]])
    o = Oracle(p, pos, vel)
    hits = [it for it in range(1, 40) if o.checkrebuild(it)]
    assert hits[0] == 28


def test_forced_and_fixed_rebuild():
    p = box_params(tplot=7, rebuild_criteria=FM_REBUILD_FIXED, fixed_rebuild=5)
    o = Oracle(p, np.zeros((1, 3)), np.zeros((1, 3)))
    hits = [it for it in range(1, 22) if o.checkrebuild(it)]
    assert hits == [5, 6, 10, 13, 15, 20]   # mod(iter,5)==0 or mod(iter+1,7)==0


# ---- sendmols renumbering --------------------------------------------------------------------------
def test_sendmols_reverses_leavers_per_pass():
    """Leavers of one axis/direction pass: molecules a<b<c leave through +x; afterwards slot a
    holds c's molecule, slot c holds a's, b keeps its own (the pass list is head-inserted, packed
    from the head, received after np and copied back from the end)."""
    p = box_params(L=10.0)
    n = 10
    pos = np.zeros((n, 3))
    pos[:, 0] = np.linspace(-4, 4, n)
    vel = np.arange(3 * n, dtype=float).reshape(n, 3)
    leavers = [2, 5, 9]
    for k in leavers:
        pos[k, 0] = 5.0 + 0.01 * k   # beyond +L/2
    o = Oracle(p, pos, vel)
    o.sendmols()
    r, v, _ = o.state()
    assert np.array_equal(v[2], vel[9]) and np.array_equal(v[9], vel[2]) and np.array_equal(v[5], vel[5])
    assert r[2, 0] == pytest.approx(5.09 - 10.0, abs=1e-14)
    for k in set(range(n)) - set(leavers):
        assert np.array_equal(v[k], vel[k])


# ---- bonds ----------------------------------------------------------------------------------------
def chain_topology(nchains, nmono):
    n = nchains * nmono
    chain = np.repeat(np.arange(1, nchains + 1), nmono).astype(np.int32)
    sub = np.tile(np.arange(1, nmono + 1), nchains).astype(np.int32)
    funcy = np.where((sub == 1) | (sub == nmono), 1, 2).astype(np.int32)
    bflag = np.zeros((n, 4), np.int32)
    for m in range(n):
        for nb in (sub[m] - 1, sub[m] + 1):
            if 1 <= nb <= nmono:
                g = (nb - 1) // 31
                bflag[m, g] |= 1 << ((nb - 1) % 31)
    return chain, sub, funcy, bflag


def test_fene_bond_force_energy_and_discovery():
    k, R0, r = 30.0, 1.5, 1.1
    p = box_params(potential_flag=1, k_c=k, R_0=R0, rc=2 ** (1 / 6))
    pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0], [r, r, 0.0]])
    o = Oracle(p, pos, np.zeros((3, 3)))
    o.set_monomers(*chain_topology(1, 3))
    o.setup()
    o.compute_forces(0)
    b, s = o.bonds()
    assert sorted(b[0][b[0] >= 0]) == [1] and sorted(b[1][b[1] >= 0]) == [0, 2] and sorted(b[2][b[2] >= 0]) == [1]
    # standard FENE U = -k R0^2/2 ln(1 - r^2/R0^2): force on 0 along +x is -dU/dr ... toward 1
    f_std = k * r / (1 - r * r / R0 / R0)
    _, _, a = o.state()
    # beads 0 and 1 are also within the WCA cutoff? r = 1.1 < 1.1225: add the LJ term
    lj = 48 * (r ** -14 - 0.5 * r ** -8) * r
    assert a[0, 0] == pytest.approx(f_std - lj, rel=1e-12)
    pe, _ = o.molecular()
    # the reference's FENE energy carries the opposite sign of the standard form
    e_std = -0.5 * k * R0 * R0 * math.log(1 - r * r / R0 / R0)
    assert pe[0] == pytest.approx(-e_std + (4 * (r ** -12 - r ** -6) + 1.0), rel=1e-12)


def test_fene_bond_breaks_beyond_R0():
    """A bond stretched to R0 or beyond aborts the force call (polymer_bond_error)."""
    p = box_params(potential_flag=1, k_c=30.0, R_0=1.5, rc=2 ** (1 / 6))
    ok = Oracle(p, np.array([[0.3, 0.0, 0.0], [1.79, 0.0, 0.0]]), np.zeros((2, 3)))
    ok.set_monomers(*chain_topology(1, 2))
    ok.setup()
    ok.compute_forces(0)
    bad = Oracle(p, np.array([[0.3, 0.0, 0.0], [1.9, 0.0, 0.0]]), np.zeros((2, 3)))
    bad.set_monomers(*chain_topology(1, 2))
    bad.setup()
    assert sorted(bad.bonds()[0][0][bad.bonds()[0][0] >= 0]) == [1]
    with pytest.raises(OracleError):
        bad.compute_forces(0)


# ---- tethers -------------------------------------------------------------------------------------
def test_tether_force_is_minus_gradient_and_slides():
    from flowmol_amd._abi import FM_ENSEMBLE_TAG_MOVE, FM_TAG_TETH_SLIDE
    k2, k4, k6 = 0.0, 5000.0, 5.0e6
    def run(d):
        p = box_params(ensemble=FM_ENSEMBLE_TAG_MOVE)
        p.teth_k2, p.teth_k4, p.teth_k6 = k2, k4, k6
        p.wallslidev[0] = 1.0
        pos = np.array([[0.0, 3.0 + d, 0.0]])
        o = Oracle(p, pos, np.zeros((1, 3)))
        o.set_tags(np.array([FM_TAG_TETH_SLIDE], np.int32), rtether=np.array([[0.0, 3.0, 0.0]]))
        o.setup()
        o.compute_forces(0)
        return o
    d = 0.03
    U = lambda x: (2.0 / 3.0) * k2 * abs(x) ** 3 + k4 * x ** 4 + k6 * x ** 6
    h = 1e-7
    f_num = -(U(d + h) - U(d - h)) / (2 * h)
    o = run(d)
    assert o.state()[2][0, 1] == pytest.approx(f_num, rel=1e-6)
    # the tether point slides by slidev*dt per force evaluation, slidev = wallslidev*sign(y)
    assert o.rtether()[0, 0] == pytest.approx(1.0 * 0.005, abs=1e-15)


# ---- leap-frog (simulation_move_particles_lfv) ---------------------------------------------------
def _match(ra, rb, L):
    """Largest distance between two position sets after wrapping into the box and pairing each
    point with its nearest partner (the rebuilds renumber molecules at different steps)."""
    from scipy.spatial import cKDTree
    L = np.asarray(L)
    wa = (ra + 0.5 * L) % L
    wb = (rb + 0.5 * L) % L
    d, _ = cKDTree(wb, boxsize=L).query(wa)
    return float(d.max())


def _lfv_params(p):
    from flowmol_amd._abi import FM_INTEGRATION_LFV
    q = type(p).from_buffer_copy(p)
    q.integration_algorithm = FM_INTEGRATION_LFV
    return q


def test_lfv_positions_equal_vv_nve():
    """Leap-frog and velocity Verlet give the same Stormer-Verlet positions when the leap-frog
    velocity starts half a step back, v(-dt/2) = v(0) - dt/2 a(0); leap-frog step k records state
    k with v(k-dt/2) + dt/2 a(k) = v(k), so its energies equal the VV record of step k."""
    from tests.helpers import equilibrated
    p, r, v, a = equilibrated(config.config_fcc, 100, nunits=(5, 5, 5), density=0.8, rcutoff=2.5, seed=31,
                              tplot=5)
    dt = p.delta_t
    ov = Oracle(p, r, v, a)
    ov.setup()
    ov.compute_forces(0)
    a0 = ov.state()[2]
    mv, nrv = ov.run_vv(1, 60, n_macros=12)
    ol = Oracle(_lfv_params(p), r, v - 0.5 * dt * a0, a0)
    ol.setup()
    ml, nrl = ol.run_lfv(0, 60, n_macros=12)
    assert nrv > 0 and nrl > 0
    assert _match(ov.state()[0], ol.state()[0], p.domain[:]) < 1e-10
    # VV records steps 5..60, leap-frog steps 0..55
    for k in range(11):
        for key in ("kinenergy", "potenergy", "virial"):
            assert ml[k + 1][key] == pytest.approx(mv[k][key], rel=1e-10, abs=1e-12), (k, key)


def test_lfv_sliding_walls_equal_vv_in_the_lab_frame():
    """Tethered walls (teth bottom, teth_slide top): leap-frog moves a sliding wall molecule by
    dt*(v + slidev) with v its peculiar velocity and applies tether_force inside the move; the same
    positions come from velocity Verlet with the lab velocity v + slidev and the tethers applied
    with the forces."""
    from flowmol_amd._abi import FM_TAG_TETH_SLIDE
    p, r, v, tag, rt = config.config_couette(nunits=(6, 6, 6), tplot=5, nplanes=8)
    dt = p.delta_t
    ov = Oracle(p, r, v)
    ov.set_tags(tag, rt)
    ov.setup()
    ov.compute_forces(0)
    a0 = ov.state()[2]
    slide = tag == FM_TAG_TETH_SLIDE
    assert slide.any()
    # restart the VV run from the lab velocity
    vlab = v.copy()
    vlab[slide, 0] += p.wallslidev[0]
    ov = Oracle(p, r, vlab)
    ov.set_tags(tag, rt)
    ov.setup()
    ov.compute_forces(0)
    ov.run_vv(1, 40)
    ol = Oracle(_lfv_params(p), r, v - 0.5 * dt * a0, a0)
    ol.set_tags(tag, rt)
    ol.setup()
    ol.run_lfv(0, 40)
    assert _match(ov.state()[0], ol.state()[0], p.domain[:]) < 1e-9
    # the tether sites of the sliding wall moved by 40 dt wallslidev, the others stayed
    expect = rt.copy()
    expect[slide, 0] += 40 * dt * p.wallslidev[0]
    assert _match(ol.rtether(), expect, p.domain[:]) < 1e-12


def test_lfv_fixed_and_fixed_slide_tags():
    """fixed molecules stay put whatever the force on them; fixed_slide molecules move by
    dt*wallslidev per step (read_tag: slidev = wallslidev, no sign); free molecules feel them."""
    from flowmol_amd._abi import (FM_ENSEMBLE_TAG_MOVE, FM_TAG_FIXED, FM_TAG_FIXED_SLIDE, FM_TAG_FREE,
                                  FM_INTEGRATION_LFV)
    p = box_params(ensemble=FM_ENSEMBLE_TAG_MOVE, integration=FM_INTEGRATION_LFV, tplot=1000)
    p.wallslidev[0] = 0.7
    pos = np.array([[0.0, -5.0, 0.0], [1.0, -5.0, 0.0], [5.0, 5.0, 5.0], [-5.0, 4.0, -3.0]])
    vel = np.array([[0.3, 0.0, 0.0], [0.0, 0.0, 0.0], [0.1, 0.2, 0.3], [0.0, 0.0, 0.0]])
    tags = np.array([FM_TAG_FIXED, FM_TAG_FREE, FM_TAG_FREE, FM_TAG_FIXED_SLIDE], np.int32)
    o = Oracle(p, pos, vel)
    o.set_tags(tags)
    o.setup()
    n = 30
    o.run_lfv(0, n)
    r, v, a = o.state()
    assert np.array_equal(r[0], pos[0]) and np.array_equal(v[0], vel[0])
    assert r[1, 0] > pos[1, 0] + 0.01                    # pushed away by the fixed neighbour
    assert np.allclose(r[2], pos[2] + n * p.delta_t * vel[2], rtol=0, atol=1e-13)  # free flight
    assert np.allclose(r[3], pos[3] + n * p.delta_t * np.array([0.7, 0.0, 0.0]), rtol=0, atol=1e-13)
    # thermostatted tags are not on this path
    with pytest.raises(OracleError):
        o.set_tags(np.array([4, 0, 0, 0], np.int32))


# ---- sort_mols -----------------------------------------------------------------------------------
def _sort_setup(L, pos, flag, freq=1, bs=1):
    p = config.make_params((L, L, L), 0.5, 2.5, sort_flag=flag, sort_freq=freq, sortblocksize=bs, tplot=1000)
    return p, Oracle(p, pos, np.zeros_like(pos))


def test_sort_hilbert_curve_visits_every_block_through_faces():
    """build_hilbert on a 4^3 block grid (src/functions_lib.f90:2555-2604): a permutation of 1..64
    in which consecutive positions are face neighbours; on 5x3x4 blocks the curve restricted to
    the grid still numbers every block once."""
    pos = np.zeros((1, 3))
    for L, nb, adjacent in [(12.0, (4, 4, 4), True), (15.0, (5, 5, 5), False)]:
        p, o = _sort_setup(L, pos, 2)
        H = o.hcurve(nb)
        assert sorted(H.ravel()) == list(range(1, int(np.prod(nb)) + 1))
        if adjacent:
            where = {int(H[i, j, k]): np.array([i, j, k]) for i in range(nb[0]) for j in range(nb[1]) for k in range(nb[2])}
            steps = [np.abs(where[m + 1] - where[m]).sum() for m in range(1, int(np.prod(nb)))]
            assert all(s_ == 1 for s_ in steps)
    # orientation, by hand from hilbert3's octant formulas: on 2^3 blocks the points
    # 0.5*(+-0.5 ...) shifted by 0.25001 and scaled by 2 visit (2,2,2), (2,2,1), (1,2,1), (1,2,2),
    # (1,1,2), (1,1,1), (2,1,1), (2,1,2); on 4^3 the first octant 0.5*(0.5+zo, 0.5+xo, 0.5+yo)
    # starts at (4,4,4), (3,4,4), (3,3,4) and the last one ends at (4,1,4)
    p = config.make_params((8.4, 8.4, 8.4), 0.5, 2.5, sort_flag=2, sortblocksize=2, tplot=1000)
    H = Oracle(p, pos, np.zeros_like(pos)).hcurve((2, 2, 2))
    walk = [(1, 1, 1), (1, 1, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1), (0, 0, 0), (1, 0, 0), (1, 0, 1)]
    assert [int(H[b]) for b in walk] == list(range(1, 9))
    p, o = _sort_setup(12.0, pos, 2)
    H = o.hcurve((4, 4, 4))
    assert (H[3, 3, 3], H[2, 3, 3], H[2, 2, 3], H[3, 0, 3]) == (1, 2, 3, 64)
    # SORT_FLAG 1: blocks numbered x fastest
    p, o = _sort_setup(12.0, pos, 1)
    H = o.hcurve((4, 4, 4))
    i, j, k = np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij")
    assert np.array_equal(H, 1 + i + 4 * j + 16 * k)


@pytest.mark.parametrize("flag", [1, 2])
def test_sort_mols_renumbers_block_by_block(flag):
    """The new numbering is a stable sort of the old one by the curve position of each molecule's
    block; positions and velocities move with their molecules."""
    rng = np.random.default_rng(7)
    L = 12.0
    pos = rng.uniform(-L / 2, L / 2, (300, 3))
    vel = rng.normal(size=(300, 3))
    p = config.make_params((L, L, L), 0.5, 2.5, sort_flag=flag, sort_freq=1, sortblocksize=1, tplot=1000)
    o = Oracle(p, pos, vel)
    H = o.hcurve((4, 4, 4))
    b = np.ceil((pos + L / 2) / 3.0).astype(int) - 1
    key = H[b[:, 0], b[:, 1], b[:, 2]]
    order = np.argsort(key, kind="stable")
    o.sort_mols()
    r, v, _ = o.state()
    assert np.array_equal(r, pos[order]) and np.array_equal(v, vel[order])


def test_sort_mols_frequency_and_skipped_rounds():
    """Only every sort_freq-th call sorts (the saved counter starts at 1); a block holding more
    molecules than ceiling(int(np/blocks)*sf1 + sf2) skips the round and raises the saved safety
    factors by 0.1 and 5, so a crowded block is sorted once they allow it."""
    rng = np.random.default_rng(3)
    L = 9.0
    pos = rng.uniform(-L / 2, L / 2, (60, 3))
    p, o = _sort_setup(L, pos, 1, freq=3)
    for call in range(1, 7):
        before = o.state()[0]
        o.sort_mols()
        moved = not np.array_equal(o.state()[0], before)
        assert moved == (call == 3), call   # the 6th call finds the order already sorted
    # 40 of 41 molecules in one of 27 blocks: int(41/27) = 1, so the estimate is ceil(1.5 + 5) = 7,
    # then ceil(1.6 + 10) = 12, 17, 22, 27, 32, 38, and ceil(2.2 + 40) = 43 >= 40 at the eighth call
    crowd = np.concatenate([rng.uniform(-4.4, -1.6, (40, 3)), rng.uniform(1.6, 4.4, (1, 3))])
    crowd = crowd[::-1].copy()
    p, o = _sort_setup(L, crowd, 1, freq=1)
    for call in range(1, 9):
        o.sort_mols()
        moved = not np.array_equal(o.state()[0], crowd)
        assert moved == (call == 8), call


# ---- MOP -----------------------------------------------------------------------------------------
def test_mop_single_crossing():
    L = 12.0
    p = config.make_params((L, L, L), 0.5, 2.5, vflux_outflag=2, nplanes=4, planespacing=3.0, tplot=1)
    # planes at 3*(n-1) + shift - L/2 with shift = 0.5*(L - 3*3) = 1.5: -4.5, -1.5, 1.5, 4.5
    pos = np.array([[0.2, -2.0, 0.1], [0.5, -1.0, 0.3]])   # crosses y = -1.5 only
    o = Oracle(p, pos, np.zeros((2, 3)))
    assert np.allclose(o.planes(), [-4.5, -1.5, 1.5, 4.5])
    o.setup()
    o.compute_forces(0)
    P = o.mop()
    _, _, a = o.state()
    rij = pos[0] - pos[1]
    acc = a[0, 0] / rij[0]
    expect = np.zeros((4, 3))
    expect[1] = acc * rij * (-1.0 - 1.0)   # sign(yi-p) - sign(yj-p) = -2
    assert np.allclose(P, expect, rtol=1e-14, atol=1e-15)


def test_mop_counts_each_crossing_pair_once():
    """Every pair crossing a plane contributes once: the MOP sum equals a direct all-pairs sum.
    The cloud spans the periodic y direction but stays clear of the x and z boundaries; a pair
    whose segment leaves the box through both x (or z) and y has its y crossing evaluated in one
    end's frame only, where the intersection can fall outside the box, and the reference then
    drops it (src/simulation_record.f90:7098-7104)."""
    L = 20.0
    p = config.make_params((L, L, L), 0.5, 2.5, vflux_outflag=2, nplanes=7, planespacing=2.7, tplot=1)
    rng = np.random.default_rng(12)
    pts = []
    while len(pts) < 300:
        q = np.array([rng.uniform(-6, 6), rng.uniform(-L / 2, L / 2), rng.uniform(-6, 6)])
        if all(np.sum(((q - w + [0, L / 2, 0]) % [1e9, L, 1e9] - [0, L / 2, 0]) ** 2) > 0.81 for w in pts):
            pts.append(q)
    r = np.array(pts)
    o = Oracle(p, r, np.zeros_like(r))
    o.setup()
    o.compute_forces(0)
    P = o.mop()
    planes = o.planes()
    direct = np.zeros_like(P)
    for i in range(r.shape[0]):
        d = r[i] - r[i + 1:]
        d[:, 1] -= L * np.round(d[:, 1] / L)
        r2 = (d * d).sum(1)
        for k in np.nonzero(r2 < p.rcutoff ** 2)[0]:
            rij = d[k]
            inv = 1.0 / r2[k]
            acc = 48 * (inv ** 7 - 0.5 * inv ** 4)
            yi, yj = r[i, 1], r[i, 1] - rij[1]
            for q, pl in enumerate(planes):
                for pp in (pl - L, pl, pl + L):   # periodic copies of the plane
                    if (yi - pp) * (yj - pp) < 0:
                        direct[q] += acc * rij * 2 * np.sign(yi - pp)
    assert np.abs(P).max() > 1.0
    assert np.allclose(P, direct, rtol=1e-10, atol=1e-10)


# ---- macros --------------------------------------------------------------------------------------
def _cv_setup(L, pos, binspercell=1.0, rc=2.5):
    p = config.make_params((L, L, L), 0.5, rc, vflux_outflag=4, tplot=1, binspercell=binspercell)
    o = Oracle(p, pos, np.zeros_like(pos))
    o.setup()
    o.compute_forces(0)
    return p, o


def _lj_acc(d):
    inv = 1.0 / float(np.dot(d, d))
    return 48.0 * (inv ** 7 - 0.5 * inv ** 4)


def test_cv_bins_follow_the_cell_ratio():
    """nbins = int(binspercell*ncells), nhb = ceiling(nbins/ncells), binsize = domain/nbins
    (src/setup_set_parameters.f90:1990-2011)."""
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    for ratio, nb, nh in [(1.0, 3, 1), (2.0, 6, 2), (1.5, 4, 2), (0.5, 1, 1)]:
        _, o = _cv_setup(9.0, pos, binspercell=ratio)
        nbins, nhb, bs = o.cv_bins()
        assert list(nbins) == [nb] * 3 and list(nhb) == [nh] * 3
        assert np.allclose(bs, 9.0 / nb)


def test_cv_single_interior_crossing():
    """Two molecules straddling the bin plane x = -1.5 (3 bins of 3.0, one halo bin per side):
    face x of the bin above the plane gets crossdir * 2 * (2 acc rij) with crossdir = sign(rij_x),
    face x+3 of the bin below holds the same running total, every other entry stays 0."""
    pos = np.array([[-1.9, 0.2, 0.1], [-0.8, 0.2, 0.1]])
    _, o = _cv_setup(9.0, pos)
    F = o.cv_stress()
    assert F.shape == (5, 5, 5, 3, 6)
    d = pos[0] - pos[1]
    expect = 4.0 * _lj_acc(d) * d * np.sign(d[0])
    # bins (1-based with halo) x: ceil((-1.5+eps+4.5)/3)+1 = 3; y, z: ceil(4.7/3)+1 = 3
    assert np.allclose(F[2, 2, 2, :, 0], expect, rtol=1e-13, atol=0)
    assert np.allclose(F[1, 2, 2, :, 3], expect, rtol=1e-13, atol=0)
    G = F.copy()
    G[2, 2, 2, :, 0] = 0
    G[1, 2, 2, :, 3] = 0
    assert not G.any()


def test_cv_periodic_pair_is_met_once_per_end():
    """A pair across the x boundary is in the half list twice, against either image; each entry
    carries fij = acc rij (halo pair) and crosses the boundary plane into a halo bin: x = +4.5
    (bin 5, copy in bin 4) and x = -4.5 (bin 2, copy in bin 1)."""
    pos = np.array([[4.0, 0.3, -0.2], [-3.9, 0.3, -0.2]])
    _, o = _cv_setup(9.0, pos)
    F = o.cv_stress()
    d = pos[0] - (pos[1] + [9.0, 0, 0])
    expect = 2.0 * _lj_acc(d) * d * np.sign(d[0])
    assert np.allclose(F[4, 2, 2, :, 0], expect, rtol=1e-12, atol=0)
    assert np.allclose(F[3, 2, 2, :, 3], expect, rtol=1e-12, atol=0)
    assert np.allclose(F[1, 2, 2, :, 0], expect, rtol=1e-12, atol=0)
    assert np.allclose(F[0, 2, 2, :, 3], expect, rtol=1e-12, atol=0)
    G = F.copy()
    for b, f in [(4, 0), (3, 3), (1, 0), (0, 3)]:
        G[b, 2, 2, :, f] = 0
    assert not G.any()


def cv_bruteforce_folded(r, L, rc, nbins, acc_of):
    """Independent statement of the CV surface stress: every pair within rc (minimum image) adds
    4 acc rij sign(rij_n) to face n of the bin just above each plane of bins normal to n that its
    segment crosses, with the bins folded into the periodic box."""
    L = np.asarray(L, float)
    nbins = np.asarray(nbins)
    hd, bs = 0.5 * L, L / nbins
    F = np.zeros(tuple(nbins) + (3, 3))
    N = r.shape[0]
    iu, ju = np.triu_indices(N, 1)
    d = r[iu] - r[ju]
    d -= L * np.round(d / L)
    r2 = (d * d).sum(1)
    m = r2 < rc * rc
    i, d, r2 = iu[m], d[m], r2[m]
    acc = acc_of(r2)
    ri = r[i]
    rj = ri - d
    q = 4.0 * acc[:, None] * d
    b1 = np.ceil((ri + hd) / bs).astype(int)
    b2 = np.ceil((rj + hd) / bs).astype(int)
    for n in range(3):
        lo = np.minimum(b1[:, n], b2[:, n])
        hi = np.maximum(b1[:, n], b2[:, n])
        cross = np.sign(d[:, n])
        for k in range(int((hi - lo).max(initial=0))):
            sel = lo + k < hi
            plane = (lo[sel] + k) * bs[n] - hd[n]
            t = (plane - ri[sel, n]) / (rj[sel, n] - ri[sel, n])
            pc = ri[sel] + t[:, None] * (rj[sel] - ri[sel])
            cb = np.ceil((pc + hd) / bs).astype(int)
            cb[:, n] = lo[sel] + k + 1
            cb = (cb - 1) % nbins
            for c in range(3):
                np.add.at(F[..., c, n], (cb[:, 0], cb[:, 1], cb[:, 2]), cross[sel] * q[sel, c])
    return F


def fold_cv(F, nbins, nhb):
    """Faces 1-3 of Pxyface with the halo bins added onto their periodic images."""
    out = np.zeros(tuple(nbins) + (3, 3))
    nbo = np.asarray(nbins) + 2 * np.asarray(nhb)
    idx = [(np.arange(nbo[d]) - nhb[d]) % nbins[d] for d in range(3)]
    for a in range(nbo[0]):
        for b in range(nbo[1]):
            for c in range(nbo[2]):
                out[idx[0][a], idx[1][b], idx[2][c]] += F[a, b, c, :, :3]
    return out


@pytest.mark.parametrize("ratio", [1.0, 2.0])
def test_cv_equals_independent_bruteforce(ratio):
    """Equilibrated 2048-molecule LJ liquid (rc 2.5): the folded CV tally equals the all-pairs
    statement above; faces 4-6 are faces 1-3 moved down one bin along their normal."""
    from tests.helpers import equilibrated
    p, r, v, _ = equilibrated(config.config_fcc, 200, nunits=(8, 8, 8), density=0.8, rcutoff=2.5, seed=21)
    L = np.array(p.domain[:])
    p2 = config.make_params(L, 0.8, 2.5, vflux_outflag=4, tplot=1, binspercell=ratio)
    o = Oracle(p2, r, v)
    o.setup()
    o.compute_forces(0)
    F = o.cv_stress()
    nbins, nhb, _ = o.cv_bins()
    got = fold_cv(F, nbins, nhb)
    ref = cv_bruteforce_folded(r, L, 2.5, nbins, lambda r2: 48.0 * ((1 / r2) ** 7 - 0.5 * (1 / r2) ** 4))
    scale = np.abs(ref).max()
    assert scale > 0
    assert np.abs(got - ref).max() <= 1e-9 * scale
    for n in range(3):
        src = np.take(F[..., n], np.arange(1, F.shape[n]), axis=n)
        dst = np.take(F[..., n + 3], np.arange(0, F.shape[n] - 1), axis=n)
        assert np.array_equal(src, dst)
        assert not np.take(F[..., n + 3], [F.shape[n] - 1], axis=n).any()


def test_macros_kinetic_and_temperature():
    p, r, v = config.config_fcc((3, 3, 3), 0.8, 2 ** (1 / 6), seed=2)
    o = Oracle(p, r, v)
    o.setup()
    o.compute_forces(0)
    m = o.macros()
    assert m["temperature"] == pytest.approx(1.0, rel=1e-12)    # seeded velocities are scaled to T=1
    assert m["kinenergy"] == pytest.approx(1.5, rel=1e-12)
    assert abs(m["vsum"]) < 1e-10                               # zero net momentum


# ---- restart fixture written by the reference ---------------------------------------------------
def test_final_state_fixture_header_and_records():
    d = read_final_state(GOLDEN)
    h = d["header"]
    assert h["globalnp"] == 2635 and tuple(h["initialnunits"]) == (30, 8, 8)
    assert h["iter"] == 100000 and h["tplot"] == 25 and tuple(h["procnp"]) == (1318, 1317)
    assert h["globaldomain"][0] == pytest.approx(47.622, abs=1e-3)
    assert h["globaldomain"][1] == pytest.approx(12.699, abs=1e-3)
    assert (h["density"], h["rcutoff"], h["delta_t"], h["delta_rneighbr"]) == (1.0, 2.5, 0.005, 0.3)
    L = np.array(h["globaldomain"])
    assert (np.abs(d["r"]) <= 0.5 * L).all()


def test_final_state_roundtrip(tmp_path):
    p, r, v = config.config_fcc((3, 3, 3), 0.8, 2.5, seed=1)
    f = tmp_path / "fs"
    write_final_state(str(f), r, v, nunits=(3, 3, 3), density=0.8, rcutoff=2.5, delta_t=0.005,
                      delta_rneighbr=0.3, domain=p.domain[:])
    d = read_final_state(str(f))
    assert np.array_equal(d["r"], r) and np.array_equal(d["v"], v)


def test_fixture_liquid_vapour_lists_and_forces():
    d = read_final_state(GOLDEN)
    h = d["header"]
    p = config.make_params(h["globaldomain"], h["density"], h["rcutoff"], delta=h["delta_rneighbr"],
                           dt=h["delta_t"], tplot=h["tplot"])
    o = Oracle(p, d["r"], d["v"])
    o.setup()
    i, j, s = o.neighbourlist()
    bi, bj, bs = o.bruteforce_pairs()
    assert np.array_equal(canonical_pairs(i, j, s), canonical_pairs(bi, bj, bs))
    o.compute_forces(0)
    a1 = o.state()[2]
    scale = o.force_scale()
    o.compute_forces_allpairs(0)
    a2 = o.state()[2]
    assert (np.abs(a1 - a2).max(axis=1) / np.maximum(scale, 1e-300)).max() < 1e-13


def test_virial_tensor_trace_is_virial():
    p, o = _disordered((5, 5, 5), seed=9, steps=50, pressure_outflag=1)
    o.compute_forces(0)
    rf = o.rfmol()
    _, vir = o.molecular()
    assert np.allclose(np.trace(rf, axis1=1, axis2=2), vir, rtol=1e-12, atol=1e-10)
    assert np.allclose(rf, np.transpose(rf, (0, 2, 1)), rtol=1e-14, atol=1e-14)


def test_polymer_config_bonds_are_nearest_neighbours():
    p, r, v, mt, mono = config.config_polymer(nunits=(6, 8, 6))
    o = Oracle(p, r, v)
    o.set_moltype(mt)
    o.set_monomers(*mono)
    o.setup()
    b, s = o.bonds()
    nb = (b >= 0).sum(axis=1)
    assert np.array_equal(nb, mono[2])           # each bead finds exactly its funcy partners
    for i in range(0, r.shape[0], 97):
        for j in b[i][b[i] >= 0]:
            assert abs(abs(j - i) - 1) == 0        # partners are the chain neighbours
