"""Parity of the HIP path (through the C-ABI) against the CPU restatement of the reference.

Bars: cell membership and neighbour sets bit-exact; forces, per-molecule potential energy and
virial within 1e-12 relative (scale-aware for net forces: |da_i|_inf / sum_j |f_ij|); positions
and velocities of the routine-by-routine drop-in sequence bit-exact where the arithmetic is
exact (wrap, renumbering) and within 1e-10 absolute over short trajectories.
"""
import sys

import numpy as np
import pytest

from flowmol_amd import config
from flowmol_amd.md import MDContext
from oracle.oracle import Oracle
from tests.helpers import canonical_pairs, equilibrated, full_from_half, rel_force_err

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c1_state():
    return equilibrated(config.config_c1, 300)


def _pair(p, r, v, a=None):
    o = Oracle(p, r, v, a)
    g = MDContext(p, r, v, a)
    return o, g


def test_cells_and_neighbourlist_c1(c1_state):
    p, r, v, a = c1_state
    say("case: cells+nlist C1 2048 WCA")
    o, g = _pair(p, r, v, a)
    o.assign_to_cell()
    g.assign_to_cell()
    assert np.array_equal(g.cells(), o.cells())
    o.updateborders(1)
    o.assign_to_neighbourlist()
    g.messenger_updateborders(1)
    g.assign_to_neighbourlist()
    oi, oj, os_ = o.neighbourlist()
    gi, gj, gs = g.neighbourlist()
    assert np.array_equal(canonical_pairs(oi, oj, os_), canonical_pairs(gi, gj, gs))
    # the GPU full list is exactly the reference half list with unshifted entries mirrored
    fi, fj, fs = full_from_half(oi, oj, os_)
    ko = np.sort((fi.astype(np.int64) << 32) | (fj.astype(np.int64) << 5) | fs)
    kg = np.sort((gi.astype(np.int64) << 32) | (gj.astype(np.int64) << 5) | gs)
    assert np.array_equal(ko, kg)
    g.close()


@pytest.mark.parametrize("n,jitter", [(12, 0.0), (12, 1e-15), (12, 1e-12), (5, 0.0), (5, 1e-12)])
def test_neighbourlist_pairs_on_the_threshold(n, jitter):
    """A simple cubic lattice whose face diagonals are rneighbr apart (2 a^2 = rneighbr2): twelve
    pairs per molecule sit on the list threshold, where the FP64 rounding of rij2 (and of the
    halo images) decides membership.  The list build's quantised prefilter leaves every such pair
    to the FP64 test, so the set must still be the reference's bit for bit.  n = 5 gives three
    cells per side, where a brick's region holds each cell twice, under two image shifts."""
    rc, delta = 2.5, 0.3
    rn2 = (rc + delta) ** 2
    a_lat = np.sqrt(rn2 / 2.0)
    L = n * a_lat
    g1 = (np.arange(n) + 0.5) * a_lat - 0.5 * L
    r = np.stack(np.meshgrid(g1, g1, g1, indexing="ij"), axis=-1).reshape(-1, 3).copy()
    r += np.random.default_rng(21).uniform(-jitter, jitter, r.shape) * L
    v = np.zeros_like(r)
    p = config.make_params((L, L, L), r.shape[0] / L ** 3, rc, delta=delta, tplot=10)
    say(f"case: {r.shape[0]} molecules on the list threshold, jitter {jitter}")
    o, g = _pair(p, r, v)
    o.assign_to_cell()
    o.updateborders(1)
    o.assign_to_neighbourlist()
    g.assign_to_cell()
    g.messenger_updateborders(1)
    g.assign_to_neighbourlist()
    oi, oj, os_ = o.neighbourlist()
    gi, gj, gs = g.neighbourlist()
    # three face pairs per molecule are always in, the six face diagonals per molecule split
    # by the rounding (all twelve in would give 9N entries and more with the boundary images)
    assert 3 * r.shape[0] < len(canonical_pairs(oi, oj, os_)) < 9 * r.shape[0]
    assert np.array_equal(canonical_pairs(oi, oj, os_), canonical_pairs(gi, gj, gs))
    g.close()


def test_neighbourlist_after_moves_without_binning():
    """assign_to_neighbourlist on positions moved since the last assign_to_cell: the cells keep
    their old members (molecules may sit a little outside them, and no longer in x order), so
    the build takes the FP64 kernel without x windows; the set is still the reference's for the
    same stale cells, and the forces over it agree."""
    p, r, v, a = _lj_liquid((6, 6, 6), seed=11, steps=100, tplot=10)
    say("case: list after three pass-1 moves without binning, 864 LJ")
    o, g = _pair(p, r, v, a)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    for _ in range(3):
        o.move_particles_vv(1)
        o.updateborders(0)
        g.simulation_move_particles_vv(1)
        g.messenger_updateborders(0)
    o.assign_to_neighbourlist()
    g.assign_to_neighbourlist()
    oi, oj, os_ = o.neighbourlist()
    gi, gj, gs = g.neighbourlist()
    assert np.array_equal(canonical_pairs(oi, oj, os_), canonical_pairs(gi, gj, gs))
    o.compute_forces(3)
    g.simulation_compute_forces(3)
    assert rel_force_err(g.state()[2], o.state()[2], o.force_scale()) < FORCE_TOL
    g.close()


def test_forces_c1(c1_state):
    p, r, v, a = c1_state
    say("case: forces C1 2048 WCA")
    o, g = _pair(p, r, v, a)
    o.setup()
    g.setup(0)
    o.compute_forces(0)
    _, _, ao = o.state()
    _, _, ag = g.state()
    scale = o.force_scale()
    assert rel_force_err(ag, ao, scale) < FORCE_TOL
    peo, viro = o.molecular()
    peg, virg = g.molecular()
    np.testing.assert_allclose(peg, peo, rtol=FORCE_TOL, atol=1e-14)
    np.testing.assert_allclose(virg, viro, rtol=FORCE_TOL, atol=1e-12)
    mo = o.macros()
    mg = g.evaluate_macroscopic_properties()
    for k in ("potenergy", "kinenergy", "pressure", "virial", "temperature"):
        assert abs(mg[k] - mo[k]) <= 1e-12 * max(1.0, abs(mo[k])), (k, mg[k], mo[k])
    # a step without tallies (short, partly filled list words)
    o.compute_forces(3)
    g.simulation_compute_forces(3)
    assert rel_force_err(g.state()[2], o.state()[2], scale) < FORCE_TOL
    g.close()


def _lj_liquid(nunits, seed, steps=200, **kw):
    p, r, v = config.config_fcc(nunits, 0.8, 2.5, seed=seed, **kw)
    # lattice planes can coincide with cell faces (e.g. 7 units, 4 cells): the reference's halo
    # binning then indexes past its halo layer, so start from a slightly perturbed lattice
    r = r + np.random.default_rng(seed).uniform(-1e-3, 1e-3, r.shape)
    o = Oracle(p, r, v)
    o.setup()
    o.compute_forces(0)
    o.run_vv(1, steps)
    o.rebuild()
    r2, v2, a2 = o.state()
    return p, r2, v2, a2


def test_forces_lj_rc25():
    p, r, v, a = _lj_liquid((6, 6, 6), seed=7, tplot=10)
    say("case: forces LJ rc2.5 864")
    o, g = _pair(p, r, v, a)
    o.setup()
    g.setup(0)
    o.compute_forces(0)
    _, _, ao = o.state()
    _, _, ag = g.state()
    assert rel_force_err(ag, ao, o.force_scale()) < FORCE_TOL
    # independent check: the reference's all-pairs minimum-image routine
    o.compute_forces_allpairs(0)
    _, _, aap = o.state()
    assert rel_force_err(ag, aap, o.force_scale()) < FORCE_TOL
    peo, viro = o.molecular()
    peg, virg = g.molecular()
    np.testing.assert_allclose(peg, peo, rtol=FORCE_TOL, atol=1e-13)
    np.testing.assert_allclose(virg, viro, rtol=FORCE_TOL, atol=1e-11)
    # a step without tallies
    o.compute_forces(3)
    g.simulation_compute_forces(3)
    assert rel_force_err(g.state()[2], o.state()[2], o.force_scale()) < FORCE_TOL
    g.close()


@pytest.mark.parametrize("nunits,density,rc", [((8, 8, 8), 0.8, 4.0), ((8, 8, 8), 1.1, 3.0), ((6, 6, 6), 0.3, 2.5)])
def test_other_cutoffs_and_densities(nunits, density, rc):
    """Away from the rc 2.5, density 0.8 liquid: a 4.0 cutoff (3 cells of 4.56 a side, about 76
    molecules per cell, more than one wavefront sorts per cell, and about 350 full-list entries
    per molecule), a dense 1.1 state at rc 3.0 (the same cell count) and a 0.3 gas (5 cells of 2.83).
    The neighbour set is the reference's, forces, PE and virial match on a record step and on a
    step without tallies, 60 routine-by-routine steps take the oracle's rebuild steps and stay
    within 1e-10, and the fused loop equals the routine sequence bit for bit."""
    p, r, v = config.config_fcc(nunits, density, rc, seed=17, tplot=10)
    r = r + np.random.default_rng(17).uniform(-1e-3, 1e-3, r.shape)
    o = Oracle(p, r, v)
    o.setup()
    o.compute_forces(0)
    o.run_vv(1, 100)
    o.rebuild()
    r, v, a = o.state()
    say(f"case: LJ {r.shape[0]} density {density} rc {rc}")
    o, g = _pair(p, r, v, a)
    o.setup()
    g.setup(0)
    assert np.array_equal(g.cells(), o.cells())
    oi, oj, os_ = o.neighbourlist()
    gi, gj, gs = g.neighbourlist()
    assert np.array_equal(canonical_pairs(oi, oj, os_), canonical_pairs(gi, gj, gs))
    for it in (0, 3):  # record step, and a step without tallies
        o.compute_forces(it)
        g.simulation_compute_forces(it)
        assert rel_force_err(g.state()[2], o.state()[2], o.force_scale()) < FORCE_TOL
        if it == 0:
            peo, viro = o.molecular()
            peg, virg = g.molecular()
            np.testing.assert_allclose(peg, peo, rtol=FORCE_TOL, atol=1e-13)
            np.testing.assert_allclose(virg, viro, rtol=FORCE_TOL, atol=1e-12)
    g.close()
    o, g = _pair(p, r, v, a)
    f = MDContext(p, r, v, a)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    f.setup(0)
    nrb = 0
    for it in range(1, 61):
        o.move_particles_vv(1)
        o.updateborders(0)
        rbo = o.checkrebuild(it)
        if rbo:
            o.rebuild()
        o.compute_forces(it)
        o.move_particles_vv(2)
        rbg = g.md_advance_vv(it)
        assert rbo == rbg, f"rebuild decision differs at iter {it}"
        nrb += bool(rbg)
    assert nrb >= 6  # before each record step, and on the gas's displacement checks
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-10
    assert np.abs(g.state()[1] - o.state()[1]).max() < 1e-9
    _, nrf = f.run_vv(1, 60)
    assert nrf == nrb
    for x, y in zip(f.state(), g.state()):
        assert np.array_equal(x, y)
    f.close()
    g.close()


def test_sendmols_renumbering_bit_exact():
    """Molecules pushed outside the box are wrapped and renumbered exactly as the single-rank
    sendmols does (leavers of each axis/direction pass get their numbers reversed)."""
    p, r, v = config.config_fcc((5, 5, 5), 0.8, 2.5, seed=3)
    rng = np.random.default_rng(11)
    L = np.array(p.domain[:])
    r = r + rng.uniform(-0.15, 0.15, r.shape)
    # push a few molecules out through faces, edges and corners
    idx = rng.choice(r.shape[0], 60, replace=False)
    r[idx] += rng.choice([-1, 0, 1], (60, 3)) * (0.5 * L + 0.01)
    say("case: sendmols renumbering 500")
    o, g = _pair(p, r, v)
    o.sendmols()
    g.sendmols()
    ro, vo, _ = o.state()
    rg, vg, _ = g.state()
    assert np.array_equal(ro, rg)
    assert np.array_equal(vo, vg)
    assert (np.abs(rg) <= 0.5 * L).all()
    g.close()


def test_drop_in_sequence_short_trajectory(c1_state):
    """md_advance_vv routine by routine on both sides: identical rebuild steps, states close."""
    p, r, v, a = c1_state
    say("case: 100-step drop-in trajectory C1")
    o, g = _pair(p, r, v, a)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    for it in range(1, 101):
        o.move_particles_vv(1)
        o.updateborders(0)
        rbo = o.checkrebuild(it)
        if rbo:
            o.rebuild()
        o.compute_forces(it)
        o.move_particles_vv(2)
        rbg = g.md_advance_vv(it)
        assert rbo == rbg, f"rebuild decision differs at iter {it}"
    ro, vo, ao = o.state()
    rg, vg, ag = g.state()
    assert np.abs(rg - ro).max() < 1e-10
    assert np.abs(vg - vo).max() < 1e-9
    g.close()


def test_fused_loop_equals_routine_sequence(c1_state):
    p, r, v, a = c1_state
    say("case: fused run_vv vs routine sequence C1")
    g1 = MDContext(p, r, v, a)
    g2 = MDContext(p, r, v, a)
    g1.setup(0)
    g2.setup(0)
    for it in range(1, 51):
        g1.md_advance_vv(it)
    g2.run_vv(1, 50)
    r1, v1, a1 = g1.state()
    r2, v2, a2 = g2.state()
    assert np.array_equal(r1, r2) and np.array_equal(v1, v2) and np.array_equal(a1, a2)
    g1.close()
    g2.close()


@pytest.mark.parametrize("tplot", [7, 10])
def test_fused_loop_frequent_rebuilds_equals_routine_sequence(tplot):
    """A thin skin (delta 0.1) makes the displacement check fire every few steps, at every
    position in the queued windows, including the pass 1 carried by a window's last force launch;
    the forced rebuilds before record steps (TPLOT 7 and 10) open other windows.  300 steps of the
    fused loop equal the routine-by-routine sequence bit for bit, with the same rebuild steps."""
    p, r, v = config.config_fcc((6, 6, 6), 0.8, 2.5, seed=91, tplot=tplot, delta=0.1)
    r = r + np.random.default_rng(91).uniform(-0.01, 0.01, r.shape)
    say("case: fused loop vs routine sequence, delta 0.1, TPLOT %d, 864 LJ, 300 steps" % tplot)
    g = MDContext(p, r, v)
    f = MDContext(p, r, v)
    g.setup(0)
    f.setup(0)
    rg = [it for it in range(1, 301) if g.md_advance_vv(it)]
    _, nrb = f.run_vv(1, 300)
    assert nrb == len(rg) and len(rg) > 60
    for x, y in zip(f.state(), g.state()):
        assert np.array_equal(x, y)
    f.close()
    g.close()


def test_energy_conservation_c1():
    p, r, v = config.config_c1()
    say("case: NVE drift C1 1000 steps")
    g = MDContext(p, r, v)
    g.setup(0)
    ms, nrb = g.run_vv(1, 1000, n_macros=10)
    e = np.array([m["totenergy"] for m in ms])
    assert nrb > 0
    assert np.abs(e - e[0]).max() < 1e-3 * abs(e[0])
    g.close()


@pytest.mark.parametrize("brick", [(1, 1, 1), (2, 3, 1), (3, 3, 3), (4, 2, 5)])
def test_brick_shapes_same_sets_and_forces(brick):
    """The LDS brick decomposition (including partial bricks and regions that wrap onto the same
    cell twice) gives the reference's neighbour set and forces for any brick shape."""
    p, r, v, a = _lj_liquid((7, 7, 7), seed=9, steps=100, tplot=10)
    p.brick[0], p.brick[1], p.brick[2] = brick
    say(f"case: brick {brick} LJ 1372")
    o, g = _pair(p, r, v, a)
    o.setup()
    g.setup(0)
    oi, oj, os_ = o.neighbourlist()
    gi, gj, gs = g.neighbourlist()
    assert np.array_equal(canonical_pairs(oi, oj, os_), canonical_pairs(gi, gj, gs))
    for it in (0, 3):  # record step, and a step without tallies
        o.compute_forces(it)
        g.simulation_compute_forces(it)
        assert rel_force_err(g.state()[2], o.state()[2], o.force_scale()) < FORCE_TOL
    g.close()


@pytest.mark.skip(reason="in-place brick retuning of a live context (MDContext.autotune_brick) ended in a "
                  "host segmentation fault on its first GPU run and no cause was found by reading the code; "
                  "it is switched off and shapes are tuned on fresh contexts (tune_brick)")
def test_autotuned_brick_keeps_sets_and_forces():
    """The brick-shape autotuner only changes how cells are grouped: the neighbour set, the forces
    and the trajectory stay the reference's."""
    p, r, v, a = _lj_liquid((8, 8, 8), seed=13, steps=100, tplot=10)
    say("case: autotuned brick, 2048 LJ")
    o, g = _pair(p, r, v, a)
    o.setup()
    g.setup(0)
    shape = g.autotune_brick(0)
    assert all(s >= 1 for s in shape)
    oi, oj, os_ = o.neighbourlist()
    gi, gj, gs = g.neighbourlist()
    assert np.array_equal(canonical_pairs(oi, oj, os_), canonical_pairs(gi, gj, gs))
    o.compute_forces(0)
    assert rel_force_err(g.state()[2], o.state()[2], o.force_scale()) < FORCE_TOL
    for it in range(1, 31):
        o.move_particles_vv(1)
        o.updateborders(0)
        if o.checkrebuild(it):
            o.rebuild()
        o.compute_forces(it)
        o.move_particles_vv(2)
        g.md_advance_vv(it)
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-10
    g.close()


def test_tuned_brick_keeps_sets_and_forces():
    """The brick-shape tuner times candidate shapes on fresh contexts; the shape it returns gives
    the reference's neighbour set, forces and trajectory."""
    from flowmol_amd.md import brick_candidates, tune_brick

    p, r, v, a = _lj_liquid((8, 8, 8), seed=13, steps=100, tplot=10)
    say("case: tuned brick, 2048 LJ")
    shape, times = tune_brick(p, r, v, a, steps=10, warmup=2, max_candidates=3)
    assert len(times) == 3 and shape in times and all(t > 0 for t in times.values())
    assert tuple(p.brick) == (0, 0, 0)  # the caller's parameters are left alone
    p.brick[0], p.brick[1], p.brick[2] = shape
    say(f"case: tuned brick {shape}")
    o, g = _pair(p, r, v, a)
    o.setup()
    g.setup(0)
    assert shape in brick_candidates(g.geometry()["ncells"], r.shape[0], 3)
    oi, oj, os_ = o.neighbourlist()
    gi, gj, gs = g.neighbourlist()
    assert np.array_equal(canonical_pairs(oi, oj, os_), canonical_pairs(gi, gj, gs))
    o.compute_forces(0)
    assert rel_force_err(g.state()[2], o.state()[2], o.force_scale()) < FORCE_TOL
    for it in range(1, 31):
        o.move_particles_vv(1)
        o.updateborders(0)
        if o.checkrebuild(it):
            o.rebuild()
        o.compute_forces(it)
        o.move_particles_vv(2)
        g.md_advance_vv(it)
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-10
    g.close()


@pytest.mark.parametrize("integ", ["vv", "lfv"])
def test_read_once_lists_keep_every_sum(integ):
    """With TPLOT 1 every step rebuilds, so inside a fused loop each list is read by one force
    call at the positions it was built from, and the loop keeps only the pairs inside rc.  After
    30 steps positions, velocities and accelerations, and every record step's sums, are
    bit-identical to the routine-by-routine sequence (full lists); the list the loop leaves behind
    is the full list."""
    from flowmol_amd._abi import FM_INTEGRATION_LFV

    p, r, v = config.config_fcc((8, 8, 8), 0.8, 2.5, seed=31, tplot=1)
    r = r + np.random.default_rng(31).uniform(-0.05, 0.05, r.shape)
    if integ == "lfv":
        p.integration_algorithm = FM_INTEGRATION_LFV
    say(f"case: read-once lists, 2048 LJ, {integ}, TPLOT 1, 30 steps")
    g1 = MDContext(p, r, v)
    g2 = MDContext(p, r, v)
    g1.setup(0)
    g2.setup(0)
    m1 = []
    if integ == "vv":
        for it in range(1, 31):
            g1.md_advance_vv(it)
            m1.append(g1.evaluate_macroscopic_properties())
        m2, nrb = g2.run_vv(1, 30, n_macros=30)
    else:
        for it in range(0, 30):
            m1.append(g1.md_advance_lfv(it, macros=True)[1])
        m2, nrb = g2.run_lfv(0, 30, n_macros=30)
    assert nrb == 30
    for x, y in zip(g1.state(), g2.state()):
        assert np.array_equal(x, y)
    assert m1 == m2
    for x, y in zip(g1.neighbourlist(), g2.neighbourlist()):
        assert np.array_equal(x, y)
    g1.close()
    g2.close()
