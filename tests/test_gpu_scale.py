"""Parity at the headline size, and the fused loop on grids with far more bricks than the device
holds resident workgroups.

The 1,048,576-molecule LJ case (FCC 64^3, density 0.8, rc 2.5, delta 0.3, dt 0.005, TPLOT 10) is
the timed workload: 39^3 cells, 13^3 bricks of 3x3x3 cells, regions of about 2,200 molecules.
Here it is checked against the oracle at its full size:
  * cell membership and the neighbour set bit-exact;
  * forces on a record step (PE, virial) and on a step without tallies within 1e-12 scale-aware
    (|da_i|_inf / sum_j |f_ij|), per-molecule PE and virial within 1e-12 relative;
  * a short drop-in trajectory through a forced rebuild with the same rebuild steps;
  * the fused loop (force launches that carry the next step's pass 1 and displacement check)
    bit-identical to the routine-by-routine sequence over a window whose check fires mid-window,
    its record-step sums (reduced on the device inside the windows) equal to those of the
    routine-by-routine sequence.
The input is the seeded lattice disordered by 200 fused steps on the device, ended by the
reference's rebuild sequence (sendmols wraps every molecule back into the box); both sides are
then fed the same positions and velocities.
"""
import sys

import numpy as np
import pytest

from flowmol_amd import config
from flowmol_amd.md import MDContext
from oracle.oracle import Oracle
from tests.helpers import canonical_pairs, rel_force_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c2_state():
    p, r, v = config.config_c2()
    say("case: C2 1M LJ, 200 fused steps to disorder the lattice")
    g = MDContext(p, r, v)
    g.setup(0)
    g.run_vv(1, 200)
    g.rebuild()
    r2, v2, a2 = g.state()
    g.close()
    return p, r2, v2, a2


def test_c2_cells_lists_forces(c2_state):
    p, r, v, a = c2_state
    say("case: C2 1M cells, neighbour set, record and plain forces vs oracle")
    o = Oracle(p, r, v, a)
    g = MDContext(p, r, v, a)
    geo = g.geometry()
    assert list(geo["ncells"]) == [39, 39, 39]
    o.setup()
    g.assign_to_cell()
    g.messenger_updateborders(1)
    g.assign_to_neighbourlist()
    assert np.array_equal(g.cells(), o.cells())
    oi, oj, os_ = o.neighbourlist()
    ko = canonical_pairs(oi, oj, os_)
    del oi, oj, os_
    gi, gj, gs = g.neighbourlist()
    kg = canonical_pairs(gi, gj, gs)
    del gi, gj, gs
    assert ko.size > 30 * r.shape[0]
    assert np.array_equal(ko, kg)
    del ko, kg
    for it in (0, 3):  # record step (PE, virial), then a step without tallies
        o.compute_forces(it)
        g.simulation_compute_forces(it)
        err = rel_force_err(g.state()[2], o.state()[2], o.force_scale())
        say(f"  step {it}: scale-aware force error {err:.2e}")
        assert err < TOL
        if it == 0:
            peo, viro = o.molecular()
            peg, virg = g.molecular()
            assert np.abs(peg - peo).max() <= TOL * np.abs(peo).max()
            assert np.abs(virg - viro).max() <= TOL * np.abs(viro).max()
            mo, mg = o.macros(), g.evaluate_macroscopic_properties()
            for key in ("kinenergy", "potenergy", "virial", "pressure", "temperature", "totenergy"):
                assert mg[key] == pytest.approx(mo[key], rel=1e-12, abs=1e-14), key
    g.close()


def test_c2_drop_in_trajectory_vs_oracle(c2_state):
    """Twelve md_advance_vv steps routine by routine on both sides; the forced rebuild of step 9
    (TPLOT 10) runs sendmols, binning and the list build at full size."""
    p, r, v, a = c2_state
    say("case: C2 1M drop-in trajectory vs oracle, 12 steps")
    o = Oracle(p, r, v, a)
    g = MDContext(p, r, v, a)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    ro, rg = [], []
    for it in range(1, 13):
        o.move_particles_vv(1)
        o.updateborders(0)
        if o.checkrebuild(it):
            o.rebuild()
            ro.append(it)
        o.compute_forces(it)
        o.move_particles_vv(2)
        if g.md_advance_vv(it):
            rg.append(it)
    assert ro == rg and 9 in ro
    (rO, vO, aO), (rG, vG, aG) = o.state(), g.state()
    assert np.abs(rG - rO).max() < 1e-11
    assert np.abs(vG - vO).max() < 1e-10
    g.close()


def _fused_vs_routine(p, r, v, a, nsteps, label):
    g = MDContext(p, r, v, a)
    f = MDContext(p, r, v, a)
    g.setup(0)
    f.setup(0)
    rg, mg = [], []
    for it in range(1, nsteps + 1):
        if g.md_advance_vv(it):
            rg.append(it)
        if it % p.tplot == 0:
            mg.append(g.evaluate_macroscopic_properties())
    mf, nrb = f.run_vv(1, nsteps, n_macros=len(mg))
    assert mf == mg  # record-step sums reduced on the device inside the fused windows
    fired = [it for it in rg if (it + 1) % p.tplot != 0]
    say(f"  {label}: rebuilds at {rg} (displacement-fired {fired})")
    assert nrb == len(rg)
    assert fired, "no displacement-fired rebuild: the case does not exercise a fired check"
    for x, y in zip(f.state(), g.state()):
        assert np.array_equal(x, y)
    f.close()
    g.close()


def test_c2_fused_loop_equals_routine_sequence(c2_state):
    """2,197 bricks against about 512 resident workgroups: a fused force launch whose next-step
    displacement check fires must still compute every brick of its own step."""
    p, r, v, a = c2_state
    say("case: C2 1M fused run_vv vs routine sequence, 40 steps")
    _fused_vs_routine(p, r, v, a, 40, "C2")


def test_one_cell_bricks_fused_loop_equals_routine_sequence():
    """Bricks of one cell on a 32,000-molecule liquid: 729 workgroups per force launch."""
    p, r, v = config.config_fcc((20, 20, 20), 0.8, 2.5, seed=31, tplot=10)
    p.brick[0], p.brick[1], p.brick[2] = 1, 1, 1
    say("case: 32k LJ, brick (1,1,1), fused vs routine, 80 steps")
    g = MDContext(p, r, v)
    g.setup(0)
    g.run_vv(1, 100)
    g.rebuild()
    r2, v2, a2 = g.state()
    g.close()
    _fused_vs_routine(p, r2, v2, a2, 80, "brick 1x1x1")
