"""GPU parity for the remaining physics of the path: Mie types, FENE chains and bond lists,
MOP planes, the per-molecule virial tensor, tethered/sliding walls, the rebuild criteria, the
reference's own restart state, and the mixed-precision pair law (against the FP64 path)."""
import sys

import numpy as np
import pytest

from flowmol_amd import config
from flowmol_amd._abi import FM_INTEGRATION_LFV, FM_PREC_MIXED, FM_REBUILD_FIXED, FM_REBUILD_VMAX
from flowmol_amd.md import MDContext
from flowmol_amd.restart import read_final_state
from oracle.oracle import Oracle
from tests.helpers import canonical_pairs, rel_force_err

pytestmark = pytest.mark.gpu
TOL = 1e-12
GOLDEN = "tests/golden/intrinsic_interface_initial_state"


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def drop_in_steps(o, g, it0, n):
    """md_advance_vv routine by routine on both sides; returns the rebuild steps of each."""
    ro, rg = [], []
    for it in range(it0, it0 + n):
        o.move_particles_vv(1)
        o.updateborders(0)
        if o.checkrebuild(it):
            o.rebuild()
            ro.append(it)
        o.compute_forces(it)
        o.move_particles_vv(2)
        if g.md_advance_vv(it):
            rg.append(it)
    return ro, rg


@pytest.fixture(scope="module")
def polymer_state():
    p, r, v, mt, mono = config.config_polymer(nunits=(6, 8, 6))
    o = Oracle(p, r, v)
    o.set_moltype(mt)
    o.set_monomers(*mono)
    o.setup()
    o.compute_forces(0)
    o.run_vv(1, 100)
    o.rebuild()
    r2, v2, a2 = o.state()
    return p, r2, v2, a2, mt, mono


def test_polymer_mie_bonds_lists_forces(polymer_state):
    p, r, v, a, mt, mono = polymer_state
    say("case: FENE + WCA/Mie14-6 melt 1152 beads")
    o = Oracle(p, r, v, a)
    o.set_moltype(mt)
    o.set_monomers(*mono)
    g = MDContext(p, r, v, a)
    g.set_moltype(mt)
    g.set_monomers(*mono)
    o.setup()
    g.setup(0)
    oi, oj, os_ = o.neighbourlist()
    gi, gj, gs = g.neighbourlist()
    assert np.array_equal(canonical_pairs(oi, oj, os_), canonical_pairs(gi, gj, gs))
    ob, obs = o.bonds()
    gb, gbs = g.bonds()
    canon = lambda b, s: sorted((i, int(x), int(y)) for i in range(b.shape[0]) for x, y in zip(b[i], s[i]) if x >= 0)
    assert canon(ob, obs) == canon(gb, gbs)
    assert sum(1 for row in gb for x in row if x >= 0) == 2 * (r.shape[0] - r.shape[0] // 8)
    o.compute_forces(0)
    _, _, ao = o.state()
    _, _, ag = g.state()
    assert rel_force_err(ag, ao, o.force_scale()) < TOL
    peo, viro = o.molecular()
    peg, virg = g.molecular()
    np.testing.assert_allclose(peg, peo, rtol=TOL, atol=1e-12)
    np.testing.assert_allclose(virg, viro, rtol=TOL, atol=1e-10)
    mo, mg = o.macros(), g.evaluate_macroscopic_properties()
    for k in ("potenergysum_LJ", "potenergysum_POLY", "virial"):
        assert abs(mg[k] - mo[k]) <= 1e-11 * max(1.0, abs(mo[k])), k
    g.close()


def test_polymer_short_trajectory(polymer_state):
    p, r, v, a, mt, mono = polymer_state
    say("case: polymer 40-step drop-in trajectory")
    o = Oracle(p, r, v, a)
    o.set_moltype(mt)
    o.set_monomers(*mono)
    g = MDContext(p, r, v, a)
    g.set_moltype(mt)
    g.set_monomers(*mono)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    ro, rg = drop_in_steps(o, g, 1, 40)
    assert ro == rg
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-9
    g.close()


def _lj(nunits, seed, steps=150, **kw):
    p, r, v = config.config_fcc(nunits, 0.8, 2.5, seed=seed, **kw)
    r = r + np.random.default_rng(seed).uniform(-1e-3, 1e-3, r.shape)
    o = Oracle(p, r, v)
    o.setup()
    o.compute_forces(0)
    o.run_vv(1, steps)
    o.rebuild()
    r2, v2, a2 = o.state()
    return p, r2, v2, a2


def test_mop_planes_and_virial_tensor():
    p, r, v, a = _lj((8, 8, 8), seed=21, tplot=1, vflux_outflag=2, nplanes=16,
                     planespacing=config.fcc_globaldomain((8, 8, 8), 0.8)[1] / 16, pressure_outflag=1)
    say("case: MOP 16 planes + rfmol, 2048 LJ")
    o = Oracle(p, r, v, a)
    g = MDContext(p, r, v, a)
    o.setup()
    g.setup(0)   # iter 0 is a record step: one MOP tally
    o.reset_mop()
    o.compute_forces(0)
    Po, Pg = o.mop(), g.mop()
    assert np.abs(Po).max() > 1.0
    assert np.abs(Pg - Po).max() <= 1e-11 * np.abs(Po).max()
    rfo, rfg = o.rfmol(), g.rfmol()
    assert np.abs(rfg - rfo).max() <= 1e-11 * np.abs(rfo).max()
    # virial tensor trace equals the per-molecule virial (both sides of each pair)
    _, virg = g.molecular()
    assert np.allclose(np.trace(rfg, axis1=1, axis2=2), virg, rtol=1e-12, atol=1e-10)
    g.close()


def _report_worst(Fo, Fg, what, n=8):
    d = np.abs(Fg - Fo)
    if d.max() <= 1e-10 * np.abs(Fo).max():
        return
    say(f"CV mismatch {what}: faces 1-3 max {d[..., :3].max():.3e}, faces 4-6 max {d[..., 3:].max():.3e}")
    for idx in np.argsort(d, axis=None)[::-1][:n]:
        b = np.unravel_index(idx, d.shape)
        say(f"  bin {b[:3]} comp {b[3]} face {b[4] + 1}: oracle {Fo[b]:.6e} gpu {Fg[b]:.6e}")


@pytest.mark.parametrize("ratio", [1.0, 2.0])
def test_cv_surface_stress_lj(ratio):
    """Control-volume surface stresses (vflux_outflag 4): every Pxyface entry, halo bins and the
    face n+3 copies included, at a record step and accumulated over a 12-step trajectory."""
    p, r, v, a = _lj((8, 8, 8), seed=23, tplot=1)
    p.vflux_outflag = 4
    p.tplot = 4
    p.teval = 4
    for d in range(3):
        p.binspercell[d] = ratio
    say(f"case: CV stress 2048 LJ, {ratio} bins per cell")
    o = Oracle(p, r, v, a)
    g = MDContext(p, r, v, a)
    assert [list(x) for x in o.cv_bins()[:2]] == [list(x) for x in g.cv_bins()[:2]]
    o.setup()
    o.reset_cv_stress()
    o.compute_forces(0)
    g.setup(0)
    Fo, Fg = o.cv_stress(), g.cv_stress()
    assert Fo.shape == Fg.shape and np.abs(Fo).max() > 1.0
    assert np.abs(Fg - Fo).max() <= 1e-11 * np.abs(Fo).max()
    for k in range(3):   # record steps 4, 8, 12
        ro, rg = drop_in_steps(o, g, 1 + 4 * k, 4)
        assert ro == rg
        Fo, Fg = o.cv_stress(), g.cv_stress()
        _report_worst(Fo, Fg, f"after step {4 * k + 4}")
        assert np.abs(Fg - Fo).max() <= 1e-10 * np.abs(Fo).max(), k
    g.close()


def test_cv_surface_stress_polymer(polymer_state):
    """CV stresses with FENE bonds (each bond met from both ends) and Mie/WCA pairs."""
    p, r, v, a, mt, mono = polymer_state
    p = type(p).from_buffer_copy(p)   # the fixture's parameters stay as they are
    p.vflux_outflag = 4
    for d in range(3):
        p.binspercell[d] = 1.0
    say("case: CV stress, FENE + Mie melt")
    o = Oracle(p, r, v, a)
    o.set_moltype(mt)
    o.set_monomers(*mono)
    g = MDContext(p, r, v, a)
    g.set_moltype(mt)
    g.set_monomers(*mono)
    o.setup()
    o.reset_cv_stress()
    o.compute_forces(0)
    g.setup(0)
    Fo, Fg = o.cv_stress(), g.cv_stress()
    assert np.abs(Fo).max() > 1.0
    assert np.abs(Fg - Fo).max() <= 1e-11 * np.abs(Fo).max()
    g.close()


def test_couette_walls_short_trajectory():
    p, r, v, tag, rt = config.config_couette(nunits=(8, 8, 8), tplot=5, nplanes=16)
    say("case: tethered walls, sliding top, 30 steps")
    o = Oracle(p, r, v)
    o.set_tags(tag, rt)
    g = MDContext(p, r, v)
    g.set_tags(tag, rt)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    _, _, ao = o.state()
    _, _, ag = g.state()
    assert rel_force_err(ag, ao, o.force_scale() + 1.0) < TOL
    ro, rg = drop_in_steps(o, g, 1, 30)
    assert ro == rg
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-9
    assert np.abs(g.rtether() - o.rtether()).max() < 1e-12
    top = tag == 6
    assert np.allclose(g.rtether()[top, 0] - rt[top, 0], 30 * p.delta_t * 1.0 + p.delta_t, atol=1e-9)
    Po, Pg = o.mop(), g.mop()
    assert np.abs(Pg - Po).max() <= 1e-10 * max(1.0, np.abs(Po).max())
    g.close()


def test_mie_tethered_walls_trajectory():
    """Mie types (12-6 and 14-6, per-pair cutoffs) with tethered sliding walls: the branch-free Mie
    launch with its tether pass on steps without tallies (TPLOT 5; record steps with MOP take the
    general loop) against the oracle, at a plain step and over a 30-step drop-in run."""
    p, r, v, tag, rt = config.config_couette(nunits=(8, 8, 8), tplot=5, nplanes=16)
    config.set_mie_types(p, {(0, 0): dict(lr=12.0, la=6.0), (0, 1): dict(lr=12.0, la=6.0, rc=2 ** (1 / 6)),
                             (1, 1): dict(lr=14.0, la=6.0, eps=1.2)})
    mt = np.random.default_rng(95).integers(0, 2, r.shape[0]).astype(np.int32)
    say("case: Mie types + tethered walls, sliding top, 30 steps")
    o = Oracle(p, r, v)
    o.set_tags(tag, rt)
    o.set_moltype(mt)
    g = MDContext(p, r, v)
    g.set_tags(tag, rt)
    g.set_moltype(mt)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    o.compute_forces(1)   # a step without tallies
    g.simulation_compute_forces(1)
    assert rel_force_err(g.state()[2], o.state()[2], o.force_scale() + 1.0) < TOL
    ro, rg = drop_in_steps(o, g, 1, 30)
    assert ro == rg
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-9
    assert np.abs(g.rtether() - o.rtether()).max() < 1e-12
    g.close()


def _lfv(p):
    q = type(p).from_buffer_copy(p)
    q.integration_algorithm = FM_INTEGRATION_LFV
    return q


def lfv_steps(o, g, it0, n):
    """md_advance_lfv routine by routine on both sides; returns the rebuild steps of each."""
    ro, rg = [], []
    for it in range(it0, it0 + n):
        o.compute_forces(it)
        o.move_particles_lfv()
        o.updateborders(0)
        if o.checkrebuild(it):
            o.rebuild()
            ro.append(it)
        if g.md_advance_lfv(it)[0]:
            rg.append(it)
    return ro, rg


def test_lfv_couette_walls_trajectory():
    """Leap-frog with tag_move (the reference's Couette path): tethers inside the move, sliding
    top wall, MOP every record step; the fused loop repeats the routine-by-routine one bit for bit."""
    p, r, v, tag, rt = config.config_couette(nunits=(8, 8, 8), tplot=5, nplanes=16)
    p = _lfv(p)
    say("case: leap-frog tethered walls, 40 steps")
    o = Oracle(p, r, v)
    o.set_tags(tag, rt)
    g = MDContext(p, r, v)
    g.set_tags(tag, rt)
    o.setup()
    o.compute_forces(0)   # the initial force call of setup (src/setup_initial_record.f90:88)
    g.setup(0)
    ro, rg = lfv_steps(o, g, 0, 40)
    assert ro == rg and len(ro) > 0
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-9
    assert np.abs(g.rtether() - o.rtether()).max() < 1e-12
    Po, Pg = o.mop(), g.mop()
    assert np.abs(Pg - Po).max() <= 1e-10 * max(1.0, np.abs(Po).max())
    # fused loop from the same start
    f = MDContext(p, r, v)
    f.set_tags(tag, rt)
    f.setup(0)
    _, nrb = f.run_lfv(0, 40)
    assert nrb == len(rg)
    for x, y in zip(f.state(), g.state()):
        assert np.array_equal(x, y)
    assert np.array_equal(f.rtether(), g.rtether())
    f.close()
    g.close()


def test_lfv_nve_macros():
    """Leap-frog NVE: record-step energies use v + dt/2 a (src/simulation_record.f90:1238-1247)."""
    p, r, v, a = _lj((6, 6, 6), seed=33, tplot=5)
    p = _lfv(p)
    say("case: leap-frog NVE 864 LJ, 30 steps")
    o = Oracle(p, r, v, a)
    g = MDContext(p, r, v, a)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    mo, nro = o.run_lfv(0, 30, n_macros=6)
    mg, nrg = g.run_lfv(0, 30, n_macros=6)
    assert nro == nrg
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-9
    for x, y in zip(mo, mg):
        for key in ("kinenergy", "potenergy", "pressure", "totenergy"):
            assert y[key] == pytest.approx(x[key], rel=1e-11, abs=1e-13), key
    g.close()


@pytest.mark.parametrize("crit", [FM_REBUILD_VMAX, FM_REBUILD_FIXED])
def test_rebuild_criteria_trajectory(crit):
    p, r, v, a = _lj((6, 6, 6), seed=23, tplot=1000, rebuild_criteria=crit, fixed_rebuild=7)
    say(f"case: rebuild criterion {crit}, 60 steps")
    o = Oracle(p, r, v, a)
    g = MDContext(p, r, v, a)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    ro, rg = drop_in_steps(o, g, 1, 60)
    assert ro == rg and len(ro) > 2
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-10
    g.close()


def test_reference_restart_state():
    d = read_final_state(GOLDEN)
    h = d["header"]
    p = config.make_params(h["globaldomain"], h["density"], h["rcutoff"], delta=h["delta_rneighbr"],
                           dt=h["delta_t"], tplot=h["tplot"])
    say("case: reference final_state (liquid-vapour 2635)")
    o = Oracle(p, d["r"], d["v"])
    g = MDContext(p, d["r"], d["v"])
    o.assign_to_cell()
    g.assign_to_cell()
    assert np.array_equal(o.cells(), g.cells())
    o.updateborders(1)
    o.assign_to_neighbourlist()
    g.assign_to_neighbourlist()
    oi, oj, os_ = o.neighbourlist()
    gi, gj, gs = g.neighbourlist()
    assert np.array_equal(canonical_pairs(oi, oj, os_), canonical_pairs(gi, gj, gs))
    o.compute_forces(0)
    g.simulation_compute_forces(0)
    assert rel_force_err(g.state()[2], o.state()[2], o.force_scale()) < TOL
    ro, rg = drop_in_steps(o, g, 1, 50)
    assert ro == rg
    g.close()


def test_mixed_precision_error_vs_fp64():
    """FP32 pair law: force error against the FP64 path stays at FP32 level."""
    p, r, v, a = _lj((8, 8, 8), seed=25, tplot=10)
    say("case: mixed precision forces, 2048 LJ")
    g64 = MDContext(p, r, v, a)
    g64.setup(0)
    pm, _, _ = config.config_fcc((8, 8, 8), 0.8, 2.5, seed=25, tplot=10, precision=FM_PREC_MIXED)
    g32 = MDContext(pm, r, v, a)
    g32.setup(0)
    a64, a32 = g64.state()[2], g32.state()[2]
    rms = np.sqrt(((a32 - a64) ** 2).sum() / (a64 ** 2).sum())
    say(f"  record-step force RMS relative error {rms:.2e}")
    assert rms < 1e-5
    pe64, _ = g64.molecular()
    pe32, _ = g32.molecular()
    assert abs(pe32.sum() - pe64.sum()) < 1e-5 * abs(pe64.sum())
    # a step without tallies takes the branch-free FP32 loop over the FP32 staging
    g64.simulation_compute_forces(3)
    g32.simulation_compute_forces(3)
    a64, a32 = g64.state()[2], g32.state()[2]
    rms = np.sqrt(((a32 - a64) ** 2).sum() / (a64 ** 2).sum())
    say(f"  plain-step force RMS relative error {rms:.2e}")
    assert rms < 1e-5
    g64.close()
    g32.close()


def test_mixed_precision_cutoff_band():
    """The mixed mode's FP32-staged plain step decides a pair near rc on the FP64 r^2: on an FCC
    lattice whose second shell lies exactly at rc (lattice constant rc; 6 such pairs per molecule)
    jittered by 1e-7, the FP64 path counts some of those pairs and not others; a pair that the
    mixed launch classified otherwise would move its molecule's force by |F(rc)| = 0.039, far
    above the FP32 rounding of the rest (forces below 1e-4 here)."""
    rc = 2.5
    L = 6
    p = config.make_params((L * rc, L * rc, L * rc), 4.0 / rc ** 3, rc, tplot=10)
    pm = config.make_params((L * rc, L * rc, L * rc), 4.0 / rc ** 3, rc, tplot=10, precision=FM_PREC_MIXED)
    r = config.fcc_lattice((L, L, L), 4.0 / rc ** 3)
    r = r + np.random.default_rng(77).uniform(-1e-7, 1e-7, r.shape)
    v = np.zeros_like(r)
    say("case: mixed precision, pairs at rc within 1e-7, 864 molecules")
    g64 = MDContext(p, r, v)
    g64.setup(0)
    g32 = MDContext(pm, r, v)
    g32.setup(0)
    g64.simulation_compute_forces(3)   # plain steps: the FP32-staged launch in the mixed context
    g32.simulation_compute_forces(3)
    a64, a32 = g64.state()[2], g32.state()[2]
    assert np.abs(a64).max() > 1e-3     # the counted and uncounted shell pairs leave net forces
    assert np.abs(a32 - a64).max() < 1e-4
    g64.close()
    g32.close()


def test_mixed_precision_fused_loop_equals_routine_sequence():
    """The fused loop's mixed-precision force launches (with the next step's pass 1) repeat the
    step-by-step sequence bit for bit."""
    pm, r, v = config.config_fcc((6, 6, 6), 0.8, 2.5, seed=27, tplot=10, precision=FM_PREC_MIXED)
    r = r + np.random.default_rng(27).uniform(-1e-3, 1e-3, r.shape)
    say("case: mixed precision fused loop vs routine sequence, 864 LJ, 60 steps")
    g = MDContext(pm, r, v)
    f = MDContext(pm, r, v)
    g.setup(0)
    f.setup(0)
    rg = [it for it in range(1, 61) if g.md_advance_vv(it)]
    _, nrb = f.run_vv(1, 60)
    assert nrb == len(rg) and len(rg) > 2
    for x, y in zip(f.state(), g.state()):
        assert np.array_equal(x, y)
    f.close()
    g.close()


@pytest.mark.parametrize("nunits,density,rc", [((8, 8, 8), 0.8, 4.0), ((8, 8, 8), 1.1, 3.0), ((6, 6, 6), 0.3, 2.5)])
def test_mixed_precision_other_cutoffs_and_densities(nunits, density, rc):
    """The mixed mode's FP32 staging at a 4.0 cutoff (about 76 molecules per cell, cells of 4.56
    a side), a 1.1 state at rc 3.0 and a 0.3 gas: forces within 1e-5 RMS of the FP64 path on a
    record step and on a step without tallies, PE within 1e-5, and the fused loop equal to the
    routine sequence over 60 steps with the same rebuild steps."""
    p, r, v = config.config_fcc(nunits, density, rc, seed=19, tplot=10)
    pm, _, _ = config.config_fcc(nunits, density, rc, seed=19, tplot=10, precision=FM_PREC_MIXED)
    r = r + np.random.default_rng(19).uniform(-1e-3, 1e-3, r.shape)
    o = Oracle(p, r, v)
    o.setup()
    o.compute_forces(0)
    o.run_vv(1, 100)
    o.rebuild()
    r, v, a = o.state()
    say(f"case: mixed precision, LJ {r.shape[0]} density {density} rc {rc}")
    g64 = MDContext(p, r, v, a)
    g32 = MDContext(pm, r, v, a)
    g64.setup(0)
    g32.setup(0)
    for it in (0, 3):  # setup's record-step forces, then a step without tallies
        if it:
            g64.simulation_compute_forces(it)
            g32.simulation_compute_forces(it)
        a64, a32 = g64.state()[2], g32.state()[2]
        rms = np.sqrt(((a32 - a64) ** 2).sum() / (a64 ** 2).sum())
        say(f"  step {it} force RMS relative error {rms:.2e}")
        assert rms < 1e-5
        if it == 0:
            pe64, _ = g64.molecular()
            pe32, _ = g32.molecular()
            assert abs(pe32.sum() - pe64.sum()) < 1e-5 * abs(pe64.sum())
    g64.close()
    g32.close()
    g = MDContext(pm, r, v, a)
    f = MDContext(pm, r, v, a)
    g.setup(0)
    f.setup(0)
    rg = [it for it in range(1, 61) if g.md_advance_vv(it)]
    _, nrb = f.run_vv(1, 60)
    assert nrb == len(rg) and len(rg) >= 6
    for x, y in zip(f.state(), g.state()):
        assert np.array_equal(x, y)
    f.close()
    g.close()


# 2-cell blocks on 8 lattice units: 4 cells a side, so the 2 blocks a side split the box evenly.
# On 6 units (3 cells a side) the first block holds 8/27 of the molecules, above the estimate
# ceil(1.5*108 + 5) = 167, so the first calls skip and raise the safety factors instead.
@pytest.mark.parametrize("flag,freq,bs,nu,moves", [(1, 1, 1, 6, True), (2, 2, 1, 6, False),
                                                   (2, 1, 2, 8, True), (2, 1, 2, 6, False)])
def test_sort_mols_trajectory(flag, freq, bs, nu, moves):
    """sort_mols inside the rebuild branch renumbers the molecules along the block curve
    (src/linklist.f90:2744-2865): both sides must agree on the new numbering, so positions
    compare molecule by molecule after every rebuild; the fused loop repeats the routine one."""
    p, r, v, a = _lj((nu, nu, nu), seed=29, steps=100, tplot=1000, sort_flag=flag, sort_freq=freq,
                     sortblocksize=bs)
    # start from a shuffled numbering so that the first sort has work to do
    perm = np.random.default_rng(31).permutation(r.shape[0])
    r, v, a = r[perm].copy(), v[perm].copy(), a[perm].copy()
    say(f"case: sort_mols flag {flag} freq {freq} block {bs}, {r.shape[0]} LJ, 60 steps")
    o = Oracle(p, r, v, a)
    g = MDContext(p, r, v, a)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    # a direct call right after setup: the counter starts at 1, so freq 1 sorts here
    o.sort_mols()
    g.sort_mols()
    assert moves == (not np.array_equal(o.state()[0], r))
    for x, y in zip(o.state()[:2], g.state()[:2]):
        assert np.array_equal(x, y)
    o.rebuild()
    g.rebuild()
    o.compute_forces(0)
    g.simulation_compute_forces(0)
    ro, rg = drop_in_steps(o, g, 1, 60)
    assert ro == rg and len(ro) > 2
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-10
    assert np.abs(g.state()[1] - o.state()[1]).max() < 1e-10
    f = MDContext(p, r, v, a)
    f.setup(0)
    f.sort_mols()
    f.rebuild()
    f.simulation_compute_forces(0)
    _, nrb = f.run_vv(1, 60)
    assert nrb == len(rg)
    for x, y in zip(f.state(), g.state()):
        assert np.array_equal(x, y)
    f.close()
    g.close()


@pytest.mark.parametrize("flag", [1, 2])
def test_sort_mols_many_blocks(flag):
    """343 one-cell blocks (6912 molecules, 7 cells a side): block numbers up to 343 take two
    8-bit passes of the radix sort.  The first direct call skips (a block above the estimate) and
    raises the safety factors, the second sorts; both sides hold the same numbering after each,
    and along 40 steps whose rebuilds sort again."""
    p, r, v, a = _lj((12, 12, 12), seed=29, steps=100, tplot=1000, sort_flag=flag, sort_freq=1, sortblocksize=1)
    perm = np.random.default_rng(31).permutation(r.shape[0])
    r, v, a = r[perm].copy(), v[perm].copy(), a[perm].copy()
    say(f"case: sort_mols flag {flag}, 343 blocks, {r.shape[0]} LJ, 40 steps")
    o = Oracle(p, r, v, a)
    g = MDContext(p, r, v, a)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    for moves in (False, True):
        o.sort_mols()
        g.sort_mols()
        assert moves == (not np.array_equal(o.state()[0], r))
        for x, y in zip(o.state()[:2], g.state()[:2]):
            assert np.array_equal(x, y)
    o.rebuild()
    g.rebuild()
    o.compute_forces(0)
    g.simulation_compute_forces(0)
    ro, rg = drop_in_steps(o, g, 1, 40)
    assert ro == rg
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-10
    assert np.abs(g.state()[1] - o.state()[1]).max() < 1e-10
    g.close()


def test_sort_mols_tethered_walls_lfv():
    """Tags, tether sites and sliding velocities travel with the renumbered molecules."""
    p, r, v, tag, rt = config.config_couette(nunits=(8, 8, 8), tplot=5, nplanes=16, sort_flag=2,
                                             sort_freq=1, sortblocksize=1)
    p = _lfv(p)
    say("case: sort_mols flag 2 with leap-frog tethered walls, 40 steps")
    o = Oracle(p, r, v)
    o.set_tags(tag, rt)
    g = MDContext(p, r, v)
    g.set_tags(tag, rt)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    ro, rg = lfv_steps(o, g, 0, 40)
    assert ro == rg and len(ro) > 0
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-9
    assert np.abs(g.rtether() - o.rtether()).max() < 1e-12
    Po, Pg = o.mop(), g.mop()
    assert np.abs(Pg - Po).max() <= 1e-10 * max(1.0, np.abs(Po).max())
    g.close()


def test_queued_rebuild_grows_capacities(monkeypatch):
    """The fused loop queues its rebuilds without waiting on the device.  A queued build that meets
    a list longer than the list capacity, or a staged region larger than the region capacity,
    stops the launches queued after it; the loop grows the capacity and builds again.  A converging
    flow packs the molecules together so that both happen, and the fused trajectory stays the
    step-by-step one (which grows its capacities at each build)."""
    p, r, v = config.config_fcc((6, 6, 6), 0.8, 2.5, seed=41, tplot=1000)
    r = r + np.random.default_rng(41).uniform(-1e-3, 1e-3, r.shape)
    v = v - 0.3 * r
    p.max_neighbours = 78   # the lattice's count; capacity 80, exceeded once the flow packs them
    say("case: queued rebuilds under a converging flow, 864 LJ, 60 steps")
    o = Oracle(p, r, v)
    o.setup()
    o.compute_forces(0)
    g = MDContext(p, r, v)
    g.setup(0)
    monkeypatch.setenv("FLOWMOL_TIGHT_REGIONS", "1")
    f = MDContext(p, r, v)
    monkeypatch.delenv("FLOWMOL_TIGHT_REGIONS")
    f.setup(0)
    nb0, reg0 = f.capacity()
    assert nb0 == 80
    ro, rg = drop_in_steps(o, g, 1, 60)
    _, nrb = f.run_vv(1, 60)
    nb1, reg1 = f.capacity()
    assert nb1 > nb0 and reg1 > reg0, (nb0, nb1, reg0, reg1)
    assert ro == rg and nrb == len(rg)
    for x, y in zip(f.state(), g.state()):
        assert np.array_equal(x, y)
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-10
    f.close()
    g.close()


def test_queued_leapfrog_grows_capacities(monkeypatch):
    """The leap-frog loop at TPLOT 1 queues a rebuild every step without waiting on the device.  A
    converging flow packs the molecules until a staged region exceeds the region capacity; the
    launches queued after that build do nothing, the loop grows the capacity and goes on from the
    halted step, and the trajectory, the rebuild count and every record step's sums stay those of
    the step-by-step sequence."""
    from flowmol_amd._abi import FM_INTEGRATION_LFV

    p, r, v = config.config_fcc((6, 6, 6), 0.8, 2.5, seed=43, tplot=1)
    p.integration_algorithm = FM_INTEGRATION_LFV
    r = r + np.random.default_rng(43).uniform(-1e-3, 1e-3, r.shape)
    v = v - 0.3 * r
    say("case: queued leap-frog rebuilds under a converging flow, 864 LJ, 60 steps")
    g = MDContext(p, r, v)
    g.setup(0)
    monkeypatch.setenv("FLOWMOL_TIGHT_REGIONS", "1")
    f = MDContext(p, r, v)
    monkeypatch.delenv("FLOWMOL_TIGHT_REGIONS")
    f.setup(0)
    _, reg0 = f.capacity()
    mg, rg = [], []
    for it in range(0, 60):
        rb, m = g.md_advance_lfv(it, macros=True)
        mg.append(m)
        rg.append(rb)
    mf, nrb = f.run_lfv(0, 60, n_macros=60)
    _, reg1 = f.capacity()
    assert reg1 > reg0, (reg0, reg1)
    assert nrb == sum(bool(x) for x in rg) == 60
    for x, y in zip(f.state(), g.state()):
        assert np.array_equal(x, y)
    assert mf == mg
    f.close()
    g.close()


def test_polymer_random_walk_start():
    """The C4 initial state (seeded random-walk chains) on a small box: neighbour set, bond lists,
    forces and a 20-step drop-in trajectory against the oracle."""
    p, r, v, mt, mono = config.config_polymer(nunits=(6, 8, 6), init="random_walk", seed=4)
    say("case: random-walk FENE + WCA/Mie14-6 melt 1152 beads")
    o = Oracle(p, r, v)
    o.set_moltype(mt)
    o.set_monomers(*mono)
    g = MDContext(p, r, v)
    g.set_moltype(mt)
    g.set_monomers(*mono)
    o.setup()
    g.setup(0)
    oi, oj, os_ = o.neighbourlist()
    gi, gj, gs = g.neighbourlist()
    assert np.array_equal(canonical_pairs(oi, oj, os_), canonical_pairs(gi, gj, gs))
    ob, obs = o.bonds()
    gb, gbs = g.bonds()
    canon = lambda b, s: sorted((i, int(x), int(y)) for i in range(b.shape[0]) for x, y in zip(b[i], s[i]) if x >= 0)
    assert canon(ob, obs) == canon(gb, gbs)
    o.compute_forces(0)
    assert rel_force_err(g.state()[2], o.state()[2], o.force_scale()) < TOL
    ro, rg = drop_in_steps(o, g, 1, 20)
    assert ro == rg
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-9
    g.close()


MIE_CASES = {
    # four types, exponents up to x^18 (five bits)
    "lr_to_34": {(0, 0): dict(lr=12.0, la=6.0), (0, 1): dict(lr=14.0, la=6.0, eps=0.8, rc=2.0),
                 (0, 2): dict(lr=20.0, la=6.0, rc=1.8), (0, 3): dict(lr=12.0, la=6.0, rc=2 ** (1 / 6)),
                 (1, 1): dict(lr=34.0, la=8.0, eps=1.2, alpha=0.9), (1, 2): dict(lr=16.0, la=6.0, rc=2.2),
                 (1, 3): dict(lr=12.0, la=6.0, rc=1.5), (2, 2): dict(lr=14.0, la=6.0),
                 (2, 3): dict(lr=12.0, la=6.0, eps=0.5), (3, 3): dict(lr=24.0, la=12.0)},
    # four types, la = 6 and lr = 12..18 throughout (the products class)
    "la6_lr12_to_18": {(0, 0): dict(lr=12.0, la=6.0), (0, 1): dict(lr=14.0, la=6.0, eps=0.8, rc=2.0),
                       (0, 2): dict(lr=16.0, la=6.0, rc=1.8), (0, 3): dict(lr=18.0, la=6.0, rc=2 ** (1 / 6)),
                       (1, 1): dict(lr=18.0, la=6.0, eps=1.2, alpha=0.9), (1, 2): dict(lr=12.0, la=6.0, rc=2.2),
                       (1, 3): dict(lr=16.0, la=6.0, rc=1.5), (2, 2): dict(lr=14.0, la=6.0),
                       (2, 3): dict(lr=12.0, la=6.0, eps=0.5), (3, 3): dict(lr=18.0, la=6.0)},
    # two types, la = 8 (four bits)
    "two_types_la8": {(0, 0): dict(lr=12.0, la=8.0), (0, 1): dict(lr=14.0, la=8.0, eps=0.8, rc=2.0),
                      (1, 1): dict(lr=16.0, la=8.0, eps=1.2, alpha=0.9, rc=1.6)},
    # one pair at lr = 13 (x^7.5): the general loop
    "half_exponent": {(0, 0): dict(lr=12.0, la=6.0), (0, 1): dict(lr=14.0, la=6.0, eps=0.8, rc=2.0),
                      (1, 1): dict(lr=13.0, la=8.0, eps=1.2, alpha=0.9), (2, 2): dict(lr=14.0, la=6.0),
                      (0, 2): dict(lr=12.0, la=6.0), (1, 2): dict(lr=12.0, la=6.0), (3, 3): dict(lr=12.0, la=6.0),
                      (0, 3): dict(lr=12.0, la=6.0), (1, 3): dict(lr=12.0, la=6.0), (2, 3): dict(lr=12.0, la=6.0)},
}


@pytest.mark.parametrize("case", sorted(MIE_CASES))
def test_mie_plain_step_forces(case):
    """Forces of a step without tallies against the oracle, then a 20-step drop-in run (record
    steps of PE and virial every 10), for Mie mixtures with per-pair cutoffs (2^(1/6) to 2.5) and
    masses 0.5 to 2.  With every exponent lr/2+1, la/2+1 whole the steps take the branch-free
    loop: the products class (la = 6, lr = 12..18), four and five exponent bits, two and four
    types; with one pair at lr = 13 (x^7.5) the general loop."""
    p, r, v = config.config_fcc((6, 6, 6), 0.8, 2.5, seed=71, tplot=10)
    config.set_mie_types(p, MIE_CASES[case])
    ntypes = p.ntypes
    for t, m in enumerate((1.0, 2.0, 0.5, 1.5)):
        p.mass_lookup[t] = m
    rng = np.random.default_rng(71)
    r = r + rng.uniform(-0.05, 0.05, r.shape)
    mt = rng.integers(0, ntypes, r.shape[0]).astype(np.int32)
    say("case: Mie plain step, %d types, %s, 864 molecules" % (ntypes, case))
    o = Oracle(p, r, v)
    o.set_moltype(mt)
    g = MDContext(p, r, v)
    g.set_moltype(mt)
    o.setup()
    o.compute_forces(0)
    g.setup(0)
    o.compute_forces(1)   # a step without tallies (TPLOT 10)
    g.simulation_compute_forces(1)
    ao, ag = o.state()[2], g.state()[2]
    assert np.abs(ao).max() > 1.0
    assert rel_force_err(ag, ao, o.force_scale()) < TOL
    ro, rg = drop_in_steps(o, g, 1, 20)
    assert ro == rg
    assert np.abs(g.state()[0] - o.state()[0]).max() < 1e-9
    mo, mg = o.macros(), g.evaluate_macroscopic_properties()   # step 20: a record step
    for k in ("potenergysum_LJ", "virial"):
        assert abs(mg[k] - mo[k]) <= 1e-10 * max(1.0, abs(mo[k])), k
    g.close()


def test_polymer_fused_loop_equals_routine_sequence(polymer_state):
    """The queued loop's Mie force launches with bonds (the next step's pass 1 carried by the
    force launch, record steps of PE and virial every 10 steps) repeat the step-by-step sequence
    bit for bit: state, rebuild steps and record sums; the record sums also match the oracle."""
    p, r, v, a, mt, mono = polymer_state
    say("case: polymer fused loop vs routine sequence, 1152 beads, 60 steps")
    g = MDContext(p, r, v, a)
    f = MDContext(p, r, v, a)
    o = Oracle(p, r, v, a)
    for x in (g, f, o):
        x.set_moltype(mt)
        x.set_monomers(*mono)
    g.setup(0)
    f.setup(0)
    o.setup()
    o.compute_forces(0)
    mg, rg = [], []
    for it in range(1, 61):
        if g.md_advance_vv(it):
            rg.append(it)
        if it % 10 == 0:
            mg.append(g.evaluate_macroscopic_properties())
    mf, nrb = f.run_vv(1, 60, n_macros=6)
    mo, _ = o.run_vv(1, 60, n_macros=6)
    assert nrb == len(rg) and len(rg) > 2
    assert len(mf) == len(mg) == len(mo) == 6
    for x, y in zip(mf, mg):
        assert x == y
    for x, y in zip(mf, mo):
        for k in ("potenergysum_LJ", "potenergysum_POLY", "virial", "kinenergy"):
            assert abs(x[k] - y[k]) <= 1e-9 * max(1.0, abs(y[k])), k
    for x, y in zip(f.state(), g.state()):
        assert np.array_equal(x, y)
    f.close()
    g.close()
