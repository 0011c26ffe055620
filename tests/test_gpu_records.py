"""Record-step sums and outputs through the C ABI.

  * the fused loop reduces each record step's sums on the device (a ring read at the window's
    synchronisation); they equal, bit for bit, the sums of the routine-by-routine sequence, whose
    record follows pass 2 (src/md.f90:159-161);
  * results/macroscopic_properties written from them matches the oracle's sums column by column;
  * results/pplane: the device's configurational MOP tally plus the kinetic crossings, normalised
    by Nstress_ave and 4*Lx*Lz, equals the oracle's tally plus the same crossings, normalised by
    hand;
  * max(potenergymol) > 1e4 at a record step aborts (src/simulation_record.f90:1271-1282);
  * labels (tag, moltype, monomer fields) move with their molecules when sendmols renumbers them.
"""
import sys

import numpy as np
import pytest

from flowmol_amd import config, records
from flowmol_amd._abi import FM_EABORT, FM_ENSEMBLE_TAG_MOVE
from flowmol_amd._lib import FlowmolHipError
from flowmol_amd.md import MDContext
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _liquid(nunits, seed, tplot, **kw):
    p, r, v = config.config_fcc(nunits, 0.8, 2.5, seed=seed, tplot=tplot, **kw)
    r = r + np.random.default_rng(seed).uniform(-0.02, 0.02, r.shape)
    return p, r, v


def test_fused_record_sums_equal_routine_sequence():
    p, r, v = _liquid((8, 8, 8), 51, 5)
    say("case: record sums, fused ring vs routine sequence, 2048 LJ, 60 steps, TPLOT 5")
    g = MDContext(p, r, v)
    f = MDContext(p, r, v)
    g.setup(0)
    f.setup(0)
    mg = []
    for it in range(1, 61):
        g.md_advance_vv(it)
        if it % 5 == 0:
            mg.append(g.evaluate_macroscopic_properties())
    mf, _ = f.run_vv(1, 60, n_macros=12)
    assert len(mf) == len(mg) == 12
    for a, b in zip(mf, mg):
        assert a == b
    for x, y in zip(f.state(), g.state()):
        assert np.array_equal(x, y)
    f.close()
    g.close()


def test_macroscopic_properties_file_matches_oracle(tmp_path):
    p, r, v = _liquid((8, 8, 8), 52, 5)
    say("case: results/macroscopic_properties vs oracle sums, 2048 LJ, 40 steps")
    o = Oracle(p, r, v)
    o.setup()
    o.compute_forces(0)
    mo, _ = o.run_vv(1, 40, n_macros=8)
    g = MDContext(p, r, v)
    g.setup(0)
    out, _ = records.run_with_records(g, 1, 40, str(tmp_path))
    cols = records.read_macroscopic_properties(str(tmp_path / "results" / "macroscopic_properties"))
    assert list(cols["iter"]) == [5, 10, 15, 20, 25, 30, 35, 40]
    assert np.allclose(cols["simtime"], 0.005 * np.arange(5, 41, 5), atol=1e-4)
    for name, key, dec in (("KE", "kinenergy", 15), ("PE", "potenergy", 15), ("TE", "totenergy", 15),
                           ("Temp", "temperature", 4), ("Pressure", "pressure", 4), ("V^2Sum", "mv2sum", 4)):
        want = np.array([m[key] for m in mo])
        assert np.abs(cols[name] - want).max() <= 0.5 * 10.0 ** -dec + 1e-11 * np.abs(want).max(), name
    g.close()


def test_pplane_matches_hand_normalised_oracle_tally(tmp_path):
    p, r, v, tag, rt = config.config_couette(nunits=(8, 8, 8), tplot=5, nplanes=16)
    say("case: results/pplane vs the oracle's MOP tally normalised by hand, Couette 2048, 20 steps")
    Nv = 2
    o = Oracle(p, r, v)
    o.set_tags(tag, rt)
    o.setup()
    o.compute_forces(0)
    o.reset_mop()
    g = MDContext(p, r, v)
    g.set_tags(tag, rt)
    g.setup(0)
    g.reset_mop()
    records.run_with_records(g, 1, 20, str(tmp_path), Nvflux_ave=Nv)
    # the oracle: configurational tally per record step plus the kinetic crossings of its state
    dom = [p.domain[0], p.domain[1], p.domain[2]]
    acc = np.zeros((3, p.nplanes))
    want = {}
    cnt = 0
    for it in range(1, 21):
        o.move_particles_vv(1)
        o.updateborders(0)
        if o.checkrebuild(it):
            o.rebuild()
        o.compute_forces(it)
        o.move_particles_vv(2)
        if it % p.tplot == 0:
            ro, vo, _ = o.state()
            acc += o.mop().T
            o.reset_mop()
            acc += records.mop_kinetic(ro, vo, 1.0, dom, p.nplanes, p.planespacing, p.delta_t, 2)
            cnt += 1
            if cnt == Nv:
                want[it] = acc / Nv / (4 * dom[0] * dom[2])
                acc[:] = 0.0
                cnt = 0
    raw = np.fromfile(str(tmp_path / "results" / "pplane"), dtype="<f8")
    n = p.nplanes
    for it, P in want.items():
        m = it // p.tplot
        got = raw[(m - 1) * 3 * n:(m - 1) * 3 * n + n]
        ref = P.ravel(order="F")[:n]
        assert np.abs(got - ref).max() <= 1e-10 * max(1e-300, np.abs(ref).max()), it
    g.close()


def test_max_potential_energy_aborts():
    """Two molecules 0.52 apart: U = 4(r^-12 - r^-6) - potshift > 1e4; dt 1e-6 keeps them there."""
    L = 10.0
    pos = np.array([[0.1, 0.2, 0.3], [0.62, 0.2, 0.3]])
    p = config.make_params((L, L, L), 2 / L ** 3, 2.5, dt=1e-6, tplot=1)
    say("case: max(potenergymol) > 1e4 aborts")
    g = MDContext(p, pos, np.zeros((2, 3)))
    g.setup(0)
    with pytest.raises(FlowmolHipError) as e:
        g.evaluate_macroscopic_properties()
    assert e.value.code == FM_EABORT
    with pytest.raises(FlowmolHipError) as e:
        g.run_vv(1, 2, n_macros=2)
    assert e.value.code == FM_EABORT
    g.close()
    # below the limit nothing is raised
    pos[1, 0] = 0.1 + 0.6
    g = MDContext(p, pos, np.zeros((2, 3)))
    g.setup(0)
    m = g.evaluate_macroscopic_properties()
    assert 1000.0 < m["max_potenergymol"] < 1e4
    g.close()


def test_labels_move_with_renumbered_molecules():
    """Molecules put outside the box are wrapped and renumbered by sendmols; the labels read back
    belong to the molecule now carrying each number (matched through the wrapped positions)."""
    p, r, v = config.config_fcc((6, 6, 6), 0.8, 2.5, seed=61, tplot=10, potential_flag=1,
                                ensemble=FM_ENSEMBLE_TAG_MOVE)
    config.set_mie_types(p, {(0, 0): dict(lr=12.0, la=6.0), (0, 1): dict(lr=12.0, la=6.0),
                             (1, 1): dict(lr=14.0, la=6.0)})
    n = r.shape[0]
    rng = np.random.default_rng(61)
    L = np.array([p.domain[0], p.domain[1], p.domain[2]])
    r = r + rng.uniform(-0.01, 0.01, r.shape)
    out = rng.random(n) < 0.2
    r[out, 0] += L[0]          # outside on +x: sendmols wraps and renumbers these
    tag = rng.choice([0, 1, 2, 3, 6], n).astype(np.int32)
    mt = rng.integers(0, 2, n).astype(np.int32)
    chain = np.arange(1000, 1000 + n, dtype=np.int32)
    sub = rng.integers(1, 9, n).astype(np.int32)
    funcy = rng.integers(0, 3, n).astype(np.int32)
    bflag = rng.integers(0, 1 << 20, (n, 4)).astype(np.int32)
    say("case: labels after sendmols renumbering, 864 molecules, 20% outside the box")
    g = MDContext(p, r, v)
    g.set_moltype(mt)
    g.set_tags(tag, None)
    g.set_monomers(chain, sub, funcy, bflag)
    g.linklist_deallocateall()
    g.sendmols()
    r1, _, _ = g.state()
    t1, m1, (c1, s1, f1, b1, gn1) = g.labels()
    wrapped = r.copy()
    wrapped[out, 0] -= L[0]
    # chainID was unique: it names the molecule each number now carries
    src = c1 - 1000
    assert np.array_equal(np.sort(src), np.arange(n))
    assert not np.array_equal(src, np.arange(n))   # some molecules were renumbered
    assert np.array_equal(r1, wrapped[src])
    assert np.array_equal(t1, tag[src]) and np.array_equal(m1, mt[src])
    assert np.array_equal(s1, sub[src]) and np.array_equal(f1, funcy[src]) and np.array_equal(b1, bflag[src])
    assert np.array_equal(gn1, src + 1)   # glob_no defaults to the number at upload and moves too
    g.close()
