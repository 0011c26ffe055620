"""Replica batches and the serving loop: each replica's result is the same alone and in the mix."""
import sys

import numpy as np
import pytest

from flowmol_amd import config
from flowmol_amd.md import MDContext
from flowmol_amd.serve import Batch, ReplicaServer, couette_request, make_context

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def alone(req, nsteps):
    ctx = make_context(req)
    ctx.run_vv(1, nsteps)
    r, v, a = ctx.state()
    out = (r, v, ctx.mop(), ctx.rtether())
    ctx.close()
    return out


@pytest.mark.parametrize("graphs", [True, False])
def test_batch_members_match_alone(graphs):
    """Three Couette replicas of different sizes and wall speeds; one joins 20 steps after the
    others.  Positions, velocities and tethers are bit-identical to the replica advanced alone;
    the MOP tallies (FP64 atomics) agree to 1e-12 relative."""
    reqs = [couette_request(0, 0.0, 4 * 7 ** 3, 0.7, 60, seed=11),
            couette_request(1, 0.0, 4 * 9 ** 3, 1.5, 60, seed=12),
            couette_request(2, 0.0, 4 * 8 ** 3, 1.1, 40, seed=13)]
    say(f"case: batch of 3 replicas, graphs={graphs}")
    ref = [alone(q, q.nsteps) for q in reqs]
    b = Batch(graphs=graphs)
    ctx = [make_context(q) for q in reqs]
    b.add(ctx[0], 1)
    b.add(ctx[1], 1)
    b.run(20)
    b.add(ctx[2], 1)      # arrives while the others are on the GPU
    b.run(40)
    # rebuild decisions on the device: one host synchronisation per window of queued steps
    nrb, syncs = b.counters()
    say(f"  60 steps: {nrb} member rebuilds, {syncs} host synchronisations")
    assert 0 < syncs <= 30 and nrb > 0
    for k, (c, q) in enumerate(zip(ctx, reqs)):
        r, v, _ = c.state()
        assert np.array_equal(r, ref[k][0]), k
        assert np.array_equal(v, ref[k][1]), k
        assert np.array_equal(c.rtether(), ref[k][3]), k
        P = c.mop()
        assert np.abs(P - ref[k][2]).max() <= 1e-12 * max(1.0, np.abs(ref[k][2]).max()), k
    b.close()
    for c in ctx:
        c.close()


def test_server_completes_and_reports_latency():
    reqs = [couette_request(k, 0.02 * k, n, 0.5 + 0.3 * k, 100, seed=20 + k)
            for k, n in enumerate([2000, 9000, 4000, 30000, 2500])]
    say("case: server, 5 replicas, budget for 2 at a time")
    budget = sorted(q.footprint for q in reqs)[-1] + sorted(q.footprint for q in reqs)[-2]
    srv = ReplicaServer(hbm_budget_bytes=budget, chunk=25)
    results, stats = srv.serve(reqs)
    srv.close()
    assert stats["replicas"] == 5
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0
    # replicas that had to wait for memory show it in their queue time
    assert max(r["queue_wait"] for r in results.values()) > 0
    # each result equals the replica run alone
    for q in reqs[:2]:
        ctx = make_context(q)
        ctx.run_vv(1, q.nsteps)
        m = ctx.evaluate_macroscopic_properties()
        assert results[q.rid]["macros"]["totenergy"] == m["totenergy"]
        ctx.close()


@pytest.mark.parametrize("kind", ["lj", "couette"])
def test_checkpoint_resume_bit_identical(kind, tmp_path):
    """N steps, a checkpoint (the rebuild sequence, then the state read back), N more steps; a
    context resumed from the checkpoint written to disk and read again runs the same N steps to the
    same bits (positions, velocities, accelerations, tether sites, simtime, record sums)."""
    from flowmol_amd import checkpoint

    if kind == "lj":
        p, r, v = config.config_fcc((8, 8, 8), 0.8, 2.5, seed=71, tplot=10)
        r = r + np.random.default_rng(71).uniform(-0.02, 0.02, r.shape)
        req = None
    else:
        req = couette_request(0, 0.0, 4 * 8 ** 3, 1.3, 60, seed=72)
    say(f"case: checkpoint/resume, {kind}, 30 + 30 steps")

    def fresh():
        if req is not None:
            return make_context(req)
        c = MDContext(p, r, v)
        c.setup(0)
        return c

    a = fresh()
    a.run_vv(1, 30)
    snap = checkpoint.take(a, 31)
    ma, _ = a.run_vv(31, 30, n_macros=3)
    path = str(tmp_path / "snap.npz")
    snap.save(path)
    b = checkpoint.resume(checkpoint.Snapshot.load(path))
    assert b.simtime == snap.simtime
    mb, _ = b.run_vv(31, 30, n_macros=3)
    for x, y in zip(a.state(), b.state()):
        assert np.array_equal(x, y)
    assert ma == mb
    assert a.simtime == b.simtime
    if req is not None:
        assert np.array_equal(a.rtether(), b.rtether())
    a.close()
    b.close()


def test_server_abort_and_resume():
    """A replica aborted by the server (at a chunk boundary) leaves a checkpoint; resumed for the
    steps it had left, its record sums equal those of the same replica run alone with the
    checkpoint taken at the same step."""
    from flowmol_amd import checkpoint
    from flowmol_amd.serve import Request

    short = couette_request(0, 0.0, 4 * 7 ** 3, 0.9, 20, seed=81)
    long_ = couette_request(1, 0.0, 4 * 8 ** 3, 1.4, 100, seed=82)
    say("case: server abort at a chunk boundary, resume from the checkpoint")
    srv = ReplicaServer(chunk=20)

    def on_result(res):
        if res["rid"] == 0:
            long_.abort()

    results, stats = srv.serve([short, long_], realtime=False, on_result=on_result)
    assert stats["aborted"] == 1 and results[1]["status"] == "aborted"
    done = results[1]["steps_done"]
    assert done == 20
    snap = results[1]["snapshot"]
    assert snap.next_iter == 21
    res2, _ = srv.serve([Request.resumed(2, 0.0, snap, 100 - done)], realtime=False)
    srv.close()
    # the same replica alone, with the checkpoint at the same step
    c = make_context(long_)
    c.run_vv(1, done)
    checkpoint.take(c, done + 1)
    c.run_vv(done + 1, 100 - done)
    m = c.evaluate_macroscopic_properties()
    assert res2[2]["macros"] == m
    assert np.abs(res2[2]["mop"] - c.mop()).max() <= 1e-12 * max(1.0, np.abs(c.mop()).max())
    c.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lj_batch_members_match_alone(precision):
    """Plain LJ replicas (no walls or planes) of different sizes in one batch: the packed plain
    launches (only the plain path compiled), the packed record steps of PE and virial, and in
    mixed precision the FP32-staged launch.  Positions, velocities and accelerations, and the
    record sums, are bit-identical to each replica advanced alone by the fused loop."""
    from flowmol_amd._abi import FM_PREC_MIXED

    prec = FM_PREC_MIXED if precision == "mixed" else 0
    cases = []
    for k, n in enumerate((6, 8, 7)):
        p, r, v = config.config_fcc((n, n, n), 0.8, 2.5, seed=60 + k, tplot=10, precision=prec)
        r = r + np.random.default_rng(60 + k).uniform(-0.03, 0.03, r.shape)
        cases.append((p, r, v))
    say(f"case: batch of 3 LJ replicas, {precision}, 45 steps")
    ref = []
    for p, r, v in cases:
        c = MDContext(p, r, v)
        c.setup(0)
        c.run_vv(1, 45)
        ref.append(c.state() + (c.evaluate_macroscopic_properties(),))
        c.close()
    b = Batch(graphs=False)
    ctx = []
    for p, r, v in cases:
        c = MDContext(p, r, v)
        c.setup(0)
        ctx.append(c)
        b.add(c, 1)
    b.run(20)
    b.run(25)
    for k, c in enumerate(ctx):
        for x, y in zip(c.state(), ref[k][:3]):
            assert np.array_equal(x, y), k
        assert c.evaluate_macroscopic_properties() == ref[k][3], k
    b.close()
    for c in ctx:
        c.close()


@pytest.mark.parametrize("kind", ["lj", "couette"])
def test_memory_footprint_bounds_the_context(kind):
    """fm_memory_footprint (the server's admission estimate) is not below the device memory a
    context holds after its setup and 60 steps with rebuilds (lists, regions and staging tables
    sized by then): measured as the drop in free device memory, after a first context has paid
    the runtime's one-time cost, within 4 MiB."""
    import torch

    from flowmol_amd.serve import memory_footprint

    if kind == "lj":
        p, r, v = config.config_fcc((20, 20, 20), 0.8, 2.5, seed=81, tplot=10)
        r = r + np.random.default_rng(81).uniform(-0.02, 0.02, r.shape)
        tag = rt = None
    else:
        p, r, v, tag, rt = config.config_couette(nunits=(20, 20, 20), nplanes=16, seed=82, tplot=10)
    say("case: memory footprint vs device memory, %s, %d molecules" % (kind, r.shape[0]))
    torch.cuda.init()
    # the runtime's one-time device memory (code objects and the like): a small context of the
    # same kind, stepped
    if kind == "lj":
        pw, rw, vw = config.config_fcc((8, 8, 8), 0.8, 2.5, seed=80, tplot=10)
        tw = rtw = None
    else:
        pw, rw, vw, tw, rtw = config.config_couette(nunits=(8, 8, 8), nplanes=16, seed=80, tplot=10)
    w = MDContext(pw, rw, vw)
    if tw is not None:
        w.set_tags(tw, rtw)
    w.setup(0)
    w.run_vv(1, 20)
    w.close()
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    g = MDContext(p, r, v)
    if tag is not None:
        g.set_tags(tag, rt)
    g.setup(0)
    g.run_vv(1, 60)
    g.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    used = free0 - free1
    est = memory_footprint(p, r.shape[0])
    say("used %.1f MiB, footprint %.1f MiB" % (used / 2**20, est / 2**20))
    assert used <= est + 4 * 2**20
    g.close()


def test_polymer_batch_members_match_alone():
    """Two FENE + WCA/Mie 14-6 melts of different sizes advanced in one batch: the packed launches
    take each member's branch-free Mie body on the steps where its one-domain launch does, so
    positions, velocities and accelerations are bit-identical to each melt advanced alone (50
    steps, record steps every 10)."""
    states = []
    for nu, seed in (((6, 8, 6), 4), ((8, 8, 6), 5)):
        p, r, v, mt, mono = config.config_polymer(nunits=nu, init="random_walk", seed=seed)
        states.append((p, r, v, mt, mono))
    say("case: batch of 2 polymer melts (1152 and 1536 beads), 50 steps")

    def make(st):
        p, r, v, mt, mono = st
        g = MDContext(p, r, v)
        g.set_moltype(mt)
        g.set_monomers(*mono)
        g.setup(0)
        return g

    ref = []
    for st in states:
        g = make(st)
        g.run_vv(1, 50)
        ref.append(g.state())
        g.close()
    b = Batch()
    ctx = [make(st) for st in states]
    for g in ctx:
        b.add(g, 1)
    b.run(50)
    for g, want in zip(ctx, ref):
        for x, y in zip(g.state(), want):
            assert np.array_equal(x, y)
    b.close()
    for g in ctx:
        g.close()
