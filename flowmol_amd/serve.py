"""Replica serving: many independent NEMD/NVE replicas of uneven size on one GPU.

Continuous batching over libflowmol_hip's replica batches (fm_batch_*):
  * a request is one replica (parameters, initial state, walls, number of steps) with an arrival
    time;
  * the loop admits every request that has arrived while the device footprint of the admitted
    replicas (fm_memory_footprint) stays within the HBM budget;
  * admitted replicas keep their state on the device and advance together, every step one packed
    launch for all of them, in chunks of at most `chunk` steps;
  * a finished replica leaves at the end of its chunk, its results are read back and its
    memory is released, so waiting requests can enter at the next chunk boundary;
  * a replica with snapshot_every > 0 checkpoints every that many steps (flowmol_amd.checkpoint);
    Request.abort() stops a replica at the next chunk boundary with a checkpoint, and
    Request.resumed() continues a checkpoint, bit-identical to the run that took it.
Latency of a replica = time its results are available - its arrival time.
"""
import ctypes as C
import math
import time

import numpy as np

from . import config
from ._abi import FmParams
from ._lib import check, load
from .md import MDContext


def memory_footprint(params, np_):
    b = C.c_int64()
    check(load().fm_memory_footprint(C.byref(params), int(np_), C.byref(b)))
    return b.value


class Batch:
    """Contexts advanced together by packed launches (fm_batch)."""

    def __init__(self, graphs=True):
        h = C.c_void_p()
        check(load().fm_batch_create(C.byref(h)))
        self._h = h
        self.members = []
        if not graphs:
            check(load().fm_batch_use_graphs(self._h, 0))

    def add(self, ctx, iter0):
        check(load().fm_batch_add(self._h, ctx._h, int(iter0)))
        self.members.append(ctx)

    def remove(self, ctx):
        check(load().fm_batch_remove(self._h, ctx._h))
        self.members.remove(ctx)

    def run(self, nsteps):
        nrb = C.c_int32()
        check(load().fm_batch_run(self._h, int(nsteps), C.byref(nrb)))
        return nrb.value

    def counters(self):
        """(member list rebuilds, host synchronisations) since the batch was created."""
        rb, ns = C.c_int64(), C.c_int64()
        check(load().fm_batch_counters(self._h, C.byref(rb), C.byref(ns)))
        return rb.value, ns.value

    def iter_of(self, ctx):
        it = C.c_int32()
        check(load().fm_batch_get_iter(self._h, ctx._h, C.byref(it)))
        return it.value

    def close(self):
        if self._h is not None and self._h.value:
            for c in list(self.members):
                self.remove(c)
            load().fm_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Request:
    """One replica.  snapshot_every > 0: a checkpoint (flowmol_amd.checkpoint) every that many
    steps; abort() stops the replica at the next chunk boundary with a final checkpoint; a request
    made by resumed() continues a checkpoint for the steps left."""

    def __init__(self, rid, arrival, params, r, v, nsteps, tag=None, rtether=None, snapshot_every=0):
        self.rid = rid
        self.arrival = float(arrival)
        self.params = params
        self.r = r
        self.v = v
        self.tag = tag
        self.rtether = rtether
        self.nsteps = int(nsteps)
        self.np = r.shape[0]
        self.footprint = memory_footprint(params, self.np)
        self.snapshot_every = int(snapshot_every)
        self.snapshot = None      # the checkpoint a resumed request starts from
        self.iter0 = 1
        self.aborted = False

    def abort(self):
        self.aborted = True

    @staticmethod
    def resumed(rid, arrival, snap, nsteps, snapshot_every=0):
        q = Request(rid, arrival, snap.params, snap.r, snap.v, nsteps, tag=snap.tag, rtether=snap.rtether,
                    snapshot_every=snapshot_every)
        q.snapshot = snap
        q.iter0 = snap.next_iter
        return q


def couette_request(rid, arrival, n_target, wall_speed, nsteps, seed, density=0.8, tplot=10, precision=0):
    """One Couette NEMD replica: FCC cube of 4*n^3 molecules (n from the target size), tethered
    walls 3 units thick, top wall sliding at wall_speed, MOP on 16 planes every record step."""
    n = max(7, int(round((n_target / 4.0) ** (1.0 / 3.0))))
    p, r, v, tag, rt = config.config_couette(nunits=(n, n, n), density=density, wallslidev=wall_speed,
                                             nplanes=16, seed=seed, tplot=tplot, precision=precision)
    return Request(rid, arrival, p, r, v, nsteps, tag=tag, rtether=rt)


def poisson_workload(n_requests, rate, seed=5, nmin=4096, nmax=262144, nsteps=2000, umin=0.5, umax=2.0,
                     precision=0):
    """Seeded Poisson arrivals; sizes log-uniform in [nmin, nmax]; wall speed uniform in
    [umin, umax]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for k in range(n_requests):
        t += rng.exponential(1.0 / rate)
        size = math.exp(rng.uniform(math.log(nmin), math.log(nmax)))
        u = rng.uniform(umin, umax)
        reqs.append(couette_request(k, t, size, u, nsteps, seed=1000 + k, precision=precision))
    return reqs


class Active:
    def __init__(self, req, ctx, t_admit):
        self.req = req
        self.ctx = ctx
        self.remaining = req.nsteps
        self.done = 0
        self.t_admit = t_admit
        self.last_snapshot = None


def grid_key(req):
    """The parameters that fix a request's cell grid (the brick tuner's table key)."""
    p = req.params
    return (tuple(round(float(p.domain[d]), 12) for d in range(3)), round(float(p.rcutoff), 12),
            round(float(p.delta_rneighbr[0]), 12), req.np)


def make_context(req, brick=None):
    """A request's context with its initial lists and forces; brick: the brick shape to use
    (None: the library's default for its grid)."""
    if req.snapshot is not None:
        from .checkpoint import resume
        return resume(req.snapshot)
    params = req.params
    if brick is not None:
        params = type(params).from_buffer_copy(params)
        params.brick[0], params.brick[1], params.brick[2] = brick
    ctx = MDContext(params, req.r, req.v)
    if req.tag is not None:
        ctx.set_tags(req.tag, req.rtether)
    ctx.setup(0)
    return ctx


def result_of(ctx):
    m = ctx.evaluate_macroscopic_properties()
    out = {"macros": m}
    if ctx.params.vflux_outflag in (1, 2, 3):
        out["mop"] = ctx.mop()
    return out


class ReplicaServer:
    """Continuous-batching request loop for replicas (see the module docstring)."""

    def __init__(self, hbm_budget_bytes=200 * 2**30, chunk=50, graphs=True, bricks=None):
        self.budget = int(hbm_budget_bytes)
        self.chunk = int(chunk)
        self.batch = Batch(graphs=graphs)
        self.bricks = {} if bricks is None else dict(bricks)   # grid_key -> brick shape

    def tune(self, requests, steps=10, warmup=3, max_candidates=3):
        """Brick shapes for the requests' cell grids, chosen by the brick tuner (flowmol_amd.md.
        tune_brick) on fresh contexts of one request per grid, kept for the contexts the server
        creates; returns the table (grid_key -> shape)."""
        from .md import tune_brick

        for q in requests:
            if q.snapshot is not None or grid_key(q) in self.bricks:
                continue

            def prepare(ctx, q=q):
                if q.tag is not None:
                    ctx.set_tags(q.tag, q.rtether)

            best, _ = tune_brick(q.params, q.r, q.v, steps=steps, warmup=warmup, max_candidates=max_candidates,
                                 prepare=prepare)
            self.bricks[grid_key(q)] = best
        return dict(self.bricks)

    def serve(self, requests, realtime=True, on_result=None):
        """Run the requests to completion.  realtime: arrivals follow the wall clock from the
        start of the call; otherwise every request counts as arrived at t=0."""
        pending = sorted(requests, key=lambda q: q.arrival)
        active = []
        used = 0
        results = {}
        t0 = time.perf_counter()
        atom_steps = 0
        batch_steps = 0
        _, syncs0 = self.batch.counters()
        while pending or active:
            now = time.perf_counter() - t0
            # admit what has arrived, in arrival order, while the budget allows
            while pending and (not realtime or pending[0].arrival <= now) and used + pending[0].footprint <= self.budget:
                q = pending.pop(0)
                ctx = make_context(q, self.bricks.get(grid_key(q)))
                self.batch.add(ctx, q.iter0)
                used += q.footprint
                active.append(Active(q, ctx, time.perf_counter() - t0))
            if not active:
                if pending and realtime:
                    time.sleep(max(0.0, min(pending[0].arrival - now, 0.01)))
                    continue
                if pending:
                    raise RuntimeError("a request does not fit the HBM budget on its own")
                break
            # replicas aborted while their chunk ran stop here, with a checkpoint to resume from
            for a in [a for a in active if a.req.aborted]:
                it = self.batch.iter_of(a.ctx)
                self.batch.remove(a.ctx)
                from .checkpoint import take
                snap = take(a.ctx, it)
                t_done = time.perf_counter() - t0
                arrival = a.req.arrival if realtime else 0.0
                res = dict(rid=a.req.rid, np=a.req.np, status="aborted", steps_done=a.done, snapshot=snap,
                           latency=t_done - arrival, t_admit=a.t_admit, t_done=t_done, queue_wait=a.t_admit - arrival)
                results[a.req.rid] = res
                if on_result:
                    on_result(res)
                a.ctx.close()
                used -= a.req.footprint
                active.remove(a)
            if not active:
                continue
            k = min(self.chunk, min(a.remaining for a in active))
            for a in active:
                if a.req.snapshot_every > 0:
                    k = min(k, a.req.snapshot_every - a.done % a.req.snapshot_every)
            self.batch.run(k)
            batch_steps += k
            atom_steps += k * sum(a.req.np for a in active)
            done = [a for a in active if a.remaining - k == 0]
            for a in active:
                a.remaining -= k
                a.done += k
                if a.req.snapshot_every > 0 and a.done % a.req.snapshot_every == 0 and a.remaining > 0:
                    # a periodic checkpoint: out of the batch for the rebuild and the read-back
                    from .checkpoint import take
                    it = self.batch.iter_of(a.ctx)
                    self.batch.remove(a.ctx)
                    a.last_snapshot = take(a.ctx, it)
                    self.batch.add(a.ctx, it)
            for a in done:
                res = result_of(a.ctx)
                if a.req.snapshot is not None and a.req.snapshot.mop is not None and "mop" in res:
                    res["mop"] = a.req.snapshot.mop + res["mop"]  # the tally before the checkpoint
                t_done = time.perf_counter() - t0
                arrival = a.req.arrival if realtime else 0.0
                res.update(rid=a.req.rid, np=a.req.np, status="done", steps_done=a.done, latency=t_done - arrival,
                           t_admit=a.t_admit, t_done=t_done, queue_wait=a.t_admit - arrival,
                           last_snapshot=a.last_snapshot)
                results[a.req.rid] = res
                if on_result:
                    on_result(res)
                self.batch.remove(a.ctx)
                a.ctx.close()
                used -= a.req.footprint
                active.remove(a)
        wall = time.perf_counter() - t0
        lat = np.array([results[k]["latency"] for k in sorted(results) if results[k]["status"] == "done"])
        stats = {"replicas": len(results), "aborted": sum(r["status"] == "aborted" for r in results.values()),
                 "wall_s": wall, "atom_steps": atom_steps,
                 "atom_steps_per_s": atom_steps / wall if wall > 0 else 0.0,
                 "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
                 "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
                 "batch_steps": batch_steps, "host_syncs": self.batch.counters()[1] - syncs0}
        return results, stats

    def close(self):
        self.batch.close()
