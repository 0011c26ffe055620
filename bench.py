#!/usr/bin/env python3
"""Throughput of the velocity Verlet step (md_advance_vv) on one MI355X, in atom-steps/s.

Workload (default): 1,048,576-molecule LJ fluid, FCC 64^3 at density 0.8, rc 2.5,
delta_rneighbr 0.3, dt 0.005, seeded T=1.0 velocities, NVE, displacement-triggered rebuilds
(REBUILD_CRITERIA 2) plus the forced rebuild every TPLOT=10 steps, FP64.

A step is one full md_advance_vv: VV pass 1, rebuild check, (rebuild: sendmols, cell binning,
neighbour list), forces, VV pass 2.  Inputs are resident in HBM before the timed region; the
timed region is exactly --steps steps bracketed by device synchronisation.

Prints ONE JSON line.  The force kernel's average duration is measured with HIP events on the
context's stream over a second pass of the same number of steps right after the timed region
(the events lengthen a step, so they stay out of the pass that gives `value`); cpu_baseline times
the CPU restatement of the reference on a bounded sample of the same workload on this host.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_VALU_PEAK_TFLOPS = 78.6   # MI355X vector FP64
FP32_VALU_PEAK_TFLOPS = 157.3  # MI355X vector FP32


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nunits", type=int, default=64, help="FCC units per side (64 -> 1,048,576 molecules)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--cpu-steps", type=int, default=10, help="steps of the one-core CPU sample (after its setup)")
    ap.add_argument("--no-timing", action="store_true", help="skip per-kernel HIP event timing")
    ap.add_argument("--autotune", action="store_true", help="pick the brick shape with the brick tuner (fresh contexts) first")
    ap.add_argument("--drift-steps", type=int, default=10000, help="drift: NVE steps per precision")
    ap.add_argument("--c3-steps", type=int, default=100, help="c3: Couette steps per precision")
    ap.add_argument("--c4-steps", type=int, default=200, help="c4: polymer steps (after as many warm-up steps / 10)")
    ap.add_argument("--workload", choices=["c2", "serve", "drift", "c3", "c4"], default="c2",
                    help="c2: 1M-molecule LJ step throughput (default); serve: replica stream latency")
    ap.add_argument("--replicas", type=int, default=48, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=6.0, help="serve: Poisson arrival rate (1/s)")
    ap.add_argument("--replica-steps", type=int, default=2000, help="serve: steps per replica")
    ap.add_argument("--serve-graphs", choices=["on", "off"], default="off",
                    help="serve: replay each window's packed launches (after its rebuilds) as one HIP graph")
    ap.add_argument("--serve-tune", choices=["on", "off"], default="on",
                    help="serve: brick shape per replica grid from the brick tuner, before the stream starts")
    ap.add_argument("--brick", default="", help="c2, c3, c4: brick shape BX,BY,BZ (cells) instead of the default (c4: instead of the tuner)")
    ap.add_argument("--extras", choices=["on", "off"], default="on",
                    help="c2 (FP64): also the mixed-precision run and the replica stream, in the same JSON line")
    ap.add_argument("--c4-tune", choices=["on", "off"], default="on",
                    help="c4: brick shape from the brick tuner (fresh contexts) first")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run a plain device-to-device copy of this size (PMC byte calibration)")
    return ap.parse_args()


def copy_calibration(mb):
    """A plain device copy of known size through the HIP runtime (its blit kernel calibrates the
    FETCH_SIZE / WRITE_SIZE counters)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mb * 2**20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def traffic_from_profile():
    """HBM bytes per force launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE), each
    scaled by the same counter on a device copy of known size; the summary with its commands is
    profiles/forces_c2_traffic.md."""
    path = os.path.join(ROOT, "profiles", "forces_c2_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch"):
            return int(line.split(":")[1].split()[0])
    return None


def serve_main(a):
    print(json.dumps(serve_result(a)), flush=True)


def serve_result(a):
    """The replica stream (batched serving of uneven Couette replicas): Poisson arrivals,
    continuous batching; p50/p99 latency and aggregate throughput."""
    import numpy as np

    from flowmol_amd.serve import ReplicaServer, poisson_workload

    reqs = poisson_workload(a.replicas, a.rate, seed=5, nsteps=a.replica_steps)
    # warm-up: one small replica through the same path (kernel loading, graph capture)
    warm = poisson_workload(1, 100.0, seed=99, nmin=4096, nmax=4096, nsteps=50)
    srv = ReplicaServer(chunk=50, graphs=a.serve_graphs == "on")
    t_tune = 0.0
    if a.serve_tune == "on":
        t_tune = time.perf_counter()
        srv.tune(reqs)
        t_tune = time.perf_counter() - t_tune
    srv.serve(warm, realtime=False)
    drift = []

    def on_result(res):
        drift.append(res["macros"]["totenergy"])

    results, stats = srv.serve(reqs, realtime=True, on_result=on_result)
    srv.close()
    sizes = [q.np for q in reqs]
    out = {
        "metric": "ensemble replica p50/p99 latency (Couette NEMD replica stream, 1 GPU)",
        "value": stats["p99_latency_s"],
        "unit": "s (p99 replica latency)",
        "higher_is_better": False,
        "n_gpus": 1,
        "p50_latency_s": stats["p50_latency_s"],
        "p99_latency_s": stats["p99_latency_s"],
        "aggregate_atom_steps_per_s": stats["atom_steps_per_s"],
        "wall_s": stats["wall_s"],
        "replicas": stats["replicas"],
        "batch_steps": stats["batch_steps"],
        "brick_tuning": {"on": a.serve_tune == "on", "grids": len(srv.bricks), "seconds_before_stream": t_tune},
        "host_syncs_per_batch_step": stats["host_syncs"] / max(1, stats["batch_steps"]),
        "dtype": "f64",
        "data": "synthetic",
        "vs_baseline": None,
        "config": {"workload": f"{a.replicas} Couette replicas, Poisson arrivals rate {a.rate}/s (seed 5), "
                               f"sizes log-uniform 4096-262144 (FCC cubes), U in [0.5,2.0], "
                               f"{a.replica_steps} steps each, MOP 16 planes every 10 steps, FP64",
                   "mean_molecules": float(np.mean(sizes)), "chunk_steps": 50},
    }
    return out


def drift_main(a):
    """NVE total-energy drift of the 1M-molecule LJ workload over --drift-steps steps, FP64 and
    mixed precision from the same start (sums every TPLOT=10 steps)."""
    import numpy as np

    from flowmol_amd import config
    from flowmol_amd._abi import FM_PREC_FP64, FM_PREC_MIXED
    from flowmol_amd.md import MDContext

    n_u, K = a.nunits, a.drift_steps
    res = {}
    for name, prec in (("fp64", FM_PREC_FP64), ("mixed", FM_PREC_MIXED)):
        p, r, v = config.config_fcc((n_u, n_u, n_u), 0.8, 2.5, seed=2, tplot=10, precision=prec)
        N = r.shape[0]
        g = MDContext(p, r, v)
        g.setup(0)
        m0 = g.evaluate_macroscopic_properties()
        g.synchronize()
        t0 = time.perf_counter()
        ms, nrb = g.run_vv(1, K, n_macros=K // 10)
        g.synchronize()
        dt = time.perf_counter() - t0
        e = np.array([m0["totenergy"]] + [m["totenergy"] for m in ms])
        t = np.arange(len(e)) * 10 * p.delta_t
        slope = np.polyfit(t, e, 1)[0]
        res[name] = {"atom_steps_per_s": N * K / dt, "rebuilds": nrb, "e0": e[0], "e_end": e[-1],
                     "max_rel_dev": float(np.max(np.abs(e - e[0])) / abs(e[0])),
                     "drift_per_time_unit": float(slope / abs(e[0]))}
        g.close()
    out = {
        "metric": f"NVE total-energy drift over {K} steps (1M LJ, 1 GPU)",
        "value": res["fp64"]["max_rel_dev"],
        "unit": "max |E(t)-E(0)|/|E(0)|",
        "higher_is_better": False,
        "n_gpus": 1,
        "steps": K,
        "fp64": res["fp64"],
        "mixed": res["mixed"],
        "mixed_over_fp64_drift": res["mixed"]["max_rel_dev"] / max(res["fp64"]["max_rel_dev"], 1e-300),
        "dtype": "f64 (and f32-pair/f64-state)",
        "data": "synthetic",
        "vs_baseline": None,
        "config": {"workload": f"LJ NVE FCC {n_u}^3, density 0.8, rc 2.5, delta 0.3, dt 0.005, T 1.0, TPLOT 10"},
    }
    print(json.dumps(out), flush=True)


def c4_main(a):
    """65,536 FENE chains of 8 beads (524,288 beads) grown as seeded random walks in the box of
    the 64x32x64-unit FCC lattice at density 0.85 (dt 0.002), WCA between all beads and Mie 14-6
    between the beads of the second species (half of the chains); velocity Verlet, displacement
    rebuild, TPLOT 10, FP64."""
    from flowmol_amd import config
    from flowmol_amd.md import MDContext

    t_gen = time.perf_counter()
    p, r, v, moltype, mono = config.config_polymer(nunits=(64, 32, 64), density=0.85, nmono=8, seed=4,
                                                   init="random_walk")
    t_gen = time.perf_counter() - t_gen
    N = r.shape[0]

    def prepare(ctx):
        ctx.set_moltype(moltype)
        ctx.set_monomers(*mono)

    brick = "default"
    if a.brick:
        p.brick[0], p.brick[1], p.brick[2] = (int(x) for x in a.brick.split(","))
        brick = "given %s" % a.brick
    elif a.autotune or a.c4_tune == "on":
        from flowmol_amd.md import tune_brick
        best, tuned = tune_brick(p, r, v, prepare=prepare, max_candidates=6)
        p.brick[0], p.brick[1], p.brick[2] = best
        brick = "autotuned %dx%dx%d (%s)" % (best + (", ".join("%dx%dx%d %.3f ms" % (b + (t * 1e3,))
                                                               for b, t in sorted(tuned.items(), key=lambda x: x[1])),))
    g = MDContext(p, r, v)
    prepare(g)
    g.setup(0)
    K = a.c4_steps
    W = max(1, K // 10)
    g.run_vv(1, W)
    g.synchronize()
    t0 = time.perf_counter()
    _, nrb = g.run_vv(1 + W, K)
    g.synchronize()
    dt = time.perf_counter() - t0
    g.close()
    print(json.dumps({
        "metric": "atom-steps/s (524k-bead FENE chains, WCA + Mie 14-6, 1 GPU)",
        "value": N * K / dt, "unit": "atom-steps/s", "higher_is_better": True, "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": dt / K * 1e3, "rebuilds_in_timed_steps": nrb,
        "dtype": "f64", "data": "synthetic", "vs_baseline": None,
        "config": {"workload": "FENE k 30 R0 1.5, 65,536 chains of 8 beads, random-walk initial state from seed 4 "
                               "(bond 0.97; close non-bonded pairs relaxed to >= 0.8) in the box of FCC 64x32x64 "
                               "at density 0.85, "
                               "WCA all pairs, Mie 14-6 (rc 2.5) within species 1 (half the chains), dt 0.002, "
                               "TPLOT 10, displacement rebuild", "molecules": N,
                   "init_generation_s": round(t_gen, 1), "brick": brick},
    }), flush=True)


def c3_main(a):
    """Couette flow between tethered FCC walls (3 lattice units, top wall at U = 1), MOP stress on
    64 planes normal to y every step, leap-frog (the reference's tag_move path): FP64 and mixed
    precision from the same start; the mixed run's force and MOP errors against FP64."""
    import numpy as np

    from flowmol_amd import config
    from flowmol_amd._abi import FM_ENSEMBLE_NVE, FM_INTEGRATION_LFV, FM_PREC_FP64, FM_PREC_MIXED
    from flowmol_amd.md import MDContext

    n_u, K = a.nunits, a.c3_steps
    res, mop, end = {}, {}, {}
    for name, prec in (("fp64", FM_PREC_FP64), ("mixed", FM_PREC_MIXED)):
        p, r, v, tag, rt = config.config_couette(nunits=(n_u, n_u, n_u), nplanes=64, tplot=1, precision=prec)
        p.integration_algorithm = FM_INTEGRATION_LFV
        if a.brick:
            p.brick[0], p.brick[1], p.brick[2] = (int(x) for x in a.brick.split(","))
        N = r.shape[0]
        g = MDContext(p, r, v)
        g.set_tags(tag, rt)
        g.setup(0)
        g.reset_mop()
        g.synchronize()
        t0 = time.perf_counter()
        _, nrb = g.run_lfv(0, K)
        g.synchronize()
        dt = time.perf_counter() - t0
        mop[name] = g.mop()
        end[name] = (p, g.state())
        res[name] = {"atom_steps_per_s": N * K / dt, "rebuilds": nrb, "ms_per_step": dt / K * 1e3}
        g.close()
    # pair-force error on the sheared state the FP64 run reached (the lattice start has zero
    # forces); tethers act in the leap-frog move, not in the force call, so NVE parameters serve
    acc = {}
    r1, v1, _ = end["fp64"][1]
    for name in ("fp64", "mixed"):
        q = type(end[name][0]).from_buffer_copy(end[name][0])
        q.ensemble = FM_ENSEMBLE_NVE
        g = MDContext(q, r1, v1)
        g.setup(0)
        acc[name] = g.state()[2]
        g.close()
    a64, a32 = acc["fp64"], acc["mixed"]
    force_rms = float(np.sqrt(((a32 - a64) ** 2).sum() / (a64 ** 2).sum()))
    pxy64, pxy32 = mop["fp64"][:, 0], mop["mixed"][:, 0]
    mop_err = float(np.linalg.norm(pxy32 - pxy64) / np.linalg.norm(pxy64))
    out = {
        "metric": "atom-steps/s (1M Couette, tethered walls, MOP 64 planes every step, 1 GPU)",
        "value": res["fp64"]["atom_steps_per_s"],
        "unit": "atom-steps/s",
        "higher_is_better": True,
        "n_gpus": 1,
        "steps": K,
        "fp64": res["fp64"],
        "mixed": res["mixed"],
        "mixed_force_rms_rel_error": force_rms,
        "mixed_mop_pxy_rel_error": mop_err,
        "dtype": "f64",
        "dtype_note": "every step is a record step with the MOP tally, which both precisions compute in FP64; "
                      "the mixed-precision context's run is beside it",
        "data": "synthetic",
        "vs_baseline": None,
        "config": {"workload": f"Couette FCC {n_u}^3 (density 0.8, rc 2.5), walls 3 units tethered "
                               f"(k4 5000, k6 5e6), top wall U = 1.0, leap-frog dt 0.005, TPLOT 1 "
                               f"(record and rebuild every step), MOP 64 planes"},
    }
    print(json.dumps(out), flush=True)


def mixed_result(a, logged64):
    """The same 1M LJ run in mixed precision (FP32 pair law, FP64 state and sums): throughput
    over the same steps with the same sums, the force error against FP64 on a plain and a
    record step of a perturbed lattice, and the largest relative difference of the logged total
    energies from the FP64 run's."""
    import numpy as np

    from flowmol_amd import config
    from flowmol_amd._abi import FM_PREC_MIXED
    from flowmol_amd.md import MDContext

    n_u = a.nunits
    p64, r, v = config.config_fcc((n_u, n_u, n_u), 0.8, 2.5, seed=2, tplot=10)
    p32, _, _ = config.config_fcc((n_u, n_u, n_u), 0.8, 2.5, seed=2, tplot=10, precision=FM_PREC_MIXED)
    N = r.shape[0]
    err = {}
    # the force error on the lattice moved off its sites by a seeded uniform offset in
    # [-0.05, 0.05] per component (on the perfect lattice the forces cancel to rounding)
    rp = r + np.random.default_rng(2).uniform(-0.05, 0.05, r.shape)
    g64 = MDContext(p64, rp, v)
    g32 = MDContext(p32, rp, v)
    g64.setup(0)
    g32.setup(0)
    for name, it in (("record", 10), ("plain", 3)):
        g64.simulation_compute_forces(it)
        g32.simulation_compute_forces(it)
        a64, a32 = g64.state()[2], g32.state()[2]
        err[name] = float(np.sqrt(((a32 - a64) ** 2).sum() / (a64 ** 2).sum()))
    g64.close()
    g32.close()
    g = MDContext(p32, r, v)
    g.setup(0)
    g.run_vv(1, a.warmup)
    g.synchronize()
    nrec = a.steps // 10 + 1
    t0 = time.perf_counter()
    macros, _ = g.run_vv(1 + a.warmup, a.steps, n_macros=nrec)
    g.synchronize()
    dt = time.perf_counter() - t0
    g.close()
    logged = [m for m in macros if m["msum"] > 0]
    de = max((abs(m["totenergy"] - m64["totenergy"]) / abs(m64["totenergy"]) for m, m64 in zip(logged, logged64)),
             default=None)
    return {"value": N * a.steps / dt, "unit": "atom-steps/s", "ms_per_step": dt / a.steps * 1e3,
            "dtype": "f32-pair/f64-state", "force_rms_rel_error_plain_step": err["plain"],
            "force_rms_rel_error_record_step": err["record"], "max_rel_totenergy_difference": de,
            "energy_records_compared": min(len(logged), len(logged64))}


def _oracle_run(args):
    """One oracle process: setup, then `steps` md_advance_vv steps with the record-step sums every
    TPLOT steps; returns (molecules, seconds of the time loop)."""
    nunits, steps, seed = args
    from flowmol_amd import config
    from oracle.oracle import Oracle

    p, r, v = config.config_fcc(nunits, 0.8, 2.5, seed=seed, tplot=10)
    o = Oracle(p, r, v)
    o.setup()
    o.compute_forces(0)
    t0 = time.perf_counter()
    o.run_vv(1, steps, n_macros=steps // 10 + 1)
    return r.shape[0], time.perf_counter() - t0


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(nunits, steps):
    """The CPU restatement of the reference (oracle) on this host's cores, two legs:
    * one core on the same 1M-molecule workload (`steps` steps after its setup);
    * every core: C processes at once, each on a 1/C share of the molecules (an FCC box of
      64 x 64/cy x 64/cz units with cy*cz = C, the share one rank of a C-rank domain decomposition
      holds), aggregate
      atom-steps/s = sum of molecules x steps / the slowest process's time.  No halo exchange is
      modelled, so this leg bounds an MPI run of the reference from above."""
    import multiprocessing as mp

    n1, t1 = _oracle_run(((nunits, nunits, nunits), steps, 2))
    one = n1 * steps / t1
    C = host_cores()
    # C = cy * cz sub-boxes, each at least 12 units (4 cells) along every side
    split = None
    while C > 1 and split is None:
        for cy in range(int(math.isqrt(C)), 0, -1):
            cz = C // cy
            if cy * cz == C and nunits % cy == 0 and nunits % cz == 0 and nunits // max(cy, cz) >= 12:
                split = (cy, cz)
                break
        if split is None:
            C -= 1
    cy, cz = split if split else (1, 1)
    share = (nunits, nunits // cy, nunits // cz)
    steps_c = max(steps, 10 * (int(1.0e7 / (4 * nunits * nunits * (nunits // C)) + 9) // 10))
    with mp.get_context("spawn").Pool(C) as pool:
        res = pool.map(_oracle_run, [(share, steps_c, 100 + k) for k in range(C)])
    n_all = sum(n for n, _ in res)
    t_all = max(t for _, t in res)
    allc = n_all * steps_c / t_all
    return {"value": allc, "unit": "atom-steps/s", "cores": C, "kind": "port",
            "sample": f"{C} concurrent single-threaded processes of the C restatement (gcc -O3), each on "
                      f"a {4 * share[0] * share[1] * share[2]}-molecule share (FCC {share[0]}x{share[1]}x{share[2]}) "
                      f"for {steps_c} md_advance_vv steps with sums every 10, {t_all:.1f} s; no halo exchange",
            "single_core": {"value": one, "cores": 1,
                            "sample": f"the {n1}-molecule workload, {steps} steps after setup, {t1:.1f} s"}}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (the path has no multi-GPU form here); got --gpus {a.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration_mb:
        copy_calibration(a.copy_calibration_mb)
    if a.workload == "serve":
        serve_main(a)
        return
    if a.workload == "drift":
        drift_main(a)
        return
    if a.workload == "c3":
        c3_main(a)
        return
    if a.workload == "c4":
        c4_main(a)
        return
    import numpy as np

    from flowmol_amd import config
    from flowmol_amd._abi import FM_PREC_FP64, FM_PREC_MIXED
    from flowmol_amd.md import MDContext

    prec = FM_PREC_MIXED if a.precision == "mixed" else FM_PREC_FP64
    n_u = a.nunits
    p, r, v = config.config_fcc((n_u, n_u, n_u), 0.8, 2.5, seed=2, tplot=10, precision=prec)
    N = r.shape[0]
    brick = None
    if a.brick and not a.autotune:
        p.brick[0], p.brick[1], p.brick[2] = (int(x) for x in a.brick.split(","))
        brick = tuple(p.brick)
    if a.autotune:
        from flowmol_amd.md import tune_brick

        brick, tuned = tune_brick(p, r, v)
        print("brick tuner (s/step): " + ", ".join("%dx%dx%d %.3e" % (b + (t,)) for b, t in tuned.items()),
              file=sys.stderr, flush=True)
        p.brick[0], p.brick[1], p.brick[2] = brick
    g = MDContext(p, r, v)
    g.setup(0)
    g.run_vv(1, a.warmup)
    g.synchronize()
    it0 = 1 + a.warmup
    # the energy/virial log: the record-step sums every TPLOT = 10 steps, reduced on the device
    nrec = a.steps // 10 + 1
    t0 = time.perf_counter()
    macros, nrb = g.run_vv(it0, a.steps, n_macros=nrec)
    g.synchronize()
    t1 = time.perf_counter()
    logged = [m for m in macros if m["msum"] > 0]
    dt = t1 - t0
    value = N * a.steps / dt
    # kernel durations: the next K steps again, with HIP events around every launch on the
    # context's stream (the events themselves lengthen a step by several per cent, so the pass
    # that gives `value` runs without them)
    kt = {}
    if not a.no_timing:
        g.kernel_timing(True)
        g.run_vv(it0 + a.steps, a.steps, n_macros=nrec)
        g.synchronize()
        kt = g.get_kernel_timing()

    # neighbour statistics of the current list (full list: two entries per pair)
    import ctypes as C
    from flowmol_amd._lib import load
    cnt = C.c_int64()
    load().fm_download_neighbourlist(g._h, None, None, None, 0, C.byref(cnt))
    p_half = cnt.value / 2.0 / N
    geo = g.geometry()
    rn = math.sqrt(geo["rneighbr2"])
    p_rc = p_half * (2.5 / rn) ** 3     # uniform-density share of list pairs inside rc

    roof = None
    # the dominant kernel: the force launch that also applies the next step's VV pass 1 (most
    # steps), else the plain force launch
    total = {k: kt[k][0] * kt[k][1] for k in ("forces", "forces_pass1") if k in kt and kt[k][1] > 0}
    dom = max(total, key=total.get) if total else None
    if dom:
        avg_s = kt[dom][0] * 1e-3
        if dom == "forces_pass1":
            # one whole step's streaming: r, v, a read and written, rdisp read and written, list
            bytes_per = N * (144 + 48 + 4 * p_half)
            flops_per = N * (8 * p_half + 16 * p_rc + 24 + 6)   # pairs, VV, displacement check
        else:
            bytes_per = N * (24 + 24 + 4 * p_half)      # r read, a written, list words read
            flops_per = N * (8 * p_half + 16 * p_rc)    # per half-list pair: 8 (+16 inside rc)
        gbs = bytes_per / avg_s / 1e9
        tfl = flops_per / avg_s / 1e12
        peak_tf = FP64_VALU_PEAK_TFLOPS if prec == FM_PREC_FP64 else FP32_VALU_PEAK_TFLOPS
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile(),
                "kernel": dom, "avg_ms": round(kt[dom][0], 5), "launches": kt[dom][1],
                "bytes_per_launch": round(bytes_per), "flops_per_launch": round(flops_per),
                "valu": {"achieved": round(tfl, 2), "peak": peak_tf, "unit": "TFLOP/s",
                         "frac": round(tfl / peak_tf, 4),
                         "dtype": "f64" if prec == FM_PREC_FP64 else "f32"}}
    out = {
        "metric": "atom-steps/s (1M LJ, 1 GPU)",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == FM_PREC_FP64 else "f32-pair/f64-state",
        "data": "synthetic",
        "config": {"workload": f"LJ NVE FCC {n_u}^3 ({N} molecules), density 0.8, rc 2.5, delta 0.3, "
                               f"dt 0.005, T 1.0, TPLOT 10, displacement rebuild, energy/virial sums every "
                               f"10 steps",
                   "molecules": N, "rebuilds_in_timed_steps": nrb, "p_half": round(p_half, 2),
                   "energy_log": {"records": len(logged),
                                  "totenergy_first": logged[0]["totenergy"] if logged else None,
                                  "totenergy_last": logged[-1]["totenergy"] if logged else None,
                                  "pressure_last": logged[-1]["pressure"] if logged else None},
                   "brick": ("autotuned %dx%dx%d" % brick) if a.autotune else ("%dx%dx%d" % brick if brick else "default"),
                   "precision": a.precision},
        "kernels_ms": {k: round(vv[0], 5) for k, vv in kt.items() if vv[1] > 0},
        "roofline": roof,
        "cpu_baseline": None,
    }
    g.close()
    if a.extras == "on" and prec == FM_PREC_FP64:
        # the rest of the headline metric, measured in the same run: the mixed-precision form of
        # this workload, and the replica stream's p50/p99 latency
        out["mixed_precision"] = mixed_result(a, logged)
        sv = serve_result(a)
        out["replica_serving"] = {k: sv[k] for k in ("p50_latency_s", "p99_latency_s", "aggregate_atom_steps_per_s",
                                                     "replicas", "host_syncs_per_batch_step", "config")}
    if a.cpu_baseline == "on":
        out["cpu_baseline"] = cpu_baseline(n_u, a.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
