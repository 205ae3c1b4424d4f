#!/usr/bin/env python3
"""Single-GPU benchmark of the LLG step on MI355X.

Workload: cuboid of 256 x 256 x 16 cells (4 nm), Ms 1e6 A/m, A 1.5e-11 J/m,
Ku1 5e5 J/m^3 along the thickness, alpha 0.1, seeded random m0, zero field,
fixed-dt (1e-14 s) RK4, demag+exchange convolution padded to 512 x 512 x 32.
One "step" is one RK4 step (4 field evaluations) of the whole grid, with the
state resident in HBM before the timed region starts.

Prints ONE JSON line.  `value` is Mcell-steps/s (cells x steps / s) of the
timed region, measured with HIP events on the engine's own stream.  The
dominant kernel class is the one with the largest total time in a profiled
pass (events around every launch); the `roofline` object reports its
algorithmic bytes per launch over its average launch time, measured as n
back-to-back launches of that pass between two HIP events on the engine's
stream.  `traffic` is the HBM bytes per launch of that kernel from the
committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary for this workload (scaled
by the copy-kernel calibration recorded there), or null if none matches.
`cpu_baseline` is the FP64 CPU oracle timed on this host on a bounded sample
of the same workload; `parity` is the GPU engine against that oracle run
(max |dm|, relative energy error) from the same m0 for the same steps.

The same line carries two more objects, each against the same simulations run
one at a time on this GPU: `sweep` (512 independent simulations of the sweep
workload advanced in batches; serial rate from an evenly spaced sample of 32
members) and `serve` (a Poisson stream of 200 requests served by continuous
batching; the serial FIFO from an evenly spaced sample of 40 requests).
`--no-sweep` / `--no-serve` leave them out; `--workload sweep|serve` runs one
of them alone with the full serial side.
"""
import argparse
import ctypes as C
import importlib
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
HERE = os.path.dirname(os.path.abspath(__file__))


def committed_traffic(kernel_class, workload):
    """HBM bytes per launch from profiles/*traffic*.md: a line
    `traffic_bytes_per_launch <class> <workload> <bytes>`."""
    d = os.path.join(HERE, "profiles")
    if not os.path.isdir(d):
        return None
    for fn in sorted(os.listdir(d)):
        if "traffic" not in fn:
            continue
        with open(os.path.join(d, fn)) as f:
            for line in f:
                t = line.strip().strip("`").split()
                if len(t) == 4 and t[0] == "traffic_bytes_per_launch" and t[1] == kernel_class and t[2] == workload:
                    return int(t[3])
    return None
CLASSES = ["z_fwd_r2c", "y_fwd", "x_fwd_kmul_x_inv", "y_inv", "z_inv_c2r_torque", "torque", "rk_final", "control"]


def model_bytes(N, P, Nx, Ny, nK, plane=False):
    """Algorithmic bytes of one field evaluation, per kernel class, and of one
    RK4 step (B_step = 4 B_conv + 12 N 27, the fixed model of the reference's
    pass structure).  The pointwise RK work (12 N 27 per step: stage builds,
    torques, final combination) runs in the epilogue of the Z inverse pass, so
    that class carries a quarter of it per evaluation and the classes of the 4
    evaluations add up to B_step.  With the plane pass (`plane`), Y forward,
    X.K.X and Y inverse are one pass that reads and writes B1 once and the
    kernel (48 A1 + 4 nK A3, the kernel counted uncompressed as in the model);
    `b_passes` is then the bytes of one step through these passes."""
    Q = P[2] // 2 + 1
    A1 = Nx * Ny * Q
    A2 = Nx * P[1] * Q
    A3 = P[0] * P[1] * Q
    per = {
        "z_fwd_r2c": 3 * (4 * N + 8 * A1),
        "y_fwd": 3 * (8 * A1 + 8 * A2),
        "x_fwd_kmul_x_inv": 3 * 16 * A2 + 4 * nK * A3,
        "y_inv": 3 * (8 * A2 + 8 * A1),
        "z_inv_c2r_torque": 3 * (8 * A1 + 4 * N) + 12 * N * 27 // 4,
    }
    b_conv = 3 * (8 * N + 32 * A1 + 32 * A2) + 4 * nK * A3
    b_step = 4 * b_conv + 12 * N * 27
    b_passes = b_step
    if plane:
        del per["y_fwd"], per["y_inv"], per["x_fwd_kmul_x_inv"]
        per["y_xkx_y_plane"] = 3 * 16 * A1 + 4 * nK * A3
        b_passes = 4 * (3 * (8 * N + 32 * A1) + 4 * nK * A3) + 12 * N * 27
    return per, b_conv, b_step, b_passes


def cpu_baseline(cfg, seconds):
    """FP64 oracle (C, OpenMP) on the same workload: time whole RK4 steps until
    about `seconds` of CPU work have run (at least one step)."""
    from oracle import Oracle

    orc = Oracle("f64")
    kw = importlib.import_module("1_amd").workloads.oracle_kwargs(cfg)
    t0 = time.perf_counter()
    s = orc.sim(cfg["m0"], **kw)
    setup = time.perf_counter() - t0
    n = 0
    t1 = time.perf_counter()
    while True:
        s.step()
        n += 1
        el = time.perf_counter() - t1
        if el >= seconds or (el / n) * (n + 1) > 2.0 * seconds:
            break
    N = int(np.prod(cfg["size"]))
    # the FP32 transcription of the same path (the oracle compiled in float), a
    # shorter sample
    o32 = Oracle("f32").sim(cfg["m0"], **kw)
    n32 = 0
    t2 = time.perf_counter()
    while True:
        o32.step()
        n32 += 1
        el32 = time.perf_counter() - t2
        if el32 >= seconds / 3 or (el32 / n32) * (n32 + 1) > seconds:
            break
    del o32
    cores = used_cores()
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": N * n / el / 1e6,
        "unit": "Mcell-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": "%d RK4 step(s) of the same 256x256x16 workload in %.1f s (FP64 oracle, OpenMP, kernel setup "
                  "%.1f s excluded); cpu: %s" % (n, el, setup, model),
        "fp32_transcription": {"value": N * n32 / el32 / 1e6, "unit": "Mcell-steps/s",
                               "sample": "%d RK4 step(s) in %.1f s (the oracle compiled in float32)" % (n32, el32)},
    }, s, n


def used_cores():
    """Threads the OpenMP oracle runs on: OMP_NUM_THREADS if set, bounded by
    the CPUs this process may run on (its affinity mask)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(omp, avail) if omp > 0 else avail


def member_parity(mm, cfg, batched_m, seconds):
    """One member of a batched workload against the FP64 oracle: its final m
    in the batch against the same simulation run alone (bit for bit), and the
    first `seconds` of it run alone with its trial log (dt, accepted) replayed
    in the oracle: max |dm| and the relative energy error at the end."""
    from oracle import Oracle

    s = mm.workloads.make_sim(cfg)
    s.run(cfg["duration"])
    same = bool(np.array_equal(s.m.view(np.uint32), batched_m.view(np.uint32)))
    s.free()
    s = mm.workloads.make_sim(cfg)
    s.run(seconds)
    dts, acc = s.trial_log()
    replay = []
    for prec in ("f64", "f32"):  # the FP64 oracle, and its FP32 twin on the same trials
        o = Oracle(prec).sim(cfg["m0"], **mm.workloads.oracle_kwargs(cfg))
        i = 0
        while i < len(dts):
            j = i
            while j < len(dts) - 1 and not acc[j]:
                j += 1
            o.step_replay(dts[i:j + 1], acc[i:j + 1])
            i = j + 1
        replay.append(o)
    o, o32 = replay
    e_gpu, e_o = s.energy(), o.energy()
    out = {"member": cfg["name"], "size_internal_xyz": list(cfg["size"]), "batch_equals_alone_bitwise": same,
           "vs": "FP64 CPU oracle replaying the member's first %.3g ns alone (%d trials, %d steps)" % (
               seconds * 1e9, len(dts), int(s.state()["steps"])),
           "max_abs_dm": float(np.abs(s.m.reshape(-1) - o.m).max()),
           "rel_energy_err": float(abs(e_gpu - e_o) / abs(e_o)),
           "fp32_twin_max_abs_dm": float(np.abs(o32.m - o.m).max())}
    s.free()
    return out


def parity_vs_oracle(mm, cfg, o, n):
    """The GPU engine run from the same m0 for the same n RK4 steps as the
    oracle sim `o` of the CPU leg: max |dm| and the relative energy error of
    the final state (FP64 device sum against the FP64 oracle)."""
    s = mm.workloads.make_sim(cfg)
    s.steps(n)
    m = s.m
    e_gpu = s.energy()
    s.free()
    e_o = o.energy()
    return {"vs": "FP64 CPU oracle, same m0 and dt, %d RK4 step(s)" % n,
            "max_abs_dm": float(np.abs(m - o.m.reshape(m.shape)).max()),
            "rel_energy_err": float(abs(e_gpu - e_o) / abs(e_o))}


def calibrate_copy(lib, n=32 << 20, reps=3):
    """A streaming kernel of known traffic for scaling the FETCH_SIZE /
    WRITE_SIZE counters: gpu_linear_combination(a, b, 1, 2, n) reads a and b
    and writes a, i.e. 8n bytes fetched and 4n bytes written per launch."""
    a = lib.new_gpu_array(n)
    b = lib.new_gpu_array(n)
    for _ in range(reps):
        lib.gpu_linear_combination(a, b, 1.0, 2.0, n)
    lib.free_gpu_array(a)
    lib.free_gpu_array(b)
    return {"kernel": "gpu_linear_combination", "launches": reps, "fetch_bytes": 8 * n, "write_bytes": 4 * n}


def dp_trial_bytes(size):
    """Algorithmic bytes of one Dormand-Prince trial of one simulation with an
    FSAL torque in hand: 6 field evaluations (B_conv each) plus the pointwise
    RK work 12 N 58 (stage builds, torques, final and error combination,
    normalisation)."""
    X, Y, Z = size
    N = X * Y * Z
    P = [2 * X if X > 1 else 1, 2 * Y, 2 * Z]
    Q = P[2] // 2 + 1
    nK = 6 if X > 1 else 4
    A1, A2, A3 = X * Y * Q, X * P[1] * Q, P[0] * P[1] * Q
    if X > 1:
        b_conv = 3 * (8 * N + 32 * A1 + 32 * A2) + 4 * nK * A3
    else:
        b_conv = 3 * (8 * N + 32 * A1) + 4 * nK * A3
    return 6 * b_conv + 12 * N * 58


def run_groups(mm, sims, groups, poll, graph, clock, t1, done_at, lane_factor=3.0, lane_floor=1024, lane_after=4):
    """Advance simulations (already started) in size groups, each in its own
    batch on its own stream, `poll` trials per asynchronous enqueue, until all
    are done; a group's batch is rebuilt from its running members when half
    have finished.  After `lane_after` enqueues of a group, its members whose
    remaining trials (``mm_batch_remaining``) exceed `lane_factor` times the
    group's median move to that group's lane, a batch of their own (as
    serve.Server does).  Fills done_at (seconds since t1); returns (batched
    trials, rebuilds, members moved to lanes)."""
    sv = importlib.import_module("1_amd.serve")

    def group(members, lane=None):
        # members: the group's simulations; bm: those its current batch holds
        return {"members": np.asarray(members, dtype=int), "bm": None, "batch": None, "busy": False, "dirty": True,
                "polls": 0, "lane": lane}

    state = []
    for members in groups:
        lane = group([]) if lane_factor else None
        state.append(group(members, lane))
        if lane is not None:
            state.append(lane)
    trials = rebuilds = moved = 0
    while any(len(g["members"]) or g["busy"] for g in state):
        launched = False
        for g in state:
            if g["busy"]:
                if not g["batch"].ready():
                    continue
                d = g["batch"].flags()
                g["busy"] = False
                fin = g["bm"][d]
                done_at[fin[np.isnan(done_at[fin])]] = clock() - t1
                if 2 * int(d.sum()) >= len(d):
                    g["dirty"] = True
                if g["lane"] is not None and g["polls"] >= lane_after:
                    rem = g["batch"].remaining()
                    rem[d] = 0
                    pick = sv.long_runners(rem, lane_factor, lane_floor)
                    if pick:
                        lane = g["lane"]
                        go = g["bm"][pick]
                        lane["members"] = np.concatenate([lane["members"], go])
                        g["members"] = g["members"][~np.isin(g["members"], go)]
                        g["dirty"] = lane["dirty"] = True
                        moved += len(pick)
            if g["dirty"]:
                if g["batch"] is not None:
                    g["batch"].free()
                    g["batch"] = None
                run = g["members"][np.isnan(done_at[g["members"]])]
                g["members"] = g["bm"] = run
                g["dirty"] = False
                g["polls"] = 0
                if len(run):
                    g["batch"] = mm.Batch([sims[j] for j in run], graph=graph)
                    rebuilds += 1
            if g["batch"] is not None and len(g["members"]):
                g["batch"].enqueue(poll)
                g["busy"] = True
                g["polls"] += 1
                trials += poll
                launched = True
        if not launched and any(len(g["members"]) or g["busy"] for g in state):
            time.sleep(0.0002)
    for g in state:
        if g["batch"] is not None:
            g["batch"].free()
    return trials, rebuilds, moved


def bench_sweep(args, mm):
    """Batched sweep against members run one at a time on the same GPU.  Runs
    the batched phase and returns a function that runs the serial sample and
    returns the result object.  All
    members start together (t = 0); after a first stretch of trials with every
    member in one batch (which gives the full-batch trial time), they are
    advanced in size groups (cells <= 20480, <= 81920, above), each group a
    batch on its own stream enqueued asynchronously, so that the many short
    trials of small members are not held to the cadence of large ones.  Setup
    (kernel spectra) is excluded from both sides; latency = time from the start
    to each member's done flag."""
    cfgs = mm.workloads.sweep(args.sweep_sims, seed=4, duration=args.sweep_ns * 1e-9)
    n = len(cfgs)
    t0 = time.perf_counter()
    sims = [mm.workloads.make_sim(c) for c in cfgs]
    for s in sims:
        s._init()
    b = mm.Batch(sims, graph=not args.no_graph)
    setup = time.perf_counter() - t0
    b.start(seconds=[c["duration"] for c in cfgs])
    done_at = np.full(n, np.nan)
    members = np.arange(n)
    trials = rebuilds = 0
    t1 = time.perf_counter()
    b.trials(1)  # captures the batched trial
    # full-batch trial time: the next `poll` trials, every member running
    ta = time.perf_counter()
    b.trials(args.poll)
    t_full = (time.perf_counter() - ta) / args.poll
    trials += 1 + args.poll
    d = b.done()
    done_at[d] = time.perf_counter() - t1
    b.free()
    cells = np.array([np.prod(c["size"]) for c in cfgs])
    bounds = (20480, 81920)
    grp = np.searchsorted(np.asarray(bounds), cells, side="left")
    groups = [np.nonzero((grp == k) & ~d)[0] for k in range(len(bounds) + 1)]
    tr, rebuilds, moved = run_groups(mm, sims, groups, args.poll, not args.no_graph, time.perf_counter, t1, done_at,
                                     lane_factor=None if args.no_lanes else args.lane_factor)
    trials += tr
    t_batch = time.perf_counter() - t1
    b.free()
    full_bytes = sum(dp_trial_bytes(c["size"]) for c in cfgs)
    steps = np.array([s.state()["steps"] for s in sims])
    cells = np.array([np.prod(c["size"]) for c in cfgs])
    j_par = int(np.argmin(cells))
    m_par = sims[j_par].m
    for s in sims:
        s.free()
    print("bench.py sweep: %d simulations batched in %.2f s" % (n, t_batch), file=sys.stderr, flush=True)

    def finish():
        # serial: an evenly spaced sample of members, each alone
        k = min(n, args.serial_sample)
        pick = np.unique(np.linspace(0, n - 1, k).astype(int))
        t_alone = []
        for j in pick:
            s = mm.workloads.make_sim(cfgs[j])
            s._init()
            s.sync()
            ta = time.perf_counter()
            s.run(cfgs[j]["duration"])
            t_alone.append(time.perf_counter() - ta)
            s.free()
        t_alone = np.array(t_alone)
        serial_rate = 1.0 / t_alone.mean()
        rate = n / t_batch
        return {
            "metric": "sweep solves/s",
            "value": round(rate, 3),
            "unit": "solves/s",
            "n_gpus": 1,
            "higher_is_better": True,
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic: seeded sweep (SplitMix64 seed 4), random m0 on the sphere",
            "config": {"workload": "sweep_%d_sims_%gns_rkdp" % (n, args.sweep_ns), "sims": n,
                       "cells_total": int(cells.sum()), "solver": "rkdp", "max_error": 1e-5},
            "batch_seconds": round(t_batch, 4),
            "batched_trials": trials,
            "batch_rebuilds": rebuilds,
            "groups_cells_upto": [20480, 81920, None],
            "moved_to_lanes": moved,
            "full_batch_trial": {"ms": round(t_full * 1e3, 4), "model_bytes": int(full_bytes),
                                 "hbm_frac": round(full_bytes / t_full / (HBM_PEAK_GBS * 1e9), 4)},
            "accepted_steps_total": int(steps.sum()),
            "mcell_steps_per_s": round(float((steps * cells).sum()) / t_batch / 1e6, 2),
            "latency_s": {"p50": round(float(np.percentile(done_at, 50)), 4),
                          "p99": round(float(np.percentile(done_at, 99)), 4)},
            "serial": {"sample": int(len(pick)), "mean_s_per_sim": round(float(t_alone.mean()), 5),
                       "solves_per_s": round(serial_rate, 3),
                       "latency_s": {"p50": round(float(np.percentile(t_alone, 50)), 4),
                                     "p99": round(float(np.percentile(t_alone, 99)), 4)}},
            "speedup_vs_serial": round(rate / serial_rate, 2),
            "parity": None if args.no_cpu else member_parity(mm, cfgs[j_par], m_par, 0.02e-9),
            "setup_seconds": round(setup, 2),
        }

    return finish


def bench_serve(args, mm):
    """Continuous batching of a Poisson request stream against one-at-a-time
    service in arrival order.  Runs the served stream and returns a function
    that runs the serial side and returns the result object.  The serial side runs `--serve-serial` requests
    (all by default) one at a time; when that is a sample, every request's
    service time is taken from the sample member with the nearest cells x
    duration and the FIFO queue is computed from the arrivals (stated in
    the output)."""
    sv = importlib.import_module("1_amd.serve")
    reqs = sv.request_stream(args.serve_requests, rate=args.serve_rate, seed=5)
    n = len(reqs)
    finished = [0]
    j_par = int(np.argmin([np.prod(r.cfg["size"]) for r in reqs]))
    m_par = [None]

    def progress(r):
        finished[0] += 1
        if r.rid == reqs[j_par].rid:
            m_par[0] = r.sim.m  # the smallest request's result, for the parity check
        if finished[0] % 25 == 0:
            print("bench.py serve: %d of %d requests served" % (finished[0], n), file=sys.stderr, flush=True)

    for r in reqs:
        r.on_finish = progress
    srv = sv.Server(poll=args.poll, graph=not args.no_graph, lane_factor=None if args.no_lanes else args.lane_factor,
                    lane_poll=args.lane_poll)
    stats = srv.serve(reqs, max_seconds=args.serve_max_seconds)
    print("bench.py serve: %d requests served in batches in %.2f s" % (len(reqs), stats["elapsed"]), file=sys.stderr,
          flush=True)
    lat = sv.latency_stats(reqs)
    rate = n / stats["elapsed"]
    work = np.array([np.prod(r.cfg["size"]) * r.duration for r in reqs])

    def finish():
        k = n if args.serve_serial <= 0 else min(n, args.serve_serial)
        pick = np.unique(np.linspace(0, n - 1, k).astype(int))
        service = {}
        for i, j in enumerate(pick):
            r = reqs[j]
            s = mm.workloads.make_sim(r.cfg)
            ta = time.perf_counter()
            s._init()
            s.run(r.duration)
            s.sync()
            service[j] = time.perf_counter() - ta
            s.free()
            if (i + 1) % 32 == 0:
                print("bench.py serve: %d of %d requests run alone" % (i + 1, len(pick)), file=sys.stderr, flush=True)
        svc = np.array([service[j] if j in service else service[min(service, key=lambda i: abs(np.log(work[i] / work[j])))]
                        for j in range(n)])
        t_free, fin = 0.0, np.zeros(n)
        arr = np.array([r.arrival for r in reqs])
        for j in np.argsort(arr, kind="stable"):
            t_free = max(t_free, arr[j]) + svc[j]
            fin[j] = t_free
        slat = fin - arr
        serial_rate = n / fin.max()
        return {
            "metric": "served solves/s",
            "value": round(rate, 3),
            "unit": "solves/s",
            "n_gpus": 1,
            "higher_is_better": True,
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic: seeded Poisson request stream (SplitMix64 seed 5)",
            "config": {"workload": "serve_%d_requests_poisson_%g_per_s" % (n, args.serve_rate), "requests": n,
                       "cells_total": int(sum(np.prod(r.cfg["size"]) for r in reqs)), "solver": "rkdp",
                       "hbm_budget_bytes": sv.HBM_BYTES},
            "elapsed_s": round(stats["elapsed"], 3),
            "latency_s": {"p50": round(lat["p50"], 4), "p99": round(lat["p99"], 4), "mean": round(lat["mean"], 4)},
            "batch_rebuilds": stats["rebuilds"],
            "moved_to_lanes": stats["moved_to_lanes"],
            "max_concurrent": stats["max_active"],
            "max_device_bytes": stats["max_bytes"],
            "serial": {"measured": int(len(pick)), "of": n,
                       "service_s_total": round(float(svc.sum()), 3),
                       "solves_per_s": round(serial_rate, 3),
                       "latency_s": {"p50": round(float(np.percentile(slat, 50)), 4),
                                     "p99": round(float(np.percentile(slat, 99)), 4)},
                       "note": "one request at a time in arrival order (FIFO)" +
                               ("" if len(pick) == n else "; unmeasured service times from the nearest measured request")},
            "speedup_vs_serial": round(rate / serial_rate, 2),
            "parity": None if args.no_cpu or m_par[0] is None else member_parity(
                mm, dict(reqs[j_par].cfg, duration=reqs[j_par].duration), m_par[0], 0.02e-9),
        }

    return finish


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--pass-launches", type=int, default=200,
                    help="back-to-back launches of the dominant pass for its average duration")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["cuboid", "sweep", "serve"], default="cuboid",
                    help="cuboid: the single-simulation RK4 step (default); sweep: the batched sweep; "
                         "serve: continuous batching of a request stream")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-serial", type=int, default=None,
                    help="requests run alone for the serial side (0: all; default 0 for --workload serve, "
                         "40 in the combined default run)")
    ap.add_argument("--serve-max-seconds", type=float, default=900.0)
    ap.add_argument("--sweep-sims", type=int, default=512)
    ap.add_argument("--sweep-ns", type=float, default=0.5)
    ap.add_argument("--poll", type=int, default=16, help="batched trials between done-flag polls")
    ap.add_argument("--no-graph", action="store_true", help="batched trials as eager launches (no HIP graph)")
    ap.add_argument("--no-lanes", action="store_true",
                    help="keep long-running members in their size group (no lanes of their own)")
    ap.add_argument("--lane-factor", type=float, default=3.0,
                    help="members with more than this times their group's median remaining trials move to a lane")
    ap.add_argument("--lane-poll", type=int, default=4, help="serving: a lane enqueues this many times --poll trials")
    ap.add_argument("--serial-sample", type=int, default=32, help="sweep members timed alone for the serial rate")
    ap.add_argument("--no-sweep", action="store_true", help="cuboid workload: leave out the sweep object")
    ap.add_argument("--no-serve", action="store_true", help="cuboid workload: leave out the serve object")
    ap.add_argument("--calibrate", action="store_true",
                    help="also launch a streaming kernel of known traffic (for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs each simulation on one MI355X; runs on %d GPUs are not part of this build "
              "(use --gpus 1)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    mm = importlib.import_module("1_amd")
    lib = mm.load()
    if args.workload == "sweep":
        print(json.dumps(bench_sweep(args, mm)()), flush=True)
        return
    if args.workload == "serve":
        if args.serve_serial is None:
            args.serve_serial = 0
        print(json.dumps(bench_serve(args, mm)()), flush=True)
        return
    if args.serve_serial is None:
        args.serve_serial = 40
    cfg = mm.workloads.cuboid_rk4()
    N = int(np.prod(cfg["size"]))
    sim = mm.workloads.make_sim(cfg)
    sim.steps(args.warmup)  # also builds the kernel and allocates everything
    st = sim.state()
    P = (int(st["px"]), int(st["py"]), int(st["pz"]))

    ms = lib.mm_sim_time_steps(sim._ptr, args.steps)
    ms_per_step = ms / args.steps
    value = N * args.steps / (ms / 1e3) / 1e6

    prof_ms = np.zeros(8)
    prof_n = np.zeros(8, np.int32)
    if args.profile_steps > 0:
        lib.mm_sim_profile_steps(sim._ptr, args.profile_steps, prof_ms.ctypes.data, prof_n.ctypes.data)
    names = list(CLASSES)
    if sim.tune_params()["pp"]:
        names[2] = "y_xkx_y_plane"
    avg = {names[i]: (prof_ms[i] / prof_n[i] if prof_n[i] else 0.0) for i in range(8) if prof_n[i]}
    # per pass, back-to-back launches between two events (no per-launch events): these add up to the step
    b2b = {}
    for i in range(5):
        if (i in (1, 3)) and names[2] == "y_xkx_y_plane":
            continue  # the plane pass does the Y passes' work
        b2b[names[i]] = lib.mm_sim_time_pass(sim._ptr, i, args.pass_launches) / args.pass_launches
    tune = sim.tune_params()
    plane = bool(tune["pp"])
    per, b_conv, b_step, b_passes = model_bytes(N, P, cfg["size"][0], cfg["size"][1], 6 if P[0] > 1 else 4, plane)
    cls_of = {k: CLASSES.index(k) for k in per if k in CLASSES}
    cls_of["y_xkx_y_plane"] = 2  # the plane pass is profiled as class 2
    dominant = max(per, key=lambda k: prof_ms[cls_of[k]])
    dom_ms = lib.mm_sim_time_pass(sim._ptr, cls_of[dominant], args.pass_launches) / args.pass_launches
    ach = per[dominant] / (dom_ms / 1e3) / 1e9 if dom_ms > 0 else 0.0
    traffic = committed_traffic(dominant, cfg["name"])
    step_gbs = b_step / (ms_per_step / 1e3) / 1e9

    m = sim.m
    norm_err = float(np.abs(np.linalg.norm(m, axis=0) - 1).max())

    out = {
        "metric": "Mcell-steps/s",
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded (SplitMix64 seed 2) random m0 uniform on the sphere",
        "config": {
            "workload": cfg["name"],
            "cells": N,
            "grid_internal_xyz": list(cfg["size"]),
            "padded": list(P),
            "solver": "rk4",
            "dt_s": cfg["dt"],
            "kernel": "FaceKernel6 (8x8 charges/face) + 6-neighbour exchange, FP64 setup",
        },
        "hbm_pct_peak_model": round(100 * step_gbs / HBM_PEAK_GBS, 2),
        "step_model_bytes": b_step,
        "step_pass_bytes": b_passes,
        "hbm_pct_peak_passes": round(100 * b_passes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 2),
        "launch_shapes": tune,
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": per[dominant],
            "avg_launch_ms": round(dom_ms, 5),
            "launches_timed": args.pass_launches,
        },
        "kernel_avg_ms": {k: round(v, 5) for k, v in b2b.items()},
        "step_ms_from_passes": round(4 * sum(b2b.values()), 5),
        "kernel_avg_ms_event_bracketed": {k: round(v, 5) for k, v in avg.items()},
        "norm_max_dev": norm_err,
    }
    sim.free()
    if args.calibrate:
        out["calibration"] = calibrate_copy(lib)
    if not args.no_cpu:
        out["cpu_baseline"], o, n = cpu_baseline(cfg, args.cpu_seconds)
        out["parity"] = parity_vs_oracle(mm, cfg, o, n)
        del o
    print("bench.py cuboid: %.2f Mcell-steps/s" % value, file=sys.stderr, flush=True)
    # the batched sweep and the served request stream, then the same
    # simulations run one at a time on this GPU.  Each batched workload starts
    # as in a fresh process: the engine's cached blocks and idle streams of the
    # earlier ones are released (mm_memory_trim).  A stream keeps the hardware
    # queue it got at creation, so streams inherited from the sweep put the
    # served stream's concurrent batches on other queues (~9 % slower).
    lib.mm_memory_trim()
    sweep_serial = bench_sweep(args, mm) if not args.no_sweep else None
    lib.mm_memory_trim()
    serve_serial = bench_serve(args, mm) if not args.no_serve else None
    if sweep_serial:
        out["sweep"] = sweep_serial()
    if serve_serial:
        out["serve"] = serve_serial()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
