"""Continuous-batching request loop for LLG simulations on one GPU.

Requests (a simulation set-up plus a simulated duration) arrive over time.
The server keeps the running simulations in batches (``Batch``, i.e.
``mm_batch_*``: one launch per pass and padded length per RK stage, captured
into a HIP graph), one per size group on its own stream, and advances each
`poll` batched trials per enqueue without waiting for the others.  Between
chunks it retires the simulations that reached their stop time, admits the
requests that have arrived, in arrival order, while their device footprint
(``mm_sim_footprint``, exact) fits the HBM budget, and rebuilds a batch when
its membership changed.  A new member starts with ``mm_sim_begin_run`` (Run()'s
set-up) and is then stepped only by the batch, so requests that arrive while
others are on the GPU overlap with them.

Every simulation computes the same numbers alone and in any mix (the batched
kernels run the single-simulation kernels' device bodies and step control is
per simulation), which the tests check.
"""
import ctypes as C
import time

import numpy as np

from . import workloads as wl
from .sim import Batch, InputErr

HBM_BYTES = 288 * 10 ** 9


class Request:
    """A simulation to run for cfg["duration"] seconds from `arrival` on.  The
    simulation is built from the workload dict cfg, or by `make()` (an
    un-initialised ``Sim``; cfg then needs only "size" and "duration").
    `on_finish(request)` runs when it has finished, before it is freed.
    `status` is "ok" once it reached its stop time, "failed" if it stopped on
    a NaN dt or a non-finite torque or error estimate (the reference stops a
    run there, rk.go:448-452, simrun.go:48-50)."""

    __slots__ = ("rid", "arrival", "cfg", "duration", "sim", "start", "finish", "footprint", "make", "on_finish",
                 "status")

    def __init__(self, rid, arrival, cfg, make=None, on_finish=None):
        self.rid = rid
        self.arrival = arrival
        self.cfg = cfg
        self.duration = cfg["duration"]
        self.sim = None
        self.start = None
        self.finish = None
        self.footprint = None
        self.make = make
        self.on_finish = on_finish
        self.status = None

    def build(self):
        return self.make() if self.make is not None else wl.make_sim(self.cfg)


def request_stream(n, rate=200.0, seed=5, min_cells=(16, 16, 1), max_cells=(512, 512, 4), dmin=0.1e-9,
                   dmax=2e-9, cell=3e-9):
    """n requests with Poisson arrivals (rate per second): sizes log-uniform from
    min_cells to max_cells (user x, y, z; z rounded to 1, 2 or 4), material and
    field drawn as in the batched sweep, simulated duration ~ U[dmin, dmax].
    Half are relaxations (alpha 0.5, no field, random m0), half hysteresis
    points (m0 along +x with N(0, 0.02) noise, field along -x of U[0, 100] mT)."""
    u = wl.uniform01(seed, 8 * n).reshape(n, 8)
    gaps = -np.log(1.0 - u[:, 0]) / rate
    arrivals = np.cumsum(gaps) - gaps[0]
    out = []
    for j in range(n):
        base = wl.sweep_member(seed * 7919 + 1, j, cell=cell)
        lo, hi = np.log(min_cells), np.log(max_cells)
        nx = int(round(np.exp(lo[0] + u[j, 1] * (hi[0] - lo[0]))))
        ny = int(round(np.exp(lo[1] + u[j, 2] * (hi[1] - lo[1]))))
        nzf = np.exp(lo[2] + u[j, 3] * (hi[2] - lo[2]))
        nz = min((1, 2, 4), key=lambda v: abs(np.log(v) - np.log(nzf)))
        X, Y, Z = nz, ny, nx
        N = X * Y * Z
        key = (seed * 1000003 + 77 * j + 11) & wl.MASK64
        relax = u[j, 4] < 0.5
        if relax:
            m = wl.normalize(wl.normal(key, 3 * N).reshape(3, X, Y, Z))
            alpha, field = 0.5, None
        else:
            m = np.zeros((3, X, Y, Z))
            m[2] = 1.0  # internal Z = user x
            m = wl.normalize(m + wl.normal(key, 3 * N).reshape(3, X, Y, Z) * 0.02)
            alpha, field = base["alpha"], ("static", (0.0, 0.0, -u[j, 5] * 100e-3))
        cfg = dict(base, name="req%d" % j, size=(X, Y, Z), alpha=alpha, field=field,
                   duration=dmin + u[j, 6] * (dmax - dmin), m0=m.astype(np.float32),
                   kind="relax" if relax else "hysteresis")
        out.append(Request(j, float(arrivals[j]), cfg))
    return out


def footprint(cfg_or_request):
    """Device bytes the simulation of a workload dict or a Request allocates
    (mm_sim_footprint)."""
    r = cfg_or_request
    s = r.build() if isinstance(r, Request) else wl.make_sim(r)
    p = s.params()
    return int(s._lib.mm_sim_footprint(C.byref(p)))


class _Group:
    """Simulations advanced by one batch on its own stream."""

    def __init__(self, lane=False):
        self.members = []   # requests whose simulations the batch references
        self.batch = None
        self.dirty = False  # rebuild before the next enqueue
        self.busy = False   # an enqueue is in flight
        self.lane = lane    # holds the long-running members of a size class
        self.polls = 0      # enqueues since the last rebuild


def long_runners(rem, factor, floor):
    """Indices of the members whose remaining trials exceed `factor` times the
    median of the running members' and `floor` (at most half of them)."""
    run = np.nonzero(rem > 0)[0]
    if len(run) < 4:
        return []
    med = float(np.median(rem[run]))
    pick = [int(i) for i in run if rem[i] > factor * med and rem[i] > floor]
    return pick if len(pick) <= len(run) // 2 else []


class Server:
    """Continuous batching over a request list whose arrivals are in seconds
    from the start of `serve`.

    Requests are split by cell count into groups (`split_cells` boundaries),
    each advanced by its own batch on its own stream, `poll` trials per
    enqueue, without waiting for the others: a batched trial costs about the
    work of its largest members, so light simulations that need many short
    steps are not held to the cadence of heavy ones, and their launches
    overlap the heavy batch on the device.

    Every size class also has a lane for its long runners: after `lane_after`
    enqueues of a group, members whose remaining trials at their current dt
    (``mm_batch_remaining``) exceed `lane_factor` times the group's median
    move to the class's lane, a batch of their own whose trials are cheap, so
    the few simulations that need many more steps than the rest no longer set
    the pace of everything else (lane_factor=None: no lanes).  A lane enqueues
    `lane_poll` times `poll` trials at a time: its few members' trials are
    cheap, and they set the end of the run."""

    def __init__(self, hbm_bytes=HBM_BYTES, poll=16, graph=True, keep_results=False, split_cells=(16384, 131072),
                 lane_factor=3.0, lane_floor=1024, lane_after=4, lane_poll=4):
        self.hbm_bytes = hbm_bytes
        self.poll = poll
        self.graph = graph
        self.keep_results = keep_results
        self.split_cells = tuple(split_cells)
        self.lane_factor = lane_factor
        self.lane_floor = lane_floor
        self.lane_after = lane_after
        self.lane_poll = max(1, int(lane_poll))
        self.results = {}

    def _group_of(self, r):
        n = int(np.prod(r.cfg["size"]))
        return sum(n > b for b in self.split_cells)

    def serve(self, requests, clock=time.perf_counter, max_seconds=None):
        """Run every request to completion; returns per-request (arrival, start,
        finish) in seconds relative to the start, and loop statistics."""
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            r.footprint = footprint(r)
            if r.footprint > self.hbm_bytes:
                raise InputErr("request %d needs %d bytes, more than the budget" % (r.rid, r.footprint))
        nc = len(self.split_cells) + 1
        groups = [_Group() for _ in range(nc)] + [_Group(lane=True) for _ in range(nc)]
        moved = 0
        t0 = clock()
        used = 0
        rebuilds = trials = 0
        max_active = max_used = 0
        qi = 0
        while qi < len(pending) or any(g.members for g in groups):
            now = clock() - t0
            if max_seconds is not None and now > max_seconds:
                raise RuntimeError("serving did not finish within %g s" % max_seconds)
            # admit arrived requests in order while they fit
            while qi < len(pending) and pending[qi].arrival <= now and used + pending[qi].footprint <= self.hbm_bytes:
                r = pending[qi]
                qi += 1
                r.sim = r.build()
                r.sim._init()
                r.sim._lib.mm_sim_begin_run(r.sim._ptr, r.duration)
                r.start = clock() - t0
                used += r.footprint
                g = groups[self._group_of(r)]
                g.members.append(r)
                g.dirty = True
            max_active = max(max_active, sum(len(g.members) for g in groups))
            max_used = max(max_used, used)
            if qi < len(pending) and pending[qi].arrival <= now:
                # an arrived request waits for memory that finished members still hold
                for g in groups:
                    if any(r.finish is not None for r in g.members):
                        g.dirty = True
            launched = False
            for g in groups:
                if g.busy:
                    if not g.batch.ready():
                        continue
                    done = g.batch.flags()
                    status = g.batch.status()
                    g.busy = False
                    t = clock() - t0
                    for r, d, stat in zip(g.members, done, status):
                        if d and r.finish is None:
                            r.finish = t
                            r.status = "failed" if stat == 2 else "ok"
                            if self.keep_results:
                                self.results[r.rid] = r.sim.m
                            if r.on_finish is not None:
                                r.on_finish(r)
                    # finished members are skipped by the batched kernels; the batch is
                    # rebuilt without them once they are half of it (or at an admission)
                    nfin = sum(r.finish is not None for r in g.members)
                    if 2 * nfin >= len(g.members):
                        g.dirty = True
                    if self.lane_factor and not g.lane and g.polls >= self.lane_after:
                        # the batch holds the first len(done) members (admissions append)
                        rem = g.batch.remaining()
                        rem[[r.finish is not None for r in g.members[:len(done)]]] = 0
                        pick = set(long_runners(rem, self.lane_factor, self.lane_floor))
                        if pick:
                            lane = groups[nc + groups.index(g)]
                            lane.members += [g.members[i] for i in sorted(pick)]
                            g.members = [r for i, r in enumerate(g.members) if i not in pick]
                            g.dirty = lane.dirty = True
                            moved += len(pick)
                if g.dirty:
                    if g.batch is not None:
                        g.batch.free()
                        g.batch = None
                    keep = []
                    for r in g.members:
                        if r.finish is not None:
                            r.sim.free()
                            r.sim = None
                            used -= r.footprint
                        else:
                            keep.append(r)
                    g.members = keep
                    if keep:
                        g.batch = Batch([r.sim for r in keep], graph=self.graph)
                        rebuilds += 1
                    g.dirty = False
                    g.polls = 0
                if g.batch is not None:
                    n = self.poll * (self.lane_poll if g.lane else 1)
                    g.batch.enqueue(n)
                    g.busy = True
                    g.polls += 1
                    trials += n
                    launched = True
            if not launched:
                if any(g.busy for g in groups):
                    time.sleep(0.0002)
                elif qi < len(pending):
                    wait = pending[qi].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
        t_end = clock() - t0
        return dict(elapsed=t_end, rebuilds=rebuilds, batched_trials=trials, max_active=max_active,
                    max_bytes=max_used, moved_to_lanes=moved)


def serve_serial(requests, clock=time.perf_counter, keep_results=False):
    """The same requests one at a time, in arrival order (each starts when it
    has arrived and the previous one finished)."""
    results = {}
    t0 = clock()
    for r in sorted(requests, key=lambda r: r.arrival):
        now = clock() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        s = r.build()
        s._init()
        r.start = clock() - t0
        s.run(r.duration)
        r.finish = clock() - t0
        r.status = "ok"
        if keep_results:
            results[r.rid] = s.m
        s.free()
    return dict(elapsed=clock() - t0), results


def latency_stats(requests):
    lat = np.array([r.finish - r.arrival for r in requests])
    return dict(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)), mean=float(lat.mean()))
