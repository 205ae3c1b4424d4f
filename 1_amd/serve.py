"""Continuous-batching request loop for LLG simulations on one GPU.

Requests (a simulation set-up plus a simulated duration) arrive over time.
The server keeps the running simulations in one batch (``Batch``, i.e.
``mm_batch_*``: one launch per pass and padded length per RK stage, captured
into a HIP graph) and advances it `poll` batched trials at a time.  Between
chunks it retires the simulations that reached their stop time, admits the
requests that have arrived, in arrival order, while their device footprint
(``mm_sim_footprint``, exact) fits the HBM budget, and rebuilds the batch when
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
    __slots__ = ("rid", "arrival", "cfg", "duration", "sim", "start", "finish", "footprint")

    def __init__(self, rid, arrival, cfg):
        self.rid = rid
        self.arrival = arrival
        self.cfg = cfg
        self.duration = cfg["duration"]
        self.sim = None
        self.start = None
        self.finish = None
        self.footprint = None


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


def footprint(cfg):
    """Device bytes the simulation of cfg allocates (mm_sim_footprint)."""
    s = wl.make_sim(cfg)
    p = s.params()
    return int(s._lib.mm_sim_footprint(C.byref(p)))


class Server:
    """Continuous batching over a request list whose arrivals are in seconds
    from the start of `serve`."""

    def __init__(self, hbm_bytes=HBM_BYTES, poll=16, graph=True, keep_results=False):
        self.hbm_bytes = hbm_bytes
        self.poll = poll
        self.graph = graph
        self.keep_results = keep_results
        self.results = {}

    def serve(self, requests, clock=time.perf_counter, max_seconds=None):
        """Run every request to completion; returns per-request (arrival, start,
        finish) in seconds relative to the start, and loop statistics."""
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            r.footprint = footprint(r.cfg)
            if r.footprint > self.hbm_bytes:
                raise InputErr("request %d needs %d bytes, more than the budget" % (r.rid, r.footprint))
        t0 = clock()
        active, batch = [], None
        dirty = False
        used = 0
        rebuilds = trials = 0
        max_active = max_used = 0
        qi = 0
        while qi < len(pending) or active:
            now = clock() - t0
            if max_seconds is not None and now > max_seconds:
                raise RuntimeError("serving did not finish within %g s" % max_seconds)
            # admit arrived requests in order while they fit
            joined = []
            while qi < len(pending) and pending[qi].arrival <= now and used + pending[qi].footprint <= self.hbm_bytes:
                r = pending[qi]
                qi += 1
                r.sim = wl.make_sim(r.cfg)
                r.sim._init()
                r.sim._lib.mm_sim_begin_run(r.sim._ptr, r.duration)
                r.start = clock() - t0
                used += r.footprint
                joined.append(r)
            if joined:
                active.extend(joined)
                dirty = True
                max_active = max(max_active, len(active))
                max_used = max(max_used, used)
            if dirty:
                if batch is not None:
                    batch.free()
                batch = Batch([r.sim for r in active], graph=self.graph) if active else None
                rebuilds += batch is not None
                dirty = False
            if not active:
                # idle until the next arrival
                wait = pending[qi].arrival - (clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
                continue
            batch.trials(self.poll)
            trials += self.poll
            done = batch.done()
            if done.any():
                t = clock() - t0
                keep = []
                for r, d in zip(active, done):
                    if d:
                        r.finish = t
                        if self.keep_results:
                            self.results[r.rid] = r.sim.m
                        r.sim.free()
                        r.sim = None
                        used -= r.footprint
                    else:
                        keep.append(r)
                active = keep
                dirty = True
        if batch is not None:
            batch.free()
        t_end = clock() - t0
        return dict(elapsed=t_end, rebuilds=rebuilds, batched_trials=trials, max_active=max_active,
                    max_bytes=max_used)


def serve_serial(requests, clock=time.perf_counter, keep_results=False):
    """The same requests one at a time, in arrival order (each starts when it
    has arrived and the previous one finished)."""
    results = {}
    t0 = clock()
    for r in sorted(requests, key=lambda r: r.arrival):
        now = clock() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        s = wl.make_sim(r.cfg)
        s._init()
        r.start = clock() - t0
        s.run(r.duration)
        r.finish = clock() - t0
        if keep_results:
            results[r.rid] = s.m
        s.free()
    return dict(elapsed=clock() - t0), results


def latency_stats(requests):
    lat = np.array([r.finish - r.arrival for r in requests])
    return dict(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)), mean=float(lat.mean()))
