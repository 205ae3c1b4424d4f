"""Continuous batching (serve.Server): requests that arrive while others run
are admitted into the running batch, and every result equals the same request
run alone (bit for bit); admission respects the HBM budget."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sv():
    return importlib.import_module("1_amd.serve")


def small_stream(sv, n, seed):
    return sv.request_stream(n, rate=400.0, seed=seed, max_cells=(96, 96, 4), dmin=0.004e-9, dmax=0.02e-9)


def test_each_result_same_alone_and_in_the_mix(mm, sv):
    reqs = small_stream(sv, 10, seed=21)
    srv = sv.Server(poll=4, keep_results=True)
    stats = srv.serve(reqs, max_seconds=300)
    assert stats["max_active"] > 1  # requests overlapped on the device
    assert all(r.finish is not None and r.finish >= r.arrival for r in reqs)
    for r in reqs:
        s = mm.workloads.make_sim(r.cfg)
        s.run(r.duration)
        assert np.array_equal(s.m, srv.results[r.rid]), r.rid
        s.free()


def test_admission_stays_within_the_hbm_budget(mm, sv):
    reqs = small_stream(sv, 8, seed=22)
    fp = sorted(sv.footprint(r.cfg) for r in reqs)
    budget = fp[-1] + fp[-2]  # at most about two requests at a time
    srv = sv.Server(hbm_bytes=budget, poll=4)
    stats = srv.serve(reqs, max_seconds=300)
    assert stats["max_bytes"] <= budget
    assert all(r.finish is not None for r in reqs)


def test_lanes_move_long_runners_and_keep_results(mm, sv):
    """Members moved to their size class's lane (a batch of their own) mid-run
    still compute what they compute alone."""
    reqs = small_stream(sv, 12, seed=23)
    srv = sv.Server(poll=2, keep_results=True, lane_factor=1.2, lane_floor=0, lane_after=1)
    stats = srv.serve(reqs, max_seconds=300)
    print("moved to lanes:", stats["moved_to_lanes"])
    assert stats["moved_to_lanes"] > 0
    for r in reqs:
        s = mm.workloads.make_sim(r.cfg)
        s.run(r.duration)
        assert np.array_equal(s.m, srv.results[r.rid]), r.rid
        s.free()


def test_long_runners_rule(sv):
    rem = np.array([0, 10, 12, 11, 100, 9, 0, 50], np.float32)
    assert sv.long_runners(rem, 3.0, 0) == [4, 7]
    assert sv.long_runners(rem, 3.0, 60) == [4]
    assert sv.long_runners(np.array([5, 100, 0], np.float32), 3.0, 0) == []      # fewer than 4 running
    assert sv.long_runners(np.array([1, 1, 9, 9, 9, 9], np.float32), 1.5, 0) == []  # more than half


def test_config5_size_range_same_alone_and_in_the_mix(mm, sv):
    """A Poisson stream of 20 requests over the serving workload's size range
    (log-uniform from 16x16x1; its largest request made 512x512x4, padded
    8 x 1024 x 1024) with short simulated durations: the requests overlap on
    the device, each finishes "ok", and each result equals the same request
    run alone, bit for bit."""
    wl = mm.workloads
    reqs = sv.request_stream(20, rate=400.0, seed=34, dmin=0.002e-9, dmax=0.01e-9)
    big = max(reqs, key=lambda r: int(np.prod(r.cfg["size"])))
    size = (4, 512, 512)
    m0 = wl.normalize(wl.normal(3434, 3 * int(np.prod(size))).reshape((3,) + size)).astype(np.float32)
    big.cfg = dict(big.cfg, size=size, m0=m0)
    cells = sorted(int(np.prod(r.cfg["size"])) for r in reqs)
    assert cells[0] <= 1024 and cells[-1] == 512 * 512 * 4
    srv = sv.Server(poll=4, keep_results=True)
    stats = srv.serve(reqs, max_seconds=300)
    print("serve 20 requests, %d to %d cells: %.2f s, max concurrent %d, rebuilds %d"
          % (cells[0], cells[-1], stats["elapsed"], stats["max_active"], stats["rebuilds"]))
    assert stats["max_active"] > 1
    for r in reqs:
        assert r.status == "ok" and r.finish >= r.arrival
        s = mm.workloads.make_sim(r.cfg)
        s.run(r.duration)
        assert np.array_equal(s.m, srv.results[r.rid]), r.rid
        s.free()


def test_failed_request_is_reported(mm, sv):
    """A request whose m0 holds a NaN finishes with status "failed"; the
    others finish "ok" with the results they have alone."""
    reqs = small_stream(sv, 6, seed=24)
    m0 = reqs[2].cfg["m0"].copy()
    m0[1, 0, 1, 1] = np.nan
    reqs[2].cfg = dict(reqs[2].cfg, m0=m0)
    srv = sv.Server(poll=4, keep_results=True)
    srv.serve(reqs, max_seconds=300)
    assert [r.status for r in reqs] == ["ok", "ok", "failed", "ok", "ok", "ok"]
    for r in reqs:
        if r.status == "ok":
            s = mm.workloads.make_sim(r.cfg)
            s.run(r.duration)
            assert np.array_equal(s.m, srv.results[r.rid]), r.rid
            s.free()
