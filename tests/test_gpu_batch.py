"""Batched engine (mm_batch_*): every simulation of a batch against the same
simulation run alone.

The batched kernels run the same device bodies as the single-simulation
kernels and the step control is per simulation on the device, so a member's
trial log (dt, accepted) and final m must be identical, bit for bit, alone
and in the mix.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = dict(sizes=(32, 48, 64, 80), thick=(1, 2, 4))


def alone(mm, cfg, seconds):
    s = mm.workloads.make_sim(cfg)
    s.run(seconds)
    out = (s.m, s.trial_log(), s.state())
    s.free()
    return out


def check_same(mm, cfgs, seconds, graph=True):
    ref = [alone(mm, c, t) for c, t in zip(cfgs, seconds)]
    sims = [mm.workloads.make_sim(c) for c in cfgs]
    b = mm.Batch(sims, graph=graph)
    nb = b.run(seconds)
    assert nb > 0 and b.done().all()
    for j, ((m, (dt, acc), st), s) in enumerate(zip(ref, sims)):
        dt2, acc2 = s.trial_log()
        assert len(dt2) == len(dt), (j, len(dt2), len(dt))
        assert np.array_equal(dt2, dt) and np.array_equal(acc2, acc), j
        assert np.array_equal(s.m, m), (j, float(np.abs(s.m - m).max()))
        assert s.state()["steps"] == st["steps"]
    b.free()
    for s in sims:
        s.free()
    return ref


def test_batch_dp_mix_equals_alone(mm):
    """Dormand-Prince (FSAL, adaptive) over 2-D and 3-D members of different
    padded lengths, materials and fields."""
    cfgs = [mm.workloads.sweep_member(7, j, duration=0.01e-9, **SMALL) for j in range(8)]
    assert len({c["size"][0] for c in cfgs}) > 1  # 2-D and 3-D members
    ref = check_same(mm, cfgs, [c["duration"] for c in cfgs])
    assert all(r[2]["steps"] > 5 for r in ref)


def test_batch_graph_replay_equals_eager_launches(mm):
    cfgs = [mm.workloads.sweep_member(8, j, duration=0.005e-9, **SMALL) for j in range(4)]
    check_same(mm, cfgs, [c["duration"] for c in cfgs], graph=False)


def test_batch_uneven_stop_times(mm):
    """Members stop at different times; finished members stay untouched while
    the others go on."""
    cfgs = [mm.workloads.sweep_member(9, j, duration=0.01e-9, **SMALL) for j in range(5)]
    check_same(mm, cfgs, [0.002e-9, 0.01e-9, 0.004e-9, 0.001e-9, 0.007e-9])


def test_batch_rk4_fixed_step_equals_alone(mm):
    cfgs = []
    for j in range(4):
        c = mm.workloads.sweep_member(10, j, duration=0.0, **SMALL)
        c.update(solver="rk4", dt=2e-15, maxerror=0.0)
        cfgs.append(c)
    ref = [None] * len(cfgs)
    for i, c in enumerate(cfgs):
        s = mm.workloads.make_sim(c)
        s.steps(30)
        ref[i] = s.m
        s.free()
    sims = [mm.workloads.make_sim(c) for c in cfgs]
    b = mm.Batch(sims)
    b.start(steps=30)
    while b.trials(8):
        pass
    for s, m in zip(sims, ref):
        assert np.array_equal(s.m, m)
        assert s.state()["steps"] == 30
