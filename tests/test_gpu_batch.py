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


@pytest.mark.parametrize("merge", [0, 1 << 30])
def test_batch_merged_and_per_length_launches_equal_alone(mm, lib, merge):
    """Members of 2-D, thin and 3-D plans and many length classes, with every
    small class of every pass merged into one dispatch
    launch (1 << 30) or each launched alone (0): the same, bit for bit, as
    alone."""
    old = lib.mm_batch_set_merge_blocks(merge)
    try:
        cfgs = [mm.workloads.sweep_member(12, j, duration=0.008e-9, sizes=(32, 40, 48, 64, 80, 96),
                                          thick=(1, 2, 4, 16)) for j in range(12)]
        assert len({c["size"] for c in cfgs}) > 6
        check_same(mm, cfgs, [c["duration"] for c in cfgs])
    finally:
        lib.mm_batch_set_merge_blocks(old)


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


def test_sweep_length_classes_batched_equal_alone_and_oracle(mm, orc):
    """The 512-simulation sweep (seed 4) at its own sizes: one member per
    length class of the batched passes (every padded Z length, every padded
    (X, Y) pair) plus the 256x256x4 member (padded 8 x 512 x 512, the largest
    plane-pass class), 20 Dormand-Prince steps each: batched == alone bit for
    bit (trial log and m), and the largest member's trials replayed in the
    FP64 oracle within 1e-5."""
    cfgs = mm.workloads.sweep(512, seed=4)
    pick = mm.workloads.sweep_length_cover(cfgs)
    big = [j for j, c in enumerate(cfgs) if c["size"] == (4, 256, 256)]
    assert big
    pick = sorted(set(pick) | {big[0]})
    sel = [cfgs[j] for j in pick]
    nsteps = 20
    ref = []
    for c in sel:
        s = mm.workloads.make_sim(c)
        s.steps(nsteps)
        st = s.state()
        assert (int(st["px"]), int(st["py"]), int(st["pz"])) == tuple(mm.workloads.padded_length(n) for n in c["size"])
        ref.append((s.m, s.trial_log()))
        s.free()
    sims = [mm.workloads.make_sim(c) for c in sel]
    b = mm.Batch(sims)
    b.start(steps=nsteps)
    while b.trials(8):
        pass
    assert (b.status() == 1).all()
    for j, (s, (m, (dt, acc))) in enumerate(zip(sims, ref)):
        dt2, acc2 = s.trial_log()
        assert np.array_equal(dt2, dt) and np.array_equal(acc2, acc), pick[j]
        assert np.array_equal(s.m, m), (pick[j], float(np.abs(s.m - m).max()))
        assert s.state()["steps"] == nsteps
    b.free()
    # the largest member in the FP64 oracle, replaying its trials
    k = pick.index(big[0])
    c = sel[k]
    dts, acc = ref[k][1]
    o = orc.sim(c["m0"], **mm.workloads.oracle_kwargs(c))
    i = 0
    while i < len(dts):
        j = i
        while not acc[j] and j - i < 9:
            j += 1
        o.step_replay(dts[i:j + 1], acc[i:j + 1])
        i = j + 1
    d = float(np.abs(ref[k][0] - o.m.reshape(ref[k][0].shape)).max())
    print("sweep cover: %d members %s; 256x256x4 member %d after %d steps (%d trials): max|dm| vs FP64 %.3g"
          % (len(pick), pick, big[0], nsteps, len(dts), d))
    assert d < 1e-5
    for s in sims:
        s.free()


def test_failed_member_is_reported_and_batch_mates_unaffected(mm):
    """A member whose m0 holds a NaN cell stops as failed (the reference stops
    a run whose dt turns NaN, rk.go:448-452); its batch-mates compute what they
    compute alone, bit for bit."""
    cfgs = [mm.workloads.sweep_member(12, j, duration=0.004e-9, **SMALL) for j in range(4)]
    bad = dict(cfgs[1])
    m0 = bad["m0"].copy()
    m0[0, 0, 0, 3] = np.nan
    bad["m0"] = m0
    cfgs[1] = bad
    ref = {j: alone(mm, c, c["duration"]) for j, c in enumerate(cfgs) if j != 1}
    sims = [mm.workloads.make_sim(c) for c in cfgs]
    b = mm.Batch(sims)
    b.run([c["duration"] for c in cfgs])
    st = b.status()
    assert st[1] == 2 and list(np.delete(st, 1)) == [1, 1, 1], st
    assert b.failed().tolist() == [False, True, False, False]
    assert sims[1].state()["steps"] <= 2
    for j, (m, (dt, acc), _) in ref.items():
        assert np.array_equal(sims[j].trial_log()[0], dt)
        assert np.array_equal(sims[j].m, m), j
    b.free()
    for s in sims:
        s.free()


def test_batch_keeps_its_field_table_when_the_simulation_replaces_it(mm):
    """A batch copies each member's pointwise field table when it is built
    (mm_sim_set_field then replaces the simulation's own table in place): a
    member whose table is replaced after the batch was built runs the ramp the
    batch was built with, bit for bit as alone; a batch built afterwards runs
    the new table, again as alone.  The second table is longer than the
    first, so the simulation's buffer grows as well."""
    base = mm.workloads.sweep_member(11, 3, duration=0.03e-9, **SMALL)  # the ramps part at 10 ps
    ramp1 = [(0.0, (0.0, 0.0, 0.0)), (0.01e-9, (20e-3, 5e-3, 0.0))]   # (t, internal (X, Y, Z) tesla)
    ramp2 = ramp1 + [(0.02e-9, (-30e-3, 0.0, 10e-3)), (0.05e-9, (0.0, 40e-3, 0.0))]

    def with_ramp(ramp):
        return mm.workloads.make_sim(dict(base, field=("pointwise", ramp)))

    def run_alone(ramp, seconds):
        s = with_ramp(ramp)
        s.run(seconds)
        out = s.m, s.trial_log()
        s.free()
        return out

    m1, (dt1, _) = run_alone(ramp1, base["duration"])
    s = with_ramp(ramp1)
    s._init()
    b = mm.Batch([s])
    for t, (bX, bY, bZ) in ramp2[2:]:
        s.applypointwise("field", t, bZ, bY, bX)   # replaces the simulation's table in place
    b.run([base["duration"]])
    b.free()
    assert np.array_equal(s.trial_log()[0], dt1) and np.array_equal(s.m, m1)
    s.free()

    m2, (dt2, _) = run_alone(ramp2, base["duration"])
    s = with_ramp(ramp2)
    b = mm.Batch([s])
    b.run([base["duration"]])
    b.free()
    assert np.array_equal(s.trial_log()[0], dt2) and np.array_equal(s.m, m2)
    assert not np.array_equal(m1, m2)
    s.free()
