"""Thermal field in the fused engine (Temperature, temperature.go:29-56;
drawn once per step, rk.go:289-292) against the FP64 oracle and against
statistical physics.

The engine and the oracle draw the same counter-based normals (seed, step,
cell), so fields and trajectories compare cell by cell; the standard
deviation sqrt(2 alpha kB T / (gamma0 mu0 Ms V dt)) / Ms is pinned by the
equilibrium of non-interacting macrospins, whose mean moment along the field
is the Langevin function L(xi) = coth(xi) - 1/xi of xi = Ms V B / (kB T).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KB = 1.380650424e-23


def film_cfg(mm, T, seed=3, size=(1, 16, 32)):
    m = np.zeros((3,) + size, np.float32)
    m[2] = 1.0
    m[0] = 0.1
    m /= np.linalg.norm(m, axis=0)
    return dict(size=size, cellsize=(3e-9, 4e-9, 4e-9), msat=8e5, aexch=1.3e-11, alpha=0.1, solver="rk12",
                maxerror=1e-3, dt=2e-13, field=("static", (0.0, 5e-3, 20e-3)), temperature=T, seed=seed,
                m0=m)


def test_thermal_field_matches_the_oracle_stream(mm, orc):
    """H_eff(T) - H_eff(0) of the same m0: the engine's thermal field equals the
    oracle's (the same normals, FP32 transform) to 2e-6 of its size."""
    print("thermal: heff of a 16x32 film at 300 K", flush=True)
    cfg = film_cfg(mm, 300.0)
    cfg0 = dict(cfg, temperature=0.0)
    s, s0 = mm.workloads.make_sim(cfg), mm.workloads.make_sim(cfg0)
    d = s.heff(cfg["m0"]) - s0.heff(cfg["m0"])
    o, o0 = (orc.sim(cfg["m0"], **mm.workloads.oracle_kwargs(c)) for c in (cfg, cfg0))
    dref = (o.heff(cfg["m0"]) - o0.heff(cfg["m0"])).reshape(d.shape)
    err = float(np.abs(d - dref).max() / np.abs(dref).max())
    print("thermal field: max |dref| %.3g, relative difference %.3g" % (np.abs(dref).max(), err))
    assert np.abs(dref).max() > 0.1
    assert err < 2e-6
    s.free()
    s0.free()


def test_thermal_run_replayed_in_oracle(mm, orc):
    """40 RK12 steps at 300 K; the engine's trial log (dt, accepted) replayed in
    the FP64 oracle, which draws each step's field from the dt of that step's
    first trial: max |dm| <= 1e-5."""
    print("thermal: 40 RK12 steps of a 16x32 film at 300 K", flush=True)
    cfg = film_cfg(mm, 300.0)
    s = mm.workloads.make_sim(cfg)
    s.steps(40)
    dts, acc = s.trial_log()
    o = orc.sim(cfg["m0"], **mm.workloads.oracle_kwargs(cfg))
    i = 0
    while i < len(dts):
        j = i
        while not acc[j]:
            j += 1
        o.step_replay(dts[i:j + 1], acc[i:j + 1])
        i = j + 1
    st = s.state()
    assert st["steps"] == 40 == o.state()["steps"]
    dm = float(np.abs(s.m.reshape(-1) - o.m).max())
    print("thermal run: %d trials, %d rejected, max|dm| vs FP64 replay %.3g" % (len(dts), len(dts) - acc.sum(), dm))
    assert dm < 1e-5
    s.free()


def test_thermal_batch_equals_alone(mm):
    """Members at different temperatures and seeds: bit-identical trial logs
    and m alone and in a batch (the stream does not depend on the launch)."""
    cfgs = [film_cfg(mm, T, seed) for T, seed in [(300.0, 1), (77.0, 2), (0.0, 0)]]
    ref = []
    for c in cfgs:
        s = mm.workloads.make_sim(c)
        s.run(5e-12)
        ref.append((s.m, s.trial_log()[0]))
        s.free()
    sims = [mm.workloads.make_sim(c) for c in cfgs]
    b = mm.Batch(sims)
    b.run([5e-12] * len(sims))
    b.free()
    for (m, dt), s in zip(ref, sims):
        assert np.array_equal(s.trial_log()[0], dt)
        assert np.array_equal(s.m, m)
        s.free()
    assert not np.array_equal(ref[0][0], ref[2][0])


def test_macrospins_reach_the_langevin_equilibrium(mm):
    """64 x 64 cubes of 10 nm with Ms = 1e3 A/m (dipolar coupling 0.4 % of the
    applied field), negligible exchange (A = 1e-20 J/m), alpha 0.5, B = 0.1 T,
    T chosen for xi = Ms V B / (kB T) = 2.  Heun (RK12) at a fixed 0.5 ps
    (max_dt, a tolerance no step exceeds): the thermal kick per step is about
    0.07 rad, the relaxation time about 110 ps.  After 1 ns, the moment along
    the field averaged over 40 snapshots 0.2 ns apart equals
    L(2) = coth(2) - 1/2 = 0.5373 within 0.01 (the FP64 oracle at 32 x 32
    gives 0.5404 +- 0.0018 on the same schedule); a sigma off by a factor
    sqrt(2) would give 0.41 or 0.65."""
    Ms, V, B, xi = 1e3, (10e-9) ** 3, 0.1, 2.0
    T = Ms * V * B / (KB * xi)
    size = (1, 64, 64)
    m = np.zeros((3,) + size, np.float32)
    m[2] = 1.0
    cfg = dict(size=size, cellsize=(10e-9, 10e-9, 10e-9), msat=Ms, aexch=1e-20, alpha=0.5, solver="rk12",
               maxerror=1.0, dt=5e-13, max_dt=5e-13, field=("static", (0.0, 0.0, B)), temperature=T, seed=11,
               m0=m)
    s = mm.workloads.make_sim(cfg)
    print("thermal: Langevin test at %.4f K" % T, flush=True)
    s.run(1e-9)
    print("thermal: equilibrated after %d steps" % s.state()["steps"], flush=True)
    mz = []
    for _ in range(40):
        s.run(0.2e-9)
        mz.append(s.average_m()[2])
    got = float(np.mean(mz))
    want = 1 / np.tanh(xi) - 1 / xi
    st = s.state()
    print("Langevin: <m.B/|B|> = %.4f (snapshots %.4f..%.4f), L(%g) = %.4f, %d steps, %d rejected" % (
        got, min(mz), max(mz), xi, want, st["steps"], st.get("rejects", 0)))
    assert abs(got - want) < 0.01
    s.free()
