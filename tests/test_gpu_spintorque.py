"""Spin-transfer torque in the engine's ABI mode (CurrentDensity,
SpinPolarization, Xi, CurrentMask; torque.go:40-91, gpu_spintorque.cu) against
the FP64 oracle with the same current: the torque of a given m, 20 RK4 steps and
a Dormand-Prince run replayed, on a film and a 3-D grid; a current changed on a
running simulation; the fused mode refuses a current."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def smooth_m(size, seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = size
    x, y, z = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    m = np.zeros((3,) + size)
    for c in range(3):
        a = rng.standard_normal(4)
        m[c] = a[0] + a[1] * np.cos(2 * np.pi * y / Y + a[3]) + a[2] * np.sin(2 * np.pi * z / Z)
    m[2] += 1.0
    return (m / np.linalg.norm(m, axis=0)).astype(np.float32)


def mask_of(size, seed):
    mk = np.random.default_rng(seed).uniform(0.5, 1.5, (3,) + size).astype(np.float32)
    mk[:, :, : size[1] // 4, :] = 0    # no current in the first quarter of Y
    return mk


J = (4e11, -1e12, 2e12)     # A/m^2, internal order
P, XI = 0.5, 0.05


def pair(mm, orc, size, seed, solver="rk4", dt=2e-14, alpha=0.02):
    cell = (3e-9, 4e-9, 4e-9)
    m0 = smooth_m(size, seed)
    mk = mask_of(size, seed + 1)
    s = mm.Sim("abi")
    X, Y, Z = size
    s.gridsize(Z, Y, X)
    s.cellsize(cell[2], cell[1], cell[0])
    s.msat(8e5)
    s.aexch(1.3e-11)
    s.alpha(alpha)
    s.solvertype(solver)
    s.maxerror(1e-5)
    if dt:
        s.dt(dt)
    s.applystatic("field", 0.02, 0.0, 0.01)
    s.currentdensity(J[2], J[1], J[0])
    s.spinpolarization(P)
    s.xi(XI)
    s.currentmask_array(mk)
    s.setm_array(m0)
    o = orc.sim(m0, size=size, cellsize=cell, msat=8e5, aexch=1.3e-11, alpha=alpha, solver=solver, dt=dt,
                maxerror=1e-5, field=(np.zeros(1), np.array([0.01, 0.0, 0.02])))
    o.set_spintorque(J, P, XI, mk)
    return s, o, m0


@pytest.mark.parametrize("size", [(1, 32, 64), (6, 16, 24)])
def test_spintorque_torque_and_rk4_vs_oracle(mm, orc, size):
    print("spin-transfer torque: torque and 20 RK4 steps of %s" % (size,), flush=True)
    s, o, m0 = pair(mm, orc, size, 3)
    k = s.torque(m0)
    kref = o.torque(m0).reshape(k.shape)
    rel = float(np.abs(k - kref).max() / np.abs(kref).max())
    # the current's part is not small: without it the torque differs
    o0 = orc.sim(m0, size=size, cellsize=(3e-9, 4e-9, 4e-9), msat=8e5, aexch=1.3e-11, alpha=0.02, solver="rk4",
                 dt=2e-14, field=(np.zeros(1), np.array([0.01, 0.0, 0.02])))
    stt = float(np.abs(kref - o0.torque(m0).reshape(k.shape)).max() / np.abs(kref).max())
    print("torque vs oracle %.3g relative; the current's share %.3g" % (rel, stt))
    assert rel < 5e-6 and stt > 1e-2
    s.steps(20)
    for _ in range(20):
        o.step()
    d = float(np.abs(s.m.reshape(-1) - o.m).max())
    print("max |dm| after 20 RK4 steps: %.3g" % d)
    assert d < 1e-5
    s.free()


def test_spintorque_dormand_prince_run_replayed_and_live_current(mm, orc):
    size = (1, 32, 64)
    print("spin-transfer torque: RKDP run of %s, then a current change" % (size,), flush=True)
    s, o, m0 = pair(mm, orc, size, 5, solver="rkdp", dt=0.0)
    o.m = orc.normalize(o.m)
    s.run(4e-12)
    s.currentdensity(0.0, 3e12, 0.0)       # live: the next torque uses it (CurrentDensity does not re-init)
    s.run(2e-12)
    dts, acc = s.trial_log()
    t_switch = None
    i = 0
    while i < len(dts):
        j = i
        while not acc[j]:
            j += 1
        if t_switch is None and o.state()["time"] * o.state()["unit_time"] >= 4e-12:
            o.set_spintorque((0.0, 3e12, 0.0), P, XI, mask_of(size, 6))
            o.m = orc.normalize(o.m)        # the second Run() normalises first (simrun.go:24)
            t_switch = i
        o.step_replay(dts[i:j + 1], acc[i:j + 1])
        i = j + 1
    assert t_switch is not None
    d = float(np.abs(s.m.reshape(-1) - o.m).max())
    print("%d trials, max |dm| vs FP64 replay %.3g" % (len(dts), d))
    assert d < 1e-5
    s.free()


def test_fused_mode_refuses_a_current(mm):
    s = mm.Sim("fused")
    s.gridsize(16, 8, 1)
    s.cellsize(4e-9, 4e-9, 4e-9)
    s.msat(8e5)
    s.aexch(1.3e-11)
    s.currentdensity(1e12, 0, 0)
    s.currentmask_array(np.ones((3, 1, 8, 16), np.float32))
    s.uniform(1, 0, 0)
    with pytest.raises(mm.InputErr):
        s.steps(1)


def test_applied_pointwise_current_run_replayed(mm, orc):
    """ApplyPointwise("j", ...) (simfield.go:61-70): a current ramp evaluated at
    each stage's time replaces CurrentDensity (torque.go:42-44); the
    Dormand-Prince run replayed in the oracle with the same table."""
    size = (1, 32, 64)
    print("spin-transfer torque: RKDP run of %s under a pointwise current" % (size,), flush=True)
    s, o, m0 = pair(mm, orc, size, 7, solver="rkdp", dt=0.0)
    tt = [-1e-12, 3e-12, 6e-12]
    jj = np.array([(0.0, 0.0, 0.0), (0.0, 3e12, 1e12), (0.0, -1e12, 2e12)])   # internal (jx, jy, jz)
    for t, (jx, jy, jz) in zip(tt, jj):
        s.applypointwise("j", t, jz, jy, jx)
    o.set_applied_current(tt, jj)
    o.m = orc.normalize(o.m)
    s.run(5e-12)
    dts, acc = s.trial_log()
    i = 0
    while i < len(dts):
        j = i
        while not acc[j]:
            j += 1
        o.step_replay(dts[i:j + 1], acc[i:j + 1])
        i = j + 1
    d = float(np.abs(s.m.reshape(-1) - o.m).max())
    print("%d trials, max |dm| vs FP64 replay %.3g" % (len(dts), d))
    assert d < 1e-5
    assert np.allclose(s.current_at(4.5e-12), jj[1] + 0.5 * (jj[2] - jj[1]))
    s.free()
