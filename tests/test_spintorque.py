"""Spin-transfer torque (CurrentDensity, SpinPolarization, Xi, CurrentMask;
simspintorque.go, torque.go:40-91, gpu_spintorque.cu:24-211) in the FP64 oracle,
and the Sim commands.  CPU only.

The STT torque is the LL torque plus dtg (eps m x Hs - beta m x (m x Hs)), Hs the
directional difference of m along u.  With alpha = xi = 0 (eps = 0, beta = b) on a
spin spiral m = (cos kz, sin kz, 0) the central difference is exactly perpendicular
to m, so the added torque is b Hs: the spiral turns at the Zhang-Li drift rate
P muB j / (e Ms) * sin(k)/h (phase k per cell of size h), in the engine's time unit
1/(gamma0 Ms).  That SI route
pins the oracle's b, u, the one-half and the sign without its internal units."""
import numpy as np
import pytest

MUB, QE, GAMMA0 = 9.2740091523e-24, 1.60217646e-19, 2.211e5
MS, AEX = 8e5, 1.3e-11


def spiral(n, k):
    z = np.arange(n)
    m = np.zeros((3, 1, 1, n))
    m[0, 0, 0] = np.cos(k * z)
    m[1, 0, 0] = np.sin(k * z)
    return m


def sim(orc, size, m0, alpha=0.0, cell=(2e-9, 2e-9, 2e-9)):
    return orc.sim(m0, size=size, cellsize=cell, msat=MS, aexch=AEX, alpha=alpha, solver="rk4", dt=1e-14)


@pytest.mark.parametrize("J", [2e12, -5e11])
def test_spiral_turns_at_the_drift_rate(orc, J):
    n, k, h, P = 64, 0.3, 2e-9, 0.5
    m0 = spiral(n, k)
    a, b = sim(orc, (1, 1, n), m0), sim(orc, (1, 1, n), m0)
    b.set_spintorque((0.0, 0.0, J), P, 0.0, np.ones((3, 1, 1, n), np.float32))
    d = (b.torque(m0) - a.torque(m0)).reshape(3, n)
    # k is the phase per cell: m(z+1) - m(z-1) = 2 sin(k) (tangent), over 2 h
    rate = P * MUB * J / (QE * MS) * np.sin(k) / h / (GAMMA0 * MS)
    z = np.arange(1, n - 1)
    tangent = np.stack([-np.sin(k * z), np.cos(k * z), 0 * z])
    assert np.allclose(d[:, 1:-1], rate * tangent, rtol=1e-9, atol=1e-12 * abs(rate))
    # the edge cell takes the one-sided difference with itself (gpu_spintorque.cu:58-83),
    # no longer perpendicular to m: the added torque is b u (D - m (m . D))
    m_0, D = m0[:, 0, 0, 0], m0[:, 0, 0, 1] - m0[:, 0, 0, 0]
    assert np.allclose(d[:, 0], rate / (2 * np.sin(k)) * (D - m_0 * (m_0 @ D)), rtol=1e-9, atol=1e-12 * abs(rate))


def test_zero_current_and_uniform_m_give_the_ll_torque(orc):
    rng = np.random.default_rng(4)
    size = (2, 6, 10)
    m0 = rng.standard_normal((3,) + size)
    m0 /= np.linalg.norm(m0, axis=0)
    a, b = sim(orc, size, m0, alpha=0.1), sim(orc, size, m0, alpha=0.1)
    b.set_spintorque((0.0, 0.0, 0.0), 0.5, 0.05, None)             # no current: the LL form (torque.go:40)
    assert np.array_equal(a.torque(m0), b.torque(m0))
    u = np.zeros((3,) + size)
    u[0], u[2] = 0.6, 0.8
    b.set_spintorque((1e12, -2e12, 3e12), 0.5, 0.05, None)
    assert np.allclose(b.torque(u), a.torque(u), rtol=0, atol=1e-12)  # uniform m: Hs = 0


def test_current_mask_scales_each_component(orc):
    n, k = 32, 0.2
    m0 = spiral(n, k)
    a, b, c = (sim(orc, (1, 1, n), m0) for _ in range(3))
    mask = np.ones((3, 1, 1, n), np.float32)
    mask[2, 0, 0, n // 2:] = 0.0           # no current in the second half
    b.set_spintorque((0.0, 0.0, 1e12), 0.5, 0.0, mask)
    c.set_spintorque((0.0, 0.0, 1e12), 0.5, 0.0, None)
    d_b = (b.torque(m0) - a.torque(m0)).reshape(3, n)
    d_c = (c.torque(m0) - a.torque(m0)).reshape(3, n)
    assert np.array_equal(d_b[:, : n // 2], d_c[:, : n // 2])
    assert np.abs(d_b[:, n // 2:]).max() < 1e-15
    mask[0], mask[1] = 0.0, 0.0            # components of the mask along axes without current change nothing
    mask[2] = 1.0
    b.set_spintorque((0.0, 0.0, 1e12), 0.5, 0.0, mask)
    assert np.array_equal(b.torque(m0), c.torque(m0))


# ------------------------------------------------------------------ the Sim commands (no engine)
def test_spintorque_commands_and_their_checks(mm, tmp_path):
    fe = __import__("importlib").import_module("1_amd.frontend")
    for name in ("currentdensity", "spinpolarization", "xi", "currentmask"):
        assert name in fe.COMMANDS and name not in fe.UNSUPPORTED
    s = mm.Sim("abi")
    s.gridsize(8, 4, 2)
    s.cellsize(2e-9, 2e-9, 2e-9)
    s.msat(MS)
    s.aexch(AEX)
    s.alpha(0.02)
    s.currentdensity(3e12, 2e12, 1e12)       # (jz, jy, jx) as CurrentDensity: internal (jx, jy, jz)
    s.spinpolarization(0.4)
    s.xi(0.05)
    s.setm_array(np.ones((3, 2, 4, 8), np.float32))
    with pytest.raises(mm.InputErr):         # the reference stops without a current mask (torque.go:68-71)
        s.params()
    s.currentmask_array(np.ones((3, 2, 4, 8), np.float32))
    p = s.params()
    f32 = [float(np.float32(v)) for v in (1e12, 2e12, 3e12)]   # float32, as the reference's input.j
    assert list(p.current) == f32 and p.spinPol == pytest.approx(0.4) and bool(p.currentMask)
    f = mm.Sim("fused")
    f.gridsize(8, 4, 2)
    f.cellsize(2e-9, 2e-9, 2e-9)
    f.msat(MS)
    f.aexch(AEX)
    f.currentdensity(1e12, 0, 0)
    f.currentmask_array(np.ones((3, 2, 4, 8), np.float32))
    with pytest.raises(mm.InputErr):         # the fused engine has no STT: mode='abi'
        f.params()


# ------------------------------------------------------------------ an applied current (Apply*("j", ...))
def test_applied_current_replaces_the_current_density(orc):
    """simfield.go:61-70, torque.go:42-44: the applied current, evaluated at the
    torque's time, is the current the torque uses; a static one equals
    CurrentDensity of the same value, a pointwise one its interpolation."""
    n, k = 32, 0.25
    m0 = spiral(n, k)
    J0, J1 = np.array([0.0, 0.0, 1e12]), np.array([0.0, 0.0, 3e12])
    a, b = sim(orc, (1, 1, n), m0), sim(orc, (1, 1, n), m0)
    a.set_spintorque(J0, 0.5, 0.0, None)
    b.set_spintorque((0.0, 0.0, 0.0), 0.5, 0.0, None)
    b.set_applied_current([0.0], J0)
    assert np.array_equal(a.torque(m0), b.torque(m0))
    t0, t1 = -2e-12, 6e-12                       # the torque is taken at t = 0
    b.set_applied_current([t0, t1], np.stack([J0, J1]))
    w = (0.0 - t0) / (t1 - t0)
    a.set_spintorque(J0 + w * (J1 - J0), 0.5, 0.0, None)
    assert np.allclose(a.torque(m0), b.torque(m0), rtol=1e-12, atol=0)
    b.set_applied_current()                      # off: back to CurrentDensity (0 here): the LL torque
    c = sim(orc, (1, 1, n), m0)
    assert np.array_equal(b.torque(m0), c.torque(m0))


def test_applied_current_commands(mm):
    s = mm.Sim("abi")
    s.gridsize(8, 4, 1)
    s.cellsize(2e-9, 2e-9, 2e-9)
    s.applypointwise("j", 0.0, 1e12, 0.0, 0.0)
    s.applypointwise("currentdensity", 1e-9, 3e12, 0.0, 0.0)   # (jz, jy, jx): the first is internal Z
    assert np.allclose(s.current_at(0.5e-9), [0.0, 0.0, 2e12])
    s.applystatic("j", 0.0, 0.0, 5e11)
    assert np.allclose(s.current_at(1.0), [5e11, 0.0, 0.0])
    with pytest.raises(mm.InputErr):
        s.applystatic("h", 0, 0, 0)
