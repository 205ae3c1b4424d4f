"""The semi-analytic solvers (SemiAnal1 / SemiAnal2, semianal.go; solver.go:46-49)
in the FP64 oracle, and the Sim's checks.  CPU only.

The step kernel (gpu_anal.cu:23-84) is the exact solution of the LL equation in a
constant field; it is pinned against an RK4 integration of that equation.  The
solvers advance m with the field of a second state half a step apart (a
leapfrog), so they are second order: against a fine RK4 run of the same film,
the error after a fixed time falls about fourfold when dt halves.  (Their steps
do not advance the time, so the applied field here is static.)"""
import numpy as np
import pytest


def ll_rk4(m, h, alpha, dt, n=4000):
    # dm/dt = (-m x h - alpha m x (m x h)) / (1 + alpha^2): the kernel's precession sense
    def f(m):
        mxh = np.cross(m.T, h.T).T
        return (-mxh - alpha * np.cross(m.T, mxh.T).T) / (1 + alpha * alpha)

    d = dt / n
    for _ in range(n):
        k1 = f(m)
        k2 = f(m + d / 2 * k1)
        k3 = f(m + d / 2 * k2)
        k4 = f(m + d * k3)
        m = m + d / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return m


def test_anal_step_is_the_constant_field_solution(orc):
    rng = np.random.default_rng(3)
    N = 6
    m = rng.standard_normal((3, N))
    m /= np.linalg.norm(m, axis=0)
    h = rng.standard_normal((3, N)) * 2.0
    h[:2, 0] = 0.0                      # a field along z: the fixed frame
    for alpha, dt in [(0.0, 0.7), (0.1, 0.3), (0.5, 1.1)]:
        out = orc.anal_step(m.reshape(-1), h.reshape(-1), dt, alpha).reshape(3, N)
        assert np.abs(out - ll_rk4(m, h, alpha, dt)).max() < 1e-9
    z = np.zeros(3 * N)
    assert not orc.anal_step(z, h.reshape(-1), 0.5, 0.1).any()   # m = 0 stays 0


def film(orc, solver, dt, m0):
    return orc.sim(m0, size=(1, 8, 12), cellsize=(3e-9, 4e-9, 4e-9), msat=8e5, aexch=1.3e-11, alpha=0.1,
                   solver=solver, dt=dt, field=(np.zeros(1), np.array([0.0, 0.05, 0.1])))


def ll_film_rk4(o, m, alpha, dt, n):
    """RK4 of dm/dt = (-m x H - alpha m x (m x H)) / (1 + alpha^2) with H = H_eff(m)
    of oracle sim o: the LL equation in the kernel's precession sense (the
    reference's semi-analytic step turns m the other way from its RK torque)."""
    shape = m.shape

    def f(m):
        h = o.heff(m.astype(np.float32)).reshape(3, -1)
        mm = m.reshape(3, -1)
        mxh = np.cross(mm.T, h.T).T
        return ((-mxh - alpha * np.cross(mm.T, mxh.T).T) / (1 + alpha * alpha)).reshape(shape)

    for _ in range(n):
        k1 = f(m)
        k2 = f(m + dt / 2 * k1)
        k3 = f(m + dt / 2 * k2)
        k4 = f(m + dt * k3)
        m = m + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return m


@pytest.mark.parametrize("solver", ["semianal1", "semianal2"])
def test_semianal_solvers_are_second_order(orc, solver):
    rng = np.random.default_rng(8)
    m0 = np.zeros((3, 1, 8, 12))
    m0[0], m0[2] = 1.0, 0.4
    m0 += 0.2 * rng.standard_normal(m0.shape)
    m0 /= np.linalg.norm(m0, axis=0)
    T, alpha = 1e-12, 0.1
    o = film(orc, "rk4", 1e-15, m0)
    unit_t = o.state()["unit_time"]
    ref = ll_film_rk4(o, m0.astype(np.float64), alpha, 2.5e-15 / unit_t, round(T / 2.5e-15))
    errs = []
    for dt in (2e-14, 1e-14):
        s = film(orc, solver, dt, m0)
        for _ in range(round(T / dt)):
            s.step()
        assert s.state()["time"] == 0.0           # the time is not advanced (semianal.go)
        errs.append(float(np.abs(s.m - ref.reshape(-1)).max()))
    print(solver, "max |dm| at dt 2e-14, 1e-14:", errs)
    assert errs[1] < 2e-3 and 3.0 < errs[0] / errs[1] < 5.0


def test_sim_refuses_run_and_the_fused_mode(mm):
    s = mm.Sim("abi")
    s.gridsize(8, 4, 1)
    s.cellsize(2e-9, 2e-9, 2e-9)
    s.solvertype("semianal1")
    with pytest.raises(mm.InputErr):
        s.run(1e-12)
    with pytest.raises(mm.InputErr):
        s.relax()
    f = mm.Sim("fused")
    f.gridsize(8, 4, 1)
    f.cellsize(2e-9, 2e-9, 2e-9)
    f.msat(8e5)
    f.aexch(1.3e-11)
    f.solvertype("semianal2")
    with pytest.raises(mm.InputErr):
        f.params()
