"""The semi-analytic solvers in the engine's ABI mode (SemiAnal1 / SemiAnal2,
semianal.go: calcHeff and the drop-in gpu_anal_fw_step) against the FP64 oracle
running the same solver: m after 40 steps of a film and a 3-D grid."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", ["semianal1", "semianal2"])
@pytest.mark.parametrize("size", [(1, 32, 64), (4, 16, 24)])
def test_semianal_steps_vs_oracle(mm, orc, solver, size):
    print("semi-analytic: 40 %s steps of %s" % (solver, size), flush=True)
    rng = np.random.default_rng(2)
    m0 = np.zeros((3,) + size)
    m0[0], m0[2] = 1.0, 0.4
    m0 += 0.2 * rng.standard_normal(m0.shape)
    m0 = (m0 / np.linalg.norm(m0, axis=0)).astype(np.float32)
    cell = (3e-9, 4e-9, 4e-9)
    s = mm.Sim("abi")
    X, Y, Z = size
    s.gridsize(Z, Y, X)
    s.cellsize(cell[2], cell[1], cell[0])
    s.msat(8e5)
    s.aexch(1.3e-11)
    s.alpha(0.1)
    s.solvertype(solver)
    s.dt(2e-14)
    s.applystatic("field", 0.1, 0.05, 0.0)
    s.setm_array(m0)
    s.steps(40)
    o = orc.sim(m0, size=size, cellsize=cell, msat=8e5, aexch=1.3e-11, alpha=0.1, solver=solver, dt=2e-14,
                field=(np.zeros(1), np.array([0.0, 0.05, 0.1])))
    for _ in range(40):
        o.step()
    d = float(np.abs(s.m.reshape(-1) - o.m).max())
    moved = float(np.abs(o.m - m0.reshape(-1)).max())
    print("max |dm| vs the FP64 oracle %.3g (m moved by %.3g)" % (d, moved))
    assert d < 1e-5 and moved > 1e-2
    assert s.state()["steps"] == 40 and s.state()["time"] == 0.0
    s.free()
