"""The autotuner (mm_sim_tune, include/mumax_engine.h) and the launch shapes it
chooses from.

Launch shapes change which block computes which lines, never the arithmetic of
a line or a cell, so every shape must give the same m bit for bit; the plane
pass (for the grids it fits) computes the same transforms in another order,
so it agrees to FP32 rounding.  Each case
runs fixed-step RK4 steps with the default shapes, with every alternative
shape forced through the tuning table (Z forward rows halved and doubled,
fused Z inverse rows halved, 2 blocks per CU for the grid-strided passes), and
after the tuner's own choice.  The three grids take the three convolution
plans: separate 3-D passes, the thin-plane pass and the 2-D pass.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [((12, 32, 32), (4e-9, 4e-9, 4e-9)),    # padded 24 x 64 x 64: Z, Y, X.K.X, Y inverse, Z inverse
         ((16, 256, 24), (4e-9, 4e-9, 4e-9)),   # padded 32 x 512 x 48: the plane pass fits as well
         ((4, 16, 64), (4e-9, 4e-9, 4e-9)),     # padded 8 x 32 x 128: thin-plane pass
         ((1, 32, 128), (3e-9, 3.9e-9, 3.9e-9))]  # 2-D


def make(mm, size, cell):
    s = mm.Sim("fused")
    X, Y, Z = size
    s.gridsize(Z, Y, X)
    s.cellsize(cell[2], cell[1], cell[0])
    s.msat(8e5)
    s.aexch(1.3e-11)
    s.alpha(0.05)
    s.anisuniaxial(0, 0, 1)
    s.k1(5e4)
    s.solvertype("rk4")
    s.dt(2e-14)
    rng = np.random.default_rng(5)
    m = rng.standard_normal((3,) + size)
    s.setm_array((m / np.linalg.norm(m, axis=0)).astype(np.float32))
    return s


def run(mm, size, cell, n=4):
    s = make(mm, size, cell)
    s.steps(n)
    m, shapes = s.m, s.tune_params()
    s.free()
    return m, shapes


@pytest.mark.parametrize("size,cell", CASES)
def test_every_launch_shape_gives_the_same_steps(mm, lib, tmp_path, size, cell):
    print("tune: shapes of grid %s" % (size,), flush=True)
    lib.mm_tune_clear()
    try:
        ref, shapes = run(mm, size, cell)
        assert shapes == {"zr": 0, "er": 0, "yb": 8, "xb": 4, "y2b": 8, "pp": 0, "xv": 0}
        for zr, er, grid, pp, xv in [(1, 1, 2, 0, 1), (2, 0, 16, 1, 0), (0, 0, 4, 0, 2), (0, 0, 8, 0, 3)]:
            table = tmp_path / "tune.txt"
            table.write_text("# test\n%d %d %d %d %d %d %d %d %d %d\n" % (size + (zr, er, grid, grid, grid, pp, xv)))
            lib.mm_tune_clear()
            assert lib.mm_tune_load(str(table).encode()) == 1
            m, shapes = run(mm, size, cell)
            assert shapes["yb"] == shapes["xb"] == shapes["y2b"] == grid
            if shapes["pp"] or shapes["xv"] >= 2:
                # the plane pass and the half-element X.K.X compute the same transforms in another
                # order: FP32 rounding apart
                assert np.abs(m - ref).max() < 1e-6, (zr, er, grid, shapes)
            else:
                assert np.array_equal(m.view(np.uint32), ref.view(np.uint32)), (zr, er, grid, shapes)
        lib.mm_tune_clear()
        s = make(mm, size, cell)
        res = s.tune(reps=3)
        assert "zr" in res and "er" in res
        for k, val in res.items():
            if k == "plane_vs_three":
                assert min(val) > 0
                continue
            v, t_default, t_tuned = val
            assert 0 < t_tuned <= t_default, (k, res)
        s.steps(4)
        if s.tune_params()["pp"] or s.tune_params()["xv"] >= 2:
            assert np.abs(s.m - ref).max() < 1e-6
        else:
            assert np.array_equal(s.m.view(np.uint32), ref.view(np.uint32))
        s.free()
        # the choice is in the table: a new simulation of this size takes it
        s = make(mm, size, cell)
        assert s.tune_params() == {k: res.get(k, (d,))[0] for k, d in zip(s.TUNABLES, (0, 0, 8, 4, 8, 0, 0))}
        s.free()
    finally:
        lib.mm_tune_clear()
        table = os.path.join(os.path.dirname(mm.gpukern.__file__), "tune_gfx950.txt")
        if os.path.exists(table):
            lib.mm_tune_load(table.encode())
