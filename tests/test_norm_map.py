"""The norm map (SetMsat / SetMsatRange / SetMsatEllips, simgeom.go:266-340):
every normalisation scales a cell to its own norm, 0 leaving it empty
(Sim.Normalize, sim.go:230-239; gpu_normalize.cu:43-71).  CPU only: the FP64
oracle's map normalisation and dynamics, and the Sim setters' cell choice.

The oracle's map path is pinned by the normalisation's invariants (norm, direction,
empty cells) and by an identity of the dynamics: the demag and the folded exchange
are linear convolutions whose kernel depends on the displacement only, so a grid
whose map empties every cell outside a sub-box evolves exactly as the sub-box alone
(same trials, m and energy; FP64 rounding apart)."""
import numpy as np
import pytest

KW = dict(cellsize=(3e-9, 4e-9, 4e-9), msat=8e5, aexch=1.3e-11, alpha=0.05, solver="rkdp", maxerror=1e-5,
          field=(np.zeros(1), np.array([0.01, -0.02, 0.03])))


def test_normalize_map_norms_directions_and_empty_cells(orc):
    rng = np.random.default_rng(3)
    N = 600
    m = rng.standard_normal((3, N)) * rng.uniform(0.1, 10.0, N)
    nmap = rng.uniform(0.05, 1.0, N).astype(np.float32)
    nmap[::7] = 0
    m[:, ::11] = 0          # m = 0 in empty cells: 0, not 0/0
    nmap[::11] = 0
    out = orc.normalize_map(m.reshape(-1), nmap).reshape(3, N)
    live = nmap != 0
    norm = np.linalg.norm(out, axis=0)
    assert np.allclose(norm[live], nmap[live].astype(float), rtol=1e-13)
    cos = (out[:, live] * m[:, live]).sum(0) / (norm[live] * np.linalg.norm(m[:, live], axis=0))
    assert np.allclose(cos, 1.0, atol=1e-13)
    assert np.all(out[:, ~live] == 0)
    # a map of ones is the uniform normalisation, bit for bit
    m1 = rng.standard_normal(3 * N)
    assert np.array_equal(orc.normalize_map(m1, np.ones(N, np.float32)), orc.normalize(m1))


def _masked(orc, big, box, m0, nmap_inside=1.0):
    nmap = np.zeros(big, np.float32)
    nmap[box] = nmap_inside
    return orc.sim(m0, size=big, norm_map=nmap, **KW), nmap


def test_masked_grid_evolves_as_its_sub_box(orc):
    rng = np.random.default_rng(1)
    big = (2, 12, 16)
    box = (slice(0, 2), slice(2, 10), slice(3, 13))
    m0 = rng.standard_normal((3,) + big)     # the cells outside are emptied by Run's normalisation
    o, _ = _masked(orc, big, box, m0)
    small = m0[(slice(None),) + box]
    s = orc.sim(small, size=small.shape[1:], **KW)
    for _ in range(3):
        o.run(2e-12)
        s.run(2e-12)
        mb = o.m.reshape((3,) + big)
        inside = mb[(slice(None),) + box]
        assert o.state()["steps"] == s.state()["steps"]
        assert np.abs(inside - s.m.reshape(inside.shape)).max() < 1e-10
        outside = mb.copy()
        outside[(slice(None),) + box] = 0
        assert not outside.any()
        assert abs(o.energy() - s.energy()) < 1e-10 * abs(s.energy())
    # without the map the emptied cells take part: the dynamics differ
    u = orc.sim(m0, size=big, **KW)
    u.run(2e-12)
    assert np.abs(u.m.reshape((3,) + big)[(slice(None),) + box] - o.m.reshape((3,) + big)[(slice(None),) + box]).max() > 1e-3


def test_fractional_map_holds_after_steps_and_runs(orc):
    """Steps() does not normalise first (simrun.go:85-92), each step's end does
    (rk.go:444); Run() also normalises first (simrun.go:24)."""
    rng = np.random.default_rng(5)
    size = (2, 8, 12)
    nmap = rng.uniform(0.2, 1.0, size).astype(np.float32)
    nmap[:, :2, :] = 0
    m0 = rng.standard_normal((3,) + size)
    m0 /= np.linalg.norm(m0, axis=0)
    o = orc.sim(m0, size=size, norm_map=nmap, **KW)
    o.step()
    m = o.m.reshape((3,) + size)
    assert np.allclose(np.linalg.norm(m, axis=0), nmap, rtol=1e-12, atol=0)
    o.run(1e-12)
    m = o.m.reshape((3,) + size)
    assert np.allclose(np.linalg.norm(m, axis=0), nmap, rtol=1e-12, atol=0)
    assert not m[:, :, :2, :].any()


# ------------------------------------------------------------------ the Sim setters (no engine)
@pytest.fixture()
def sim(mm):
    s = mm.Sim("fused")
    s.gridsize(20, 12, 3)          # internal (X, Y, Z) = (3, 12, 20)
    s.cellsize(4e-9, 3e-9, 2e-9)   # internal cell (2, 3, 4) nm
    return s


def test_setmsat_and_range_cells(sim, mm):
    sim.setmsat(5, 4, 1, 0.25)                 # indexed as SetAlpha: internal [1, 4, 5]
    n = sim._norm_map
    assert n.shape == (3, 12, 20) and n[1, 4, 5] == np.float32(0.25)
    assert np.count_nonzero(n != 1) == 1
    sim.setmsatrange(2, 3, 0, 6, 5, 1, 0.0)    # internal [0..1][3..5][2..6]
    assert not n[0:2, 3:6, 2:7].any()
    assert np.count_nonzero(n != 1) == 2 * 3 * 5
    whole = np.random.default_rng(2).uniform(0, 1, (3, 12, 20)).astype(np.float32)
    fresh = mm.Sim("fused")
    fresh.gridsize(20, 12, 3)
    fresh.setmsat_array(whole)
    assert np.array_equal(fresh._norm_map, whole)
    with pytest.raises(mm.InputErr):
        sim.setmsat(20, 0, 0, 0.5)
    with pytest.raises(mm.InputErr):
        sim.setmsatrange(0, 0, 0, 0, 12, 0, 0.5)


def test_setmsatellips_is_the_ellipse_through_every_layer(sim, mm):
    """Cells whose corner (k c_Z, j c_Y) lies inside the ellipse, every layer;
    the reference's float32 range and test give the brute-force set here
    (no corner within 1e-4 of the boundary)."""
    cx, cy, sx, sy = 41e-9, 17e-9, 50e-9, 22e-9
    sim.setmsatellips(cx, cy, sx, sy, 0.5)
    n = sim._norm_map
    j, k = np.meshgrid(np.arange(12), np.arange(20), indexing="ij")
    q = ((k * 4e-9 - cx) / (sx / 2)) ** 2 + ((j * 3e-9 - cy) / (sy / 2)) ** 2
    assert np.abs(q - 1).min() > 1e-4
    want = np.where(q < 1, np.float32(0.5), np.float32(1))
    assert 20 < np.count_nonzero(q < 1) < 12 * 20
    for x in range(3):
        assert np.array_equal(n[x], want)
    with pytest.raises(mm.InputErr):
        sim.setmsatellips(2e-9, 17e-9, 20e-9, 10e-9, 0.5)   # reaches past the grid's first column


def test_setmsat_commands_reach_the_sim(tmp_path):
    fe = __import__("importlib").import_module("1_amd.frontend")
    for name in ("setmsat", "setmsatrange", "setmsatellips"):
        assert name in fe.COMMANDS and name not in fe.UNSUPPORTED
    sh = fe.Shell(str(tmp_path))
    for line in ("gridsize 8 4 2", "cellsize 2e-9 2e-9 2e-9", "setmsat 3 2 1 0.5", "setmsatrange 0 0 0 1 0 0 0"):
        w = line.split()
        sh.call(w[0], w[1:])
    n = sh.sim._norm_map
    assert n[1, 2, 3] == np.float32(0.5) and not n[0, 0, 0:2].any()
