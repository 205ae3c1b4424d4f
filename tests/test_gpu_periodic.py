"""Periodic boundaries in the engine (Periodic, simgeom.go:84-88) against the
FP64 oracle.

The reference keeps a periodic axis at its size in the kernel grid (padSize,
kernel.go:79-90), adds the periodic images into the kernel (demag.go:48-79)
and convolves circularly over the period.  The fused engine keeps its padded
lengths (>= 2N) and folds the wrap into the kernel instead: the padded index of
displacement d holds the reference's kernel at d mod N, so the linear
convolution over the padded grid is the circular one.  The ABI mode replays
the reference's own layout (the drop-in FFT plan of the unpadded period).
Fields are compared relative to their maximum (FP32 transforms), trajectories
by max |dm| against the oracle replaying the same steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# size (internal X, Y, Z), periodic images, cell (m)
CASES = [
    ((8, 16, 32), (0, 2, 1), (3e-9, 4e-9, 4e-9)),    # 3-D, periodic Y and Z (thin-plane pass)
    ((4, 8, 16), (1, 1, 1), (4e-9, 4e-9, 4e-9)),     # periodic along every axis
    ((1, 32, 64), (0, 3, 3), (3e-9, 4e-9, 4e-9)),    # film, periodic in plane (2-D plan)
    ((16, 64, 24), (2, 0, 0), (4e-9, 4e-9, 4e-9)),   # periodic along the thickness only (separate 3-D passes)
    ((2, 48, 12), (0, 1, 2), (4e-9, 3e-9, 5e-9)),    # 3 * 2^a padded lengths
]


def random_m(size, seed):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((3,) + size)
    return (m / np.linalg.norm(m, axis=0)).astype(np.float32)


def cfg_of(size, per, cell, m0, **kw):
    c = dict(size=size, periodic=per, cellsize=cell, msat=8e5, aexch=1.3e-11, alpha=0.05, solver="rkdp",
             maxerror=1e-5, field=("static", (0.01, -0.02, 0.03)), m0=m0)
    c.update(kw)
    return c


@pytest.mark.parametrize("mode", ["fused", "abi"])
@pytest.mark.parametrize("size,per,cell", CASES)
def test_periodic_heff_vs_oracle(mm, orc, size, per, cell, mode):
    print("periodic: heff %s %s images %s" % (mode, size, per), flush=True)
    m0 = random_m(size, 2)
    cfg = cfg_of(size, per, cell, m0)
    s = mm.workloads.make_sim(cfg, mode)
    o = orc.sim(m0, **mm.workloads.oracle_kwargs(cfg))
    h = s.heff(m0)
    href = o.heff(m0).reshape(h.shape)
    err = float(np.abs(h - href).max() / np.abs(href).max())
    print("periodic heff: relative difference %.3g" % err)
    assert err < 5e-6
    # the images change the field: the same grid without them differs
    o0 = orc.sim(m0, **mm.workloads.oracle_kwargs(dict(cfg, periodic=None)))
    assert np.abs(o0.heff(m0).reshape(h.shape) - href).max() > 1e-3 * np.abs(href).max()
    s.free()


def test_periodic_odd_period_in_abi_mode(mm, orc):
    """Odd periods have asymmetric image ranges (demag.go:42-53) and kernels
    without parity along the axis: the ABI mode (the reference's own layout)
    computes them; the fused mode refuses them."""
    size, per, cell = (3, 5, 12), (1, 1, 0), (4e-9, 4e-9, 4e-9)
    m0 = random_m(size, 4)
    cfg = cfg_of(size, per, cell, m0)
    s = mm.workloads.make_sim(cfg, "abi")
    o = orc.sim(m0, **mm.workloads.oracle_kwargs(cfg))
    h = s.heff(m0)
    href = o.heff(m0).reshape(h.shape)
    assert np.abs(h - href).max() < 5e-6 * np.abs(href).max()
    s.free()
    f = mm.workloads.make_sim(cfg, "fused")
    with pytest.raises(mm.InputErr):
        f.heff(m0)


@pytest.mark.parametrize("size,per,cell", [CASES[0], CASES[2]])
def test_periodic_dormand_prince_run_replayed(mm, orc, size, per, cell):
    """30 adaptive Dormand-Prince steps; the engine's trial log replayed in the
    FP64 oracle: max |dm| <= 1e-5."""
    print("periodic: 30 RKDP steps of %s images %s" % (size, per), flush=True)
    m0 = random_m(size, 5)
    cfg = cfg_of(size, per, cell, m0, maxerror=1e-4)
    s = mm.workloads.make_sim(cfg)
    s.steps(30)
    dts, acc = s.trial_log()
    o = orc.sim(m0, **mm.workloads.oracle_kwargs(cfg))
    i = 0
    while i < len(dts):
        j = i
        while not acc[j]:
            j += 1
        o.step_replay(dts[i:j + 1], acc[i:j + 1])
        i = j + 1
    assert s.state()["steps"] == 30 == o.state()["steps"]
    dm = float(np.abs(s.m.reshape(-1) - o.m).max())
    print("periodic run: %d trials, max|dm| vs FP64 replay %.3g" % (len(dts), dm))
    assert dm < 1e-5
    s.free()


def test_periodic_member_in_a_batch_equals_alone(mm):
    """A periodic and a free simulation of one size share the padded length
    class but not the kernel: each computes the same alone and in a batch."""
    size, per, cell = CASES[0]
    cfgs = [cfg_of(size, per, cell, random_m(size, 7)), cfg_of(size, (0, 0, 0), cell, random_m(size, 7))]
    ref = []
    for c in cfgs:
        s = mm.workloads.make_sim(c)
        s.run(2e-12)
        ref.append((s.m, s.trial_log()[0]))
        s.free()
    sims = [mm.workloads.make_sim(c) for c in cfgs]
    b = mm.Batch(sims)
    b.run([2e-12] * len(sims))
    b.free()
    for (m, dt), s in zip(ref, sims):
        assert np.array_equal(s.trial_log()[0], dt)
        assert np.array_equal(s.m, m)
        s.free()
    assert not np.array_equal(ref[0][0], ref[1][0])


def test_periodic_kernel_is_its_own_spectrum(mm):
    """Spectra are shared per (size, cell, kernel, images): a periodic
    simulation does not take the free one's spectrum."""
    lib = mm.load()
    size, per, cell = CASES[1]
    n0 = lib.mm_kernel_cache_size()
    a = mm.workloads.make_sim(cfg_of(size, per, cell, random_m(size, 1)))
    b = mm.workloads.make_sim(cfg_of(size, (0, 0, 0), cell, random_m(size, 1)))
    c = mm.workloads.make_sim(cfg_of(size, per, cell, random_m(size, 2)))
    for s in (a, b, c):
        s.heff(random_m(size, 3))
    assert lib.mm_kernel_cache_size() == n0 + 2
    for s in (a, b, c):
        s.free()
    assert lib.mm_kernel_cache_size() == n0
