"""The norm map in the engine (SetMsat, simgeom.go:266-340; Sim.Normalize,
sim.go:230-239; gpu_normalize.cu:43-71) against the FP64 oracle with the same
map.  Fixed-step runs normalise in the fused Z inverse epilogue, adaptive runs in
the commit kernel, Run() first in the start normalisation; the ABI mode calls
gpu_normalize_map.  Every plan kind: thin plane, separate 3-D passes, 2-D."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CELL = (4e-9, 4e-9, 4e-9)


def smooth_m(size, seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = size
    x, y, z = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    m = np.zeros((3,) + size)
    for c in range(3):
        a = rng.standard_normal(4)
        m[c] = a[0] + a[1] * np.cos(2 * np.pi * y / Y + a[3]) + a[2] * np.sin(2 * np.pi * z / Z)
    m[2] += 2.0
    return (m / np.linalg.norm(m, axis=0)).astype(np.float32)


def norm_map(size, seed):
    """norms in [0.3, 1], a band of empty cells along the first Y rows"""
    n = np.random.default_rng(seed).uniform(0.3, 1.0, size).astype(np.float32)
    n[:, : max(1, size[1] // 8), :] = 0
    return n


def cfg_of(size, m0, nmap, **kw):
    c = dict(size=size, cellsize=CELL, msat=8e5, aexch=1.3e-11, alpha=0.05, solver="rkdp", maxerror=1e-5,
             field=("static", (0.01, -0.02, 0.03)), m0=m0, norm_map=nmap)
    c.update(kw)
    return c


def check_norms(m, nmap):
    n = np.linalg.norm(m.reshape((3,) + nmap.shape), axis=0)
    live = nmap != 0
    assert np.abs(n[live] - nmap[live]).max() < 2e-6
    assert not m.reshape((3,) + nmap.shape)[:, ~live].any()


@pytest.mark.parametrize("mode", ["fused", "abi"])
@pytest.mark.parametrize("size", [(2, 16, 32), (12, 16, 32), (1, 32, 64)])
def test_norm_map_rk4_steps_vs_oracle(mm, orc, size, mode):
    print("norm map: 20 RK4 steps of %s, %s" % (size, mode), flush=True)
    m0 = smooth_m(size, 1)
    nmap = norm_map(size, 2)
    cfg = cfg_of(size, m0, nmap, solver="rk4", dt=2e-14)
    s = mm.workloads.make_sim(cfg, mode)
    s.steps(20)
    o = orc.sim(m0, **mm.workloads.oracle_kwargs(cfg))
    for _ in range(20):
        o.step()
    d = float(np.abs(s.m.reshape(-1) - o.m).max())
    print("max |dm| vs the FP64 oracle: %.3g" % d)
    assert d < 1e-5
    check_norms(s.m, nmap)
    s.free()


@pytest.mark.parametrize("mode", ["fused", "abi"])
@pytest.mark.parametrize("size", [(2, 16, 32), (1, 32, 64)])
def test_norm_map_dormand_prince_run_replayed(mm, orc, size, mode):
    """Run() normalises to the map first (simrun.go:24), then each accepted
    step (rk.go:444); the trial log replayed in the oracle."""
    print("norm map: RKDP run of %s, %s" % (size, mode), flush=True)
    m0 = smooth_m(size, 3)
    nmap = norm_map(size, 4)
    cfg = cfg_of(size, m0, nmap)
    s = mm.workloads.make_sim(cfg, mode)
    s.run(5e-12)
    dts, acc = s.trial_log()
    o = orc.sim(m0, **mm.workloads.oracle_kwargs(cfg))
    o.m = orc.normalize_map(o.m, nmap)
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
    check_norms(s.m, nmap)
    s.free()


def test_masked_grid_runs_as_its_sub_box(mm):
    """A grid whose map empties every cell outside a sub-box runs as the sub-box
    alone (the demag and exchange kernels depend on the displacement only):
    within the FP32 envelope of the two FFT sizes, the outside exactly 0."""
    big, box = (4, 24, 40), (slice(0, 4), slice(4, 20), slice(6, 34))
    m0 = np.random.default_rng(7).standard_normal((3,) + big).astype(np.float32)
    nmap = np.zeros(big, np.float32)
    nmap[box] = 1
    a = mm.workloads.make_sim(cfg_of(big, m0, nmap, solver="rk4", dt=2e-14))
    small = np.ascontiguousarray(m0[(slice(None),) + box])
    b = mm.workloads.make_sim(cfg_of(small.shape[1:], small, None, solver="rk4", dt=2e-14))
    a.run(1e-12)
    b.run(1e-12)
    ma = a.m.reshape((3,) + big)
    d = float(np.abs(ma[(slice(None),) + box] - b.m.reshape(small.shape)).max())
    print("masked grid vs its sub-box: max |dm| %.3g" % d)
    assert d < 1e-5
    ma[(slice(None),) + box] = 0
    assert not ma.any()
    a.free()
    b.free()


def test_norm_map_member_in_a_batch_equals_alone_and_ones_is_no_map(mm):
    size = (2, 16, 32)
    cfgs = [cfg_of(size, smooth_m(size, 8), norm_map(size, 9)),
            cfg_of((1, 32, 64), smooth_m((1, 32, 64), 10), norm_map((1, 32, 64), 11)),
            cfg_of(size, smooth_m(size, 12), None)]
    ref = []
    for c in cfgs:
        s = mm.workloads.make_sim(c)
        s.run(3e-12)
        ref.append(s.m)
        s.free()
    sims = [mm.workloads.make_sim(c) for c in cfgs]
    b = mm.Batch(sims)
    b.run([3e-12] * len(sims))
    b.free()
    for m, s in zip(ref, sims):
        assert np.array_equal(s.m, m)
        s.free()
    ones = mm.workloads.make_sim(dict(cfgs[2], norm_map=np.ones(size, np.float32)))
    ones.run(3e-12)
    assert np.array_equal(ones.m, ref[2])
    ones.free()
