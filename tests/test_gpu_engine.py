"""GPU parity of the device-resident engine (include/mumax_engine.h) against
the FP64 oracle.

Tolerances: the engine computes in FP32 (like the reference); fields and
torques are compared relative to their maximum (FP32 FFT round-off of a few
1e-7 x log2 P), trajectories by max |dm| with the oracle replaying the GPU's
accepted-dt schedule (trajectories are compared at the same accepted-dt sequence), bar 1e-5.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def random_m(size, seed):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((3,) + size)
    return (m / np.linalg.norm(m, axis=0)).astype(np.float32)


def smooth_m(size, seed):
    """A smooth, exchange-friendly state (random low-order modes)."""
    rng = np.random.default_rng(seed)
    X, Y, Z = size
    x, y, z = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    m = np.zeros((3,) + size)
    for c in range(3):
        a = rng.standard_normal(4)
        m[c] = a[0] + a[1] * np.cos(2 * np.pi * y / Y + a[3]) + a[2] * np.sin(2 * np.pi * z / Z)
    m[2] += 2.0
    return (m / np.linalg.norm(m, axis=0)).astype(np.float32)


def engine(mm, size, cell, mode="fused", solver="rkdp", dt=0.0, maxerror=1e-5, alpha=0.05, k1=0.0, axis=None,
           field=None, m0=None, msat=8e5, aexch=1.3e-11):
    s = mm.Sim(mode)
    X, Y, Z = size
    s.gridsize(Z, Y, X)
    s.cellsize(cell[2], cell[1], cell[0])
    s.msat(msat)
    s.aexch(aexch)
    s.alpha(alpha)
    if axis is not None:
        s.anisuniaxial(axis[2], axis[1], axis[0])
        s.k1(k1)
    s.solvertype(solver)
    s.maxerror(maxerror)
    if dt:
        s.dt(dt)
    if field is not None:
        s.applystatic("field", field[2], field[1], field[0])
    s.setm_array(m0)
    return s


def oracle(orc, size, cell, m0, solver="rkdp", dt=0.0, maxerror=1e-5, alpha=0.05, k1=0.0, axis=None, field=None,
           msat=8e5, aexch=1.3e-11):
    kw = {}
    if axis is not None:
        kw.update(k1=k1, anis_axis=axis)
    if field is not None:
        kw["field"] = (np.zeros(1), np.asarray(field, float))
    return orc.sim(m0, size=size, cellsize=cell, msat=msat, aexch=aexch, alpha=alpha, solver=solver, dt=dt,
                   maxerror=maxerror, **kw)


CASES = [
    ((1, 32, 128), (3e-9, 3.90625e-9, 3.90625e-9)),   # film, 2-D path
    ((4, 16, 32), (4e-9, 4e-9, 4e-9)),                # cuboid, 3-D path
    ((3, 12, 20), (2e-9, 3e-9, 2.5e-9)),              # non-power-of-two data sizes
    ((1, 24, 40), (2e-9, 2e-9, 2e-9)),                # padded 48 x 96: 3 * 2^a lengths
    ((2, 48, 80), (3e-9, 3e-9, 3e-9)),                # padded 4 x 96 x 192
    ((6, 40, 24), (3e-9, 2e-9, 2.5e-9)),              # padded 12 x 96 x 48
    ((1, 96, 384), (3e-9, 3e-9, 3e-9)),               # padded 192 x 768
    # full 3-D path (padded X >= 24: Z forward, Y forward, X.K.X with the
    # compressed kernel's sign restoration, Y inverse, fused Z inverse)
    ((12, 32, 32), (4e-9, 4e-9, 4e-9)),               # padded 24 x 64 x 64 (X = 3 * 2^3)
    ((16, 64, 64), (4e-9, 4e-9, 4e-9)),               # padded 32 x 128 x 128
    ((10, 20, 48), (2.5e-9, 3e-9, 2e-9)),             # padded 24 x 48 x 96, anisotropic cells
    ((24, 12, 16), (3e-9, 3e-9, 3e-9)),               # padded 48 x 24 x 32 (thick)
]
CASES_3D = CASES[7:]


def test_padded_lengths():
    """Each axis is padded to the smallest 2^a or 3 * 2^a >= 2N (1 for N = 1):
    the same linear convolution as the reference's 2N (kernel.go:79-90)."""
    import importlib

    mm = importlib.import_module("1_amd")
    for size, P in (((1, 24, 40), (1, 48, 96)), ((3, 12, 20), (6, 24, 48)), ((4, 16, 32), (8, 32, 64)),
                    ((6, 40, 24), (12, 96, 48)), ((7, 100, 129), (16, 256, 384))):
        s = engine(mm, size, (3e-9, 3e-9, 3e-9), m0=random_m(size, 0))
        st = s.state()
        assert (int(st["px"]), int(st["py"]), int(st["pz"])) == P
        s.free()


@pytest.mark.parametrize("size,cell", CASES)
def test_kernel_spectrum_vs_oracle(mm, orc, size, cell):
    m0 = random_m(size, 1)
    s = engine(mm, size, cell, m0=m0)
    st = s.state()
    P = (int(st["px"]), int(st["py"]), int(st["pz"]))
    got = s.kernel_spectrum()
    u = orc.units(8e5, 1.3e-11)
    ci = np.asarray(cell) / u["length"]
    K = orc.facekernel6(P, ci)
    orc.add_exch6_kernel(K, ci)
    ref = orc.conv_new(size, P, K).kernel_spectrum()
    nK = got.shape[0]
    ref = ref[:nK]
    assert np.abs(got - ref).max() < 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("mode", ["fused", "abi"])
@pytest.mark.parametrize("size,cell", CASES)
def test_heff_and_torque_vs_oracle(mm, orc, size, cell, mode):
    m0 = random_m(size, 2)
    field = (0.01, -0.02, 0.03)
    axis = (0.0, 0.6, 0.8)
    s = engine(mm, size, cell, mode=mode, m0=m0, field=field, axis=axis, k1=3e5)
    o = oracle(orc, size, cell, m0, field=field, axis=axis, k1=3e5)
    h = s.heff(m0)
    href = o.heff(m0).reshape(h.shape)
    assert np.abs(h - href).max() < 5e-6 * np.abs(href).max()
    k = s.torque(m0)
    kref = o.torque(m0).reshape(k.shape)
    assert np.abs(k - kref).max() < 5e-6 * np.abs(href).max()


@pytest.mark.parametrize("size,cell", CASES[:2] + CASES_3D)
def test_rk4_steps_vs_oracle(mm, orc, size, cell):
    m0 = smooth_m(size, 3)
    dt = 2e-14
    axis = (1.0, 0.0, 0.0)
    s = engine(mm, size, cell, solver="rk4", dt=dt, m0=m0, axis=axis, k1=5e5, alpha=0.1, msat=1e6, aexch=1.5e-11)
    o = oracle(orc, size, cell, m0, solver="rk4", dt=dt, axis=axis, k1=5e5, alpha=0.1, msat=1e6, aexch=1.5e-11)
    s.steps(20)
    for _ in range(20):
        o.step()
    got = s.m
    ref = o.m.reshape(got.shape)
    assert np.abs(got - ref).max() < 1e-5
    st = s.state()
    assert st["steps"] == 20
    assert st["time"] == pytest.approx(o.state()["time"], rel=1e-6)


def pass_counts(mm, s, steps=1):
    """Launch counts per pass class over `steps` profiled steps of `s`:
    [Z forward, Y forward, X.K.X / plane / 2-D, Y inverse, Z inverse, ...]."""
    lib = mm.load()
    s._init()  # the device simulation is created lazily
    assert s._ptr
    ms = np.zeros(8)
    n = np.zeros(8, np.int32)
    lib.mm_sim_profile_steps(s._ptr, steps, ms.ctypes.data, n.ctypes.data)
    return n


@pytest.mark.parametrize("size,cell", CASES_3D)
def test_full_3d_path_runs_separate_passes(mm, size, cell):
    """The CASES_3D grids take the separate Y forward, X.K.X and Y inverse
    passes (not the thin-plane or 2-D pass), so the parity tests above pin them."""
    s = engine(mm, size, cell, solver="rk4", dt=1e-14, m0=random_m(size, 0))
    n = pass_counts(mm, s)
    assert n[1] == n[2] == n[3] == 4 and n[0] == n[4] == 4, n
    s.free()


def test_cuboid_256x256x16_vs_oracle(mm, lib, orc, tmp_path):
    """The benchmark's workload at full size (256x256x16 cells of 4 nm, padded
    32 x 512 x 512) through the separate passes (k_zfwd / k_ypass<512> /
    k_xmul<32> / k_zinv_epi): H_eff and the torque of the seeded random m0
    (5e-6 relative to max |H|), then 20 fixed-dt RK4 steps (1e-14 s) against
    the FP64 oracle (max |dm| <= 1e-5)."""
    cfg = mm.workloads.cuboid_rk4()
    force_plane_pass(mm, lib, tmp_path, cfg["size"], on=0)
    try:
        _cuboid_separate_passes(mm, orc, cfg)
    finally:
        restore_tuning(mm, lib)


def _cuboid_separate_passes(mm, orc, cfg):
    s = mm.workloads.make_sim(cfg)
    o = orc.sim(cfg["m0"], **mm.workloads.oracle_kwargs(cfg))
    m0 = cfg["m0"]
    h = s.heff(m0)
    href = o.heff(m0).reshape(h.shape)
    eh = float(np.abs(h - href).max() / np.abs(href).max())
    k = s.torque(m0)
    kref = o.torque(m0).reshape(k.shape)
    ek = float(np.abs(k - kref).max() / np.abs(href).max())
    s.steps(20)
    for _ in range(20):
        o.step()
    dm = float(np.abs(s.m - o.m.reshape(s.m.shape)).max())
    print("cuboid 256x256x16: H_eff rel %.3g, torque rel %.3g, 20 RK4 steps max|dm| %.3g" % (eh, ek, dm))
    assert eh < 5e-6 and ek < 5e-6
    assert dm < 1e-5
    assert s.state()["time"] == pytest.approx(o.state()["time"], rel=1e-6)
    n = pass_counts(mm, s)
    assert n[1] == n[3] == 4, n
    s.free()


# grids the register-resident plane pass fits (padded Y 512, padded X 16 or 32):
# Y forward, X.K.X and Y inverse in one pass, block 0 packing the real planes
# kz = 0 and kz = Q - 1 through their Hermitian symmetry
PLANE_CASES = [
    ((16, 256, 32), (4e-9, 4e-9, 4e-9)),       # padded 32 x 512 x 64 (X 32: two X sub-rounds)
    ((13, 200, 20), (3e-9, 2.5e-9, 4e-9)),     # padded 32 x 512 x 48: Nx < PX/2, Ny < 256, Z 3 * 2^a
    ((8, 256, 64), (4e-9, 4e-9, 4e-9)),        # padded 16 x 512 x 128 (X 16)
    ((7, 193, 6), (3e-9, 3e-9, 3e-9)),         # padded 16 x 512 x 12: few planes
]


def force_plane_pass(mm, lib, tmp_path, size, on=1):
    """New simulations of `size` take the plane pass (tuning table entry)."""
    t = tmp_path / "plane.txt"
    t.write_text("%d %d %d 0 0 8 4 8 %d\n" % (tuple(size) + (on,)))
    lib.mm_tune_clear()
    assert lib.mm_tune_load(str(t).encode()) == 1


def restore_tuning(mm, lib):
    import os

    lib.mm_tune_clear()
    if os.path.exists(mm.gpukern.TUNE_TABLE):
        lib.mm_tune_load(mm.gpukern.TUNE_TABLE.encode())


@pytest.mark.parametrize("size,cell", PLANE_CASES)
def test_plane_pass_vs_oracle(mm, lib, orc, tmp_path, size, cell):
    """H_eff and the torque through the plane pass (5e-6 relative to max |H|),
    then 10 RK4 steps (1e-5), against the FP64 oracle; the pass counts show
    that the separate Y passes did not run."""
    print("plane pass: grid %s" % (size,), flush=True)
    force_plane_pass(mm, lib, tmp_path, size)
    try:
        m0 = random_m(size, 7)
        s = engine(mm, size, cell, solver="rk4", dt=2e-14, m0=m0, k1=3e5, axis=(0, 0, 1))
        o = oracle(orc, size, cell, m0, solver="rk4", dt=2e-14, k1=3e5, axis=(0, 0, 1))
        h = s.heff(m0)
        assert s.tune_params()["pp"] == 1
        href = o.heff(m0).reshape(h.shape)
        k = s.torque(m0)
        kref = o.torque(m0).reshape(k.shape)
        eh = float(np.abs(h - href).max() / np.abs(href).max())
        ek = float(np.abs(k - kref).max() / np.abs(href).max())
        s.steps(10)
        for _ in range(10):
            o.step()
        dm = float(np.abs(s.m - o.m.reshape(s.m.shape)).max())
        print("plane pass %s: H_eff rel %.3g, torque rel %.3g, 10 RK4 steps max|dm| %.3g" % (size, eh, ek, dm))
        assert eh < 5e-6 and ek < 5e-6
        assert dm < 1e-5
        n = pass_counts(mm, s)
        assert n[1] == n[3] == 0 and n[2] == n[0] == n[4] == 4, n
        s.free()
    finally:
        restore_tuning(mm, lib)


def test_cuboid_256x256x16_plane_pass_vs_oracle(mm, lib, orc, tmp_path):
    """The benchmark's cuboid through the plane pass (256 blocks for its 257
    planes): H_eff and torque (5e-6), 20 RK4 steps (1e-5) against the FP64 oracle."""
    cfg = mm.workloads.cuboid_rk4()
    force_plane_pass(mm, lib, tmp_path, cfg["size"])
    try:
        s = mm.workloads.make_sim(cfg)
        o = orc.sim(cfg["m0"], **mm.workloads.oracle_kwargs(cfg))
        m0 = cfg["m0"]
        h = s.heff(m0)
        assert s.tune_params()["pp"] == 1
        href = o.heff(m0).reshape(h.shape)
        k = s.torque(m0)
        kref = o.torque(m0).reshape(k.shape)
        eh = float(np.abs(h - href).max() / np.abs(href).max())
        ek = float(np.abs(k - kref).max() / np.abs(href).max())
        s.steps(20)
        for _ in range(20):
            o.step()
        dm = float(np.abs(s.m - o.m.reshape(s.m.shape)).max())
        print("cuboid 256x256x16, plane pass: H_eff rel %.3g, torque rel %.3g, 20 RK4 steps max|dm| %.3g" % (eh, ek, dm))
        assert eh < 5e-6 and ek < 5e-6
        assert dm < 1e-5
        s.free()
    finally:
        restore_tuning(mm, lib)


def replay_log(o, dts, acc, start=0, stop=None):
    """Replay trials [start, stop) of a GPU trial log (dt, accepted) step by
    step in oracle sim `o`; returns the index reached."""
    i, stop = start, len(dts) if stop is None else stop
    while i < stop:
        j = i
        while not acc[j] and j - i < 9:  # a step: rejected trials, then the accepted one (MAX_STEP_TRIALS 10)
            j += 1
        o.step_replay(dts[i:j + 1], acc[i:j + 1])
        i = j + 1
    return i


def normalise_oracle(o):
    m3 = o.m.reshape(3, -1)
    o.m = (m3 / np.linalg.norm(m3, axis=0)).reshape(-1)


def replay_in_oracle(mm, orc, cfg, s):
    """Replay the GPU's trial log step by step in the oracle `orc` started
    from the same m0 (normalised, as Run() does first, simrun.go:24); returns
    the oracle sim."""
    dts, acc = s.trial_log()
    o = orc.sim(cfg["m0"], **mm.workloads.oracle_kwargs(cfg))
    normalise_oracle(o)
    replay_log(o, dts, acc)
    return o


@pytest.mark.parametrize("mode", ["fused", "abi"])
def test_dp_run_replayed_in_oracle(mm, orc, mode):
    cfg = mm.workloads.film_sp4_like(seed=1)
    s = mm.workloads.make_sim(cfg, mode)
    n = s.run(0.05e-9)
    assert n > 5
    o = replay_in_oracle(mm, orc, cfg, s)
    got = s.m
    ref = o.m.reshape(got.shape)
    assert np.abs(got - ref).max() < 1e-5
    assert s.state()["steps"] == o.state()["steps"]


def _f32_next_dt(dt, err, max_error, errororder):
    """The step-size update rk.go:404-420 in the reference's float32 host
    arithmetic: factor = float32((maxError/err)^(1/errororder)) clamped to
    [0.01, 2], dt' = dt * factor."""
    q = np.float32(max_error) / np.float32(err)
    f = np.float32(np.power(np.float64(q), 1.0 / errororder))
    f = np.float32(min(max(f, np.float32(0.01)), np.float32(2.0)))
    return np.float32(np.float32(dt) * f)


CTRL_CASES = [("film_128x32x1", None), ("cuboid_4x32x32", (4, 32, 32))]


def _ctrl_cfg(mm, which):
    if which[1] is None:
        return mm.workloads.film_sp4_like(seed=1)
    cfg = mm.workloads.cuboid_rk4(size=which[1])
    cfg.update(solver="rkdp", maxerror=1e-5, dt=0.0)
    return cfg


@pytest.mark.parametrize("mode", ["fused", "abi"])
@pytest.mark.parametrize("which", CTRL_CASES, ids=[c[0] for c in CTRL_CASES])
def test_device_controller_follows_the_reference_update(mm, which, mode):
    """Every logged trial of the engine's own Dormand-Prince run (no replay):
    accepted exactly when err <= 2 maxError (rk.go:421-428), and the next
    trial's dt is the float32 update of rk.go:404-420 from that trial's error
    estimate (growth clamp 2, shrink clamp 0.01, errororder 4), to 1 ulp."""
    cfg = _ctrl_cfg(mm, which)
    s = mm.workloads.make_sim(cfg, mode)
    s.steps(200)
    dts, acc = s.trial_log()
    errs = s.trial_errors()
    assert len(dts) == len(errs) >= 200
    assert dts[0] == np.float32(1e-4)
    n_clamped = 0
    for i in range(len(dts) - 1):
        assert bool(acc[i]) == (errs[i] <= np.float32(2e-5)), i
        want = _f32_next_dt(dts[i], errs[i], 1e-5, 4)
        got = np.float32(dts[i + 1])
        assert abs(int(got.view(np.int32)) - int(want.view(np.int32))) <= 1, (i, got, want, errs[i])
        n_clamped += bool(np.float32(1e-5) / errs[i] >= 16)
    assert n_clamped >= 5  # the doubling phase ran through the growth clamp


@pytest.mark.parametrize("which", CTRL_CASES, ids=[c[0] for c in CTRL_CASES])
def test_device_controller_own_run_vs_oracle_own_run(mm, orc, which):
    """The engine and the FP64 oracle each lead their own Dormand-Prince run
    from the same m0 over the first 200 trials (no schedule replayed): the
    same accept/reject decisions, and dt within twice the deviation of the
    FP32 twin leading its own run (or 1e-5 relative): dt is chosen from an
    error estimate of ~1e-5, in which float32 fields leave a relative noise of
    ~1e-4 that the controller carries into the next dt."""
    from oracle import Oracle

    cfg = _ctrl_cfg(mm, which)
    s = mm.workloads.make_sim(cfg, "fused")
    s.steps(200)
    dg, ag = s.trial_log()
    runs = []
    for O in (orc, Oracle("f32")):
        o = O.sim(cfg["m0"], **mm.workloads.oracle_kwargs(cfg))
        while len(o.trial_log()[0]) < 200:
            o.step()
        runs.append(o.trial_log())
    (d64, a64), (d32, a32) = runs
    n = 200
    e_gpu = float(np.abs(dg[:n] / d64[:n] - 1).max())
    e_32 = float(np.abs(d32[:n] / d64[:n] - 1).max())
    print("%s: first %d trials, rejects GPU %d, FP64 %d, FP32 twin %d; max rel. dt deviation from FP64: GPU %.3g, "
          "FP32 twin %.3g" % (which[0], n, n - ag[:n].sum(), n - a64[:n].sum(), n - a32[:n].sum(), e_gpu, e_32))
    assert np.array_equal(ag[:n], a64[:n])
    assert e_gpu <= max(1e-5, 2 * e_32)


def test_film_1ns_dp_inside_fp32_envelope_of_fp64_oracle(mm, orc):
    """The 128x32x1 film (Dormand-Prince, maxError 1e-5, B = (-24.6, 4.3, 0) mT)
    for the full 1 ns; the FP64 oracle and its FP32 twin (same algorithm,
    float32 arithmetic) replay the GPU's accepted-dt schedule.

    The film reverses during this nanosecond, and the reversal amplifies
    rounding: the FP32 twin itself drifts from FP64 by ~3e-6 at 0.05 ns and
    ~1e-2 at 1 ns, so no float32 engine meets 1e-5 after 1 ns on this
    trajectory (the 0.05 ns replay test holds the 1e-5 bar).  Bar here: the
    GPU's max |dm| and max |d<m>| against FP64 stay within twice the FP32
    twin's (or 1e-5, whichever is larger)."""
    from oracle import Oracle

    cfg = mm.workloads.film_sp4_like(seed=1)
    s = mm.workloads.make_sim(cfg, "fused")
    s.run(1e-9)
    st = s.state()
    assert st["time"] * st["unit_time"] >= 1e-9 * (1 - 1e-6)  # internal time units
    o64 = replay_in_oracle(mm, orc, cfg, s)
    o32 = replay_in_oracle(mm, Oracle("f32"), cfg, s)
    got = s.m.reshape(3, -1).astype(np.float64)
    r64 = o64.m.reshape(3, -1)
    r32 = o32.m.reshape(3, -1)
    e_gpu, e_32 = float(np.abs(got - r64).max()), float(np.abs(r32 - r64).max())
    a_gpu = float(np.abs(got.mean(axis=1) - r64.mean(axis=1)).max())
    a_32 = float(np.abs(r32.mean(axis=1) - r64.mean(axis=1)).max())
    print("film 1 ns: steps %d, rejects %d; vs FP64 oracle: GPU max|dm| %.3g, max|d<m>| %.3g; "
          "FP32 twin max|dm| %.3g, max|d<m>| %.3g" % (st["steps"], st["rejects"], e_gpu, a_gpu, e_32, a_32))
    assert st["steps"] == o64.state()["steps"] == o32.state()["steps"]
    assert e_gpu <= max(1e-5, 2 * e_32)
    assert a_gpu <= max(1e-5, 2 * a_32)


def _crossing(ts, ys, level):
    """First time the sampled curve ys(ts) reaches `level` (linear interpolation), or None."""
    for i in range(1, len(ys)):
        if (ys[i - 1] - level) * (ys[i] - level) <= 0 and ys[i] != ys[i - 1]:
            return ts[i - 1] + (level - ys[i - 1]) / (ys[i] - ys[i - 1]) * (ts[i] - ts[i - 1])
    return None


def test_vortex_ramp_precision_vs_fp64_oracle(mm, orc):
    """Precision study on a reduced vortex film (64x64x1 of 2 nm cubes, alpha 0.5,
    Dormand-Prince 1e-5, in-plane field ramp 0 -> 50 mT over 0.5 ns; the same
    physics as the 1024x1024 study at an oracle-sized grid).  The GPU runs in 40
    chunks (each Run() renormalises first, simrun.go:24); the FP64 oracle replays
    the same trials chunk by chunk, and so does its FP32 twin.  Reported and
    checked at every chunk end: max |dm| (bar: within twice the FP32 twin's,
    or 1e-5; the switch amplifies rounding, the twin drifts by ~3e-4), the
    relative error of the FP64-reduced energy (1e-5), <m>, and the switching
    time (<m> along the field reaching 0.5; 1e-3 of the ramp)."""
    from oracle import Oracle

    T, nch = 0.5e-9, 40
    cfg = mm.workloads.film_vortex_ramp(seed=3, size=(1, 64, 64), ramp_time=T)
    s = mm.workloads.make_sim(cfg, "fused")
    ts, marks, avg_g, e_g, m_g = [], [], [], [], []
    for k in range(nch):
        s.run(T / nch)
        st = s.state()
        ts.append(st["time"] * st["unit_time"])
        marks.append(len(s.trial_log()[0]))
        avg_g.append(s.average_m())
        e_g.append(s.energy())
        m_g.append(s.m.reshape(3, -1).astype(np.float64))
    dts, acc = s.trial_log()
    o = orc.sim(cfg["m0"], **mm.workloads.oracle_kwargs(cfg))
    o32 = Oracle("f32").sim(cfg["m0"], **mm.workloads.oracle_kwargs(cfg))
    i, err, rel_e, d_avg, err32 = 0, 0.0, 0.0, 0.0, 0.0
    avg_o = []
    for k in range(nch):
        normalise_oracle(o)  # each Run() normalises first
        normalise_oracle(o32)
        replay_log(o32, dts, acc, i, marks[k])
        i = replay_log(o, dts, acc, i, marks[k])
        mo = o.m.reshape(3, -1)
        err32 = max(err32, float(np.abs(o32.m.reshape(3, -1) - mo).max()))
        avg_o.append(mo.mean(axis=1))
        err = max(err, float(np.abs(m_g[k] - mo).max()))
        eo = o.energy()
        rel_e = max(rel_e, abs(e_g[k] - eo) / abs(eo))
        d_avg = max(d_avg, float(np.abs(np.asarray(avg_g[k]) - avg_o[k]).max()))
    tsw_g = _crossing(ts, [a[2] for a in avg_g], 0.5)
    tsw_o = _crossing(ts, [a[2] for a in avg_o], 0.5)
    print("vortex 64x64: trials %d; vs FP64 oracle: max|dm| %.3g (FP32 twin %.3g), max rel. energy error %.3g, "
          "max|d<m>| %.3g; switching time GPU %s ns, oracle %s ns" % (
              len(dts), err, err32, rel_e, d_avg, tsw_g and "%.5f" % (tsw_g * 1e9), tsw_o and "%.5f" % (tsw_o * 1e9)))
    assert err <= max(1e-5, 2 * err32)
    assert rel_e <= 1e-5
    assert tsw_o is not None and tsw_g is not None
    assert abs(tsw_g - tsw_o) <= 1e-3 * T


def test_vortex_1024_precision_study(mm, orc):
    """The 1024x1024x1 vortex film (2 nm cubes, alpha 0.5, Dormand-Prince 1e-5,
    in-plane field ramp 0 -> 50 mT over 2 ns), at full size:
    - the first 5 steps replayed in the FP64 oracle: max |dm| <= 1e-5;
    - the 2 ns run in 20 chunks: <m> and energy from the FP64 device reductions,
      the switching time (<m> along the field reaching 0.5), energy decreasing
      at fixed field is not required (the field ramps), the state stays unit;
    - the energy of the final state (zero field, t = 0) from the GPU pipeline
      (FP32 fields, FP64 sum) against the FP64 oracle: relative error <= 1e-5."""
    cfg = mm.workloads.film_vortex_ramp(seed=3)
    T, nch = cfg["duration"], 20
    s5 = mm.workloads.make_sim(cfg, "fused")
    s5.steps(5)
    m5 = s5.m.reshape(3, -1).astype(np.float64)
    dts5, acc5 = s5.trial_log()
    s5.free()
    s = mm.workloads.make_sim(cfg, "fused")
    ts, avg = [], []
    for k in range(nch):
        s.run(T / nch)
        st = s.state()
        ts.append(st["time"] * st["unit_time"])
        avg.append(s.average_m())
    mf = s.m
    st = s.state()
    assert np.abs(np.linalg.norm(mf, axis=0) - 1).max() < 1e-5
    tsw = _crossing(ts, [a[2] for a in avg], 0.5)
    s.free()
    s2 = mm.workloads.make_sim(cfg, "fused")
    s2.setm_array(mf)
    e_gpu = s2.energy()
    s2.free()
    o = orc.sim(cfg["m0"], **mm.workloads.oracle_kwargs(cfg))
    o.m = mf.reshape(-1).astype(np.float64)
    e_o = o.energy()
    rel_e = abs(e_gpu - e_o) / abs(e_o)
    o.m = cfg["m0"].reshape(-1).astype(np.float64)
    normalise_oracle(o)
    replay_log(o, dts5, acc5)
    err5 = float(np.abs(m5 - o.m.reshape(3, -1)).max())
    print("vortex 1024x1024, 2 ns: steps %d, trials %d, rejects %d; switching time %s ns; <m> at 2 ns %s; "
          "first 5 steps max|dm| vs FP64 oracle %.3g; final-state energy GPU %.9g vs FP64 %.9g (rel. %.3g)"
          % (st["steps"], st["trials"], st["rejects"], tsw and "%.4f" % (tsw * 1e9), np.round(avg[-1], 4), err5,
             e_gpu, e_o, rel_e))
    assert err5 <= 1e-5
    assert rel_e <= 1e-5


def test_save_and_loadm_round_trip(mm, tmp_path):
    """Save("m", ...) in every format and LoadM of the file give back m; the
    table holds time, <m> and B in user order."""
    from importlib import import_module

    ovf = import_module("1_amd.ovf")
    cfg = mm.workloads.film_sp4_like(seed=1)
    s = mm.workloads.make_sim(cfg)
    s.run(0.002e-9)
    for fmt in ("binary", "omf_2", "text"):
        p = s.save("m", fmt, outdir=str(tmp_path))
        m, meta = ovf.read_ovf(p)
        assert np.array_equal(m, s.m)
        assert np.allclose(meta["step"], cfg["cellsize"]) and float(meta["desc"]["msat"]) == cfg["msat"]
        s2 = mm.workloads.make_sim(cfg)
        s2.loadm(p)
        assert np.array_equal(s2.m, s.m)
        s2.free()
    t = s.save("table", outdir=str(tmp_path))
    s.close_table()
    row = [float(x) for x in open(t).read().splitlines()[5].split()]
    avg = s.average_m()
    assert row[1:4] == pytest.approx([avg[2], avg[1], avg[0]], abs=1e-6)
    assert row[4:7] == pytest.approx([-24.6e-3, 4.3e-3, 0.0], abs=1e-9)


# ------------------------------------------------------------------ live field, Relax, Hysteresis
def _relax_film(mm, alpha=1.0):
    size, cell = (1, 32, 64), (3e-9, 3.90625e-9, 3.90625e-9)
    m0 = np.zeros((3,) + size, np.float32)
    m0[2], m0[1] = 1.0, 0.3
    m0 /= np.linalg.norm(m0, axis=0)
    s = engine(mm, size, cell, m0=m0, alpha=alpha, maxerror=1e-4)
    return s, size, cell, m0


def test_applystatic_on_running_sim_replayed_in_oracle(mm, orc):
    """ApplyStatic between two Run()s replaces the field without re-initialising
    (simfield.go:87-90); the oracle replays the same trials with the field
    swapped at the same point."""
    cfg = mm.workloads.film_sp4_like(seed=1)
    s = mm.workloads.make_sim(cfg)
    s.run(0.02e-9)
    n1 = len(s.trial_log()[0])
    s.applystatic("field", 10e-3, -5e-3, 2e-3)   # user (x, y, z) = internal (Z, Y, X)
    s.run(0.02e-9)
    dts, acc = s.trial_log()
    assert len(dts) > n1 + 3
    o = orc.sim(cfg["m0"], **mm.workloads.oracle_kwargs(cfg))
    normalise_oracle(o)
    assert replay_log(o, dts, acc, 0, n1) == n1
    o.set_field(np.zeros(1), [[2e-3, -5e-3, 10e-3]])
    normalise_oracle(o)
    replay_log(o, dts, acc, n1)
    got = s.m
    assert np.abs(got - o.m.reshape(got.shape)).max() < 1e-5
    assert s.state()["steps"] == o.state()["steps"]


def _replay(O, size, cell, m0, dts, acc, events, marks=()):
    """Replay a trial log with energy control on in oracle precision O;
    events = [(trial index, internal B)] applies a static field before that
    trial; returns the sim and {mark: (m, energy)} at the given trial marks."""
    o = oracle(O, size, cell, m0, alpha=1.0, maxerror=1e-4)
    O.lib.or_sim_set_want_energy(o.ptr, 1)
    ev = dict(events)
    out, i = {}, 0
    for mk in sorted(set(ev) | set(marks)):
        i = replay_log(o, dts, acc, i, mk)
        assert i == mk
        if mk in marks:
            out[mk] = (o.m.copy(), o.state()["energy"])
        if mk in ev:
            o.set_field(np.zeros(1), [ev[mk]])
    replay_log(o, dts, acc, i)
    return o, out


def _twin_led_envelope(orc, size, cell, m0, fields):
    """The FP32 twin runs the same relaxations (fields: internal B per point)
    with its own accept/reject decisions; its trials are replayed in FP64.
    Returns the largest max|dm| and |dE| between the two over all steps.

    An adaptive FP32 run accepts trials whose error estimate -- dominated,
    when dt cycles at the stability edge, by the stiffest modes -- fell below
    the bar in its own rounding; replayed in other arithmetic, those modes
    show the rounding it selected, up to a few 1e-4 here.  That is the
    envelope any FP32 engine leading its own run lies in."""
    from oracle import Oracle

    O32 = Oracle("f32")
    t = oracle(O32, size, cell, m0, alpha=1.0, maxerror=1e-4)
    snaps, events = [], []

    def snap():
        snaps.append((len(t.trial_log()[0]), t.m.copy(), t.state()["energy"]))

    for b in fields:
        if b is not None:
            events.append((len(t.trial_log()[0]), b))
            t.set_field(np.zeros(1), [b])
        t.relax(on_step=snap)
    dts, acc = t.trial_log()
    _, ref = _replay(orc, size, cell, m0, dts, acc, events, [mk for mk, _, _ in snaps])
    dm = max(float(np.abs(m - ref[mk][0]).max()) for mk, m, _ in snaps)
    de = max(abs(e - ref[mk][1]) for mk, _, e in snaps)
    return dm, de


def test_relax_replayed_in_oracle(mm, orc):
    """Relax() (relax.go:14-25): the GPU's energy-controlled trials, up to the
    step at which it stopped by itself, replayed in the FP64 oracle.  Bar: the
    GPU's max|dm| and energy error within twice the FP32 envelope of this
    relaxation (the twin leading its own run, replayed in FP64); the GPU
    stops at an energy within the FP32/FP64 spread of the oracle's own
    Relax()."""
    s, size, cell, m0 = _relax_film(mm)
    n = s.relax(max_steps=20000)
    assert 10 < n < 20000
    st = s.state()
    dts, acc = s.trial_log()
    o64, _ = _replay(orc, size, cell, m0, dts, acc, [])
    e_gpu = float(np.abs(s.m.reshape(-1) - o64.m).max())
    de_gpu = abs(st["energy"] - o64.state()["energy"])
    env_m, env_e = _twin_led_envelope(orc, size, cell, m0, [None])
    print("relax: %d steps, %d trials; vs FP64 replay: GPU max|dm| %.3g, |dE| %.3g; FP32 envelope max|dm| %.3g, "
          "|dE| %.3g" % (n, len(dts), e_gpu, de_gpu, env_m, env_e))
    assert o64.state()["steps"] == n
    assert e_gpu <= max(1e-5, 2 * env_m)
    assert de_gpu <= max(1e-6 * abs(st["energy"]), 2 * env_e)
    # the oracle's own Relax from the same start: same minimum to the FP32/FP64 spread
    o2 = oracle(orc, size, cell, m0, alpha=1.0, maxerror=1e-4)
    o2.relax()
    assert st["energy"] == pytest.approx(o2.state()["energy"], rel=1e-4)
    mo = o2.m.reshape(3, -1).mean(1)
    assert np.abs(s.average_m() - mo).max() < 3e-2


def test_hysteresis_replayed_in_oracle(mm, orc, tmp_path):
    """Hysteresis() (simhysteresis.go:16-27): ApplyStatic + Relax per field
    point on one live simulation, replayed in the FP64 oracle with the field
    swapped at the same trials; at every point's end max|dm| and the energy
    error stay within twice the FP32 envelope of the same scan.  Hysteresis()
    itself gives bit-identical m and writes one m file and one table row per
    point."""
    s, size, cell, m0 = _relax_film(mm)
    pts = [(-20e-3, 5e-3, 0.0), (0.0, 2.5e-3, 0.0), (20e-3, 0.0, 0.0)]   # user (x, y, z)
    fields = [(bz, by, bx) for bx, by, bz in pts]                       # internal
    events, ends = [], []
    for (bx, by, bz), b in zip(pts, fields):
        events.append((len(s.trial_log()[0]), b))
        s.applystatic("field", bx, by, bz)
        s.relax(max_steps=20000)
        ends.append((len(s.trial_log()[0]), s.m.reshape(-1).copy(), s.state()["energy"]))
    dts, acc = s.trial_log()
    _, ref = _replay(orc, size, cell, m0, dts, acc, events, [mk for mk, _, _ in ends])
    e_gpu = max(float(np.abs(m - ref[mk][0]).max()) for mk, m, _ in ends)
    de_gpu = max(abs(e - ref[mk][1]) for mk, _, e in ends)
    env_m, env_e = _twin_led_envelope(orc, size, cell, m0, fields)
    print("hysteresis: %d trials; vs FP64 replay at the 3 points: GPU max|dm| %.3g, |dE| %.3g; FP32 envelope "
          "max|dm| %.3g, |dE| %.3g" % (len(dts), e_gpu, de_gpu, env_m, env_e))
    assert e_gpu <= max(1e-5, 2 * env_m)
    assert de_gpu <= max(1e-6 * 3e3, 2 * env_e)
    got = s.m

    h, _, _, _ = _relax_film(mm)
    h.hysteresis(*pts[0], *pts[2], 2, outdir=str(tmp_path))
    h.close_table()
    assert np.array_equal(h.m, got)
    assert sorted(p.name for p in tmp_path.glob("m*.omf")) == ["m%08d.omf" % k for k in range(3)]
    rows = [ln for ln in open(tmp_path / "datatable.odt").read().splitlines() if not ln.startswith("#")]
    assert len(rows) == 3
    assert float(rows[-1].split()[4]) == pytest.approx(20e-3, abs=1e-9)


# ------------------------------------------------------------------ per-cell damping (SetAlpha)
def _alpha_map(size, seed):
    return np.random.default_rng(seed).uniform(0.3, 3.0, size).astype(np.float32)


@pytest.mark.parametrize("mode", ["fused", "abi"])
def test_alpha_map_torque_and_rk4_vs_oracle(mm, orc, mode):
    """SetAlpha's damping multiplier (simgeom.go:246-264, torque.go:49-51):
    the torque and 20 RK4 steps (the fused path applies it in the Z inverse
    epilogue) against the oracle with the same map."""
    size, cell = (2, 16, 32), (4e-9, 4e-9, 4e-9)
    m0 = smooth_m(size, 5)
    amap = _alpha_map(size, 6)
    s = engine(mm, size, cell, mode=mode, solver="rk4", dt=2e-14, m0=m0, alpha=0.1, msat=1e6, aexch=1.5e-11)
    s._alpha_map = amap.copy()
    o = oracle(orc, size, cell, m0, solver="rk4", dt=2e-14, alpha=0.1, msat=1e6, aexch=1.5e-11)
    o.set_alpha_map(amap)
    k = s.torque(m0)
    kref = o.torque(m0).reshape(k.shape)
    assert np.abs(k - kref).max() < 5e-6 * np.abs(kref).max()
    s.steps(20)
    for _ in range(20):
        o.step()
    d = np.abs(s.m - o.m.reshape(s.m.shape)).max()
    print(mode, "max |dm| after 20 RK4 steps with an alpha map:", d)
    assert d < 1e-5
    # the map changes the dynamics
    p = engine(mm, size, cell, mode=mode, solver="rk4", dt=2e-14, m0=m0, alpha=0.1, msat=1e6, aexch=1.5e-11)
    p.steps(20)
    assert np.abs(p.m - s.m).max() > 1e-4


def test_alpha_map_of_ones_is_no_map_and_batched_equals_alone(mm):
    size, cell = (1, 32, 64), (3e-9, 3.90625e-9, 3.90625e-9)
    m0 = random_m(size, 7)
    a = engine(mm, size, cell, m0=m0, alpha=0.2)
    b = engine(mm, size, cell, m0=m0, alpha=0.2)
    b._alpha_map = np.ones(size, np.float32)
    a.run(5e-12), b.run(5e-12)
    assert np.array_equal(a.m, b.m)
    # a member with a map computes the same in a batch as alone
    alone = engine(mm, size, cell, m0=m0, alpha=0.2)
    alone._alpha_map = _alpha_map(size, 8)
    alone.run(8e-12)
    sims = [engine(mm, size, cell, m0=m0, alpha=0.2), engine(mm, (2, 16, 24), (3e-9, 3e-9, 3e-9),
                                                             m0=random_m((2, 16, 24), 9), alpha=0.1)]
    sims[0]._alpha_map = _alpha_map(size, 8)
    bt = mm.Batch(sims)
    bt.run([8e-12, 5e-12])
    assert np.array_equal(sims[0].m, alone.m)
    bt.free()


# ------------------------------------------------------------------ applied-field functions of time
@pytest.mark.parametrize("kind,args", [("rf", (10e9,)), ("pulse", (20e-12, 10e-12)), ("rotating", (8e9, 0.0, 1.0, 2.0))])
def test_field_function_run_vs_oracle(mm, orc, kind, args):
    """ApplyRf / ApplyPulse / ApplyRotating on a film: the Dormand-Prince run
    replayed in the FP64 oracle with the same field function, within 1e-5 or
    twice the drift of the FP32 twin replaying the same trials (this state
    drifts by 2-4e-5 in FP32 over 40 ps even under a static field); the field
    Sim reports equals the oracle's."""
    size, cell = (1, 32, 64), (3e-9, 3.90625e-9, 3.90625e-9)
    m0 = smooth_m(size, 12)
    s = engine(mm, size, cell, m0=m0, alpha=0.05)
    getattr(s, "apply" + kind)("field", 0.0, 40e-3, -25e-3, *args)   # user (x, y, z) = (0, 40, -25) mT
    s.run(40e-12)
    code = {"pulse": 1, "rf": 2, "rotating": 5}[kind]
    o = oracle(orc, size, cell, m0, alpha=0.05)
    o.set_field_function(code, s._fn[1], list(s._fn[2]))
    for t in (3e-12, 17e-12, 31e-12):
        assert np.allclose(s.field_at(t), o.field_at(t), rtol=1e-6, atol=1e-12)
    normalise_oracle(o)
    dts, acc = s.trial_log()
    replay_log(o, dts, acc)
    d = np.abs(s.m - o.m.reshape(s.m.shape)).max()
    from oracle import Oracle

    tw = oracle(Oracle("f32"), size, cell, m0, alpha=0.05)
    tw.set_field_function(code, s._fn[1], list(s._fn[2]))
    normalise_oracle(tw)
    replay_log(tw, dts, acc)
    dt32 = np.abs(tw.m - o.m).max()
    print(kind, "trials", len(dts), "max |dm| vs FP64 replay: GPU %.3g, FP32 twin %.3g" % (d, dt32))
    assert d <= max(1e-5, 2 * dt32)
