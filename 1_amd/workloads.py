"""Seeded synthetic workloads: the film, cuboid and sweep configurations.

Random numbers come from SplitMix64 (Steele, Lea & Flood, 2014) turned into
uniforms with 53 bits and into normals with the Box-Muller transform, so the
same seed gives the same initial state everywhere (numpy here, any C port).
All vectors are returned in the internal planar layout ``[3][X][Y][Z]`` with
X = film thickness (internal axes are the user's (z, y, x), simgeom.go:19-39).
"""
import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(seed, n):
    """n outputs of SplitMix64 started at `seed`, as uint64."""
    with np.errstate(over="ignore"):
        x = np.uint64(seed & MASK64) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed, n):
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n):
    """n standard normals (Box-Muller on consecutive uniform pairs)."""
    u = uniform01(seed, 2 * n)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def normalize(m):
    m = np.asarray(m, np.float64)
    return m / np.linalg.norm(m, axis=0, keepdims=True)


def film_sp4_like(seed=1):
    """Thin film 128x32x1 of 3.90625x3.90625x3 nm, Ms 8e5, A 1.3e-11, alpha 0.02,
    m0 = normalize((1,0.25,0.1)/|.| + N(0,0.01)), B = (-24.6, 4.3, 0) mT, DP 1e-5, 1 ns."""
    X, Y, Z = 1, 32, 128
    N = X * Y * Z
    u_user = np.array([1.0, 0.25, 0.1])
    u_user /= np.linalg.norm(u_user)
    u_int = u_user[::-1]  # internal (X,Y,Z) = user (z,y,x)
    noise = normal(seed, 3 * N).reshape(3, X, Y, Z) * 0.01
    m = normalize(u_int[:, None, None, None] + noise)
    return dict(
        name="film_128x32x1_dp",
        size=(X, Y, Z),
        cellsize=(3e-9, 3.90625e-9, 3.90625e-9),
        msat=8.0e5, aexch=1.3e-11, alpha=0.02,
        solver="rkdp", maxerror=1e-5, dt=0.0,
        field=("static", (0.0, 4.3e-3, -24.6e-3)),
        duration=1e-9,
        m0=m.astype(np.float32),
    )


def cuboid_rk4(seed=2, size=(16, 256, 256), steps=1000):
    """Cuboid 256x256x16 cells of 4 nm, Ms 1e6, A 1.5e-11, Ku1 5e5 along user z
    (internal X), alpha 0.1, random m0 uniform on the sphere, zero field,
    fixed-dt RK4 with dt = 1e-14 s."""
    X, Y, Z = size
    N = X * Y * Z
    g = normal(seed, 3 * N).reshape(3, X, Y, Z)
    m = normalize(g)
    return dict(
        name="cuboid_%dx%dx%d_rk4" % (Z, Y, X),
        size=(X, Y, Z),
        cellsize=(4e-9, 4e-9, 4e-9),
        msat=1.0e6, aexch=1.5e-11, alpha=0.1,
        k1=5e5, anis_axis=(1.0, 0.0, 0.0),
        solver="rk4", maxerror=0.0, dt=1e-14,
        field=None,
        steps=steps,
        m0=m.astype(np.float32),
    )


def film_vortex_ramp(seed=3, size=(1, 1024, 1024), ramp_time=2e-9, bmax=50e-3, core_radius=10e-9):
    """Film 1024x1024x1 of 2 nm cubes, Ms 8e5, A 1.3e-11, alpha 0.5: a vortex
    (counter-clockwise, core up) centred at a seeded position within the middle
    half of the film plus N(0, 0.05) noise; in-plane field along user x ramped
    0 -> 50 mT over 2 ns; DP."""
    X, Y, Z = size
    N = X * Y * Z
    cell = 2e-9
    c = uniform01(seed, 2)
    cy = (0.25 + 0.5 * c[0]) * Y * cell
    cz = (0.25 + 0.5 * c[1]) * Z * cell
    yy = (np.arange(Y) + 0.5) * cell - cy
    zz = (np.arange(Z) + 0.5) * cell - cz
    YY, ZZ = np.meshgrid(yy, zz, indexing="ij")
    r = np.sqrt(YY ** 2 + ZZ ** 2) + 1e-30
    mX = np.exp(-(r / core_radius) ** 2)
    s = np.sqrt(1 - mX ** 2)
    # user in-plane (mx, my) = (-dy, dx)/r  ->  internal Z = user x, Y = user y
    mZ = -YY / r * s
    mY = ZZ / r * s
    m = np.stack([mX, mY, mZ])[:, None, :, :] * np.ones((1, X, 1, 1))
    m = normalize(m + normal(seed + 1000, 3 * N).reshape(3, X, Y, Z) * 0.05)
    return dict(
        name="film_%dx%dx%d_vortex_ramp" % (Z, Y, X),
        size=(X, Y, Z),
        cellsize=(cell, cell, cell),
        msat=8.0e5, aexch=1.3e-11, alpha=0.5,
        solver="rkdp", maxerror=1e-5, dt=0.0,
        field=("pointwise", [(0.0, (0.0, 0.0, 0.0)), (ramp_time, (0.0, 0.0, bmax))]),
        duration=ramp_time,
        m0=m.astype(np.float32),
    )


SWEEP_SIZES = tuple(range(32, 257, 16))  # user Nx, Ny
SWEEP_THICK = (1, 2, 4)                  # user Nz


def sweep_member(seed, j, duration=0.5e-9, sizes=SWEEP_SIZES, thick=SWEEP_THICK, cell=3e-9):
    """Simulation j of the batched sweep: user Nx, Ny uniform in `sizes`, Nz in
    `thick`, cubes of `cell`; Ms ~ U[4e5, 1.2e6] A/m, A ~ U[0.8e-11, 2e-11] J/m,
    alpha ~ U[0.01, 0.5], B_ext of isotropic direction with |B| ~ U[0, 100] mT;
    m0 uniform on the sphere per cell; Dormand-Prince, maxError 1e-5.  Every
    draw comes from SplitMix64 streams keyed by (seed, j)."""
    key = (seed * 1000003 + j) & MASK64
    u = uniform01(key, 10)
    nx = sizes[int(u[0] * len(sizes)) % len(sizes)]
    ny = sizes[int(u[1] * len(sizes)) % len(sizes)]
    nz = thick[int(u[2] * len(thick)) % len(thick)]
    msat = 4e5 + u[3] * 8e5
    aexch = 0.8e-11 + u[4] * 1.2e-11
    alpha = 0.01 + u[5] * 0.49
    cz = 2 * u[6] - 1  # isotropic direction: cos(theta) uniform, phi uniform
    ph = 2 * np.pi * u[7]
    sz = np.sqrt(max(0.0, 1 - cz * cz))
    b = u[8] * 100e-3
    bx, by, bz = b * sz * np.cos(ph), b * sz * np.sin(ph), b * cz  # user axes
    X, Y, Z = nz, ny, nx
    N = X * Y * Z
    m = normalize(normal(key ^ 0x5DEECE66D, 3 * N).reshape(3, X, Y, Z))
    return dict(
        name="sweep%d_%d" % (seed, j),
        size=(X, Y, Z),
        cellsize=(cell, cell, cell),
        msat=msat, aexch=aexch, alpha=alpha,
        solver="rkdp", maxerror=1e-5, dt=0.0,
        field=("static", (bz, by, bx)),  # internal (X, Y, Z) = user (z, y, x)
        duration=duration,
        m0=m.astype(np.float32),
    )


def padded_length(n):
    """The engine's padded length of an axis of n cells: the smallest 2^a or
    3 * 2^a >= 2n (1 for n = 1), the same linear convolution as the reference's
    2n (kernel.go:79-90)."""
    if n == 1:
        return 1
    best, a = None, 1
    while a < 4 * n:
        for c in (a, 3 * a):
            if c >= 2 * n and (best is None or c < best):
                best = c
        a *= 2
    return best


def sweep_length_cover(cfgs):
    """Indices of members of a sweep that together take every padded Z length
    and every (padded X, padded Y) pair the sweep holds -- the length classes
    of the batched passes (greedy, smallest members first)."""
    keys = []
    for c in cfgs:
        P = [padded_length(n) for n in c["size"]]
        keys.append({("M", P[2] // 2), ("XY", P[0], P[1])})
    need = set().union(*keys)
    pick = []
    while need:
        j = max(range(len(cfgs)), key=lambda j: (len(keys[j] & need), -int(np.prod(cfgs[j]["size"]))))
        pick.append(j)
        need -= keys[j]
    return sorted(pick)


def sweep(n=512, seed=4, duration=0.5e-9, **kw):
    """The batched sweep: n independent simulations (see sweep_member)."""
    return [sweep_member(seed, j, duration, **kw) for j in range(n)]


def make_sim(cfg, mode="fused"):
    """Build a ``Sim`` from a workload dict using the reference's commands."""
    from .sim import Sim

    s = Sim(mode)
    X, Y, Z = cfg["size"]
    s.gridsize(Z, Y, X)
    cx, cy, cz = cfg["cellsize"]
    s.cellsize(cz, cy, cx)
    if cfg.get("periodic"):
        pX, pY, pZ = cfg["periodic"]
        s.periodic(pZ, pY, pX)
    s.msat(cfg["msat"])
    s.aexch(cfg["aexch"])
    s.alpha(cfg["alpha"])
    if cfg.get("anis_axis") is not None:
        ax = cfg["anis_axis"]
        s.anisuniaxial(ax[2], ax[1], ax[0])
        s.k1(cfg["k1"])
    s.solvertype(cfg["solver"])
    if cfg.get("maxerror"):
        s.maxerror(cfg["maxerror"])
    if cfg.get("dt"):
        s.dt(cfg["dt"])
    if cfg.get("max_dt"):
        s.maxdt(cfg["max_dt"])
    if cfg.get("min_dt"):
        s.mindt(cfg["min_dt"])
    if cfg.get("temperature"):
        s.temperature(cfg["temperature"])
        s.thermalseed(cfg.get("seed", 0))
    f = cfg.get("field")
    if f is not None:
        if f[0] == "static":
            bX, bY, bZ = f[1]
            s.applystatic("field", bZ, bY, bX)
        else:
            for t, (bX, bY, bZ) in f[1]:
                s.applypointwise("field", t, bZ, bY, bX)
    if cfg.get("norm_map") is not None:
        s.setmsat_array(cfg["norm_map"])
    s.setm_array(cfg["m0"])
    return s


def oracle_kwargs(cfg):
    """Keyword arguments for the oracle's Sim (internal order)."""
    kw = dict(size=cfg["size"], cellsize=cfg["cellsize"], msat=cfg["msat"], aexch=cfg["aexch"],
              alpha=cfg["alpha"], solver=cfg["solver"], maxerror=cfg.get("maxerror") or 1e-4,
              dt=cfg.get("dt") or 0.0)
    if cfg.get("anis_axis") is not None:
        kw.update(k1=cfg["k1"], anis_axis=cfg["anis_axis"])
    if cfg.get("temperature"):
        kw.update(temperature=cfg["temperature"], seed=cfg.get("seed", 0))
    if cfg.get("periodic"):
        kw["periodic"] = tuple(cfg["periodic"])
    if cfg.get("norm_map") is not None:
        kw["norm_map"] = cfg["norm_map"]
    for k in ("max_dt", "min_dt"):
        if cfg.get(k):
            kw[k] = cfg[k]
    f = cfg.get("field")
    if f is not None:
        if f[0] == "static":
            kw["field"] = (np.zeros(1), np.asarray(f[1], float))
        else:
            kw["field"] = (np.asarray([p[0] for p in f[1]], float), np.asarray([p[1] for p in f[1]], float))
    return kw
