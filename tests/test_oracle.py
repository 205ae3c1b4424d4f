"""Pins of the CPU oracle (oracle/oracle.c) without the reference binary.

Each test checks the oracle against something it does not restate: the
reference's own known answers (redux_test.go, fft_test.go), closed forms
(exact demag factors, dipole far field, spin-wave dispersion, analytic
macrospin precession and damping), brute force (direct DFT, direct
convolution) or an independent library (numpy.fft).  A plausible slip in the
oracle (a dropped term, a sign, a transposed index) fails at least one.
"""
import numpy as np
import pytest

from oracle import REDUCE_ADD, REDUCE_MAX, REDUCE_MAXABS, REDUCE_MIN, REDUCE_SUMABS


# ------------------------------------------------------------------ units (material.go:36-110)
def test_units_closed_form(orc):
    u = orc.units(8.0e5, 1.3e-11)
    mu0 = 4e-7 * np.pi
    assert u["length"] == pytest.approx(np.sqrt(2 * 1.3e-11 / (mu0 * 8e5 ** 2)), rel=1e-15)
    assert u["length"] == pytest.approx(5.6858e-9, rel=1e-4)   # exchange length of permalloy-like film
    assert u["time"] == pytest.approx(1 / (2.211e5 * 8e5), rel=1e-15)
    assert u["time"] == pytest.approx(5.6535e-12, rel=1e-4)
    assert u["field"] == pytest.approx(mu0 * 8e5, rel=1e-15)
    assert u["edens"] == pytest.approx(mu0 * 8e5 ** 2 / 2, rel=1e-12)  # A/L^2 = mu0 Ms^2 / 2


# ------------------------------------------------------------------ FaceKernel6 (demag.go:26-183)
def test_facekernel_cube_self_term(orc):
    K = orc.facekernel6((2, 2, 2), (1.0, 1.0, 1.0))
    diag = [K[i, 0, 0, 0] for i in range(3)]
    # exact demag factor of a cube is 1/3; 8x8 point charges per face err by < 1 %
    for d in diag:
        assert d == pytest.approx(-1 / 3, rel=1e-2)
    assert diag[0] == diag[1] == diag[2]
    # off-diagonal self terms vanish by symmetry
    for i in (3, 4, 5):
        assert abs(K[i, 0, 0, 0]) < 1e-15


def test_facekernel_quadrature_converges_to_exact_cube(orc):
    # the face-charge sum converges to the exact self-demag of a cube as n grows
    errs = []
    for n in (2, 4, 8, 16):
        K = orc.facekernel6((2, 2, 2), (1.0, 1.0, 1.0), accuracy=n)
        errs.append(abs(K[0, 0, 0, 0] + 1 / 3))
    assert errs[0] > errs[1] > errs[2] > errs[3]


def test_facekernel_thin_film_limit(orc):
    # film of 128x128 cubic cells, one cell thick: the field at the centre of a
    # uniformly magnetised (255x255) platelet tends to -M along the normal (internal X)
    K = orc.facekernel6((1, 256, 256), (1.0, 1.0, 1.0))
    assert K[0].sum() == pytest.approx(-1.0, abs=5e-3)
    assert abs(K[1].sum()) < 5e-3 and abs(K[2].sum()) < 5e-3
    # in-plane and normal factors add up to the (quadrature) trace of a closed body
    assert K[0].sum() + K[1].sum() + K[2].sum() == pytest.approx(-1.0, abs=5e-3)


def _face_solid(h, v1, v2, w1, w2):
    """Integral of h / r^3 over the rectangle [v1,v2] x [w1,w2] at height h (the
    solid angle it subtends): the corner sum of atan(v w / (h r))."""
    F = lambda v, w: np.arctan(v * w / (h * np.sqrt(h * h + v * v + w * w)))
    return F(v2, w2) - F(v1, w2) - F(v2, w1) + F(v1, w1)


def _box_demag(lo, hi, cell):
    """-H_u / M_u at the centre of cell 0 of the box of cells lo..hi (per axis)
    uniformly magnetised along u, from the charges of its two faces normal to u."""
    out = []
    for u in range(3):
        v, w = (u + 1) % 3, (u + 2) % 3
        e = [((lo[a] - 0.5) * cell[a], (hi[a] + 0.5) * cell[a]) for a in range(3)]
        out.append(-(_face_solid(e[u][1], *e[v], *e[w]) + _face_solid(-e[u][0], *e[v], *e[w])) / (4 * np.pi))
    return np.array(out)


PERIODIC_CASES = [
    # size, periodic images, cell, image range per axis (cells): demag.go:35-53 scaled by images + 1
    ((4, 6, 8), (2, 1, 3), (1.0, 1.3, 0.8), ((-3, 3), (-4, 4), (-12, 12))),
    ((5, 3, 7), (1, 2, 1), (1.0, 1.0, 1.0), ((-4, 2), (-3, 0), (-6, 4))),   # odd periods: asymmetric
    ((1, 8, 8), (0, 4, 4), (0.5, 1.0, 1.0), ((0, 0), (-15, 15), (-15, 15))),  # film, periodic in plane
    ((2, 6, 4), (0, 2, 0), (1.0, 1.0, 1.0), ((-1, 1), (-6, 6), (-3, 3))),    # free X and Z
]


@pytest.mark.parametrize("size,per,cell,box", PERIODIC_CASES)
def test_facekernel_periodic_images_sum_to_the_box_field(orc, size, per, cell, box):
    """A periodic axis keeps its size (padSize, kernel.go:79-90) and every
    displacement of the image range adds into its wrapped cell: the kernel's
    sum is the field at the centre cell of the box of all image cells, uniformly
    magnetised -- the closed form of its face charges.  The face-charge
    quadrature converges to it (x4 per doubled accuracy, <= 2.5e-5 at 8)."""
    P = [n if (k or n == 1) else 2 * n for n, k in zip(size, per)]
    want = _box_demag([b[0] for b in box], [b[1] for b in box], cell)
    err = []
    for acc in (8, 16):
        got = orc.facekernel6(P, cell, acc, per).reshape(6, -1).sum(axis=1)
        err.append(np.abs(got[:3] - want).max())
        if box[0][0] == -box[0][1] and box[1][0] == -box[1][1] and box[2][0] == -box[2][1]:
            assert np.abs(got[3:]).max() < 1e-12   # symmetric box: no shear field at its centre
    assert err[0] < 2.5e-5 and err[1] < err[0] / 3, err


def test_periodic_uniform_magnetisation_gives_the_box_field(orc):
    """Fully periodic grid (sizes 4, 6, 8; 2, 1, 3 images): the convolution over
    the unpadded period turns a uniform m into the uniform field of the image
    box (exchange adds nothing to a uniform state), along each axis."""
    size, per, cell, box = PERIODIC_CASES[0]
    want = _box_demag([b[0] for b in box], [b[1] for b in box], cell)
    N = int(np.prod(size))
    for u in range(3):
        m = np.zeros((3, N))
        m[u] = 1.0
        s = orc.sim(m.astype(np.float32), size=size, cellsize=tuple(c * 1e-9 for c in cell), msat=8e5,
                    aexch=1.3e-11, alpha=0.1, periodic=per, demag_accuracy=16)
        h = s.heff(m).reshape(3, N)
        assert np.abs(h[u] - h[u].mean()).max() < 1e-9
        assert h[u].mean() == pytest.approx(want[u], abs=1e-5)
        assert np.abs(np.delete(h, u, axis=0)).max() < 1e-9


@pytest.mark.parametrize("size,per", [((4, 6, 8), (2, 1, 3)), ((1, 6, 10), (0, 3, 2)), ((3, 5, 6), (1, 0, 1))])
def test_periodic_field_commutes_with_shifts_along_periodic_axes(orc, size, per):
    """Along a periodic axis the field of a random m shifted by one cell is the
    field shifted by one cell (a circular convolution, demag + exchange); along a
    free axis it is not."""
    rng = np.random.default_rng(3)
    m = rng.standard_normal((3,) + size)
    m /= np.linalg.norm(m, axis=0)
    s = orc.sim(m.astype(np.float32), size=size, cellsize=(2e-9, 3e-9, 2.5e-9), msat=8e5, aexch=1.3e-11, alpha=0.1,
                periodic=per)
    h = s.heff(m).reshape((3,) + size)
    for a in range(3):
        if size[a] == 1:
            continue
        hs = s.heff(np.roll(m, 1, axis=1 + a)).reshape((3,) + size)
        d = np.abs(hs - np.roll(h, 1, axis=1 + a)).max()
        if per[a]:
            assert d < 1e-9 * np.abs(h).max(), (a, d)
        else:
            assert d > 1e-3 * np.abs(h).max(), (a, d)


def test_facekernel_far_field_is_point_dipole(orc):
    P = (2, 2, 64)
    cell = (1.0, 1.0, 1.0)
    K = orc.facekernel6(P, cell)
    for d in (10, 20, 30):
        # dipole of volume 1 along z at distance d on the z axis: Bz = 2/(4 pi d^3), Bx = By = 0
        assert K[2, 0, 0, d] == pytest.approx(2 / (4 * np.pi * d ** 3), rel=2e-3)
        # transverse: Bx from an x dipole = -1/(4 pi d^3)
        assert K[0, 0, 0, d] == pytest.approx(-1 / (4 * np.pi * d ** 3), rel=2e-3)


def test_facekernel_symmetry(orc):
    P = (4, 6, 8)
    K = orc.facekernel6(P, (0.7, 1.0, 1.3))

    def mirror(a, axes):
        for ax in axes:
            a = np.roll(np.flip(a, ax), 1, ax)
        return a

    # diagonal comps even in every axis; xy odd in x,y; xz odd in x,z; yz odd in y,z
    parity = {0: (1, 1, 1), 1: (1, 1, 1), 2: (1, 1, 1), 3: (1, -1, -1), 4: (-1, 1, -1), 5: (-1, -1, 1)}
    for c, par in parity.items():
        for ax in range(3):
            m = mirror(K[c], [ax])
            # the row at index P/2 carries no displacement and stays zero
            sl = [slice(None)] * 3
            sl[ax] = P[ax] // 2
            assert np.all(K[c][tuple(sl)] == 0)
            assert np.allclose(m, par[ax] * K[c], atol=1e-14)


def test_facekernel_float32_twin(orc):
    P = (2, 8, 8)
    cell = (0.5, 0.7, 0.7)
    K64 = orc.facekernel6(P, cell)
    K32 = orc.facekernel6_f32(P, np.asarray(cell, np.float32))
    assert np.abs(K32 - K64).max() <= 2e-6 * np.abs(K64).max()


# ------------------------------------------------------------------ exchange (exchange.go:44-69)
def test_exchange_kernel_spin_wave(orc):
    # 3-D interior cells: h = -(2/hz^2)(1 - cos k) m for a spin wave along z
    N = (4, 4, 64)
    P = (8, 8, 128)
    cell = (1.0, 0.8, 0.5)
    K = np.zeros((6,) + P)
    orc.add_exch6_kernel(K, cell)
    c = orc.conv_new(N, P, K)
    k = 2 * np.pi * 5 / 64
    z = np.arange(N[2])
    m = np.zeros((3,) + N)
    m[0] = np.cos(k * z)[None, None, :]
    h = c.apply(m)
    expect = -(2 / cell[2] ** 2) * (1 - np.cos(k)) * m[0]
    sl = (slice(1, 3), slice(1, 3), slice(1, N[2] - 1))
    assert np.allclose(h[0][sl], expect[sl], atol=1e-10)


def test_exchange_2d_centre_quirk(orc):
    P = (1, 8, 8)
    K = np.zeros((6,) + P)
    orc.add_exch6_kernel(K, (2.0, 1.0, 1.0))
    # wrap(+-1, 1) = 0 overwrites the centre with 1/hx^2 (exchange.go:58-61)
    assert K[0, 0, 0, 0] == pytest.approx(1 / 4.0)
    assert K[0, 0, 1, 0] == 1.0 and K[0, 0, 0, 1] == 1.0
    assert K[3:].sum() == 0


def test_conv_exchange_equals_stencil_3d(orc):
    N = (3, 5, 6)
    P = (6, 10, 12)
    cell = (0.9, 1.1, 1.3)
    K = np.zeros((6,) + P)
    orc.add_exch6_kernel(K, cell)
    rng = np.random.default_rng(0)
    m = rng.standard_normal((3,) + N)
    h_conv = orc.conv_new(N, P, K).apply(m)
    h_sten = orc.exch6_stencil(m, np.zeros_like(m), N, cell)
    assert np.allclose(h_conv, h_sten, atol=1e-11)


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 11, 12, 13, 16, 30, 64, 96, 100])
def test_fft_brute_force(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    j = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(j, j) / n)
    assert np.allclose(orc.fft(x.copy()), x @ W.T, atol=1e-10 * n)
    assert np.allclose(orc.fft(x.copy(), +1), x @ W.conj().T, atol=1e-10 * n)


def test_fft_round_trip_parseval_impulse(orc):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 256)) + 1j * rng.standard_normal((3, 256))
    X = orc.fft(x.copy())
    assert np.allclose(orc.fft(X.copy(), +1) / 256, x, atol=1e-12)
    assert np.allclose((np.abs(X) ** 2).sum(axis=1) / 256, (np.abs(x) ** 2).sum(axis=1))
    d = np.zeros((1, 64), complex)
    d[0, 0] = 1
    assert np.allclose(orc.fft(d), 1.0)


def test_fft3_matches_numpy(orc):
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 6, 10)) + 1j * rng.standard_normal((4, 6, 10))
    assert np.allclose(orc.fft3(x.copy()), np.fft.fftn(x), atol=1e-11)


# ------------------------------------------------------------------ convolution (conv.go)
def _symmetric(K):
    R = K.copy()
    for ax in (1, 2, 3):
        R = np.roll(np.flip(R, ax), 1, ax)
    return 0.5 * (K + R)


@pytest.mark.parametrize("N,P", [((2, 3, 4), (4, 6, 8)), ((1, 5, 7), (1, 10, 14)), ((3, 4, 5), (6, 8, 10))])
def test_conv_matches_direct(orc, N, P):
    rng = np.random.default_rng(3)
    K = _symmetric(rng.standard_normal((6,) + P))
    m = rng.standard_normal((3,) + N)
    h1 = orc.conv_new(N, P, K).apply(m)
    h2 = orc.conv_direct(N, P, K, m)
    assert np.allclose(h1, h2, atol=1e-11)


def test_conv_unit_and_zero_kernel(orc):
    N, P = (2, 4, 4), (4, 8, 8)
    m = np.random.default_rng(4).standard_normal((3,) + N)
    K = np.zeros((6,) + P)
    assert np.abs(orc.conv_new(N, P, K).apply(m)).max() == 0
    for c in range(3):
        K[c, 0, 0, 0] = 1.0  # UnitKernel6 (kernel.go:47-58)
    assert np.allclose(orc.conv_new(N, P, K).apply(m), m, atol=1e-14)


def test_padded_forward_equals_manual_padding(orc):
    # fft_test.go:79-165: padded transform == transform of the manually padded array
    N, P = (1, 16, 16), (1, 32, 32)
    rng = np.random.default_rng(5)
    x = rng.random(N)
    c = orc.conv_new(N, P, np.zeros((6,) + P))
    got = c.forward1(x)
    xp = np.zeros(P)
    xp[: N[0], : N[1], : N[2]] = x
    ref = np.fft.fftn(xp)[:, :, : P[2] // 2 + 1]
    assert np.allclose(got, ref, atol=1e-12)


def test_kernel_spectrum_real_and_symmetric(orc):
    P = (4, 16, 16)
    K = orc.facekernel6(P, (1.0, 1.0, 1.0))
    c = orc.conv_new((2, 8, 8), P, K)
    S = c.kernel_spectrum()
    full = np.fft.fftn(K, axes=(1, 2, 3)).real / np.prod(P)
    assert np.allclose(S, full[..., : P[2] // 2 + 1], atol=1e-14)
    assert np.abs(np.fft.fftn(K, axes=(1, 2, 3)).imag).max() < 1e-12


# ------------------------------------------------------------------ pointwise
def test_local_fields_closed_form(orc):
    rng = np.random.default_rng(6)
    m = orc.normalize(rng.standard_normal((3, 10)))
    u = np.array([0.0, 0.6, 0.8])
    h = orc.local_fields(m, np.zeros_like(m), np.array([0.1, -0.2, 0.3]), 1, 0.7, u)
    mu = (m * u[:, None]).sum(0)
    assert np.allclose(h, np.array([0.1, -0.2, 0.3])[:, None] + 0.7 * mu * u[:, None])


def test_torque_closed_form(orc):
    rng = np.random.default_rng(7)
    m = orc.normalize(rng.standard_normal((3, 50)))
    h = rng.standard_normal((3, 50))
    a = 0.3
    k = orc.deltaM(m, h, a, 1 / (1 + a * a))
    mxh = np.cross(m.T, h.T).T
    ref = (mxh - a * np.cross(m.T, mxh.T).T) / (1 + a * a)
    assert np.allclose(k, ref)
    assert np.abs((m * k).sum(0)).max() < 1e-14  # m . k = 0
    # |k| = |m x h| sqrt(1 + a^2) / (1 + a^2) for |m| = 1
    assert np.allclose(np.linalg.norm(k, axis=0), np.linalg.norm(mxh, axis=0) / np.sqrt(1 + a * a))


def test_sim_torque_with_alpha_map(orc):
    """The simulation's torque with a per-cell damping multiplier: damping
    alpha * map[i], prefactor 1/(1+alpha^2) of the global alpha (torque.go:49-51,
    gpu_torque.cu's alpha_map); a map of ones is no map."""
    size, cell = (2, 6, 8), (3e-9, 3e-9, 3e-9)
    rng = np.random.default_rng(11)
    m = orc.normalize(rng.standard_normal((3,) + size).reshape(3, -1)).reshape((3,) + size).astype(np.float32)
    amap = rng.uniform(0.2, 3.0, size).astype(np.float32)
    a = 0.3
    s = orc.sim(m, size=size, cellsize=cell, msat=8e5, aexch=1.3e-11, alpha=a, solver="rk4", dt=1e-14)
    k0 = s.torque(m).reshape(3, -1)
    h = s.heff(m).reshape(3, -1)
    s.set_alpha_map(np.ones(size, np.float32))
    assert np.array_equal(s.torque(m).reshape(3, -1), k0)
    s.set_alpha_map(amap)
    k = s.torque(m).reshape(3, -1)
    m2 = m.reshape(3, -1).astype(np.float64)
    mxh = np.cross(m2.T, h.T).T
    ref = (mxh - a * amap.reshape(-1).astype(np.float64) * np.cross(m2.T, mxh.T).T) / (1 + a * a)
    assert np.allclose(k, ref, rtol=1e-12, atol=1e-14)
    s.set_alpha_map(None)
    assert np.array_equal(s.torque(m).reshape(3, -1), k0)


def test_field_functions_special_values(orc):
    """ApplyPulse / ApplyRf / ApplyRfRamp / ApplySawtooth / ApplyRotating
    (simfield.go:155-260) at points where each formula has a known value."""
    size, cell = (1, 4, 4), (3e-9, 3e-9, 3e-9)
    m = np.zeros((3,) + size, np.float32)
    m[2] = 1
    s = orc.sim(m, size=size, cellsize=cell, msat=8e5, aexch=1.3e-11, alpha=0.1, solver="rk4", dt=1e-14)
    b = np.array([0.25, -0.5, 0.125])            # exact in float32
    f = 2e9
    s.set_field_function(2, b, [f])              # RF: sin(2 pi f t)
    assert np.allclose(s.field_at(1 / (4 * f)), b) and np.allclose(s.field_at(3 / (4 * f)), -b)
    assert np.abs(s.field_at(1 / f)).max() < 1e-7
    s.set_field_function(4, b, [f])              # sawtooth: a triangle wave of amplitude 1
    assert np.allclose(s.field_at(1 / (4 * f)), b) and np.allclose(s.field_at(1 / (8 * f)), 0.5 * b, rtol=1e-6)
    assert np.allclose(s.field_at(5 / (8 * f)), -0.5 * b, rtol=1e-6)
    s.set_field_function(3, b, [f, 1e-9])        # RF ramp: linear envelope up to 1 ns
    assert np.allclose(s.field_at(1 / (4 * f)), b * (1 / (4 * f)) / 1e-9, rtol=1e-6)
    assert np.allclose(s.field_at(2e-9 + 1 / (4 * f)), b)
    s.set_field_function(1, b, [2e-9, 1e-9])     # pulse: 1 ns cosine rise, 2 ns flat, 1 ns fall
    for t, v in ((-1e-9, 0.0), (0.5e-9, 0.5), (1e-9, 1.0), (2.9e-9, 1.0), (3.5e-9, 0.5), (4.2e-9, 0.0)):
        assert np.allclose(s.field_at(t), v * b, atol=1e-7), t
    s.set_field_function(5, b, [f, 0.0, np.pi / 2, np.pi])   # rotating, per-component phase
    assert np.allclose(s.field_at(0.0), [0.0, b[1], 0.0], atol=1e-7)
    assert np.allclose(s.field_at(1 / (4 * f)), [b[0], 0.0, -b[2]], atol=1e-7)


def test_normalize(orc):
    m = np.random.default_rng(8).standard_normal((3, 40)) * 3
    assert np.allclose(np.linalg.norm(orc.normalize(m), axis=0), 1.0)


# ------------------------------------------------------------------ reductions (redux_test.go:13-102)
NS = [2, 8, 16, 31, 33, 128, 255, 511, 1023, 1024, 1025, 20000]


@pytest.mark.parametrize("N", NS)
def test_reduce_known_answers(orc, N):
    assert orc.reduce(REDUCE_ADD, np.ones(N)) == N
    assert orc.reduce(REDUCE_SUMABS, -np.ones(N)) == N
    x = np.ones(N)
    x[12349 % N] = 10.0
    assert orc.reduce(REDUCE_MAX, x) == 10.0
    x[12349 % N] = -10.0
    assert orc.reduce(REDUCE_MAXABS, x) == 10.0
    assert orc.reduce(REDUCE_MIN, x) == -10.0


# ------------------------------------------------------------------ RK stepping (rk.go)
def _macrospin(orc, solver, dt, alpha=0.1, B=0.1, t_end_internal=50.0, maxerror=1e-6):
    msat = 8e5
    u = orc.units(msat, 1.3e-11)
    th0 = 0.3
    m0 = np.zeros((3, 1, 1, 2), np.float32)
    m0[0], m0[1], m0[2] = np.sin(th0), 0.0, np.cos(th0)
    s = orc.sim(m0, size=(1, 1, 2), cellsize=(1e-9, 1e-9, 1e-9), msat=msat, aexch=1.3e-11, alpha=alpha,
                solver=solver, dt=dt * u["time"], maxerror=maxerror, want_demag=False, exch_type=0,
                field=(np.zeros(1), np.array([0.0, 0.0, B])))
    s.run(t_end_internal * u["time"])
    st = s.state()
    t = st["time"]
    H = B / u["field"]
    dtg = 1 / (1 + alpha ** 2)
    # internal-coordinate LL (the reference's closed form, dm/dt = -m x h sense): mz = tanh(a dtg H t + atanh(mz0)), phi = phi0 - dtg H t
    mz = np.tanh(alpha * dtg * H * t + np.arctanh(np.cos(th0)))
    phi = -dtg * H * t
    r = np.sqrt(1 - mz ** 2)
    exact = np.array([r * np.cos(phi), r * np.sin(phi), mz])
    got = s.m.reshape(3, 2)[:, 0]
    return np.abs(got - exact).max(), st


def test_rk4_macrospin_fourth_order(orc):
    e1, _ = _macrospin(orc, "rk4", dt=2.0)
    e2, _ = _macrospin(orc, "rk4", dt=1.0)
    assert e2 < 1e-6
    assert 10 < e1 / e2 < 24  # 2^4 = 16


def test_rk2_macrospin_second_order(orc):
    e1, _ = _macrospin(orc, "rk2", dt=0.5)
    e2, _ = _macrospin(orc, "rk2", dt=0.25)
    assert 3 < e1 / e2 < 5.5


@pytest.mark.parametrize("solver", ["rkdp", "rkck", "rk23", "rk12"])
def test_adaptive_macrospin(orc, solver):
    err, st = _macrospin(orc, solver, dt=0.0, maxerror=1e-7)
    assert err < 2e-5
    assert st["steps"] > 3


def test_forced_reject_path(orc):
    # a huge first dt with a tiny maxError must be rejected and retried (rk.go:425-428)
    err, st = _macrospin(orc, "rkdp", dt=40.0, maxerror=1e-9, t_end_internal=50.0)
    assert st["rejects"] >= 1
    assert err < 1e-5


# ------------------------------------------------------------------ step-size controller (rk.go:404-420)
def _ring(orc, solver, n=64, maxerror=1e-5, dt=0.0):
    """n uncoupled cells (no demag, no exchange) precessing undamped about
    internal X at one tilt, their azimuths spread evenly over the circle: the
    max-abs of the error estimate over all cells no longer depends on the
    precession phase, so err is a function of dt alone, err ~ C dt^(p+1) for an
    embedded estimate of local order p+1."""
    size = (1, 1, n)
    ph = 2 * np.pi * np.arange(n) / n
    m0 = np.zeros((3,) + size, np.float32)
    m0[0, 0, 0] = 0.6
    m0[1, 0, 0] = 0.8 * np.cos(ph)
    m0[2, 0, 0] = 0.8 * np.sin(ph)
    u = orc.units(8e5, 1.3e-11)
    return orc.sim(m0, size=size, cellsize=(3e-9,) * 3, msat=8e5, aexch=1.3e-11, alpha=0.0, solver=solver,
                   maxerror=maxerror, dt=dt * u["time"], want_demag=False, exch_type=0,
                   field=(np.zeros(1), np.array([[0.1, 0.0, 0.0]])))


@pytest.mark.parametrize("solver", ["rkdp", "rkck", "rk23", "rk12"])
def test_controller_growth_is_clamped_at_two(orc, solver):
    """From the default first dt (1e-4 internal, sim.go:30) the error is far
    below maxError, so the first trials double dt exactly: the factor
    (maxError/err)^(1/errororder) is clamped to 2 (rk.go:409-411)."""
    s = _ring(orc, solver)
    for _ in range(9):
        s.step()
    dt, acc = s.trial_log()
    assert dt[0] == 1e-4
    assert np.all(acc[:9] == 1)
    assert np.array_equal(dt[1:9], dt[:8] * 2.0)


def test_controller_shrink_is_clamped_at_one_hundredth(orc):
    """A first dt far beyond the error bar is rejected and the retry uses
    exactly 0.01 of it: the factor's lower clamp (rk.go:412-414) and the reject
    test err > 2 maxError (rk.go:421-428)."""
    s = _ring(orc, "rkdp", dt=5000.0)
    s.step()
    dt, acc = s.trial_log()
    assert acc[0] == 0 and dt[1] == dt[0] * 0.01


@pytest.mark.parametrize("solver,order,lo,hi", [("rkdp", 4, -0.30, -0.22), ("rkck", 4, -0.33, -0.22),
                                                ("rk23", 2, -0.56, -0.44)])
def test_controller_log_dt_contracts_by_one_minus_local_order_over_errororder(orc, solver, order, lo, hi):
    """With err = C dt^(order+1) the update l' = l + (log maxError - log err) /
    errororder of l = log dt moves a deviation from the fixed point by the
    factor 1 - (order+1)/errororder: -1/4 for Dormand-Prince and Cash-Karp
    (errororder 4, rk.go:196-212, 170-195) and -1/2 for Bogacki-Shampine
    (errororder 2, rk.go:160-169).  An errororder of 5 would give 0 (the
    first correction lands on the fixed point) and one of 3 -2/3.  Checked as
    the ratio of successive log-dt steps once the doubling phase is over,
    while the steps stand well above rounding; the controller then sits at
    err = maxError (no safety factor)."""
    s = _ring(orc, solver)
    for _ in range(40):
        s.step()
    dt, acc = s.trial_log()
    d = np.diff(np.log(dt))
    k = int(np.argmax(np.abs(d - np.log(2.0)) > 1e-12))  # first step that is not a doubling
    big = d[k:][np.abs(d[k:]) > 2e-3]
    n = int(np.argmax(np.abs(d[k:]) <= 2e-3))
    ratios = d[k + 1:k + n] / d[k:k + n - 1]
    assert len(ratios) >= 3, (big, ratios)
    assert np.all((ratios > lo) & (ratios < hi)), ratios
    assert abs(s.state()["error"] / 1e-5 - 1) < 2e-3


# ------------------------------------------------------------------ energy (energy.go, field.go:64-87)
def _energy_sim(orc, m0, **kw):
    cell = kw.pop("cellsize", (2e-9, 2e-9, 2e-9))
    return orc.sim(np.asarray(m0, np.float32), size=m0.shape[1:], cellsize=cell, msat=8e5, aexch=1.3e-11,
                   alpha=0.1, **kw)


def test_energy_demag_is_half_m_dot_h(orc):
    # phi = -m.H_demag / 2 per unit volume (calcEDensDemagExch, energy.go:20-23); H itself is
    # pinned by the kernel and convolution tests above.  (A single cell, or a row one cell wide
    # in y or z, has no demag kernel at all: FaceKernel6's loop is empty there, demag.go:42-46.)
    rng = np.random.default_rng(5)
    m0 = rng.standard_normal((3, 2, 3, 4))
    m0 /= np.linalg.norm(m0, axis=0)
    s = _energy_sim(orc, m0, exch_type=0, exch_in_conv=False)
    h = s.heff(m0.astype(np.float32)).reshape(3, -1)
    m = m0.astype(np.float32).astype(float).reshape(3, -1)
    u = orc.units(8e5, 1.3e-11)
    vol = (2e-9 / u["length"]) ** 3
    assert abs(h).max() > 0.05
    assert s.energy() == pytest.approx(-0.5 * (m * h).sum() * vol, rel=1e-9)


def test_energy_zeeman_and_sign(orc):
    # phi -= m.Hext (no factor 1/2), summed times the cell volume (energy.go:48-51)
    m0 = np.zeros((3, 1, 2, 3))
    m0[0], m0[2] = 0.6, 0.8
    B = np.array([0.01, 0.02, -0.03])
    s = _energy_sim(orc, m0, want_demag=False, exch_type=0, field=(np.zeros(1), B))
    u = orc.units(8e5, 1.3e-11)
    vol = (2e-9 / u["length"]) ** 3
    mdotb = float(np.float32(0.6)) * B[0] + float(np.float32(0.8)) * B[2]
    assert s.energy() == pytest.approx(-6 * vol * mdotb / u["field"], rel=1e-7)


def test_energy_anisotropy_is_half_m_dot_h(orc):
    # uniaxial: phi -= k (m.u)^2 / 2, i.e. -m.H_anis / 2 (gpu_local_contr2.cu:156)
    rng = np.random.default_rng(3)
    m0 = rng.standard_normal((3, 1, 1, 5))
    m0 /= np.linalg.norm(m0, axis=0)
    s = _energy_sim(orc, m0, want_demag=False, exch_type=0, k1=4e5, anis_axis=(0.0, 0.6, 0.8))
    h = s.heff(m0.astype(np.float32)).reshape(3, -1)
    m = m0.astype(np.float32).astype(float).reshape(3, -1)
    u = orc.units(8e5, 1.3e-11)
    vol = (2e-9 / u["length"]) ** 3
    assert abs(h).max() > 0.1
    assert s.energy() == pytest.approx(-0.5 * (m * h).sum() * vol, rel=1e-6)


def test_energy_exchange_two_cells(orc):
    # 6-neighbour stencil, two cells at angle t along Z, a film (no X term).  A missing
    # neighbour counts as m = 0 (gpu_exch1.cu), so H_i = -(4/h^2) m_i + m_j / h^2 and
    # E = vol sum_i -m_i.H_i / 2 = vol (4 - cos t) / h^2
    t = 0.7
    m0 = np.zeros((3, 1, 1, 2))
    m0[2, 0, 0, 0] = 1.0
    m0[2, 0, 0, 1], m0[1, 0, 0, 1] = np.cos(t), np.sin(t)
    s = _energy_sim(orc, m0, want_demag=False, exch_type=6, exch_in_conv=False)
    u = orc.units(8e5, 1.3e-11)
    h = 2e-9 / u["length"]
    m = m0.astype(np.float32).astype(float)
    cos_t = (m[:, 0, 0, 0] * m[:, 0, 0, 1]).sum()
    assert s.energy() == pytest.approx(h ** 3 * (4 - cos_t) / h ** 2, rel=1e-6)


# ------------------------------------------------------------------ Relax / live field (relax.go, simfield.go)
def _film(orc, field=None, alpha=1.0):
    size = (1, 16, 32)
    m0 = np.zeros((3,) + size, np.float32)
    m0[2], m0[1] = 1.0, 0.3
    m0 /= np.linalg.norm(m0, axis=0)
    kw = {} if field is None else dict(field=(np.zeros(1), np.asarray(field, float)))
    return orc.sim(m0, size=size, cellsize=(3e-9, 3.90625e-9, 3.90625e-9), msat=8e5, aexch=1.3e-11, alpha=alpha,
                   solver="rkdp", maxerror=1e-4, **kw)


def test_set_field_equals_field_at_creation(orc):
    B = (0.0, 0.004, -0.02)
    a = _film(orc, field=B)
    b = _film(orc)
    b.set_field(np.zeros(1), [B])
    for _ in range(20):
        a.step()
        b.step()
    assert np.array_equal(a.m, b.m)


def test_relax_lowers_energy_and_stops(orc):
    s = _film(orc)
    e0 = s.energy()
    n = s.relax()
    st = s.state()
    assert 5 < n < 2000 and st["steps"] == n
    assert st["energy"] < e0
    # every accepted step kept the energy from rising; the last one raised it
    s2 = _film(orc)
    s2.orc.lib.or_sim_set_want_energy(s2.ptr, 1)
    es = []
    for _ in range(n):
        s2.step()
        es.append(s2.state()["energy"])
    assert np.array_equal(s2.m, s.m)
    assert all(es[i + 1] <= es[i] for i in range(n - 2))
    assert es[-1] > es[-2]


def test_float32_coefficients_are_equivalent(orc):
    # the reference's tableau literals are float32 (rk.go:62-212); the oracle's exact
    # rationals give the same trajectory to well below the FP32 engine's own rounding
    import ctypes

    orc.lib.or_set_tableau_f32.argtypes = [ctypes.c_int]
    ms = []
    for f32 in (0, 1):
        orc.lib.or_set_tableau_f32(f32)
        s = _film(orc, field=(0.0, 0.004, -0.02), alpha=0.1)
        for _ in range(40):
            s.step()
        ms.append(s.m)
    orc.lib.or_set_tableau_f32(0)
    d = np.abs(ms[0] - ms[1]).max()
    assert 0 < d < 1e-6


# ------------------------------------------------------------------ thermal field (temperature.go)
def test_thermal_normals_are_standard_independent_and_reproducible(orc):
    """The thermal stream's normals: mean 0 and variance 1 within 5 standard
    errors, a normal distribution by Kolmogorov-Smirnov, no correlation
    between components, cells or steps, and the same draw for the same (seed,
    step, cell)."""
    from scipy import stats

    n = 20000
    a = orc.thermal_normals(7, 3, range(n))          # step 3, cells 0..n
    b = orc.thermal_normals(7, 4, range(n))          # the next step
    c = orc.thermal_normals(8, 3, range(n))          # another seed
    x = a.reshape(-1)
    se = 1 / np.sqrt(x.size)
    assert abs(x.mean()) < 5 * se
    assert abs(x.var() - 1) < 5 * np.sqrt(2 / x.size)
    assert stats.kstest(x, "norm").pvalue > 1e-3
    for u, v in [(a[:, 0], a[:, 1]), (a[:, 0], a[:, 2]), (a[:, 1], a[:, 2]), (a[:-1, 0], a[1:, 0]),
                 (a[:, 0], b[:, 0]), (a[:, 2], c[:, 2])]:
        assert abs(np.corrcoef(u, v)[0, 1]) < 5 / np.sqrt(n)
    assert np.array_equal(orc.thermal_normals(7, 3, [0, 123, 19999]), a[[0, 123, 19999]])


def _thermal_film(orc, T, dt=2e-13, seed=1):
    size, cell = (1, 8, 16), (3e-9, 4e-9, 4e-9)
    m0 = np.zeros((3,) + size, np.float32)
    m0[2] = 1
    return orc.sim(m0, size=size, cellsize=cell, msat=8e5, aexch=1.3e-11, alpha=0.1, solver="rk12",
                   maxerror=1e-3, dt=dt, temperature=T, seed=seed), m0


def test_thermal_field_in_heff_scales_with_temperature_and_dt(orc):
    """calcHeff adds the thermal field (field.go:74-80): H(T) - H(0) is the
    stream's normals times one standard deviation, which grows as sqrt(T) and
    as 1/sqrt(dt) (updateTempNoise, temperature.go:29-36)."""
    s0, m0 = _thermal_film(orc, 0.0)
    h0 = s0.heff(m0)
    d = {}
    for T, dt in [(300.0, 2e-13), (1200.0, 2e-13), (300.0, 0.5e-13)]:
        s, _ = _thermal_film(orc, T, dt)
        d[(T, dt)] = (s.heff(m0) - h0).reshape(3, -1)
    n = orc.thermal_normals(1, 0, range(d[(300.0, 2e-13)].shape[1])).T   # step 0, seed 1
    sd = d[(300.0, 2e-13)] / n
    assert np.allclose(sd, sd.flat[0], rtol=1e-9)                         # one sigma for every cell and component
    assert sd.flat[0] > 0
    assert np.allclose(d[(1200.0, 2e-13)], 2 * d[(300.0, 2e-13)], rtol=1e-9)   # sqrt(4 T)
    assert np.allclose(d[(300.0, 0.5e-13)], 2 * d[(300.0, 2e-13)], rtol=1e-9)  # 1 / sqrt(dt / 4)


def test_thermal_field_is_redrawn_per_step_and_kept_through_a_step(orc):
    """The field is drawn at the start of each step (rk.go:289-292): after one
    step it is the next step's draw; a step replayed with a rejected first
    trial keeps one draw, from the dt of its first trial."""
    s, m0 = _thermal_film(orc, 300.0)
    s.step()
    m = s.m.reshape(3, -1).astype(np.float32)
    s0, _ = _thermal_film(orc, 0.0)
    s0.m = s.m
    ratio = (s.heff(m) - s0.heff(m)).reshape(3, -1) / orc.thermal_normals(1, 0, range(m.shape[1])).T
    assert np.allclose(ratio, ratio.flat[0], rtol=1e-9)    # heff still shows step 0's draw until the next step starts
    s2, _ = _thermal_film(orc, 300.0)
    s3, _ = _thermal_film(orc, 300.0)
    s2.step_replay([2e-13 / s2.state()["unit_time"]] * 2, [0, 1])
    s3.step_replay([2e-13 / s3.state()["unit_time"]], [1])
    assert np.array_equal(s2.m, s3.m)                       # the rejected trial left no trace, same draw
