"""GPU parity of the drop-in entry points (include/gpukern.h) through the C ABI.

Each entry point is called exactly as gpu.go calls it and compared with the
oracle, with numpy, or with the reference's own known answers
(redux_test.go, fft_test.go, zeropad_test.cpp).  Integer/index work
(padding, reductions of exact values) must be bit-exact; float32 arithmetic
is held to the tolerances written in each test.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32


class Dev:
    """A device array owned through new_gpu_array/free_gpu_array."""

    def __init__(self, lib, host=None, n=None):
        self.lib = lib
        if host is not None:
            host = np.ascontiguousarray(host, F32).reshape(-1)
            n = host.size
        self.n = int(n)
        self.p = lib.new_gpu_array(self.n)
        assert self.p
        if host is not None:
            lib.memcpy_to_gpu(host.ctypes.data, self.p, self.n)

    def get(self):
        out = np.empty(self.n, F32)
        self.lib.memcpy_from_gpu(self.p, out.ctypes.data, self.n)
        return out

    def free(self):
        if self.p:
            self.lib.free_gpu_array(self.p)
            self.p = None

    def __del__(self):
        self.free()


def f3(*v):
    return (C.c_float * len(v))(*v)


def i3(*v):
    return (C.c_int * len(v))(*v)


# ------------------------------------------------------------------ memory (gpu_mem.cu)
def test_memory(lib):
    before = lib.gpu_usedmem()
    d = Dev(lib, n=1000)
    assert lib.gpu_usedmem() - before == 4000
    assert np.all(d.get() == 0)  # new_gpu_array returns zeroed memory
    x = np.arange(1000, dtype=F32)
    lib.memcpy_gpu_dir(x.ctypes.data, d.p, 1000, 1)
    assert np.array_equal(d.get(), x)
    e = Dev(lib, n=1000)
    lib.memcpy_gpu_dir(d.p, e.p, 1000, 2)
    y = np.empty(1000, F32)
    lib.memcpy_gpu_dir(e.p, y.ctypes.data, 1000, 3)
    assert np.array_equal(y, x)
    assert lib.gpu_array_get(d.p, 17) == 17.0
    lib.gpu_array_set(d.p, 17, F32(-3.5))
    assert lib.gpu_array_get(d.p, 17) == -3.5
    lib.gpu_zero(d.p, 1000)
    assert np.all(d.get() == 0)
    assert lib.gpu_stride_float() >= 1
    lib.gpu_override_stride(7)
    assert lib.gpu_stride_float() == 7
    assert lib.gpu_maxthreads() >= 256
    d.free()
    e.free()
    assert lib.gpu_usedmem() == before


# ------------------------------------------------------------------ linalg (gpu_linalg.cu)
def test_linalg(lib):
    rng = np.random.default_rng(0)
    n = 3000
    a, b, c = (rng.standard_normal(n).astype(F32) for _ in range(3))
    da, db, dc = Dev(lib, a), Dev(lib, b), Dev(lib, c)
    lib.gpu_add(da.p, db.p, n)
    assert np.array_equal(da.get(), a + b)
    da = Dev(lib, a)
    lib.gpu_madd(da.p, F32(0.37), db.p, n)
    assert np.allclose(da.get(), a + F32(0.37) * b, rtol=1e-6, atol=1e-6)
    da = Dev(lib, a)
    lib.gpu_madd2(da.p, db.p, dc.p, n)
    assert np.allclose(da.get(), a + b * c, rtol=1e-6, atol=1e-6)
    da = Dev(lib, a)
    lib.gpu_add_constant(da.p, F32(2.5), n)
    assert np.array_equal(da.get(), a + F32(2.5))
    da = Dev(lib, a)
    lib.gpu_linear_combination(da.p, db.p, F32(0.5), F32(-2.0), n)
    assert np.allclose(da.get(), 0.5 * a - 2.0 * b, rtol=1e-6, atol=1e-6)
    da = Dev(lib, a)
    vecs = (C.c_void_p * 2)(db.p, dc.p)
    w = np.array([0.25, -1.5], F32)
    lib.gpu_linear_combination_many(da.p, C.cast(vecs, C.c_void_p), w.ctypes.data, 2, n)
    assert np.allclose(da.get(), a + 0.25 * b - 1.5 * c, rtol=1e-6, atol=1e-6)
    N = n // 3
    r = Dev(lib, n=N)
    lib.gpu_scale_dot_product(r.p, db.p, dc.p, F32(-0.5), N)
    ref = -0.5 * (b.reshape(3, N) * c.reshape(3, N)).sum(0)
    assert np.allclose(r.get(), ref, rtol=1e-5, atol=1e-6)


# ------------------------------------------------------------------ reductions (redux_test.go:13-102)
NS = [2, 8, 16, 31, 33, 128, 255, 511, 1023, 1024, 1025, 20000]


def reductor(lib, N):
    # Reductor.init (redux.go:103-121)
    threads = lib.gpu_maxthreads() // 2
    while N <= threads:
        threads //= 2
    blocks = -(-N // (threads * 2))
    return blocks, threads


@pytest.mark.parametrize("N", NS)
def test_reduce_known_answers(lib, N):
    blocks, threads = reductor(lib, N)
    dev = Dev(lib, n=blocks)
    host = np.zeros(blocks, F32)

    def red(op, x):
        d = Dev(lib, x)
        return lib.gpu_reduce(op, d.p, dev.p, host.ctypes.data, blocks, threads, N)

    assert red(1, np.ones(N)) == N
    assert red(5, -np.ones(N)) == N
    x = np.ones(N, F32)
    x[12349 % N] = 10.0
    assert red(2, x) == 10.0
    x[12349 % N] = -10.0
    assert red(3, x) == 10.0
    assert red(4, x) == -10.0


def test_reduce_random_vs_oracle(lib, orc):
    rng = np.random.default_rng(1)
    N = 123457
    x = rng.standard_normal(N).astype(F32)
    blocks, threads = reductor(lib, N)
    dev = Dev(lib, n=blocks)
    host = np.zeros(blocks, F32)
    d = Dev(lib, x)
    for op in (1, 2, 3, 4, 5):
        got = lib.gpu_reduce(op, d.p, dev.p, host.ctypes.data, blocks, threads, N)
        ref = orc.reduce(op, x.astype(np.float64))
        if op in (2, 3, 4):
            assert got == F32(ref)  # max/min are exact
        else:
            assert got == pytest.approx(ref, rel=1e-6, abs=1e-3)


# ------------------------------------------------------------------ normalize / torque
def test_normalize(lib):
    rng = np.random.default_rng(2)
    m = (rng.standard_normal((3, 777)) * 3).astype(F32)
    d = Dev(lib, m)
    lib.gpu_normalize_uniform(d.p, 777)
    got = d.get().reshape(3, 777)
    assert np.allclose(got, m / np.linalg.norm(m.astype(np.float64), axis=0), atol=1e-6)
    mp = rng.random(777).astype(F32)
    mp[::7] = 0
    d = Dev(lib, m)
    dm = Dev(lib, mp)
    lib.gpu_normalize_map(d.p, dm.p, 777)
    ref = m / np.linalg.norm(m.astype(np.float64), axis=0) * mp
    assert np.allclose(d.get().reshape(3, 777), ref, atol=1e-6)


def test_deltaM_vs_oracle(lib, orc):
    rng = np.random.default_rng(3)
    N = 5000
    m = orc.normalize(rng.standard_normal((3, N))).astype(F32)
    h = rng.standard_normal((3, N)).astype(F32)
    a = 0.02
    dm, dh = Dev(lib, m), Dev(lib, h)
    lib.gpu_deltaM(dm.p, dh.p, F32(a), None, F32(1 / (1 + a * a)), N)
    ref = orc.deltaM(m.astype(np.float64), h.astype(np.float64), a, 1 / (1 + a * a))
    assert np.abs(dh.get().reshape(3, N) - ref).max() < 2e-6 * np.abs(ref).max() + 1e-7
    # alpha map
    amap = rng.random(N).astype(F32)
    dh = Dev(lib, h)
    da = Dev(lib, amap)
    lib.gpu_deltaM(dm.p, dh.p, F32(0.5), da.p, F32(1.0), N)
    mm = m.astype(np.float64)
    hh = h.astype(np.float64)
    mxh = np.cross(mm.T, hh.T).T
    ref = mxh - 0.5 * amap * np.cross(mm.T, mxh.T).T
    assert np.abs(dh.get().reshape(3, N) - ref).max() < 1e-5


def test_local_fields_vs_oracle(lib, orc):
    rng = np.random.default_rng(4)
    N = 4000
    m = orc.normalize(rng.standard_normal((3, N))).astype(F32)
    h = rng.standard_normal((3, N)).astype(F32)
    hext = np.array([0.1, -0.2, 0.05], F32)
    u = np.array([0.6, 0.0, 0.8], F32)
    k = F32(0.7958)
    for anis in (0, 1):
        dm, dh = Dev(lib, m), Dev(lib, h)
        kk = np.array([k, 0], F32)
        lib.gpu_add_local_fields(dm.p, dh.p, N, hext.ctypes.data, None, anis, kk.ctypes.data, u.ctypes.data)
        ref = orc.local_fields(m.astype(np.float64), h.astype(np.float64), hext, anis, float(k), u)
        assert np.abs(dh.get().reshape(3, N) - ref).max() < 1e-6
    # H and phi (gpu_local_contr3.cu:238-256)
    dm, dh = Dev(lib, m), Dev(lib, h)
    phi0 = rng.standard_normal(N).astype(F32)
    dp = Dev(lib, phi0)
    kk = np.array([k, 0], F32)
    lib.gpu_add_local_fields_H_and_phi(dm.p, dh.p, dp.p, N, hext.ctypes.data, None, 1, kk.ctypes.data, u.ctypes.data)
    mu = (m * u[:, None]).sum(0)
    ref_phi = phi0 - ((m * hext[:, None]).sum(0) + 0.5 * k * mu * mu)
    assert np.abs(dp.get() - ref_phi).max() < 2e-6


def test_exchange_stencil_vs_oracle(lib, orc):
    rng = np.random.default_rng(5)
    for size in [(3, 20, 70), (1, 37, 130), (5, 9, 64)]:
        for periodic in [(0, 0, 0), (0, 1, 1)]:
            N = int(np.prod(size))
            m = rng.standard_normal((3,) + size).astype(F32)
            h = rng.standard_normal((3,) + size).astype(F32)
            cell = np.array([0.8, 1.1, 1.3], F32)
            dm, dh = Dev(lib, m), Dev(lib, h)
            lib.gpu_add_exch(dm.p, dh.p, i3(*size), i3(*periodic), i3(0, 0, 0), cell.ctypes.data, 6)
            ref = orc.exch6_stencil(m.astype(np.float64), h.astype(np.float64), size, cell.astype(np.float64),
                                    periodic)
            assert np.abs(dh.get().reshape(ref.shape) - ref).max() < 1e-5 * np.abs(ref).max(), (size, periodic)
            # exchInConv on all components: no-op
            dh2 = Dev(lib, h)
            lib.gpu_add_exch(dm.p, dh2.p, i3(*size), i3(*periodic), i3(1, 1, 1), cell.ctypes.data, 6)
            assert np.array_equal(dh2.get(), h.reshape(-1))
            _ = N


def test_kernelmul(lib):
    rng = np.random.default_rng(6)
    n = 1000
    M = (rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))).astype(np.complex64)
    K = rng.standard_normal((6, n)).astype(F32)
    d = [Dev(lib, M[i].view(F32)) for i in range(3)]
    dk = [Dev(lib, K[i]) for i in range(6)]
    lib.gpu_kernelmul6(*[x.p for x in d], *[x.p for x in dk], 2 * n)
    xx, yy, zz, yz, xz, xy = K
    ref = [xx * M[0] + xy * M[1] + xz * M[2], xy * M[0] + yy * M[1] + yz * M[2], xz * M[0] + yz * M[1] + zz * M[2]]
    for i in range(3):
        assert np.allclose(d[i].get().view(np.complex64), ref[i], atol=1e-5)
    d = [Dev(lib, M[i].view(F32)) for i in range(3)]
    lib.gpu_kernelmul4(*[x.p for x in d], *[x.p for x in dk[:4]], 2 * n)
    ref = [xx * M[0], yy * M[1] + yz * M[2], yz * M[1] + zz * M[2]]
    for i in range(3):
        assert np.allclose(d[i].get().view(np.complex64), ref[i], atol=1e-5)


def test_zeropad_exact(lib):
    # zeropad_test.cpp:20-59: 129x255 -> 258x510, exact
    rng = np.random.default_rng(7)
    S = (2, 129, 255)
    D = (3, 258, 510)
    src = rng.standard_normal(S).astype(F32)
    ds, dd = Dev(lib, src), Dev(lib, n=int(np.prod(D)))
    lib.gpu_copy_pad(ds.p, dd.p, *S, *D)
    ref = np.zeros(D, F32)
    ref[: S[0], : S[1], : S[2]] = src
    assert np.array_equal(dd.get().reshape(D), ref)
    back = Dev(lib, n=int(np.prod(S)))
    lib.gpu_copy_unpad(dd.p, back.p, *D, *S)
    assert np.array_equal(back.get().reshape(S), src)


# ------------------------------------------------------------------ FFT plan (gpu_fft0.cu, fft_test.go)
@pytest.mark.parametrize("size", [
    (1, 16, 16), (1, 32, 128), (4, 16, 32), (2, 8, 4), (16, 64, 64),   # powers of two
    (1, 96, 96), (3, 5, 12), (1, 160, 224), (2, 48, 80), (5, 7, 9),     # 2^a 3^b 5^c 7^d (fft_test.go's 32 i)
    (1, 11, 13), (2, 37, 22), (3, 17, 40), (1, 101, 64),                 # prime factors: Bluestein
])
def test_fft_plan_padded(lib, size):
    """The drop-in plan against numpy at the Go driver's padded sizes 2N
    (kernel.go:80-90), any N."""
    P = tuple(2 * s if s > 1 else 1 for s in size)
    rng = np.random.default_rng(8)
    x = rng.random(size).astype(F32)
    plan = lib.new_gpuFFT3dPlan_padded(i3(*size), i3(*P))
    Q = P[2] // 2 + 1
    din = Dev(lib, x)
    dout = Dev(lib, n=P[0] * P[1] * (P[2] + 2))
    lib.gpuFFT3dPlan_forward(plan, din.p, dout.p)
    got = dout.get().view(np.complex64).reshape(P[0], P[1], Q)
    xp = np.zeros(P)
    xp[: size[0], : size[1], : size[2]] = x
    ref = np.fft.rfftn(xp)
    assert np.abs(got - ref).max() < 2e-6 * np.abs(ref).max()
    back = Dev(lib, n=x.size)
    lib.gpuFFT3dPlan_inverse(plan, dout.p, back.p)
    norm = lib.gpuFFT3dPlan_normalization(plan)
    assert norm == np.prod(P)
    assert np.abs(back.get().reshape(size) / norm - x).max() < 1e-4  # fft_test.go tolerance
    lib.delete_gpuFFT3dPlan(plan)


@pytest.mark.parametrize("size,P", [((2, 9, 7), (2, 9, 7)), ((1, 6, 5), (1, 12, 5)), ((3, 4, 13), (6, 8, 13))])
def test_fft_plan_periodic_odd_sizes(lib, size, P):
    """Periodic axes keep their size (padSize, kernel.go:83-87), so the real
    transform can have odd length."""
    rng = np.random.default_rng(10)
    x = rng.random(size).astype(F32)
    plan = lib.new_gpuFFT3dPlan_padded(i3(*size), i3(*P))
    Q = P[2] // 2 + 1
    din = Dev(lib, x)
    dout = Dev(lib, n=P[0] * P[1] * 2 * Q)
    lib.gpuFFT3dPlan_forward(plan, din.p, dout.p)
    got = dout.get().view(np.complex64).reshape(P[0], P[1], Q)
    xp = np.zeros(P)
    xp[: size[0], : size[1], : size[2]] = x
    ref = np.fft.rfftn(xp)
    assert np.abs(got - ref).max() < 2e-6 * np.abs(ref).max()
    back = Dev(lib, n=x.size)
    lib.gpuFFT3dPlan_inverse(plan, dout.p, back.p)
    assert np.abs(back.get().reshape(size) / np.prod(P) - x).max() < 1e-4
    lib.delete_gpuFFT3dPlan(plan)


# ------------------------------------------------------------------ semi-analytical step, STT, noise
def _rk_macrospin(m, h, alpha, dt, n=4000, precession=+1):
    def f(m):
        mxh = np.cross(m.T, h.T).T
        return (precession * mxh - alpha * np.cross(m.T, mxh.T).T) / (1 + alpha * alpha)

    d = dt / n
    for _ in range(n):
        k1 = f(m)
        k2 = f(m + d / 2 * k1)
        k3 = f(m + d / 2 * k2)
        k4 = f(m + d * k3)
        m = m + d / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return m


def test_anal_fw_step_is_exact_constant_field_solution(lib, orc):
    rng = np.random.default_rng(9)
    N = 64
    m = orc.normalize(rng.standard_normal((3, N)))
    h = rng.standard_normal((3, N)) * 0.5
    h[:, 0] = [0.0, 0.0, 0.7]  # field along the third axis only (special frame)
    alpha, dt = 0.1, 0.8
    dm, dh, do = Dev(lib, m), Dev(lib, h), Dev(lib, n=3 * N)
    lib.gpu_anal_fw_step(F32(dt), F32(alpha), N, dm.p, do.p, dh.p)
    # gpu_anal.cu:63-81 rotates (e2,e3) by +|h| qt in the right-handed frame (h, e2, e3):
    # dm/dt = -m x h - alpha m x (m x h), the opposite precession sense to gpu_deltaM's
    # internal-coordinate torque (gpu_torque.cu:40-52)
    ref = _rk_macrospin(m, h, alpha, dt, precession=-1)
    assert np.abs(do.get().reshape(3, N) - ref).max() < 2e-5


def test_spintorque_reduces_to_deltaM_without_current(lib, orc):
    rng = np.random.default_rng(10)
    size = (2, 8, 16)
    N = int(np.prod(size))
    m = orc.normalize(rng.standard_normal((3, N))).astype(F32)
    h = rng.standard_normal((3, N)).astype(F32)
    dm, dh = Dev(lib, m), Dev(lib, h)
    u = np.zeros(3, F32)
    lib.gpu_spintorque_deltaM(dm.p, dh.p, F32(0.3), F32(0.1), F32(0.2), u.ctypes.data, None, F32(0.9), *size)
    ref = 0.9 * orc.deltaM(m.astype(np.float64), h.astype(np.float64), 0.3, 1.0)
    assert np.abs(dh.get().reshape(3, N) - ref).max() < 1e-5


def test_spintorque_directional_derivative(lib):
    rng = np.random.default_rng(11)
    size = (3, 6, 10)
    N = int(np.prod(size))
    m = rng.standard_normal((3,) + size).astype(F32)
    h = np.zeros((3,) + size, F32)
    u = np.array([0.3, -0.2, 0.5], F32)
    eps, beta = 1.0, 0.0
    dm, dh = Dev(lib, m), Dev(lib, h)
    lib.gpu_spintorque_deltaM(dm.p, dh.p, F32(0.0), F32(beta), F32(eps), u.ctypes.data, None, F32(1.0), *size)
    # Hspin = sum_d u_d (m(+1) - m(-1)), the edge cell standing in for a missing neighbour
    mm = m.astype(np.float64)
    hs = np.zeros_like(mm)
    for ax in range(3):
        idx = np.arange(size[ax])
        ip = np.minimum(idx + 1, size[ax] - 1)
        im = np.maximum(idx - 1, 0)
        hs += u[ax] * (np.take(mm, ip, axis=ax + 1) - np.take(mm, im, axis=ax + 1))
    ref = np.cross(mm.reshape(3, N).T, (eps * hs).reshape(3, N).T).T
    assert np.abs(dh.get().reshape(3, N) - ref).max() < 1e-5


def test_gaussian_noise_statistics(lib):
    n = 1 << 20
    d = Dev(lib, n=n)
    lib.gpu_gaussian_noise(d.p, F32(0.5), F32(2.0), n)
    x = d.get().astype(np.float64)
    assert abs(x.mean() - 0.5) < 0.01
    assert abs(x.std() - 2.0) < 0.01
    lib.gpu_gaussian_noise(d.p, F32(0.0), F32(1.0), n)
    y = d.get()
    assert not np.array_equal(x, y)


def test_cubic_anisotropy_is_minus_gradient_of_its_energy(lib):
    """anisType 2 (gpu_local_contr3.cu's cubic branch): the field equals minus
    the gradient, by central differences, of
    E = k0/2 (a0^2 a1^2 + a1^2 a2^2 + a0^2 a2^2) + k1/2 a0^2 a1^2 a2^2 with
    a_i = m . u_i; with k1 = 0 the energy density phi is E (-m . H / 4 = E for a
    quartic)."""
    rng = np.random.default_rng(21)
    N = 300
    m = rng.standard_normal((3, N)).astype(F32)
    m /= np.linalg.norm(m, axis=0)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    u = np.ascontiguousarray(q.T, F32)            # rows u_0, u_1, u_2, orthonormal
    k0, k1 = 0.7, 0.4
    hext = np.zeros(3, F32)

    def energy(mm):
        a = u.astype(np.float64) @ mm
        return 0.5 * k0 * (a[0] ** 2 * a[1] ** 2 + a[1] ** 2 * a[2] ** 2 + a[0] ** 2 * a[2] ** 2) + \
            0.5 * k1 * a[0] ** 2 * a[1] ** 2 * a[2] ** 2

    m64, eps = m.astype(np.float64), 1e-6
    grad = np.zeros((3, N))
    for c in range(3):
        d = np.zeros((3, 1))
        d[c] = eps
        grad[c] = (energy(m64 + d) - energy(m64 - d)) / (2 * eps)
    dm, dh = Dev(lib, m), Dev(lib, np.zeros_like(m))
    kk = np.array([k0, k1], F32)
    lib.gpu_add_local_fields(dm.p, dh.p, N, hext.ctypes.data, None, 2, kk.ctypes.data, u.reshape(-1).ctypes.data)
    h = dh.get().reshape(3, N)
    assert np.abs(h + grad).max() < 1e-5
    kk0 = np.array([k0, 0.0], F32)
    dm, dh, dp = Dev(lib, m), Dev(lib, np.zeros_like(m)), Dev(lib, np.zeros(N, F32))
    lib.gpu_add_local_fields_H_and_phi(dm.p, dh.p, dp.p, N, hext.ctypes.data, None, 2, kk0.ctypes.data,
                                       u.reshape(-1).ctypes.data)
    a = u.astype(np.float64) @ m64
    e4 = 0.5 * k0 * (a[0] ** 2 * a[1] ** 2 + a[1] ** 2 * a[2] ** 2 + a[0] ** 2 * a[2] ** 2)
    assert np.abs(dp.get() - e4).max() < 1e-6
