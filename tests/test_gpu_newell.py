"""The Newell demag kernel (an option beside the reference's FaceKernel6) and
the shared kernel spectra.

The Newell tensor is not in the reference; it is pinned here by mathematics:
  * the self term equals Aharoni's closed form for a rectangular prism
    (J. Appl. Phys. 83, 3432 (1998)), for several aspect ratios;
  * outside the source cell the kernel is traceless (the field of a uniformly
    magnetised cell is divergence- and curl-free there) and the self term has
    trace -1;
  * at near and far displacements it equals the cell average of the point-dipole
    field over both cells, integrated here independently by high-order
    Gauss-Legendre quadrature in numpy;
  * the average demag field of a uniformly magnetised cuboid of cells equals
    minus the cuboid's own Aharoni factor (cell averages add up exactly), run
    through the whole FP32 engine.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def aharoni(A, B, Cc):
    """Demag factor along the third side of an A x B x Cc prism (Aharoni 1998, eq. 1)."""
    a, b, c = A / 2.0, B / 2.0, Cc / 2.0
    r = np.sqrt(a * a + b * b + c * c)
    rab, rbc, rac = np.sqrt(a * a + b * b), np.sqrt(b * b + c * c), np.sqrt(a * a + c * c)
    L = np.log
    t = ((b * b - c * c) / (2 * b * c) * L((r - a) / (r + a)) + (a * a - c * c) / (2 * a * c) * L((r - b) / (r + b))
         + b / (2 * c) * L((rab + a) / (rab - a)) + a / (2 * c) * L((rab + b) / (rab - b))
         + c / (2 * a) * L((rbc - b) / (rbc + b)) + c / (2 * b) * L((rac - a) / (rac + a))
         + 2 * np.arctan(a * b / (c * r)) + (a ** 3 + b ** 3 - 2 * c ** 3) / (3 * a * b * c)
         + (a * a + b * b - 2 * c * c) / (3 * a * b * c) * r + c / (a * b) * (rac + rbc)
         - (rab ** 3 + rbc ** 3 + rac ** 3) / (3 * a * b * c))
    return t / np.pi


def aharoni_diag(d):
    return np.array([aharoni(d[1], d[2], d[0]), aharoni(d[2], d[0], d[1]), aharoni(d[0], d[1], d[2])])


def kernel(lib, P, cell, kind=1, acc=8):
    P = np.ascontiguousarray(P, np.int32)
    cell = np.ascontiguousarray(cell, np.float64)
    out = np.empty((6,) + tuple(P), np.float64)
    lib.mm_demag_kernel(P.ctypes.data, cell.ctypes.data, kind, acc, out.ctypes.data)
    return out   # xx yy zz yz xz xy


def box_average(R, d, n):
    """V E[D(R + a - b)] for a, b uniform in the cell (n-point Gauss-Legendre per
    coordinate of each cell), D the point-dipole field: the kernel K = -N."""
    x, w = np.polynomial.legendre.leggauss(n)
    x, w = x / 2, w / 2
    u = [((x[:, None] - x[None, :]) * d[a]).ravel() for a in range(3)]
    wu = (w[:, None] * w[None, :]).ravel()
    X = R[0] + u[0][:, None, None]
    Y = R[1] + u[1][None, :, None]
    Z = R[2] + u[2][None, None, :]
    W = wu[:, None, None] * wu[None, :, None] * wu[None, None, :]
    r2 = X * X + Y * Y + Z * Z
    ir3 = W / (r2 * np.sqrt(r2))
    ir5 = 3 * ir3 / r2
    V = d[0] * d[1] * d[2] / (4 * np.pi)
    comps = [ir5 * X * X - ir3, ir5 * Y * Y - ir3, ir5 * Z * Z - ir3, ir5 * Y * Z, ir5 * X * Z, ir5 * X * Y]
    return V * np.array([c.sum() for c in comps])


@pytest.mark.parametrize("cell", [(1.0, 1.0, 1.0), (3e-9, 3.9e-9, 4.2e-9), (1.0, 2.0, 5.0), (10.0, 1.0, 1.0)])
def test_self_term_is_aharoni(lib, cell):
    K = kernel(lib, (4, 4, 4), cell)
    got = -K[:3, 0, 0, 0]
    want = aharoni_diag(np.array(cell) / max(cell))
    print(cell, got, want)
    assert np.allclose(got, want, rtol=1e-11, atol=1e-13)
    assert np.abs(K[3:, 0, 0, 0]).max() < 1e-13


def test_trace_and_parity(lib):
    P, cell = (16, 64, 48), (3.0, 4.0, 4.5)
    K = kernel(lib, P, cell)
    tr = K[0] + K[1] + K[2]
    assert tr[0, 0, 0] == pytest.approx(-1.0, abs=1e-13)
    tr[0, 0, 0] = 0
    print("max |trace| off the centre:", np.abs(tr).max())
    assert np.abs(tr).max() < 1e-9
    # K(-R): diagonal even, yz odd in y and z, xz odd in x and z, xy odd in x and y
    for x, y, z in ((1, 2, 3), (5, 17, 9), (7, 30, 20)):
        a, b = K[:, x, y, z], K[:, -x, -y, -z]
        assert np.allclose(a[:3], b[:3], rtol=1e-12)
        assert np.allclose(a[3:], b[3:], rtol=1e-12)            # two sign flips each
        c = K[:, x, -y, z]
        assert np.allclose(c[3], -a[3], rtol=1e-12) and np.allclose(c[5], -a[5], rtol=1e-12)
        assert np.allclose(c[4], a[4], rtol=1e-12)


@pytest.mark.parametrize("disp", [(2, 1, 3), (0, 3, 1), (1, 6, 5), (4, 20, 11), (7, 31, 23)])
def test_matches_cell_average_of_dipole_field(lib, disp):
    """Near displacements (27-term Newell sums) and far ones (the kernel's own
    3-point quadrature) against 8-point quadrature of the same double average."""
    P, cell = (16, 64, 48), np.array([3.0, 4.0, 4.5])
    K = kernel(lib, P, cell)
    R = np.array(disp) * cell
    want = box_average(R, cell, 8)
    got = K[:, disp[0], disp[1], disp[2]]
    scale = np.abs(want).max()
    print(disp, np.abs(got - want).max() / scale)
    assert np.abs(got - want).max() <= 2e-8 * scale


def test_cuboid_average_field_is_its_demag_factor(mm, lib):
    """Uniform m in a cuboid of 12 x 20 x 36 cells with the Newell kernel and no
    exchange: the average demag field along each axis is minus the cuboid's
    demag factor (Aharoni), through the FP32 engine."""
    gk = mm.gpukern
    size, cell = (12, 20, 36), (2e-9, 3e-9, 2.5e-9)
    body = np.array(size) * np.array(cell)
    want = aharoni_diag(body / body.max())
    for c in range(3):
        p = gk.Params()
        p.size[:] = list(size)
        p.cellsize[:] = list(cell)
        p.msat, p.aexch, p.alpha = 8e5, 1.3e-11, 0.1
        p.exchType, p.exchInConv, p.wantDemag, p.demagAccuracy = 0, 0, 1, 8
        p.solver, p.maxError = 54, 1e-5
        p.demagKernel = 1
        m = np.zeros((3,) + size, np.float32)
        m[c] = 1.0
        s = lib.mm_sim_new(C.byref(p), m.ctypes.data, 0)
        h = np.empty_like(m)
        lib.mm_sim_heff(s, m.ctypes.data, h.ctypes.data)
        lib.mm_sim_free(s)
        avg = h.reshape(3, -1).mean(axis=1, dtype=np.float64)
        print(c, avg, -want[c])
        assert avg[c] == pytest.approx(-want[c], abs=5e-6)
        assert np.abs(np.delete(avg, c)).max() < 5e-6


def test_kernel_spectra_are_shared(mm, lib):
    """Simulations of one padded size, cell and kernel share one spectrum; it is
    freed with the last of them; a different kernel or cell is another."""
    def film(nx, kernel="facekernel6", cell=3.90625e-9):
        s = mm.Sim()
        s.msat(8e5), s.aexch(1.3e-11), s.alpha(0.1)
        s.gridsize(nx, 32, 1), s.cellsize(cell, 3.71e-9, 3e-9)   # a cell no other test uses
        s.kerneltype(kernel)
        s.uniform(1, 0, 0)
        s._init()
        return s

    n0 = lib.mm_kernel_cache_size()
    a, b = film(48), film(40)          # both pad to 96 x 64
    assert lib.mm_kernel_cache_size() == n0 + 1
    c = film(48, "newell")
    d = film(48, cell=4e-9)
    assert lib.mm_kernel_cache_size() == n0 + 3
    ha, hb = a.heff(a.m[:, :, :, :]), b.heff(b.m)
    assert np.isfinite(ha).all() and np.isfinite(hb).all()
    a.free()
    assert lib.mm_kernel_cache_size() == n0 + 3
    b.free()
    assert lib.mm_kernel_cache_size() == n0 + 2
    spec_c = c.kernel_spectrum()
    c.free(), d.free()
    assert lib.mm_kernel_cache_size() == n0
    assert np.isfinite(spec_c).all()


def test_newell_film_runs_close_to_facekernel6(mm):
    """A film relaxing under the two kernels: same dynamics to within the
    kernels' difference (target-cell average versus centre value)."""
    def run(kernel):
        s = mm.Sim()
        s.msat(8e5), s.aexch(1.3e-11), s.alpha(0.5)
        s.gridsize(128, 32, 1), s.cellsize(3.90625e-9, 3.90625e-9, 3e-9)
        s.solvertype("rkdp"), s.maxerror(1e-5)
        s.kerneltype(kernel)
        s.uniform(1, 0.25, 0.1)
        s.applystatic("field", -24.6e-3, 4.3e-3, 0)
        s.run(50e-12)
        m = s.m
        s.free()
        return m

    a, b = run("facekernel6"), run("newell")
    d = np.abs(a - b).max()
    print("max |m_face6 - m_newell| after 50 ps:", d)
    assert 0 < d < 0.05
