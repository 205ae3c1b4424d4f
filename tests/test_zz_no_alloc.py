"""gpuFFT3dPlan_forward_noalloc (include/gpukern.h) against gpuFFT3dPlan_forward.

The no-allocation variant runs the same kernels with their temporaries in the
library's fixed device scratch array, so its output must equal the allocating
forward transform's bit for bit.  Two calls in a row, the second on a smaller
plan, check that what the first left in the shared array does not leak into
the second.  The sizes make every scratch user run: Z rows of half-length 13
(packed rows plus a Bluestein buffer), Y lines of length 22 (Bluestein), Z
rows of half-length 7 (packed rows, radix-7 Stockham).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32


def i3(*v):
    return (C.c_int * 3)(*v)


def upload(lib, x):
    x = np.ascontiguousarray(x, F32).reshape(-1)
    p = lib.new_gpu_array(x.size)
    assert p
    lib.memcpy_to_gpu(x.ctypes.data, p, x.size)
    return p


def download(lib, p, n):
    out = np.empty(n, F32)
    lib.memcpy_from_gpu(p, out.ctypes.data, n)
    return out


def forward_both(lib, size, P, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(size).astype(F32)
    plan = lib.new_gpuFFT3dPlan_padded(i3(*size), i3(*P))
    n_out = P[0] * P[1] * 2 * (P[2] // 2 + 1)
    din = upload(lib, x)
    d_ref = lib.new_gpu_array(n_out)
    d_na = lib.new_gpu_array(n_out)
    lib.gpuFFT3dPlan_forward(plan, din, d_ref)
    lib.gpuFFT3dPlan_forward_noalloc(plan, din, d_na)
    ref = download(lib, d_ref, n_out)
    got = download(lib, d_na, n_out)
    for p in (din, d_ref, d_na):
        lib.free_gpu_array(p)
    lib.delete_gpuFFT3dPlan(plan)
    xp = np.zeros(P)
    xp[: size[0], : size[1], : size[2]] = x
    return ref, got, np.fft.rfftn(xp)


def test_noalloc_forward_equals_forward_twice(lib):
    print("no-alloc FFT: plan (3,11,13) padded (6,22,26), then (2,5,7) padded (4,10,14)", flush=True)
    for size, P, seed in [((3, 11, 13), (6, 22, 26), 1), ((2, 5, 7), (4, 10, 14), 2)]:
        ref, got, npref = forward_both(lib, size, P, seed)
        assert np.array_equal(ref.view(np.uint32), got.view(np.uint32)), size
        c = got.view(np.complex64).reshape(npref.shape)
        assert np.abs(c - npref).max() < 2e-6 * np.abs(npref).max(), size


def test_noalloc_forward_power_of_two_needs_no_scratch(lib):
    print("no-alloc FFT: plan (4,16,32) padded (8,32,64)", flush=True)
    ref, got, npref = forward_both(lib, (4, 16, 32), (8, 32, 64), 3)
    assert np.array_equal(ref.view(np.uint32), got.view(np.uint32))
