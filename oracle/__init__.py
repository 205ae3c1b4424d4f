"""CPU oracle for the MuMax 1.1 LLG hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg import it; the product library
(``libgpukern.so``) never loads it, and no product code path falls back to it.

It wraps ``liboracle.so`` (FP64) and ``liboracle_f32.so`` (FP32 twin), built
from ``oracle.c`` by ``oracle/Makefile``.  All arrays use the reference's
internal layout: planar ``[3][X][Y][Z]`` with Z fastest, X = film thickness.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

XX, YY, ZZ, YZ, XZ, XY = range(6)
REDUCE_ADD, REDUCE_MAX, REDUCE_MAXABS, REDUCE_MIN, REDUCE_SUMABS = 1, 2, 3, 4, 5
ANIS_NONE, ANIS_UNIAXIAL = 0, 1
SOLVERS = {"rk2": 2, "rk12": 12, "rk3": 3, "rk23": 23, "rk4": 4, "rkck": 45, "rkdp": 54, "rk45": 54,
           "semianal1": 101, "semianal2": 102}


def build():
    """Compile liboracle*.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class Params(C.Structure):
    _fields_ = [
        ("size", C.c_int * 3),
        ("periodic", C.c_int * 3),
        ("cellsize", C.c_double * 3),
        ("msat", C.c_double),
        ("aexch", C.c_double),
        ("alpha", C.c_double),
        ("anisType", C.c_int),
        ("k1", C.c_double),
        ("anisAxis", C.c_double * 3),
        ("exchType", C.c_int),
        ("exchInConv", C.c_int),
        ("wantDemag", C.c_int),
        ("demagAccuracy", C.c_int),
        ("solver", C.c_int),
        ("maxError", C.c_double),
        ("dt", C.c_double),
        ("minDm", C.c_double),
        ("maxDm", C.c_double),
        ("minDt", C.c_double),
        ("maxDt", C.c_double),
        ("nField", C.c_int),
        ("fieldT", C.POINTER(C.c_double)),
        ("fieldB", C.POINTER(C.c_double)),
        ("wantEnergy", C.c_int),
        ("temperature", C.c_double),
        ("seed", C.c_ulonglong),
    ]


class Oracle:
    """ctypes binding of one oracle precision (``f64`` or ``f32``)."""

    def __init__(self, precision="f64"):
        name = "liboracle.so" if precision == "f64" else "liboracle_f32.so"
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.real = np.float64 if self.lib.or_real_size() == 8 else np.float32
        self._bind()

    def _bind(self):
        L = self.lib
        P = C.c_void_p
        L.or_units.argtypes = [C.c_double, C.c_double, P]
        for f in ("or_facekernel6", "or_facekernel6_f32"):
            getattr(L, f).argtypes = [P, P, C.c_int, P, P]
        L.or_add_exch6_kernel.argtypes = [P, P, P]
        L.or_fft_c2c.argtypes = [P, C.c_int, C.c_int, C.c_int]
        L.or_fft3_c2c.argtypes = [P, P, C.c_int]
        L.or_conv_new.argtypes = [P, P, P]
        L.or_conv_new.restype = P
        L.or_conv_free.argtypes = [P]
        L.or_conv_apply.argtypes = [P, P, P]
        L.or_conv_kernel_spectrum.argtypes = [P, P]
        L.or_conv_forward1.argtypes = [P, P, P]
        L.or_conv_direct.argtypes = [P, P, P, P, P]
        L.or_local_fields.argtypes = [P, P, C.c_int, P, C.c_int, self._creal(), P]
        L.or_local_fields_phi.argtypes = [P, P, P, C.c_int, P, C.c_int, self._creal(), P]
        L.or_deltaM.argtypes = [P, P, self._creal(), self._creal(), C.c_int]
        L.or_normalize.argtypes = [P, C.c_int]
        L.or_reduce.argtypes = [C.c_int, P, C.c_int]
        L.or_reduce.restype = C.c_double
        L.or_add_exch6_stencil.argtypes = [P, P, P, P, P]
        L.or_sim_new.argtypes = [C.POINTER(Params), P]
        L.or_sim_new.restype = P
        L.or_sim_free.argtypes = [P]
        L.or_sim_step.argtypes = [P]
        L.or_sim_step_replay.argtypes = [P, P, P, C.c_int]
        L.or_sim_run.argtypes = [P, C.c_double]
        L.or_sim_get_m.argtypes = [P, P]
        L.or_sim_set_m.argtypes = [P, P]
        L.or_sim_state.argtypes = [P, P]
        L.or_sim_heff.argtypes = [P, P, P]
        L.or_sim_torque.argtypes = [P, P, P]
        L.or_sim_set_field.argtypes = [P, C.c_int, P, P]
        L.or_sim_set_want_energy.argtypes = [P, C.c_int]
        L.or_sim_set_alpha_map.argtypes = [P, P]
        L.or_sim_set_norm_map.argtypes = [P, P]
        L.or_sim_set_spintorque.argtypes = [P, P, C.c_double, C.c_double, P]
        L.or_sim_set_applied_current.argtypes = [P, C.c_int, P, P, C.c_int, P]
        L.or_anal_step.argtypes = [P, P, P, C.c_double, C.c_double, C.c_int]
        L.or_normalize_map.argtypes = [P, P, C.c_int]
        L.or_sim_set_field_function.argtypes = [P, C.c_int, P, P]
        L.or_sim_field_at.argtypes = [P, C.c_double, P]
        L.or_sim_trial_log.argtypes = [P, P, P, C.c_long]
        L.or_sim_trial_log.restype = C.c_long
        L.or_sim_energy.argtypes = [P]
        L.or_thermal_normals.argtypes = [C.c_ulonglong, C.c_longlong, C.c_long, P]
        L.or_sim_energy.restype = C.c_double

    def _creal(self):
        return C.c_double if self.lib.or_real_size() == 8 else C.c_float

    def _arr(self, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or self.real)
        return a, a.ctypes.data

    # ---- units / kernels ----
    def units(self, msat, aexch):
        out = np.zeros(5)
        self.lib.or_units(msat, aexch, out.ctypes.data)
        return dict(length=out[0], time=out[1], field=out[2], energy=out[3], edens=out[4])

    def facekernel6(self, P, cell, accuracy=8, periodic=(0, 0, 0)):
        P = np.asarray(P, np.int32)
        per = np.asarray(periodic, np.int32)
        cell, cp = self._arr(cell)
        K = np.zeros((6,) + tuple(P), self.real)
        self.lib.or_facekernel6(P.ctypes.data, cp, accuracy, per.ctypes.data, K.ctypes.data)
        return K

    def facekernel6_f32(self, P, cell, accuracy=8, periodic=(0, 0, 0)):
        P = np.asarray(P, np.int32)
        per = np.asarray(periodic, np.int32)
        cell = np.ascontiguousarray(cell, np.float32)
        K = np.zeros((6,) + tuple(P), np.float32)
        self.lib.or_facekernel6_f32(P.ctypes.data, cell.ctypes.data, accuracy, per.ctypes.data, K.ctypes.data)
        return K

    def add_exch6_kernel(self, K, cell):
        P = np.asarray(K.shape[1:], np.int32)
        cell, cp = self._arr(cell)
        assert K.dtype == self.real and K.flags.c_contiguous
        self.lib.or_add_exch6_kernel(P.ctypes.data, cp, K.ctypes.data)
        return K

    # ---- FFT ----
    def fft(self, x, sign=-1):
        """FFT along the last axis of a complex array."""
        x = np.ascontiguousarray(x, np.complex128 if self.real == np.float64 else np.complex64)
        n = x.shape[-1]
        self.lib.or_fft_c2c(x.ctypes.data, n, x.size // n, sign)
        return x

    def fft3(self, x, sign=-1):
        x = np.ascontiguousarray(x, np.complex128 if self.real == np.float64 else np.complex64)
        n = np.asarray(x.shape, np.int32)
        self.lib.or_fft3_c2c(x.ctypes.data, n.ctypes.data, sign)
        return x

    # ---- convolution ----
    def conv_new(self, N, P, K):
        N = np.asarray(N, np.int32)
        P = np.asarray(P, np.int32)
        K, kp = self._arr(K)
        return Conv(self, self.lib.or_conv_new(N.ctypes.data, P.ctypes.data, kp), N, P)

    def conv_direct(self, N, P, K, m):
        N = np.asarray(N, np.int32)
        P = np.asarray(P, np.int32)
        K, kp = self._arr(K)
        m, mp = self._arr(m)
        h = np.zeros_like(m)
        self.lib.or_conv_direct(N.ctypes.data, P.ctypes.data, kp, mp, h.ctypes.data)
        return h

    # ---- pointwise ----
    def local_fields(self, m, h, hext, anis_type=0, k=0.0, axis=(0, 0, 0)):
        m, mp = self._arr(m)
        h = np.ascontiguousarray(h, self.real).copy()
        hx, hp = self._arr(hext)
        ax, ap = self._arr(axis)
        self.lib.or_local_fields(mp, h.ctypes.data, m.size // 3, hp, anis_type, k, ap)
        return h

    def deltaM(self, m, h, alpha, dtgilb):
        m, mp = self._arr(m)
        h = np.ascontiguousarray(h, self.real).copy()
        self.lib.or_deltaM(mp, h.ctypes.data, alpha, dtgilb, m.size // 3)
        return h

    def normalize(self, m):
        m = np.ascontiguousarray(m, self.real).copy()
        self.lib.or_normalize(m.ctypes.data, m.size // 3)
        return m

    def anal_step(self, m, h, dt, alpha):
        m = np.ascontiguousarray(m, self.real).copy()
        h = np.ascontiguousarray(h, self.real)
        self.lib.or_anal_step(m.ctypes.data, m.ctypes.data, h.ctypes.data, dt, alpha, m.size // 3)
        return m

    def normalize_map(self, m, nmap):
        m = np.ascontiguousarray(m, self.real).copy()
        nmap = np.ascontiguousarray(nmap, self.real).reshape(-1)
        assert nmap.size == m.size // 3
        self.lib.or_normalize_map(m.ctypes.data, nmap.ctypes.data, m.size // 3)
        return m

    def reduce(self, op, x):
        x, xp = self._arr(x)
        return self.lib.or_reduce(op, xp, x.size)

    def exch6_stencil(self, m, h, size, cell, periodic=(0, 0, 0)):
        m, mp = self._arr(m)
        h = np.ascontiguousarray(h, self.real).copy()
        s = np.asarray(size, np.int32)
        per = np.asarray(periodic, np.int32)
        cell, cp = self._arr(cell)
        self.lib.or_add_exch6_stencil(mp, h.ctypes.data, s.ctypes.data, per.ctypes.data, cp)
        return h

    # ---- simulation ----
    def sim(self, m0, **kw):
        return Sim(self, m0, **kw)

    def thermal_normals(self, seed, step, cells):
        """The thermal field's standard normals (3 per cell) of `cells` at `step`."""
        out = np.empty((len(cells), 3))
        buf = np.empty(3)
        for j, i in enumerate(cells):
            self.lib.or_thermal_normals(int(seed), int(step), int(i), buf.ctypes.data)
            out[j] = buf
        return out


class Conv:
    def __init__(self, orc, ptr, N, P):
        self.orc, self.ptr, self.N, self.P = orc, ptr, N, P

    def __del__(self):
        if getattr(self, "ptr", None):
            self.orc.lib.or_conv_free(self.ptr)
            self.ptr = None

    def apply(self, m):
        m, mp = self.orc._arr(m)
        h = np.zeros_like(m)
        self.orc.lib.or_conv_apply(self.ptr, mp, h.ctypes.data)
        return h

    def kernel_spectrum(self):
        P = self.P
        out = np.zeros((6, P[0], P[1], P[2] // 2 + 1), self.orc.real)
        self.orc.lib.or_conv_kernel_spectrum(self.ptr, out.ctypes.data)
        return out

    def forward1(self, x):
        x, xp = self.orc._arr(x)
        P = self.P
        out = np.zeros((P[0], P[1], P[2] // 2 + 1, 2), self.orc.real)
        self.orc.lib.or_conv_forward1(self.ptr, xp, out.ctypes.data)
        return out[..., 0] + 1j * out[..., 1]


class Sim:
    """One oracle simulation.  Keyword arguments mirror the reference's input
    commands in internal axis order (see ``Params``)."""

    def __init__(self, orc, m0, size, cellsize, msat, aexch, alpha, solver="rkdp", maxerror=1e-4, dt=0.0,
                 k1=0.0, anis_axis=None, field=None, periodic=(0, 0, 0), exch_type=6, exch_in_conv=True,
                 want_demag=True, min_dm=0.0, max_dm=0.0, min_dt=0.0, max_dt=0.0, want_energy=False,
                 demag_accuracy=8, temperature=0.0, seed=0, norm_map=None):
        self.orc = orc
        p = Params()
        p.size[:] = list(size)
        p.periodic[:] = list(periodic)
        p.cellsize[:] = list(cellsize)
        p.msat, p.aexch, p.alpha = msat, aexch, alpha
        if anis_axis is not None:
            p.anisType = ANIS_UNIAXIAL
            ax = np.asarray(anis_axis, float)
            p.anisAxis[:] = list(ax / np.linalg.norm(ax))
        p.k1 = k1
        p.exchType = exch_type
        p.exchInConv = int(exch_in_conv)
        p.wantDemag = int(want_demag)
        p.demagAccuracy = demag_accuracy
        p.solver = SOLVERS[solver]
        p.maxError, p.dt = maxerror, dt
        p.minDm, p.maxDm, p.minDt, p.maxDt = min_dm, max_dm, min_dt, max_dt
        self._fieldT = self._fieldB = None
        if field is not None:
            t, b = field
            self._fieldT = np.ascontiguousarray(t, np.float64)
            self._fieldB = np.ascontiguousarray(b, np.float64).reshape(-1, 3)
            p.nField = len(self._fieldT)
            p.fieldT = self._fieldT.ctypes.data_as(C.POINTER(C.c_double))
            p.fieldB = self._fieldB.ctypes.data_as(C.POINTER(C.c_double))
        p.wantEnergy = int(want_energy)
        p.temperature = float(temperature)
        p.seed = int(seed) & ((1 << 64) - 1)
        self.params = p
        self.N = int(np.prod(size))
        m0 = np.ascontiguousarray(m0, np.float32).reshape(-1)
        assert m0.size == 3 * self.N
        self.ptr = orc.lib.or_sim_new(C.byref(p), m0.ctypes.data)
        if norm_map is not None:
            self.set_norm_map(norm_map)

    def __del__(self):
        if getattr(self, "ptr", None):
            self.orc.lib.or_sim_free(self.ptr)
            self.ptr = None

    def step(self):
        return self.orc.lib.or_sim_step(self.ptr)

    def step_replay(self, dts, accepts):
        dts = np.ascontiguousarray(dts, np.float64)
        acc = np.ascontiguousarray(accepts, np.int32)
        return self.orc.lib.or_sim_step_replay(self.ptr, dts.ctypes.data, acc.ctypes.data, len(dts))

    def run(self, seconds):
        return self.orc.lib.or_sim_run(self.ptr, seconds)

    @property
    def m(self):
        out = np.zeros(3 * self.N)
        self.orc.lib.or_sim_get_m(self.ptr, out.ctypes.data)
        return out

    @m.setter
    def m(self, v):
        v = np.ascontiguousarray(v, np.float64).reshape(-1)
        self.orc.lib.or_sim_set_m(self.ptr, v.ctypes.data)

    def state(self):
        st = np.zeros(9)
        self.orc.lib.or_sim_state(self.ptr, st.ctypes.data)
        keys = ("time", "dt", "steps", "error", "max_torque", "trials", "rejects", "energy", "unit_time")
        return dict(zip(keys, st))

    def heff(self, m):
        m = np.ascontiguousarray(m, np.float64).reshape(-1)
        h = np.zeros_like(m)
        self.orc.lib.or_sim_heff(self.ptr, m.ctypes.data, h.ctypes.data)
        return h

    def torque(self, m):
        m = np.ascontiguousarray(m, np.float64).reshape(-1)
        k = np.zeros_like(m)
        self.orc.lib.or_sim_torque(self.ptr, m.ctypes.data, k.ctypes.data)
        return k

    def energy(self):
        return self.orc.lib.or_sim_energy(self.ptr)

    def trial_log(self):
        """(dt used, accepted) of every trial so far, like the GPU engine's log"""
        n = self.orc.lib.or_sim_trial_log(self.ptr, None, None, 0)
        dt = np.zeros(max(n, 1))
        acc = np.zeros(max(n, 1), np.int32)
        self.orc.lib.or_sim_trial_log(self.ptr, dt.ctypes.data, acc.ctypes.data, n)
        return dt[:n], acc[:n]

    def set_field_function(self, kind, b, params):
        """applied field as a function of time: kind 1 pulse, 2 RF, 3 RF ramp,
        4 sawtooth, 5 rotating; b internal (X, Y, Z) tesla; params[4]"""
        b = np.ascontiguousarray(b, np.float64)
        p = np.zeros(4)
        p[: len(params)] = params
        self.orc.lib.or_sim_set_field_function(self.ptr, int(kind), b.ctypes.data, p.ctypes.data)

    def field_at(self, t):
        out = np.zeros(3)
        self.orc.lib.or_sim_field_at(self.ptr, float(t), out.ctypes.data)
        return out

    def set_alpha_map(self, amap):
        """per-cell damping multiplier (internal [X][Y][Z] order), None: none"""
        self._amap = None if amap is None else np.ascontiguousarray(amap, np.float32).reshape(-1)
        self.orc.lib.or_sim_set_alpha_map(self.ptr, None if self._amap is None else self._amap.ctypes.data)

    def set_spintorque(self, j, spin_pol, xi, jmask=None):
        """spin-transfer torque: current density j (A/m^2, internal order), spin
        polarisation, non-adiabaticity xi, mask (3, X, Y, Z) or None (1)"""
        jj = np.ascontiguousarray(j, np.float64)
        self._jmask = None if jmask is None else np.ascontiguousarray(jmask, np.float32).reshape(-1)
        self.orc.lib.or_sim_set_spintorque(self.ptr, jj.ctypes.data, float(spin_pol), float(xi),
                                           None if self._jmask is None else self._jmask.ctypes.data)

    def set_applied_current(self, t=None, j=None, kind=0, params=(0.0, 0.0, 0.0, 0.0)):
        """an applied current density: points t (s) and j (n x 3 A/m^2, internal
        order; one point: static; none: off) or, kind > 0, a function of time
        with amplitude j (3,)"""
        self._jT = None if t is None else np.ascontiguousarray(t, np.float64)
        self._jV = None if j is None else np.ascontiguousarray(j, np.float64).reshape(-1)
        self._jP = np.ascontiguousarray(params, np.float64)
        n = 0 if self._jT is None else len(self._jT)
        self.orc.lib.or_sim_set_applied_current(self.ptr, n, None if self._jT is None else self._jT.ctypes.data,
                                                None if self._jV is None else self._jV.ctypes.data, int(kind),
                                                self._jP.ctypes.data)

    def set_norm_map(self, nmap):
        """per-cell norm of m after each normalisation (internal [X][Y][Z] order), None: 1"""
        self._nmap = None if nmap is None else np.ascontiguousarray(nmap, np.float32).reshape(-1)
        self.orc.lib.or_sim_set_norm_map(self.ptr, None if self._nmap is None else self._nmap.ctypes.data)

    def set_field(self, t, b):
        """replace the applied field table (t: n times in s, b: n x 3 tesla)"""
        self._fieldT = np.ascontiguousarray(t, np.float64)
        self._fieldB = np.ascontiguousarray(b, np.float64).reshape(-1, 3)
        self.orc.lib.or_sim_set_field(self.ptr, len(self._fieldT), self._fieldT.ctypes.data,
                                      self._fieldB.ctypes.data)

    def relax(self, on_step=None):
        """Relax() (relax.go:14-25): with energy control on, step until a step
        raises the energy; returns the number of steps.  on_step() is called
        after every step."""
        self.orc.lib.or_sim_set_want_energy(self.ptr, 1)
        self.step()
        n = 1
        if on_step:
            on_step()
        prev = self.state()["energy"] + 1
        while self.state()["energy"] <= prev:
            prev = self.state()["energy"]
            self.step()
            n += 1
            if on_step:
                on_step()
        return n
