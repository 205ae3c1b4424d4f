"""Python mirror of the reference's simulation commands over libgpukern.

``Sim`` keeps the reference's command names and argument order (sim*.go,
lib/mumax.py): vector arguments are given in the user's (z, y, x) order and
stored in internal (X, Y, Z) order, exactly as ``GridSize(z, y, x)``,
``CellSize``, ``ApplyStatic(what, hz, hy, hx)`` and ``AnisUniaxial(uz, uy, ux)``
do (simgeom.go:25-55, simfield.go:87-153, simanis.go:32-45).  Magnetisation
arrays are in the internal planar layout ``[3][X][Y][Z]``.

The commands build an ``mmParams`` and the first ``run``/``steps`` creates the
device-resident engine (``mm_sim_new``), like ``Sim.init()`` (init.go:36-77).
"""
import ctypes as C

import numpy as np

from . import gpukern

MODES = {"fused": 0, "abi": 1}
SOLVERS = {"rk2": 2, "rk12": 12, "rk3": 3, "rk23": 23, "rk4": 4, "rkck": 45, "rkdp": 54, "rk45": 54}


class InputErr(ValueError):
    """Invalid command argument (the reference panics with InputErr)."""


class Sim:
    def __init__(self, mode="fused"):
        if mode not in MODES:
            raise InputErr("mode must be one of %s" % list(MODES))
        self.mode = mode
        self._size = None
        self._cell = None
        self._msat = 0.0
        self._aexch = 0.0
        self._alpha = 0.0
        self._k1 = 0.0
        self._axis = None
        self._field = []  # list of (t, [bX, bY, bZ]); static: single entry with t=None
        self._static = None
        self._solver = "rk12"       # DEFAULT_SOLVERTYPE (sim.go:24)
        self._maxerror = 1e-4       # DEFAULT_MAXERROR (sim.go:25)
        self._dt = 0.0
        self._mindm = self._maxdm = self._mindt = self._maxdt = 0.0
        self._exchinconv = True
        self._demag_accuracy = 8
        self._m0 = None
        self._ptr = None
        self._lib = gpukern.load()
        self._keep = []

    # ------------------------------------------------------------ material (simmaterial)
    def msat(self, v):
        self._invalidate()
        self._msat = float(v)

    def aexch(self, v):
        self._invalidate()
        self._aexch = float(v)

    def alpha(self, v):
        self._invalidate()
        self._alpha = float(v)

    def k1(self, v):
        self._invalidate()
        self._k1 = float(v)

    def anisuniaxial(self, uz, uy, ux):
        self._invalidate()
        n = np.sqrt(ux * ux + uy * uy + uz * uz)
        if n == 0:
            raise InputErr("Anisotropy axis should not be 0")
        self._axis = (ux / n, uy / n, uz / n)

    # ------------------------------------------------------------ geometry (simgeom.go)
    def gridsize(self, z, y, x):
        self._invalidate()
        if x <= 0 or y <= 0 or z <= 0:
            raise InputErr("Size should be > 0")
        if z == 1:
            raise InputErr("For a 2D geometry, use (X, Y, 1) cells, not (1, X, Y)")
        self._size = (int(x), int(y), int(z))

    def cellsize(self, z, y, x):
        self._invalidate()
        if x <= 0 or y <= 0 or z <= 0:
            raise InputErr("Cell size should be > 0")
        self._cell = (float(x), float(y), float(z))

    def partsize(self, z, y, x):
        if self._size is None:
            raise InputErr("gridsize must be set before partsize")
        X, Y, Z = self._size
        self.cellsize(z / Z, y / Y, x / X)

    # ------------------------------------------------------------ fields (simfield.go)
    # Applying a field does not invalidate a running simulation
    # (simfield.go:87-110): the engine's field table is replaced in place.
    def applystatic(self, what, hz, hy, hx):
        if what.lower() not in ("field", "b"):
            raise InputErr("only 'field' can be applied")
        self._static = (hx, hy, hz)
        self._field = []
        self._push_field()

    def applypointwise(self, what, t, hz, hy, hx):
        if what.lower() not in ("field", "b"):
            raise InputErr("only 'field' can be applied")
        if self._static is not None:
            self._field = []
        self._static = None
        self._field.append((float(t), (hx, hy, hz)))
        self._push_field()

    def _field_table(self):
        if self._static is not None:
            return np.zeros(1), np.ascontiguousarray(self._static, np.float64)
        if self._field:
            t = np.ascontiguousarray([f[0] for f in self._field], np.float64)
            b = np.ascontiguousarray(np.asarray([f[1] for f in self._field], np.float64).reshape(-1))
            return t, b
        return None, None

    def _push_field(self):
        # a pointwise field needs two points before it can be evaluated
        # (simfield.go:123-125); the engine keeps the previous field until then
        if self._ptr and (self._static is not None or len(self._field) != 1):
            t, b = self._field_table()
            n = 0 if t is None else len(t)
            self._lib.mm_sim_set_field(self._ptr, n, None if t is None else t.ctypes.data,
                                       None if b is None else b.ctypes.data)

    # ------------------------------------------------------------ solver (simsolver.go)
    def solvertype(self, s):
        self._invalidate()
        s = s.lower()
        if s not in SOLVERS:
            raise InputErr("Unknown solver type: %s" % s)
        self._solver = s

    def maxerror(self, e):
        self._maxerror = float(e)

    def dt(self, t):
        self._invalidate()
        self._dt = float(t)

    def mindm(self, v):
        self._mindm = float(v)

    def maxdm(self, v):
        self._maxdm = float(v)

    def mindt(self, v):
        self._mindt = float(v)

    def maxdt(self, v):
        self._maxdt = float(v)

    def exchinconv(self, flag):
        self._invalidate()
        self._exchinconv = bool(flag)

    # ------------------------------------------------------------ magnetisation (simconfig.go)
    def uniform(self, mz, my, mx):
        self._invalidate()
        X, Y, Z = self._need_size()
        m = np.empty((3, X, Y, Z), np.float32)
        m[0], m[1], m[2] = mx, my, mz
        self._m0 = m

    def setm_array(self, m):
        """Initial magnetisation in the internal layout (3, X, Y, Z)."""
        self._invalidate()
        m = np.ascontiguousarray(m, np.float32).reshape((3,) + self._need_size())
        self._m0 = m.copy()

    def loadm(self, file):
        """Initial magnetisation from an OVF 1.0 / 2.0 file of the grid size
        (simconfig.go:291-297; the header's multiplier is not applied)."""
        from . import ovf

        m, meta = ovf.read_ovf(file)
        if meta["size"] != tuple(self._need_size()):
            raise InputErr("%s holds a %s grid, the simulation is %s" % (file, meta["size"], self._size))
        self.setm_array(m)

    # ------------------------------------------------------------ output (simoutput.go)
    def field_at(self, t):
        """Applied field in tesla (internal X, Y, Z) at time t (simfield.go:87-153)."""
        if self._static is not None:
            return np.array(self._static, float)
        if not self._field:
            return np.zeros(3)
        ts = [f[0] for f in self._field]
        i = int(np.searchsorted(ts, t, side="left"))
        if i == 0 or i >= len(ts):
            return np.zeros(3)
        (t0, b0), (t1, b1) = self._field[i - 1], self._field[i]
        w = (t - t0) / (t1 - t0)
        return np.asarray(b0, float) + w * (np.asarray(b1, float) - np.asarray(b0, float))

    def save(self, what, format="binary", outdir="."):
        """Save(what, format) (simoutput.go:56-61, 90-117): "m" as
        outdir/m%08d.omf in OVF 1.0 binary ("binary", "binary4", "omf"),
        OVF 2.0 binary ("binary_2", "binary4_2", "omf_2") or OVF 1.0 text
        ("text", "ascii"); "table" appends time, <m> and B (user x, y, z) to
        outdir/datatable.odt."""
        from . import ovf

        import os

        self._init()
        if what == "m":
            fmts = {"binary": (1, True), "binary4": (1, True), "omf": (1, True), "binary_2": (2, True),
                    "binary4_2": (2, True), "omf_2": (2, True), "text": (1, False), "ascii": (1, False)}
            if format not in fmts:
                raise InputErr("unknown format %s. options are: %s" % (format, ", ".join(fmts)))
            version, binary = fmts[format]
            idx = getattr(self, "_autosave_idx", 0)
            path = os.path.join(outdir, "m%08d.omf" % idx)
            self._autosave_idx = idx + 1
            st = self.state()
            t = st["time"] * st["unit_time"]
            desc = {"time": "%g" % t, "msat": "%g" % self._msat, "aexch": "%g" % self._aexch,
                    "alpha": "%g" % self._alpha, "torque": "%g" % st["max_torque"]}
            ovf.write_ovf(path, self.m, self._cell, version=version,
                          binary=binary, multiplier=self._msat, desc=desc, total_time=t)
            return path
        if what == "table":
            path = os.path.join(outdir, "datatable.odt")
            st = self.state()
            t = st["time"] * st["unit_time"]
            m = self.average_m()
            b = self.field_at(t)
            if getattr(self, "_table", None) is None:
                self._table = ovf.OdtTable(path, "mumax data", [("Time", "s"), ("Mx/Ms", ""), ("My/Ms", ""),
                                                               ("Mz/Ms", ""), ("Bx", "T"), ("By", "T"),
                                                               ("Bz", "T")])
            self._table.row([t, m[2], m[1], m[0], b[2], b[1], b[0]])
            return path
        raise InputErr("unknown output quantity %s. options are: m, table" % what)

    def close_table(self):
        if getattr(self, "_table", None) is not None:
            self._table.close()
            self._table = None

    # ------------------------------------------------------------ run (simrun.go)
    def run(self, seconds):
        self._init()
        return int(self._lib.mm_sim_run(self._ptr, float(seconds)))

    def relax(self, max_steps=0):
        """Relax() (relax.go:14-25): steps with energy control until a step
        raises the total energy (or max_steps > 0 steps were taken); returns
        the number of steps taken."""
        self._init()
        return int(self._lib.mm_sim_relax(self._ptr, int(max_steps)))

    def hysteresis(self, bz0, by0, bx0, bz1, by1, bx1, steps, save=True, outdir="."):
        """Hysteresis() (simhysteresis.go:16-27): field scan in `steps`
        intervals, relaxing and saving m (omf) and a table row at each point."""
        for i in range(steps + 1):
            f = i / steps
            self.applystatic("field", bz0 + (bz1 - bz0) * f, by0 + (by1 - by0) * f, bx0 + (bx1 - bx0) * f)
            self.relax()
            if save:
                self.save("m", "omf", outdir)
                self.save("table", "ascii", outdir)

    def steps(self, n):
        self._init()
        self._lib.mm_sim_steps(self._ptr, int(n))

    def step(self):
        self.steps(1)

    def steps_async(self, n):
        self._init()
        self._lib.mm_sim_steps_async(self._ptr, int(n))

    def sync(self):
        self._lib.mm_sim_sync(self._ptr)

    # ------------------------------------------------------------ outputs
    @property
    def m(self):
        self._init()
        out = np.empty((3,) + self._size, np.float32)
        self._lib.mm_sim_get_m(self._ptr, out.ctypes.data)
        return out

    @m.setter
    def m(self, v):
        self._init()
        v = np.ascontiguousarray(v, np.float32).reshape((3,) + self._size)
        self._lib.mm_sim_set_m(self._ptr, v.ctypes.data)

    def state(self):
        self._init()
        st = np.zeros(13)
        self._lib.mm_sim_state(self._ptr, st.ctypes.data)
        keys = ("time", "dt", "steps", "error", "max_torque", "trials", "rejects", "energy", "unit_time",
                "unit_field", "px", "py", "pz")
        return dict(zip(keys, st))

    def trial_log(self, max_entries=1 << 16):
        self._init()
        dt = np.zeros(max_entries)
        acc = np.zeros(max_entries, np.int32)
        n = self._lib.mm_sim_trial_log(self._ptr, dt.ctypes.data, acc.ctypes.data, max_entries)
        return dt[:n], acc[:n]

    def heff(self, m):
        self._init()
        m = np.ascontiguousarray(m, np.float32)
        h = np.empty_like(m)
        self._lib.mm_sim_heff(self._ptr, m.ctypes.data, h.ctypes.data)
        return h

    def torque(self, m):
        self._init()
        m = np.ascontiguousarray(m, np.float32)
        k = np.empty_like(m)
        self._lib.mm_sim_torque(self._ptr, m.ctypes.data, k.ctypes.data)
        return k

    def energy(self):
        """Total energy in internal units x cell volume (energy.go:38-52)."""
        self._init()
        return float(self._lib.mm_sim_energy(self._ptr))

    def energy_si(self):
        st = self.state()
        L = np.sqrt(2 * self._aexch / (4e-7 * np.pi * self._msat ** 2))
        return self.energy() * self._aexch * L

    def average_m(self):
        self._init()
        out = np.zeros(3)
        self._lib.mm_sim_average_m(self._ptr, out.ctypes.data)
        return out

    def kernel_spectrum(self):
        self._init()
        n = self._lib.mm_sim_kernel_spectrum(self._ptr, None)
        out = np.empty(n, np.float32)
        self._lib.mm_sim_kernel_spectrum(self._ptr, out.ctypes.data)
        st = self.state()
        P = (int(st["px"]), int(st["py"]), int(st["pz"]))
        return out.reshape(-1, P[0], P[1], P[2] // 2 + 1)

    def device_bytes(self):
        self._init()
        return int(self._lib.mm_sim_device_bytes(self._ptr))

    def free(self):
        if self._ptr:
            self._lib.mm_sim_free(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # ------------------------------------------------------------ internals
    def _need_size(self):
        if self._size is None:
            raise InputErr("The gridsize should be set first. E.g.: gridsize 32 32 4")
        return self._size

    def _invalidate(self):
        if self._ptr:
            # init.go:20-26: re-initialisation is not allowed
            raise InputErr("This parameter could only be set once.")

    def params(self):
        p = gpukern.Params()
        if self._msat == 0:
            raise InputErr("Saturation magnetization should first be set. E.g. msat 800E3")
        if self._aexch == 0:
            raise InputErr("Exchange constant should first be set. E.g. aexch 12E-13")
        if self._cell is None:
            raise InputErr("The cell size should be set first. E.g. cellsize 1E-9 1E-9 1E-9")
        p.size[:] = list(self._need_size())
        p.cellsize[:] = list(self._cell)
        p.msat, p.aexch, p.alpha = self._msat, self._aexch, self._alpha
        if self._axis is not None:
            p.anisType = 1
            p.anisAxis[:] = list(self._axis)
        p.k1 = self._k1
        p.exchType = 6
        p.exchInConv = int(self._exchinconv)
        p.wantDemag = 1
        p.demagAccuracy = self._demag_accuracy
        p.solver = SOLVERS[self._solver]
        p.maxError, p.dt = self._maxerror, self._dt
        p.minDm, p.maxDm, p.minDt, p.maxDt = self._mindm, self._maxdm, self._mindt, self._maxdt
        if self._static is None and len(self._field) == 1:
            raise InputErr("Pointwise field/current needs at least two points")
        t, b = self._field_table()
        if t is not None:
            self._keep = [t, b]
            p.nField = len(t)
            p.fieldT = t.ctypes.data_as(C.POINTER(C.c_double))
            p.fieldB = b.ctypes.data_as(C.POINTER(C.c_double))
        return p

    def _init(self):
        if self._ptr:
            return
        if self._m0 is None:
            raise InputErr("Initial magnetization was not set")
        p = self.params()
        m0 = np.ascontiguousarray(self._m0, np.float32)
        self._ptr = self._lib.mm_sim_new(C.byref(p), m0.ctypes.data, MODES[self.mode])
        if not self._ptr:
            raise RuntimeError("mm_sim_new failed")


class Batch:
    """Many fused-mode simulations advanced together (mm_batch_*): one launch
    per pass and padded length per RK stage for the whole batch, per-simulation
    step control on the device, one captured HIP graph per batched trial.

    The simulations stay usable alone; a simulation computes the same numbers
    in a batch as alone.  `run(seconds)` is Run() for every member (seconds:
    one value or one per simulation)."""

    def __init__(self, sims, graph=True):
        if not sims:
            raise InputErr("empty batch")
        self.sims = list(sims)
        for s in self.sims:
            if s.mode != "fused":
                raise InputErr("a batch needs fused-mode simulations")
            s._init()
        self._lib = self.sims[0]._lib
        ptrs = (C.c_void_p * len(self.sims))(*[s._ptr for s in self.sims])
        self._ptr = self._lib.mm_batch_new(ptrs, len(self.sims))
        if not self._ptr:
            raise RuntimeError("mm_batch_new failed")
        self._lib.mm_batch_set_graph(self._ptr, int(bool(graph)))

    def __len__(self):
        return len(self.sims)

    def _seconds(self, seconds):
        v = np.broadcast_to(np.asarray(seconds, np.float64), (len(self.sims),))
        return np.ascontiguousarray(v)

    def start(self, seconds=None, steps=None):
        sec = None if seconds is None else self._seconds(seconds)
        st = None if steps is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(steps, np.int64), (len(self.sims),)))
        self._lib.mm_batch_start(self._ptr, None if sec is None else sec.ctypes.data,
                                 None if st is None else st.ctypes.data)

    def trials(self, n):
        """Enqueue n batched trials and wait; returns the simulations not yet done."""
        return int(self._lib.mm_batch_trials(self._ptr, int(n)))

    def done(self):
        out = np.zeros(len(self.sims), np.int32)
        self._lib.mm_batch_done(self._ptr, out.ctypes.data)
        return out.astype(bool)

    def enqueue(self, n):
        """Enqueue n batched trials (and the done flags' copy) without waiting."""
        self._lib.mm_batch_enqueue(self._ptr, int(n))

    def ready(self):
        """Whether the work of the last enqueue has finished."""
        return bool(self._lib.mm_batch_ready(self._ptr))

    def flags(self):
        """Done flags after the last enqueue (waits for it if needed)."""
        out = np.zeros(len(self.sims), np.int32)
        self._lib.mm_batch_flags(self._ptr, out.ctypes.data)
        return out.astype(bool)

    def run(self, seconds, poll=16):
        sec = self._seconds(seconds)
        return int(self._lib.mm_batch_run(self._ptr, sec.ctypes.data, int(poll)))

    def free(self):
        if self._ptr:
            self._lib.mm_batch_free(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
