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
SOLVERS = {"rk2": 2, "rk12": 12, "rk3": 3, "rk23": 23, "rk4": 4, "rkck": 45, "rkdp": 54, "rk45": 54,
           "semianal1": 101, "semianal2": 102}


class InputErr(ValueError):
    """Invalid command argument (the reference panics with InputErr)."""


# columns of the data table, simoutput.go:133-147
TABULATE = ("time", "m", "b", "j", "id", "e", "maxdm/dt", "minmaxmz", "corepos")


def field_function(kind, params, b, t):
    """Value (tesla, internal X, Y, Z) of a field function of time at t seconds,
    as the engine evaluates it (simfield.go:155-260; scale rounded to float32)."""
    f, pi = params[0], np.pi
    if kind == 1:
        d, r = params[0], params[1]
        sc = 0.0
        if 0 < t < r:
            sc = 0.5 - 0.5 * np.cos(t * pi / r)
        elif r <= t < d + r:
            sc = 1.0
        elif d + r <= t < d + 2.0 * r:
            sc = 0.5 + 0.5 * np.cos((t - d - r) * pi / r)
        s = [sc] * 3
    elif kind == 2:
        s = [np.sin(f * 2 * pi * t)] * 3
    elif kind == 3:
        s = [np.sin(f * 2 * pi * t) * (t / params[1] if t < params[1] else 1.0)] * 3
    elif kind == 4:
        s = [np.arcsin(np.sin(f * 2 * pi * t)) / (pi / 2)] * 3
    else:
        s = [np.sin(f * 2 * pi * t + params[1 + c]) for c in range(3)]
    return np.array([np.float32(s[c]) * np.float32(b[c]) for c in range(3)], np.float64)


def _component(c):
    c = int(c)
    if c not in (0, 1, 2):
        raise InputErr("component must be 0 (x), 1 (y) or 2 (z)")
    return c


def _parabola_peak(f0, f1, f2):
    # interpolate_maxpos (corepos.go:36-40) with the neighbours at -1 and +1, in float32
    f0, f1, f2 = np.float32(f0), np.float32(f1), np.float32(f2)
    b = (f2 - f1) / np.float32(2)
    a = ((f2 - f0) - (f0 - f1)) / np.float32(2)
    return -b / (np.float32(2) * a)


class _Output:
    """A saved quantity with its period (Periodic, simoutput.go:64-87)."""

    def __init__(self, kind, version=1, binary=True):
        self.kind, self.version, self.binary = kind, version, binary
        self.period = np.float32(0)
        self.since = np.float32(0)

    def need_save(self, t32):
        return t32 == 0 or np.float32(t32 - self.since) >= self.period

    def due_internal(self, unit32):
        return float(np.float32(self.since + self.period)) / unit32

    def save(self, sim):
        import os

        from . import ovf

        t32 = sim._time_f32()
        if self.kind == "table":
            path = sim._table_row()
        else:
            st = sim.state()
            path = os.path.join(sim._outdir, "%s%08d.omf" % (self.kind, sim._autosave_idx))
            if self.kind == "m":
                data, mult, unit = sim.m, sim._msat, "A/m"
            else:   # torque: dm/dt of the current m in units gamma*Msat
                data, mult, unit = sim.torque(sim.m), 1.0, ""
            ovf.write_ovf(path, data, sim._cell, version=self.version, binary=self.binary, multiplier=mult,
                          valueunit=unit, desc=sim._ovf_desc(), total_time=st["time"] * st["unit_time"],
                          stage_time=float(self.period))
        self.since = t32
        return path


def _resolve_output(what, format):
    # resolve (simoutput.go:93-131)
    what, format = str(what), str(format)
    if what == "m":
        fmts = {"binary": (1, True), "binary4": (1, True), "omf": (1, True), "binary_2": (2, True),
                "binary4_2": (2, True), "omf_2": (2, True), "text": (1, False), "ascii": (1, False)}
        if format not in fmts:
            raise InputErr("unknown format %s. options are: binary (=binary4,omf,binary4_2, omf_2), text (=ascii)"
                           % format)
        return _Output("m", *fmts[format])
    if what == "table":
        return _Output("table")
    if what == "torque":
        if format not in ("binary", "text"):
            raise InputErr("unknown format %s for torque. options are: binary, text" % format)
        return _Output("torque", 1, format == "binary")
    if what in ("phi", "energydensity"):
        raise InputErr("saving %s is not supported by this engine" % what)
    raise InputErr("unknown output quantity %s. options are: m, table, torque" % what)


class Sim:
    def __init__(self, mode="fused"):
        if mode not in MODES:
            raise InputErr("mode must be one of %s" % list(MODES))
        self.mode = mode
        self._size = None
        self._cell = None
        self._periodic = (0, 0, 0)
        self._msat = 0.0
        self._aexch = 0.0
        self._alpha = 0.0
        self._k1 = 0.0
        self._axis = None
        self._field = []  # list of (t, [bX, bY, bZ]); static: single entry with t=None
        self._static = None
        self._fn = None   # (kind, (bX, bY, bZ), params): a field function of time
        self._solver = "rk12"       # DEFAULT_SOLVERTYPE (sim.go:24)
        self._maxerror = 1e-4       # DEFAULT_MAXERROR (sim.go:25)
        self._dt = 0.0
        self._mindm = self._maxdm = self._mindt = self._maxdt = 0.0
        self._exchinconv = True
        self._demag_accuracy = 8
        self._want_energy = False
        self._kernel = 0
        self._alpha_map = None
        self._norm_map = None
        self._temperature = 0.0
        self._thermal_seed = 0
        self._current = (0.0, 0.0, 0.0)   # A/m^2, internal order
        self._spin_pol = 0.0
        self._xi = 0.0
        self._jmask = None
        self._jstatic = None   # applied current density (Apply*("j", ...)): static, points, function
        self._jfield = []
        self._jfn = None
        self._m0 = None
        # output (sim.go:179-209): schedule, table columns, output id
        self._outdir = "."
        self._schedule = []
        self._tab = {k: k in ("time", "m", "b", "id") for k in TABULATE}
        self._table = None
        self._table_path = None
        self._autosave_idx = -1
        self._m_fresh = False
        self._user_desc = {}
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

    def periodic(self, z, y, x):
        """Periodic(z, y, x) (simgeom.go:84-88): the number of periodic images
        along each axis (0: free boundary), in the order of gridsize."""
        self._invalidate()
        v = (int(x), int(y), int(z))
        if min(v) < 0:
            raise InputErr("Periodic images should be >= 0")
        self._periodic = v

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
    @staticmethod
    def _is_current(what):
        # apply (simfield.go:61-70): the field, or the current density
        w = str(what).lower()
        if w in ("currentdensity", "j"):
            return True
        if w in ("field", "b"):
            return False
        raise InputErr("Illegal Argument: %s Options are: [field (or B), currentdensity (or j)]" % what)

    def applystatic(self, what, hz, hy, hx):
        if self._is_current(what):
            f = np.float32
            self._jstatic, self._jfield, self._jfn = (float(f(hx)), float(f(hy)), float(f(hz))), [], None
            self._push_current()
            return
        self._static = (hx, hy, hz)
        self._field = []
        self._fn = None
        self._push_field()

    # functions of time (simfield.go:155-260); amplitude in the reference's (hz, hy, hx) order
    def _apply_fn(self, what, kind, hz, hy, hx, params):
        if self._is_current(what):
            self._jstatic, self._jfield = None, []
            self._jfn = (kind, (float(hx), float(hy), float(hz)), tuple(float(p) for p in params) + (0.0,) * (4 - len(params)))
            self._push_current()
            return
        self._static = None
        self._field = []
        self._fn = (kind, (float(hx), float(hy), float(hz)), tuple(float(p) for p in params) + (0.0,) * (4 - len(params)))
        self._push_field()

    def applypulse(self, what, hz, hy, hx, duration, risetime):
        """ApplyPulse: raised-cosine rise and fall of `risetime`, `duration` at full amplitude."""
        self._apply_fn(what, 1, hz, hy, hx, [duration, risetime])

    def applyrf(self, what, hz, hy, hx, freq):
        """ApplyRf: B sin(2 pi f t)."""
        self._apply_fn(what, 2, hz, hy, hx, [freq])

    def applyrframp(self, what, hz, hy, hx, freq, ramptime):
        """ApplyRfRamp: B sin(2 pi f t) min(t / ramptime, 1)."""
        self._apply_fn(what, 3, hz, hy, hx, [freq, ramptime])

    def applysawtooth(self, what, hz, hy, hx, freq):
        """ApplySawtooth: B asin(sin(2 pi f t)) / (pi/2)."""
        self._apply_fn(what, 4, hz, hy, hx, [freq])

    def applyrotating(self, what, hz, hy, hx, freq, phasex, phasey, phasez):
        """ApplyRotating: component c is B_c sin(2 pi f t + phase_c)."""
        self._apply_fn(what, 5, hz, hy, hx, [freq, phasex, phasey, phasez])

    def applypointwise(self, what, t, hz, hy, hx):
        if self._is_current(what):
            if self._jstatic is not None or self._jfn is not None:
                self._jfield = []
            self._jstatic, self._jfn = None, None
            self._jfield.append((float(t), (float(hx), float(hy), float(hz))))
            self._push_current()
            return
        if self._static is not None:
            self._field = []
        self._static = None
        self._fn = None
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
        if self._ptr and self._fn is not None:
            kind, b, prm = self._fn
            b, prm = np.ascontiguousarray(b, np.float64), np.ascontiguousarray(prm, np.float64)
            self._lib.mm_sim_set_field_function(self._ptr, kind, b.ctypes.data, prm.ctypes.data)
            return
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
        """ExchInConv (simfield.go): the engine folds exchange into the convolution."""
        self._invalidate()
        if not flag:
            raise InputErr("exchinconv false is not supported: exchange is part of the convolution kernel")
        self._exchinconv = True

    def exchtype(self, neighbours):
        """ExchType (simfield.go:41-44): only the 6-neighbour Laplacian."""
        self._invalidate()
        if int(neighbours) != 6:
            raise InputErr("exchange type %d is not supported (only 6)" % neighbours)

    def _alpha_mask(self):
        # initAlphaMask (simgeom.go:226-240): a multiplier of 1 in every cell
        self._invalidate()
        if self._alpha_map is None:
            self._alpha_map = np.ones(tuple(self._need_size()), np.float32)
        return self._alpha_map

    def setalpha(self, z, y, x, alpha):
        """SetAlpha (simgeom.go:246-250): damping multiplier of the cell of user
        indices (x, y, z) = the arguments' (z, y, x) order of the reference."""
        self._alpha_mask()[x, y, z] = alpha

    def setalpharange(self, x1, y1, z1, x2, y2, z2, alpha):
        """SetAlphaRange (simgeom.go:253-264): multiplier over user x1..x2,
        y1..y2, z1..z2 (inclusive)."""
        self._alpha_mask()[z1:z2 + 1, y1:y2 + 1, x1:x2 + 1] = alpha

    def _msat_mask(self):
        # initMsatMask (simgeom.go:266-278): a norm of 1 in every cell
        self._invalidate()
        if self._norm_map is None:
            self._norm_map = np.ones(tuple(self._need_size()), np.float32)
        return self._norm_map

    def _msat_box(self, lo, hi, msat):
        # the reference's Set panics outside the grid: an input error here
        n = self._msat_mask()
        for a, b, size in zip(lo, hi, n.shape):
            if a <= b and (a < 0 or b >= size):
                raise InputErr("cells %s..%s are outside the grid %s" % (tuple(lo), tuple(hi), n.shape))
        n[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1] = np.float32(msat)

    def setmsat(self, z, y, x, msat):
        """SetMsat (simgeom.go:279-284): the normalised saturation magnetisation
        of one cell, i.e. the norm every normalisation gives m there (0: an
        empty cell); indexed as SetAlpha."""
        x, y, z = int(x), int(y), int(z)
        self._msat_box((x, y, z), (x, y, z), msat)

    def setmsatrange(self, z1, y1, x1, z2, y2, x2, msat):
        """SetMsatRange (simgeom.go:287-299): the norm over cells x1..x2 (first
        internal index), y1..y2, z1..z2 (third), inclusive."""
        self._msat_box((int(x1), int(y1), int(z1)), (int(x2), int(y2), int(z2)), msat)

    def setmsat_array(self, nmap):
        """The whole norm map at once, internal layout (X, Y, Z)."""
        n = self._msat_mask()
        n[...] = np.asarray(nmap, np.float32).reshape(n.shape)

    def setmsatellips(self, cx, cy, sx, sy, msat):
        """SetMsatEllips (simgeom.go:301-340): the norm inside the ellipse of
        centre (cx, cy) and axes (sx, sy) (metres; cx along the last internal
        axis, cy along the middle one) through every layer, in the reference's
        float32 arithmetic over its cell range."""
        if self._cell is None:
            raise InputErr("The cell size should be set first. E.g. cellsize 1E-9 1E-9 1E-9")
        f = np.float32
        n = self._msat_mask()
        cx, cy, sx, sy = f(cx), f(cy), f(sx), f(sy)
        rx, ry = sx / f(2.0), sy / f(2.0)
        c1, c2 = f(self._cell[1]), f(self._cell[2])
        start_x = int(np.floor(np.float64((cx - rx) / c2)))
        start_y = int(np.floor(np.float64((cy - ry) / c1)))
        stop_x = int(np.ceil(np.float64((cx + rx) / c2)))
        stop_y = int(np.ceil(np.float64((cy + ry) / c1)))
        for j in range(start_y, stop_y + 1):
            y = f(j) * c1 - cy
            for k in range(start_x, stop_x + 1):
                x = f(k) * c2 - cx
                if x * x / rx / rx + y * y / ry / ry < f(1):
                    if not (0 <= j < n.shape[1] and 0 <= k < n.shape[2]):
                        raise InputErr("the ellipse covers cell (%d, %d) outside the grid %s" % (j, k, n.shape))
                    n[:, j, k] = f(msat)

    def temperature(self, T):
        """Temperature(T) (temperature.go:47-56): a thermal field of standard
        deviation sqrt(2 alpha kB T / (gamma0 mu0 Ms V dt)) / Ms per component,
        drawn at the start of each step.  Before the simulation exists it also
        selects the RK12 solver, as the reference does; afterwards the solver
        must already be RK12, and the new temperature applies from the next
        trial."""
        T = float(np.float32(T))
        if T < 0:
            raise InputErr("temperature must be >= 0")
        if not self._ptr:
            self._solver = "rk12"
        elif self._solver != "rk12":
            raise InputErr("Finite temperature is only compatible with the RK12 solver type")
        self._temperature = T
        if self._ptr:
            self._lib.mm_sim_set_temperature(self._ptr, T)

    def thermalseed(self, seed):
        """Seed of the thermal field's counter-based stream (the reference seeds
        cuRAND with 0 at its first use, gpu_temperature.cu:31-39)."""
        self._invalidate()
        self._thermal_seed = int(seed) & ((1 << 64) - 1)

    def kerneltype(self, name):
        """KernelType (simfield.go): the demag kernel.  "facekernel6" (also the
        reference's generator names "mumaxkern-go", "mumaxkern") is the
        reference's FaceKernel6; "newell" the Newell tensor (cell-averaged)."""
        self._invalidate()
        key = str(name).lower()
        if key in ("facekernel6", "mumaxkern-go", "mumaxkern"):
            self._kernel = 0
        elif key == "newell":
            self._kernel = 1
        else:
            raise InputErr("unknown kernel type %s. options are: facekernel6, newell" % name)

    def demagaccuracy(self, accuracy):
        """DemagAccuracy (simfield.go:280-287)."""
        self._invalidate()
        self._demag_accuracy = int(accuracy)

    # ------------------------------------------------------------ magnetisation (simconfig.go)
    def uniform(self, mz, my, mx):
        self._invalidate()
        X, Y, Z = self._need_size()
        m = np.empty((3, X, Y, Z), np.float32)
        m[0], m[1], m[2] = mx, my, mz
        self._m0 = m

    def _mlocal(self):
        # initMLocal: a zero magnetisation of the grid size until one is set
        self._invalidate()
        if self._m0 is None:
            self._m0 = np.zeros((3,) + tuple(self._need_size()), np.float32)
        return self._m0

    def setmcell(self, z, y, x, mz, my, mx):
        """SetMCell (simconfig.go:36-43): cell of user indices (x, y, z) = internal (Z, Y, X)."""
        m = self._mlocal()
        m[0, x, y, z], m[1, x, y, z], m[2, x, y, z] = mx, my, mz

    def vortex(self, circulation, polarization):
        """Vortex (simconfig.go:74-95): in-plane circulation around the grid centre,
        core along the thickness with the given polarization."""
        m = self._mlocal()
        X, Y, Z = m.shape[1:]
        cy, cz = Y // 2, Z // 2
        yy, zz = np.meshgrid(np.arange(Y) - cy, np.arange(Z) - cz, indexing="ij")
        m[0] = 0.0
        m[1] = zz * circulation
        m[2] = -yy * circulation
        m[:, :, cy, cz] = 0.0
        m[0, :, cy, cz] = polarization

    def setmrange(self, z1, y1, x1, z2, y2, x2, mz, my, mx):
        """SetMRange (simconfig.go:20-32): cells of user indices x1..x2, y1..y2,
        z1..z2 (inclusive; internal Z, Y, X) set to (mx, my, mz)."""
        m = self._mlocal()
        sl = (slice(x1, x2 + 1), slice(y1, y2 + 1), slice(z1, z2 + 1))
        m[(0,) + sl], m[(1,) + sl], m[(2,) + sl] = mx, my, mz

    def _walls(self, phi_of_j):
        # SBW/SNW/ABW/ANW (simconfig.go:212-289): theta over the Z (user x) index
        m = self._mlocal()
        X, Y, Z = m.shape[1:]
        k = np.arange(Z)
        theta = np.pi * ((k - Z // 2) + 0.5) / Z
        phi = np.broadcast_to(phi_of_j(np.arange(Y), Y), (Y,))
        th, ph = np.meshgrid(theta, phi, indexing="xy")   # [Y][Z]
        m[0] = np.sin(th).astype(np.float32)
        m[1] = (np.cos(th) * np.cos(ph)).astype(np.float32)
        m[2] = (np.cos(th) * np.sin(ph)).astype(np.float32)

    def sbw(self):
        """SBW: starting state of a symmetric Bloch wall."""
        self._walls(lambda j, Y: np.zeros(Y))

    def snw(self):
        """SNW: starting state of a symmetric Neel wall."""
        self._walls(lambda j, Y: np.full(Y, np.pi / 2.0))

    def abw(self):
        """ABW: starting state of an asymmetric Bloch wall (phi over the user y index)."""
        self._walls(lambda j, Y: np.pi * ((j - Y // 2) + 0.5) / Y)

    def anw(self):
        """ANW (simconfig.go:272-289, phi = pi/4 as there)."""
        self._walls(lambda j, Y: np.full(Y, np.pi / 4.0))

    def _vortex_block(self, cx, cy, sx, sy, circulation, polarization, ellipse):
        # SetVortexEllips / SetVortexRectangle (simconfig.go:159-209), float32 arithmetic
        m = self._mlocal()
        f = np.float32
        c1, c2 = f(self._cell[1]), f(self._cell[2])
        cxi, cyi = int(f(cx) / c2), int(f(cy) / c1)
        rx, ry = int(f(sx) / c2 / f(2)), int(f(sy) / c1 / f(2))
        X, Y, Z = m.shape[1:]
        for j in range(-ry, ry + 1):
            for k in range(-rx, rx + 1):
                if ellipse:
                    with np.errstate(divide="ignore", invalid="ignore"):
                        inside = f(j * j) / f(ry * ry) + f(k * k) / f(rx * rx) <= f(1)
                    if not inside:
                        continue
                y, z = j + cyi, k + cxi
                if not (0 <= y < Y and 0 <= z < Z):
                    raise InputErr("the vortex extends beyond the grid")
                m[0, :, y, z] = 0.0
                m[1, :, y, z] = k * circulation
                m[2, :, y, z] = -j * circulation
        if not (0 <= cyi < Y and 0 <= cxi < Z):
            raise InputErr("the vortex centre is outside the grid")
        m[:, :, cyi, cxi] = 0.0
        m[0, :, cyi, cxi] = polarization

    def setvortexellips(self, cx, cy, sx, sy, circulation, polarization):
        """SetVortexEllips: a vortex inside the ellipse of centre (cx, cy) and
        axes (sx, sy) metres (user x, y)."""
        self._vortex_block(cx, cy, sx, sy, int(circulation), int(polarization), True)

    def setvortexrectangle(self, cx, cy, sx, sy, circulation, polarization):
        """SetVortexRectangle: the same over the rectangle sx x sy."""
        self._vortex_block(cx, cy, sx, sy, int(circulation), int(polarization), False)

    def vortexinarray(self, vi, vj, usx, usy, sepx, sepy, circulation, polarization):
        """VortexInArray (simconfig.go:127-157): the vortex of unit (vi, vj) of
        an array of unit size (usx, usy) and separation (sepx, sepy), metres."""
        m = self._mlocal()
        f = np.float32
        c1, c2 = f(self._cell[1]), f(self._cell[2])
        bx, sxc = f(usx) / c2, f(sepx) / c2
        by, syc = f(usy) / c1, f(sepy) / c1
        cx = f(vi) * bx + bx / f(2) - f(0.5)
        cy = f(vj) * by + by / f(2) - f(0.5)
        X, Y, Z = m.shape[1:]
        j0, j1 = int(f(vj) * by + syc / f(2)), int(f(vj + 1) * by - syc / f(2))
        k0, k1 = int(f(vi) * bx + sxc / f(2)), int(f(vi + 1) * bx - sxc / f(2))
        if j0 < 0 or k0 < 0 or j1 > Y or k1 > Z or not (0 <= int(cy) < Y and 0 <= int(cx) < Z):
            raise InputErr("the vortex unit extends beyond the grid")
        for j in range(j0, j1):
            for k in range(k0, k1):
                m[0, :, j, k] = 0.0
                m[1, :, j, k] = (f(k) - cx) * f(circulation)
                m[2, :, j, k] = -(f(j) - cy) * f(circulation)
        m[:, :, int(cy), int(cx)] = 0.0
        m[0, :, int(cy), int(cx)] = polarization

    def setrandom(self, seed=None):
        """SetRandom (simconfig.go:314-321): every component uniform on [-0.5, 0.5)
        (numpy's generator, not Go's math/rand); `seed` is a seed or a Generator."""
        m = self._mlocal()
        m[...] = (np.random.default_rng(seed).random(m.shape) - 0.5).astype(np.float32)

    def addnoise(self, amplitude, seed=None):
        """AddNoise (simconfig.go:303-311): each component += amplitude * U[-1, 1)."""
        m = self._mlocal()
        m += (2 * amplitude * (np.random.default_rng(seed).random(m.shape) - 0.5)).astype(np.float32)

    def run_until(self, stop):
        """Run()'s step loop up to the absolute internal time `stop`, without
        Run()'s normalisation at the start (runs interrupted at output times)."""
        self._init()
        return int(self._lib.mm_sim_run_until(self._ptr, float(stop)))

    def setm_array(self, m):
        """Initial magnetisation in the internal layout (3, X, Y, Z)."""
        self._invalidate()
        m = np.ascontiguousarray(m, np.float32).reshape((3,) + self._need_size())
        self._m0 = m.copy()

    # ------------------------------------------------------------ spin-transfer torque (simspintorque.go)
    def _push_spintorque(self):
        if self._ptr:
            j = np.asarray(self._current, np.float64)
            if j.any() and self._jmask is None:
                raise InputErr("Spin-transfer torque requires a currentMask to be set")
            self._lib.mm_sim_set_spintorque(self._ptr, j.ctypes.data, self._spin_pol, self._xi)

    def _applied_current(self):
        return self._jfn is not None or self._jstatic is not None or len(self._jfield) >= 2

    def _push_current(self):
        if not self._ptr:
            return
        if len(self._jfield) == 1 and self._jstatic is None and self._jfn is None:
            return  # like the field: a pointwise current needs two points (simfield.go:123-125)
        if self._applied_current() and self._jmask is None:
            raise InputErr("Spin-transfer torque requires a currentMask to be set")
        prm = np.zeros(4)
        if self._jfn is not None:
            kind, amp, p4 = self._jfn
            j, t, n, prm = np.asarray(amp, np.float64), None, 0, np.asarray(p4, np.float64)
        elif self._jstatic is not None:
            kind, j, t, n = 0, np.asarray(self._jstatic, np.float64), np.zeros(1), 1
        else:
            kind, n = 0, len(self._jfield)
            t = np.asarray([p[0] for p in self._jfield], np.float64)
            j = np.asarray([p[1] for p in self._jfield], np.float64).reshape(-1)
            if n == 0:
                t = j = None
        self._lib.mm_sim_set_current(self._ptr, n, None if t is None else t.ctypes.data,
                                     None if j is None else j.ctypes.data, kind, prm.ctypes.data)

    def current_at(self, t):
        """The current density in A/m^2 (internal X, Y, Z) the torque uses at
        time t: the applied one (simfield.go:87-260) or CurrentDensity's."""
        if self._jfn is not None:
            return field_function(self._jfn[0], self._jfn[2], self._jfn[1], t)
        if self._jstatic is not None:
            return np.array(self._jstatic, float)
        if len(self._jfield) < 2:
            return np.array(self._current, float)
        ts = [f[0] for f in self._jfield]
        i = int(np.searchsorted(ts, t, side="left"))
        if i == 0 or i >= len(ts):
            return np.zeros(3)
        (t0, b0), (t1, b1) = self._jfield[i - 1], self._jfield[i]
        w = (t - t0) / (t1 - t0)
        return np.asarray(b0, float) + w * (np.asarray(b1, float) - np.asarray(b0, float))

    def currentdensity(self, jz, jy, jx):
        """CurrentDensity(jz, jy, jx) (simspintorque.go:30-36), A/m^2: a nonzero
        current makes every torque the spin-transfer (Zhang-Li) form
        (torque.go:40-91); the ABI mode runs it.  Takes effect at the next
        torque on a running simulation, as in the reference."""
        f = np.float32
        self._current = (float(f(jx)), float(f(jy)), float(f(jz)))
        self._push_spintorque()

    def spinpolarization(self, p):
        """SpinPolarization(p) (simspintorque.go:24-27)."""
        self._spin_pol = float(np.float32(p))
        self._push_spintorque()

    def xi(self, xi):
        """Xi(xi), the degree of non-adiabaticity (simspintorque.go:19-22)."""
        self._xi = float(np.float32(xi))
        self._push_spintorque()

    def currentmask_array(self, mask):
        """The current mask, internal layout (3, X, Y, Z): component d scales the
        current along axis d per cell (gpu_spintorque.cu:39-47)."""
        self._invalidate()
        self._jmask = np.ascontiguousarray(mask, np.float32).reshape((3,) + tuple(self._need_size())).copy()

    def currentmask(self, file):
        """CurrentMask(file) (simspintorque.go:39-52): the mask from an OVF file
        of the grid size (the reference resamples other sizes; here they are an
        input error)."""
        from . import ovf

        mask, meta = ovf.read_ovf(file)
        if meta["size"] != tuple(self._need_size()):
            raise InputErr("%s holds a %s grid, the simulation is %s" % (file, meta["size"], self._size))
        self.currentmask_array(mask)

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
        """Applied field in tesla (internal X, Y, Z) at time t (simfield.go:87-260)."""
        if self._fn is not None:
            return field_function(self._fn[0], self._fn[2], self._fn[1], t)
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

    def outputdir(self, path):
        """Directory of the saved files (simoutput.go:30-40); created if absent."""
        import os

        os.makedirs(path, exist_ok=True)
        self._outdir = str(path)

    def autosave(self, what, format, interval):
        """Autosave(what, format, interval) (simoutput.go:46-52): save every
        `interval` seconds of simulated time, checked before each step of
        Run/Step and after each step of Relax."""
        out = _resolve_output(what, format)
        out.period = np.float32(interval)
        self._schedule.append(out)

    def tabulate(self, what, want):
        """Tabulate(what, bool) (simoutput.go:210-215): columns of the table
        (time, m, b, j, id, e, maxdm/dt, minmaxmz, corepos)."""
        if self._table is not None:
            raise InputErr("tabulate must be called before the datatable is opened (Before run, desc, autosave, ...).")
        key = str(what).lower()
        if key not in TABULATE:
            raise InputErr("Illegal argument: %s options are: %s" % (what, list(TABULATE)))
        self._tab[key] = bool(want)

    def desc(self, key, value):
        """Desc(key, value) (simmetadata.go:19-30): a header entry of the table
        (which it opens) and of every OVF file saved afterwards."""
        self._user_desc[str(key)] = str(value)
        if self._table is None:
            self._open_table()

    def save(self, what, format="binary", outdir=None):
        """Save(what, format) (simoutput.go:56-61, 90-117): "m" as
        m%08d.omf in OVF 1.0 binary ("binary", "binary4", "omf"), OVF 2.0
        binary ("binary_2", "binary4_2", "omf_2") or OVF 1.0 text ("text",
        "ascii"); "torque" as torque%08d.omf; "table" appends a row to
        datatable.odt.  The number in the file name is the output id, which
        advances once per state of m that is saved (simrun.go:117-133).
        Returns the path written."""
        self._init()
        if outdir is not None:
            self.outputdir(outdir)
        out = _resolve_output(what, format)
        self._assure_output()
        return out.save(self)

    def close_table(self):
        if self._table is not None:
            self._table.close()
            self._table = None

    def _assure_output(self):
        # assureOutputUpToDate (simrun.go:117-133): a new output id per saved state
        if not self._m_fresh:
            self._autosave_idx += 1
            self._m_fresh = True

    def _time_f32(self):
        # float32(s.time) * s.UnitTime(): the output schedule works in float32 seconds
        st = self.state()
        return np.float32(np.float32(st["time"]) * np.float32(st["unit_time"]))

    def _save_scheduled(self):
        if not self._schedule:
            return
        t = self._time_f32()
        for out in self._schedule:
            if out.need_save(t):
                self._assure_output()
                out.save(self)

    def _ovf_desc(self):
        st = self.state()
        b = self.field_at(st["time"] * st["unit_time"])
        d = {"msat": self._msat, "aexch": self._aexch, "alpha": self._alpha,
             "time": st["time"] * st["unit_time"], "Bx": b[2], "By": b[1], "Bz": b[0],
             "torque": st["max_torque"], "id": self._autosave_idx, "iteration": int(st["steps"])}
        d = {k: "%.10g" % v for k, v in d.items()}
        d.update(self._user_desc)
        return d

    def _open_table(self):
        import os

        from . import ovf

        cols = []
        if self._tab["time"]:
            cols += [("Time", "s")]
        if self._tab["m"]:
            cols += [("Mx/Ms", ""), ("My/Ms", ""), ("Mz/Ms", "")]
        if self._tab["b"]:
            cols += [("Bx", "T"), ("By", "T"), ("Bz", "T")]
        if self._tab["j"]:
            cols += [("jx", "A/m2"), ("jy", "A/m2"), ("jz", "A/m2")]
        if self._tab["e"]:
            cols += [("Energy", "J")]
        if self._tab["maxdm/dt"]:
            cols += [("max_dm/dt", "gammaMs")]
        if self._tab["minmaxmz"]:
            cols += [("min_Mz/Ms", ""), ("max_Mz/Ms", "")]
        if self._tab["corepos"]:
            cols += [("core_x", "m"), ("core_y", "m")]
        if self._tab["id"]:
            cols += [("id", "")]
        self._table_path = os.path.join(self._outdir, "datatable.odt")
        self._table = ovf.OdtTable(self._table_path, "mumax data", cols, desc=self._user_desc)

    def _table_row(self):
        # Table.Save (simoutput.go:217-272)
        if self._table is None:
            self._open_table()
        st = self.state()
        t = st["time"] * st["unit_time"]
        row = []
        if self._tab["time"]:
            row += [t]
        if self._tab["m"]:
            m = self.average_m()
            row += [m[2], m[1], m[0]]
        if self._tab["b"]:
            b = self.field_at(t)
            row += [b[2], b[1], b[0]]
        if self._tab["j"]:
            j = self.current_at(t)
            row += [j[2], j[1], j[0]]  # simoutput.go:244
        if self._tab["e"]:
            row += [self.energy_si()]
        if self._tab["maxdm/dt"]:
            row += [st["max_torque"]]
        if self._tab["minmaxmz"]:
            mz = self.m[0]
            row += [float(mz.min()), float(mz.max())]
        if self._tab["corepos"]:
            row += list(self.getcorepos())
        if self._tab["id"]:
            row += [self._autosave_idx]
        self._table.row(row)
        return self._table_path

    # ------------------------------------------------------------ getters (simget.go)
    def getm(self, component):
        """Average of a user component (0 = x, 1 = y, 2 = z) of m (simget.go:25-33)."""
        return float(np.float32(self.average_m()[2 - _component(component)]))

    def getmpos(self, component, z, y, x):
        """m component at cell (user indices x = z arg, ..., simget.go:36-38)."""
        return float(self.m[2 - _component(component), x, y, z])

    def getmaxm(self, component):
        return float(self.m[2 - _component(component)].max())

    def getminm(self, component):
        return float(self.m[2 - _component(component)].min())

    def getmaxtorque(self):
        """Maximum torque of the last step in units gamma*Msat (simget.go:64-67)."""
        return float(np.float32(self.state()["max_torque"]))

    def getcorepos(self):
        """Vortex core position in metres, centre = (0, 0) (corepos.go:11-34):
        the cell of largest |m_z| in the first layer away from the edges,
        refined by a parabola through its neighbours."""
        mz = np.abs(self.m[0, 0]).astype(np.float32)   # [Y][Z]
        Y, Z = mz.shape
        if Y < 3 or Z < 3:
            raise InputErr("the core position needs at least 3 x 3 cells")
        inner = mz[1:-1, 1:-1]
        iy, iz = np.unravel_index(int(np.argmax(inner)), inner.shape)
        iy, iz = iy + 1, iz + 1
        f0 = mz[iy, iz]
        with np.errstate(divide="ignore", invalid="ignore"):
            px = np.float32(iz) + _parabola_peak(f0, mz[iy, iz - 1], mz[iy, iz + 1]) - np.float32(Z) / 2
            py = np.float32(iy) + _parabola_peak(f0, mz[iy - 1, iz], mz[iy + 1, iz]) - np.float32(Y) / 2
        return float(np.float32(px * np.float32(self._cell[2]))), float(np.float32(py * np.float32(self._cell[1])))

    def gete(self):
        """Total energy in SI units (simget.go:74-77)."""
        return float(np.float32(self.energy_si()))

    def gettime(self):
        st = self.state()
        return st["time"] * st["unit_time"]

    # ------------------------------------------------------------ run (simrun.go)
    def run(self, seconds):
        """Run(seconds) (simrun.go:19-57): normalise m, then step until the
        time has advanced by `seconds`, saving scheduled output before each
        step.  Returns the number of steps taken."""
        self._no_semianal("run")
        self._init()
        if not self._schedule:
            n = int(self._lib.mm_sim_run(self._ptr, float(seconds)))
        else:
            st = self.state()
            stop = st["time"] + float(seconds) / st["unit_time"]
            self._lib.mm_sim_run(self._ptr, 0.0)      # Run()'s normalisation
            self._m_fresh = False
            n = self._run_scheduled(stop)
        self._m_fresh = False
        self._assure_output()
        return n

    def _run_scheduled(self, stop):
        # The steps between two outputs run on the device in one stretch: it
        # ends at the first step boundary at or after a time just below the
        # next output's due time; from there single steps find the boundary
        # where the reference's float32 check first passes.
        n = 0
        while True:
            st = self.state()
            t = st["time"]
            if not t < stop:
                return n
            self._save_scheduled()
            unit = float(np.float32(st["unit_time"]))
            due = min(out.due_internal(unit) for out in self._schedule)
            target = min(stop, due * (1.0 - 1e-5))
            if target > t:
                n += self.run_until(target)
            else:
                self._lib.mm_sim_steps(self._ptr, 1)
                n += 1
            self._m_fresh = False

    def relax(self, max_steps=0):
        """Relax() (relax.go:14-25): steps with energy control until a step
        raises the total energy (or max_steps > 0 steps were taken); scheduled
        output is saved after each step.  Returns the number of steps taken."""
        self._no_semianal("relax")
        self._init()
        self._m_fresh = False
        if not self._schedule:
            return int(self._lib.mm_sim_relax(self._ptr, int(max_steps)))
        n, prev = 0, np.inf
        while True:
            self._lib.mm_sim_relax(self._ptr, 1)
            n += 1
            self._m_fresh = False
            self._save_scheduled()
            e = self.state()["energy"]
            if e > prev or (max_steps > 0 and n >= max_steps):
                return n
            prev = e

    def hysteresis(self, bz0, by0, bx0, bz1, by1, bx1, steps, save=True, outdir=None):
        """Hysteresis() (simhysteresis.go:16-27): field scan in `steps`
        intervals, relaxing and saving m (omf) and a table row at each point."""
        f32 = np.float32

        def point(b0, b1, i):
            # float32 in the reference's order: b0 + (b1 - b0) * float32(i) / float32(steps)
            return float(f32(b0) + (f32(b1) - f32(b0)) * f32(i) / f32(steps))

        for i in range(steps + 1):
            self.applystatic("field", point(bz0, bz1, i), point(by0, by1, i), point(bx0, bx1, i))
            self.relax()
            if save:
                self.save("m", "omf", outdir)
                self.save("table", "ascii", outdir)

    def steps(self, n):
        """Steps(n) (simrun.go:88-92): n times Step()."""
        self._init()
        if not self._schedule:
            self._lib.mm_sim_steps(self._ptr, int(n))
        else:
            for _ in range(int(n)):
                self._save_scheduled()
                self._lib.mm_sim_steps(self._ptr, 1)
        self._m_fresh = False

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

    def trial_errors(self, max_entries=1 << 16):
        """The error estimate of each logged trial (0 for fixed-step solvers)."""
        self._init()
        err = np.zeros(max_entries, np.float32)
        n = self._lib.mm_sim_trial_errors(self._ptr, err.ctypes.data, max_entries)
        return err[:n]

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

    TUNABLES = ("zr", "er", "yb", "xb", "y2b", "pp", "xv")

    def tune(self, reps=20):
        """Autotune this simulation's launch shapes on the device
        (mm_sim_tune): returns {tunable: (chosen value, ms per launch at the
        default, ms at the chosen value)} for the tunables its plan has, and
        for plans the plane pass fits "plane_vs_three": (ms of the three
        separate passes, ms of the plane pass); the choice is also stored in
        the process's tuning table."""
        self._init()
        par = np.zeros(7, np.int32)
        md = np.zeros(8)
        mt = np.zeros(8)
        self._lib.mm_sim_tune(self._ptr, int(reps), par.ctypes.data, md.ctypes.data, mt.ctypes.data)
        out = {k: (int(par[i]), float(md[i]), float(mt[i])) for i, k in enumerate(self.TUNABLES) if md[i] >= 0}
        if md[7] >= 0:
            out["plane_vs_three"] = (float(md[7]), float(mt[7]))  # ms: three passes, plane pass
        return out

    def tune_params(self):
        """The launch shapes this simulation runs with (mm_sim_get_tune)."""
        self._init()
        par = np.zeros(7, np.int32)
        self._lib.mm_sim_get_tune(self._ptr, par.ctypes.data)
        return dict(zip(self.TUNABLES, (int(v) for v in par)))

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

    def _no_semianal(self, what):
        # SemiAnal1/2 (semianal.go) never advance the time: only Steps ends with them
        if self._solver.startswith("semianal"):
            raise InputErr("%s with %s would not end (its steps do not advance the time): use steps" % (what, self._solver))

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
        p.periodic[:] = list(self._periodic)
        if self.mode == "fused":
            for n, k in zip(p.size, self._periodic):
                if k and n % 2 and n > 1:
                    raise InputErr("a periodic axis of odd size (%d) needs mode='abi'" % n)
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
        p.wantEnergy = int(self._want_energy)
        p.demagKernel = self._kernel
        p.temperature = self._temperature
        p.seed = self._thermal_seed
        if self._alpha_map is not None:
            amap = np.ascontiguousarray(self._alpha_map, np.float32)
            self._keep_alpha = amap
            p.alphaMap = amap.ctypes.data_as(C.POINTER(C.c_float))
        if self._solver.startswith("semianal") and self.mode != "abi":
            raise InputErr("the semi-analytic solvers run in the ABI mode (Sim(mode='abi'))")
        if any(self._current) or self._applied_current():
            if self.mode != "abi":
                raise InputErr("spin-transfer torque runs in the ABI mode (Sim(mode='abi'))")
            if self._jmask is None:
                raise InputErr("Spin-transfer torque requires a currentMask to be set")
            if self._alpha_map is not None:
                raise InputErr("spin-transfer torque with a space-dependent alpha is not implemented (torque.go:66)")
        p.current[:] = list(self._current)
        p.spinPol, p.xi = self._spin_pol, self._xi
        if self._jmask is not None:
            self._keep_jmask = self._jmask
            p.currentMask = self._jmask.ctypes.data_as(C.POINTER(C.c_float))
        if self._norm_map is not None:
            nmap = np.ascontiguousarray(self._norm_map, np.float32)
            self._keep_norm = nmap
            p.normMap = nmap.ctypes.data_as(C.POINTER(C.c_float))
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
        if self._fn is not None:
            self._push_field()
        if self._applied_current():
            self._push_current()


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
        self._status = np.zeros(len(self.sims), np.int32)

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
        self._lib.mm_batch_enqueue(self._ptr, int(n))
        return int((~self.flags()).sum())

    def done(self):
        """Done flags now (a failed member is done too; see status())."""
        out = np.zeros(len(self.sims), np.int32)
        self._lib.mm_batch_done(self._ptr, out.ctypes.data)
        self._status = out
        return out.astype(bool)

    def enqueue(self, n):
        """Enqueue n batched trials (and the done flags' copy) without waiting."""
        self._lib.mm_batch_enqueue(self._ptr, int(n))

    def ready(self):
        """Whether the work of the last enqueue has finished."""
        return bool(self._lib.mm_batch_ready(self._ptr))

    def flags(self):
        """Done flags after the last enqueue (waits for it if needed); a
        failed member is done too (see status())."""
        out = np.zeros(len(self.sims), np.int32)
        self._lib.mm_batch_flags(self._ptr, out.ctypes.data)
        self._status = out
        return out.astype(bool)

    def status(self):
        """Per-member status as of the last flags() / done(): 0 running,
        1 done, 2 failed (stopped on a NaN dt, rk.go:448-452, or a NaN or
        infinite torque or error estimate)."""
        return self._status.copy()

    def failed(self):
        return self._status == 2

    def remaining(self):
        """Each member's remaining trials at its current dt after the last
        enqueue (0 once done)."""
        out = np.zeros(len(self.sims), np.float32)
        self._lib.mm_batch_remaining(self._ptr, out.ctypes.data)
        return out

    def run(self, seconds, poll=16):
        sec = self._seconds(seconds)
        n = int(self._lib.mm_batch_run(self._ptr, sec.ctypes.data, int(poll)))
        self.done()
        return n

    def free(self):
        if self._ptr:
            self._lib.mm_batch_free(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
