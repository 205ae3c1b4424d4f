"""OVF 1.0 / 2.0, .tensor and ODT files through libgpukern's C++ readers and
writers (include/mumax_io.h), in the engine's internal layout: m[3][X][Y][Z]
with internal X = user z, Z = user x (mumax/omf, mumax/tensor)."""
import ctypes as C

import numpy as np

from . import gpukern


class OvfInfo(C.Structure):
    _fields_ = [("size", C.c_int * 3), ("step", C.c_double * 3), ("version", C.c_int), ("binary", C.c_int),
                ("multiplier", C.c_double), ("valueunit", C.c_char * 32), ("meshunit", C.c_char * 16),
                ("title", C.c_char * 64), ("total_time", C.c_double), ("stage_time", C.c_double),
                ("ndesc", C.c_int)]


class IOErr(Exception):
    pass


def _check(lib, rc):
    if rc != 0:
        raise IOErr(lib.mm_io_error().decode())


def _strs(items):
    arr = (C.c_char_p * max(1, len(items)))(*[str(x).encode() for x in items])
    return arr


def read_ovf(path):
    """Returns (m[3][X][Y][Z] float32, info dict with the header and desc entries)."""
    lib = gpukern.load()
    info = OvfInfo()
    p = str(path).encode()
    _check(lib, lib.mm_ovf_info(p, C.byref(info)))
    m = np.zeros((3,) + tuple(info.size), np.float32)
    _check(lib, lib.mm_ovf_read(p, m.ctypes.data, C.byref(info)))
    desc = {}
    k, v = C.create_string_buffer(256), C.create_string_buffer(256)
    for i in range(info.ndesc):
        _check(lib, lib.mm_ovf_desc(p, i, k, v, 256))
        desc[k.value.decode()] = v.value.decode()
    meta = dict(size=tuple(info.size), step=tuple(info.step), version=info.version, binary=info.binary,
                multiplier=info.multiplier, valueunit=info.valueunit.decode(), meshunit=info.meshunit.decode(),
                title=info.title.decode(), total_time=info.total_time, stage_time=info.stage_time, desc=desc)
    return m, meta


def write_ovf(path, m, step, version=2, binary=True, multiplier=1.0, valueunit="A/m", meshunit="m",
              title="mumax data", desc=None, total_time=0.0, stage_time=0.0):
    """Write m[3][X][Y][Z] (internal layout) with internal cell size `step`."""
    lib = gpukern.load()
    m = np.ascontiguousarray(m, np.float32)
    info = OvfInfo()
    info.size[:] = list(m.shape[1:])
    info.step[:] = [float(s) for s in step]
    info.version, info.binary, info.multiplier = version, 4 if binary else 0, multiplier
    info.valueunit, info.meshunit, info.title = valueunit.encode(), meshunit.encode(), title.encode()
    info.total_time, info.stage_time = total_time, stage_time
    desc = desc or {}
    keys, vals = _strs(list(desc)), _strs(list(desc.values()))
    _check(lib, lib.mm_ovf_write(str(path).encode(), m.ctypes.data, C.byref(info), keys, vals, len(desc)))


def read_tensor(path):
    lib = gpukern.load()
    rank = C.c_int()
    size = (C.c_int * 16)()
    p = str(path).encode()
    _check(lib, lib.mm_tensor_read(p, C.byref(rank), size, None))
    shape = tuple(size[: rank.value])
    data = np.zeros(shape, np.float32)
    _check(lib, lib.mm_tensor_read(p, C.byref(rank), size, data.ctypes.data))
    return data


def write_tensor(path, data):
    lib = gpukern.load()
    data = np.ascontiguousarray(data, np.float32)
    size = (C.c_int * max(1, data.ndim))(*data.shape)
    _check(lib, lib.mm_tensor_write(str(path).encode(), data.ndim, size, data.ctypes.data))


class OdtTable:
    """ODT 1.0 table: OdtTable(path, title, [(name, unit), ...], desc={key: value});
    row(values); close()."""

    def __init__(self, path, title, columns, desc=None):
        self._lib = gpukern.load()
        self.ncols = len(columns)
        desc = desc or {}
        self._t = self._lib.mm_odt_open_desc(str(path).encode(), title.encode(), self.ncols,
                                             _strs([c[0] for c in columns]), _strs([c[1] for c in columns]),
                                             len(desc), _strs(list(desc)), _strs(list(desc.values())))
        if not self._t:
            raise IOErr(self._lib.mm_io_error().decode())

    def row(self, values):
        v = np.ascontiguousarray(values, np.float64)
        assert v.size == self.ncols
        _check(self._lib, self._lib.mm_odt_row(self._t, v.ctypes.data))

    def close(self):
        if self._t:
            _check(self._lib, self._lib.mm_odt_close(self._t))
            self._t = None
