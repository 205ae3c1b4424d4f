"""MI355X-native micromagnetic LLG engine (drop-in libgpukern.so).

Import with ``importlib.import_module("1_amd")``.  The compute lives entirely in
``libgpukern.so`` (HIP, gfx950); this package only loads it (``gpukern``),
mirrors the reference's simulation commands (``sim.Sim``) and defines the
seeded workloads (``workloads``).
"""
from . import gpukern, workloads  # noqa: F401
from .gpukern import LIB_PATH, load  # noqa: F401
from .sim import Batch, InputErr, Sim  # noqa: F401
