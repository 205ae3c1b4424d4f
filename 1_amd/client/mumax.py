"""Client side of the line protocol between a Python input script and the
simulation process (the API of lib/mumax.py:1-470).

A script run by ``python -m 1_amd.frontend script.py`` imports this module
(``from mumax import *``).  Every function writes one command line, the
command name followed by its arguments separated by spaces, to stdout and
flushes it; a getter then reads one reply line from stdin, which starts with
"%" and holds a number (simget.go:17-21, lib/mumax.py:447-492).  Vector
arguments are in the user's (x, y, z) order, as in lib/mumax.py.
"""
import sys

inf = float("inf")


def _send(command, *args):
    sys.stdout.write(" ".join([command] + [str(a) for a in args]) + "\n")
    sys.stdout.flush()


def _recv():
    line = sys.stdin.readline()
    if not line.startswith("%"):
        sys.stderr.write("mumax client: expected a '%%' reply, got %r\n" % line)
        sys.exit(10)
    return float(line[1:])


def _command(name, nargs):
    def f(*args):
        if len(args) != nargs:
            raise TypeError("%s takes %d arguments (%d given)" % (name, nargs, len(args)))
        _send(name, *args)
    return f


def _getter(name, nargs, nreplies=1):
    def f(*args):
        if len(args) != nargs:
            raise TypeError("%s takes %d arguments (%d given)" % (name, nargs, len(args)))
        _send(name, *args)
        vals = tuple(_recv() for _ in range(nreplies))
        return vals[0] if nreplies == 1 else vals
    return f


# Python name: (command sent, number of arguments); lib/mumax.py:24-398
_COMMANDS = {
    "msat": ("msat", 1), "aexch": ("aexch", 1), "alpha": ("alpha", 1), "k1": ("k1", 1), "k2": ("k2", 1),
    "anisUniaxial": ("anisuniaxial", 3), "anisCubic": ("aniscubic", 6),
    "spinPolarization": ("spinpolarization", 1), "xi": ("xi", 1), "temperature": ("temperature", 1),
    "gridSize": ("gridsize", 3), "partSize": ("partsize", 3), "cellSize": ("cellsize", 3),
    "maxCellSize": ("maxcellsize", 3), "periodic": ("periodic", 3), "ellipsoid": ("ellipsoid", 3),
    "mask": ("mask", 1), "setMsat": ("setmsat", 4), "setMsatRange": ("SetMsatRange", 7),
    "setMsatEllips": ("SetMsatEllips", 5), "setAlpha": ("setalpha", 4), "setAlphaRange": ("SetAlphaRange", 7),
    "loadm": ("loadm", 1), "uniform": ("uniform", 3), "addNoise": ("addnoise", 1), "setRandom": ("setrandom", 0),
    "vortex": ("vortex", 2), "setVortexEllips": ("SetVortexEllips", 6),
    "vortexInArray": ("VortexInArray", 8), "dotArray": ("DotArray", 3),
    "dotArrayRectangle": ("DotArrayRectangle", 6), "dotArrayEllips": ("DotArrayEllips", 6),
    "antiDotArrayRectangle": ("AntiDotArrayRectangle", 6), "antiDotArrayEllips": ("AntiDotArrayEllips", 6),
    "setmCell": ("setmcell", 6), "setmRange": ("setmrange", 9), "seed": ("seed", 1),
    "save": ("save", 2), "savem": ("savem", 2), "saveh": ("saveh", 2), "autosave": ("autosave", 3),
    "tabulate": ("tabulate", 2), "subsampleOutput": ("subsampleoutput", 1),
    "solvertype": ("solvertype", 1), "maxError": ("maxerror", 1), "maxdt": ("maxdt", 1), "mindt": ("mindt", 1),
    "maxdm": ("maxdm", 1), "mindm": ("mindm", 1),
    "applyPointwise": ("applypointwise", 5), "currentMask": ("currentmask", 1), "fieldMask": ("fieldmask", 1),
    "applyStatic": ("applystatic", 4), "applyRF": ("applyrf", 5),
    "relax": ("relax", 0), "run": ("run", 1), "step": ("step", 0), "steps": ("steps", 1),
    "desc": ("desc", 2), "saveBenchmark": ("savebenchmark", 1),
    "exchInConv": ("exchinconv", 1), "exchType": ("exchtype", 1), "kernelType": ("kerneltype", 1),
    "demag": ("demag", 1), "energy": ("energy", 1),
}

# getters: (command sent, number of arguments, number of replies); lib/mumax.py:401-445
_GETTERS = {
    "getm": ("getm", 1, 1), "getmPos": ("getmPos", 4, 1), "getMaxm": ("getmaxm", 1, 1),
    "getMinm": ("getminm", 1, 1), "getCorepos": ("getcorepos", 0, 2), "getMaxTorque": ("getmaxtorque", 0, 1),
    "getE": ("getE", 0, 1), "getTime": ("getTime", 0, 1),
}

for _name, (_cmd, _n) in _COMMANDS.items():
    globals()[_name] = _command(_cmd, _n)
for _name, (_cmd, _n, _r) in _GETTERS.items():
    globals()[_name] = _getter(_cmd, _n, _r)


def applyFunction(what, func, duration, timestep):
    """Samples func(t) -> (bx, by, bz) every timestep up to duration as a
    pointwise field (lib/mumax.py:303-309)."""
    t = 0
    while t <= duration:
        bx, by, bz = func(t)
        applyPointwise(what, t, bx, by, bz)  # noqa: F821 (defined from the table above)
        t += timestep


__all__ = ["inf", "applyFunction"] + list(_COMMANDS) + list(_GETTERS)
