"""Command-line front end: the reference's input language (refsh), its Python
pipe protocol and its daemon, driving ``Sim``.

Input language (refsh/readline.go:17-110, refsh/resolve.go, refsh/parse.go):
a line is a command followed by its arguments; words are separated by
spaces, tabs or ':'; a line ends at a newline or ';'; '#' starts a comment
that runs to the end of the line.  Command names are case-insensitive and
are the reference's ``Sim`` method names (refsh.AddAllMethods,
main_slave.go:75), arguments in the order of those methods (user z, y, x for
vectors given to GridSize, CellSize, ApplyStatic, Uniform, ...).  Getters
answer with one line "%<value>" on the reply stream (simget.go:17-21).

Modes (main.go, main_master.go, main_slave.go, main_python.go, daemon.go):

  python -m 1_amd.frontend file.in          run a script; output in file.out/
  python -m 1_amd.frontend --stdin name.in  read the script from stdin
  python -m 1_amd.frontend file.py          run a Python script with the
        client module (1_amd/client/mumax.py, "from mumax import *") in its
        path, in the working directory file.out/; its stdout is this
        process's command stream and getter replies go to its stdin
  python -m 1_amd.frontend --daemon [dirs]  run every *.in / *.py whose
        .out directory does not exist yet, each in its own process
  python -m 1_amd.frontend --daemon --batch [dirs]  the same, but the .in
        scripts found together that are set-up commands, one run and saves
        run as one continuously batched workload (1_amd/serve.py)

Errors end the process like the reference's refsh with CrashOnError: an
unknown command exits with status 251 (-5), a wrong argument count with 255
(-1), an unparsable argument with 253 (-3), an invalid value (InputErr) with 1.
"""
import argparse
import importlib
import os
import random
import re
import subprocess
import sys
import time

import numpy as np

_sim = importlib.import_module(__package__ + ".sim") if __package__ else importlib.import_module("1_amd.sim")
InputErr = _sim.InputErr

EXIT_UNKNOWN_COMMAND = 251   # os.Exit(-5), refsh.go:147
EXIT_NARGS = 255             # os.Exit(-1), parse.go:30
EXIT_PARSE = 253             # os.Exit(-3), parse.go:68-125
EXIT_INPUT = 1
EXIT_SUBPROCESS = 4

KNOWN_EXTENSIONS = (".in", ".py")   # main_master.go:27


class ShellExit(Exception):
    """Ends the interpreter with an exit status (the reference calls os.Exit)."""

    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


# ------------------------------------------------------------------ tokenizer
def read_line(stream):
    """One line of words from a character stream (ReadLine, readline.go:17-44).
    Returns (words, eof); eof is True only at the end of input with no words."""
    words, cur = [], []
    while True:
        ch = stream.read(1)
        if ch == "#":                       # ReadCharNoComment: skip to the line end
            while ch not in ("\n", ""):
                ch = stream.read(1)
        if ch in ("", "\n", ";"):
            if cur:
                words.append("".join(cur))
            return words, (ch == "" and not words)
        if ch in (" ", "\t", ":"):
            if cur:
                words.append("".join(cur))
                cur = []
            continue
        cur.append(ch)


def read_nonempty_line(stream):
    line, eof = read_line(stream)
    while not line and not eof:
        line, eof = read_line(stream)
    return line, eof


# ------------------------------------------------------------------ argument parsing (parse.go)
_INT = re.compile(r"[+-]?[0-9]+\Z")
_BOOLS = {"1": True, "t": True, "T": True, "TRUE": True, "true": True, "True": True,
          "0": False, "f": False, "F": False, "FALSE": False, "false": False, "False": False}


def _parse_float(s):
    if s == "inf":
        return float("inf")
    if s == "-inf":
        return float("-inf")
    s = s.lower()
    if "_" in s:
        raise ValueError(s)
    return float(s)


def parse_arg(arg, kind):
    """Value of one argument: 'i' int, 'l' int64, 'f' float32, 'd' float64,
    'b' bool, 's' string (parseArg, parse.go:44-66)."""
    try:
        if kind in "il":
            if not _INT.match(arg):
                raise ValueError(arg)
            return int(arg)
        if kind == "f":
            return float(np.float32(_parse_float(arg)))
        if kind == "d":
            return _parse_float(arg)
        if kind == "b":
            return _BOOLS[arg]
        return arg
    except (ValueError, KeyError, OverflowError):
        name = {"i": "int", "l": "int64", "f": "float32", "d": "float64", "b": "bool"}[kind]
        raise ShellExit(EXIT_PARSE, "Could not parse to %s: %s" % (name, arg))


# ------------------------------------------------------------------ reply formatting
def go_float(v, bits=32):
    """A float as Go's fmt.Println prints it (shortest digits that read back
    as the same float32/float64; %e form below 1e-4 and from 1e6 on)."""
    x = np.float32(v) if bits == 32 else np.float64(v)
    if np.isnan(x):
        return "NaN"
    if np.isinf(x):
        return "+Inf" if x > 0 else "-Inf"
    if x == 0:
        return "-0" if np.signbit(x) else "0"
    sci = np.format_float_scientific(x, unique=True, trim="-", exp_digits=2)
    e = int(sci.split("e")[1])
    if e < -4 or e >= 6:
        return sci
    return np.format_float_positional(x, unique=True, trim="-")


# ------------------------------------------------------------------ commands
# refsh command name -> (Sim method or Shell handler, argument kinds); the
# names and argument orders of the reference's Sim methods (sim*.go).
COMMANDS = {
    # material (simmaterial.go, simanis.go)
    "msat": ("msat", "f"), "aexch": ("aexch", "f"), "alpha": ("alpha", "f"), "k1": ("k1", "f"),
    "temperature": ("temperature", "f"),
    "anisuniaxial": ("anisuniaxial", "fff"),
    # geometry (simgeom.go)
    "gridsize": ("gridsize", "iii"), "periodic": ("periodic", "iii"), "cellsize": ("cellsize", "fff"), "partsize": ("partsize", "fff"),
    "demagaccuracy": ("demagaccuracy", "i"), "exchinconv": ("exchinconv", "b"), "exchtype": ("exchtype", "i"),
    "demag": ("_demag", "b"), "energy": ("_energy", "b"), "kerneltype": ("kerneltype", "s"),
    "setalpha": ("setalpha", "iiif"), "setalpharange": ("setalpharange", "iiiiiif"),
    "setmsat": ("setmsat", "iiif"), "setmsatrange": ("setmsatrange", "iiiiiif"),
    "setmsatellips": ("setmsatellips", "fffff"),
    # spin-transfer torque (simspintorque.go)
    "currentdensity": ("currentdensity", "fff"), "spinpolarization": ("spinpolarization", "f"),
    "xi": ("xi", "f"), "currentmask": ("currentmask", "s"),
    # excitation (simfield.go)
    "applystatic": ("applystatic", "sfff"), "applypointwise": ("applypointwise", "sdfff"),
    "applypulse": ("applypulse", "sfffdd"), "applyrf": ("applyrf", "sfffd"), "applyrframp": ("applyrframp", "sfffdd"),
    "applysawtooth": ("applysawtooth", "sfffd"), "applyrotating": ("applyrotating", "sfffdddd"),
    # solver (simsolver.go)
    "solvertype": ("solvertype", "s"), "maxerror": ("maxerror", "f"), "dt": ("dt", "f"),
    "mindm": ("mindm", "f"), "maxdm": ("maxdm", "f"), "mindt": ("mindt", "f"), "maxdt": ("maxdt", "f"),
    "relaxmaxtorque": ("_relaxmaxtorque", "f"),
    # initial magnetisation (simconfig.go)
    "uniform": ("uniform", "fff"), "setmcell": ("setmcell", "iiifff"), "setmrange": ("setmrange", "iiiiiifff"),
    "vortex": ("vortex", "ii"), "setrandom": ("_setrandom", ""), "addnoise": ("_addnoise", "f"),
    "sbw": ("sbw", ""), "snw": ("snw", ""), "abw": ("abw", ""), "anw": ("anw", ""),
    "setvortexellips": ("setvortexellips", "ffffii"), "setvortexrectangle": ("setvortexrectangle", "ffffii"),
    "vortexinarray": ("vortexinarray", "ffffffii"),
    "randomseed": ("_seed", "l"), "seed": ("_seed", "l"), "loadm": ("loadm", "s"),
    # run (simrun.go, relax.go, simhysteresis.go)
    "run": ("run", "d"), "step": ("step", ""), "steps": ("steps", "i"), "relax": ("relax", ""),
    "hysteresis": ("_hysteresis", "ffffffi"),
    # output (simoutput.go, simdebug.go, simmetadata.go)
    "save": ("save", "ss"), "autosave": ("autosave", "ssf"), "tabulate": ("tabulate", "sb"),
    "savem": ("_savem", "ss"), "saveh": ("_saveh", "ss"), "desc": ("desc", "ss"),
    "subsampleoutput": ("_subsample", "i"),
    # getters (simget.go)
    "getm": ("_getm", "i"), "getmpos": ("_getmpos", "iiii"), "getmaxm": ("_getmaxm", "i"),
    "getminm": ("_getminm", "i"), "getmaxtorque": ("_getmaxtorque", ""), "getcorepos": ("_getcorepos", ""),
    "gete": ("_gete", ""), "gettime": ("_gettime", ""),
    # refsh built-in (refsh.go:209-218)
    "include": ("_include", "s"),
}

# Commands of the reference whose physics or geometry this engine does not
# implement; they stop the script with an InputErr naming the command.
UNSUPPORTED = (
    "k2", "aniscubic", "fieldmask",
    "maxcellsize", "ellipsoid", "mask", "dotarray", "dotarrayrectangle", "dotarrayellips", "antidotarrayrectangle",
    "antidotarrayellips", "squareholearray", "vortex_in_array", "setm", "applyrotatingburst", "edgecorrection",
    "savebenchmark", "exchangetype",
)


class Shell:
    """Executes commands on one ``Sim`` (refsh.Refsh with the Sim's methods).

    `reply` receives the getters' "%value" lines; `log` the echo of every
    command (refsh.go:141) unless it is None."""

    def __init__(self, outdir, reply=None, log=None, mode="fused"):
        self.sim = _sim.Sim(mode)
        self.sim.outputdir(outdir)
        self.outdir = outdir
        self.reply = reply if reply is not None else sys.stdout
        self.log = log
        self.calls = 0
        self.ran = False

    # ---------------------------------------------------------- dispatch (refsh.go:133-160)
    def call(self, fname, argv):
        if self.log is not None:
            self.log.write(">>>  %s \t %s\n" % (fname, argv))
            self.log.flush()
        self.calls += 1
        name = fname.lower()
        if name not in COMMANDS:
            if name in UNSUPPORTED:
                raise InputErr("%s is not supported by this engine" % fname)
            raise ShellExit(EXIT_UNKNOWN_COMMAND, 'Unknown command: "%s". Options are: %s'
                            % (fname, sorted(COMMANDS)))
        target, kinds = COMMANDS[name]
        if len(argv) != len(kinds):
            raise ShellExit(EXIT_NARGS, "Error calling %s %s : needs %d arguments." % (fname, argv, len(kinds)))
        args = [parse_arg(a, k) for a, k in zip(argv, kinds)]
        fn = getattr(self, target) if target.startswith("_") else getattr(self.sim, target)
        if name in ("run", "step", "steps", "relax", "hysteresis"):
            self.ran = True
        return fn(*args)

    def execute(self, stream):
        """Every command of the stream (Refsh.Exec)."""
        line, eof = read_nonempty_line(stream)
        while not eof:
            if line:
                self.call(line[0], line[1:])
            line, eof = read_nonempty_line(stream)

    def close(self):
        self.sim.close_table()
        self.sim.free()

    # ---------------------------------------------------------- replies
    def _send(self, *vals):
        self.reply.write("%" + " ".join(vals) + "\n")
        self.reply.flush()

    def _getm(self, c):
        self._send(go_float(self.sim.getm(c)))

    def _getmpos(self, c, z, y, x):
        self._send(go_float(self.sim.getmpos(c, z, y, x)))

    def _getmaxm(self, c):
        self._send(go_float(self.sim.getmaxm(c)))

    def _getminm(self, c):
        self._send(go_float(self.sim.getminm(c)))

    def _getmaxtorque(self):
        self._send(go_float(self.sim.getmaxtorque()))

    def _getcorepos(self):
        x, y = self.sim.getcorepos()
        self._send(go_float(x))
        self._send(go_float(y))

    def _gete(self):
        self._send(go_float(self.sim.gete()))

    def _gettime(self):
        self._send(go_float(self.sim.gettime(), 64))

    # ---------------------------------------------------------- commands with front-end state
    def _seed(self, s):
        # RandomSeed (simconfig.go:325-327): seeds the generator of setrandom / addnoise
        self._rng = np.random.default_rng(s)

    def _rng_get(self):
        if getattr(self, "_rng", None) is None:
            self._rng = np.random.default_rng(0)
        return self._rng

    def _setrandom(self):
        self.sim.setrandom(self._rng_get())

    def _addnoise(self, a):
        self.sim.addnoise(a, self._rng_get())

    def _hysteresis(self, bz0, by0, bx0, bz1, by1, bx1, steps):
        self.sim.hysteresis(bz0, by0, bx0, bz1, by1, bx1, steps)

    def _save_named(self, fname, format, data):
        # saveOmf (simoutput.go:414-426): OVF 1.0, multiplier 1, unit "Msat", to
        # the file name as given (relative to the working directory)
        from . import ovf

        kinds = {"binary": True, "text": False}
        if format not in kinds:
            raise InputErr("unknown format %s. options are: binary, text" % format)
        ovf.write_ovf(fname, data, self.sim._cell, version=1, binary=kinds[format], multiplier=1.0,
                      valueunit="Msat", desc=self.sim._ovf_desc())

    def _savem(self, fname, format):
        # SaveM (simdebug.go:26-30): m in units of Msat
        self.sim._init()
        self._save_named(fname, format, self.sim.m)

    def _saveh(self, fname, format):
        # SaveH (simdebug.go:18-23): the effective field of the current m, in units of Msat
        self.sim._init()
        self._save_named(fname, format, self.sim.heff(self.sim.m))

    def _demag(self, on):
        if not on:
            raise InputErr("demag false is not supported: the demagnetising field is always computed")

    def _energy(self, on):
        self.sim._want_energy = bool(on)

    def _relaxmaxtorque(self, v):
        # RelaxMaxTorque (simhysteresis.go:29-31) sets a value Relax() does not read
        pass

    def _subsample(self, factor):
        if factor != 1:
            raise InputErr("subsampleoutput %d is not supported (only 1)" % factor)

    def _include(self, path):
        with open(path) as f:
            self.execute(f)


# ------------------------------------------------------------------ modes
def out_dir(infile):
    return os.path.splitext(infile)[0] + ".out"


# commands that need the engine or the step loop: a script is served in a batch
# only if none of them comes before its one `run`
_STEPPING = {"run", "step", "steps", "relax", "hysteresis", "autosave", "save", "savem", "include", "getm",
             "getmpos", "getmaxm", "getminm", "getmaxtorque", "getcorepos", "gete", "gettime"}


def read_script(stream):
    lines = []
    line, eof = read_nonempty_line(stream)
    while not eof:
        if line:
            lines.append(line)
        line, eof = read_nonempty_line(stream)
    return lines


def batch_plan(lines):
    """(set-up lines, run duration in s, save lines) when the script is set-up
    commands, one `run` and then only `save` commands -- the shape the batched
    server can run without stopping the batch -- else None."""
    names = [ln[0].lower() for ln in lines]
    if names.count("run") != 1:
        return None
    i = names.index("run")
    if any(n in _STEPPING for n in names[:i]) or any(n != "save" for n in names[i + 1:]):
        return None
    if len(lines[i]) != 2:
        return None
    try:
        duration = parse_arg(lines[i][1], "d")
    except ShellExit:
        return None
    return lines[:i], duration, lines[i + 1:]


def _finish_out(outdir):
    running = os.path.join(outdir, "running")
    if os.path.exists(running):
        os.replace(running, os.path.join(outdir, "finished"))


def _start_out(outdir):
    os.makedirs(outdir, exist_ok=True)
    with open(os.path.join(outdir, "running"), "w") as f:
        f.write("This simulation was started on:\n %s\nThis file is renamed to \"finished\" when the simulation "
                "is ready.\n" % time.ctime())


class _NullStream:
    """A reply stream that drops what it is given (batched scripts have no getters)."""

    def write(self, s):
        return len(s)

    def flush(self):
        pass


def serve_scripts(infiles, mode="fused", silent=True, poll=16):
    """Runs .in scripts of the batchable shape (batch_plan) together through the
    continuous-batching server (serve.Server): all set-ups first, every `run`
    advanced in the batches of its size group, each script's saves as soon as
    its own run has finished.  Every script writes what it writes when run
    alone (a simulation computes the same numbers alone and in any mix).
    Returns {infile: exit status}; scripts of another shape are not run and
    are missing from the result."""
    from . import serve

    reqs, status = [], {}
    for k, f in enumerate(infiles):
        with open(f) as fh:
            plan = batch_plan(read_script(fh))
        if plan is None:
            continue
        setup, duration, saves = plan
        outdir = out_dir(f)
        _start_out(outdir)
        shell = Shell(outdir, reply=_NullStream(), log=None if silent else sys.stderr, mode=mode)
        try:
            for ln in setup:
                shell.call(ln[0], ln[1:])
            shell.sim.params()
            if shell.sim._m0 is None:
                raise InputErr("Initial magnetization was not set")
        except (ShellExit, InputErr) as e:
            sys.stderr.write("%s: %s\n" % (f, e))
            status[f] = e.status if isinstance(e, ShellExit) else EXIT_INPUT
            shell.close()
            continue

        def done(_request, shell=shell, saves=saves, f=f):
            try:
                shell.ran = True
                for ln in saves:
                    shell.call(ln[0], ln[1:])
                status[f] = 0
            except (ShellExit, InputErr) as e:
                sys.stderr.write("%s: %s\n" % (f, e))
                status[f] = e.status if isinstance(e, ShellExit) else EXIT_INPUT
            shell.sim.close_table()
            if status[f] == 0:
                _finish_out(out_dir(f))

        cfg = dict(size=shell.sim._size, duration=duration)
        reqs.append(serve.Request(k, 0.0, cfg, make=lambda shell=shell: shell.sim, on_finish=done))
    if reqs:
        serve.Server(poll=poll).serve(reqs)
    return status



def run_script(infile, stream=None, reply=None, mode="fused", silent=False):
    """The slave's work for one input (main_slave.go:14-96): output directory
    infile.out with a "running" file renamed to "finished" at the end.
    Returns the exit status."""
    outdir = out_dir(infile)
    os.makedirs(outdir, exist_ok=True)
    running = os.path.join(outdir, "running")
    with open(running, "w") as f:
        f.write("This simulation was started on:\n %s\nThis file is renamed to \"finished\" when the simulation "
                "is ready.\n" % time.ctime())
    shell = Shell(outdir, reply=reply, log=None if silent else sys.stderr, mode=mode)
    own = stream is None
    if own:
        stream = open(infile)
    try:
        shell.execute(stream)
    except ShellExit as e:
        sys.stderr.write("%s\n" % e)
        return e.status
    except InputErr as e:
        sys.stderr.write("InputErr: %s\n" % e)
        return EXIT_INPUT
    finally:
        if own:
            stream.close()
        shell.close()
    os.replace(running, os.path.join(outdir, "finished"))
    if shell.calls == 0:
        sys.stderr.write("Input contained no commands.\n")
    elif not shell.ran:
        sys.stderr.write("Input contained no commands to make the simulation run (e.g., \"run\").\n")
    return 0


def client_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "client")


def run_python(infile, mode="fused", silent=False):
    """main_python.go:18-90: the script runs as a child process in infile.out/
    with the client module importable; its stdout is the command stream, the
    getters' replies go to its stdin.  Returns the exit status."""
    outdir = out_dir(infile)
    os.makedirs(outdir, exist_ok=True)
    env = dict(os.environ)
    env["PYTHONPATH"] = client_path() + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
    env["PYTHONUNBUFFERED"] = "1"
    script = os.path.relpath(os.path.abspath(infile), os.path.abspath(outdir))
    proc = subprocess.Popen([sys.executable, script], cwd=outdir, env=env, stdin=subprocess.PIPE,
                            stdout=subprocess.PIPE, text=True, bufsize=1)
    try:
        status = run_script(infile, stream=proc.stdout, reply=proc.stdin, mode=mode, silent=silent)
    except BaseException:
        proc.kill()
        raise
    if status != 0:
        proc.kill()                          # main_python.go:70-80: end the sibling
        proc.wait()
        return status
    try:
        proc.stdin.close()
    except BrokenPipeError:
        pass
    rc = proc.wait()
    if rc != 0:
        sys.stderr.write("python exited with status %d\n" % rc)
        return EXIT_SUBPROCESS
    return 0


def find_input(dirs, rand=False):
    """An input file with a known extension whose .out directory does not exist
    (daemon.go:110-160)."""
    found = find_inputs(dirs, rand, first=True)
    return found[0] if found else None


def find_inputs(dirs, rand=False, first=False):
    out = []
    for d in dirs:
        try:
            names = os.listdir(d)
        except OSError as e:
            sys.stderr.write("%s\n" % e)
            continue
        names = sorted(names)
        if rand:
            random.shuffle(names)
        have = set(names)
        for n in names:
            base, ext = os.path.splitext(n)
            if ext in KNOWN_EXTENSIONS and base + ".out" not in have:
                out.append(os.path.join(d, n))
                if first:
                    return out
    return out


def _run_child(f, argv_extra):
    print("[daemon] Starting simulation: %s" % f, file=sys.stderr)
    cmd = [sys.executable, "-m", __package__ + ".frontend", *argv_extra, f]
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ)
    env["PYTHONPATH"] = root + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
    rc = subprocess.call(cmd, env=env)
    print("[daemon] Finished simulation %s (status %d)" % (f, rc), file=sys.stderr)
    return rc


def daemon(dirs, watch=60, walltime=0, rand=False, argv_extra=(), sleep=time.sleep, batch=False, mode="fused"):
    """DaemonMain (daemon.go:29-100): claim an input by creating its .out
    directory, run it in a child process, repeat; with watch == 0 exit when no
    input is left, else look again every `watch` seconds; walltime > 0 hours
    ends the daemon (a running simulation is not aborted).

    With batch, every input found in one look is claimed at once; the .in
    scripts of the batchable shape (batch_plan) run together in this process
    through the continuous-batching server, the others in child processes as
    above."""
    start = time.time()
    dirs = list(dirs) or ["."]
    sleeping = False
    while True:
        if walltime > 0 and time.time() - start > walltime * 3600:
            print("[daemon] Reached maximum walltime: exiting", file=sys.stderr)
            return 0
        if batch:
            claimed = []
            for f in find_inputs(dirs, rand):
                try:
                    os.mkdir(out_dir(f))
                    claimed.append(f)
                except OSError as e:
                    sys.stderr.write("%s\n" % e)
            if claimed:
                sleeping = False
                ins = [f for f in claimed if f.endswith(".in")]
                print("[daemon] Serving %d input(s) in batches" % len(ins), file=sys.stderr)
                served = serve_scripts(ins, mode=mode)
                for f, rc in sorted(served.items()):
                    print("[daemon] Finished simulation %s (status %d)" % (f, rc), file=sys.stderr)
                for f in claimed:
                    if f not in served:
                        _run_child(f, argv_extra)
                continue
            f = None
        else:
            f = find_input(dirs, rand)
        if f is None:
            if watch == 0:
                print("[daemon] No new input files and -watch=0: exiting", file=sys.stderr)
                return 0
            if not sleeping:
                print("[daemon] Looking for new input files every %d seconds" % watch, file=sys.stderr)
            sleeping = True
            sleep(watch)
            continue
        sleeping = False
        try:
            os.mkdir(out_dir(f))
        except OSError as e:
            sys.stderr.write("%s\n" % e)
            continue
        _run_child(f, argv_extra)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m 1_amd.frontend", description=__doc__.split("\n\n")[0])
    ap.add_argument("files", nargs="*")
    ap.add_argument("--stdin", action="store_true",
                    help="read the input from stdin; the file name only names the output directory")
    ap.add_argument("--slave", action="store_true", help="accepted for compatibility (this process is the slave)")
    ap.add_argument("--silent", action="store_true", help="do not echo the commands on stderr")
    ap.add_argument("--daemon", action="store_true", help="watch directories for new input files")
    ap.add_argument("--watch", type=int, default=60)
    ap.add_argument("--walltime", type=int, default=0)
    ap.add_argument("--random", action="store_true")
    ap.add_argument("--batch", action="store_true",
                    help="with --daemon: serve the inputs found together in batches (set-up, one run, saves)")
    ap.add_argument("--mode", choices=sorted(_sim.MODES), default="fused")
    a = ap.parse_args(argv)
    if a.daemon:
        extra = ["--mode", a.mode] + (["--silent"] if a.silent else [])
        return daemon(a.files, a.watch, a.walltime, a.random, extra, batch=a.batch, mode=a.mode)
    if not a.files:
        sys.stderr.write("no input files\n")
        return 255
    for f in a.files:
        ext = os.path.splitext(f)[1]
        if a.stdin:
            rc = run_script(f, stream=sys.stdin, reply=sys.stdout, mode=a.mode, silent=a.silent)
        elif ext == ".in":
            rc = run_script(f, mode=a.mode, silent=a.silent)
        elif ext == ".py":
            rc = run_python(f, mode=a.mode, silent=a.silent)
        else:
            sys.stderr.write("Unknown file format: %s (known: %s)\n" % (ext, ", ".join(KNOWN_EXTENSIONS)))
            return 2
        if rc != 0:
            return rc
    return 0


if __name__ == "__main__":
    sys.exit(main())
