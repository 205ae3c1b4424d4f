"""Front end: the input language (tokenizer, argument parsing, dispatch and
its exit statuses), the Python pipe protocol, the output schedule and the
daemon.  The CPU tests run no simulation (no command here creates the
engine); the GPU tests drive whole scripts and compare with ``Sim`` calls."""
import io
import os
import subprocess
import sys
import textwrap
from importlib import import_module

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def fe():
    return import_module("1_amd.frontend")


@pytest.fixture(scope="module")
def simmod():
    return import_module("1_amd.sim")


def _env():
    env = dict(os.environ)
    env["PYTHONPATH"] = ROOT + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
    return env


# ------------------------------------------------------------------ tokenizer (refsh/readline.go)
def test_read_line_words_comments_separators(fe):
    src = io.StringIO("msat 800E3 # a comment; not a command\n\n  gridsize\t64:32 1;run 1e-9\ninclude x")
    lines = []
    line, eof = fe.read_nonempty_line(src)
    while not eof:
        lines.append(line)
        line, eof = fe.read_nonempty_line(src)
    assert lines == [["msat", "800E3"], ["gridsize", "64", "32", "1"], ["run", "1e-9"], ["include", "x"]]


def test_read_line_eof_and_empty(fe):
    assert fe.read_line(io.StringIO("")) == ([], True)
    assert fe.read_line(io.StringIO("\n")) == ([], False)
    assert fe.read_line(io.StringIO("# only a comment")) == ([], True)
    assert fe.read_line(io.StringIO("step")) == (["step"], False)


# ------------------------------------------------------------------ arguments (refsh/parse.go)
def test_parse_args(fe):
    assert fe.parse_arg("42", "i") == 42 and fe.parse_arg("-3", "i") == -3
    assert fe.parse_arg("1E-9", "d") == 1e-9
    assert fe.parse_arg("inf", "f") == float("inf") and fe.parse_arg("-inf", "d") == float("-inf")
    assert fe.parse_arg("0.1", "f") == float(np.float32(0.1))
    assert fe.parse_arg("true", "b") is True and fe.parse_arg("F", "b") is False and fe.parse_arg("1", "b")
    assert fe.parse_arg("binary", "s") == "binary"
    for bad, kind in (("1.5", "i"), ("x", "f"), ("yes", "b"), ("1_0", "d")):
        with pytest.raises(fe.ShellExit) as e:
            fe.parse_arg(bad, kind)
        assert e.value.status == fe.EXIT_PARSE


def test_go_float_format(fe):
    """Replies print floats as Go's fmt.Println: shortest digits, %e below 1e-4
    and from 1e6 on, two-digit exponents, +Inf/-Inf/NaN."""
    cases = [(0.5, 32, "0.5"), (np.float32(0.1), 32, "0.1"), (1e-9, 64, "1e-09"), (1e6, 64, "1e+06"),
             (123456.0, 64, "123456"), (1.5e-5, 32, "1.5e-05"), (-0.25, 32, "-0.25"), (0.0001, 64, "0.0001"),
             (float("inf"), 32, "+Inf"), (float("-inf"), 64, "-Inf"), (float("nan"), 32, "NaN"), (0.0, 32, "0"),
             (2.5e-10, 64, "2.5e-10"), (np.float32(1 / 3), 32, "0.33333334")]
    for v, bits, want in cases:
        assert fe.go_float(v, bits) == want, (v, bits)
    for v in (np.float32(0.7), np.float32(-1.2345e-7), np.float32(3.4e38)):
        assert np.float32(float(fe.go_float(v))) == v


# ------------------------------------------------------------------ dispatch
def test_command_table_targets_exist(fe, simmod):
    for name, (target, kinds) in fe.COMMANDS.items():
        owner = fe.Shell if target.startswith("_") else simmod.Sim
        assert callable(getattr(owner, target, None)), name
        assert set(kinds) <= set("ilfdbs"), name
    assert not set(fe.COMMANDS) & set(fe.UNSUPPORTED)


def test_client_commands_resolve(fe):
    """Every command the client sends is a known command of the shell (or one
    it rejects by name), with the same number of arguments."""
    sys.path.insert(0, os.path.join(ROOT, "1_amd", "client"))
    try:
        client = import_module("mumax")
    finally:
        sys.path.pop(0)
    for table in (client._COMMANDS, client._GETTERS):
        for pyname, spec in table.items():
            cmd, nargs = spec[0].lower(), spec[1]
            assert cmd in fe.COMMANDS or cmd in fe.UNSUPPORTED, pyname
            if cmd in fe.COMMANDS:
                assert len(fe.COMMANDS[cmd][1]) == nargs, pyname


def test_shell_errors_and_exit_status(fe, tmp_path):
    sh = fe.Shell(str(tmp_path / "o.out"), reply=io.StringIO())
    sh.call("MSAT", ["800e3"])                   # case-insensitive
    sh.call("GridSize", ["64", "32", "1"])
    assert sh.sim._msat == float(np.float32(800e3)) and sh.sim._size == (1, 32, 64)
    with pytest.raises(fe.ShellExit) as e:
        sh.call("nosuchcommand", [])
    assert e.value.status == fe.EXIT_UNKNOWN_COMMAND
    with pytest.raises(fe.ShellExit) as e:
        sh.call("msat", [])
    assert e.value.status == fe.EXIT_NARGS
    with pytest.raises(fe.ShellExit) as e:
        sh.call("steps", ["1.5"])
    assert e.value.status == fe.EXIT_PARSE
    with pytest.raises(fe.InputErr, match="aniscubic"):
        sh.call("aniscubic", ["1e3", "0"])       # physics this engine leaves out
    sh.call("temperature", ["300"])              # Temperature before set-up selects RK12 (temperature.go:47-53)
    assert sh.sim._temperature == 300.0 and sh.sim._solver == "rk12"
    sh.call("periodic", ["3", "2", "0"])         # Periodic(z, y, x) in gridsize's order (simgeom.go:84-88)
    assert sh.sim._periodic == (0, 2, 3)
    with pytest.raises(fe.InputErr, match="images"):
        sh.call("periodic", ["-1", "0", "0"])
    with pytest.raises(fe.InputErr, match="output quantity"):
        sh.sim.autosave("h", "binary", 1e-12)
    sh.close()


def test_initial_state_commands(fe, tmp_path):
    """setmcell / setmrange / uniform / vortex / setrandom / addnoise in user
    index and component order (simconfig.go:20-95, 303-321)."""
    sh = fe.Shell(str(tmp_path / "o.out"), reply=io.StringIO())
    sh.call("gridsize", ["5", "4", "2"])        # user x=5, y=4, z=2 -> internal (2, 4, 5)
    sh.call("uniform", ["1", "0", "0"])
    m = sh.sim._m0
    assert m.shape == (3, 2, 4, 5) and np.all(m[2] == 1) and np.all(m[:2] == 0)
    sh.call("setmcell", ["4", "1", "0", "0.1", "0.2", "0.3"])   # cell user (x=4, y=1, z=0)
    assert tuple(m[:, 0, 1, 4]) == (np.float32(0.3), np.float32(0.2), np.float32(0.1))
    sh.call("setmrange", ["0", "0", "1", "1", "3", "1", "0", "0", "-1"])   # x 0..1, y 0..3, z = 1
    assert np.all(m[0, 1, :, :2] == -1) and np.all(m[2, 1, :, :2] == 0) and np.all(m[2, 1, :, 2:] == 1)
    sh.call("vortex", ["1", "-1"])
    cy, cz = 2, 2
    assert np.all(m[0, :, cy, cz] == -1) and np.all(m[1:, :, cy, cz] == 0)
    assert m[1, 0, 0, 4] == 2 and m[2, 0, 0, 4] == 2     # (x - cx) * c, -(y - cy) * c
    sh.call("setalpha", ["4", "1", "0", "2.5"])                   # cell user (x=4, y=1, z=0)
    sh.call("setalpharange", ["0", "0", "1", "1", "3", "1", "0.5"])  # x 0..1, y 0..3, z = 1
    am = sh.sim._alpha_map
    assert am.shape == (2, 4, 5) and am[0, 1, 4] == 2.5 and np.all(am[1, :, :2] == 0.5)
    assert (am == 1).sum() == am.size - 1 - 8
    sh.call("seed", ["7"])
    sh.call("setrandom", [])
    a = m.copy()
    assert a.min() >= -0.5 and a.max() < 0.5 and np.std(a) > 0.2
    sh.call("addnoise", ["0.1"])
    assert np.abs(m - a).max() <= 0.1 and np.abs(m - a).max() > 0.05
    sh.close()


def test_wall_and_vortex_array_states(fe, tmp_path):
    """SBW/SNW/ABW/ANW, SetVortexEllips/Rectangle, VortexInArray
    (simconfig.go:127-289) in user index order."""
    sh = fe.Shell(str(tmp_path / "o.out"), reply=io.StringIO())
    sh.call("gridsize", ["16", "8", "1"])
    sh.call("cellsize", ["2e-9", "2e-9", "2e-9"])
    sh.call("sbw", [])
    m = sh.sim._m0
    assert np.allclose(np.linalg.norm(m, axis=0), 1, atol=1e-6)
    assert np.allclose(m[0, 0, 3, :], np.sin(np.pi * (np.arange(16) - 8 + 0.5) / 16), atol=1e-7)
    assert np.abs(m[2]).max() == 0 and m[0, 0, 0, 0] < -0.99 and m[0, 0, 0, 15] > 0.99
    sh.call("snw", [])
    assert np.abs(sh.sim._m0[1]).max() < 1e-7
    sh.call("abw", [])
    m = sh.sim._m0
    assert np.allclose(m[2, 0, :, 8] / np.cos(np.pi * 0.5 / 16), np.sin(np.pi * (np.arange(8) - 4 + 0.5) / 8), atol=1e-6)
    sh.call("uniform", ["1", "0", "0"])
    sh.call("setvortexellips", ["16e-9", "8e-9", "12e-9", "8e-9", "1", "-1"])   # centre cell (8, 4), radii 3, 2
    m = sh.sim._m0
    assert tuple(m[:, 0, 4, 8]) == (-1, 0, 0)
    assert tuple(m[:, 0, 4, 11]) == (0, 3, 0) and tuple(m[:, 0, 6, 8]) == (0, 0, -2)
    assert tuple(m[:, 0, 6, 11]) == (0, 0, 1)              # outside the ellipse: untouched (+x)
    sh.call("setvortexrectangle", ["16e-9", "8e-9", "12e-9", "8e-9", "-1", "1"])
    assert tuple(sh.sim._m0[:, 0, 6, 11]) == (0, -3, 2)
    sh.call("uniform", ["0", "0", "1"])
    sh.call("vortexinarray", ["1", "0", "16e-9", "16e-9", "4e-9", "4e-9", "1", "1"])   # unit 8 x 8 cells, k 9..14
    m = sh.sim._m0
    assert tuple(m[:, 0, 3, 11]) == (1, 0, 0) and tuple(m[:, 0, 2, 9]) == (0, np.float32(9 - 11.5), np.float32(-(2 - 3.5)))
    assert tuple(m[:, 0, 2, 15]) == (1, 0, 0) and tuple(m[:, 0, 0, 9]) == (1, 0, 0)   # outside the unit
    with pytest.raises(fe.InputErr):
        sh.call("setvortexellips", ["2e-9", "8e-9", "12e-9", "8e-9", "1", "1"])
    sh.close()


def test_field_function_commands(fe, simmod, tmp_path):
    """applyrf / applypulse / applyrframp / applysawtooth / applyrotating take the
    reference's argument order (what, hz, hy, hx, ...) and replace the field;
    applystatic replaces them again."""
    sh = fe.Shell(str(tmp_path / "o.out"), reply=io.StringIO())
    sh.call("applyrf", ["field", "0.01", "0.02", "0.03", "1e9"])
    kind, b, prm = sh.sim._fn
    assert kind == 2 and b == tuple(float(np.float32(x)) for x in (0.03, 0.02, 0.01)) and prm[0] == 1e9
    assert sh.sim.field_at(0.25e-9) == pytest.approx(np.array(b), rel=1e-6)     # sin(pi/2) = 1
    sh.call("applypulse", ["field", "0", "0", "0.1", "2e-9", "1e-9"])
    assert sh.sim._fn[0] == 1 and sh.sim.field_at(2e-9)[0] == pytest.approx(float(np.float32(0.1)))
    sh.call("applyrotating", ["field", "0.1", "0.1", "0.1", "1e9", "0", "1.5707963267948966", "0"])
    assert sh.sim.field_at(0.0) == pytest.approx([0.0, float(np.float32(0.1)), 0.0], abs=1e-7)
    sh.call("applystatic", ["field", "0", "0", "0.5"])
    assert sh.sim._fn is None and sh.sim.field_at(1.0)[0] == 0.5
    with pytest.raises(fe.InputErr):
        sh.call("applyrf", ["current", "0", "0", "1", "1e9"])
    sh.close()


def test_output_schedule_rule(simmod):
    """Periodic.NeedSave (simoutput.go:80-82): due at t == 0 and when float32
    time - last save >= period."""
    out = simmod._resolve_output("m", "binary")
    out.period = np.float32(1e-12)
    assert out.need_save(np.float32(0))
    out.since = np.float32(3e-12)
    assert not out.need_save(np.float32(3.9e-12)) and out.need_save(np.float32(4e-12))
    assert out.due_internal(1.0) == pytest.approx(4e-12)
    assert simmod._resolve_output("table", "anything").kind == "table"
    with pytest.raises(simmod.InputErr):
        simmod._resolve_output("m", "png")


def test_corepos_interpolation(simmod):
    """interpolate_maxpos (corepos.go:36-40): the vertex of the parabola
    through (-1, f1), (0, f0), (1, f2)."""
    for a, x0, c in ((-1.0, 0.3, 2.0), (-0.5, -0.2, 1.0)):
        f = [np.float32(a * (x - x0) ** 2 + c) for x in (0, -1, 1)]
        assert simmod._parabola_peak(*f) == pytest.approx(x0, abs=1e-5)


# ------------------------------------------------------------------ script modes, CPU only
def test_run_script_without_run(fe, tmp_path, capsys):
    inp = tmp_path / "a.in"
    inp.write_text("msat 800e3\naexch 1.3e-11\ngridsize 64 32 1\n")
    assert fe.run_script(str(inp)) == 0
    assert (tmp_path / "a.out" / "finished").exists() and not (tmp_path / "a.out" / "running").exists()
    assert "no commands to make the simulation run" in capsys.readouterr().err
    bad = tmp_path / "b.in"
    bad.write_text("msat 800e3\nfoo 1\n")
    assert fe.run_script(str(bad)) == fe.EXIT_UNKNOWN_COMMAND
    assert (tmp_path / "b.out" / "running").exists()


def test_client_protocol(tmp_path):
    """The client writes one command line per call and reads '%' replies."""
    script = tmp_path / "c.py"
    script.write_text("from mumax import *\nimport sys\nmsat(800e3)\ngridSize(64, 32, 1)\n"
                      "v = getm(0)\nx, y = getCorepos()\nsys.stderr.write('%r %r %r' % (v, x, y))\n")
    env = _env()
    env["PYTHONPATH"] = os.path.join(ROOT, "1_amd", "client")
    r = subprocess.run([sys.executable, str(script)], input="%0.5\n%1e-09\n%-2.5e-08\n", capture_output=True,
                       text=True, env=env, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.splitlines() == ["msat 800000.0", "gridsize 64 32 1", "getm 0", "getcorepos"]
    assert r.stderr == "0.5 1e-09 -2.5e-08"
    r = subprocess.run([sys.executable, str(script)], input="no reply\n", capture_output=True, text=True, env=env,
                       timeout=60)
    assert r.returncode == 10


def test_python_mode_setters_only(fe, tmp_path):
    """A .py input runs in name.out/ with the client importable; its commands
    reach the shell (which here creates no engine)."""
    script = tmp_path / "p.py"
    script.write_text("from mumax import *\nimport os\nmsat(800e3)\naexch(1.3e-11)\ngridSize(8, 4, 1)\n"
                      "open('cwd.txt', 'w').write(os.getcwd())\n")
    assert fe.run_python(str(script), silent=True) == 0
    out = tmp_path / "p.out"
    assert (out / "finished").exists()
    assert os.path.samefile(open(out / "cwd.txt").read(), out)
    bad = tmp_path / "q.py"
    bad.write_text("from mumax import *\nmsat(800e3)\nsteps(1.5)\nimport time\ntime.sleep(30)\n")
    assert fe.run_python(str(bad), silent=True) == fe.EXIT_PARSE


def test_batch_plan(fe):
    """Scripts of set-up commands, one run and then only saves are batchable."""
    def plan(text):
        return fe.batch_plan(fe.read_script(io.StringIO(text)))

    p = plan("msat 8e5\ngridsize 8 8 1\nuniform 1 0 0\nrun 1e-9\nsave m binary\nsave table ascii\n")
    assert p is not None and p[1] == 1e-9 and [ln[0] for ln in p[0]] == ["msat", "gridsize", "uniform"]
    assert [ln[1] for ln in p[2]] == ["m", "table"]
    assert plan("msat 8e5\nrun 1e-9\n")[2] == []
    for bad in ("msat 8e5\n", "run 1e-9\nrun 1e-9\n", "autosave m binary 1e-12\nrun 1e-9\n",
                "run 1e-9\ngetm 0\n", "save m binary\nrun 1e-9\n", "run 1e-9\nsteps 3\n", "run x\n",
                "include a.in\nrun 1e-9\n"):
        assert plan(bad) is None, bad


def test_daemon_runs_new_inputs_once(tmp_path):
    (tmp_path / "a.in").write_text("msat 800e3\n")
    (tmp_path / "b.in").write_text("msat 800e3\n")
    (tmp_path / "b.out").mkdir()                 # already claimed: skipped
    (tmp_path / "c.txt").write_text("not an input")
    r = subprocess.run([sys.executable, "-m", "1_amd.frontend", "--daemon", "--watch", "0", "--silent",
                        str(tmp_path)], capture_output=True, text=True, env=_env(), timeout=120, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "a.out" / "finished").exists()
    assert not (tmp_path / "b.out" / "finished").exists()
    assert not (tmp_path / "c.out").exists()
    assert "Starting simulation: %s" % (tmp_path / "a.in") in r.stderr


# ------------------------------------------------------------------ GPU: whole scripts
FILM = """msat 800e3
aexch 1.3e-11
alpha 0.02
gridsize 128 32 1
cellsize 3.90625e-9 3.90625e-9 3e-9
solvertype rkdp
maxerror 1e-5
uniform 1 0.25 0.1
seed 3
addnoise 0.01
applystatic field -24.6e-3 4.3e-3 0
"""


def f32(*v):
    # the script's float32 arguments (refsh parse.go:115-125)
    return [float(np.float32(x)) for x in v]


def _film_sim(simmod, alpha=0.02, nx=128):
    s = simmod.Sim()
    s.msat(*f32(800e3)), s.aexch(*f32(1.3e-11)), s.alpha(*f32(alpha))
    s.gridsize(nx, 32, 1), s.cellsize(*f32(3.90625e-9, 3.90625e-9, 3e-9))
    s.solvertype("rkdp"), s.maxerror(*f32(1e-5))
    s.uniform(*f32(1, 0.25, 0.1))
    s.addnoise(*f32(0.01), np.random.default_rng(3))
    s.applystatic("field", *f32(-24.6e-3, 4.3e-3, 0))
    return s


@pytest.mark.gpu
def test_script_output_schedule_matches_stepwise(fe, simmod, tmp_path):
    """A script with autosave/tabulate: the device runs the steps between
    outputs in one stretch; the saved states, their ids and times are those
    of the reference's loop (check every output before every step,
    simrun.go:29-49), replayed here one step at a time."""
    inp = tmp_path / "film.in"
    inp.write_text(FILM + "tabulate e true\ntabulate maxdm/dt true\nautosave m binary 20e-12\n"
                   "autosave table ascii 5e-12\nrun 60e-12\nsave m omf\n")
    assert fe.run_script(str(inp), silent=True) == 0
    out = tmp_path / "film.out"
    ovf = import_module("1_amd.ovf")

    s = _film_sim(simmod)
    s._init()
    lib, ptr = s._lib, s._ptr
    st = s.state()
    stop = st["time"] + 60e-12 / st["unit_time"]
    lib.mm_sim_run(ptr, 0.0)
    outs = [["m", np.float32(20e-12), np.float32(0)], ["table", np.float32(5e-12), np.float32(0)]]
    idx, fresh, saved, rows = -1, False, {}, []
    while True:
        st = s.state()
        if not st["time"] < stop:
            break
        t32 = np.float32(np.float32(st["time"]) * np.float32(st["unit_time"]))
        for o in outs:
            if t32 == 0 or np.float32(t32 - o[2]) >= o[1]:
                if not fresh:
                    idx, fresh = idx + 1, True
                if o[0] == "m":
                    saved[idx] = s.m
                else:
                    rows.append((st["time"] * st["unit_time"], idx))
                o[2] = t32
        lib.mm_sim_steps(ptr, 1)
        fresh = False
    final = s.m
    idx += 1
    files = sorted(p.name for p in out.glob("m*.omf"))
    assert files == ["m%08d.omf" % k for k in sorted(saved)] + ["m%08d.omf" % idx]
    for k, m in saved.items():
        got, meta = ovf.read_ovf(out / ("m%08d.omf" % k))
        assert np.array_equal(got, m), k
        assert int(meta["desc"]["id"]) == k
    got, _ = ovf.read_ovf(out / ("m%08d.omf" % idx))
    assert np.array_equal(got, final)
    lines = open(out / "datatable.odt").read().splitlines()
    assert [c for c in lines[4].split(":")[1].split()] == ["Time", "Mx/Ms", "My/Ms", "Mz/Ms", "Bx", "By", "Bz",
                                                          "Energy", "max_dm/dt", "id"]
    table = [[float(x) for x in ln.split()] for ln in lines if not ln.startswith("#")]
    assert len(table) == len(rows) and len(rows) >= 12
    for r, (t, k) in zip(table, rows):
        assert r[0] == pytest.approx(t, rel=1e-9) and r[-1] == k
    s.free()


@pytest.mark.gpu
def test_python_script_getters(fe, simmod, tmp_path):
    """A Python input through the pipe protocol: every getter's reply equals
    the value read from a Sim given the same commands."""
    script = tmp_path / "vort.py"
    script.write_text(textwrap.dedent("""\
        from mumax import *
        import json
        msat(800e3); aexch(1.3e-11); alpha(0.5)
        gridSize(64, 64, 1); cellSize(4e-9, 4e-9, 4e-9)
        solvertype('rkdp'); maxError(1e-5)
        vortex(1, 1)
        run(20e-12)
        vals = dict(m=[getm(i) for i in range(3)], maxm=getMaxm(2), minm=getMinm(0), mpos=getmPos(1, 40, 20, 0),
                    torque=getMaxTorque(), e=getE(), t=getTime(), core=list(getCorepos()))
        json.dump(vals, open('vals.json', 'w'))
        """))
    assert fe.run_python(str(script), silent=True) == 0
    import json

    got = json.load(open(tmp_path / "vort.out" / "vals.json"))
    s = simmod.Sim()
    s.msat(*f32(800e3)), s.aexch(*f32(1.3e-11)), s.alpha(*f32(0.5))
    s.gridsize(64, 64, 1), s.cellsize(*f32(4e-9, 4e-9, 4e-9))
    s.solvertype("rkdp"), s.maxerror(*f32(1e-5))
    s.vortex(1, 1)
    s.run(20e-12)
    # float32 replies print the shortest digits of the float32 value
    assert f32(*got["m"]) == f32(*[s.getm(i) for i in range(3)])
    assert f32(got["maxm"], got["minm"]) == f32(s.getmaxm(2), s.getminm(0))
    assert f32(got["mpos"]) == f32(s.getmpos(1, 40, 20, 0))
    assert f32(got["torque"]) == f32(s.getmaxtorque()) and got["t"] == s.gettime()
    assert got["e"] == pytest.approx(s.gete(), rel=1e-6)
    assert f32(*got["core"]) == f32(*s.getcorepos())
    assert abs(got["core"][0]) < 2 * 4e-9 and abs(got["core"][1]) < 2 * 4e-9
    s.free()


@pytest.mark.gpu
def test_relax_with_autosave_same_steps(simmod, tmp_path):
    """Relax with scheduled output steps on the host (relax.go:27-43) and takes
    the same steps as the relax that runs on the device."""
    a = _film_sim(simmod, alpha=1.0, nx=64)
    a.run(0)                      # Run()'s normalisation (Relax does not normalise)
    na = a.relax()
    b = _film_sim(simmod, alpha=1.0, nx=64)
    b.outputdir(str(tmp_path))
    b.autosave("table", "ascii", 2e-12)
    b.run(0)
    nb = b.relax()
    b.close_table()
    assert na == nb and na > 10
    assert np.array_equal(a.m, b.m)
    rows = [ln for ln in open(tmp_path / "datatable.odt").read().splitlines() if not ln.startswith("#")]
    t = [float(r.split()[0]) for r in rows]
    assert len(t) >= 2 and all(np.diff(t) >= 2e-12 * (1 - 1e-6))
    print("relax: %d steps, %d table rows" % (na, len(t)))
    a.free(), b.free()


@pytest.mark.gpu
def test_daemon_batch_serves_scripts_like_serial(fe, tmp_path):
    """The daemon's batch mode runs the batchable scripts together through the
    continuous-batching server and writes, for each, what the script writes
    when run alone; a script of another shape runs in a child process."""
    def script(nx, ny, t, extra=""):
        return ("msat 800e3\naexch 1.3e-11\nalpha 0.1\ngridsize %d %d 1\ncellsize 3.90625e-9 3.90625e-9 3e-9\n"
                "solvertype rkdp\nmaxerror 1e-5\nuniform 1 0.25 0.1\nseed %d\naddnoise 0.02\n"
                "applystatic field -24.6e-3 4.3e-3 0\n%srun %g\nsave m binary\nsave table ascii\n"
                % (nx, ny, nx, extra, t))
    texts = {"a": script(64, 32, 40e-12), "b": script(128, 48, 25e-12), "c": script(40, 24, 60e-12),
             "d": script(32, 32, 10e-12, extra="tabulate e true\n")}
    serial, batched = tmp_path / "serial", tmp_path / "batched"
    serial.mkdir(), batched.mkdir()
    for k, txt in texts.items():
        (serial / (k + ".in")).write_text(txt)
        (batched / (k + ".in")).write_text(txt)
    (batched / "e.in").write_text(script(32, 16, 5e-12) + "getm 0\n")   # not batchable
    for k in texts:
        assert fe.run_script(str(serial / (k + ".in")), silent=True) == 0
    r = subprocess.run([sys.executable, "-m", "1_amd.frontend", "--daemon", "--batch", "--watch", "0", "--silent",
                        str(batched)], capture_output=True, text=True, env=_env(), timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr
    assert "Serving 5 input(s) in batches" in r.stderr
    ovf = import_module("1_amd.ovf")
    for k in texts:
        a, _ = ovf.read_ovf(serial / (k + ".out") / "m00000000.omf")
        b, _ = ovf.read_ovf(batched / (k + ".out") / "m00000000.omf")
        assert np.array_equal(a, b), k
        ta = open(serial / (k + ".out") / "datatable.odt").read()
        tb = open(batched / (k + ".out") / "datatable.odt").read()
        assert ta == tb, k
        assert (batched / (k + ".out") / "finished").exists()
    assert (batched / "e.out" / "finished").exists()
    assert "Starting simulation: %s" % (batched / "e.in") in r.stderr


@pytest.mark.gpu
def test_savem_and_saveh(fe, simmod, tmp_path):
    """SaveM / SaveH (simdebug.go:18-30): m and the effective field of the
    current m, OVF 1.0, to the file names as given."""
    inp = tmp_path / "f.in"
    mf, hf = tmp_path / "mfinal.omf", tmp_path / "hfinal.omf"
    inp.write_text(FILM + "run 5e-12\nsavem %s binary\nsaveh %s binary\n" % (mf, hf))
    assert fe.run_script(str(inp), silent=True) == 0
    ovf = import_module("1_amd.ovf")
    s = _film_sim(simmod)
    s.run(5e-12)
    m, meta = ovf.read_ovf(mf)
    assert np.array_equal(m, s.m) and meta["valueunit"] == "Msat" and meta["multiplier"] == 1.0
    h, _ = ovf.read_ovf(hf)
    assert np.array_equal(h, s.heff(s.m))
    s.free()
