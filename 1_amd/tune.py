"""Autotune the engine's launch shapes on this GPU and write the tuning table.

For each workload shape, builds the simulation, lets ``mm_sim_tune`` time every
candidate shape of each pass (Z forward rows per block, fused Z inverse rows
per block, blocks per CU of the Y, X.K.X and 2-D passes, and for the grids it
fits the register-resident plane pass in place of the Y, X.K.X and Y inverse
passes) and keeps the fastest.  The table (``tune_gfx950.txt`` beside this file by default) is read
when the library is loaded, so new simulations of a listed size run with the
tuned shapes.  The reference has no tuner: one launch rule (make1dconf,
``src/gpukern/gpu_conf.cu:68-97``) sizes every kernel.

    python 1_amd/tune.py [--reps 20] [--out PATH] [--report PATH]

Prints (and optionally writes) a markdown table of default vs tuned ms per
launch for every tunable.
"""
import argparse
import importlib
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

DESCR = {
    "zr": "Z forward rows per block (0 default, 1 half, 2 double)",
    "er": "fused Z inverse rows per block (0 default, 1 half)",
    "yb": "3-D Y passes, blocks per CU (both passes)",
    "xb": "X.K.X pass, blocks per CU",
    "y2b": "2-D Y.K.Y pass, blocks per CU",
    "pp": "3-D plans with padded Y 512, padded X 16/32: 1 = one register-resident plane pass for Y, X.K.X, Y "
          "inverse; 0 = three passes (ms: the passes it replaces)",
    "xv": "X.K.X variant, padded X 16-64: 0 default, 1 no prefetch of the next item, 2 half the elements per lane "
          "over twice the lanes, 3 both",
}


def shapes(mm):
    """The single-simulation workloads and a few more thick and thin shapes."""
    w = mm.workloads
    out = [("cuboid 256x256x16 RK4", w.cuboid_rk4()),
           ("film 128x32x1 RK45", w.film_sp4_like()),
           ("film 1024x1024x1 RK45", w.film_vortex_ramp())]
    for size in [(32, 128, 128), (8, 512, 512), (64, 64, 64)]:
        cfg = w.cuboid_rk4(size=size)
        out.append(("cuboid %dx%dx%d RK4" % (size[2], size[1], size[0]), cfg))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--out", default=os.path.join(HERE, "tune_gfx950.txt"))
    ap.add_argument("--report", default=None, help="also write the markdown report here")
    args = ap.parse_args()
    mm = importlib.import_module("1_amd")
    lib = mm.load()
    lib.mm_tune_clear()  # every shape starts from the defaults
    lines = ["| workload (internal X,Y,Z) | tunable | default | chosen | ms/launch default | ms/launch tuned | gain |",
             "|---|---|---|---|---|---|---|"]
    default = {"zr": 0, "er": 0, "yb": 8, "xb": 4, "y2b": 8, "pp": 0, "xv": 0}
    extra = ["", "Plane pass against the three passes it replaces (ms per evaluation, tuned shapes):", "",
             "| workload (internal X,Y,Z) | Y fwd + X.K.X + Y inv | plane pass |", "|---|---|---|"]
    for name, cfg in shapes(mm):
        s = mm.workloads.make_sim(cfg)
        res = s.tune(args.reps)
        s.free()
        for k, val in res.items():
            if k == "plane_vs_three":
                extra.append("| %s %s | %.5f | %.5f |" % (name, tuple(cfg["size"]), val[0], val[1]))
                continue
            v, md, mt = val
            lines.append("| %s %s | %s | %d | %d | %.5f | %.5f | %.1f%% |" % (
                name, tuple(cfg["size"]), k, default[k], v, md, mt, 100 * (md - mt) / md if md > 0 else 0))
        print("tune.py: %s done" % name, file=sys.stderr, flush=True)
    n = lib.mm_tune_save(args.out.encode())
    if n < 0:
        raise SystemExit("cannot write %s" % args.out)
    text = "\n".join(lines + (extra if len(extra) > 5 else []))
    print(text)
    print("\n%d entries written to %s" % (n, args.out))
    print("\n" + "\n".join("- `%s`: %s" % kv for kv in DESCR.items()))
    if args.report:
        with open(args.report, "w") as f:
            f.write(text + "\n")


if __name__ == "__main__":
    main()
