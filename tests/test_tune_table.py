"""The tuning table's file format (mm_tune_save / mm_tune_load): CPU only, no
kernel runs.  Lines are "Nx Ny Nz zr er yb xb y2b pp xv" (internal order; 8 or 9
columns read the missing ones as 0); comments and malformed lines are skipped,
and a variant that does not exist for the size (a row shape for its Z length,
the plane pass, an X.K.X variant for its padded X) falls back to the default 0."""
import ctypes as C
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "1_amd", "libgpukern.so")


def _lib():
    lib = C.CDLL(LIB)
    lib.mm_tune_load.argtypes = lib.mm_tune_save.argtypes = [C.c_char_p]
    return lib


def test_table_round_trip_and_validation(tmp_path):
    lib = _lib()
    src = tmp_path / "in.txt"
    src.write_text("# comment\n"
                   "16 256 256 2 1 6 3 8 1 2\n"    # Z length 512: both variants exist; plane pass and xv 2 fit
                   "1 32 2 2 1 4 4 12\n"           # Z length 4 (M = 2): one shape only -> 0 0; no pp, xv: 0
                   "not a line\n"
                   "4 4 4 0 0 0 4 8\n"             # zero blocks per CU: rejected
                   "1 1024 1024 1 0 8 4 16 1 3\n"  # 2-D: no plane pass, no X.K.X variant -> 0 0
                   "48 64 64 0 0 8 4 8 0 1\n")     # padded X 96: no X.K.X variant -> 0
    lib.mm_tune_clear()
    assert lib.mm_tune_load(str(src).encode()) == 4
    out = tmp_path / "out.txt"
    assert lib.mm_tune_save(str(out).encode()) == 4
    rows = [l.split() for l in out.read_text().splitlines() if not l.startswith("#")]
    assert rows == [["1", "32", "2", "0", "0", "4", "4", "12", "0", "0"],
                    ["1", "1024", "1024", "1", "0", "8", "4", "16", "0", "0"],
                    ["16", "256", "256", "2", "1", "6", "3", "8", "1", "2"],
                    ["48", "64", "64", "0", "0", "8", "4", "8", "0", "0"]]
    lib.mm_tune_clear()
    assert lib.mm_tune_save(str(out).encode()) == 0
    assert lib.mm_tune_load(str(tmp_path / "missing.txt").encode()) == -1


def test_committed_table_parses():
    table = os.path.join(ROOT, "1_amd", "tune_gfx950.txt")
    if not os.path.exists(table):
        return
    lib = _lib()
    lib.mm_tune_clear()
    n = sum(1 for l in open(table) if l.strip() and not l.startswith("#"))
    assert lib.mm_tune_load(table.encode()) == n
    lib.mm_tune_clear()
