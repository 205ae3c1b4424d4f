"""OVF / .tensor / ODT readers and writers (C++ in libgpukern, host only).

Pinned by the reference's own fixture tests/golden/s-state.omf (the standard
problem 4 S-state, copied from test/stdprobl4/s-state.omf): its header, and the
torque its header records (9.903189e-05, the reference's maxabs(k0)) against
the FP64 oracle's torque of the state read back, which fails for any other
axis or component order (reversed components give ~0.23).
"""
import importlib
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "s-state.omf")


@pytest.fixture(scope="module")
def ovf():
    return importlib.import_module("1_amd.ovf")


def test_sp4_s_state_header_and_state(ovf, orc):
    m, meta = ovf.read_ovf(FIXTURE)
    assert meta["size"] == (1, 32, 128) and meta["version"] == 1 and meta["binary"] == 4
    assert np.allclose(meta["step"], (3e-9, 3.90625e-9, 3.90625e-9))
    assert meta["multiplier"] == 800000 and meta["valueunit"] == "A/m"
    assert meta["desc"]["msat"] == "800000" and meta["desc"]["aexch"] == "1.3e-11"
    assert np.abs(np.linalg.norm(m, axis=0) - 1).max() < 1e-6
    avg = m.reshape(3, -1).mean(axis=1)
    assert avg[2] > 0.9 and 0.05 < avg[1] < 0.2  # S-state: along user x with some user y
    o = orc.sim(m, size=meta["size"], cellsize=meta["step"], msat=8e5, aexch=1.3e-11, alpha=1.0, solver="rkdp",
                maxerror=1e-5)
    k = np.abs(o.torque(m.reshape(-1).astype(np.float64))).max()
    assert abs(k / float(meta["desc"]["torque"]) - 1) < 0.05


@pytest.mark.parametrize("version", [1, 2])
@pytest.mark.parametrize("binary", [True, False])
def test_ovf_round_trip(ovf, tmp_path, version, binary):
    rng = np.random.default_rng(3)
    m = rng.standard_normal((3, 2, 5, 7)).astype(np.float32)
    p = tmp_path / "m.ovf"
    ovf.write_ovf(p, m, (4e-9, 2e-9, 1e-9), version=version, binary=binary, desc={"alpha": 0.1, "note": "x"},
                  total_time=1e-9, stage_time=2e-10)
    back, meta = ovf.read_ovf(p)
    assert np.array_equal(back, m)
    assert meta["size"] == (2, 5, 7) and np.allclose(meta["step"], (4e-9, 2e-9, 1e-9))
    assert meta["version"] == version and meta["binary"] == (4 if binary else 0)
    assert meta["desc"]["alpha"] == "0.1" and meta["desc"]["note"] == "x"
    if version == 2:
        assert meta["total_time"] == pytest.approx(1e-9) and meta["stage_time"] == pytest.approx(2e-10)


def test_rewritten_fixture_has_identical_data_bytes(ovf, tmp_path):
    m, meta = ovf.read_ovf(FIXTURE)
    p = tmp_path / "s.omf"
    ovf.write_ovf(p, m, meta["step"], version=1, binary=True, multiplier=meta["multiplier"], desc=meta["desc"])
    def data(path):
        b = open(path, "rb").read()
        i = b.index(b"# Begin: Data Binary 4\n") + len(b"# Begin: Data Binary 4\n")
        return b[i:i + 4 * (1 + 3 * m[0].size)]
    assert data(p) == data(FIXTURE)


def test_tensor_round_trip_and_magic(ovf, tmp_path):
    rng = np.random.default_rng(4)
    t = rng.standard_normal((3, 1, 4, 6)).astype(np.float32)
    p = tmp_path / "m.tensor"
    ovf.write_tensor(p, t)
    raw = open(p, "rb").read()
    assert raw[:4] == b"#t1\n" and np.frombuffer(raw[4:24], "<i4").tolist() == [4, 3, 1, 4, 6]
    assert np.array_equal(ovf.read_tensor(p), t)


def test_odt_table(ovf, tmp_path):
    p = tmp_path / "t.odt"
    t = ovf.OdtTable(p, "mumax", [("t", "s"), ("mx", ""), ("my", "")])
    rows = [[0.0, 1.0, 0.0], [1e-12, 0.99, 0.1], [2e-12, 0.98, 0.2]]
    for r in rows:
        t.row(r)
    t.close()
    lines = open(p).read().splitlines()
    assert lines[0] == "# ODT 1.0" and lines[-1] == "# Table End"
    assert lines[3].startswith("# Units:") and "{s}" in lines[3]
    assert lines[4].startswith("# Columns:") and "mx" in lines[4]
    vals = np.array([[float(x) for x in l.split()] for l in lines[5:-1]])
    assert np.allclose(vals, rows)
