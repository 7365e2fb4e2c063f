"""The reference-mirroring harness (ppopp27-brdc-ae_amd/brdc/harness.py) and
the .de.bin format on the CPU: the file the reference ships
(data/inputs/reduced_dense_49152.de.bin, kept as a fixture) parses, round
trips, and the oracle solves it to the DSTERF fixture made from it."""
import io
import os

import numpy as np
import pytest

from brdc import harness
from brdc.generators import read_de_bin, write_de_bin

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DEBIN = os.path.join(GOLD, "reduced_dense_49152.de.bin")


def test_de_bin_fixture_parses_and_round_trips(tmp_path):
    d, e = read_de_bin(DEBIN)
    assert d.shape == (49152,) and e.shape == (49151,)
    assert np.all(np.isfinite(d)) and np.all(np.isfinite(e))
    p = tmp_path / "x.de.bin"
    write_de_bin(str(p), d, e)
    assert p.read_bytes() == open(DEBIN, "rb").read()
    # a dense n*n payload after the header is also accepted (column-major)
    n = 5
    a = np.arange(n * n, dtype=np.float64).reshape(n, n)
    a = a + a.T
    q = tmp_path / "dense.bin"
    q.write_bytes(np.int64(n).tobytes() + np.asfortranarray(a).tobytes(order="F"))
    dd, ee = read_de_bin(str(q))
    np.testing.assert_array_equal(dd, np.diag(a))
    np.testing.assert_array_equal(ee, np.diag(a, 1))
    r = tmp_path / "bad.bin"
    r.write_bytes(np.int64(4).tobytes() + b"\0" * 16)
    with pytest.raises(ValueError):
        read_de_bin(str(r))


def test_oracle_solves_the_shipped_input(orc):
    d, e = read_de_bin(DEBIN)
    w, info = orc.dstebr(d, e)
    ref = np.load(os.path.join(GOLD, "reduced_dense_49152_dsterf.npz"))["w"]
    assert info == 0
    assert np.max(np.abs(w - ref)) / np.max(np.abs(ref)) <= 1e-10


def test_bench_mode_distribution():
    d, e = harness.bench_data(10000, 7)
    i = np.arange(10000)
    assert np.all(np.abs(d - (10000 + 0.25 * i)) <= 0.5)
    assert np.all((e >= 0.05) & (e <= 0.95))
    assert abs(np.mean(d - (10000 + 0.25 * i))) < 0.02 and abs(np.mean(e) - 0.5) < 0.02


def test_harness_cpu_column_and_table(tmp_path):
    out = io.StringIO()
    rows = harness.main(["--sizes=100,257", "--data=normal,bench", "--run-self=0"], out=out)
    lines = out.getvalue().splitlines()
    assert lines[0].split() == ["mode", "n", "prep_ms", "cpu_ms"]
    assert [r["mode"] for r in rows] == ["normal", "normal", "bench", "bench"]
    assert all(r["cpu_ms"] >= 0 and "cpu_info" not in r for r in rows)
    p = tmp_path / "o.de.bin"
    out = io.StringIO()
    harness.main([f"--input={DEBIN}", "--input-tridiag=1", "--run-cpu=0", "--run-self=0",
                  f"--output-tridiag={p}"], out=out)
    assert "wrote_tridiag=" in out.getvalue()
    assert p.read_bytes() == open(DEBIN, "rb").read()


def test_harness_argument_errors():
    with pytest.raises(SystemExit):
        harness.main(["--run-cpu=0", "--run-self=0"], out=io.StringIO())
    with pytest.raises(SystemExit):
        harness.main([f"--input={DEBIN}", "--run-self=0"], out=io.StringIO())  # dense path
    with pytest.raises(SystemExit):
        harness.main(["--data=nosuchmode", "--run-self=0"], out=io.StringIO())
