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


def _stdlib_generator(tmp_path_factory):
    """Builds tests/cpu_harness/stdlib_generators.cpp with g++ (plain x86-64
    flags, like the reference bench's host code)."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("stdgen") / "stdlib_generators"
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_harness",
                       "stdlib_generators.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), src], check=True)
    return str(exe)


@pytest.fixture(scope="module")
def stdgen(tmp_path_factory):
    return _stdlib_generator(tmp_path_factory)


def _stdlib_values(exe, *args):
    import subprocess
    out = subprocess.run([exe, *map(str, args)], check=True, capture_output=True,
                         text=True).stdout.split()
    return np.array([int(x, 16) for x in out], dtype=np.uint64)


@pytest.mark.parametrize("mode", ["normal", "uniform", "toeplitz", "clustered"])
@pytest.mark.parametrize("n", [1, 2, 5, 1001, 4096])
def test_paper_modes_match_the_standard_library_bit_for_bit(stdgen, mode, n):
    """--data=normal|uniform|toeplitz|clustered are the reference bench's
    make_paper_tridiagonal_data inputs (tridiag_dc_bench.cu:257-308): the same
    bits as std::mt19937_64 with std::normal_distribution /
    std::uniform_real_distribution built with g++."""
    from brdc.paper_data import paper_data
    d, e = paper_data(mode, n)
    assert d.shape == (n,) and e.shape == (max(n - 1, 0),)
    np.testing.assert_array_equal(np.concatenate([d, e]).view(np.uint64),
                                  _stdlib_values(stdgen, mode, n))


def test_bench_and_test_modes_match_the_standard_library_bit_for_bit(stdgen):
    """--data=bench (std::mt19937 seeded 1234 + the size's index in --sizes,
    tridiag_dc_bench.cu:229-255, 1513) and --data=test (seed 1300 + n,
    311-335, 1510)."""
    from brdc.paper_data import bench_data, test_data
    for n, idx in ((1, 0), (6, 1), (3000, 2)):
        d, e = bench_data(n, idx)
        np.testing.assert_array_equal(np.concatenate([d, e]).view(np.uint64),
                                      _stdlib_values(stdgen, "bench", n, idx))
    for n in (1, 4, 150):
        d, e = test_data(n)
        np.testing.assert_array_equal(np.concatenate([d, e]).view(np.uint64),
                                      _stdlib_values(stdgen, "test", n))


def test_paper_modes_first_entries():
    """Closed forms of the deterministic modes and the leading draws of the
    random ones, written out."""
    from brdc.paper_data import paper_data
    d, e = paper_data("toeplitz", 4)
    assert d.tolist() == [2.0] * 4 and e.tolist() == [0.25] * 3
    d, e = paper_data("clustered", 3)
    assert d.tolist() == [1.0 - 1e-12, 1.0, 1.0 + 1e-12]
    assert e[0] == 1e-4 * (1.0 + 0.1 * np.cos(0.33)) and e[1] == 1e-4 * (1.0 + 0.1 * np.cos(0.66))
    d, e = paper_data("normal", 2)
    assert round(d[0], 8) == 0.23527131 and round(d[1], 8) == 1.30436513 and e[0] > 0
    d, e = paper_data("uniform", 2)
    assert round(d[0], 8) == 0.21609729 and round(d[1], 8) == -0.99298487
    assert 0.1 <= e[0] < 0.3


def test_harness_modes_are_distinct_from_splitmix_generators():
    """The reference's mode names give the reference's matrices; this
    package's generators carry their own splitmix- names."""
    from brdc.generators import generate, seed_for
    from brdc.paper_data import paper_data
    d_ref, _ = harness.make_input("normal", 300)
    d_pkg, _ = harness.make_input("splitmix-normal", 300)
    np.testing.assert_array_equal(d_ref, paper_data("normal", 300)[0])
    np.testing.assert_array_equal(d_pkg, generate("normal", 300, seed_for(2, 300))[0])
    assert not np.array_equal(d_ref, d_pkg)
    d0, _ = harness.make_input("bench", 500, 0)
    d1, _ = harness.make_input("bench", 500, 1)
    assert not np.array_equal(d0, d1)  # the bench seed follows the size's index


def test_error_columns():
    w = np.array([1.0, 2.0, 3.0])
    ref = np.array([1.0, 2.0, 3.0 + 1e-9])
    assert harness.rel_err_2norm(w, ref) == pytest.approx(1e-9 / np.sqrt(14.0 + 6e-9), rel=1e-6)
    assert harness.rel_err_max(w, ref) == pytest.approx(1e-9 / 3.0, rel=1e-6)


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


def test_autotune_search_keeps_the_start_unless_a_candidate_wins_by_the_margin():
    """The coordinate search of brdc.autotune: a candidate within the noise
    margin of the kept configuration does not replace it; a clear win does,
    and later knobs are searched from the new configuration."""
    from brdc import autotune

    def time_of(cfg):
        t = 10.0
        t -= {8: 0.0, 12: 0.05, 16: 0.5}[cfg["leaf_max"]]      # 16 wins by 5 %
        t -= 0.05 if cfg["stream_min_k"] == 1024 else 0.0    # 0.5 %: noise
        t -= 0.3 if cfg["wide_max_waves"] == 4 else 0.0      # 3 %: a win
        return t

    best, timings = autotune._search(time_of, (8, 12, 16), knobs=True)
    assert best["leaf_max"] == 16
    assert best["stream_min_k"] == autotune.KNOB_CANDIDATES["stream_min_k"][0]
    assert best["wide_max_waves"] == 4
    assert best["secular_lanes_log2"] == autotune.KNOB_CANDIDATES["secular_lanes_log2"][0]
    # every candidate was timed once, from the configuration kept at its turn
    assert len(timings) == 3 + sum(len(v) - 1 for v in autotune.KNOB_CANDIDATES.values())
    best0, _ = autotune._search(time_of, (8, 12, 16), knobs=True, margin=0.0)
    assert best0["stream_min_k"] == 1024
