"""Command-line benchmark harness with the reference's flags and table
(gpu/bench/tridiag_dc_bench.cu: main, --sizes/--warmup/--iters/--data/--input
/--input-tridiag/--output-tridiag/--run-cpu/--run-self/--run-cusolver/
--tridiag-diagnostics), for the tridiagonal eigenvalue path only:

    python -m brdc.harness --sizes=4096,16384 --data=normal,toeplitz
    python -m brdc.harness --input=reduced_dense_49152.de.bin --input-tridiag=1

Columns: mode, n, prep_ms (input preparation), cpu_ms (LAPACK DSTERF on the
host, the reference's acceptance oracle; the reference puts its CPU solver
there), self_ms (this GPU solve, HIP-event timed, median of --iters after
--warmup), self_err (the reference's error column: the 2-norm relative
difference ||w - w_cpu|| / ||w_cpu||, tridiag_dc_bench.cu:747-757),
self_maxerr (max |w - w_cpu| / max |w_cpu|, the acceptance measure of
scripts/cpu/smoke_br.f90:32-33), rslv_ms (rocSOLVER STEDC values-only on the
same GPU, one run: the vendor column the reference fills with cuSOLVER),
rslv/self and rslv_err (2-norm).  --run-rocsolver replaces --run-cusolver
(also accepted).  Dense inputs (--input without --input-tridiag) need the
two-stage reduction, which this path does not include.

Data modes (--data, comma-separated, default bench):
* the reference bench's own inputs, bit for bit (brdc.paper_data):
  bench, test, normal, uniform, toeplitz, clustered;
* this package's splitmix64 generators (brdc.generators, seed
  seed_for(2, n)): splitmix-uniform, splitmix-normal, splitmix-toeplitz,
  splitmix-clustered.

--tridiag-diagnostics=1 prints, per merge level, the fields of the
reference's TridiagMergeLevelDiagnostics that this path has: merges, rows,
node sizes, non-deflated poles, zero (tiny-z) and close (Givens) deflations,
secular evaluations (sum, most for one root, failed roots) and the level's
milliseconds by kernel class.
"""
from __future__ import annotations

import sys
import time

import numpy as np

from . import paper_data
from .generators import GENERATORS, generate, read_de_bin, seed_for, write_de_bin

SPLITMIX_PREFIX = "splitmix-"
MODES = ("bench", "test") + paper_data.PAPER_MODES + tuple(SPLITMIX_PREFIX + g
                                                          for g in GENERATORS)


def _flag(argv, name, default=None):
    pre = name + "="
    for a in argv:
        if a.startswith(pre):
            return a[len(pre):]
    return default


def _int_flag(argv, name, default):
    v = _flag(argv, name)
    return default if v is None else int(v)


def make_input(mode: str, n: int, size_index: int = 0):
    """(D, E) of one --data mode at order n; size_index is n's position in
    --sizes (the bench mode's seed, tridiag_dc_bench.cu:1513)."""
    if mode == "bench":
        return paper_data.bench_data(n, size_index)
    if mode == "test":
        return paper_data.test_data(n)
    if mode in paper_data.PAPER_MODES:
        return paper_data.paper_data(mode, n)
    if mode.startswith(SPLITMIX_PREFIX) and mode[len(SPLITMIX_PREFIX):] in GENERATORS:
        return generate(mode[len(SPLITMIX_PREFIX):], n, seed_for(2, n))
    raise SystemExit(f"unknown --data mode {mode!r} (expected one of {', '.join(MODES)})")


def rel_err_2norm(w, ref):
    """||w - ref|| / ||ref|| (tridiag_dc_bench.cu:747-757)."""
    num = float(np.sum((w - ref) ** 2))
    den = float(np.sum(ref * ref))
    return float(np.sqrt(num / max(den, 1.0e-30)))


def rel_err_max(w, ref):
    return float(np.max(np.abs(w - ref)) / max(np.max(np.abs(ref)), 1e-300))


def _median_gpu_ms(fn, warmup, iters):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    times = []
    for _ in range(max(1, iters)):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        times.append(a.elapsed_time(b))
    return float(np.median(times))


def run_case(mode, d, e, prep_ms, opts, out=sys.stdout):
    row = {"mode": mode, "n": d.shape[0], "prep_ms": prep_ms}
    w_cpu = None
    if opts["run_cpu"]:
        # the host LAPACK's DSTERF (scipy's OpenBLAS), the reference's
        # acceptance oracle (scripts/cpu/smoke_br.f90:30)
        from scipy.linalg.lapack import dsterf
        t0 = time.perf_counter()
        w_cpu, info = dsterf(np.array(d, copy=True), np.array(e, copy=True))[:2]
        row["cpu_ms"] = (time.perf_counter() - t0) * 1e3
        if info:
            row["cpu_info"] = int(info)
    w_self = None
    if opts["run_self"]:
        import torch
        from . import Solver
        dev = torch.device("cuda:0")
        n = d.shape[0]
        dt = torch.from_numpy(np.ascontiguousarray(d)).to(dev)
        et = torch.from_numpy(np.ascontiguousarray(e if n > 1 else np.zeros(1))).to(dev)
        s = Solver(n, device=dev, leaf_max=opts["leaf"])
        w = torch.empty_like(dt)

        def solve():
            w.copy_(dt)
            s.solve(w, et)
        row["self_ms"] = _median_gpu_ms(solve, opts["warmup"], opts["iters"])
        w_self = w.cpu().numpy()
        row["self_info"] = int(s.info.item())
        if w_cpu is not None:
            row["self_err"] = rel_err_2norm(w_self, w_cpu)
            row["self_maxerr"] = rel_err_max(w_self, w_cpu)
        if opts["diagnostics"]:
            from . import profile_levels
            row["merges"] = s.merge_stats()
            row["level_ms"] = profile_levels(s, w, dt, et)
    if opts["run_rocsolver"]:
        import torch
        from .comparators import rocsolver_eigvals
        dev = torch.device("cuda:0")
        dt = torch.from_numpy(np.ascontiguousarray(d)).to(dev)
        et = torch.from_numpy(np.ascontiguousarray(e)).to(dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        wr, info = rocsolver_eigvals(dt, et, "stedc")
        torch.cuda.synchronize()
        row["rslv_ms"] = (time.perf_counter() - t0) * 1e3
        wr = wr.cpu().numpy()
        ref = w_cpu if w_cpu is not None else w_self
        if ref is not None:
            row["rslv_err"] = rel_err_2norm(wr, ref)
        if "self_ms" in row:
            row["rslv/self"] = row["rslv_ms"] / row["self_ms"]
    return row


COLUMNS = ("mode", "n", "prep_ms", "cpu_ms", "self_ms", "self_err", "self_maxerr", "rslv_ms",
           "rslv/self", "rslv_err")
WIDTH = {"mode": 20, "n": 8, "prep_ms": 12, "cpu_ms": 14, "self_ms": 14, "self_err": 16,
         "self_maxerr": 16, "rslv_ms": 14, "rslv/self": 12, "rslv_err": 16}


def format_row(row, cols):
    out = []
    for c in cols:
        v = row.get(c, "")
        if isinstance(v, float):
            v = f"{v:.3e}" if c.endswith("err") else f"{v:.3f}"
        out.append(str(v).ljust(WIDTH[c]))
    return "".join(out)


def format_merges(stats, level_ms=None):
    """Per-level table of the merge diagnostics (and kernel milliseconds)."""
    classes = ("prepare", "deflate", "secular", "secular_mid", "secular_stream", "wprod",
               "rows", "finalize")
    head = ("  level  merges      rows  size_min  size_max     K_sum  zero_defl  close_defl"
            "  iter_sum  iter_max  failed")
    if level_ms is not None:
        head += "  ms(" + " ".join(classes) + ")"
    lines = [head]
    for lev in sorted(set(int(x) for x in stats["level"])):
        m = stats[stats["level"] == lev]
        line = (f"  {lev:5d} {len(m):7d} {int(m['size'].sum()):9d} {int(m['size'].min()):9d} "
                f"{int(m['size'].max()):9d} {int(m['k'].sum()):9d} {int(m['tiny'].sum()):10d} "
                f"{int(m['givens'].sum()):11d} {int(m['evals'].sum()):9d} "
                f"{int(m['evals_max'].max()):9d} {int(m['failed'].sum()):7d}")
        if level_ms is not None:
            per = level_ms.get(lev, {})
            line += "  " + " ".join(f"{per.get(c, 0.0):.4f}" for c in classes)
        lines.append(line)
    if level_ms is not None and -1 in level_ms:
        lines.append("  before the merges (ms): " + " ".join(
            f"{k}={v:.4f}" for k, v in level_ms[-1].items()))
    return "\n".join(lines)


def main(argv=None, out=sys.stdout):
    argv = list(sys.argv[1:] if argv is None else argv)
    sizes = [int(x) for x in (_flag(argv, "--sizes", "512,1024,2048,4096")).split(",") if x]
    opts = {
        "warmup": _int_flag(argv, "--warmup", 2),
        "iters": _int_flag(argv, "--iters", 10),
        "run_cpu": _int_flag(argv, "--run-cpu", 1) != 0,
        "run_self": _int_flag(argv, "--run-self", 1) != 0,
        "run_rocsolver": (_int_flag(argv, "--run-rocsolver",
                                    _int_flag(argv, "--run-cusolver", 0)) != 0),
        "diagnostics": _int_flag(argv, "--tridiag-diagnostics", 0) != 0,
        "leaf": _int_flag(argv, "--leaf", 0),
    }
    modes = [m for m in _flag(argv, "--data", "bench").split(",") if m]
    input_path = _flag(argv, "--input")
    output_path = _flag(argv, "--output-tridiag")
    input_tridiag = _int_flag(argv, "--input-tridiag", 0) != 0
    if not (opts["run_cpu"] or opts["run_self"] or opts["run_rocsolver"] or output_path):
        raise SystemExit("at least one of --run-cpu, --run-self, --run-rocsolver or "
                         "--output-tridiag must be enabled")
    cols = [c for c in COLUMNS if
            (c not in ("cpu_ms",) or opts["run_cpu"]) and
            (c not in ("self_ms",) or opts["run_self"]) and
            (c not in ("self_err", "self_maxerr") or (opts["run_cpu"] and opts["run_self"])) and
            (c not in ("rslv_ms",) or opts["run_rocsolver"]) and
            (c != "rslv/self" or (opts["run_self"] and opts["run_rocsolver"])) and
            (c != "rslv_err" or (opts["run_rocsolver"] and (opts["run_cpu"] or opts["run_self"])))]
    print("".join(c.ljust(WIDTH[c]) for c in cols), file=out, flush=True)
    rows = []
    cases = []
    if input_path:
        if not input_tridiag:
            raise SystemExit("dense --input needs the two-stage reduction, which this "
                             "eigenvalue path does not include; pass --input-tridiag=1 for "
                             "a tridiagonal .de.bin")
        t0 = time.perf_counter()
        d, e = read_de_bin(input_path)
        cases.append(("tridiag_bin", d, e, (time.perf_counter() - t0) * 1e3))
    else:
        for mode in modes:
            for idx, n in enumerate(sizes):
                t0 = time.perf_counter()
                d, e = make_input(mode, n, idx)
                cases.append((mode, d, e, (time.perf_counter() - t0) * 1e3))
    for mode, d, e, prep in cases:
        if output_path:
            write_de_bin(output_path, d, e)
            print(f"wrote_tridiag={output_path} n={d.shape[0]} "
                  f"bytes={8 + 8 * (2 * d.shape[0] - 1)}", file=out, flush=True)
        if not (opts["run_cpu"] or opts["run_self"] or opts["run_rocsolver"]):
            continue
        row = run_case(mode, d, e, prep, opts, out)
        rows.append(row)
        print(format_row(row, cols), file=out, flush=True)
        if opts["diagnostics"] and "merges" in row:
            print(format_merges(row["merges"], row.get("level_ms")), file=out, flush=True)
    return rows


if __name__ == "__main__":
    main()
