"""Command-line benchmark harness with the reference's flags and table
(gpu/bench/tridiag_dc_bench.cu: main, --sizes/--warmup/--iters/--data/--input
/--input-tridiag/--output-tridiag/--run-cpu/--run-self/--run-cusolver/
--tridiag-diagnostics), for the tridiagonal eigenvalue path only:

    python -m brdc.harness --sizes=4096,16384 --data=normal,toeplitz
    python -m brdc.harness --input=reduced_dense_49152.de.bin --input-tridiag=1

Columns: mode, n, prep_ms (input preparation), cpu_ms (LAPACK DSTERF on the
host, the reference's acceptance oracle; the reference puts its CPU solver
there), self_ms (this GPU solve, HIP-event timed, median of --iters after
--warmup), self_err (max |w - w_cpu| / max |w_cpu|), rslv_ms (rocSOLVER
STEDC values-only on the same GPU, one run: the vendor column the reference
fills with cuSOLVER), rslv/self and rslv_err.  --run-rocsolver replaces
--run-cusolver (also accepted).  Dense inputs (--input without
--input-tridiag) need the two-stage reduction, which this path does not
include.

Data modes: normal, uniform, toeplitz, clustered (the generators of
brdc.generators) and bench (D_i = n + i/4 + U(-1/2, 1/2), E_i = U(0.05, 0.95):
the distribution of the reference's bench mode, drawn from this package's
splitmix64 stream rather than std::mt19937).
"""
from __future__ import annotations

import sys
import time

import numpy as np

from .generators import GENERATORS, SplitMix64, generate, read_de_bin, seed_for, write_de_bin


def _flag(argv, name, default=None):
    pre = name + "="
    for a in argv:
        if a.startswith(pre):
            return a[len(pre):]
    return default


def _int_flag(argv, name, default):
    v = _flag(argv, name)
    return default if v is None else int(v)


def bench_data(n: int, seed: int):
    g = SplitMix64(seed)
    d = float(n) + 0.25 * np.arange(n) + (g.uniform(n) - 0.5)
    e = 0.05 + 0.9 * g.uniform(max(n - 1, 0))
    return d, e


def make_input(mode: str, n: int):
    if mode == "bench":
        return bench_data(n, seed_for(0, n))
    if mode in GENERATORS:
        return generate(mode, n, seed_for(2, n))
    raise SystemExit(f"unknown --data mode {mode!r} (bench, {', '.join(GENERATORS)})")


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
            row["self_err"] = float(np.max(np.abs(w_self - w_cpu)) /
                                    max(np.max(np.abs(w_cpu)), 1e-300))
        if opts["diagnostics"]:
            row["merges"] = s.merge_stats()
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
            row["rslv_err"] = float(np.max(np.abs(wr - ref)) / max(np.max(np.abs(ref)), 1e-300))
        if "self_ms" in row:
            row["rslv/self"] = row["rslv_ms"] / row["self_ms"]
    return row


COLUMNS = ("mode", "n", "prep_ms", "cpu_ms", "self_ms", "self_err", "rslv_ms", "rslv/self",
           "rslv_err")
WIDTH = {"mode": 12, "n": 8, "prep_ms": 12, "cpu_ms": 14, "self_ms": 14, "self_err": 16,
         "rslv_ms": 14, "rslv/self": 12, "rslv_err": 16}


def format_row(row, cols):
    out = []
    for c in cols:
        v = row.get(c, "")
        if isinstance(v, float):
            v = f"{v:.3e}" if c.endswith("err") else f"{v:.3f}"
        out.append(str(v).ljust(WIDTH[c]))
    return "".join(out)


def format_merges(stats):
    lines = ["  level  merges      rows     K_sum   K_max  givens  evals/root"]
    for lev in sorted(set(int(x) for x in stats["level"])):
        m = stats[stats["level"] == lev]
        ks = int(m["k"].sum())
        lines.append(f"  {lev:5d} {len(m):7d} {int(m['size'].sum()):9d} {ks:9d} "
                     f"{int(m['k'].max()):7d} {int(m['givens'].sum()):7d} "
                     f"{(m['evals'].sum() / max(ks, 1)):11.2f}")
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
            (c != "self_err" or (opts["run_cpu"] and opts["run_self"])) and
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
            for n in sizes:
                t0 = time.perf_counter()
                d, e = make_input(mode, n)
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
            print(format_merges(row["merges"]), file=out, flush=True)
    return rows


if __name__ == "__main__":
    main()
