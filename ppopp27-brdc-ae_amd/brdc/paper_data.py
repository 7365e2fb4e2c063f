"""The reference bench's built-in tridiagonal inputs, bit for bit.

gpu/bench/tridiag_dc_bench.cu builds its inputs with the C++ standard
library's engines and distributions (libstdc++, x86-64 without FMA
contraction, the build of gpu/CMakeLists.txt):

* ``--data=normal|uniform|toeplitz|clustered``: make_paper_tridiagonal_data
  (tridiag_dc_bench.cu:257-308).  One std::mt19937_64 seeded with
  ``n * 0x9E3779B97F4A7C15 ^ h`` (h = 0xD1B54A32D192ED03 for normal, else
  0x94D049BB133111EB).  normal: D ~ N(0, 1), E = |N(0, 0.2)| (two
  normal_distribution objects, each with its own cached second value);
  uniform: D ~ U(-1, 1), E ~ U(0.1, 0.3); toeplitz: D = 2, E = 0.25;
  clustered: D_i = 1 + 1e-12 (i - (n+1)/2), E_i = 1e-4 (1 + 0.1 cos(0.33 i))
  with i 1-based.
* ``--data=bench``: make_bench_tridiagonal_data (229-255), std::mt19937
  seeded with 1234 + the size's position in --sizes: D_i = n + i/4 +
  U(-1/2, 1/2), E_i ~ U(0.05, 0.95).
* ``--data=test``: make_test_tridiagonal_data (311-335), std::mt19937 seeded
  with 1300 + n, n*n uniform draws in column order of which the diagonal
  (plus n) and the first subdiagonal are kept.

The engines, std::generate_canonical<double, 53> and the distributions are
restated here from the C++ standard's definitions and libstdc++'s documented
algorithms (uniform_real: canonical * (b - a) + a; normal: Marsaglia's polar
method, returning y * m and caching x * m).  tests/test_harness.py checks the
results against a small program of its own built with g++ on the same
standard library.
"""
from __future__ import annotations

import math

import numpy as np

# --------------------------------------------------------------- engines
_MT32 = dict(n=624, m=397, a=0x9908B0DF, upper=0x80000000, lower=0x7FFFFFFF)
_MT64 = dict(n=312, m=156, a=0xB5026F5AA96619E9, upper=0xFFFFFFFF80000000,
             lower=0x7FFFFFFF)


class _MersenneTwister:
    """std::mt19937 (w = 32) or std::mt19937_64 (w = 64) with the single-value
    seeding of the standard ([rand.eng.mt]); ``raw(k)`` returns the next k
    outputs.  The twist runs vectorised in slices whose inputs are all
    already final."""

    def __init__(self, seed: int, w: int):
        self.w = w
        p = _MT64 if w == 64 else _MT32
        self.n, self.m = p["n"], p["m"]
        self.dt = np.uint64 if w == 64 else np.uint32
        mask = (1 << w) - 1
        f, shift = (6364136223846793005, 62) if w == 64 else (1812433253, 30)
        st = [seed & mask]
        for i in range(1, self.n):
            prev = st[-1]
            st.append((f * (prev ^ (prev >> shift)) + i) & mask)
        self.mt = np.array(st, dtype=self.dt)
        self.a = self.dt(p["a"])
        self.upper = self.dt(p["upper"])
        self.lower = self.dt(p["lower"])
        self.idx = self.n
        self.buf = np.zeros(0, dtype=self.dt)

    def _twist(self):
        n, m, mt = self.n, self.m, self.mt
        one = self.dt(1)
        zero = self.dt(0)

        def update(lo, hi, src_lo):
            i = np.arange(lo, hi)
            y = (mt[i] & self.upper) | (mt[(i + 1) % n] & self.lower)
            mag = np.where((y & one) != zero, self.a, zero)
            mt[i] = mt[src_lo + (i - lo)] ^ (y >> one) ^ mag

        # i in [0, n-m) reads mt[i+m], not yet updated; then slices of n-m
        # whose mt[i+m-n] an earlier slice has updated; the last element reads
        # the new mt[0]
        step = n - m
        update(0, step, m)
        lo = step
        while lo < n - 1:
            hi = min(lo + step, n - 1)
            update(lo, hi, lo + m - n)
            lo = hi
        update(n - 1, n, m - 1)

    def _temper(self, y):
        if self.w == 64:
            y = y ^ ((y >> np.uint64(29)) & np.uint64(0x5555555555555555))
            y = y ^ ((y << np.uint64(17)) & np.uint64(0x71D67FFFEDA60000))
            y = y ^ ((y << np.uint64(37)) & np.uint64(0xFFF7EEE000000000))
            return y ^ (y >> np.uint64(43))
        y = y ^ (y >> np.uint32(11))
        y = y ^ ((y << np.uint32(7)) & np.uint32(0x9D2C5680))
        y = y ^ ((y << np.uint32(15)) & np.uint32(0xEFC60000))
        return y ^ (y >> np.uint32(18))

    def raw(self, k: int) -> np.ndarray:
        out = []
        while k > 0:
            if self.idx >= self.n:
                self._twist()
                self.buf = self._temper(self.mt.copy())
                self.idx = 0
            take = min(k, self.n - self.idx)
            out.append(self.buf[self.idx:self.idx + take])
            self.idx += take
            k -= take
        return np.concatenate(out) if out else np.zeros(0, dtype=self.dt)

    def canonical(self, k: int) -> np.ndarray:
        """std::generate_canonical<double, 53>: one 64-bit output, or two
        32-bit outputs u0 + u1 * 2^32, over 2^64 (sums in double), clamped
        below 1."""
        if self.w == 64:
            s = self.raw(k).astype(np.float64)
        else:
            r = self.raw(2 * k)
            s = r[0::2].astype(np.float64) + r[1::2].astype(np.float64) * 4294967296.0
        c = s / 18446744073709551616.0
        return np.where(c >= 1.0, np.nextafter(1.0, 0.0), c)


def _uniform(eng, k, a, b):
    """std::uniform_real_distribution<double>(a, b): canonical * (b - a) + a."""
    return eng.canonical(k) * (b - a) + a


def _normal(eng, k, mean, stddev):
    """k draws of one std::normal_distribution<double>(mean, stddev) object
    (Marsaglia polar: each accepted pair returns y * mult, then the cached
    x * mult); an odd k leaves the last cached value unused."""
    out = []
    while len(out) < k:
        x = 2.0 * float(eng.canonical(1)[0]) - 1.0
        y = 2.0 * float(eng.canonical(1)[0]) - 1.0
        r2 = x * x + y * y
        if r2 > 1.0 or r2 == 0.0:
            continue
        mult = math.sqrt(-2.0 * math.log(r2) / r2)
        out.append(y * mult)
        out.append(x * mult)
    v = np.array(out[:k], dtype=np.float64)
    return v * stddev + mean


# ------------------------------------------------------------- generators
PAPER_MODES = ("normal", "uniform", "toeplitz", "clustered")


def paper_data(mode: str, n: int) -> tuple[np.ndarray, np.ndarray]:
    """make_paper_tridiagonal_data (tridiag_dc_bench.cu:257-308)."""
    if mode not in PAPER_MODES:
        raise ValueError(f"unsupported paper mode {mode!r}")
    ne = max(n - 1, 0)
    seed = ((n * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF) ^ \
        (0xD1B54A32D192ED03 if mode == "normal" else 0x94D049BB133111EB)
    eng = _MersenneTwister(seed, 64)
    if mode == "normal":
        d = _normal(eng, n, 0.0, 1.0)
        e = np.abs(_normal(eng, ne, 0.0, 0.20))
    elif mode == "uniform":
        d = _uniform(eng, n, -1.0, 1.0)
        e = _uniform(eng, ne, 0.10, 0.30)
    elif mode == "toeplitz":
        d = np.full(n, 2.0)
        e = np.full(ne, 0.25)
    else:
        i1 = np.arange(1, n + 1, dtype=np.float64)
        d = 1.0 + 1.0e-12 * (i1 - 0.5 * (float(n) + 1.0))
        e = np.array([1.0e-4 * (1.0 + 0.1 * math.cos(0.33 * float(i + 1))) for i in range(ne)],
                     dtype=np.float64)
    return np.ascontiguousarray(d), np.ascontiguousarray(e)


def bench_data(n: int, size_index: int) -> tuple[np.ndarray, np.ndarray]:
    """make_bench_tridiagonal_data (tridiag_dc_bench.cu:229-255) with the
    seed the bench uses, 1234 + the size's position in --sizes (1513)."""
    eng = _MersenneTwister((1234 + size_index) & 0xFFFFFFFF, 32)
    noise = _uniform(eng, n, -0.5, 0.5)
    d = float(n) + 0.25 * np.arange(n, dtype=np.float64) + noise
    e = _uniform(eng, max(n - 1, 0), 0.05, 0.95)
    return np.ascontiguousarray(d), np.ascontiguousarray(e)


def test_data(n: int) -> tuple[np.ndarray, np.ndarray]:
    """make_test_tridiagonal_data (tridiag_dc_bench.cu:311-335), seed
    1300 + n (1510): n*n draws of U(-1, 1) in column order; the diagonal
    entry gets + n, the entry below it is the subdiagonal."""
    eng = _MersenneTwister((1300 + n) & 0xFFFFFFFF, 32)
    d = np.zeros(n)
    e = np.zeros(max(n - 1, 0))
    for col in range(n):
        v = _uniform(eng, n, -1.0, 1.0)
        d[col] = v[col] + float(n)
        if col + 1 < n:
            e[col] = v[col + 1]
    return d, e
