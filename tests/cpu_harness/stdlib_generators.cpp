// Prints, as hexadecimal bit patterns, the diagonal then the off-diagonal of
// the reference bench's built-in inputs for one mode and order, drawn with
// the C++ standard library's engines and distributions the way
// gpu/bench/tridiag_dc_bench.cu draws them (229-335): the check that the
// Python restatement (brdc/paper_data.py) reproduces them bit for bit.
// Usage: stdlib_generators <normal|uniform|toeplitz|clustered|bench|test> <n> [size index]
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

static void put(double v) {
    uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    std::printf("%016llx\n", (unsigned long long)b);
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string mode = argv[1];
    const long n = std::atol(argv[2]);
    const long idx = argc > 3 ? std::atol(argv[3]) : 0;
    std::vector<double> d(n), e(n > 1 ? n - 1 : 0);
    if (mode == "bench") {
        std::mt19937 rng((uint32_t)(1234 + idx));
        std::uniform_real_distribution<double> noise(-0.5, 0.5), off(0.05, 0.95);
        for (long i = 0; i < n; i++) d[i] = (double)n + 0.25 * (double)i + noise(rng);
        for (long i = 0; i + 1 < n; i++) e[i] = off(rng);
    } else if (mode == "test") {
        std::mt19937 rng((uint32_t)(1300 + n));
        std::uniform_real_distribution<double> u(-1.0, 1.0);
        for (long c = 0; c < n; c++)
            for (long r = 0; r < n; r++) {
                const double v = u(rng);
                if (r == c) d[c] = v + (double)n;
                if (r == c + 1) e[c] = v;
            }
    } else {
        const uint64_t base = (uint64_t)n * 0x9E3779B97F4A7C15ULL;
        std::mt19937_64 rng(base ^ (mode == "normal" ? 0xD1B54A32D192ED03ULL
                                                     : 0x94D049BB133111EBULL));
        std::normal_distribution<double> nd(0.0, 1.0), no(0.0, 0.20);
        std::uniform_real_distribution<double> ud(-1.0, 1.0), uo(0.10, 0.30);
        for (long i = 0; i < n; i++) {
            if (mode == "normal") d[i] = nd(rng);
            else if (mode == "uniform") d[i] = ud(rng);
            else if (mode == "toeplitz") d[i] = 2.0;
            else d[i] = 1.0 + 1.0e-12 * ((double)(i + 1) - 0.5 * ((double)n + 1.0));
        }
        for (long i = 0; i + 1 < n; i++) {
            if (mode == "normal") e[i] = std::abs(no(rng));
            else if (mode == "uniform") e[i] = uo(rng);
            else if (mode == "toeplitz") e[i] = 0.25;
            else e[i] = 1.0e-4 * (1.0 + 0.1 * std::cos(0.33 * (double)(i + 1)));
        }
    }
    for (double v : d) put(v);
    for (double v : e) put(v);
    return 0;
}
