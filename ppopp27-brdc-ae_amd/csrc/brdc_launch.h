// brdc_launch.h -- launch-shape helpers of the secular kernel, shared by the
// host enqueue code and the device (no kernels here, so host-only builds can
// include it).
#pragma once

#include <hip/hip_runtime.h>

namespace brdc {

constexpr int kWave = 64;
// ints per merge record: K, G, T, secular evaluations (the rest spare)
constexpr int kCountStride = 8;

// Lanes per root: about 4+ poles per lane per pass, and no more lanes than
// the level needs to fill the machine (~target lanes in flight; kTargetLanes
// unless tuned).
constexpr int kTargetLanes = 1 << 18;
__host__ __device__ inline int secular_lanes_max(int K, int merges, int target = kTargetLanes) {
    int L = 1;
    while (L < kWave && 4 * 2 * L <= K && (long long)merges * K * (2 * L) <= target) L <<= 1;
    return L;
}
// A level of short passes (fewer than 32 poles per lane after a doubling)
// doubles only while its waves still fit two per SIMD (1024 SIMDs on the
// MI355X): there a pass is short next to the root solver's step, whose
// dependent chain slows as more waves share a SIMD.
constexpr long long kTwoWavesPerSimd = 2 * 1024LL * kWave;
__host__ __device__ inline int secular_lanes(int K, int merges, int target = kTargetLanes) {
    int L = 1;
    while (L < kWave && 4 * 2 * L <= K && (long long)merges * K * (2 * L) <= target &&
           (K >= 64 * L || (long long)merges * K * (2 * L) <= kTwoWavesPerSimd))
        L <<= 1;
    return L;
}

// waves a merge of K poles needs at most (secular_lanes never exceeds
// secular_lanes_max)
__host__ __device__ inline int secular_waves(int K, int merges, int target = kTargetLanes) {
    const int L = secular_lanes_max(K, merges, target);
    return (K * L + kWave - 1) / kWave;
}

// An upper bound on the waves any merge of K <= maxsz poles needs: the most
// at secular_lanes_max, where secular_waves is not
// monotone in K (a smaller K can be given more lanes per root), but it grows
// with K while the lane count holds, so its maximum sits at the upper end of
// a constant-lane run: K = maxsz, K = 8L - 1 (the next K doubles L) or
// K = target / (merges * L) (the next K halves L).
inline int secular_slots(int maxsz, int merges, int target = kTargetLanes) {
    int best = secular_waves(maxsz, merges, target);
    for (int L = 1; L <= kWave; L <<= 1) {
        const long long cand[2] = {8LL * L - 1, target / ((long long)merges * L)};
        for (long long k : cand)
            if (k >= 1 && k <= maxsz) {
                const int w = secular_waves((int)k, merges, target);
                if (w > best) best = w;
            }
    }
    return best;
}

}  // namespace brdc
