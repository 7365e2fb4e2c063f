// brdc_deflate_wide.h -- deflation of the few huge merges at the top of a
// tree, spread over many blocks.
//
// k_deflate runs one block per merge, which leaves a level of one to eight
// merges of thousands of rows on as many compute units.  Here each merge is
// cut into chunks of kWideChunk positions (or, for the Givens stage, of
// kWideChunk compacted poles) with one block per (merge, chunk), and the
// stages that need a prefix over the whole merge meet at kernel boundaries
// through small per-chunk records:
//   W1  tiny-z flags and kept count per chunk        (dlaed7_br.f:240-264)
//   W2  kept positions compacted into P (prefix of the chunk counts)
//   W3  speculative close-pole Givens runs from every t, per-chunk overlap
//       summary and run lengths                       (dlaed7_br.f:296-335)
//   W4  one block per merge: exact sequential resolution when any runs
//       overlap (as k_deflate), Givens offsets per chunk
//   W5  run replay (rotations, deflated values)
//   W6  kept / tiny counts per chunk after the replay
//   W7  compaction into the K and T lists, counts
// The arithmetic per position is k_deflate's; only the work split differs,
// so the outputs are identical.
#pragma once

#include "brdc_kernels.h"

namespace brdc {

constexpr int kWideChunk = 4096;    // positions (or poles) per block
constexpr int kWideThreads = 1024;  // threads per block, 4 positions each
constexpr int kWideMaxMerges = 4;   // levels with at most this many merges ...
constexpr int kWideMinSize = 4096;  // ... of at least this many rows use it
constexpr int kWideFields = 8;      // ints per chunk record

// record fields
enum WideField { kWKeep, kWCarry, kWFirst, kWBad, kWRuns, kWGoff, kWNa, kWNt };

struct WideGeo {
    Geo g;
    int chunks;                  // chunks per merge (of the level's largest merge)
    int* __restrict__ rec;       // [merge][chunk][kWideFields]
    double* __restrict__ tol;    // per merge: the deflation tolerance
    __device__ int* at(int merge, int c) const { return rec + ((size_t)merge * chunks + c) * kWideFields; }
};

// chunks that start inside a merge of sz rows (only their records are written)
__device__ __forceinline__ int wide_chunks_in(const WideGeo& wg, int sz) {
    return min(wg.chunks, (sz + kWideChunk - 1) / kWideChunk);
}

// Field `field` of a merge's chunk records q < n summed, in total and over
// q < c.  Lane l reads record q0 + l, 64 records to a load (every wave of
// the block alike): a loop over the records would wait a memory round trip
// per record.
struct WideSums {
    int below, total;
};
__device__ __forceinline__ WideSums wide_field_sums(const WideGeo& wg, int merge, int n, int c,
                                                    int field) {
    const int lane = threadIdx.x & (kWave - 1);
    int below = 0, total = 0;
    for (int q0 = 0; q0 < n; q0 += kWave) {
        const int q = q0 + lane;
        int vt = (q < n) ? wg.at(merge, q)[field] : 0;
        int vb = (q < c) ? vt : 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            vt += __shfl_xor(vt, o, kWave);
            vb += __shfl_xor(vb, o, kWave);
        }
        total += vt;
        below += vb;
    }
    return WideSums{__builtin_amdgcn_readfirstlane(below), __builtin_amdgcn_readfirstlane(total)};
}

// kept poles of a merge: the sum over its chunks
__device__ __forceinline__ int wide_kept(const WideGeo& wg, int merge, int sz) {
    return wide_field_sums(wg, merge, wide_chunks_in(wg, sz), 0, kWKeep).total;
}

// block-wide exclusive scan of one int per thread (kWideThreads threads);
// returns the exclusive prefix and writes the block total to *total
__device__ __forceinline__ int wide_scan(int v, int* sh, int* total) {
    constexpr int NW = kWideThreads / kWave;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int incl = wave_incl_sum(v);
    if (lane == kWave - 1) sh[w] = incl;
    __syncthreads();
    int off = 0, tot = 0;
    for (int q = 0; q < NW; q++) {
        const int x = sh[q];
        if (q < w) off += x;
        tot += x;
    }
    __syncthreads();  // sh reuse by the caller
    *total = tot;
    return off + incl - v;
}

__device__ __forceinline__ int wide_sum(int v, int* sh) {
    int total;
    (void)wide_scan(v, sh, &total);
    return total;
}

// W1: flags, kept count per chunk; chunk 0 also records the tolerance
__global__ __launch_bounds__(kWideThreads) void k_wide_flags(WideGeo wg, const double* __restrict__ ew,
                                                            const double* __restrict__ md,
                                                            const double* __restrict__ mz,
                                                            int* __restrict__ fl) {
    __shared__ int sh[kWideThreads / kWave];
    const int merge = blockIdx.x / wg.chunks, c = blockIdx.x % wg.chunks;
    const Item it = wg.g.item(merge);
    const int s = it.s, sz = it.sz;
    const int p0 = c * kWideChunk;
    if (p0 >= sz) return;
    const double rho = fabs(2.0 * ew[s + sz / 2 - 1]);
    const double tol = 8.0 * kEps * fmax(fabs(md[s]), fabs(md[s + sz - 1]));
    if (c == 0 && threadIdx.x == 0) wg.tol[merge] = tol;
    int cnt = 0;
    for (int p = p0 + (int)threadIdx.x; p < min(sz, p0 + kWideChunk); p += kWideThreads) {
        const bool keep = rho * fabs(mz[s + p]) > tol;
        fl[s + p] = keep ? 0 : 2;
        cnt += keep;
    }
    const int tot = wide_sum(cnt, sh);
    if (threadIdx.x == 0) wg.at(merge, c)[kWKeep] = tot;
}

// W2: compaction of the kept positions into P, in position order
__global__ __launch_bounds__(kWideThreads) void k_wide_compact(WideGeo wg, const int* __restrict__ fl,
                                                              int* __restrict__ P) {
    __shared__ int sh[kWideThreads / kWave];
    const int merge = blockIdx.x / wg.chunks, c = blockIdx.x % wg.chunks;
    const Item it = wg.g.item(merge);
    const int s = it.s, sz = it.sz;
    const int p0 = c * kWideChunk;
    if (p0 >= sz) return;
    const int base = wide_field_sums(wg, merge, c, c, kWKeep).below;
    // each thread owns 4 consecutive positions, so the chunk keeps its order
    constexpr int per = kWideChunk / kWideThreads;
    const int pa = p0 + (int)threadIdx.x * per;
    int f[per], cnt = 0;
#pragma unroll
    for (int u = 0; u < per; u++) {
        const int p = pa + u;
        f[u] = (p < sz) ? fl[s + p] : 1;
        cnt += (f[u] == 0);
    }
    int total;
    int off = base + wide_scan(cnt, sh, &total);
#pragma unroll
    for (int u = 0; u < per; u++)
        if (f[u] == 0) P[s + off++] = pa + u;
}

// W3: speculative run ends for the poles t of this chunk of [1, m), and the
// chunk's overlap summary: largest end (carry), first run start, whether a
// walk hit the cap or two runs overlap inside the chunk, and the run lengths
__global__ __launch_bounds__(kWideThreads) void k_wide_spec(WideGeo wg, const int* __restrict__ P,
                                                           const double* __restrict__ md,
                                                           const double* __restrict__ mz,
                                                           int* __restrict__ rend) {
    constexpr int NW = kWideThreads / kWave;
    __shared__ int sh[NW], shc[NW], shf[NW];
    __shared__ int s_bad;
    const int merge = blockIdx.x / wg.chunks, c = blockIdx.x % wg.chunks;
    const Item it = wg.g.item(merge);
    const int s = it.s;
    const int m = wide_kept(wg, merge, it.sz);
    const int t0 = c * kWideChunk;
    if (t0 >= m) return;
    const double tol = wg.tol[merge];
    const int tb = max(1, t0), te = min(m, t0 + kWideChunk);
    for (int t = tb + (int)threadIdx.x; t < te; t += kWideThreads)
        rend[s + t] = speculative_end(s, t, m, kSpecCap, tol, P, md, mz);
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    // each wave scans a contiguous slice of the chunk in order
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int R = ((te - tb + NW - 1) / NW + kWave - 1) / kWave * kWave;
    const int wb = min(te, tb + w * R), we = min(te, wb + R);
    int carry = INT_MIN, first = INT_MAX, runs = 0;
    bool bad = false;
    for (int base = wb; base < we; base += kWave) {
        const int t = base + lane;
        const int e = (t < we) ? rend[s + t] : -1;
        if (e == -2) bad = true;
        const int incl = wave_incl_max(e == -2 ? INT_MAX : e);
        int excl = __shfl_up(incl, 1, kWave);
        if (lane == 0) excl = INT_MIN;
        excl = max(excl, carry);
        if (e >= 0 && excl >= t) bad = true;
        if (e >= 0) {
            first = min(first, t);
            runs += e - t;
        }
        carry = max(carry, __shfl(incl, kWave - 1, kWave));
    }
    first = wave_min(first);
    for (int o = 32; o > 0; o >>= 1) runs += __shfl_xor(runs, o, kWave);
    if (__any(bad) && lane == 0) s_bad = 1;
    if (lane == 0) {
        shc[w] = carry;
        shf[w] = first;
        sh[w] = runs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int cin = INT_MIN, cf = INT_MIN, ff = INT_MAX, rr = 0, b = s_bad;
        for (int q = 0; q < NW; q++) {
            if (shf[q] != INT_MAX && cin >= shf[q]) b = 1;  // a run of an earlier slice reaches in
            cin = max(cin, shc[q]);
            cf = max(cf, shc[q]);
            ff = min(ff, shf[q]);
            rr += sh[q];
        }
        int* r = wg.at(merge, c);
        r[kWCarry] = cf;
        r[kWFirst] = ff;
        r[kWBad] = b;
        r[kWRuns] = rr;
    }
}

// W4: one block per merge.  Runs that overlap across chunks (or a capped
// walk) take k_deflate's exact sequential resolution, and the chunk run
// lengths are recounted; then the Givens offset of every chunk.
__global__ __launch_bounds__(kWideThreads) void k_wide_resolve(WideGeo wg, const int* __restrict__ P,
                                                              const double* __restrict__ md,
                                                              const double* __restrict__ mz,
                                                              int* __restrict__ rend,
                                                              int* __restrict__ counts) {
    __shared__ int sh[kWideThreads / kWave];
    __shared__ int s_fix;
    const int merge = blockIdx.x;
    const Item it = wg.g.item(merge);
    const int s = it.s;
    const int m = wide_kept(wg, merge, it.sz);
    const int nc = (m + kWideChunk - 1) / kWideChunk;  // chunks of poles
    const int lane = threadIdx.x & (kWave - 1);
    if (threadIdx.x < kWave) {
        // a chunk needs the sequential resolution if a walk hit the cap or a
        // run of an earlier chunk reaches its first run start (lane l reads
        // record q0 + l; cin: the largest run end of the chunks before)
        bool fix = false;
        int cin = INT_MIN;
        for (int q0 = 0; q0 < nc; q0 += kWave) {
            const int q = q0 + lane;
            int bad = 0, first = INT_MAX, carry = INT_MIN;
            if (q < nc) {
                const int* r = wg.at(merge, q);
                bad = r[kWBad];
                first = r[kWFirst];
                carry = r[kWCarry];
            }
            const int incl = wave_incl_max(carry);
            int excl = __shfl_up(incl, 1, kWave);
            if (lane == 0) excl = INT_MIN;
            excl = max(excl, cin);
            if (bad || (first != INT_MAX && excl >= first)) fix = true;
            cin = max(cin, __shfl(incl, kWave - 1, kWave));
        }
        fix = __any(fix);
        if (lane == 0) s_fix = fix ? 1 : 0;
        if (fix && lane == 0) {  // k_deflate's sequential resolution
            const double tol = wg.tol[merge];
            int cover = -1;
            for (int t = 1; t < m; t++) {
                int e = rend[s + t];
                if (e == -1) continue;
                if (t <= cover) {
                    rend[s + t] = -1;
                    continue;
                }
                if (e == -2) {
                    e = speculative_end(s, t, m, m, tol, P, md, mz);
                    rend[s + t] = e;
                }
                cover = e;
            }
        }
    }
    __syncthreads();
    if (s_fix) {  // recount the run lengths of every chunk
        for (int q = 0; q < nc; q++) {
            const int tb = max(1, q * kWideChunk), te = min(m, (q + 1) * kWideChunk);
            int runs = 0;
            for (int t = tb + (int)threadIdx.x; t < te; t += kWideThreads) {
                const int e = rend[s + t];
                runs += (e >= 0) ? e - t : 0;
            }
            const int tot = wide_sum(runs, sh);
            if (threadIdx.x == 0) wg.at(merge, q)[kWRuns] = tot;
            __syncthreads();
        }
    }
    if (threadIdx.x < kWave) {
        // Givens offsets: the exclusive prefix sum of the chunks' run lengths
        int g = 0;
        for (int q0 = 0; q0 < nc; q0 += kWave) {
            const int q = q0 + lane;
            const int v = (q < nc) ? wg.at(merge, q)[kWRuns] : 0;
            const int incl = wave_incl_sum(v);
            if (q < nc) wg.at(merge, q)[kWGoff] = g + incl - v;
            g += __shfl(incl, kWave - 1, kWave);
        }
        if (lane == 0) counts[kCountStride * merge + 1] = g;
    }
}

// W5: replay of the runs that start in this chunk of poles
__global__ __launch_bounds__(kWideThreads) void k_wide_replay(
    WideGeo wg, const int* __restrict__ P, const int* __restrict__ rend, double* __restrict__ md,
    double* __restrict__ mz, double* __restrict__ mq1, double* __restrict__ mq2,
    int* __restrict__ fl, double* __restrict__ gd, double* __restrict__ gq1,
    double* __restrict__ gq2) {
    __shared__ int sh[kWideThreads / kWave];
    const int merge = blockIdx.x / wg.chunks, c = blockIdx.x % wg.chunks;
    const Item it = wg.g.item(merge);
    const int s = it.s;
    const int m = wide_kept(wg, merge, it.sz);
    const int t0 = c * kWideChunk;
    if (t0 >= m) return;
    const double tol = wg.tol[merge];
    const int goff = wg.at(merge, c)[kWGoff];
    constexpr int per = kWideChunk / kWideThreads;
    const int ta = t0 + (int)threadIdx.x * per;
    int e[per], len = 0;
#pragma unroll
    for (int u = 0; u < per; u++) {
        const int t = ta + u;
        e[u] = (t >= 1 && t < m) ? rend[s + t] : -1;
        len += (e[u] >= 0) ? e[u] - t : 0;
    }
    int total;
    int off = goff + wide_scan(len, sh, &total);
#pragma unroll
    for (int u = 0; u < per; u++) {
        const int t = ta + u;
        if (e[u] >= 0) {
            replay_run(s, t, e[u], off, tol, P, md, mz, mq1, mq2, fl, gd, gq1, gq2);
            off += e[u] - t;
        }
    }
}

// W6: kept (0) and tiny (2) counts per chunk after the replay
__global__ __launch_bounds__(kWideThreads) void k_wide_count(WideGeo wg, const int* __restrict__ fl) {
    __shared__ int sh[kWideThreads / kWave];
    const int merge = blockIdx.x / wg.chunks, c = blockIdx.x % wg.chunks;
    const Item it = wg.g.item(merge);
    const int s = it.s, sz = it.sz;
    const int p0 = c * kWideChunk;
    if (p0 >= sz) return;
    int na = 0, nt = 0;
    for (int p = p0 + (int)threadIdx.x; p < min(sz, p0 + kWideChunk); p += kWideThreads) {
        const int f = fl[s + p];
        na += (f == 0);
        nt += (f == 2);
    }
    const int ta = wide_sum(na, sh);
    const int tt = wide_sum(nt, sh);
    if (threadIdx.x == 0) {
        int* r = wg.at(merge, c);
        r[kWNa] = ta;
        r[kWNt] = tt;
    }
}

// W7: compaction into the secular list (K) and the tiny-deflated list (T)
__global__ __launch_bounds__(kWideThreads) void k_wide_scatter(
    WideGeo wg, const int* __restrict__ fl, const double* __restrict__ md,
    const double* __restrict__ mz, const double* __restrict__ mq1,
    const double* __restrict__ mq2, double* __restrict__ ad, double* __restrict__ az,
    double2* __restrict__ adz, double2* __restrict__ aq, double* __restrict__ td,
    double* __restrict__ tq1, double* __restrict__ tq2, int* __restrict__ counts) {
    __shared__ int sh[kWideThreads / kWave];
    const int merge = blockIdx.x / wg.chunks, c = blockIdx.x % wg.chunks;
    const Item it = wg.g.item(merge);
    const int s = it.s, sz = it.sz;
    const int p0 = c * kWideChunk;
    if (p0 >= sz) return;
    const int nq = wide_chunks_in(wg, sz);
    const WideSums sa = wide_field_sums(wg, merge, nq, c, kWNa);
    const WideSums st = wide_field_sums(wg, merge, nq, c, kWNt);
    const int abase = sa.below, tbase = st.below, K = sa.total, Tn = st.total;
    constexpr int per = kWideChunk / kWideThreads;
    const int pa = p0 + (int)threadIdx.x * per;
    int f[per], na = 0, nt = 0;
#pragma unroll
    for (int u = 0; u < per; u++) {
        const int p = pa + u;
        f[u] = (p < sz) ? fl[s + p] : 1;
        na += (f[u] == 0);
        nt += (f[u] == 2);
    }
    int tot;
    int ao = abase + wide_scan(na, sh, &tot);
    int to = tbase + wide_scan(nt, sh, &tot);
#pragma unroll
    for (int u = 0; u < per; u++) {
        const int p = pa + u;
        if (f[u] == 0) {
            const int o = s + ao++;
            const double dv = md[s + p], zv = mz[s + p];
            ad[o] = dv;
            az[o] = zv;
            adz[o] = make_double2(dv, zv * zv);
            aq[o] = make_double2(mq1[s + p], mq2[s + p]);
        } else if (f[u] == 2) {
            const int o = s + to++;
            td[o] = md[s + p];
            tq1[o] = mq1[s + p];
            tq2[o] = mq2[s + p];
        }
    }
    if (c == 0 && threadIdx.x == 0) {
        counts[kCountStride * merge + 0] = K;
        counts[kCountStride * merge + 2] = Tn;
        counts[kCountStride * merge + 3] = 0;
        counts[kCountStride * merge + 4] = 0;
        counts[kCountStride * merge + 5] = 0;
        counts[kCountStride * merge + 6] = 0;
    }
}

}  // namespace brdc
