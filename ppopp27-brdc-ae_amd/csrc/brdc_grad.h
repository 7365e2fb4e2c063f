// brdc_grad.h -- backward of the eigenvalues of symmetric tridiagonals.
//
// For T = tridiag(E, D, E) with unit eigenvectors v_i, d lambda_i / d D_j =
// v_ij^2 and d lambda_i / d E_j = 2 v_ij v_i,j+1, so an upstream gradient g
// gives
//   dL/dD_j = sum_i g_i v_ij^2,   dL/dE_j = 2 sum_i g_i v_ij v_i,j+1.
// No eigenvector is stored by the forward pass: each is recomputed at its
// eigenvalue mu by twisted factorization of T - mu I (Dhillon-Parlett):
//   forward  D+_1 = D_1 - mu,  l_k = E_k / D+_k,  D+_{k+1} = D_{k+1} - mu - l_k E_k
//   backward D-_n = D_n - mu,  u_k = E_k / D-_{k+1}, D-_k = D_k - mu - u_k E_k
//   twist    gamma_k = D+_k + D-_k - (D_k - mu),  r = argmin |gamma_k|
//   vector   z_r = 1,  z_k = -l_k z_{k+1} (k < r),  z_{k+1} = -u_k z_k (k >= r)
// with pivots below pivmin replaced by -pivmin (the LAPACK DLAR1V guard).
//
// Layout: one lane per eigenvalue, 64 eigenvalues of one matrix per wave (a
// task).  Every lane walks the same row k at the same time, so D and E come
// in as scalar loads, and the per-lane pivots live in the wave's slot of two
// scratch arrays [row][lane] (coalesced).  Each task writes its partial sums
// for its matrix's rows; a second kernel adds the tasks of each matrix in a
// fixed order (deterministic results).
#pragma once

#include "brdc_kernels.h"

namespace brdc {

constexpr int kGradWaves = 4;   // waves per block of the vector kernel
constexpr int kSeg = 16;        // rows per checkpoint segment
constexpr int kWin = 256;       // rows of D and E staged in LDS per wave

// Task t of a batch: matrix b with tprefix[b] <= t < tprefix[b+1], first
// eigenvalue 64 * (t - tprefix[b]).  One matrix: tprefix == nullptr.
__device__ __forceinline__ int grad_task_matrix(const int* tprefix, int nmat, int t) {
    if (!tprefix) return 0;
    int lo = 0, hi = nmat;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tprefix[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ double pivot_guard(double x, double pivmin) {
    return (fabs(x) < pivmin) ? -pivmin : x;
}

// Reciprocal on the factorizations' dependent chains: v_rcp_f64 and one
// Newton step (relative error below 2^-50).  Each step of the stationary and
// progressive factorizations is then exact for data perturbed by a few ulp,
// which moves the recomputed eigenvector by O(n eps / gap) as rounding does.
__device__ __forceinline__ double rcp_chain(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// Sum over the wave's 64 lanes of each of NV values (NV a power of two up to
// 64), scattered: returns in lane L the total of value L / (64 / NV) (the
// 64 / NV lanes that share a value agree).  Halving on lane bits 5, 4, ...
// costs NV/2 + NV/4 + ... + 1 exchanges, plus log2(64 / NV) for the lanes
// that share a value.
template <int NV>
__device__ __forceinline__ double reduce_scatter(double* v, int lane) {
#pragma unroll
    for (int h = NV / 2, bit = 32; h >= 1; h >>= 1, bit >>= 1) {
        const bool hi = (lane & bit) != 0;
#pragma unroll
        for (int i = 0; i < h; i++) {
            const double keep = hi ? v[h + i] : v[i];
            const double send = hi ? v[i] : v[h + i];
            v[i] = keep + __shfl_xor(send, bit, kWave);
        }
    }
    double t = v[0];
#pragma unroll
    for (int bit = kWave / NV / 2; bit >= 1; bit >>= 1) t += __shfl_xor(t, bit, kWave);
    return t;
}

// One task: 64 eigenvalues of one matrix, one per lane.  The pivots are not
// stored: pass F keeps the stationary pivot D+ at every segment start (every
// kSeg rows), pass B the progressive pivot D- likewise, and every later pass
// recomputes a segment's pivots in registers from its checkpoints.  Per lane
// the scratch holds three checkpoint rows per segment (D+, D-, z), so memory
// traffic is O(n / kSeg) per eigenvalue and the passes are bound by their
// dependent recurrences, not by memory.  All lanes walk the same segment at
// the same time (per-lane differences are selects), so D and E arrive as
// scalar loads and the last pass can sum each row over the wave's lanes.
//   cp: per wave 3 arrays of (segments + 1) x 64 doubles: cpF[s] = D+ and
//       cpB[s] = D- at row s*kSeg, zc[s] = z at row s*kSeg.
//   partial: per task 2 * n_b doubles at poff[b] + (t - tprefix[b]) * 2 * n_b.
//   torder: the tasks longest first (a batch's tasks by descending matrix
//       order), or nullptr for task order.  Wave w takes positions w, then
//       2 nwave - 1 - w, then 2 nwave + w, ... of that list, so the longest
//       tasks go one per wave and the second round starts on the waves that
//       hold the shortest first tasks.
__global__ __launch_bounds__(kWave * kGradWaves) void k_grad_vectors(
    int nmat, const int* __restrict__ moff, const int* __restrict__ tprefix,
    const long long* __restrict__ poff, const int* __restrict__ torder, int ntask, int n_one,
    const double* __restrict__ d, const double* __restrict__ e, const double* __restrict__ w,
    const double* __restrict__ g, double* __restrict__ cpbuf, long long slot,
    double* __restrict__ partial) {
    __shared__ double stage[kGradWaves][2][kWin];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = uniform((int)(threadIdx.x / kWave));
    const int wave = uniform((int)(blockIdx.x * kGradWaves + wv));
    const int nwave = (int)gridDim.x * kGradWaves;
    const double safmin = 0x1p-1022;
    double* sD = stage[wv][0];
    double* sE = stage[wv][1];
    for (int round = 0;; round++) {
        const int q = round * nwave + ((round & 1) ? nwave - 1 - wave : wave);
        if (q >= ntask) break;  // positions grow with the round
        const int t = torder ? uniform(torder[q]) : q;
        const int b = uniform(grad_task_matrix(tprefix, nmat, t));
        const int off = moff ? uniform(moff[b]) : 0;
        const int n = moff ? uniform(moff[b + 1]) - off : n_one;
        const int t0 = tprefix ? uniform(tprefix[b]) : 0;
        const long long po = poff ? poff[b] : 0;
        double* __restrict__ part = partial + po + (long long)(t - t0) * 2 * n;
        const int i = (t - t0) * kWave + lane;
        const bool active = i < n;
        const CDoubles D = as_const(d + off);
        const double mu = w[off + (active ? i : 0)];
        const double gi = active ? g[off + i] : 0.0;
        // pivmin = safmin * max(1, max E^2) (DLARRD / DLAR1V)
        double emax = 0.0;
        for (int k = lane; k < n - 1; k += kWave) emax = fmax(emax, fabs(e[off + k]));
        for (int o = 32; o > 0; o >>= 1) emax = fmax(emax, __shfl_xor(emax, o, kWave));
        const double pivmin = safmin * fmax(1.0, emax * emax);
        if (n == 1) {
            if (lane == 0) {
                part[0] = gi;  // d lambda / d D = 1
                part[1] = 0.0;
            }
            continue;
        }
        const int S = (n + kSeg - 1) / kSeg;  // segments; only the last may be short
        double* __restrict__ cpF = cpbuf + (long long)wave * slot;
        double* __restrict__ cpB = cpF + (long long)(S + 1) * kWave;
        double* __restrict__ zc = cpB + (long long)(S + 1) * kWave;
        auto at = [&](int s) { return (long long)s * kWave + lane; };
        // D and E reach the segments through a window of kWin rows in this
        // wave's LDS (refilled with coalesced loads when a segment leaves it;
        // one wave's LDS operations complete in order, so no barrier).  A
        // segment reads D of rows lo..lo+16 and E of rows lo-1..lo+15, clamped
        // to the matrix (rows outside it only feed values that are never used).
        int wbase = -(1 << 30);
        auto window = [&](int lo, bool up) {
            if (lo - 1 >= wbase && lo + kSeg + 1 <= wbase + kWin) return;
            wbase = up ? lo - 1 : lo + kSeg + 1 - kWin;
            for (int m = lane; m < kWin; m += kWave) {
                const int row = min(max(wbase + m, 0), n - 1);
                sD[m] = d[off + row];
                sE[m] = e[off + min(row, n - 2)];
            }
        };
        double dv[kSeg + 1], ev[kSeg], eprev;
        auto load_seg = [&](int lo, bool up) {
            window(lo, up);
            const int o = lo - wbase;
#pragma unroll
            for (int j = 0; j <= kSeg; j++) dv[j] = sD[o + j];
#pragma unroll
            for (int j = 0; j < kSeg; j++) ev[j] = sE[o + j];
            eprev = sE[o - 1];
        };
        auto step = [&](double ek, double x, double dk) {  // dk - ek^2 / x, guarded
            return pivot_guard(fma(-ek * ek, rcp_chain(x), dk - mu), pivmin);
        };
        // P[j] = D+ of row lo+j from D+_lo
        double P[kSeg], M[kSeg];
        auto seg_dp = [&](double x) {
#pragma unroll
            for (int j = 0; j < kSeg; j++) {
                P[j] = x;
                x = step(ev[j], x, dv[j + 1]);
            }
        };
        // M[j] = D- of row lo+j: the top segment starts at D_{n-1} - mu, the
        // others from D- of row lo+16 (xhi)
        auto seg_dm = [&](int len, bool top, double xhi) {
            double x = xhi;
#pragma unroll
            for (int j = kSeg - 1; j >= 0; j--) {
                const double rec = step(ev[j], x, dv[j]);
                const double start = top ? pivot_guard(dv[j] - mu, pivmin) : rec;
                x = (j > len - 1) ? x : ((j == len - 1) ? start : rec);
                M[j] = x;
            }
        };
        // pass F: D+ chain over all rows, checkpoint at segment starts
        {
            double x = pivot_guard(D[0] - mu, pivmin);
            for (int s = 0; s < S; s++) {
                const int lo = s * kSeg;
                load_seg(lo, true);
                cpF[at(s)] = x;
                seg_dp(x);
                // D+ of row lo+16 continues the chain (P holds rows lo..lo+15)
                x = step(ev[kSeg - 1], P[kSeg - 1], dv[kSeg]);
            }
        }
        // pass B: D- chain downward; twist index r = argmin |gamma_k|
        double gbest = HUGE_VAL;
        int r = n - 1;
        {
            double xhi = 0.0;
            double cpn = cpF[at(S - 1)];
            for (int s = S - 1; s >= 0; s--) {
                const int lo = s * kSeg, len = min(kSeg, n - lo);
                const double cpc = cpn;
                if (s > 0) cpn = cpF[at(s - 1)];  // next segment's checkpoint in flight
                load_seg(lo, false);
                seg_dp(cpc);
                seg_dm(len, s == S - 1, xhi);
#pragma unroll
                for (int j = kSeg - 1; j >= 0; j--) {
                    const double gam = fabs(P[j] + M[j] - (dv[j] - mu));
                    const bool better = (j < len) && gam < gbest;
                    gbest = better ? gam : gbest;
                    r = better ? lo + j : r;
                }
                cpB[at(s)] = M[0];
                xhi = M[0];
            }
        }
        // pass Z: z with z_r = 1, |z|^2, and zc[s] = z at each segment start
        double nrm = 1.0;
        {
            // rows below r: z_k = -(E_k / D+_k) z_{k+1}, segments from the top down
            double z = 1.0;
            double cpn = cpF[at(S - 1)];
            for (int s = S - 1; s >= 0; s--) {
                const int lo = s * kSeg, len = min(kSeg, n - lo);
                const double cpc = cpn;
                if (s > 0) cpn = cpF[at(s - 1)];
                load_seg(lo, false);
                seg_dp(cpc);
#pragma unroll
                for (int j = kSeg - 1; j >= 0; j--) {
                    const int k = lo + j;
                    const double zn = -(ev[j] * rcp_chain(P[j])) * z;
                    z = (j >= len || k >= r) ? 1.0 : zn;
                    nrm += (j < len && k < r) ? zn * zn : 0.0;
                }
                if (lo <= r) zc[at(s)] = z;
            }
            // rows above r: z_k = -(E_{k-1} / D-_k) z_{k-1}, segments upward
            z = 1.0;
            double cbn = (S > 1) ? cpB[at(1)] : 0.0;
            for (int s = 0; s < S; s++) {
                const int lo = s * kSeg, len = min(kSeg, n - lo);
                const double cbc = cbn;
                if (s + 2 < S) cbn = cpB[at(s + 2)];
                load_seg(lo, true);
                seg_dm(len, s == S - 1, cbc);
#pragma unroll
                for (int j = 0; j < kSeg; j++) {
                    const int k = lo + j;
                    const double ek = (j == 0) ? eprev : ev[j - 1];
                    const double zn = -(ek * rcp_chain(M[j])) * z;
                    const bool on = (j < len) && k > r;
                    z = (k <= r) ? 1.0 : (j < len ? zn : z);
                    nrm += on ? zn * zn : 0.0;
                    if (j == 0 && lo > r) zc[at(s)] = z;
                }
            }
        }
        // pass C: contributions g_i z_k^2 / |z|^2 and 2 g_i z_k z_{k+1} / |z|^2,
        // summed over the wave's eigenvalues, one segment at a time
        const double sc = active ? gi / nrm : 0.0;
        // checkpoints of the next two segments are in flight (F: D+, B: D- of
        // the row after the segment, Z: z at the segment start)
        auto ldF = [&](int q) { return (q < S) ? cpF[at(q)] : 0.0; };
        auto ldB = [&](int q) { return (q < S) ? cpB[at(q)] : 0.0; };
        auto ldZ = [&](int q) { return (q < S) ? zc[at(q)] : 0.0; };
        double f0 = ldF(0), f1 = ldF(1), b0 = ldB(1), b1 = ldB(2), z0 = ldZ(0), z1 = ldZ(1),
               z2 = ldZ(2);
        for (int s = 0; s < S; s++) {
            const int lo = s * kSeg, len = min(kSeg, n - lo);
            const double cF = f0, cB = b0, zlo = z0, zhi = z1;  // zhi: z of row lo+16
            f0 = f1;
            b0 = b1;
            z0 = z1;
            z1 = z2;
            if (s + 2 < S) {
                f1 = ldF(s + 2);
                b1 = ldB(s + 3);
                z2 = ldZ(s + 3);
            }
            load_seg(lo, true);
            seg_dp(cF);
            seg_dm(len, s == S - 1, cB);
            // below r: downward from z_{lo+16} (or from z_r = 1 inside the segment)
            double x = zhi;
#pragma unroll
            for (int j = kSeg - 1; j >= 0; j--) {
                const int k = lo + j;
                x = (j >= len || k >= r) ? 1.0 : -(ev[j] * rcp_chain(P[j])) * x;
                P[j] = x;
            }
            // at and above r: upward from z_lo (or from z_r = 1)
            x = zlo;
#pragma unroll
            for (int j = 0; j < kSeg; j++) {
                const int k = lo + j;
                if (j > 0) x = (k <= r) ? 1.0 : -(ev[j - 1] * rcp_chain(M[j])) * x;
                else x = (k <= r) ? 1.0 : zlo;
                M[j] = (k < r) ? P[j] : x;  // z of row lo+j
            }
            // v[j] = dD contribution of row lo+j, v[16+j] = dE contribution
            double v[2 * kSeg];
#pragma unroll
            for (int j = 0; j < kSeg; j++) {
                const double zn = (j + 1 < kSeg) ? M[j + 1] : zhi;
                const double znext = (j + 1 < len) ? zn : ((j + 1 == len) ? zhi : 0.0);
                v[j] = sc * M[j] * M[j];
                v[kSeg + j] = 2.0 * sc * M[j] * znext;
            }
            // reduce-scatter over the wave: at each halving the lane keeps the
            // half of its values its lane bit selects and adds the partner's
            // copy of it; afterwards lane L holds value L / kShare
            constexpr int kShare = kWave / (2 * kSeg);  // lanes holding each total
            const double total = reduce_scatter<2 * kSeg>(v, lane);
            if ((lane & (kShare - 1)) == 0) {
                const int idx = lane / kShare;
                const int row = idx & (kSeg - 1);
                if (row < len) part[(idx < kSeg ? 0 : n) + lo + row] = total;
            }
        }
    }
}

// dD_j, dE_j of every row: the sum over the matrix's tasks, in task order.
__global__ void k_grad_reduce(int rows, int nmat, const int* __restrict__ moff,
                              const int* __restrict__ tprefix, const long long* __restrict__ poff,
                              int n_one, const double* __restrict__ partial,
                              double* __restrict__ dd, double* __restrict__ de) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= rows) return;
    int b = 0, off = 0, n = n_one, nt = (n_one + kWave - 1) / kWave;
    long long po = 0;
    if (moff) {
        int lo = 0, hi = nmat;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (moff[mid] <= j) lo = mid;
            else hi = mid;
        }
        b = lo;
        off = moff[b];
        n = moff[b + 1] - off;
        nt = tprefix[b + 1] - tprefix[b];
        po = poff[b];
    }
    const int k = j - off;
    double sd = 0.0, se = 0.0;
    for (int t = 0; t < nt; t++) {
        const double* p = partial + po + (long long)t * 2 * n;
        sd += p[k];
        se += p[n + k];
    }
    dd[j] = sd;
    de[j] = (k < n - 1) ? se : 0.0;
}

}  // namespace brdc
