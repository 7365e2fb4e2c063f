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
// through a window in the wave's LDS, and the per-lane pivots are kept only
// at every kSeg-th row (checkpoints [segment][lane], coalesced) and recomputed
// in registers between them.  Each task writes its partial sums for its
// matrix's rows; a second kernel adds the tasks of each matrix in a fixed
// order (deterministic results).
#pragma once

#include "brdc_kernels.h"

namespace brdc {

constexpr int kGradWaves = 4;   // waves per block of the vector kernel
constexpr int kSeg = 16;        // rows per checkpoint segment
constexpr int kWin = 256;       // rows of D and E staged in LDS per wave
// contribution ranges per task of an n-row matrix: about kGradRangeSegs
// segments each (one range: the chains' launch does the contributions too)
constexpr int kGradRangeSegs = 256;
__host__ __device__ inline int grad_parts(int n) {
    const int segs = (n > 0 ? n : 1) + kSeg - 1;
    const int p = (segs / kSeg) / kGradRangeSegs;
    return p < 1 ? 1 : (p > 64 ? 64 : p);
}

// Checkpoint doubles of one task of an n-row matrix: (segments + 1) x 64 per
// array; 3 arrays (D+, D-, z at segment starts), and for a long matrix 5 more
// (the z scan's segment transfers: down PD, QD; up U0, AU, QU).
__host__ __device__ inline long long grad_cp_slot(int n) {
    const long long segs = ((n > 0 ? n : 1) + kSeg - 1) / kSeg;
    return (grad_parts(n) > 1 ? 8 : 3) * (segs + 1) * 64;
}
// partial products of z above this bound mark a segment's transfer as unusable
// (their squares and sums stay finite below it); the kernel takes the bound
// as an argument (scan_big) so that a smaller one can force the row walk
constexpr double kScanBig = 0x1p400;

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

// Reciprocal of a pivot for the eigenvector's ratios: v_rcp_f64 and one
// Newton step (relative error below 2^-50; the pivot chains use the same
// refinement in the factor form r (2 - x r)).  Each step of the stationary
// and progressive factorizations is then exact for data perturbed by a few
// ulp, which moves the recomputed eigenvector by O(n eps / gap) as rounding
// does.
__device__ __forceinline__ double rcp_chain(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// Sum over the wave's 64 lanes of each of 32 values v[0..31], scattered:
// returns in lane L the total of value L / 2 (both lanes of a pair agree).
// Each halving keeps, in every lane, the half of its values that one lane bit
// selects and adds the partner lane's copy of that half.  The partners differ
// in bit 5 (half swap), bit 4 (row swap), then bit 3, 2, 1 and lower bits
// (row mirror, half-row mirror, quad swap: lane i pairs with i ^ 15, i ^ 7,
// i ^ 2), and a final quad swap adds the pair i, i ^ 1; the pairings span all
// 64 lanes, so every total is over the whole wave, and no value leaves the
// VALU (the first two halvings need no select: the swap hands every lane its
// own half and the partner's copy of it).
__device__ __forceinline__ double reduce_scatter32(double (&v)[32], int lane) {
#pragma unroll
    for (int i = 0; i < 16; i++) {  // bit 5: value i (bit clear) or 16 + i (set)
        lane_swap_f64<true>(v[i], v[16 + i]);
        v[i] = v[i] + v[16 + i];
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {  // bit 4: value ... + i or ... + 8 + i
        lane_swap_f64<false>(v[i], v[8 + i]);
        v[i] = v[i] + v[8 + i];
    }
    const bool b3 = (lane & 8) != 0, b2 = (lane & 4) != 0, b1 = (lane & 2) != 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {  // bit 3, partner i ^ 15
        const double keep = b3 ? v[4 + i] : v[i];
        const double send = b3 ? v[i] : v[4 + i];
        v[i] = keep + dpp_f64<0x140>(send);
    }
#pragma unroll
    for (int i = 0; i < 2; i++) {  // bit 2, partner i ^ 7
        const double keep = b2 ? v[2 + i] : v[i];
        const double send = b2 ? v[i] : v[2 + i];
        v[i] = keep + dpp_f64<0x141>(send);
    }
    const double keep = b1 ? v[1] : v[0];  // bit 1, partner i ^ 2
    const double send = b1 ? v[0] : v[1];
    const double t = keep + dpp_f64<0x4E>(send);
    return t + dpp_f64<0xB1>(t);  // the pair i, i ^ 1
}

// Sum over the wave's 64 lanes of each of 16 values v[0..15], scattered:
// returns in lane L the total of value L / 4 (the four lanes of a quad
// agree).  The halvings pair lanes across bit 5, 4, 3, 2 as above, then the
// quad adds its lanes i ^ 2 and i ^ 1: the same partners in the same order as
// reduce_scatter32, so each total is the same sum.
__device__ __forceinline__ double reduce_scatter16(double (&v)[16], int lane) {
#pragma unroll
    for (int i = 0; i < 8; i++) {  // bit 5
        lane_swap_f64<true>(v[i], v[8 + i]);
        v[i] = v[i] + v[8 + i];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {  // bit 4
        lane_swap_f64<false>(v[i], v[4 + i]);
        v[i] = v[i] + v[4 + i];
    }
    const bool b3 = (lane & 8) != 0, b2 = (lane & 4) != 0;
#pragma unroll
    for (int i = 0; i < 2; i++) {  // bit 3, partner i ^ 15
        const double keep = b3 ? v[2 + i] : v[i];
        const double send = b3 ? v[i] : v[2 + i];
        v[i] = keep + dpp_f64<0x140>(send);
    }
    const double keep = b2 ? v[1] : v[0];  // bit 2, partner i ^ 7
    const double send = b2 ? v[0] : v[1];
    double t = keep + dpp_f64<0x141>(send);
    t = t + dpp_f64<0x4E>(t);     // partner i ^ 2
    return t + dpp_f64<0xB1>(t);  // partner i ^ 1
}

// Work items of the eigenvector kernel (k_grad_vectors).  A task is 64
// eigenvalues of one matrix, one per lane; its passes are
//   F  the stationary pivot chain D+ (top down), checkpointed at every
//      kSeg-row segment start (cpF),
//   B  the progressive chain D- (bottom up), checkpointed likewise (cpB),
//   G  gamma_k = D+_k + D-_k - (D_k - mu) and the twist r = argmin |gamma_k|
//      (segments recomputed from the checkpoints),
//   Z  z with z_r = 1: downward through D+ below r, upward through D- above
//      r, with |z|^2 and z at every segment start (zc),
//   C  the contributions g z_k^2 / |z|^2 and 2 g z_k z_{k+1} / |z|^2, summed
//      over the wave's eigenvalues per row.
// A short matrix runs all passes in one item (kGAll).  A long one (grad_parts
// above 1) spreads them over four launches, so its independent work runs on
// separate waves: the two chains (kGChainF, kGChainB), the twist search over
// row ranges (kGGamma), the two z directions (kGZDown, kGZUp), and the
// contributions over row ranges (kGContrib).
enum GradKind { kGAll, kGChainF, kGChainB, kGGamma, kGZDown, kGZUp, kGContrib };

// One task's passes over its pivots.  The pivots are not stored: every pass
// recomputes a segment's pivots in registers from its checkpoints, so memory
// traffic is O(n / kSeg) per eigenvalue and the passes are bound by their
// dependent recurrences.  All lanes walk the same segment at the same time
// (per-lane differences are selects), so D and E come from an LDS window
// shared by the wave, and the last pass can sum each row over the lanes.
//   items: the launch's items (task, kind, a, b): a, b = segment range
//       [sa, sb) (kGGamma, kGContrib) or the task's first gamma item and its
//       count (kGZDown, kGZUp); nullptr (one matrix of n_one rows): item q of
//       launch lmode is derived from the task count and parts = grad_parts.
//   cp: per task the arrays of (segments + 1) x 64 doubles at cpoff[b] + (t -
//       tprefix[b]) * slot_b (cpoff == nullptr: t * slot_b), slot_b =
//       grad_cp_slot(n_b): cpF[s] = D+ and cpB[s] = D- at row s*kSeg, zc[s] =
//       z at row s*kSeg; a long matrix adds its segments' z transfers.
//   rarr, ndarr, nuarr: per task 64 twists and the two halves of |z|^2;
//   garr, rgarr: per gamma item 64 best |gamma| values and their rows.
//   partial: per task 2 * n_b doubles at poff[b] + (t - tprefix[b]) * 2 * n_b.
// Wave w takes items w, then 2 nwave - 1 - w, then 2 nwave + w, ... so with
// the longest items listed first they go one per wave, and a second round
// starts on the waves that hold the shortest first items.
__global__ __launch_bounds__(kWave * kGradWaves) void k_grad_vectors(
    int lmode, const int4* __restrict__ items, int nitem, int parts, int nmat,
    const int* __restrict__ moff, const int* __restrict__ tprefix,
    const long long* __restrict__ poff, const long long* __restrict__ cpoff, int n_one,
    const double* __restrict__ d, const double* __restrict__ e, const double* __restrict__ w,
    const double* __restrict__ g, double* __restrict__ cpbuf, int* __restrict__ rarr,
    double* __restrict__ ndarr, double* __restrict__ nuarr, double* __restrict__ garr,
    int* __restrict__ rgarr, double* __restrict__ partial, double scan_big) {
    __shared__ double stage[kGradWaves][4][kWin];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = uniform((int)(threadIdx.x / kWave));
    const int wave = uniform((int)(blockIdx.x * kGradWaves + wv));
    const int nwave = (int)gridDim.x * kGradWaves;
    const double safmin = 0x1p-1022;
    double* sD = stage[wv][0];
    double* sE = stage[wv][1];
    double* sE2 = stage[wv][2];  // E^2
    double* sH = stage[wv][3];   // E^2 / (-pivmin), rounded as the guard step forms it
    for (int round = 0;; round++) {
        const int q = round * nwave + ((round & 1) ? nwave - 1 - wave : wave);
        if (q >= nitem) break;  // positions grow with the round
        int t, kind, ia = 0, ib = 0;
        int S1 = (n_one + kSeg - 1) / kSeg;  // one matrix: its segments
        if (items) {
            const int4 it = items[q];
            t = uniform(it.x);
            kind = uniform(it.y);
            ia = uniform(it.z);
            ib = uniform(it.w);
        } else if (lmode == 1) {
            const bool split = parts > 1;
            t = split ? q >> 1 : q;
            kind = split ? kGChainF + (q & 1) : kGAll;
        } else if (lmode == 3) {
            t = q >> 1;
            kind = kGZDown + (q & 1);
            ia = t * parts;
            ib = parts;
        } else {
            t = q / parts;
            const int c = q - t * parts;
            kind = (lmode == 2) ? kGGamma : kGContrib;
            ia = (int)(((long long)S1 * c) / parts);
            ib = (int)(((long long)S1 * (c + 1)) / parts);
        }
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
#pragma unroll 8
        for (int k = lane; k < n - 1; k += kWave) emax = fmax(emax, fabs(e[off + k]));
        for (int o = 32; o > 0; o >>= 1) emax = fmax(emax, __shfl_xor(emax, o, kWave));
        const double pivmin = safmin * fmax(1.0, emax * emax);
        if (n == 1) {
            if (lane == 0 && kind == kGAll) {
                part[0] = gi;  // d lambda / d D = 1
                part[1] = 0.0;
            }
            continue;
        }
        const int S = (n + kSeg - 1) / kSeg;  // segments; only the last may be short
        const long long slot = grad_cp_slot(n), arr = (long long)(S + 1) * kWave;
        double* __restrict__ cpF = cpbuf + (cpoff ? cpoff[b] : 0) + (long long)(t - t0) * slot;
        double* __restrict__ cpB = cpF + arr;
        double* __restrict__ zc = cpB + arr;
        // segment transfers (long matrices): written by the twist search
        double* __restrict__ tPD = zc + arr;
        double* __restrict__ tQD = tPD + arr;
        double* __restrict__ tU0 = tQD + arr;
        double* __restrict__ tAU = tU0 + arr;
        double* __restrict__ tQU = tAU + arr;
        const long long tl = (long long)t * kWave + lane;
        auto at = [&](int s) { return (long long)s * kWave + lane; };
        // D and E reach the segments through a window of kWin rows in this
        // wave's LDS (refilled with coalesced loads when a segment leaves it;
        // wave barriers order the refill's stores against the reads).  A
        // segment reads D of rows lo..lo+16 and E of rows lo-1..lo+15: D
        // clamped to the matrix, E zero outside rows 0..n-2 (rows outside the
        // matrix only feed values that are never used, except that the zero E
        // of row n-1 starts the top segment's D- chain).
        const double rpv = __builtin_amdgcn_rcp(-pivmin);
        const double wpv = fma(pivmin, rpv, 2.0);
        int wbase = -(1 << 30);
        int so = 0;  // the loaded segment's offset in the window
        auto window = [&](int lo, bool up) {
            if (lo - 1 >= wbase && lo + kSeg + 1 <= wbase + kWin) return;
            wbase = up ? lo - 1 : lo + kSeg + 1 - kWin;
            // every load in flight before the first store (n >= 2 here, so
            // the clamped E index is in range; its value is dropped outside)
            constexpr int kPer = kWin / kWave;
            double vd[kPer], ve[kPer];
#pragma unroll
            for (int q = 0; q < kPer; q++) {
                const int row = wbase + lane + q * kWave;
                vd[q] = d[off + min(max(row, 0), n - 1)];
                const double x = e[off + min(max(row, 0), n - 2)];
                ve[q] = (row >= 0 && row <= n - 2) ? x : 0.0;
            }
            // lanes read entries other lanes stored: the wave's earlier reads
            // of the window finish before these stores, and the stores before
            // the reads that follow
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < kPer; q++) {
                const double b = ve[q] * ve[q];
                sD[lane + q * kWave] = vd[q];
                sE[lane + q * kWave] = ve[q];
                sE2[lane + q * kWave] = b;
                sH[lane + q * kWave] = b * rpv;
            }
            __builtin_amdgcn_wave_barrier();
        };
        double dv[kSeg + 1], ev[kSeg], eprev;
        auto load_seg = [&](int lo, bool up) {
            window(lo, up);
            const int o = lo - wbase;
            so = o;
#pragma unroll
            for (int j = 0; j <= kSeg; j++) dv[j] = sD[o + j];
#pragma unroll
            for (int j = 0; j < kSeg; j++) ev[j] = sE[o + j];
            eprev = sE[o - 1];
        };
        // One step of a pivot chain: dk - mu - ek^2 / guard(x), row k = so + j
        // of the window (E^2 and E^2 / (-pivmin) come from the window).  The
        // chain carries unguarded values and the guard acts as a select of the
        // step from -pivmin (rg below), decided from x while its reciprocal is
        // in flight; 1/x is v_rcp_f64 refined by one Newton factor 2 - x r.
        auto stepw = [&](int j, double x, double dk) {
            const double b = sE2[so + j], h = sH[so + j], c = dk - mu;
            const double r = __builtin_amdgcn_rcp(x);
            const double w = fma(-x, r, 2.0);
            const double y = fma(-(b * r), w, c);
            const double rg = fma(-h, wpv, c);
            return (fabs(x) < pivmin) ? rg : y;
        };
        auto guard = [&](double x) { return pivot_guard(x, pivmin); };
        // P[j] = D+ of row lo+j from D+_lo; returns D+ of row lo+16 (unguarded)
        double P[kSeg], M[kSeg];
        auto seg_dp = [&](double x) {
#pragma unroll
            for (int j = 0; j < kSeg; j++) {
                P[j] = guard(x);
                x = stepw(j, x, dv[j + 1]);
            }
            return x;
        };
        // The same step with the guard applied to x first (the same values:
        // the step from -pivmin is the guarded one), also giving 1 / guard(x):
        // fewer operations, one more on the chain.
        auto stepg = [&](int j, double x, double dk, double& rinv) {
            const double c = dk - mu;
            const double xg = guard(x);
            const double r = __builtin_amdgcn_rcp(xg);
            const double w = fma(-xg, r, 2.0);
            rinv = r * w;
            return fma(-(sE2[so + j] * r), w, c);
        };
        // seg_dp with P[j] = 1 / guard(D+ of row lo+j) instead
        auto seg_dp_r = [&](double x) {
#pragma unroll
            for (int j = 0; j < kSeg; j++) x = stepg(j, x, dv[j + 1], P[j]);
            return x;
        };
        // M[j] = D- of row lo+j: the top segment starts at D_{n-1} - mu, the
        // others from D- of row lo+16 (xhi)
        auto seg_dm = [&](int len, bool top, double xhi) {
            // the top segment: E is 0 from row n-1 on, so the step into row
            // n-1 from any finite value gives D_{n-1} - mu exactly
            double x = top ? 1.0 : xhi;
#pragma unroll
            for (int j = kSeg - 1; j >= 0; j--) {
                x = stepw(j, x, dv[j]);
                M[j] = guard(x);
            }
        };
        // seg_dm with M[j] = 1 / guard(D- of row lo+j) instead: each step's
        // reciprocal of its input is the previous row's, the last one apart
        auto seg_dm_r = [&](int len, bool top, double xhi) {
            double x = top ? 1.0 : xhi;
            double dummy;
#pragma unroll
            for (int j = kSeg - 1; j >= 0; j--) x = stepg(j, x, dv[j], (j < kSeg - 1) ? M[j + 1] : dummy);
            const double xg = guard(x);
            const double r = __builtin_amdgcn_rcp(xg);
            M[0] = r * fma(-xg, r, 2.0);
        };
        // gamma over segments [sa, sb), bottom up, rows high to low: the first
        // strictly smaller |gamma| wins (the sequential scan's tie rule).
        // store_t: also each segment's z transfers for the z scan.  Down (rows
        // below the twist): z_lo = PD z_{lo+16} with the factors f_k = -E_k /
        // D+_k, and QD = the segment's sum of (z_k / z_{lo+16})^2.  Up (rows
        // above it): z_lo = U0 z_{lo-1} and z_{lo+15} = AU z_{lo-1} with the
        // factors g_k = -E_{k-1} / D-_k, and QU = the sum of (z_k / z_{lo-1})^2.
        // A partial product above scan_big (or not finite) stores Q = -1, and
        // the scan walks the rows instead.
        auto gamma_scan = [&](int sa, int sb, double& gbest, int& rbest, bool store_b,
                              bool store_t) {
            double xhi = (sb < S) ? cpB[at(sb)] : 0.0;
            double cpn = cpF[at(sb - 1)];
            for (int s = sb - 1; s >= sa; s--) {
                const int lo = s * kSeg, len = min(kSeg, n - lo);
                const double cpc = cpn;
                if (s > sa) cpn = cpF[at(s - 1)];  // next segment's checkpoint in flight
                load_seg(lo, false);
                seg_dm(len, s == S - 1, xhi);
                // D+ top down with gamma on the fly (P[j] = 1 / guard(D+_j) for
                // the transfers): the segment's best keeps its highest row among
                // equal values (<=, rows rising), and a lower segment replaces
                // the best only when strictly smaller, as the bottom-up scan
                {
                    double x = cpc, sg = HUGE_VAL;
                    int srow = rbest;
#pragma unroll
                    for (int j = 0; j < kSeg; j++) {
                        const double xg = guard(x);
                        const double gam = fabs(xg + M[j] - (dv[j] - mu));
                        const bool better = (j < len) && gam <= sg;
                        sg = better ? gam : sg;
                        srow = better ? lo + j : srow;
                        const double r = __builtin_amdgcn_rcp(xg);
                        const double w = fma(-xg, r, 2.0);
                        P[j] = r * w;
                        x = fma(-(sE2[so + j] * r), w, dv[j + 1] - mu);
                    }
                    const bool better = sg < gbest;
                    gbest = better ? sg : gbest;
                    rbest = better ? srow : rbest;
                }
                if (store_b) cpB[at(s)] = M[0];
                xhi = M[0];
                if (store_t) {
                    double pd = 1.0, qd = 0.0;
                    bool okd = true;
#pragma unroll
                    for (int j = kSeg - 1; j >= 0; j--) {
                        pd *= -(ev[j] * P[j]);
                        qd = fma(pd, pd, qd);
                        okd = okd && fabs(pd) <= scan_big;
                    }
                    double au = 1.0, qu = 0.0, u0 = 1.0;
                    bool oku = true;
#pragma unroll
                    for (int j = 0; j < kSeg; j++) {
                        au *= -(((j == 0) ? eprev : ev[j - 1]) * rcp_chain(M[j]));
                        if (j == 0) u0 = au;
                        const bool in = j < len;
                        qu = in ? fma(au, au, qu) : qu;
                        oku = oku && (!in || fabs(au) <= scan_big);
                    }
                    tPD[at(s)] = pd;
                    tQD[at(s)] = okd ? qd : -1.0;
                    tU0[at(s)] = u0;
                    tAU[at(s)] = au;
                    tQU[at(s)] = oku ? qu : -1.0;
                }
            }
        };
        // pass F
        if (kind == kGAll || kind == kGChainF) {
            double x = D[0] - mu;
            for (int s = 0; s < S; s++) {
                const int lo = s * kSeg;
                load_seg(lo, true);
                cpF[at(s)] = guard(x);
                x = seg_dp(x);  // D+ of row lo+16 continues the chain
            }
            if (kind == kGChainF) continue;
        }
        // pass B alone (the chain; the twist search comes later)
        if (kind == kGChainB) {
            double xhi = 0.0;
            for (int s = S - 1; s >= 0; s--) {
                const int lo = s * kSeg, len = min(kSeg, n - lo);
                load_seg(lo, false);
                seg_dm(len, s == S - 1, xhi);
                cpB[at(s)] = M[0];
                xhi = M[0];
            }
            continue;
        }
        int r = n - 1;
        if (kind == kGAll) {  // pass B with the twist search
            double gbest = HUGE_VAL;
            gamma_scan(0, S, gbest, r, true, false);
        }
        if (kind == kGGamma) {  // the twist search over [ia, ib)
            double gbest = HUGE_VAL;
            int rb = n - 1;
            if (ia < ib) gamma_scan(ia, ib, gbest, rb, false, true);
            garr[(long long)q * kWave + lane] = gbest;
            rgarr[(long long)q * kWave + lane] = rb;
            continue;
        }
        if (kind == kGZDown || kind == kGZUp) {  // the twist from the range results
            double gbest = HUGE_VAL;
            for (int c = ib - 1; c >= 0; c--) {
                const long long x = (long long)(ia + c) * kWave + lane;
                const double gv = garr[x];
                const bool better = gv < gbest;
                gbest = better ? gv : gbest;
                r = better ? rgarr[x] : r;
            }
        }
        // pass Z: z with z_r = 1, |z|^2, and zc[s] = z at each segment start.
        // Segments wholly on the other side of every lane's twist hold z = 1
        // and add nothing, so the walks start at the lanes' extreme twist
        // segments.
        const int sr = r / kSeg;  // this lane's twist segment
        auto wave_max = [&](int v) {
            for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
            return uniform(v);
        };
        // rows below r in segments s_hi down to s_lo: z_k = -(E_k / D+_k) z_{k+1}
        auto walk_down = [&](int s_hi, int s_lo, double& z, double& nrm) {
            double cpn = cpF[at(s_hi)];
            for (int s = s_hi; s >= s_lo; s--) {
                const int lo = s * kSeg, len = min(kSeg, n - lo);
                const double cpc = cpn;
                if (s > s_lo) cpn = cpF[at(s - 1)];
                load_seg(lo, false);
                seg_dp_r(cpc);
#pragma unroll
                for (int j = kSeg - 1; j >= 0; j--) {
                    const int k = lo + j;
                    const double zn = -(ev[j] * P[j]) * z;
                    z = (j >= len || k >= r) ? 1.0 : zn;
                    nrm += (j < len && k < r) ? zn * zn : 0.0;
                }
                if (lo <= r) zc[at(s)] = z;
            }
        };
        // rows above r in segments s_lo up to s_hi: z_k = -(E_{k-1} / D-_k) z_{k-1}
        auto walk_up = [&](int s_lo, int s_hi, double& z, double& nrm) {
            double cbn = (s_lo + 1 < S) ? cpB[at(s_lo + 1)] : 0.0;
            for (int s = s_lo; s <= s_hi; s++) {
                const int lo = s * kSeg, len = min(kSeg, n - lo);
                const double cbc = cbn;
                if (s + 2 < S) cbn = cpB[at(s + 2)];
                load_seg(lo, true);
                seg_dm_r(len, s == S - 1, cbc);
#pragma unroll
                for (int j = 0; j < kSeg; j++) {
                    const int k = lo + j;
                    const double ek = (j == 0) ? eprev : ev[j - 1];
                    const double zn = -(ek * M[j]) * z;
                    const bool on = (j < len) && k > r;
                    z = (k <= r) ? 1.0 : (j < len ? zn : z);
                    nrm += on ? zn * zn : 0.0;
                    if (j == 0 && lo > r) zc[at(s)] = z;
                }
            }
        };
        // A long matrix's scan (kGZDown, kGZUp): each lane's twist segment is
        // walked row by row (the wave takes its lanes' twist segments one
        // after another), the segments beyond it are crossed with their
        // transfers, kScanBlk at a time with the next block's loads in flight.
        // An unusable transfer (Q < 0) sends the wave back to the row walk.
        constexpr int kScanBlk = 8;
        double nrm = 1.0;
        if (kind == kGAll || kind == kGZDown) {
            const int smax = wave_max(sr);
            double z = 1.0;
            if (kind == kGAll) {
                walk_down(smax, 0, z, nrm);
            } else {
                bool todo = true;
                for (;;) {
                    const int s = wave_max(todo ? sr : -1);
                    if (s < 0) break;
                    double zs = 1.0, ns = 1.0;
                    walk_down(s, s, zs, ns);
                    if (todo && sr == s) {
                        z = zs;
                        nrm = ns;
                        todo = false;
                    }
                }
                double pb[kScanBlk], qb[kScanBlk], pn[kScanBlk], qn[kScanBlk];
                auto fetch = [&](int s0, double(&p)[kScanBlk], double(&q)[kScanBlk]) {
#pragma unroll
                    for (int i = 0; i < kScanBlk; i++) {
                        const int s = max(s0 - i, 0);
                        p[i] = tPD[at(s)];
                        q[i] = tQD[at(s)];
                    }
                };
                bool bad = false;
                fetch(smax - 1, pb, qb);
                for (int s0 = smax - 1; s0 >= 0; s0 -= kScanBlk) {
                    fetch(s0 - kScanBlk, pn, qn);
#pragma unroll
                    for (int i = 0; i < kScanBlk; i++) {
                        const int s = s0 - i;
                        if (s >= 0 && s < sr) {
                            nrm = fma(z, z * qb[i], nrm);
                            z = pb[i] * z;
                            zc[at(s)] = z;
                            bad = bad || qb[i] < 0.0;
                        }
                    }
#pragma unroll
                    for (int i = 0; i < kScanBlk; i++) {
                        pb[i] = pn[i];
                        qb[i] = qn[i];
                    }
                }
                if (__any(bad)) {
                    z = 1.0;
                    nrm = 1.0;
                    walk_down(smax, 0, z, nrm);
                }
                rarr[tl] = r;
                ndarr[tl] = nrm;
                continue;
            }
        }
        if (kind == kGAll || kind == kGZUp) {
            if (kind == kGZUp) nrm = 0.0;
            const int smin = -wave_max(-sr);
            double z = 1.0;
            if (kind == kGAll) {
                walk_up(smin, S - 1, z, nrm);
            } else {
                bool todo = true;
                for (;;) {
                    const int s = -wave_max(todo ? -sr : -S);
                    if (s >= S) break;
                    double zs = 1.0, ns = 0.0;
                    walk_up(s, s, zs, ns);
                    if (todo && sr == s) {
                        z = zs;
                        nrm = ns;
                        todo = false;
                    }
                }
                double ub[kScanBlk], ab[kScanBlk], qb[kScanBlk];
                double un[kScanBlk], an[kScanBlk], qn[kScanBlk];
                auto fetch = [&](int s0, double(&u)[kScanBlk], double(&a)[kScanBlk],
                                 double(&q)[kScanBlk]) {
#pragma unroll
                    for (int i = 0; i < kScanBlk; i++) {
                        const int s = min(s0 + i, S - 1);
                        u[i] = tU0[at(s)];
                        a[i] = tAU[at(s)];
                        q[i] = tQU[at(s)];
                    }
                };
                bool bad = false;
                fetch(smin + 1, ub, ab, qb);
                for (int s0 = smin + 1; s0 < S; s0 += kScanBlk) {
                    fetch(s0 + kScanBlk, un, an, qn);
#pragma unroll
                    for (int i = 0; i < kScanBlk; i++) {
                        const int s = s0 + i;
                        if (s < S && s > sr) {
                            nrm = fma(z, z * qb[i], nrm);
                            zc[at(s)] = ub[i] * z;
                            z = ab[i] * z;
                            bad = bad || qb[i] < 0.0;
                        }
                    }
#pragma unroll
                    for (int i = 0; i < kScanBlk; i++) {
                        ub[i] = un[i];
                        ab[i] = an[i];
                        qb[i] = qn[i];
                    }
                }
                if (__any(bad)) {
                    z = 1.0;
                    nrm = 0.0;
                    walk_up(smin, S - 1, z, nrm);
                }
                nuarr[tl] = nrm;
                continue;
            }
        }
        // pass C: all segments (kGAll) or [ia, ib) (kGContrib)
        if (kind == kGContrib) {
            r = rarr[tl];
            nrm = ndarr[tl] + nuarr[tl];
        }
        const double sc = active ? gi / nrm : 0.0;
        const int sa = (kind == kGAll) ? 0 : ia;
        const int sb = (kind == kGAll) ? S : ib;
        // contributions g_i z_k^2 / |z|^2 and 2 g_i z_k z_{k+1} / |z|^2, summed
        // over the wave's eigenvalues, one segment at a time; the checkpoints of
        // the next two segments are in flight (F: D+, B: D- of the row after
        // the segment, Z: z at the segment start)
        auto ldF = [&](int x) { return (x < S) ? cpF[at(x)] : 0.0; };
        auto ldB = [&](int x) { return (x < S) ? cpB[at(x)] : 0.0; };
        auto ldZ = [&](int x) { return (x < S) ? zc[at(x)] : 0.0; };
        double f0 = ldF(sa), f1 = ldF(sa + 1), b0 = ldB(sa + 1), b1 = ldB(sa + 2), z0 = ldZ(sa),
               z1 = ldZ(sa + 1), z2 = ldZ(sa + 2);
        for (int s = sa; s < sb; s++) {
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
            seg_dp_r(cF);
            seg_dm_r(len, s == S - 1, cB);
            // below r: downward from z_{lo+16} (or from z_r = 1 inside the segment)
            double x = zhi;
#pragma unroll
            for (int j = kSeg - 1; j >= 0; j--) {
                const int k = lo + j;
                x = (j >= len || k >= r) ? 1.0 : -(ev[j] * P[j]) * x;
                P[j] = x;
            }
            // at and above r: upward from z_lo (or from z_r = 1)
            x = zlo;
#pragma unroll
            for (int j = 0; j < kSeg; j++) {
                const int k = lo + j;
                if (j > 0) x = (k <= r) ? 1.0 : -(ev[j - 1] * M[j]) * x;
                else x = (k <= r) ? 1.0 : zlo;
                M[j] = (k < r) ? P[j] : x;  // z of row lo+j
            }
            // the row sums over the wave's eigenvalues, dD then dE, each a
            // reduce-scatter of 16 values: afterwards lane L holds the total
            // of row L / 4.  dE is summed as sc z_k z_{k+1} and doubled at the
            // end (scaling by 2 is exact, so this is the sum of the doubled
            // terms).
            static_assert(kSeg == 16, "the reduce-scatter takes 16 values");
            const int row = lane >> 2;
            const bool writer = (lane & 3) == 0 && row < len;
            {
                double v[kSeg];
#pragma unroll
                for (int j = 0; j < kSeg; j++) v[j] = (sc * M[j]) * M[j];
                const double total = reduce_scatter16(v, lane);
                if (writer) part[lo + row] = total;
            }
            {
                double v[kSeg];
#pragma unroll
                for (int j = 0; j < kSeg; j++) {
                    const double zn = (j + 1 < kSeg) ? M[j + 1] : zhi;
                    const double znext = (j + 1 < len) ? zn : ((j + 1 == len) ? zhi : 0.0);
                    v[j] = (sc * M[j]) * znext;
                }
                const double total = reduce_scatter16(v, lane);
                if (writer) part[n + lo + row] = 2.0 * total;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Numerically degenerate eigenvalues.  Twisted factorization at an eigenvalue
// of a cluster whose members agree to within rounding returns one vector for
// all of them (a matrix split into identical blocks: both copies of lambda
// pick the same twist), and those vectors are neither distinct nor
// orthogonal.  Consecutive eigenvalues closer than kClusterRel * ||T||
// (||T|| = max |lambda|) form a cluster; its members are taken out of the
// twisted-factorization pass (their weight g_i is zeroed there) and their
// vectors are recomputed by inverse iteration with Gram-Schmidt against the
// cluster's earlier members -- the method of LAPACK DSTEIN: LU with partial
// pivoting of T - x I (the DLAGTF role), solves with small pivots perturbed
// (the DLAGTS role), start vectors from a fixed hash, shifts of equal
// members spread by 10 eps |x|.  Their contributions are added to the partial
// sums of the task holding the cluster's first member, one wave per task, so
// the sums stay deterministic.  A cluster of more than kClusterMax members is
// orthogonalised in groups of kClusterMax.
constexpr double kClusterRel = 0x1p-44;
constexpr int kClusterMax = 32;
constexpr int kClusterWaves = 32;   // waves (one-wave blocks) of the cluster kernel
constexpr int kClusterIters = 4;    // inverse-iteration solves per member
constexpr int kClusterArrays = 6 + kClusterMax;  // a, b, c, d, x, pivots, vectors

// this matrix's cluster tolerance from its sorted eigenvalues
__device__ __forceinline__ double cluster_tol(const double* __restrict__ w, int n) {
    return kClusterRel * fmax(fabs(w[0]), fabs(w[n - 1]));
}

// Is eigenvalue k (of n, sorted, w the matrix's own) the first member of a
// cluster?
__device__ __forceinline__ bool cluster_starts(const double* __restrict__ w, int n, int k,
                                               double tol) {
    return (k == 0 || w[k] - w[k - 1] > tol) && (k + 1 < n && w[k + 1] - w[k] <= tol);
}

// gm[j] = g[j], or 0 for a member of a cluster (rows numbered as in the
// batch layout; moff == nullptr: one matrix of n_one rows); tflag[t] = 1 when
// a cluster starts among task t's 64 eigenvalues.  One wave per task, lane l
// taking the task's eigenvalue l.
__global__ __launch_bounds__(256) void k_grad_mask(int ntask, int nmat, const int* __restrict__ moff,
                                                   const int* __restrict__ tprefix, int n_one,
                                                   const double* __restrict__ w,
                                                   const double* __restrict__ g,
                                                   double* __restrict__ gm,
                                                   int* __restrict__ tflag) {
    const int t = uniform((int)(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave));
    const int lane = threadIdx.x & (kWave - 1);
    if (t >= ntask) return;
    const int b = uniform(grad_task_matrix(tprefix, nmat, t));
    const int off = moff ? uniform(moff[b]) : 0;
    const int n = moff ? uniform(moff[b + 1]) - off : n_one;
    const int t0 = tprefix ? uniform(tprefix[b]) : 0;
    const int k = (t - t0) * kWave + lane;
    const double* W = w + off;
    bool starts = false;
    if (k < n) {
        const double tol = cluster_tol(W, n);
        const double lo = (k > 0) ? W[k] - W[k - 1] : HUGE_VAL;
        const double hi = (k + 1 < n) ? W[k + 1] - W[k] : HUGE_VAL;
        gm[off + k] = (lo <= tol || hi <= tol) ? 0.0 : g[off + k];
        starts = lo > tol && hi <= tol;
    }
    const unsigned long long any = __ballot(starts);
    if (lane == 0) tflag[t] = any ? 1 : 0;
}

// uniform in (-1, 1) from (row, member): the start vector of an iteration
__device__ __forceinline__ double cluster_start(int k, int m) {
    unsigned x = (unsigned)k * 0x9E3779B1u ^ ((unsigned)m + 0x7F4A7C15u) * 0x85EBCA77u;
    x ^= x >> 15;
    x *= 0x2C1B3C6Du;
    x ^= x >> 12;
    x *= 0x297A2D39u;
    x ^= x >> 15;
    return ((double)(x >> 8) + 0.5) * 0x1p-23 - 1.0;
}

// One wave per task with a cluster (tflag), every cluster whose first member
// lies in the task.  Wave x looks at the tasks 64 at a time, tasks [64 x, 64 x
// + 64), then 64 (x + waves) on.  Lane 0 runs the sequential factorization
// and solves; the wave shares the dot products, normalisations and
// contributions.  scratch: kClusterArrays * nmax doubles per wave.
__global__ __launch_bounds__(kWave) void k_grad_clusters(
    int ntask, int nmat, const int* __restrict__ moff, const int* __restrict__ tprefix,
    const long long* __restrict__ poff, int n_one, int nmax, const double* __restrict__ d,
    const double* __restrict__ e, const double* __restrict__ w, const double* __restrict__ g,
    const int* __restrict__ tflag, double* __restrict__ scratch, double* __restrict__ partial) {
    const int lane = threadIdx.x;
    double* const la = scratch + (size_t)blockIdx.x * kClusterArrays * nmax;  // U diagonal
    double* const lb = la + nmax;       // U first superdiagonal
    double* const lc = lb + nmax;       // multipliers
    double* const ld = lc + nmax;       // U second superdiagonal
    double* const lx = ld + nmax;       // the iterate
    int* const lp = (int*)(lx + nmax);  // row-swap flags (ints in a double-sized slot)
    double* const vec = lx + 2 * (size_t)nmax;  // the group's finished vectors
    for (int tb = blockIdx.x * kWave; tb < ntask; tb += gridDim.x * kWave) {
      unsigned long long todo = __ballot(tb + lane < ntask && tflag[tb + lane] != 0);
      while (todo) {
        const int t = tb + __builtin_ctzll(todo);
        todo &= todo - 1;
        const int b = uniform(grad_task_matrix(tprefix, nmat, t));
        const int off = moff ? uniform(moff[b]) : 0;
        const int n = moff ? uniform(moff[b + 1]) - off : n_one;
        const int t0 = tprefix ? uniform(tprefix[b]) : 0;
        if (n < 2) continue;
        const double* __restrict__ W = w + off;
        const double* __restrict__ D = d + off;
        const double* __restrict__ E = e + off;
        double* __restrict__ part = partial + (poff ? poff[b] : 0) + (long long)(t - t0) * 2 * n;
        const double tol = cluster_tol(W, n);
        double tnorm = 0.0;  // max-row-sum norm, for the pivot threshold
        for (int k = lane; k < n; k += kWave) {
            const double s = fabs(D[k]) + (k > 0 ? fabs(E[k - 1]) : 0.0) +
                             (k + 1 < n ? fabs(E[k]) : 0.0);
            tnorm = fmax(tnorm, s);
        }
        for (int o = 32; o > 0; o >>= 1) tnorm = fmax(tnorm, __shfl_xor(tnorm, o, kWave));
        const double ptol = fmax(kEps * tnorm, 0x1p-1022);
        const int i0 = (t - t0) * kWave, i1 = min(i0 + kWave, n);
        for (int cs = i0; cs < i1; cs++) {
            if (!cluster_starts(W, n, cs, tol)) continue;
            int ce = cs + 1;
            while (ce + 1 < n && W[ce + 1] - W[ce] <= tol) ce++;
            double xprev = 0.0;
            for (int m = cs; m <= ce; m++) {
                const int grp = (m - cs) % kClusterMax;  // position in its group
                if (tnorm == 0.0) {
                    // T = 0: every vector is an eigenvector; the unit vectors
                    // e_(m - cs) make the cluster's basis orthonormal
                    const double gmv = g[off + m];
                    if (lane == 0) part[(m - cs) % n] += gmv;
                    continue;
                }
                // the shift: equal members spread by 10 eps |x| (DSTEIN)
                double x = W[m];
                if (m > cs) x = fmax(x, xprev + 10.0 * kEps * fabs(x));
                xprev = x;
                if (lane == 0) {
                    // LU of T - x I with partial pivoting: entering step k,
                    // row k holds a0 (column k) and b0 (column k + 1); the
                    // pivot is the larger of a0 and E_k relative to their
                    // rows (DLAGTF's rule).  U gets la (diagonal), lb, ld
                    // (super-diagonals), L the multipliers lc and swaps lp.
                    double a0 = D[0] - x, b0 = E[0];
                    double s1 = fabs(a0) + fabs(b0);
                    for (int k = 0; k < n - 1; k++) {
                        const double a1 = D[k + 1] - x, c = E[k];
                        const double bn = (k + 1 < n - 1) ? E[k + 1] : 0.0;
                        const double s2 = fabs(c) + fabs(a1) + fabs(bn);
                        const double piv1 = (a0 == 0.0) ? 0.0 : fabs(a0) / s1;
                        const double piv2 = (c == 0.0) ? 0.0 : fabs(c) / s2;
                        if (c == 0.0 || piv2 <= piv1) {
                            const double mult = (c == 0.0) ? 0.0 : c / a0;
                            la[k] = a0;
                            lb[k] = b0;
                            ld[k] = 0.0;
                            lc[k] = mult;
                            lp[k] = 0;
                            a0 = a1 - mult * b0;
                            b0 = bn;
                            s1 = s2;
                        } else {
                            const double mult = a0 / c;
                            la[k] = c;
                            lb[k] = a1;
                            ld[k] = bn;
                            lc[k] = mult;
                            lp[k] = 1;
                            a0 = b0 - mult * a1;
                            b0 = -mult * bn;
                        }
                    }
                    la[n - 1] = a0;
                }
                // start vector
                for (int k = lane; k < n; k += kWave) lx[k] = cluster_start(k, m - cs);
                __builtin_amdgcn_wave_barrier();
                __threadfence();
                for (int it = 0; it < kClusterIters; it++) {
                    // right-hand side of size ptol: the solution's entries stay
                    // near 1 however small the pivots, so nothing overflows
                    for (int k = lane; k < n; k += kWave) lx[k] *= ptol;
                    __builtin_amdgcn_wave_barrier();
                    __threadfence();
                    if (lane == 0) {
                        // forward: apply the row swaps and multipliers
                        for (int k = 0; k < n - 1; k++) {
                            if (lp[k]) {
                                const double tk = lx[k];
                                lx[k] = lx[k + 1];
                                lx[k + 1] = tk - lc[k] * lx[k + 1];
                            } else {
                                lx[k + 1] -= lc[k] * lx[k];
                            }
                        }
                        // back substitution, pivots below ptol perturbed
                        for (int k = n - 1; k >= 0; k--) {
                            double tmp = lx[k];
                            if (k + 1 < n) tmp -= lb[k] * lx[k + 1];
                            if (k + 2 < n) tmp -= ld[k] * lx[k + 2];
                            double ak = la[k];
                            if (fabs(ak) < ptol) ak = copysign(ptol, ak);
                            lx[k] = tmp / ak;
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                    __threadfence();
                    // Gram-Schmidt against the group's earlier vectors, twice
                    for (int pass = 0; pass < 2; pass++) {
                        for (int q = 0; q < grp; q++) {
                            const double* v = vec + (size_t)q * nmax;
                            double dot = 0.0;
                            for (int k = lane; k < n; k += kWave) dot = fma(v[k], lx[k], dot);
                            for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o, kWave);
                            for (int k = lane; k < n; k += kWave) lx[k] = fma(-dot, v[k], lx[k]);
                        }
                    }
                    // normalise: scale by the largest entry first, then the 2-norm
                    double amax = 0.0;
                    for (int k = lane; k < n; k += kWave) amax = fmax(amax, fabs(lx[k]));
                    for (int o = 32; o > 0; o >>= 1) amax = fmax(amax, __shfl_xor(amax, o, kWave));
                    const double inv = (amax > 0.0) ? 1.0 / amax : 1.0;
                    double ss = 0.0;
                    for (int k = lane; k < n; k += kWave) {
                        const double y = lx[k] * inv;
                        lx[k] = y;
                        ss = fma(y, y, ss);
                    }
                    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o, kWave);
                    const double rn = 1.0 / sqrt(ss);
                    for (int k = lane; k < n; k += kWave) lx[k] *= rn;
                    __builtin_amdgcn_wave_barrier();
                    __threadfence();
                }
                // keep the vector for the group's later members, and add
                // g_m v_k^2 and 2 g_m v_k v_{k+1} to the task's partial sums
                const double gmv = g[off + m];
                double* v = vec + (size_t)grp * nmax;
                for (int k = lane; k < n; k += kWave) v[k] = lx[k];
                __builtin_amdgcn_wave_barrier();
                __threadfence();
                for (int k = lane; k < n; k += kWave) {
                    const double vk = v[k];
                    part[k] = fma(gmv * vk, vk, part[k]);
                    if (k + 1 < n) part[n + k] = fma(2.0 * gmv * vk, v[k + 1], part[n + k]);
                }
                __builtin_amdgcn_wave_barrier();
                __threadfence();
            }
        }
      }
    }
}

// The per-matrix sums of the tasks' partial sums: block of 256 threads = 32
// rows x 8 slices of the matrix's tasks; each slice sums its contiguous run
// of tasks in order and the slices meet in LDS in slice order, so the result
// does not depend on the launch shape.
constexpr int kReduceRows = 32, kReduceSlices = 8;
__global__ __launch_bounds__(kReduceRows * kReduceSlices) void k_grad_reduce(
    int rows, int nmat, const int* __restrict__ moff, const int* __restrict__ tprefix,
    const long long* __restrict__ poff, int n_one, const double* __restrict__ partial,
    double* __restrict__ dd, double* __restrict__ de) {
    __shared__ double red[2][kReduceSlices][kReduceRows];
    const int r = threadIdx.x % kReduceRows, sl = threadIdx.x / kReduceRows;
    const int j = blockIdx.x * kReduceRows + r;
    double sd = 0.0, se = 0.0;
    int k = 0, n = 1;
    if (j < rows) {
        int b = 0, off = 0, nt = (n_one + kWave - 1) / kWave;
        long long po = 0;
        n = n_one;
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
        k = j - off;
        const int ta = (int)(((long long)nt * sl) / kReduceSlices);
        const int tb = (int)(((long long)nt * (sl + 1)) / kReduceSlices);
        // unrolled so that the loads of several tasks are in flight; the adds
        // keep the task order
#pragma unroll 4
        for (int t = ta; t < tb; t++) {
            const double* p = partial + po + (long long)t * 2 * n;
            sd += p[k];
            se += p[n + k];
        }
    }
    red[0][sl][r] = sd;
    red[1][sl][r] = se;
    __syncthreads();
    if (sl == 0 && j < rows) {
        double ad = red[0][0][r], ae = red[1][0][r];
        for (int q = 1; q < kReduceSlices; q++) {
            ad += red[0][q][r];
            ae += red[1][q][r];
        }
        dd[j] = ad;
        de[j] = (k < n - 1) ? ae : 0.0;
    }
}

}  // namespace brdc
