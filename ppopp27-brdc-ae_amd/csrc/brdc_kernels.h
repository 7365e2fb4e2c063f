// brdc_kernels.h -- gfx950 kernels of the BR values-only D&C solve.
//
// Data layout in HBM (all FP64 unless noted, one slot per matrix row):
//   state A/B  (d, blo, bhi)  node eigenvalues ascending + first/last rows of
//              the node's eigenvector matrix; ping-ponged between levels.
//              This is the whole persistent BR state (dlaed0_br.f:112-119).
//   merge scratch, reused by every level, indexed by the merge's first row s:
//     m*   merged children sorted by value (d, z, q1, q2)          [prepare]
//     a*   non-deflated poles in order (d, z, q1, q2), K entries    [deflate]
//     g*   close-pole (Givens) deflated values + rows, G entries     [deflate]
//     t*   tiny-z deflated values + rows, T entries                  [deflate]
//     r*   per root: lambda, tau, near1, near2 (+ int org)          [secular]
//     w    product-form update vector                               [wprod]
//     qo*  new boundary rows of the roots                           [rows]
//   counts: 4 ints per merge (K, G, T, secular evaluations)
// Every kernel derives its merge geometry from (n, depth, merge index), so a
// level is a fixed launch sequence with no host round trip.
#pragma once

#include <hip/hip_runtime.h>

#include "brdc_secular.h"
#include "brdc_tree.h"

namespace brdc {

constexpr int kWave = 64;
constexpr int kMaxLeaf = 64;
constexpr int kJacobiSweeps = 40;
constexpr int kSpecCap = 256;     // speculative Givens-run walk bound
constexpr int kThreads = 256;     // block size of the element/root-parallel kernels

// ---------------------------------------------------------------------------
// block-wide scans (wave64 shuffles + one LDS pass)
template <int T>
__device__ __forceinline__ int block_exclusive_sum(int v, int* total, int* lds) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    int x = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) lds[wid] = x;
    __syncthreads();
    if (wid == 0) {
        constexpr int NW = T / kWave;
        const int wv = (lane < NW) ? lds[lane] : 0;
        int ws = wv;
#pragma unroll
        for (int o = 1; o < NW; o <<= 1) {
            const int y = __shfl_up(ws, o, kWave);
            if (lane >= o) ws += y;
        }
        if (lane < NW) lds[lane] = ws - wv;
        if (lane == NW - 1) lds[NW] = ws;
    }
    __syncthreads();
    const int res = x - v + lds[wid];
    *total = lds[T / kWave];
    __syncthreads();
    return res;
}

template <int T>
__device__ __forceinline__ int block_exclusive_max(int v, int* lds) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    int x = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(x, o, kWave);
        if (lane >= o) x = max(x, y);
    }
    int excl = __shfl_up(x, 1, kWave);
    if (lane == 0) excl = INT_MIN;
    if (lane == kWave - 1) lds[wid] = x;
    __syncthreads();
    if (wid == 0) {
        constexpr int NW = T / kWave;
        const int wv = (lane < NW) ? lds[lane] : INT_MIN;
        int ws = wv;
#pragma unroll
        for (int o = 1; o < NW; o <<= 1) {
            const int y = __shfl_up(ws, o, kWave);
            if (lane >= o) ws = max(ws, y);
        }
        int we = __shfl_up(ws, 1, kWave);
        if (lane == 0) we = INT_MIN;
        if (lane < NW) lds[lane] = we;
    }
    __syncthreads();
    const int res = max(excl, lds[wid]);
    __syncthreads();
    return res;
}

__device__ __forceinline__ void report_failure(int* info, int n, int start, int size) {
    // dlaed0_br.f:302-303 encoding with a 1-based first row
    const int r = start + 1;
    const int code = r * (n + 1) + r + size - 1;
    atomicCAS(info, 0, code);
}

// ---------------------------------------------------------------------------
// max-norm, finiteness and scale (DLANST('M') + DLASCL, dstedc_dc_br.f:86-118)
__global__ __launch_bounds__(1024) void k_norm(int n, const double* __restrict__ d,
                                               const double* __restrict__ e, double* scal,
                                               int* info, int* info_user) {
    __shared__ double red[16];
    __shared__ int bad[16];
    double m = 0.0;
    int nonfinite = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double x = d[i];
        nonfinite |= !isfinite(x);
        m = fmax(m, fabs(x));
        if (i < n - 1) {
            const double y = e[i];
            nonfinite |= !isfinite(y);
            m = fmax(m, fabs(y));
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        m = fmax(m, __shfl_down(m, o, kWave));
        nonfinite |= __shfl_down(nonfinite, o, kWave);
    }
    const int wid = threadIdx.x / kWave;
    if ((threadIdx.x & 63) == 0) {
        red[wid] = m;
        bad[wid] = nonfinite;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0;
        int bb = 0;
        for (int w = 0; w < (int)(blockDim.x / kWave); w++) {
            mm = fmax(mm, red[w]);
            bb |= bad[w];
        }
        const bool ok = mm > 0.0;
        scal[0] = ok ? 1.0 / mm : 1.0;  // scale into [-1, 1]
        scal[1] = ok ? mm : 1.0;        // undo
        *info = bb ? n + 1 : 0;
        if (info_user && n <= 1) *info_user = *info;
    }
}

// scaled off-diagonal with negligible entries zeroed (dstedc_dc_br.f:94-105):
// a zero there splits the matrix, which the tree then handles as a rho = 0
// merge (everything deflates) or inside a leaf
__global__ void k_prep_e(int n, const double* __restrict__ d, const double* __restrict__ e,
                         const double* __restrict__ scal, double* ew) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const double tiny = kEps * sqrt(fabs(d[i])) * sqrt(fabs(d[i + 1]));
    const double x = e[i];
    ew[i] = (fabs(x) > tiny) ? x * scal[0] : 0.0;
}

// ---------------------------------------------------------------------------
// Leaf eigensolve: one wavefront per leaf (<= MAXM rows) holding the dense
// leaf in LDS.  Cyclic Jacobi in round-robin (tournament) order: each round
// picks m/2 disjoint (p, q) pairs, computes their rotations, and applies
// A <- J^T A J as independent 2x2 block updates -- block (a, b) of the pair
// structure becomes R_a^T A[P_a, P_b] R_b -- so a round is one LDS pass with
// no intermediate barrier.  Only the first and last rows of the accumulated
// rotation product are kept (the role of DSTEQR_BOUNDARY_BR,
// dlaed0_br.f:354-371).  The rank-one cut corrections (dlaed0_br.f:95-103)
// are applied to the leaf's end rows on load.
template <int MAXM>
__global__ __launch_bounds__(kWave) void k_leaf(int n, int depth, const double* __restrict__ d_in,
                                                const double* __restrict__ ew,
                                                const double* __restrict__ scal, double* out_d,
                                                double* out_lo, double* out_hi, int* info,
                                                int single, int* info_user) {
    constexpr int LD = MAXM + 1;
    __shared__ double A[MAXM * LD];
    __shared__ double V0[MAXM], V1[MAXM];
    __shared__ double RC[MAXM / 2], RS[MAXM / 2];
    __shared__ int RP[MAXM / 2], RQ[MAXM / 2];
    __shared__ int changed;
    int s, m;
    node_geom(n, depth, blockIdx.x, &s, &m);
    const int lane = threadIdx.x;
    const double scale = scal[0];
    for (int idx = lane; idx < MAXM * LD; idx += kWave) A[idx] = 0.0;
    __syncthreads();
    if (lane < m) {
        double di = d_in[s + lane] * scale;
        if (lane == 0 && s > 0) di -= fabs(ew[s - 1]);
        if (lane == m - 1 && s + m < n) di -= fabs(ew[s + m - 1]);
        A[lane * LD + lane] = di;
        if (lane < m - 1) {
            const double el = ew[s + lane];
            A[lane * LD + lane + 1] = el;
            A[(lane + 1) * LD + lane] = el;
        }
        V0[lane] = (lane == 0) ? 1.0 : 0.0;
        V1[lane] = (lane == m - 1) ? 1.0 : 0.0;
    }
    __syncthreads();
    // an odd m gets a zero dummy row/column m; its pairs never rotate
    const int mm = m + (m & 1);
    const int npair = mm / 2;
    const int nblk = npair * npair;
    bool converged = (m <= 1);
    for (int sweep = 0; sweep < kJacobiSweeps && !converged; sweep++) {
        if (lane == 0) changed = 0;
        __syncthreads();
        for (int r = 0; r < mm - 1; r++) {
            if (lane < npair) {
                const int a = (lane == 0) ? 0 : ((lane - 1 + r) % (mm - 1)) + 1;
                const int bs = mm - 1 - lane;
                const int b = ((bs - 1 + r) % (mm - 1)) + 1;
                const int p = min(a, b), q = max(a, b);
                double c = 1.0, sn = 0.0;
                if (q < m) {
                    const double apq = A[p * LD + q];
                    const double app = A[p * LD + p], aqq = A[q * LD + q];
                    if (apq != 0.0 && fabs(apq) > kEps * sqrt(fabs(app)) * sqrt(fabs(aqq))) {
                        const double theta = (aqq - app) / (2.0 * apq);
                        double t;
                        if (fabs(theta) > 1e150)
                            t = 0.5 / theta;
                        else
                            t = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(t * t + 1.0);
                        sn = t * c;
                        changed = 1;
                    }
                }
                RC[lane] = c;
                RS[lane] = sn;
                RP[lane] = p;
                RQ[lane] = q;
            }
            __syncthreads();
            for (int blk = lane; blk < nblk; blk += kWave) {
                const int a = blk / npair, b = blk - a * npair;
                const double sa = RS[a], sb = RS[b];
                if (sa == 0.0 && sb == 0.0) continue;
                const double ca = RC[a], cb = RC[b];
                const int pa = RP[a], qa = RQ[a], pb = RP[b], qb = RQ[b];
                const double x00 = A[pa * LD + pb], x01 = A[pa * LD + qb];
                const double x10 = A[qa * LD + pb], x11 = A[qa * LD + qb];
                // rows (J_a^T), then columns (J_b)
                const double y00 = ca * x00 - sa * x10, y01 = ca * x01 - sa * x11;
                const double y10 = sa * x00 + ca * x10, y11 = sa * x01 + ca * x11;
                double z00 = cb * y00 - sb * y01, z01 = sb * y00 + cb * y01;
                double z10 = cb * y10 - sb * y11, z11 = sb * y10 + cb * y11;
                if (a == b) {  // the annihilated pair
                    z01 = 0.0;
                    z10 = 0.0;
                }
                A[pa * LD + pb] = z00;
                A[pa * LD + qb] = z01;
                A[qa * LD + pb] = z10;
                A[qa * LD + qb] = z11;
            }
            if (lane < npair && RS[lane] != 0.0) {  // first/last rows of V <- V J
                const double c = RC[lane], sn = RS[lane];
                const int p = RP[lane], q = RQ[lane];
                double x = V0[p], y = V0[q];
                V0[p] = c * x - sn * y;
                V0[q] = sn * x + c * y;
                x = V1[p];
                y = V1[q];
                V1[p] = c * x - sn * y;
                V1[q] = sn * x + c * y;
            }
            __syncthreads();
        }
        converged = (changed == 0);
        __syncthreads();
    }
    if (!converged && lane == 0) report_failure(info, n, s, m);
    // sort ascending, carrying the two rows
    __shared__ double L[MAXM];
    if (lane < m) L[lane] = A[lane * LD + lane];
    __syncthreads();
    if (lane < m) {
        const double x = L[lane];
        int rank = 0;
        for (int k = 0; k < m; k++) {
            const double y = L[k];
            rank += (y < x) || (y == x && k < lane);
        }
        if (single) {
            out_d[rank] = x * scal[1];
        } else {
            out_d[s + rank] = x;
            out_lo[s + rank] = V0[lane];
            out_hi[s + rank] = V1[lane];
        }
    }
    if (single && lane == 0 && info_user) *info_user = *info;
}

// ---------------------------------------------------------------------------
// Merge step 1 (the DLAMRG merge + z/QBR set-up of DLAED7_BR and DLAED8_BR,
// dlaed7_br.f:95-113, 240-264): each element finds its rank in the merged
// order by binary search in the sibling (merge path; ties take the left child
// first, as DLAMRG does) and writes z = [last rows of left; sign(rho) * first
// rows of right] / sqrt(2) with the parent's two boundary rows
// q1 = [blo_left, 0], q2 = [0, bhi_right].
__device__ __forceinline__ int count_less(const double* a, int len, double x) {
    int lo = 0, hi = len;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ int count_less_equal(const double* a, int len, double x) {
    int lo = 0, hi = len;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kThreads) void k_merge_prepare(
    int n, int depth, int bpm, const double* __restrict__ sd, const double* __restrict__ slo,
    const double* __restrict__ shi, const double* __restrict__ ew, double* __restrict__ md,
    double* __restrict__ mz, double* __restrict__ mq1, double* __restrict__ mq2) {
    const int merge = blockIdx.x / bpm, part = blockIdx.x % bpm;
    int s, sz;
    node_geom(n, depth, merge, &s, &sz);
    const int idx = part * kThreads + threadIdx.x;
    if (idx >= sz) return;
    const int cut = sz / 2;
    const double rho = ew[s + cut - 1];
    const double t0 = 1.0 / sqrt(2.0);
    const double x = sd[s + idx];
    int rank;
    double z, q1, q2;
    if (idx < cut) {
        rank = idx + count_less(sd + s + cut, sz - cut, x);
        z = shi[s + idx];
        q1 = slo[s + idx];
        q2 = 0.0;
    } else {
        rank = (idx - cut) + count_less_equal(sd + s, cut, x);
        z = slo[s + idx];
        if (rho < 0.0) z = -z;
        q1 = 0.0;
        q2 = shi[s + idx];
    }
    md[s + rank] = x;
    mz[s + rank] = z * t0;
    mq1[s + rank] = q1;
    mq2[s + rank] = q2;
}

// ---------------------------------------------------------------------------
// Merge step 2: deflation (DLAED8_BR, dlaed7_br.f:266-381).
//
// The reference walks the non-tiny poles once, keeping the last kept pole
// JLAM; a pair (JLAM, J) with |t*c*s| <= tol is rotated (JLAM deflated, J
// carries the rotated state forward).  Decisions therefore chain only through
// runs of consecutive rotations.  Here the walk is resolved in parallel and
// exactly:
//   A  tiny-z flags (rho*|z| <= tol) and compaction of the m non-tiny
//      positions P[0..m) (block prefix sum);
//   B  for every t whose pair (P[t-1], P[t]) rotates on the original values,
//      a speculative run from t (as if P[t-1] were unrotated) to its end
//      e(t), the first position whose test fails on the carried state;
//      real runs are then picked exactly as the sequential walk would:
//      t is real iff t > e of the previous real run.  Segments whose
//      candidates no earlier speculative run reaches resolve locally; any
//      other case falls back to one sequential resolution pass;
//   C  real runs are replayed to write the deflated values (sorted within
//      the run, as DLAED8_BR's insertion loop keeps them) and the carried
//      state of the run's surviving pole;
//   D  compaction into [kept poles | Givens-deflated | tiny-deflated].
struct GivensStep {
    double c, s, tau;
};

__device__ __forceinline__ bool givens_test(double dl, double zl, double dj, double zj, double tol,
                                            GivensStep* g) {
    // DLAPY2 then the |t*c*s| <= tol test (dlaed7_br.f:306-314)
    const double xa = fabs(zj), ya = fabs(zl);
    const double wmax = fmax(xa, ya), wmin = fmin(xa, ya);
    double tau;
    if (wmin == 0.0 || wmax > 1.79e308) {
        tau = wmax;
    } else {
        const double q = wmin / wmax;
        tau = wmax * sqrt(1.0 + q * q);
    }
    const double t = dj - dl;
    const double c = zj / tau;
    const double sn = -zl / tau;
    g->c = c;
    g->s = sn;
    g->tau = tau;
    return fabs(t * c * sn) <= tol;
}

template <int T>
__global__ __launch_bounds__(T) void k_deflate(
    int n, int depth, const double* __restrict__ ew, double* __restrict__ md,
    double* __restrict__ mz, double* __restrict__ mq1, double* __restrict__ mq2, int* __restrict__ P,
    int* __restrict__ rend, int* __restrict__ fl, double* __restrict__ ad, double* __restrict__ az,
    double2* __restrict__ adz, double2* __restrict__ aq, double* __restrict__ gd,
    double* __restrict__ gq1, double* __restrict__ gq2, double* __restrict__ td,
    double* __restrict__ tq1, double* __restrict__ tq2, int* __restrict__ counts) {
    __shared__ int lds[T / kWave + 2];
    __shared__ int any_fix;
    const int merge = blockIdx.x;
    int s, sz;
    node_geom(n, depth, merge, &s, &sz);
    const int cut = sz / 2;
    const double rho = fabs(2.0 * ew[s + cut - 1]);
    const double tol = 8.0 * kEps * fmax(fabs(md[s]), fabs(md[s + sz - 1]));
    const int tid = threadIdx.x;

    // ---- A: tiny-z flags, non-tiny compaction
    const int seg = (sz + T - 1) / T;
    const int p0 = min(sz, tid * seg), p1 = min(sz, p0 + seg);
    int cnt = 0;
#pragma unroll 8
    for (int p = p0; p < p1; p++) {
        const bool keep = rho * fabs(mz[s + p]) > tol;
        cnt += keep;
        fl[s + p] = keep ? 0 : 2;
    }
    int m;
    int off = block_exclusive_sum<T>(cnt, &m, lds);
    for (int p = p0; p < p1; p++)
        if (fl[s + p] == 0) P[s + off++] = p;
    if (tid == 0) any_fix = 0;
    __syncthreads();

    // ---- B: speculative runs from every rotating pair
    const int tseg = (m + T - 1) / T;
    const int t1 = min(m, (tid + 1) * tseg);
    const int t0 = max(1, min(m, tid * tseg));
    int first_cand = INT_MAX, max_end = INT_MIN;
    bool unknown = false;
    for (int t = t0; t < t1; t++) {
        const int jl = P[s + t - 1], j = P[s + t];
        GivensStep g;
        const double dj = md[s + j];
        if (givens_test(md[s + jl], mz[s + jl], dj, mz[s + j], tol, &g)) {
            double cd = md[s + jl] * g.s * g.s + dj * g.c * g.c;
            double cz = g.tau;
            int u = t + 1, steps = 0;
            while (u < m && steps < kSpecCap) {
                const int jj = P[s + u];
                const double d2 = md[s + jj];
                if (!givens_test(cd, cz, d2, mz[s + jj], tol, &g)) break;
                cd = cd * g.s * g.s + d2 * g.c * g.c;
                cz = g.tau;
                u++;
                steps++;
            }
            const int e = (u < m && steps == kSpecCap) ? -2 : u;
            rend[s + t] = e;
            if (first_cand == INT_MAX) first_cand = t;
            if (e == -2) unknown = true;
            else max_end = max(max_end, e);
        } else {
            rend[s + t] = -1;
        }
    }
    const int carry = block_exclusive_max<T>(unknown ? INT_MAX : max_end, lds);
    if (unknown || (first_cand != INT_MAX && carry >= first_cand)) any_fix = 1;
    __syncthreads();
    if (any_fix) {
        // exact sequential resolution (rare: overlapping or very long runs)
        if (tid == 0) {
            int cover = -1;
            for (int t = 1; t < m; t++) {
                int e = rend[s + t];
                if (e == -1) continue;
                if (t <= cover) {
                    rend[s + t] = -1;
                    continue;
                }
                if (e == -2) {  // finish the capped walk
                    const int jl = P[s + t - 1], j = P[s + t];
                    GivensStep g;
                    givens_test(md[s + jl], mz[s + jl], md[s + j], mz[s + j], tol, &g);
                    double cd = md[s + jl] * g.s * g.s + md[s + j] * g.c * g.c;
                    double cz = g.tau;
                    int u = t + 1;
                    while (u < m) {
                        const int jj = P[s + u];
                        const double d2 = md[s + jj];
                        if (!givens_test(cd, cz, d2, mz[s + jj], tol, &g)) break;
                        cd = cd * g.s * g.s + d2 * g.c * g.c;
                        cz = g.tau;
                        u++;
                    }
                    e = u;
                    rend[s + t] = e;
                }
                cover = e;
            }
        }
    } else {
        int cover = -1;
        for (int t = t0; t < t1; t++) {
            const int e = rend[s + t];
            if (e == -1) continue;
            if (t <= cover) rend[s + t] = -1;
            else cover = e;
        }
    }
    __syncthreads();

    // ---- C: replay real runs, write deflated values and carried states
    int my_g = 0;
    for (int t = t0; t < t1; t++) {
        const int e = rend[s + t];
        if (e >= 0) my_g += e - t;
    }
    int G;
    int goff = block_exclusive_sum<T>(my_g, &G, lds);
    for (int t = t0; t < t1; t++) {
        const int e = rend[s + t];
        if (e < 0) continue;
        int cur = P[s + t - 1];
        double cd = md[s + cur], cz = mz[s + cur], c1 = mq1[s + cur], c2 = mq2[s + cur];
        // every value the run produces is a convex combination of the run's
        // original poles; clamping to their range removes last-ulp excursions
        // so the three output lists stay exactly sorted
        const double lo = cd;
        double hi = cd;
        const int base = goff;
        for (int u = t; u < e; u++) {
            const int j = P[s + u];
            const double dj = md[s + j], qj1 = mq1[s + j], qj2 = mq2[s + j];
            hi = dj;
            GivensStep g;
            givens_test(cd, cz, dj, mz[s + j], tol, &g);
            // DROT on the two selected rows (dlaed7_br.f:323-325) and the
            // D update (dlaed7_br.f:326-328)
            gd[s + goff] = fmin(fmax(cd * g.c * g.c + dj * g.s * g.s, lo), dj);
            gq1[s + goff] = g.c * c1 + g.s * qj1;
            gq2[s + goff] = g.c * c2 + g.s * qj2;
            goff++;
            fl[s + cur] = 1;
            const double nd = fmin(fmax(cd * g.s * g.s + dj * g.c * g.c, lo), dj);
            const double n1 = g.c * qj1 - g.s * c1;
            const double n2 = g.c * qj2 - g.s * c2;
            cd = nd;
            cz = g.tau;
            c1 = n1;
            c2 = n2;
            cur = j;
        }
        md[s + cur] = fmin(fmax(cd, lo), hi);
        mz[s + cur] = cz;
        mq1[s + cur] = c1;
        mq2[s + cur] = c2;
        // keep this run's deflated values ascending
        for (int a = base + 1; a < goff; a++) {
            const double x = gd[s + a], x1 = gq1[s + a], x2 = gq2[s + a];
            int b = a - 1;
            while (b >= base && gd[s + b] > x) {
                gd[s + b + 1] = gd[s + b];
                gq1[s + b + 1] = gq1[s + b];
                gq2[s + b + 1] = gq2[s + b];
                b--;
            }
            gd[s + b + 1] = x;
            gq1[s + b + 1] = x1;
            gq2[s + b + 1] = x2;
        }
    }
    __syncthreads();

    // ---- D: compaction of kept poles and tiny-deflated entries
    int na = 0, nt = 0;
#pragma unroll 8
    for (int p = p0; p < p1; p++) {
        const int f = fl[s + p];
        na += (f == 0);
        nt += (f == 2);
    }
    int K, Tn;
    int aoff = block_exclusive_sum<T>(na, &K, lds);
    int toff = block_exclusive_sum<T>(nt, &Tn, lds);
    for (int p = p0; p < p1; p++) {
        const int f = fl[s + p];
        if (f == 0) {
            const double dv = md[s + p], zv = mz[s + p];
            ad[s + aoff] = dv;
            az[s + aoff] = zv;
            adz[s + aoff] = make_double2(dv, zv * zv);
            aq[s + aoff] = make_double2(mq1[s + p], mq2[s + p]);
            aoff++;
        } else if (f == 2) {
            td[s + toff] = md[s + p];
            tq1[s + toff] = mq1[s + p];
            tq2[s + toff] = mq2[s + p];
            toff++;
        }
    }
    if (tid == 0) {
        counts[4 * merge + 0] = K;
        counts[4 * merge + 1] = G;
        counts[4 * merge + 2] = Tn;
        counts[4 * merge + 3] = 0;
    }
}

// ---------------------------------------------------------------------------
// Merge steps 3-5 share one work layout: L lanes per item (root or pole),
// L a power of two chosen from the merge's K at run time so the level's grid
// is filled and each lane streams about K/L poles or roots.  Lane c of an
// item takes the strided indices c, c+L, ... (coalesced across the L lanes);
// partial sums meet in an xor butterfly inside the item's lane group.
__device__ __forceinline__ int lanes_per_item(int K, int capacity) {
    int L = kWave;
    while (L > 1 && (K * L > capacity || K < 16 * L)) L >>= 1;
    return L;
}

template <class T>
__device__ __forceinline__ T group_sum(T v, int L) {
    for (int o = 1; o < L; o <<= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Merge step 3: secular roots (DLAED9_BR's DLAED4_BR loop, dlaed9_br.f:181-262);
// per root the compact state (org, tau, near1, near2) plus d[org].
struct LaneEval {
    const double2* __restrict__ dz;  // (d_j, z_j^2)
    int K, L, c;
    __device__ __forceinline__ double d(int j) const { return dz[j].x; }
    __device__ __forceinline__ double zsq(int j) const { return dz[j].y; }
    __device__ SecSums sums(double dorg, double tau, int ex_lo, int nex) const {
        double S = 0.0, A = 0.0, DP = 0.0, DF = 0.0, J = 0.0;
        double jd = (double)c;
        const double dL = (double)L;
#pragma unroll 2
        for (int j = c; j < K; j += L) {
            const double2 p = dz[j];
            const double delta = (p.x - dorg) - tau;
            const double r = rcp_nr(delta);
            double term = p.y * r;
            if ((unsigned)(j - ex_lo) < (unsigned)nex) term = 0.0;
            S += term;
            A += fabs(term);
            const double x = term * fabs(r);
            DP += fmin(x, 0.0);
            DF += fmax(x, 0.0);
            J = fma(jd, term, J);
            jd += dL;
        }
        SecSums out;
        out.S = group_sum(S, L);
        out.A = group_sum(A, L);
        out.dpsi = -group_sum(DP, L);
        out.dphi = group_sum(DF, L);
        out.J = group_sum(J, L);
        return out;
    }
};

__global__ __launch_bounds__(kThreads) void k_secular(
    int n, int depth, int bpm, const double* __restrict__ ew, const int* counts,
    const double* __restrict__ ad, const double* __restrict__ az,
    const double2* __restrict__ adz, double* __restrict__ rlam, double* __restrict__ rtau,
    double* __restrict__ rn1, double* __restrict__ rn2, double* __restrict__ rdorg,
    int* __restrict__ rorg, int* counts_rw, int* info) {
    const int merge = blockIdx.x / bpm, part = blockIdx.x % bpm;
    const int K = counts[4 * merge];
    const int L = lanes_per_item(K, bpm * kThreads);
    const int per_block = kThreads / L;
    if (part * per_block >= K) return;
    int s, sz;
    node_geom(n, depth, merge, &s, &sz);
    const int cut = sz / 2;
    const double rho = fabs(2.0 * ew[s + cut - 1]);
    const int c = threadIdx.x & (L - 1);
    const int j = part * per_block + threadIdx.x / L;
    int iters = 0;
    if (j < K) {
        if (K == 1) {
            const double z0 = az[s];
            rlam[s] = ad[s] + rho * z0 * z0;
            rtau[s] = rho * z0 * z0;
            rn1[s] = 1.0;  // secular eigenvector (1)
            rn2[s] = 0.0;
            rorg[s] = 0;
            rdorg[s] = ad[s];
            iters = 1;
        } else if (K == 2) {
            double lam, v0, v1, tau;
            secular_two(j, ad + s, az + s, rho, &lam, &v0, &v1, &tau);
            rlam[s + j] = lam;
            rtau[s + j] = tau;
            rn1[s + j] = v0;  // secular eigenvector (v0, v1)
            rn2[s + j] = v1;
            rorg[s + j] = 0;
            rdorg[s + j] = ad[s];
            iters = 1;
        } else {
            const LaneEval E{adz + s, K, L, c};
            const RootState r = secular_root(E, K, j, rho);
            if (c == 0) {
                rlam[s + j] = r.lambda;
                rtau[s + j] = r.tau;
                rn1[s + j] = r.near1;
                rn2[s + j] = r.near2;
                rorg[s + j] = r.org;
                rdorg[s + j] = ad[s + r.org];
                iters = r.iters;
                if (r.info) report_failure(info, n, s, sz);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) iters += __shfl_down(iters, o, kWave);
    if ((threadIdx.x & 63) == 0 && iters) atomicAdd(&counts_rw[4 * merge + 3], iters);
}

// Merge step 4: product-form update vector (dlaed9_br.f:218-266): L lanes per
// pole i streaming the roots j: w_i = sign(z_i) sqrt(-prod_j delta_j(i) /
// prod_{j != i} (d_i - d_j)).  delta_j(i) comes from root j's compact state;
// the two pole-adjacent entries are root j's near1/near2 (dlaed9_br.f:272-331).
__global__ __launch_bounds__(kThreads) void k_wprod(
    int n, int depth, int bpm, const int* __restrict__ counts, const double* __restrict__ ad,
    const double* __restrict__ az, const double* __restrict__ rtau,
    const double* __restrict__ rn1, const double* __restrict__ rn2,
    const double* __restrict__ rdorg, double* __restrict__ w) {
    const int merge = blockIdx.x / bpm, part = blockIdx.x % bpm;
    const int K = counts[4 * merge];
    if (K <= 2) return;
    const int L = lanes_per_item(K, bpm * kThreads);
    const int per_block = kThreads / L;
    if (part * per_block >= K) return;
    int s, sz;
    node_geom(n, depth, merge, &s, &sz);
    const int c = threadIdx.x & (L - 1);
    const int i = part * per_block + threadIdx.x / L;
    if (i >= K) return;  // whole lane groups leave together
    const double* __restrict__ d = ad + s;
    const double di = d[i];
    double prod = 1.0;
    for (int j = c; j < K; j += L) {
        double dl = (di - rdorg[s + j]) - rtau[s + j];
        const int a = (j == K - 1) ? K - 2 : j;
        if ((unsigned)(i - a) <= 1u) dl = (i == a) ? rn1[s + j] : rn2[s + j];
        prod *= (j == i) ? dl : dl * rcp_nr(di - d[j]);
    }
    for (int o = 1; o < L; o <<= 1) prod *= __shfl_xor(prod, o, kWave);
    if (c == 0) w[s + i] = copysign(sqrt(-prod), az[s + i]);
}

// Merge step 5: the two boundary rows of the merged eigenvector block
// (dlaed9_br.f:272-331): L lanes per root j streaming the poles i,
// x_i = w_i / delta_j(i), row_r(j) = sum_i q_r(i) x_i / ||x||.
__global__ __launch_bounds__(kThreads) void k_rows(
    int n, int depth, int bpm, const int* __restrict__ counts, const double* __restrict__ ad,
    const double2* __restrict__ aq, const double* __restrict__ w,
    const double* __restrict__ rtau, const double* __restrict__ rn1,
    const double* __restrict__ rn2, const double* __restrict__ rdorg,
    double* __restrict__ qo1, double* __restrict__ qo2) {
    const int merge = blockIdx.x / bpm, part = blockIdx.x % bpm;
    const int K = counts[4 * merge];
    const int L = lanes_per_item(K, bpm * kThreads);
    const int per_block = kThreads / L;
    if (part * per_block >= K) return;
    int s, sz;
    node_geom(n, depth, merge, &s, &sz);
    const int c = threadIdx.x & (L - 1);
    const int j = part * per_block + threadIdx.x / L;
    if (j >= K) return;
    if (K <= 2) {  // eigenvector returned directly (dlaed9_br.f:146-172); L == 1
        const double v0 = rn1[s + j], v1 = rn2[s + j];
        const double2 q0 = aq[s];
        double a1 = q0.x * v0, a2 = q0.y * v0;
        if (K == 2) {
            const double2 q1 = aq[s + 1];
            a1 += q1.x * v1;
            a2 += q1.y * v1;
        }
        qo1[s + j] = a1;
        qo2[s + j] = a2;
        return;
    }
    const double tau = rtau[s + j], nr1 = rn1[s + j], nr2 = rn2[s + j];
    const double dorg = rdorg[s + j];
    const int a = (j == K - 1) ? K - 2 : j;
    const double* __restrict__ d = ad + s;
    const double* __restrict__ wv = w + s;
    const double2* __restrict__ q = aq + s;
    double temp = 0.0, acc1 = 0.0, acc2 = 0.0;
#pragma unroll 2
    for (int i = c; i < K; i += L) {
        double dl = (d[i] - dorg) - tau;
        if ((unsigned)(i - a) <= 1u) dl = (i == a) ? nr1 : nr2;
        const double x = wv[i] * rcp_nr(dl);
        const double2 qi = q[i];
        temp = fma(x, x, temp);
        acc1 = fma(qi.x, x, acc1);
        acc2 = fma(qi.y, x, acc2);
    }
    temp = group_sum(temp, L);
    acc1 = group_sum(acc1, L);
    acc2 = group_sum(acc2, L);
    if (c == 0) {
        const double nrm = sqrt(temp);
        qo1[s + j] = acc1 / nrm;
        qo2[s + j] = acc2 / nrm;
    }
}

// Merge step 6: the node's ascending order (DLAMRG of roots with deflated
// values, dlaed7_br.f:146, and the boundary-row persist, 155-160) as a 3-way
// merge of three sorted lists: roots, Givens-deflated, tiny-deflated.
__global__ __launch_bounds__(kThreads) void k_finalize(
    int n, int depth, int bpm, const int* __restrict__ counts, const double* __restrict__ rlam,
    const double* __restrict__ qo1, const double* __restrict__ qo2, const double* __restrict__ gd,
    const double* __restrict__ gq1, const double* __restrict__ gq2, const double* __restrict__ td,
    const double* __restrict__ tq1, const double* __restrict__ tq2, double* __restrict__ od,
    double* __restrict__ olo, double* __restrict__ ohi, int root, const double* __restrict__ scal,
    const int* __restrict__ info, int* info_user) {
    const int merge = blockIdx.x / bpm, part = blockIdx.x % bpm;
    int s, sz;
    node_geom(n, depth, merge, &s, &sz);
    if (root && blockIdx.x == 0 && threadIdx.x == 0 && info_user) *info_user = *info;
    const int idx = part * kThreads + threadIdx.x;
    if (idx >= sz) return;
    const int K = counts[4 * merge], G = counts[4 * merge + 1], T = counts[4 * merge + 2];
    double x, q1, q2;
    int rank;
    if (idx < K) {
        x = rlam[s + idx];
        rank = idx + count_less(gd + s, G, x) + count_less(td + s, T, x);
        q1 = root ? 0.0 : qo1[s + idx];
        q2 = root ? 0.0 : qo2[s + idx];
    } else if (idx < K + G) {
        const int g = idx - K;
        x = gd[s + g];
        rank = g + count_less_equal(rlam + s, K, x) + count_less(td + s, T, x);
        q1 = gq1[s + g];
        q2 = gq2[s + g];
    } else {
        const int t = idx - K - G;
        x = td[s + t];
        rank = t + count_less_equal(rlam + s, K, x) + count_less_equal(gd + s, G, x);
        q1 = tq1[s + t];
        q2 = tq2[s + t];
    }
    if (root) {
        od[s + rank] = x * scal[1];
    } else {
        od[s + rank] = x;
        olo[s + rank] = q1;
        ohi[s + rank] = q2;
    }
}

}  // namespace brdc
