// brdc_subtree.h -- the bottom of the merge tree, one block per subtree.
//
// The deepest merge levels of a tree hold thousands of merges of a few dozen
// rows each.  As separate launches, each of their six merge steps costs a
// kernel boundary (L2 written back, every operand fetched again) for a few
// microseconds of work.  Here one block takes a whole subtree of at most
// kSubRows rows -- its leaves and every merge below its root, the root
// included -- and keeps the node state and all merge scratch in LDS: the
// leaves' QL iterations, then per level the merge-path prepare, the
// deflation (the same block routine as k_deflate, one wave per merge), the
// secular roots packed two lanes each across the level's merges, the product
// form, the boundary rows and the three-way finalize.  Only the subtree
// root's state (or, when the subtree is a whole matrix, its eigenvalues)
// goes back to HBM, with each merge's count record.
//
// The arithmetic is that of the level kernels (dlaed0_br.f:95-210 through
// dlaed9_br.f:181-331, as cited there) with two lanes per root and one per
// pole, so the results agree with the level-by-level path to rounding.
#pragma once

#include "brdc_kernels.h"

namespace brdc {

// Wavefronts of a subtree block and the largest subtree: two lanes per
// secular root, so 32 rows per wave.
constexpr int kSubWaves = 4;
constexpr int kSubRows = 128;
static_assert(kSubRows == 32 * kSubWaves, "two lanes per root of the subtree's largest merge");
constexpr int kSubThreads = kSubWaves * kWave;
constexpr int kSubMaxLevels = 6;     // merge levels of a subtree (leaves of >= 2 rows)
constexpr int kSubMerges = 1 << (kSubMaxLevels - 1);  // merges of one level
constexpr int kSubLeafCol = 2 * kSubRows;  // leaf column doubles: (largest leaf) x leaves

// Subtree roots of one launch.  A single matrix: the nodes at depth g.depth,
// with `levels` the tree's leaf depth; merge (r, i) of subtree b (relative
// depth r, index i) has record (2^(g.depth + r) - 1) + b 2^r + i.  A batch:
// plan entries whose flags carry the merge levels inside (bits 8..15) and,
// per subtree, kSubMaxLevels record bases (srec), merge (r, i) at
// srec[b * kSubMaxLevels + r] + i.
constexpr int kSubLevelShift = 8;

struct SubGeo {
    Geo g;
    int levels;
    const int* srec;
    __device__ __forceinline__ int height(const Item& it) const {
        return g.plan ? ((it.flags >> kSubLevelShift) & 0xFF) : levels - g.depth;
    }
    __device__ __forceinline__ int record(int b, int r, int i) const {
        if (g.plan) return srec[b * kSubMaxLevels + r] + i;
        return ((1 << (g.depth + r)) - 1) + (b << r) + i;
    }
};

struct SubtreeLds {
    double ew[kSubRows];
    double sd[kSubRows], slo[kSubRows], shi[kSubRows];
    double md[kSubRows], mz[kSubRows], mq1[kSubRows], mq2[kSubRows];
    double ad[kSubRows], az[kSubRows];
    double2 adz[kSubRows], aq[kSubRows];
    double gd[kSubRows], gq1[kSubRows], gq2[kSubRows];
    double td[kSubRows], tq1[kSubRows], tq2[kSubRows];
    double rlam[kSubRows], rn1[kSubRows], rn2[kSubRows];
    double2 rs[kSubRows], pw[kSubRows];
    double qo1[kSubRows], qo2[kSubRows];
    double leaf[4 * kSubLeafCol];
    int P[kSubRows], rend[kSubRows], fl[kSubRows];
    // per merge of the current level
    int mst[kSubMerges], msz[kSubMerges], K[kSubMerges], G[kSubMerges], T[kSubMerges];
    int roff[kSubMerges + 1];
    int ev[kSubMerges], evmax[kSubMerges], nfail[kSubMerges];
};

// The merge of relative depth r holding element e of a subtree of S rows:
// its index, first row and size (the halving of node_geom, from the top).
__device__ __forceinline__ int sub_locate(int S, int r, int e, int* ms, int* msz) {
    int s = 0, sz = S, idx = 0;
    for (int b = 0; b < r; b++) {
        const int left = sz / 2;
        idx <<= 1;
        if (e - s >= left) {
            s += left;
            sz -= left;
            idx |= 1;
        } else {
            sz = left;
        }
    }
    *ms = s;
    *msz = sz;
    return idx;
}

// Boundary rows of root j of a merge of K <= 2 poles at LDS offset ms: the
// secular eigenvector (near1, near2) applied directly (dlaed9_br.f:146-172).
__device__ __forceinline__ void sub_rows_direct(const SubtreeLds& L, int ms, int K, int j,
                                                double* q1o, double* q2o) {
    const double v0 = L.rn1[ms + j], v1 = L.rn2[ms + j];
    const double2 q0 = L.aq[ms];
    double a1 = q0.x * v0, a2 = q0.y * v0;
    if (K == 2) {
        const double2 q1 = L.aq[ms + 1];
        a1 += q1.x * v1;
        a2 += q1.y * v1;
    }
    *q1o = a1;
    *q2o = a2;
}

// The secular roots of a level (at most kSubRows) go two lanes each, dealt
// round the block's waves, and small merges deflate one per wave; a SIMD
// holds two such waves, which hide part of each other's dependent FP64
// chains.

// One block per subtree root (blockIdx.x).  out_final: the eigenvalues when
// the subtree is its whole matrix; od/olo/ohi: the subtree root's node state
// otherwise.
__global__ __launch_bounds__(kSubThreads) void k_subtree(
    SubGeo sg, const int* __restrict__ moff, const double* __restrict__ d_in,
    const double* __restrict__ ew, const double* __restrict__ scal, double* __restrict__ out_final,
    double* __restrict__ od, double* __restrict__ olo, double* __restrict__ ohi,
    int* __restrict__ counts, int* info, int* info_user) {
    __shared__ SubtreeLds L;
    __shared__ int dsh[kSubWaves][deflate_sh_ints<kWave>()];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & (kWave - 1);
    const int b = blockIdx.x;
    const Item it = sg.g.item(b);
    const int s0 = it.s, S = it.sz, H = sg.height(it);
    const bool mroot = (it.flags & kItemRoot) != 0;  // the subtree is its whole matrix
    if (tid < S) L.ew[tid] = ew[s0 + tid];

    // ---- leaves (relative depth H), one lane each
    {
        const int NL = 1 << H;
        const int maxm = max_node_size(S, H);
        if (tid < NL) {
            int ls, m;
            node_geom(S, H, tid, &ls, &m);
            double* const vd = L.leaf + tid;
            double* const ve = vd + maxm * NL;
            double* const v0 = ve + maxm * NL;
            double* const v1 = v0 + maxm * NL;
            if (!leaf_ql(s0 + ls, m, sg.g.n, scal[2 * it.mat], d_in, ew, NL, vd, ve, v0, v1))
                report_item_failure(fail_site(info, moff, sg.g.n, Item{s0 + ls, m, it.mat, 0}), m);
            leaf_emit(m, NL, vd, v0, v1, [&](int rank, double x, double lo, double hi) {
                L.sd[ls + rank] = x;
                L.slo[ls + rank] = lo;
                L.shi[ls + rank] = hi;
            });
        }
    }
    __syncthreads();

    for (int r = H - 1; r >= 0; r--) {
        const int nm = 1 << r;
        const bool top = (r == 0);
        const bool keep_rows = !(top && mroot);  // a matrix's root merge keeps no rows
        // ---- prepare: one element per thread
        if (tid < S) {
            int ms, msz;
            sub_locate(S, r, tid, &ms, &msz);
            prepare_element(ms, msz, tid - ms, L.ew[ms + msz / 2 - 1], L.sd, L.slo, L.shi, L.md,
                            L.mz, L.mq1, L.mq2);
        }
        if (tid < nm) {
            int ms, msz;
            node_geom(S, r, tid, &ms, &msz);
            L.mst[tid] = ms;
            L.msz[tid] = msz;
            L.ev[tid] = 0;
            L.evmax[tid] = 0;
            L.nfail[tid] = 0;
        }
        __syncthreads();
        // ---- deflation: merges dealt to the waves, each merge one wave
        for (int m = wave; m < nm; m += kSubWaves) {
            const int ms = L.mst[m], msz = L.msz[m];
            const DeflateCounts c = deflate_merge<kWave>(
                ms, msz, L.ew[ms + msz / 2 - 1], L.md, L.mz, L.mq1, L.mq2, L.P, L.rend, L.fl, L.ad,
                L.az, L.adz, L.aq, L.gd, L.gq1, L.gq2, L.td, L.tq1, L.tq2,
                counts + kCountStride * sg.record(b, r, m), dsh[wave]);
            if (lane == 0) {
                L.K[m] = c.K;
                L.G[m] = c.G;
                L.T[m] = c.T;
            }
        }
        __syncthreads();
        // roots (and poles) of the level numbered across its merges
        if (wave == 0) {
            const int k = (lane < nm) ? L.K[lane] : 0;
            const int incl = wave_incl_sum(k);
            if (lane < nm) L.roff[lane] = incl - k;
            if (lane == kWave - 1) L.roff[nm] = incl;
        }
        __syncthreads();
        const int total = L.roff[nm];
        // ---- secular roots: two lanes each, consecutive roots in consecutive waves
        {
            constexpr int SL = 2;
            const int j = (lane >> 1) * kSubWaves + wave, c = lane & 1;
            int m = 0;
            while (m + 1 < nm && L.roff[m + 1] <= j) m++;
            const bool has = j < total;
            const int ms = L.mst[m], K = L.K[m], i = j - L.roff[m];
            const double rho = fabs(2.0 * L.ew[ms + L.msz[m] / 2 - 1]);
            int iters = 0;
            bool bad = false;
            if (has && K <= 2 && c == 0) {
                if (K == 1) {
                    const double z0 = L.az[ms];
                    L.rlam[ms] = L.ad[ms] + rho * z0 * z0;
                    L.rs[ms] = make_double2(L.ad[ms], rho * z0 * z0);
                    L.rn1[ms] = 1.0;  // secular eigenvector (1)
                    L.rn2[ms] = 0.0;
                } else {
                    double lam, v0, v1, tau;
                    secular_two(i, L.ad + ms, L.az + ms, rho, &lam, &v0, &v1, &tau);
                    L.rlam[ms + i] = lam;
                    L.rs[ms + i] = make_double2(L.ad[ms], tau);
                    L.rn1[ms + i] = v0;  // secular eigenvector (v0, v1)
                    L.rn2[ms + i] = v1;
                }
                iters = (i == 0) ? 1 : 0;
            }
            bool active = has && K >= 3;
            const PoleView E{L.adz + ms};
            SecularSolve<PoleView> st;
            if (active) st.init(E, K, i, rho);
            // a root's two lanes share its state; lanes without a root sum nothing
            for (;;) {
                if (!__any(active)) break;
                const int exlo = active ? st.ev_exlo : 0;
                const int j2 = active ? st.ev_exlo + st.ev_nex : 0;
                const int kend = active ? K : 0;
                const double org = active ? st.ev_org : 0.0, tau = active ? st.ev_tau : 0.0;
                double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                secular_side(E.dz, c, exlo, SL, org, tau, acc[0], acc[1], acc[2]);
                secular_side(E.dz, j2 + ((c - j2) & (SL - 1)), kend, SL, org, tau, acc[3], acc[4],
                             acc[5]);
                group_sum6(acc, SL);
                SecSums sums;
                sums.S = acc[0] + acc[3];
                sums.A = acc[3] - acc[0];
                sums.dpsi = acc[1];
                sums.dphi = acc[4];
                sums.J = acc[2] + acc[5];
                if (active && st.advance(E, sums)) {
                    active = false;
                    const RootState& rr = st.out;
                    if (c == 0) {
                        L.rlam[ms + i] = rr.lambda;
                        L.rs[ms + i] = make_double2(E.d(rr.org), rr.tau);
                        L.rn1[ms + i] = rr.near1;
                        L.rn2[ms + i] = rr.near2;
                        iters = rr.iters;
                        bad = rr.info != 0;
                    }
                }
            }
            if (bad) report_item_failure(fail_site(info, moff, sg.g.n, Item{s0 + ms, L.msz[m], it.mat, 0}),
                                         L.msz[m]);
            if (iters) {
                atomicAdd(&L.ev[m], iters);
                atomicMax(&L.evmax[m], iters);
            }
            if (bad) atomicAdd(&L.nfail[m], 1);
        }
        __syncthreads();
        if (tid < nm) {
            int* rec = counts + kCountStride * sg.record(b, r, tid);
            rec[3] = L.ev[tid];
            rec[5] = L.evmax[tid];
            rec[6] = L.nfail[tid];
        }
        if (keep_rows) {
            // Product form and boundary rows: Lp lanes per pole (root), the
            // most (up to 16) that leave every item of the level a lane
            // group; lane c takes its share [K c / Lp, K (c+1) / Lp) of the
            // merge's roots (poles), and the shares meet in a butterfly (as
            // the level kernels' small merges)
            int Lp = 1;
            while (Lp < 16 && 2 * Lp * total <= kSubThreads) Lp <<= 1;
            const int c = tid & (Lp - 1), per = kSubThreads / Lp;
            // the same trip count in every lane: items past the last are
            // computed on the last item and not stored
            auto item = [&](int base, int& ms, int& K, int& i) {
                const int j = min(base + tid / Lp, total - 1);
                int m = 0;
                while (m + 1 < nm && L.roff[m + 1] <= j) m++;
                ms = L.mst[m];
                K = L.K[m];
                i = j - L.roff[m];
            };
            // ---- product form
            for (int base = 0; base < total; base += per) {
                int ms, K, i;
                item(base, ms, K, i);
                const int ra = (K * c) / Lp, rb = (K * (c + 1)) / Lp;
                double share = -1.0;
                if (K >= 3)
                    share = wprod_share(RootsLds{L.rs + ms, L.ad + ms}, L.rn1 + ms, L.rn2 + ms, i,
                                        L.ad[ms + i], i & ~(kGroup - 1), ra, rb, K);
                for (int o = 1; o < Lp; o <<= 1) share *= __shfl_xor(share, o);
                if (c == 0 && base + tid / Lp < total && K >= 3)
                    L.pw[ms + i] = make_double2(L.ad[ms + i], copysign(sqrt(-share), L.az[ms + i]));
            }
            __syncthreads();
            // ---- boundary rows
            for (int base = 0; base < total; base += per) {
                int ms, K, j;
                item(base, ms, K, j);
                const int ia = (K * c) / Lp, ib = (K * (c + 1)) / Lp;
                double temp = 0.0, acc1 = 0.0, acc2 = 0.0;
                if (K >= 3)
                    rows_share(PolesLds{L.pw + ms, L.aq + ms}, j, L.rs[ms + j], L.rn1[ms + j],
                               L.rn2[ms + j], j & ~(kGroup - 1), ia, ib, K, temp, acc1, acc2);
                for (int o = 1; o < Lp; o <<= 1) {
                    temp += __shfl_xor(temp, o);
                    acc1 += __shfl_xor(acc1, o);
                    acc2 += __shfl_xor(acc2, o);
                }
                if (c == 0 && base + tid / Lp < total) {
                    double q1, q2;
                    if (K >= 3) {
                        const double nrm = sqrt(temp);
                        q1 = acc1 / nrm;
                        q2 = acc2 / nrm;
                    } else {
                        sub_rows_direct(L, ms, K, j, &q1, &q2);
                    }
                    L.qo1[ms + j] = q1;
                    L.qo2[ms + j] = q2;
                }
            }
            __syncthreads();
        }
        // ---- finalize: one element per thread, into LDS, or to HBM at the top
        if (tid < S) {
            int fs, fsz;
            const int fm = sub_locate(S, r, tid, &fs, &fsz);
            const int idx = tid - fs;
            if (!top) {
                finalize_element(fs, idx, L.K[fm], L.G[fm], L.T[fm], false, 1.0, L.rlam,
                                 L.qo1 + fs, L.qo2 + fs, L.gd, L.gq1, L.gq2, L.td, L.tq1, L.tq2,
                                 nullptr, L.sd, L.slo, L.shi);
            } else if (mroot) {  // fs == 0: the outputs start at the subtree's first row
                finalize_element(0, idx, L.K[0], L.G[0], L.T[0], true, scal[2 * it.mat + 1],
                                 L.rlam, L.qo1, L.qo2, L.gd, L.gq1, L.gq2, L.td, L.tq1, L.tq2,
                                 out_final + s0, nullptr, nullptr, nullptr);
            } else {
                finalize_element(0, idx, L.K[0], L.G[0], L.T[0], false, 1.0, L.rlam, L.qo1, L.qo2,
                                 L.gd, L.gq1, L.gq2, L.td, L.tq1, L.tq2, nullptr, od + s0,
                                 olo + s0, ohi + s0);
            }
        }
        __syncthreads();
    }
    // a single matrix solved whole here: hand INFO to the caller
    if (mroot && info_user && tid == 0)
        *info_user = __hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace brdc
