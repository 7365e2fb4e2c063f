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
//   counts: kCountStride ints per merge (K, G, T, secular evaluations,
//           roots claimed by the streamed secular kernel, the most
//           evaluations of one root, roots that did not converge; one spare)
// Every kernel derives its merge geometry from (n, depth, merge index), so a
// level is a fixed launch sequence with no host round trip.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>

#include "brdc_launch.h"
#include "brdc_secular.h"
#include "brdc_tree.h"

namespace brdc {

constexpr int kMaxLeaf = 64;
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
// Geometry of the items (leaves or merges) of one launch.  A single matrix
// derives node (depth, idx) from the bits of idx (plan == nullptr); a batch
// of matrices -- one flattened merge forest over the concatenated rows --
// reads each item from a plan entry (first row, size, matrix, flags).
struct Item {
    int s, sz, mat, flags;
};
constexpr int kItemRoot = 1;  // merge: its matrix's root merge; leaf: the whole matrix

struct Geo {
    int n;            // rows (all matrices' rows for a batch)
    int depth;        // tree depth of the launch's items (single matrix)
    const int4* plan;  // batch: one entry per item, else nullptr
    __device__ __forceinline__ Item item(int idx) const {
        Item it;
        if (plan) {
            const int4 p = plan[idx];
            it.s = p.x;
            it.sz = p.y;
            it.mat = p.z;
            it.flags = p.w;
        } else {
            node_geom(n, depth, idx, &it.s, &it.sz);
            it.mat = 0;
            it.flags = (depth == 0) ? kItemRoot : 0;
        }
        return it;
    }
};

// Where an item reports a failure: its matrix's INFO word, in the matrix's
// own row numbering (moff: first row of each matrix, nullptr for one matrix).
struct FailSite {
    int* info;
    int n, s;
};
__device__ __forceinline__ FailSite fail_site(int* info, const int* moff, int n, const Item& it) {
    if (!moff) return FailSite{info, n, it.s};
    const int o = moff[it.mat];
    return FailSite{info + it.mat, moff[it.mat + 1] - o, it.s - o};
}
__device__ __forceinline__ void report_item_failure(const FailSite& f, int size) {
    report_failure(f.info, f.n, f.s, size);
}

// ---------------------------------------------------------------------------
// max-norm, finiteness and scale (DLANST('M') + DLASCL, dstedc_dc_br.f:86-118)
// One block per matrix (moff: first row of each matrix; nullptr: one matrix
// of n rows), or, for one large matrix, nparts blocks over row slices that
// leave (max, non-finite flag) partials in part[2b], part[2b+1]; k_prep_e
// then forms the scale from them.
constexpr int kNormRowsPerPart = 4096;
constexpr int kNormMaxParts = 32;

__device__ __forceinline__ void norm_finish(double mm, bool bad, int nb, double* scal, int* info,
                                            int b, const int* moff, int* info_user) {
    const bool ok = mm > 0.0;
    scal[2 * b] = ok ? 1.0 / mm : 1.0;  // scale into [-1, 1]
    scal[2 * b + 1] = ok ? mm : 1.0;    // undo
    info[moff ? b : 0] = bad ? nb + 1 : 0;
    if (info_user && nb <= 1) *info_user = info[0];
}

__global__ __launch_bounds__(1024) void k_norm(int n, const int* __restrict__ moff,
                                               const double* __restrict__ d,
                                               const double* __restrict__ e, double* scal,
                                               int* info, int* info_user, double* part) {
    __shared__ double red[16];
    __shared__ int bad[16];
    const int b = blockIdx.x;
    const int nparts = gridDim.x;
    const bool split = (moff == nullptr && nparts > 1);
    const int off = moff ? moff[b] : 0;
    const int nb = moff ? moff[b + 1] - off : n;
    // rows [lo, hi) of the matrix (all of it, or this block's slice)
    const int lo = split ? (int)((long long)nb * b / nparts) : 0;
    const int hi = split ? (int)((long long)nb * (b + 1) / nparts) : nb;
    d += off;
    e += off;
    double m = 0.0;
    int nonfinite = 0;
    // eight rows per thread and step, all loads issued before the maxima, so
    // one block keeps many loads in flight
    constexpr int kU = 8;
    for (int i0 = lo + (int)threadIdx.x; i0 < hi; i0 += kU * blockDim.x) {
        double x[kU], y[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int i = i0 + u * (int)blockDim.x;
            x[u] = (i < hi) ? d[i] : 0.0;
            y[u] = (i < hi && i < nb - 1) ? e[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kU; u++) {
            nonfinite |= (int)!isfinite(x[u]) | (int)!isfinite(y[u]);
            m = fmax(m, fmax(fabs(x[u]), fabs(y[u])));
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
        if (split) {
            part[2 * b] = mm;
            part[2 * b + 1] = bb ? 1.0 : 0.0;
        } else {
            norm_finish(mm, bb != 0, nb, scal, info, b, moff, info_user);
        }
    }
}

// scaled off-diagonal with negligible entries zeroed (dstedc_dc_br.f:94-105):
// a zero there splits the matrix, which the tree then handles as a rho = 0
// merge (everything deflates) or inside a leaf
// Every row gets an entry; the last row of each matrix gets 0, so a batch's
// matrices are decoupled wherever a leaf or merge reads across their border.
// With nparts > 0 (one matrix whose norm k_norm left in nparts partials),
// every block forms the scale from the partials itself and block 0 stores it
// with INFO for the kernels that follow.
__global__ void k_prep_e(int n, const int* __restrict__ moff, int nmat,
                         const double* __restrict__ d, const double* __restrict__ e,
                         double* scal, double* ew, int nparts, const double* __restrict__ part,
                         int* info) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double sc;
    if (nparts > 0) {
        double mm = 0.0;
        bool bad = false;
        for (int q = 0; q < nparts; q++) {
            mm = fmax(mm, part[2 * q]);
            bad = bad || part[2 * q + 1] != 0.0;
        }
        sc = (mm > 0.0) ? 1.0 / mm : 1.0;
        if (i == 0) norm_finish(mm, bad, n, scal, info, 0, nullptr, nullptr);
    }
    if (i >= n) return;
    int b = 0, last = n - 1;
    if (moff) {  // the matrix holding row i: moff[b] <= i < moff[b + 1]
        int lo = 0, hi = nmat;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (moff[mid] <= i) lo = mid;
            else hi = mid;
        }
        b = lo;
        last = moff[b + 1] - 1;
    }
    if (i == last) {
        ew[i] = 0.0;
        return;
    }
    const double tiny = kEps * sqrt(fabs(d[i])) * sqrt(fabs(d[i + 1]));
    const double x = e[i];
    if (nparts <= 0) sc = scal[2 * b];
    ew[i] = (fabs(x) > tiny) ? x * sc : 0.0;
}

// ---------------------------------------------------------------------------
// Leaf eigensolve, one lane per leaf (the DSTEQR_BOUNDARY_BR role,
// dlaed0_br.f:354-371): implicit QL with the Wilkinson shift on the leaf's
// tridiagonal (Bowdler, Martin, Reinsch and Wilkinson), with each rotation
// applied only to the first and last rows of the eigenvector matrix, so a
// leaf costs O(m^2) flops.  The rank-one cut corrections (dlaed0_br.f:95-103)
// are applied to the leaf's end rows on load.  A lane's d, e and two rows
// live in LDS, element k of lane l at [k][l]: no lane touches another's
// entries, so there is no barrier.  Eigenvalues leave sorted ascending with
// their rows.
__device__ __forceinline__ double lapy2(double x, double y) {
    const double xa = fabs(x), ya = fabs(y);
    const double w = fmax(xa, ya), z = fmin(xa, ya);
    if (z == 0.0 || w > 1.79e308) return w;
    const double q = z * rcp_nr(w);
    return w * sqrt(fma(q, q, 1.0));
}

// Plane rotation of the QL sweep: r = sqrt(f^2 + g^2), c = g / r, s = f / r
// from one v_rsq_f64 refined by two Newton steps (about ten dependent
// operations, against some thirty for a scaled hypotenuse and a division).
// Inside a sweep |f| and |g| are bounded by the scaled matrix's norm, so
// f^2 + g^2 cannot overflow; when it underflows to zero (|f|, |g| below
// 1e-154, far under the deflation threshold) r = 0 takes the sweep's
// underflow exit, as an exact zero does.
__device__ __forceinline__ void ql_rotation(double f, double g, double& r, double& c,
                                            double& s) {
    const double t = fma(f, f, g * g);
    double y = __builtin_amdgcn_rsq(t);
#pragma unroll
    for (int it = 0; it < 2; it++) y = fma(0.5 * y, fma(-(t * y), y, 1.0), y);
    const bool zero = (t == 0.0);
    r = zero ? 0.0 : t * y;
    c = g * y;
    s = f * y;
}

// Lanes (leaves) per block for leaves of at most maxm rows: each lane holds
// four LDS vectors of maxm doubles, and a block stays within 64 KB.
__host__ __device__ inline int leaf_lanes(int maxm) { return maxm <= 32 ? kWave : kWave / 2; }
__host__ __device__ inline size_t leaf_lds_bytes(int maxm) {
    return sizeof(double) * 4 * (size_t)maxm * leaf_lanes(maxm);
}

// One leaf of m rows on a lane's four LDS vectors (element k of each at
// [k * ls]): loads the scaled diagonal with the cut corrections, the
// off-diagonal and the unit end rows, then runs the QL iteration.  On return
// vd, v0, v1 hold the eigenvalues and the two kept rows (unordered); the
// result is false when the iteration limit was reached.
__device__ __forceinline__ bool leaf_ql(int s, int m, int n, double scale,
                                        const double* __restrict__ d_in,
                                        const double* __restrict__ ew, int ls, double* vd,
                                        double* ve, double* v0, double* v1) {
#define sd_(k) vd[(k) * ls]
#define se_(k) ve[(k) * ls]
#define s0_(k) v0[(k) * ls]
#define s1_(k) v1[(k) * ls]
    // the leaf's entries, four rows at a time with their loads in flight
    // together
    const double ecut_lo = (s > 0) ? fabs(ew[s - 1]) : 0.0;
    const double ecut_hi = (s + m < n) ? fabs(ew[s + m - 1]) : 0.0;
    auto put = [&](int k, double dv, double ev) {
        // the shift rounds after the scaling, as a separate subtraction
#pragma clang fp contract(off)
        double dk = dv * scale;
        if (k == 0 && s > 0) dk -= ecut_lo;
        if (k == m - 1 && s + m < n) dk -= ecut_hi;
        sd_(k) = dk;
        se_(k) = (k < m - 1) ? ev : 0.0;
        s0_(k) = (k == 0) ? 1.0 : 0.0;
        s1_(k) = (k == m - 1) ? 1.0 : 0.0;
    };
    int k0 = 0;
    for (; k0 + 4 <= m; k0 += 4) {
        double dv[4], ev[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            dv[u] = d_in[s + k0 + u];
            ev[u] = (k0 + u < m - 1) ? ew[s + k0 + u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) put(k0 + u, dv[u], ev[u]);
    }
    for (; k0 < m; k0++) put(k0, d_in[s + k0], (k0 < m - 1) ? ew[s + k0] : 0.0);
    bool failed = false;
    const int maxit = 30 * m;
    int iter = 0;
    // e[k] is negligible when |e[k]| <= eps (|d[k]| + |d[k+1]|)
    auto negligible = [](double ek, double dk, double dk1) {
        return fabs(ek) <= kEps * (fabs(dk) + fabs(dk1));
    };
    // the first negligible off-diagonal at or after l (m - 1 if none), four
    // entries a step so that their loads are in flight together
    auto scan = [&](int from) {
        int k = from;
        for (; k + 4 <= m - 1; k += 4) {
            double dv[5], ev[4];
#pragma unroll
            for (int u = 0; u < 5; u++) dv[u] = sd_(k + u);
#pragma unroll
            for (int u = 0; u < 4; u++) ev[u] = se_(k + u);
            int hit = 4;
#pragma unroll
            for (int u = 3; u >= 0; u--)
                if (negligible(ev[u], dv[u], dv[u + 1])) hit = u;
            if (hit < 4) return k + hit;
        }
        for (; k < m - 1; k++)
            if (negligible(se_(k), sd_(k), sd_(k + 1))) return k;
        return m - 1;
    };
    for (int l = 0; l < m && !failed; l++) {
        int mm = scan(l);
        for (;;) {
            if (mm == l) break;  // d[l] is an eigenvalue
            if (++iter > maxit) {
                failed = true;
                break;
            }
            const double dl = sd_(l), el = se_(l);
            double gg = (sd_(l + 1) - dl) * rcp_nr(2.0 * el);
            double r = lapy2(gg, 1.0);
            gg = sd_(mm) - dl + el * rcp_nr(gg + copysign(r, gg));
            double sn = 1.0, c = 1.0, p = 0.0;
            bool underflow = false;
            // d[k+1] and the two rows' entries k+1 ride in registers from one
            // step to the next; each step loads entry k and stores entry k+1
            double dk1 = sd_(mm), x1 = s0_(mm), y1 = s1_(mm);
            // The sweep rewrites d[l..mm-1] and e[l..mm-1] and leaves e[mm]
            // zero, so the next iteration's first negligible off-diagonal is
            // the lowest k in [l, mm] that tests negligible on the stored
            // values: each step tests the entries it stores (dnx = the d[k+2]
            // stored by the step before).
            int nxt = mm;
            double dnx = 0.0;
            // entry k's loads are issued one step ahead: the step stores
            // entries k + 1 only, so the prefetched entries are current
            double ek_n = se_(mm - 1), dk_n = sd_(mm - 1), x0_n = s0_(mm - 1), y0_n = s1_(mm - 1);
            for (int k = mm - 1; k >= l; k--) {
                const double ek = ek_n, dk = dk_n;
                const double x0 = x0_n, y0 = y0_n;
                if (k > l) {
                    ek_n = se_(k - 1);
                    dk_n = sd_(k - 1);
                    x0_n = s0_(k - 1);
                    y0_n = s1_(k - 1);
                }
                const double f = sn * ek, b = c * ek;
                double cn, sv;
                ql_rotation(f, gg, r, cn, sv);
                se_(k + 1) = r;
                if (r == 0.0) {  // recover from underflow
                    sd_(k + 1) = dk1 - p;
                    s0_(k + 1) = x1;
                    s1_(k + 1) = y1;
                    se_(mm) = 0.0;
                    underflow = true;
                    break;
                }
                sn = sv;
                c = cn;
                const double ek1 = r;  // the e[k+1] just stored
                gg = dk1 - p;
                r = (dk - gg) * sn + 2.0 * c * b;
                p = sn * r;
                const double dnew = gg + p;
                sd_(k + 1) = dnew;
                if (k + 1 < mm && negligible(ek1, dnew, dnx)) nxt = k + 1;
                dnx = dnew;
                gg = c * r - b;
                // the two kept rows of the eigenvector matrix
                s0_(k + 1) = sn * x0 + c * x1;
                x1 = c * x0 - sn * x1;
                s1_(k + 1) = sn * y0 + c * y1;
                y1 = c * y0 - sn * y1;
                dk1 = dk;
            }
            if (underflow) {
                mm = scan(l);
                continue;
            }
            s0_(l) = x1;
            s1_(l) = y1;
            const double dl_new = dk1 - p;
            sd_(l) = dl_new;
            se_(l) = gg;
            se_(mm) = 0.0;
            mm = negligible(gg, dl_new, dnx) ? l : nxt;
        }
    }
    return !failed;
#undef sd_
#undef se_
#undef s0_
#undef s1_
}

// Ascending order of a solved leaf by rank (ties by position): emit(rank, x,
// first row, last row) for each of its m eigenvalues.
template <class Emit>
__device__ __forceinline__ void leaf_emit(int m, int ls, const double* vd, const double* v0,
                                          const double* v1, Emit emit) {
    for (int a = 0; a < m; a++) {
        const double x = vd[a * ls];
        int rank = 0, b = 0;
        for (; b + 4 <= m; b += 4) {
            double y[4];
#pragma unroll
            for (int u = 0; u < 4; u++) y[u] = vd[(b + u) * ls];
#pragma unroll
            for (int u = 0; u < 4; u++) rank += (y[u] < x) || (y[u] == x && b + u < a);
        }
        for (; b < m; b++) {
            const double y = vd[b * ls];
            rank += (y < x) || (y == x && b < a);
        }
        emit(rank, x, v0[a * ls], v1[a * ls]);
    }
}

// maxm: the largest leaf's row count (sizes the dynamic LDS: the fewer bytes
// per leaf, the more leaves are resident per CU); LP = leaf_lanes(maxm).
template <int LP>
__global__ __launch_bounds__(LP) void k_leaf(
    Geo g, int nleaves, int maxm, const int* __restrict__ moff, const double* __restrict__ d_in,
    const double* __restrict__ ew, const double* __restrict__ scal, double* out_final,
    double* out_d, double* out_lo, double* out_hi, int* info, int* info_user) {
    extern __shared__ double leaf_lds[];
    const int lane = threadIdx.x;
    const int leaf = blockIdx.x * LP + lane;
    if (leaf >= nleaves) return;
    // element k of this lane's vectors at [k * LP + lane]
    double* const vd = leaf_lds + lane;
    double* const ve = vd + (size_t)maxm * LP;
    double* const v0 = ve + (size_t)maxm * LP;
    double* const v1 = v0 + (size_t)maxm * LP;
    const Item it = g.item(leaf);
    const int s = it.s, m = it.sz, n = g.n;
    const bool single = (it.flags & kItemRoot) != 0;  // the leaf is its whole matrix
    if (!leaf_ql(s, m, n, scal[2 * it.mat], d_in, ew, LP, vd, ve, v0, v1))
        report_item_failure(fail_site(info, moff, n, it), m);
    const double unscale = scal[2 * it.mat + 1];
    leaf_emit(m, LP, vd, v0, v1, [&](int rank, double x, double lo, double hi) {
        if (single) {
            out_final[s + rank] = x * unscale;
        } else {
            out_d[s + rank] = x;
            out_lo[s + rank] = lo;
            out_hi[s + rank] = hi;
        }
    });
    if (single && info_user) *info_user = *info;  // a one-leaf solve: one matrix
}

// ---------------------------------------------------------------------------
// Merge step 1 (the DLAMRG merge + z/QBR set-up of DLAED7_BR and DLAED8_BR,
// dlaed7_br.f:95-113, 240-264): each element finds its rank in the merged
// order by binary search in the sibling (merge path; ties take the left child
// first, as DLAMRG does) and writes z = [last rows of left; sign(rho) * first
// rows of right] / sqrt(2) with the parent's two boundary rows
// q1 = [blo_left, 0], q2 = [0, bhi_right].
// Entries of the ascending a[0..len) below x (kLE: at most x).  An 8-way
// search: each round reads seven independent probes and keeps the eighth of
// [lo, hi) they bracket, so a search of n entries waits about log8(n) + 1
// memory round trips instead of the log2(n) of a bisection.
template <bool kLE>
__device__ __forceinline__ int count_before(const double* a, int len, double x) {
    int lo = 0, hi = len;  // a[i] passes for i < lo and fails for i >= hi
    while (hi - lo > 8) {
        const int span = hi - lo;
        int p[7];
        double v[7];
#pragma unroll
        for (int k = 0; k < 7; k++) {
            p[k] = lo + (int)(((long long)span * (k + 1)) >> 3);
            v[k] = a[p[k]];
        }
        // the passing probes are a prefix of the seven: lo moves past the
        // last of them, hi to the first failing one (selects, no indexing)
        int nlo = lo, nhi = hi;
#pragma unroll
        for (int k = 0; k < 7; k++) {
            const bool pass = kLE ? (v[k] <= x) : (v[k] < x);
            if (pass) nlo = p[k] + 1;
            if (!pass && nhi == hi) nhi = p[k];
        }
        lo = nlo;
        hi = nhi;
    }
    int c = 0;  // at most eight entries left, read together
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int i = lo + k;
        if (i < hi) {
            const double v = a[i];
            c += (kLE ? (v <= x) : (v < x)) ? 1 : 0;
        }
    }
    return lo + c;
}
__device__ __forceinline__ int count_less(const double* a, int len, double x) {
    return count_before<false>(a, len, x);
}
__device__ __forceinline__ int count_less_equal(const double* a, int len, double x) {
    return count_before<true>(a, len, x);
}

// Element idx < sz of the merge at rows [s, s + sz), rho the off-diagonal at
// its cut.
__device__ __forceinline__ void prepare_element(int s, int sz, int idx, double rho,
                                                const double* __restrict__ sd,
                                                const double* __restrict__ slo,
                                                const double* __restrict__ shi,
                                                double* __restrict__ md, double* __restrict__ mz,
                                                double* __restrict__ mq1,
                                                double* __restrict__ mq2) {
    const int cut = sz / 2;
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

__global__ __launch_bounds__(kThreads) void k_merge_prepare(
    Geo g, int bpm, const double* __restrict__ sd, const double* __restrict__ slo,
    const double* __restrict__ shi, const double* __restrict__ ew, double* __restrict__ md,
    double* __restrict__ mz, double* __restrict__ mq1, double* __restrict__ mq2) {
    const int merge = blockIdx.x / bpm, part = blockIdx.x % bpm;
    const Item it = g.item(merge);
    const int s = it.s, sz = it.sz;
    const int idx = part * kThreads + threadIdx.x;
    if (idx >= sz) return;
    prepare_element(s, sz, idx, ew[s + sz / 2 - 1], sd, slo, shi, md, mz, mq1, mq2);
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

// wave-level scans (wave64)
__device__ __forceinline__ int wave_incl_sum(int v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(v, o, kWave);
        if (lane >= o) v += y;
    }
    return v;
}
__device__ __forceinline__ int wave_incl_max(int v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(v, o, kWave);
        if (lane >= o) v = max(v, y);
    }
    return v;
}
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
    return v;
}

// Replays the real Givens run starting at t (ending before e): writes the
// deflated values and rows at g[goff..), keeps them ascending within the run,
// and stores the carried state of the run's surviving pole P[e-1].
__device__ __forceinline__ void replay_run(int s, int t, int e, int goff, double tol,
                                           const int* __restrict__ P, double* __restrict__ md,
                                           double* __restrict__ mz, double* __restrict__ mq1,
                                           double* __restrict__ mq2, int* __restrict__ fl,
                                           double* __restrict__ gd, double* __restrict__ gq1,
                                           double* __restrict__ gq2) {
    int cur = P[s + t - 1];
    double cd = md[s + cur], cz = mz[s + cur], c1 = mq1[s + cur], c2 = mq2[s + cur];
    // every value the run produces is a convex combination of the run's
    // original poles; clamping to their range removes last-ulp excursions so
    // the three output lists stay exactly sorted
    const double lo = cd;
    double hi = cd;
    const int base = goff;
    for (int u = t; u < e; u++) {
        const int j = P[s + u];
        const double dj = md[s + j], qj1 = mq1[s + j], qj2 = mq2[s + j];
        hi = dj;
        GivensStep g;
        givens_test(cd, cz, dj, mz[s + j], tol, &g);
        // DROT on the two selected rows (dlaed7_br.f:323-325) and the D
        // update (dlaed7_br.f:326-328)
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

// speculative end of the Givens run that starts at t if P[t-1] is unrotated:
// the first u whose test fails on the carried state, m if none does, or -2
// when the walk exceeds cap steps; -1 if pair (P[t-1], P[t]) does not rotate
__device__ __forceinline__ int speculative_end(int s, int t, int m, int cap, double tol,
                                               const int* __restrict__ P,
                                               const double* __restrict__ md,
                                               const double* __restrict__ mz) {
    const int jl = P[s + t - 1], j = P[s + t];
    GivensStep g;
    const double dj = md[s + j];
    if (!givens_test(md[s + jl], mz[s + jl], dj, mz[s + j], tol, &g)) return -1;
    double cd = md[s + jl] * g.s * g.s + dj * g.c * g.c;
    double cz = g.tau;
    int u = t + 1, steps = 0;
    while (u < m && steps < cap) {
        const int jj = P[s + u];
        const double d2 = md[s + jj];
        if (!givens_test(cd, cz, d2, mz[s + jj], tol, &g)) break;
        cd = cd * g.s * g.s + d2 * g.c * g.c;
        cz = g.tau;
        u++;
        steps++;
    }
    return (u < m && steps == cap) ? -2 : u;
}

// Kept (K), Givens-deflated (G) and tiny-deflated (T) counts of a merge.
struct DeflateCounts {
    int K, G, T;
};

// Barrier among the T threads that share a deflation: the block, or for
// T = 64 one wavefront, whose lanes run in lockstep, so the workgroup-scope
// fences (the waits for the wave's own loads and stores) and a stop to the
// compiler's reordering make every lane's writes visible to the others.
template <int T>
__device__ __forceinline__ void group_sync() {
    if constexpr (T == kWave) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        __syncthreads();
    }
}

// LDS ints a deflation of T threads needs (deflate_merge's sh argument)
template <int T>
constexpr int deflate_sh_ints() {
    return 2 * (T / kWave) + 1;
}

// The merge at rows [s, s + sz) deflated by T threads: the block, or one
// wavefront when T = 64 (every thread of the group calls it); ecut is the
// off-diagonal at its cut, sh the group's deflate_sh_ints<T>() LDS ints.
// rec, if not null, is the merge's count record.  Each wave owns a
// contiguous range of positions (and later of t), walked in coalesced
// 64-wide tiles with ballot/popcount compaction; cross-wave offsets come
// from per-wave totals in LDS.
template <int T>
__device__ __forceinline__ DeflateCounts deflate_merge(
    int s, int sz, double ecut, double* __restrict__ md, double* __restrict__ mz,
    double* __restrict__ mq1, double* __restrict__ mq2, int* __restrict__ P,
    int* __restrict__ rend, int* __restrict__ fl, double* __restrict__ ad, double* __restrict__ az,
    double2* __restrict__ adz, double2* __restrict__ aq, double* __restrict__ gd,
    double* __restrict__ gq1, double* __restrict__ gq2, double* __restrict__ td,
    double* __restrict__ tq1, double* __restrict__ tq2, int* __restrict__ rec, int* sh) {
    constexpr int NW = T / kWave;
    int* const wa = sh;
    int* const wb = sh + NW;
    int& s_fix = sh[2 * NW];
    const double rho = fabs(2.0 * ecut);
    const double tol = 8.0 * kEps * fmax(fabs(md[s]), fabs(md[s + sz - 1]));
    const int tid = threadIdx.x & (T - 1), lane = tid & (kWave - 1), w = tid / kWave;
    const unsigned long long lt = (1ull << lane) - 1ull;

    // ---- A: tiny-z flags, non-tiny compaction into P[0..m)
    const int R = ((sz + NW - 1) / NW + kWave - 1) / kWave * kWave;
    const int pb = min(sz, w * R), pe = min(sz, pb + R);
    int cnt = 0;
    for (int base = pb; base < pe; base += kWave) {
        const int p = base + lane;
        bool keep = false;
        if (p < pe) {
            keep = rho * fabs(mz[s + p]) > tol;
            fl[s + p] = keep ? 0 : 2;
        }
        cnt += __popcll(__ballot(keep));
    }
    if (lane == 0) wa[w] = cnt;
    if (tid == 0) s_fix = 0;
    group_sync<T>();
    int off = 0, m = 0;
    for (int q = 0; q < NW; q++) {
        const int v = wa[q];
        if (q < w) off += v;
        m += v;
    }
    for (int base = pb; base < pe; base += kWave) {
        const int p = base + lane;
        const bool keep = (p < pe) && fl[s + p] == 0;
        const unsigned long long b = __ballot(keep);
        if (keep) P[s + off + __popcll(b & lt)] = p;
        off += __popcll(b);
    }
    group_sync<T>();

    // ---- B: speculative runs from every rotating pair (independent per t)
    for (int t = 1 + tid; t < m; t += T) rend[s + t] = speculative_end(s, t, m, kSpecCap, tol, P, md, mz);
    group_sync<T>();

    // ---- B2: disjoint speculative runs make every candidate a real start
    // (the sequential walk would start each of them); any overlap, or a walk
    // past the cap, goes through one exact sequential resolution instead
    const int Rt = ((m + NW - 1) / NW + kWave - 1) / kWave * kWave;
    const int tb = min(m, w * Rt), te = min(m, tb + Rt);
    int carry = INT_MIN, first = INT_MAX;
    bool bad = false;
    for (int base = tb; base < te; base += kWave) {
        const int t = base + lane;
        const int e = (t >= 1 && t < te) ? rend[s + t] : -1;
        if (e == -2) bad = true;
        const int incl = wave_incl_max(e == -2 ? INT_MAX : e);
        int excl = __shfl_up(incl, 1, kWave);
        if (lane == 0) excl = INT_MIN;
        excl = max(excl, carry);
        if (e >= 0 && excl >= t) bad = true;
        if (e >= 0) first = min(first, t);
        carry = max(carry, __shfl(incl, kWave - 1, kWave));
    }
    first = wave_min(first);
    const bool wbad = __any(bad);
    group_sync<T>();  // wa/wb reuse
    if (lane == 0) {
        wa[w] = carry;
        wb[w] = first;
        if (wbad) s_fix = 1;
    }
    group_sync<T>();
    if (lane == 0) {
        int cin = INT_MIN;
        for (int q = 0; q < w; q++) cin = max(cin, wa[q]);
        if (wb[w] != INT_MAX && cin >= wb[w]) s_fix = 1;
    }
    group_sync<T>();
    if (s_fix) {
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
                    e = speculative_end(s, t, m, m, tol, P, md, mz);
                    rend[s + t] = e;
                }
                cover = e;
            }
        }
        group_sync<T>();
    }

    // ---- C: Givens-list offsets (scan of run lengths over t) and replay
    int runs = 0;
    for (int base = tb; base < te; base += kWave) {
        const int t = base + lane;
        const int e = (t >= 1 && t < te) ? rend[s + t] : -1;
        runs += (e >= 0) ? e - t : 0;
    }
    for (int o = 32; o > 0; o >>= 1) runs += __shfl_xor(runs, o, kWave);
    group_sync<T>();  // wa reuse
    if (lane == 0) wa[w] = runs;
    group_sync<T>();
    int goff = 0, G = 0;
    for (int q = 0; q < NW; q++) {
        const int v = wa[q];
        if (q < w) goff += v;
        G += v;
    }
    for (int base = tb; base < te; base += kWave) {
        const int t = base + lane;
        const int e = (t >= 1 && t < te) ? rend[s + t] : -1;
        const int len = (e >= 0) ? e - t : 0;
        const int incl = wave_incl_sum(len);
        if (e >= 0) replay_run(s, t, e, goff + incl - len, tol, P, md, mz, mq1, mq2, fl, gd, gq1, gq2);
        goff += __shfl(incl, kWave - 1, kWave);
    }
    group_sync<T>();

    // ---- D: compaction into kept poles and tiny-deflated entries
    int na = 0, nt = 0;
    for (int base = pb; base < pe; base += kWave) {
        const int p = base + lane;
        const int f = (p < pe) ? fl[s + p] : 1;
        na += __popcll(__ballot(f == 0));
        nt += __popcll(__ballot(f == 2));
    }
    group_sync<T>();  // wa/wb reuse
    if (lane == 0) {
        wa[w] = na;
        wb[w] = nt;
    }
    group_sync<T>();
    int aoff = 0, toff = 0, K = 0, Tn = 0;
    for (int q = 0; q < NW; q++) {
        const int va = wa[q], vt = wb[q];
        if (q < w) {
            aoff += va;
            toff += vt;
        }
        K += va;
        Tn += vt;
    }
    for (int base = pb; base < pe; base += kWave) {
        const int p = base + lane;
        const int f = (p < pe) ? fl[s + p] : 1;
        const unsigned long long ba = __ballot(f == 0), bt = __ballot(f == 2);
        if (f == 0) {
            const int o = s + aoff + __popcll(ba & lt);
            const double dv = md[s + p], zv = mz[s + p];
            ad[o] = dv;
            az[o] = zv;
            adz[o] = make_double2(dv, zv * zv);
            aq[o] = make_double2(mq1[s + p], mq2[s + p]);
        } else if (f == 2) {
            const int o = s + toff + __popcll(bt & lt);
            td[o] = md[s + p];
            tq1[o] = mq1[s + p];
            tq2[o] = mq2[s + p];
        }
        aoff += __popcll(ba);
        toff += __popcll(bt);
    }
    if (tid == 0 && rec) {
        rec[0] = K;
        rec[1] = G;
        rec[2] = Tn;
        rec[3] = 0;
        rec[4] = 0;
        rec[5] = 0;
        rec[6] = 0;
    }
    group_sync<T>();  // wa/wb are free for the group's next merge
    return DeflateCounts{K, G, Tn};
}

// One block per merge.
template <int T>
__global__ __launch_bounds__(T) void k_deflate(
    Geo g, const double* __restrict__ ew, double* __restrict__ md,
    double* __restrict__ mz, double* __restrict__ mq1, double* __restrict__ mq2, int* __restrict__ P,
    int* __restrict__ rend, int* __restrict__ fl, double* __restrict__ ad, double* __restrict__ az,
    double2* __restrict__ adz, double2* __restrict__ aq, double* __restrict__ gd,
    double* __restrict__ gq1, double* __restrict__ gq2, double* __restrict__ td,
    double* __restrict__ tq1, double* __restrict__ tq2, int* __restrict__ counts) {
    __shared__ int sh[deflate_sh_ints<T>()];
    const int merge = blockIdx.x;
    const Item it = g.item(merge);
    deflate_merge<T>(it.s, it.sz, ew[it.s + it.sz / 2 - 1], md, mz, mq1, mq2, P, rend, fl, ad, az,
                     adz, aq, gd, gq1, gq2, td, tq1, tq2, counts + kCountStride * merge, sh);
}

// ---------------------------------------------------------------------------
// Merge steps 3-5 share one work layout.  A block owns 64 consecutive items
// (roots or poles) of one merge, one item per lane; every lane of a wave
// streams the same index of the other dimension at the same time, so the
// streamed data (pole or root state) is wave-uniform and comes in through
// scalar loads.  For large K the block's W waves split the stream into W
// contiguous chunks and their partial results meet once per pass in LDS.
// W follows K at run time (at most the launch's width WL): about kChunk or
// more streamed entries per wave.
constexpr int kGroup = 64;
constexpr int kChunk = 256;

// Streamed data is read through the constant address space with wave-uniform
// indices, which selects scalar (SMEM) loads into SGPRs.
using cdbl = const __attribute__((address_space(4))) double*;
struct CPairs {  // array of (x, y) double pairs viewed through the constant space
    cdbl p;
    __device__ __forceinline__ double2 operator[](int j) const {
        return make_double2(p[2 * j], p[2 * j + 1]);
    }
};
struct CDoubles {
    cdbl p;
    __device__ __forceinline__ double operator[](int j) const { return p[j]; }
};
__device__ __forceinline__ CPairs as_const(const double2* q) {
    return CPairs{(cdbl) reinterpret_cast<const double*>(q)};
}
__device__ __forceinline__ CDoubles as_const(const double* q) { return CDoubles{(cdbl)q}; }
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// K of merges m0 .. m0+63 of a level, merge m0 + l in lane l (0 past the
// last merge): one load for 64 merges, where a loop over merges reading one
// count at a time would wait a memory round trip per merge
__device__ __forceinline__ int merge_counts(const int* counts, int m0, int merges) {
    const int m = m0 + (int)(threadIdx.x & (kWave - 1));
    return (m < merges) ? counts[kCountStride * m] : 0;
}
__device__ __forceinline__ int lane_value(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
constexpr int kMaxWaves = 8;  // 512 threads: the root solver needs ~180 VGPRs per lane

__host__ __device__ inline int waves_for(int K, int Wmax) {
    int W = 1;
    while (W < Wmax && K >= 2 * W * kChunk) W <<= 1;
    return W;
}

// Merge step 3: secular roots (DLAED9_BR's DLAED4_BR loop, dlaed9_br.f:181-262);
// per root the compact state (org, tau, near1, near2) plus d[org].
//
// Layout: L lanes per root (L a power of two picked per merge from K and the
// level's merge count), 64/L roots per wavefront, waves independent (no
// barriers).  Lane c of a root sums the poles c, c+L, c+2L, ... (loads
// coalesced across the root's L lanes, shared by the wave's roots), and a
// root's lanes combine their sums in an xor butterfly.  A wave therefore
// runs only as many passes as its 64/L roots need, and a small merge is
// spread over many lanes for latency.
// A double moved between lanes by a DPP pattern within rows of 16 lanes (a
// VALU operation, no LDS round trip).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// gfx950's lane swaps of two registers a, b (VALU, no LDS round trip):
//   half swap (v_permlane32_swap): lanes 32..63 of a <-> lanes 0..31 of b,
//   row swap (v_permlane16_swap): rows 1 and 3 of a <-> rows 0 and 2 of b
// (rows of 16 lanes).  Afterwards a + b in every lane is the lane's value of
// a (b) plus its partner's across lane bit 5 (4), where a lane whose bit is
// clear keeps a's values and a lane whose bit is set keeps b's: one step of
// a reduce-scatter; with a == b it is one step of the xor butterfly.  Every
// lane of the wave must be active.
template <bool kHalf>
__device__ __forceinline__ void lane_swap_f64(double& a, double& b) {
    const long long x = __double_as_longlong(a), y = __double_as_longlong(b);
    const unsigned xl = (unsigned)x, xh = (unsigned)(x >> 32);
    const unsigned yl = (unsigned)y, yh = (unsigned)(y >> 32);
    auto swap = [](unsigned p, unsigned q) {
        if constexpr (kHalf) return __builtin_amdgcn_permlane32_swap(p, q, false, false);
        else return __builtin_amdgcn_permlane16_swap(p, q, false, false);
    };
    const auto lo = swap(xl, yl);
    const auto hi = swap(xh, yh);
    a = __longlong_as_double((long long)(unsigned)lo[0] | ((long long)hi[0] << 32));
    b = __longlong_as_double((long long)(unsigned)lo[1] | ((long long)hi[1] << 32));
}

// Sums over aligned groups of L lanes (L a power of two up to 64) of six
// values; every lane of a group gets its group's totals.  Each step adds a
// lane's value and its partner's, with the partners of an xor butterfly on
// lane bits 0, 1, 2, 3 (DPP quad swaps, then the half-row and row mirrors,
// which pair lane i with a lane of the other quad and of the other half-row)
// and bits 4, 5 (row and half swaps), so the sums are those of the butterfly.
__device__ __forceinline__ void group_sum6(double (&acc)[6], int L) {
    if (L >= 2) {
#pragma unroll
        for (int q = 0; q < 6; q++) acc[q] += dpp_f64<0xB1>(acc[q]);  // quad [1,0,3,2]
    }
    if (L >= 4) {
#pragma unroll
        for (int q = 0; q < 6; q++) acc[q] += dpp_f64<0x4E>(acc[q]);  // quad [2,3,0,1]
    }
    if (L >= 8) {
#pragma unroll
        for (int q = 0; q < 6; q++) acc[q] += dpp_f64<0x141>(acc[q]);  // row half mirror
    }
    if (L >= 16) {
#pragma unroll
        for (int q = 0; q < 6; q++) acc[q] += dpp_f64<0x140>(acc[q]);  // row mirror
    }
    if (L >= 32) {
#pragma unroll
        for (int q = 0; q < 6; q++) {
            double a = acc[q], b = acc[q];
            lane_swap_f64<false>(a, b);
            acc[q] = a + b;
        }
    }
    if (L >= 64) {
#pragma unroll
        for (int q = 0; q < 6; q++) {
            double a = acc[q], b = acc[q];
            lane_swap_f64<true>(a, b);
            acc[q] = a + b;
        }
    }
}

struct PoleView {
    const double2* __restrict__ dz;  // (d_j, z_j^2)
    __device__ __forceinline__ double d(int j) const { return dz[j].x; }
    __device__ __forceinline__ double zsq(int j) const { return dz[j].y; }
};

constexpr int kBatch = 8;

// Reciprocal of the pole terms of a secular evaluation (the poles outside
// the root's exclusion window): v_rcp_f64 and one Newton step, relative error
// below 2^-50 (a few ulp).  The stopping bound of each pass, eps * erretm
// with erretm >= 8 A + sum_j |j - ii| |term_j| >= 9 A over these poles (A =
// sum of |term_j|), stays above that rounding.
__device__ __forceinline__ double rcp_stream(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// Sums of one side of the pole list for this lane's strided poles
// j = j_start, j_start + L, ... < j_end: S += z_j^2/delta_j, D += (z_j/delta_j)^2,
// E += j z_j^2/delta_j.  Poles are sorted and every iterate stays inside its
// bracket, so all poles below the origin give delta < 0 (psi side) and all
// above give delta > 0 (phi side): the two sides are two plain loops.  Four
// poles share one reciprocal: with u = x0 x1 and v = x2 x3, 1/x0 = x1 v /
// (u v) and so on (deflation keeps every delta above about 1e-16 of the
// spectrum's width and below it, so u v neither overflows nor underflows),
// and a batch's index-weighted sum is jd T + L sum_b b t_b.
__device__ __forceinline__ void secular_side(const double2* __restrict__ dz, int j, int jend,
                                             int L, double dorg, double tau, double& S,
                                             double& D, double& E) {
    double jd = (double)j;
    const double dL = (double)L;
    for (; j + (kBatch - 1) * L < jend; j += kBatch * L) {
        double2 p[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; b++) p[b] = dz[j + b * L];
        double T = 0.0, B = 0.0;
#pragma unroll
        for (int b = 0; b < kBatch; b += 4) {
            double dl[4];
#pragma unroll
            for (int q = 0; q < 4; q++) dl[q] = (p[b + q].x - dorg) - tau;
            const double u = dl[0] * dl[1], v = dl[2] * dl[3];
            const double ruv = rcp_stream(u * v);
            const double ru = v * ruv, rv = u * ruv;
            const double rr[4] = {dl[1] * ru, dl[0] * ru, dl[3] * rv, dl[2] * rv};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double t = p[b + q].y * rr[q];
                T = fma(p[b + q].y, rr[q], T);
                D = fma(t, rr[q], D);
                if (b + q > 0) B = fma((double)(b + q), t, B);
            }
        }
        S += T;
        E = fma(jd, T, E);
        E = fma(dL, B, E);
        jd += kBatch * dL;
    }
    for (; j < jend; j += L) {
        const double2 p = dz[j];
        const double r = rcp_stream((p.x - dorg) - tau);
        const double term = p.y * r;
        S += term;
        D = fma(term, r, D);
        E = fma(jd, term, E);
        jd += dL;
    }
}

// One group of 64/L roots of one merge; dz is the merge's pole list (d_j,
// z_j^2), in global memory or staged in LDS.  Returns the lane's root's
// secular evaluation count (on the root's first lane; 0 elsewhere) and adds a
// root that did not converge to *failed.
__device__ __forceinline__ int secular_group(const FailSite& fs, int s, int sz, int K, int L,
                                             int grp, double rho, const double* __restrict__ ad,
                                             const double* __restrict__ az,
                                             const double2* __restrict__ dz,
                                             double* __restrict__ rlam, double2* __restrict__ rs,
                                             double* __restrict__ rn1,
                                             double* __restrict__ rn2, int* failed) {
    const int lane = threadIdx.x & (kWave - 1);
    const int i = grp * (kWave / L) + lane / L;  // this lane's root
    const int c = lane & (L - 1);
    if (K <= 2) {
        if (i < K && c == 0) {
            if (K == 1) {
                const double z0 = az[s];
                rlam[s] = ad[s] + rho * z0 * z0;
                rs[s] = make_double2(ad[s], rho * z0 * z0);
                rn1[s] = 1.0;  // secular eigenvector (1)
                rn2[s] = 0.0;
            } else {
                double lam, v0, v1, tau;
                secular_two(i, ad + s, az + s, rho, &lam, &v0, &v1, &tau);
                rlam[s + i] = lam;
                rs[s + i] = make_double2(ad[s], tau);
                rn1[s + i] = v0;  // secular eigenvector (v0, v1)
                rn2[s + i] = v1;
            }
        }
        return (i == 0 && c == 0) ? 1 : 0;
    }
    const PoleView E{dz};
    SecularSolve<PoleView> st;
    bool active = (i < K);
    st.init(E, K, min(i, K - 1), rho);
    for (;;) {
        if (!__any(active)) break;
        // psi side: poles [0, ex_lo); phi side: poles [ex_lo + nex, K)
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const int j2 = st.ev_exlo + st.ev_nex;
        secular_side(dz, c, st.ev_exlo, L, st.ev_org, st.ev_tau, acc[0], acc[1], acc[2]);
        secular_side(dz, j2 + ((c - j2) & (L - 1)), K, L, st.ev_org, st.ev_tau, acc[3], acc[4],
                     acc[5]);
        group_sum6(acc, L);
        SecSums sums;
        sums.S = acc[0] + acc[3];
        sums.A = acc[3] - acc[0];
        sums.dpsi = acc[1];
        sums.dphi = acc[4];
        sums.J = acc[2] + acc[5];
        if (active && st.advance(E, sums)) active = false;
    }
    int iters = 0;
    if (i < K && c == 0) {
        const RootState& r = st.out;
        rlam[s + i] = r.lambda;
        rs[s + i] = make_double2(dz[r.org].x, r.tau);
        rn1[s + i] = r.near1;
        rn2[s + i] = r.near2;
        iters = r.iters;
        if (r.info) {
            report_item_failure(fs, sz);
            *failed += 1;
        }
    }
    return iters;
}

// Adds a wave's evaluation count, its most evaluations of one root and its
// non-converged roots to a merge's counters (one atomic of each per wave).
__device__ __forceinline__ void add_secular_counts(int* __restrict__ rec, int iters, int imax,
                                                   int failed) {
    for (int o = 32; o > 0; o >>= 1) {
        iters += __shfl_down(iters, o, kWave);
        imax = max(imax, __shfl_down(imax, o, kWave));
        failed += __shfl_down(failed, o, kWave);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (iters) atomicAdd(&rec[3], iters);
        if (imax) atomicMax(&rec[5], imax);
        if (failed) atomicAdd(&rec[6], failed);
    }
}

// All roots of the level's merges with K < kstream.  Work is split into
// tasks (merge, slot), slot < slots; the task's wave solves root groups slot,
// slot + slots, ... of its merge, so every group is covered whatever the slot
// count, and the grid is capped, each wave walking tasks with a grid stride.  A merge of at most
// stage_k poles is first copied into the wave's own stage_k-entry slice of
// the dynamic LDS (WL * stage_k double2), so every pass and every pole
// lookup of the root solver reads LDS; larger merges read global memory.
// One wave's LDS operations complete in order, so the copy needs no barrier.
// Three waves per SIMD (at most 168 VGPRs): the root solver's state machine
// is latency-bound, and the occupancy hides its dependent FP64 chains better
// than the few spills it costs.
template <int WL>
__global__ __launch_bounds__(kWave * WL) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_secular(
    Geo g, int merges, int slots, int kstream, int stage_k, int target,
    const double* __restrict__ ew, const int* counts, const double* __restrict__ ad,
    const double* __restrict__ az, const double2* __restrict__ adz, double* __restrict__ rlam,
    double2* __restrict__ rs, double* __restrict__ rn1, double* __restrict__ rn2, int* counts_rw,
    const int* __restrict__ moff, int* info) {
    extern __shared__ double2 sec_stage[];
    const int ntask = merges * slots;
    const int nwave = gridDim.x * WL;
    const int lane = threadIdx.x & (kWave - 1);
    double2* const stage = sec_stage + (threadIdx.x / kWave) * stage_k;
    for (int task = uniform(blockIdx.x * WL + (int)(threadIdx.x / kWave)); task < ntask;
         task += nwave) {
        // slot-major: the tasks a merge actually needs (its first slots)
        // are contiguous across merges, so the grid stride spreads them
        // evenly over the waves whatever the slot over-provisioning
        const int slot = task / merges, merge = task - slot * merges;
        const int K = uniform(counts[kCountStride * merge]);
        const int L = secular_lanes(K, merges, target);
        const int rpw = kWave / L;  // roots per wave
        if (slot * rpw >= K || K >= kstream) continue;  // idle slot / streamed merge
        const Item it = g.item(merge);
        const int s = it.s, sz = it.sz;
        const FailSite fs = fail_site(info, moff, g.n, it);
        const double rho = fabs(2.0 * ew[s + sz / 2 - 1]);
        int iters = 0, imax = 0, failed = 0;
        if (K <= stage_k) {
            for (int j = lane; j < K; j += kWave) stage[j] = adz[s + j];
            __builtin_amdgcn_wave_barrier();
            for (int grp = slot; grp * rpw < K; grp += slots) {
                const int it = secular_group(fs, s, sz, K, L, grp, rho, ad, az, stage, rlam, rs,
                                             rn1, rn2, &failed);
                iters += it;
                imax = max(imax, it);
            }
            __builtin_amdgcn_wave_barrier();
        } else {
            for (int grp = slot; grp * rpw < K; grp += slots) {
                const int it = secular_group(fs, s, sz, K, L, grp, rho, ad, az, adz + s, rlam,
                                             rs, rn1, rn2, &failed);
                iters += it;
                imax = max(imax, it);
            }
        }
        add_secular_counts(counts_rw + kCountStride * merge, iters, imax, failed);
    }
}

// Packed secular mode for levels of many small merges: the level's roots
// (merges with K < kstream, in merge order) are numbered 0 .. total-1, root
// i of merge m being roff[m] + i, and a wave takes 64/L consecutive roots
// whatever merges they belong to, so no lane idles for a merge whose K is
// not a multiple of the group.  k_root_scan forms roff (exclusive prefix of
// the merges' K, roff[merges] = total) in one block.
constexpr int kScanThreads = 1024;
constexpr int kScanPer = 4;  // consecutive merges per thread in a tile
__global__ __launch_bounds__(kScanThreads) void k_root_scan(int merges, int kstream,
                                                             const int* __restrict__ counts,
                                                             int* __restrict__ roff) {
    constexpr int NW = kScanThreads / kWave;
    constexpr int kTile = kScanThreads * kScanPer;
    __shared__ int wsum[NW];
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    int carry = 0;  // roots of the earlier tiles
    for (int base = 0; base < merges; base += kTile) {
        // this thread's merges: base + t * kScanPer ... (+ kScanPer - 1)
        int k[kScanPer], sum = 0;
#pragma unroll
        for (int u = 0; u < kScanPer; u++) {
            const int m = base + t * kScanPer + u;
            const int K = (m < merges) ? counts[kCountStride * m] : 0;
            k[u] = (K < kstream) ? K : 0;
            sum += k[u];
        }
        const int incl = wave_incl_sum(sum);
        if (lane == kWave - 1) wsum[w] = incl;
        __syncthreads();
        int off = carry + incl - sum, tile = 0;
        for (int q = 0; q < NW; q++) {
            const int x = wsum[q];
            if (q < w) off += x;
            tile += x;
        }
#pragma unroll
        for (int u = 0; u < kScanPer; u++) {
            const int m = base + t * kScanPer + u;
            if (m < merges) roff[m] = off;
            off += k[u];
        }
        carry += tile;
        __syncthreads();  // wsum is rewritten by the next tile
    }
    if (t == 0) roff[merges] = carry;
}

// All roots of the level's merges with K < kstream, packed (see k_root_scan):
// L lanes per root from the level's mean K, each wave a contiguous run of
// roots.  The wave's 64/L root positions take the run's roots in order, and
// a position whose root has converged takes the next root of the run at the
// start of the next pass, so the wave's passes follow the roots' mean
// evaluation count rather than every group's slowest root.  A position's
// merge is found by a walk over roff from the run's first merge (found once
// by binary search); roots of merges of at most two poles take the closed
// form as they are claimed.  Poles are read from global memory (the L1 cache
// holds the few merges a wave spans).  Each root's arithmetic is that of a
// lone root at L lanes, whatever it runs beside.
template <int WL>
__global__ __launch_bounds__(kWave * WL) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_secular_packed(
    Geo g, int merges, int target, const int* __restrict__ roff, const double* __restrict__ ew,
    const int* __restrict__ counts, const double* __restrict__ ad,
    const double* __restrict__ az, const double2* __restrict__ adz, double* __restrict__ rlam,
    double2* __restrict__ rs, double* __restrict__ rn1, double* __restrict__ rn2,
    int* counts_rw, const int* __restrict__ moff, int* info) {
    const int total = uniform(roff[merges]);
    if (total == 0) return;
    const int L = secular_lanes((total + merges - 1) / merges, merges, target);
    const int rpg = kWave / L;  // root positions per wave
    const long long groups = (total + rpg - 1) / rpg;
    const int nwave = gridDim.x * WL;
    const int wave = uniform(blockIdx.x * WL + (int)(threadIdx.x / kWave));
    const int lane = threadIdx.x & (kWave - 1);
    const int c = lane & (L - 1), lead = lane - c;  // lane c of the position at lanes lead..
    const int g0 = (int)(groups * wave / nwave), g1 = (int)(groups * (wave + 1) / nwave);
    if (g0 >= g1) return;
    // this wave's run of roots [R0, R1)
    const int R0 = g0 * rpg, R1 = (int)min<long long>((long long)g1 * rpg, total);
    // the merge holding root R0: the last m with roff[m] <= R0
    int m;
    {
        int lo = 0, hi = merges;  // roff[lo] <= R0 < roff[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (roff[mid] <= R0) lo = mid;
            else hi = mid;
        }
        m = uniform(lo);
    }
    int next = R0;  // the run's next unclaimed root (wave-uniform)
    // this position's root: its merge m (walks forward), K, index i, rows s
    int K = 0, i = 0, s = 0, sz = 0;
    double rho = 0.0;
    bool active = false;
    const double2* __restrict__ dz = adz;
    SecularSolve<PoleView> st;
    for (;;) {
        const unsigned long long need = __ballot(!active && c == 0);
        if (need && next < R1) {
            // one root per idle position, in rank order
            const int rank = (int)__popcll(need & ((1ull << lead) - 1ull));
            const int R = next + rank;
            next = min(next + (int)__popcll(need), R1);
            if (!active && R < R1) {
                while (roff[m + 1] <= R) m++;
                K = counts[kCountStride * m];
                i = R - roff[m];
                const Item it = g.item(m);
                s = it.s;
                sz = it.sz;
                rho = fabs(2.0 * ew[s + sz / 2 - 1]);
                dz = adz + s;
                if (K >= 3) {
                    st.init(PoleView{dz}, K, i, rho);
                    active = true;
                } else if (c == 0) {  // two-pole closed form / single pole
                    if (K == 1) {
                        const double z0 = az[s];
                        rlam[s] = ad[s] + rho * z0 * z0;
                        rs[s] = make_double2(ad[s], rho * z0 * z0);
                        rn1[s] = 1.0;
                        rn2[s] = 0.0;
                    } else {
                        double lam, v0, v1, tau;
                        secular_two(i, ad + s, az + s, rho, &lam, &v0, &v1, &tau);
                        rlam[s + i] = lam;
                        rs[s + i] = make_double2(ad[s], tau);
                        rn1[s + i] = v0;
                        rn2[s + i] = v1;
                    }
                    if (i == 0) {
                        atomicAdd(&counts_rw[kCountStride * m + 3], 1);
                        atomicMax(&counts_rw[kCountStride * m + 5], 1);
                    }
                }
            }
        }
        if (!__any(active)) {
            if (next >= R1) break;  // the run is done
            continue;               // the positions took closed-form roots only
        }
        const PoleView E{dz};
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (active) {
            const int j2 = st.ev_exlo + st.ev_nex;
            secular_side(dz, c, st.ev_exlo, L, st.ev_org, st.ev_tau, acc[0], acc[1], acc[2]);
            secular_side(dz, j2 + ((c - j2) & (L - 1)), K, L, st.ev_org, st.ev_tau, acc[3],
                         acc[4], acc[5]);
        }
        group_sum6(acc, L);
        SecSums sums;
        sums.S = acc[0] + acc[3];
        sums.A = acc[3] - acc[0];
        sums.dpsi = acc[1];
        sums.dphi = acc[4];
        sums.J = acc[2] + acc[5];
        if (active && st.advance(E, sums)) {
            active = false;
            if (c == 0) {
                const RootState& r = st.out;
                rlam[s + i] = r.lambda;
                rs[s + i] = make_double2(dz[r.org].x, r.tau);
                rn1[s + i] = r.near1;
                rn2[s + i] = r.near2;
                if (r.iters) {
                    atomicAdd(&counts_rw[kCountStride * m + 3], r.iters);
                    atomicMax(&counts_rw[kCountStride * m + 5], r.iters);
                }
                if (r.info) {
                    report_item_failure(fail_site(info, moff, g.n, g.item(m)), sz);
                    atomicAdd(&counts_rw[kCountStride * m + 6], 1);
                }
            }
        }
    }
}

// Streamed secular solve for the merges with K >= kstream (the big merges
// of the top levels, where the O(K^2) work sits).
//
// A workgroup of W waves serves kStreamSlots * 64/L roots at a time: L lanes
// per root position and kStreamSlots roots (slots) per position, the same
// positions in every wave.  A pass cuts the K poles into kStreamSlices fixed
// slices; the waves take slices from a counter in LDS (the older waves issue
// first and finish their share sooner, so a fixed slice per wave leaves them
// idle at the barrier), and lane c of a position takes every L-th pole of a
// slice.  Every lane of a wave needs the same few poles at the same time:
// the wave fetches 128 poles with two coalesced loads and the lanes read
// them from LDS (L distinct addresses per read, broadcast within each
// position's lanes).  Each slice's partial sums have their own row in LDS
// and are added in slice order, whichever wave computed them.  Wave s holds
// the state machines of slot s, advances them and publishes the next pass's
// evaluation points, so a pass costs two barriers and one state-machine
// step.
//
// Roots are handed out dynamically: a slot whose root has converged takes
// the next root of the merge's counter.  Roots need 3 to 5 passes depending
// on where they lie in the spectrum, so a static split leaves some
// workgroups with far more passes than others.  L is a function of the
// merge alone (its size and its share of the grid): a root's arithmetic
// never depends on which workgroup or which neighbours it ran with.
constexpr int kStreamBatch = 8;
constexpr int kStageChunk = 128;  // poles staged per chunk (two per lane)

struct StreamAcc {  // one side's sums over a pole range
    double S, D, J;
};

// A load that names the global address space.  Through a generic pointer the
// compiler emits a flat load, which also counts against the LDS wait counter:
// the first LDS read after the next chunk's prefetch would then wait for
// that prefetch.
typedef double vec2d __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const vec2d global_vec2d;
__device__ __forceinline__ double2 load_global(const double2* p, int j) {
    const vec2d v = ((global_vec2d*)p)[j];
    return make_double2(v.x, v.y);
}

// Poles [ja, jb) pass through this wave's LDS slot in chunks of 128: lane l
// fetches poles c0 + l and c0 + 64 + l (the next chunk's loads are in flight
// while the current chunk is consumed).  One wave's LDS operations complete
// in order, so the slot needs no barrier.
template <class F>
__device__ __forceinline__ void staged_range(const double2* __restrict__ dz, double2* slot, int ja,
                                             int jb, F&& consume) {
    if (ja >= jb) return;
    const int lane = threadIdx.x & (kWave - 1);
    const double2 zero = make_double2(0.0, 0.0);
    double2 n0 = (ja + lane < jb) ? load_global(dz, ja + lane) : zero;
    double2 n1 = (ja + kWave + lane < jb) ? load_global(dz, ja + kWave + lane) : zero;
    for (int c0 = ja; c0 < jb; c0 += kStageChunk) {
        slot[lane] = n0;
        slot[kWave + lane] = n1;
        const int c1 = c0 + kStageChunk;
        if (c1 < jb) {
            n0 = (c1 + lane < jb) ? load_global(dz, c1 + lane) : zero;
            n1 = (c1 + kWave + lane < jb) ? load_global(dz, c1 + kWave + lane) : zero;
        }
        consume(c0, min(kStageChunk, jb - c0));
    }
}

constexpr int kStreamSlots = 1;    // roots per lane of the streamed kernel
constexpr int kStreamSlices = 12;  // pole slices per pass of the streamed kernel
// Exclusion flags: one bit per 128 poles of a merge (the stage chunk), set
// when an active root excludes a pole there; merges beyond kFlagWords * 32
// chunks run every chunk in the excluding form.
constexpr int kFlagWords = 128;

// One wave's share [q0, q1) of a pass for this lane's roots r < NR, whose
// excluded poles are j with (unsigned)(j - exlo[r]) < nex[r].  Lane c of a
// position takes poles c, c + L, ... of each batch of kStreamBatch * L
// poles; the batches lie on a grid fixed by q0.  A chunk that holds no
// active root's excluded pole (the bulk: `flags` clear) runs the plain form.
// A chunk that holds one runs the same arithmetic with the excluded terms
// replaced by exact zeros, so every root that is not excluded there gets
// bitwise the sums of the plain form: a root's sums do not depend on which
// roots share its wave.  Each root's terms go to one running sum; the part
// below the root (delta < 0, the psi side) is split off at the batch that
// holds its first excluded pole, and the rest is the phi side.
template <int L, int NR>
__device__ __forceinline__ void stream_pass(const double2* __restrict__ dz, double2* slot, int q0,
                                            int q1, int c, const double (&org)[NR],
                                            const double (&tau)[NR], const int (&exlo)[NR],
                                            const int (&nex)[NR], const unsigned* flags,
                                            bool flag_all, StreamAcc (&lo)[NR],
                                            StreamAcc (&hi)[NR]) {
    // The two batch forms spell out every fused multiply-add; the compiler
    // may fuse nothing else, so both give each root the same roundings.
#pragma clang fp contract(off)
    constexpr int kB = kStreamBatch * L;
    constexpr double dL = (double)L;
    StreamAcc cur[NR];
    bool split[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        cur[r] = StreamAcc{0.0, 0.0, 0.0};
        lo[r] = StreamAcc{0.0, 0.0, 0.0};
        hi[r] = StreamAcc{0.0, 0.0, 0.0};
        split[r] = false;
    }
    staged_range(dz, slot, q0, q1, [&](int c0, int nb) {
        bool win = flag_all;
        if (!win) {
            const int ga = c0 / kStageChunk, gb = (c0 + nb - 1) / kStageChunk;
            const unsigned fa = uniform((int)flags[ga >> 5]), fb = uniform((int)flags[gb >> 5]);
            win = (((fa >> (ga & 31)) | (fb >> (gb & 31))) & 1u) != 0;
        }
        const int nfull = nb / kB;  // batches every lane has in full
        int o = c;
        if (!win) {
            for (int bi = 0; bi < nfull; bi++, o += kB) {
                double2 p[kStreamBatch];
#pragma unroll
                for (int b = 0; b < kStreamBatch; b++) p[b] = slot[o + b * L];
                const int jl = c0 + o;  // this lane's first pole of the batch
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    // J gets sum_b (jl + b L) t_b = jl * T + L * sum_b b t_b
                    double T = 0.0, B = 0.0;
#pragma unroll
                    for (int b = 0; b < kStreamBatch; b += 4) {
                        // one reciprocal for four poles: with u = xy, v = zw,
                        // 1/u = v/(uv) and 1/v = u/(uv), then 1/x = y/u and so
                        // on.  Poles closer than the deflation tolerance were
                        // deflated, so every delta here is at least about
                        // 1e-16 and at most the spectrum's width: uv neither
                        // overflows nor underflows.
                        double dl[4];
#pragma unroll
                        for (int q = 0; q < 4; q++) dl[q] = (p[b + q].x - org[r]) - tau[r];
                        const double u = dl[0] * dl[1], v = dl[2] * dl[3];
                        const double ruv = rcp_stream(u * v);
                        const double ru = v * ruv, rv = u * ruv;
                        const double rr[4] = {dl[1] * ru, dl[0] * ru, dl[3] * rv, dl[2] * rv};
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const double t = p[b + q].y * rr[q];
                            T = fma(p[b + q].y, rr[q], T);
                            cur[r].D = fma(t, rr[q], cur[r].D);
                            if (b + q > 0) B = fma((double)(b + q), t, B);
                        }
                    }
                    cur[r].S += T;
                    cur[r].J = fma((double)jl, T, cur[r].J);
                    cur[r].J = fma(dL, B, cur[r].J);
                }
            }
        } else {
            for (int bi = 0; bi < nfull; bi++, o += kB) {
                double2 p[kStreamBatch];
#pragma unroll
                for (int b = 0; b < kStreamBatch; b++) p[b] = slot[o + b * L];
                const int j0 = c0 + bi * kB, jl = c0 + o;
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    // the root's own batch splits its terms at exlo: Tp, Dp, Bp
                    // collect the part above it
                    const bool at = (unsigned)(exlo[r] - j0) < (unsigned)kB;
                    double T = 0.0, B = 0.0, Tp = 0.0, Dp = 0.0, Bp = 0.0;
#pragma unroll
                    for (int b = 0; b < kStreamBatch; b += 4) {
                        double dl[4];
                        bool ex[4];
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            ex[q] = (unsigned)(jl + (b + q) * L - exlo[r]) < (unsigned)nex[r];
                            dl[q] = ex[q] ? 1.0 : (p[b + q].x - org[r]) - tau[r];
                        }
                        const double u = dl[0] * dl[1], v = dl[2] * dl[3];
                        const double ruv = rcp_stream(u * v);
                        const double ru = v * ruv, rv = u * ruv;
                        const double rr[4] = {dl[1] * ru, dl[0] * ru, dl[3] * rv, dl[2] * rv};
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const double y = ex[q] ? 0.0 : p[b + q].y;
                            const bool above = at && (jl + (b + q) * L > exlo[r]);
                            const double yc = above ? 0.0 : y, ya = above ? y : 0.0;
                            const double tc = yc * rr[q], ta = ya * rr[q];
                            T = fma(yc, rr[q], T);
                            cur[r].D = fma(tc, rr[q], cur[r].D);
                            Tp = fma(ya, rr[q], Tp);
                            Dp = fma(ta, rr[q], Dp);
                            if (b + q > 0) {
                                B = fma((double)(b + q), tc, B);
                                Bp = fma((double)(b + q), ta, Bp);
                            }
                        }
                    }
                    cur[r].S += T;
                    cur[r].J = fma((double)jl, T, cur[r].J);
                    cur[r].J = fma(dL, B, cur[r].J);
                    if (at) {
                        lo[r] = cur[r];
                        cur[r].S = Tp;
                        cur[r].D = Dp;
                        cur[r].J = fma(dL, Bp, (double)jl * Tp);
                        split[r] = true;
                    }
                }
            }
        }
        // the end of the share: fewer than kB poles, one at a time
        for (; o < nb; o += L) {
            const int j = c0 + o;
            const double2 p = slot[o];
#pragma unroll
            for (int r = 0; r < NR; r++) {
                if (!split[r] && j >= exlo[r]) {
                    lo[r] = cur[r];
                    cur[r] = StreamAcc{0.0, 0.0, 0.0};
                    split[r] = true;
                }
                const bool ex = (unsigned)(j - exlo[r]) < (unsigned)nex[r];
                const double rr = rcp_stream(ex ? 1.0 : (p.x - org[r]) - tau[r]);
                const double t = ex ? 0.0 : p.y * rr;
                cur[r].S += t;
                cur[r].D = fma(t, rr, cur[r].D);
                cur[r].J = fma((double)j, t, cur[r].J);
            }
        }
    });
#pragma unroll
    for (int r = 0; r < NR; r++) {
        // unsplit with the root above the share: every pole lies below it
        const bool below = !split[r] && exlo[r] >= q0;
        lo[r].S = below ? cur[r].S : lo[r].S;
        lo[r].D = below ? cur[r].D : lo[r].D;
        lo[r].J = below ? cur[r].J : lo[r].J;
        hi[r].S = below ? 0.0 : cur[r].S;
        hi[r].D = below ? 0.0 : cur[r].D;
        hi[r].J = below ? 0.0 : cur[r].J;
    }
}

// Lanes per root position of the streamed kernel for a K-pole merge whose
// workgroups take about `roots` roots per slot each: enough lanes that each
// of the 64 / L positions works through about 4 or more roots per slot, no
// more than leaves each lane at least 256 poles of each of the `slices`
// slices of a pass, at most 16.
__device__ __forceinline__ int stream_lanes(int K, int roots, int slices) {
    const int want = (256 + roots - 1) / max(roots, 1);
    int L = 1;
    while (L < 16 && L < want && (long long)K >= 2LL * slices * 128 * L) L <<= 1;
    return L;
}

// The first evaluation of every root of the streamed merges (DLAED4's
// midpoint pass): S_i = sum over j outside the root's two interval poles of
// z_j^2 / ((d_j - d_org) - tau) at tau = the interval midpoint (root i < K-1:
// org = i, tau = (d_{i+1} - d_i)/2, poles i, i+1 excluded; root K-1: org =
// K-1, tau = rho/2, poles K-2, K-1 excluded), written to rlam[i] for the
// streamed kernel's initial guess.  Only S is needed here, so four terms
// share one reciprocal and one accumulation:
//   t0 + t1 + t2 + t3 = ((y0 x1 + y1 x0) v + (y2 x3 + y3 x2) u) / (u v),
//   u = x0 x1, v = x2 x3, y = z^2, x = delta.
// A tile is 64 consecutive roots, one per lane; wave w of the workgroup
// takes the pole slice [K w/W, K (w+1)/W), the same poles for every lane
// (wave-uniform loads).  Poles below the tile's window [i0 - 1, i0 + 65) lie
// below every root of the tile and poles above it above, so four terms of
// a quad share a sign there; the window's poles are taken one at a time
// with the per-lane exclusion.  Each slice's sums meet in LDS and are added
// in slice order.
template <int W>
__global__ __launch_bounds__(kWave * W) void k_secular_mid(Geo g, int merges, int kstream,
                                                           const double* __restrict__ ew,
                                                           const int* __restrict__ counts,
                                                           const double2* __restrict__ adz,
                                                           double* __restrict__ rlam) {
    __shared__ double part[W][kWave];
    const int wave = uniform(threadIdx.x / kWave), lane = threadIdx.x & (kWave - 1);
    int t = blockIdx.x;  // this workgroup's next tile
    int tbase = 0;       // tiles of the streamed merges before `merge`
    int Kl = 0;          // counts of merges 64c .. 64c+63, c = merge / 64
    for (int merge = 0; merge < merges; merge++) {
        if ((merge & (kWave - 1)) == 0) Kl = merge_counts(counts, merge, merges);
        const int K = lane_value(Kl, merge & (kWave - 1));
        if (K < kstream) continue;
        const int nt = (K + kWave - 1) / kWave;
        for (; t < tbase + nt; t += gridDim.x) {
            const Item it = g.item(merge);
            const int s = it.s, sz = it.sz;
            const double rho = fabs(2.0 * ew[s + sz / 2 - 1]);
            const double2* __restrict__ dz = adz + s;
            const int i0 = (t - tbase) * kWave, i = i0 + lane;
            // the evaluation point of SecularSolve::init
            double org = 0.0, tau = 0.0;
            int exlo = K;
            if (i < K - 1) {
                org = dz[i].x;
                tau = (dz[i + 1].x - org) * 0.5;
                exlo = i;
            } else if (i == K - 1) {
                org = dz[K - 1].x;
                tau = rho * 0.5;
                exlo = K - 2;
            }
            const int wlo = max(0, i0 - 1), whi = min(K, i0 + kWave + 1);
            const int q0 = (int)(((long long)K * wave) / W);
            const int q1 = (int)(((long long)K * (wave + 1)) / W);
            double S = 0.0;
            // eight poles (two quads) in pole order
            auto batch8 = [&](const double2(&p)[8]) {
#pragma unroll
                for (int h = 0; h < 8; h += 4) {
                    double x[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) x[q] = (p[h + q].x - org) - tau;
                    const double u = x[0] * x[1], v = x[2] * x[3];
                    const double ruv = rcp_stream(u * v);
                    const double a = fma(p[h + 1].y, x[0], p[h].y * x[1]);
                    const double b = fma(p[h + 3].y, x[2], p[h + 2].y * x[3]);
                    S = fma(fma(b, u, a * v), ruv, S);
                }
            };
            auto load8 = [&](double2(&p)[8], int j) {
#pragma unroll
                for (int q = 0; q < 8; q++) p[q] = dz[j + q];
            };
            auto plain = [&](int ja, int jb) {
                int j = ja;
                if (j + 8 <= jb) {
                    // two buffers of eight poles in turn: a buffer's (scalar)
                    // loads are waited for before the other's are issued, so
                    // each batch's arithmetic overlaps the next batch's loads
                    // without copies between the buffers (scalar loads return
                    // out of order: any wait is a wait for all of them)
                    double2 A[8], B[8];
                    load8(A, j);
                    for (; j + 16 <= jb; j += 16) {
                        __builtin_amdgcn_sched_barrier(0);
                        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): A is in
                        load8(B, j + 8);
                        __builtin_amdgcn_sched_barrier(0);   // B in flight before A's work
                        batch8(A);
                        __builtin_amdgcn_sched_barrier(0);
                        __builtin_amdgcn_s_waitcnt(0xC07F);  // B is in
                        load8(A, min(j + 16, jb - 8));       // in range; used if j + 24 <= jb
                        __builtin_amdgcn_sched_barrier(0);
                        batch8(B);
                    }
                    if (j + 8 <= jb) {  // A holds poles j .. j+7
                        batch8(A);
                        j += 8;
                    }
                }
                for (; j < jb; j++) {
                    const double2 p = dz[j];
                    S = fma(p.y, rcp_stream((p.x - org) - tau), S);
                }
            };
            plain(q0, min(q1, wlo));
            for (int j = max(q0, wlo); j < min(q1, whi); j++) {
                const double2 p = dz[j];
                const bool ex = (unsigned)(j - exlo) < 2u;
                const double r = rcp_stream(ex ? 1.0 : (p.x - org) - tau);
                S = fma(ex ? 0.0 : p.y, r, S);
            }
            plain(max(q0, whi), q1);
            part[wave][lane] = S;
            __syncthreads();
            if (wave == 0 && i < K) {
                double sum = 0.0;
                for (int w = 0; w < W; w++) sum += part[w][lane];
                rlam[s + i] = sum;
            }
            __syncthreads();  // part[] is read before the next tile rewrites it
        }
        tbase += nt;
    }
}

template <int W>
__global__ __launch_bounds__(kWave * W)
__attribute__((amdgpu_waves_per_eu(kStreamSlots == 1 ? 4 : 2, kStreamSlots == 1 ? 4 : 2)))
void k_secular_stream(
    Geo g, int merges, int kstream, const double* __restrict__ ew, const int* counts,
    const double2* __restrict__ adz, double* __restrict__ rlam, double2* __restrict__ rs,
    double* __restrict__ rn1, double* __restrict__ rn2, int* counts_rw,
    const int* __restrict__ moff, int* info) {
    static_assert(W >= kStreamSlots, "one advancing wave per slot");
    constexpr int NR = kStreamSlots;
    __shared__ double part[kStreamSlices][NR][6][kWave];
    __shared__ double2 stage[W][kStageChunk];
    __shared__ double pass_org[NR][kWave], pass_tau[NR][kWave];
    __shared__ int pass_exlo[NR][kWave], pass_nex[NR][kWave];
    __shared__ unsigned pass_flags[2][kFlagWords];  // by pass parity
    __shared__ int pass_live[NR];
    __shared__ int pass_next;  // the next slice to take
    // the slots' root state machines, one per lane and slot: kept in LDS, so
    // no registers hold them through the pole loops
    __shared__ SecularSolve<PoleView> roots[NR][kWave];
    const int wave = uniform(threadIdx.x / kWave), lane = threadIdx.x & (kWave - 1);
    double2* slot = stage[wave];
    const bool owner = wave < NR;  // wave s advances slot s
    // Workgroups to merges: a root of a K-pole merge weighs K (its passes each
    // stream K poles), and workgroup b covers the weight interval
    // [wsum b / G, wsum (b+1) / G); it serves, in order, every merge whose
    // interval meets its own, taking roots there until the merge's counter
    // runs out.
    long long wsum = 0;
    for (int m0 = 0; m0 < merges; m0 += kWave) {
        const long long Kl = merge_counts(counts, m0, merges);
        long long wl = (Kl >= kstream) ? Kl * Kl : 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) wl += __shfl_xor(wl, o, kWave);
        wsum += ((long long)uniform((int)(wl >> 32)) << 32) | (unsigned)uniform((int)wl);
    }
    if (wsum == 0) return;  // no merge of the level is streamed
    // workgroup b takes interval v = 2 (b mod h) + b div h, h = ceil(G/2): the
    // SIMDs favour older waves and a CU's two workgroups are either b and
    // b + h (the younger) or neighbours, so every merge gets as many of the
    // younger workgroups as of the older ones, and no merge lags behind
    const long long G = gridDim.x;
    const long long h = (G + 1) / 2;
    const long long v = 2 * ((long long)blockIdx.x % h) + (long long)blockIdx.x / h;
    const long long wa = wsum * v / G;
    const long long wb = wsum * (v + 1) / G;
    long long wbase = 0;
    for (int merge = 0; merge < merges; merge++) {
        const int K = uniform(counts[kCountStride * merge]);
        if (K < kstream) continue;
        const long long w0 = wbase, w1 = wbase + (long long)K * K;
        wbase = w1;
        if (w1 <= wa || w0 >= wb) continue;
        // the workgroups that serve this merge: b with [wsum b/G, wsum (b+1)/G)
        // meeting [w0, w1)
        const long long bfirst = w0 * G / wsum;
        const long long blast = (w1 * G + wsum - 1) / wsum - 1;
        const int share = (int)max<long long>(1, K / max<long long>(1, blast - bfirst + 1));
        const Item it = g.item(merge);
        const int s = it.s, sz = it.sz;
        const FailSite fs = fail_site(info, moff, g.n, it);
        const double rho = fabs(2.0 * ew[s + sz / 2 - 1]);
        const double2* __restrict__ dz = adz + s;
        const PoleView E{dz};
        const int L = uniform(stream_lanes(K, max(1, share / NR), kStreamSlices));
        const int c = lane & (L - 1), lead = lane - c;  // lane c of the position at lanes lead..
        const bool flag_all = K > kFlagWords * 32 * kStageChunk;
        int* claim = counts_rw + kCountStride * merge + 4;
        for (int x = threadIdx.x; x < 2 * kFlagWords; x += kWave * W) pass_flags[x / kFlagWords][x % kFlagWords] = 0u;
        __syncthreads();
        // an owning wave's slot: state machine, claims, results
        SecularSolve<PoleView>& st = roots[owner ? wave : 0][lane];
        int i = -1;              // this lane's root in the slot, -1 when it needs one
        bool exhausted = false;  // the merge's counter has run out
        int iters = 0, imax = 0, failed = 0;
        int parity = 0;
        for (;;) {
            if (owner) {
                const unsigned long long need = __ballot(i < 0 && c == 0);
                if (need && !exhausted) {
                    // one root per needing position, numbered by rank
                    const int nneed = (int)__popcll(need);
                    const int rank = (int)__popcll(need & ((1ull << lead) - 1ull));
                    int b0 = 0;
                    if (lane == 0) b0 = atomicAdd(claim, nneed);
                    const int base = uniform(__shfl(b0, 0, kWave));
                    if (base + nneed >= K) exhausted = true;
                    if (i < 0 && base + rank < K) {
                        i = base + rank;
                        st.init(E, K, i, rho);
                        // the midpoint pass ran in k_secular_mid
                        SecSums s0 = {rlam[s + i], 0.0, 0.0, 0.0, 0.0};
                        st.advance(E, s0);
                    }
                }
                const bool active = (i >= 0);
                const int exlo = active ? st.ev_exlo : K, nex = active ? st.ev_nex : 0;
                pass_org[wave][lane] = active ? st.ev_org : 0.0;
                pass_tau[wave][lane] = active ? st.ev_tau : 0.0;
                pass_exlo[wave][lane] = exlo;
                pass_nex[wave][lane] = nex;
                if (active && c == 0 && !flag_all) {
                    const int ga = exlo / kStageChunk, gb = (exlo + nex - 1) / kStageChunk;
                    atomicOr(&pass_flags[parity][ga >> 5], 1u << (ga & 31));
                    atomicOr(&pass_flags[parity][gb >> 5], 1u << (gb & 31));
                }
                const bool live = __any(active);
                if (lane == 0) pass_live[wave] = live ? 1 : 0;
                if (wave == 0 && lane == 0) pass_next = 0;
            }
            __syncthreads();
            int nlive = 0;
#pragma unroll
            for (int r = 0; r < NR; r++) nlive += pass_live[r];
            if (uniform(nlive) == 0) break;  // the merge is done (same value in every wave)
            if (wave == W - 1) {
                // the other parity's flags were read in the last pass and are
                // set again only after this pass's second barrier
                for (int x = lane; x < kFlagWords; x += kWave) pass_flags[parity ^ 1][x] = 0u;
            }
            double org[NR], tau[NR];
            int exl[NR], nx[NR];
#pragma unroll
            for (int r = 0; r < NR; r++) {
                org[r] = pass_org[r][lane];
                tau[r] = pass_tau[r][lane];
                exl[r] = pass_exlo[r][lane];
                nx[r] = pass_nex[r][lane];
            }
            for (;;) {
                int k0 = 0;
                if (lane == 0) k0 = atomicAdd(&pass_next, 1);
                const int k = uniform(__shfl(k0, 0, kWave));
                if (k >= kStreamSlices) break;
                const int q0 = (int)(((long long)K * k) / kStreamSlices);
                const int q1 = (int)(((long long)K * (k + 1)) / kStreamSlices);
                StreamAcc lo[NR], hi[NR];
                auto pass = [&](auto lanes) {
                    constexpr int LL = decltype(lanes)::value;
                    stream_pass<LL, NR>(dz, slot, q0, q1, c, org, tau, exl, nx,
                                        pass_flags[parity], flag_all, lo, hi);
                };
                switch (L) {  // stream_lanes returns 1, 2, 4, 8 or 16
                    case 1: pass(std::integral_constant<int, 1>{}); break;
                    case 2: pass(std::integral_constant<int, 2>{}); break;
                    case 4: pass(std::integral_constant<int, 4>{}); break;
                    case 8: pass(std::integral_constant<int, 8>{}); break;
                    default: pass(std::integral_constant<int, 16>{}); break;
                }
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    double acc[6] = {lo[r].S, lo[r].D, lo[r].J, hi[r].S, hi[r].D, hi[r].J};
                    group_sum6(acc, L);
#pragma unroll
                    for (int q = 0; q < 6; q++) part[k][r][q][lane] = acc[q];
                }
            }
            __syncthreads();
            // the other waves wait at the next barrier; part[] and pass_*[] are
            // rewritten only after it
            if (owner && i >= 0) {
                double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                for (int k = 0; k < kStreamSlices; k++) {
#pragma unroll
                    for (int q = 0; q < 6; q++) acc[q] += part[k][wave][q][lane];
                }
                SecSums sums;
                sums.S = acc[0] + acc[3];
                sums.A = acc[3] - acc[0];
                sums.dpsi = acc[1];
                sums.dphi = acc[4];
                sums.J = acc[2] + acc[5];
                if (st.advance(E, sums)) {
                    const RootState& r = st.out;
                    if (c == 0) {
                        rlam[s + i] = r.lambda;
                        rs[s + i] = make_double2(st.dorg, r.tau);  // dorg = d_org
                        rn1[s + i] = r.near1;
                        rn2[s + i] = r.near2;
                        if (r.info) {
                            report_item_failure(fs, sz);
                            failed += 1;
                        }
                        iters += r.iters;
                        imax = max(imax, r.iters);
                    }
                    i = -1;
                }
            }
            parity ^= 1;
        }
        if (owner) add_secular_counts(counts_rw + kCountStride * merge, iters, imax, failed);
        __syncthreads();  // pass_*[] of this merge are read before the next merge rewrites them
    }
}

// Merges of at most small_max_k(maxsz) poles get their w and boundary rows
// from the first block of their finalize launch, in LDS.  Levels of merges
// of at most kSmallMaxK rows need no other launch for them; above that the
// product-form and boundary-row launches take the merges of more than
// kSmallTopK poles.
constexpr int kSmallMaxK = 256;
constexpr int kSmallTopK = 128;
__host__ __device__ inline int small_max_k(int maxsz) {
    return maxsz <= kSmallMaxK ? kSmallMaxK : kSmallTopK;
}

// Root data of the product-form loop: (d_org, tau) pairs and d.  RootsSmem
// reads them through scalar loads (one base per batch, so a batch arrives in a
// few wide loads); RootsLds reads a merge staged in LDS.
struct RootsSmem {
    cdbl rs, d;
    static constexpr bool kLazy = false;
    // two 4-root buffers with ordered scalar waits measured slower here than
    // the plain 8-root batch at all but the largest merges (a half's 20
    // VALU operations hide little of a scalar load)
    static constexpr bool kPipe = false;
    __device__ __forceinline__ void batch4(int j, double2 (&r)[4], double (&dj)[4]) const {
        const cdbl rb = rs + 2 * (size_t)j;
        const cdbl db = d + j;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            r[b] = make_double2(rb[2 * b], rb[2 * b + 1]);
            dj[b] = db[b];
        }
    }
    __device__ __forceinline__ void batch(int j, double2 (&r)[kBatch], double (&dj)[kBatch]) const {
        const cdbl rb = rs + 2 * (size_t)j;
        const cdbl db = d + j;
#pragma unroll
        for (int b = 0; b < kBatch; b++) {
            r[b] = make_double2(rb[2 * b], rb[2 * b + 1]);
            dj[b] = db[b];
        }
    }
    __device__ __forceinline__ double2 r(int j) const {
        return make_double2(rs[2 * (size_t)j], rs[2 * (size_t)j + 1]);
    }
    __device__ __forceinline__ double dj(int j) const { return d[j]; }
};
struct RootsLds {
    const double2* rs;
    const double* d;
    static constexpr bool kLazy = false;
    static constexpr bool kPipe = false;
    __device__ __forceinline__ void batch4(int, double2 (&)[4], double (&)[4]) const {}
    __device__ __forceinline__ void batch(int j, double2 (&r)[kBatch], double (&dj)[kBatch]) const {
#pragma unroll
        for (int b = 0; b < kBatch; b++) {
            r[b] = rs[j + b];
            dj[b] = d[j + b];
        }
    }
    __device__ __forceinline__ double2 r(int j) const { return rs[j]; }
    __device__ __forceinline__ double dj(int j) const { return d[j]; }
};

// Merge step 4: product-form update vector (dlaed9_br.f:218-266) at pole i
// (group of poles from g0), over the roots [j0, j1) of a K-pole merge:
// the share prod_j delta_j(i) / prod_{j != i} (d_i - d_j) of
// w_i = sign(z_i) sqrt(-prod_j delta_j(i) / prod_{j != i} (d_i - d_j)).
// delta_j(i) comes from root j's compact state; the two pole-adjacent
// entries are root j's near1/near2 (n1, n2; dlaed9_br.f:272-331).
template <class R>
__device__ __forceinline__ double wprod_share(const R& roots, const double* n1, const double* n2,
                                              int i, double di, int g0, int j0, int j1, int K) {
    // the product of ratios as a ratio of products num / den * 2^ex, with the
    // exponents moved out after every batch: no division per pair
    double num = 1.0, den = 1.0;
    int ex = 0;
    auto renorm = [&]() {
        const int en = __builtin_amdgcn_frexp_exp(num), ed = __builtin_amdgcn_frexp_exp(den);
        num = __builtin_amdgcn_ldexp(num, -en);
        den = __builtin_amdgcn_ldexp(den, -ed);
        ex += en - ed;
    };
    // roots whose near1/near2 apply to a pole of this group: j in [g0-1, g0+63]
    // (near1 at pole j, near2 at pole j+1) and the last root (poles K-2, K-1)
    const int wa = max(j0, g0 - 1), wb = min(j1, g0 + kGroup);
    // the near deltas this pole can take in the loop: near1 of root i and
    // near2 of root i-1 (read once, not per pair)
    const double a1 = (i < K - 1) ? n1[i] : 0.0;
    const double a2 = (i >= 1) ? n2[i - 1] : 0.0;
    auto term_to = [&](double& nu, double& de, double2 r, double dj, int j, bool check) {
        double dl = (di - r.x) - r.y;
        if (check) {  // j == i and the pole-adjacent roots lie in the window only
            if (i == j) dl = a1;
            else if (i == j + 1) dl = a2;
            nu *= dl;
            de *= (j == i) ? 1.0 : di - dj;
        } else {
            nu *= dl;
            de *= di - dj;
        }
    };
    auto term = [&](double2 r, double dj, int j, bool check) { term_to(num, den, r, dj, j, check); };
    auto range = [&](int ja, int jb, bool check) {
        int j = ja;
        if constexpr (R::kPipe) {
            if (j + kBatch <= jb) {
                // a batch of eight roots in two buffers of four: a buffer's
                // (scalar) loads are waited for before the other's are issued
                // (scalar loads return out of order: any wait is a wait for
                // all of them), so each half's arithmetic overlaps the next
                // half's loads; the product chains are those of the plain
                // batch (even pairs b = 0, 2, 4, 6; odd pairs 1, 3, 5, 7)
                double2 rA[4], rB[4];
                double dA[4], dB[4];
                roots.batch4(j, rA, dA);
                for (; j + kBatch <= jb; j += kBatch) {
                    __builtin_amdgcn_sched_barrier(0);
                    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): A is in
                    roots.batch4(j + 4, rB, dB);
                    __builtin_amdgcn_sched_barrier(0);   // B in flight before A's work
                    double ne = 1.0, de = 1.0, no = 1.0, dd = 1.0;
                    term_to(ne, de, rA[0], dA[0], j, check);
                    term_to(ne, de, rA[2], dA[2], j + 2, check);
                    term_to(no, dd, rA[1], dA[1], j + 1, check);
                    term_to(no, dd, rA[3], dA[3], j + 3, check);
                    __builtin_amdgcn_sched_barrier(0);
                    __builtin_amdgcn_s_waitcnt(0xC07F);  // B is in
                    // the next batch's first half (in range; used if j + 16 <= jb)
                    roots.batch4(min(j + kBatch, jb - 4), rA, dA);
                    __builtin_amdgcn_sched_barrier(0);
                    term_to(ne, de, rB[0], dB[0], j + 4, check);
                    term_to(ne, de, rB[2], dB[2], j + 6, check);
                    term_to(no, dd, rB[1], dB[1], j + 5, check);
                    term_to(no, dd, rB[3], dB[3], j + 7, check);
                    num = (num * ne) * no;
                    den = (den * de) * dd;
                    renorm();
                }
            }
        }
        for (; j + kBatch <= jb; j += kBatch) {
            double2 r[kBatch];
            double dj[kBatch];
            roots.batch(j, r, dj);
            // even and odd pairs in two product chains, joined before the
            // exponents move out
            const double n0 = num, d0 = den;
            num = 1.0;
            den = 1.0;
            auto tb = [&](int b) {
                if constexpr (R::kLazy) term(roots.r(j + b), roots.dj(j + b), j + b, check);
                else term(r[b], dj[b], j + b, check);
            };
#pragma unroll
            for (int b = 0; b < kBatch; b += 2) tb(b);
            const double n1v = num, d1v = den;
            num = 1.0;
            den = 1.0;
#pragma unroll
            for (int b = 1; b < kBatch; b += 2) tb(b);
            num = (n0 * n1v) * num;
            den = (d0 * d1v) * den;
            renorm();
        }
        for (; j < jb; j++) {
            term(roots.r(j), roots.dj(j), j, check);
            renorm();
        }
    };
    const int jlast = (K - 1 >= j0 && K - 1 < j1) ? K - 1 : -1;
    const int jend = (jlast >= 0) ? K - 1 : j1;  // the last root is handled apart
    if (wa < wb) {
        range(j0, min(wa, jend), false);
        range(wa, min(wb, jend), true);
        range(max(wb, j0), jend, false);
    } else {
        range(j0, jend, false);
    }
    if (jlast >= 0) {  // root K-1: near1 at pole K-2, near2 at pole K-1
        const double2 r = roots.r(jlast);
        double dl = (di - r.x) - r.y;
        if (i == K - 2) dl = n1[jlast];
        else if (i == K - 1) dl = n2[jlast];
        num *= dl;
        den *= (jlast == i) ? 1.0 : di - roots.dj(jlast);
        renorm();
    }
    return __builtin_amdgcn_ldexp(num / den, ex);
}

// Pole data of the boundary-row loop: product-form pairs (d_i, w_i) and the
// merged children's rows (q1_i, q2_i), through scalar loads or from LDS.
struct PolesSmem {
    cdbl pw, aq;
    static constexpr bool kLazyQ = false;
    static constexpr bool kPipe = true;  // two 4-pole buffers, ordered scalar waits
    __device__ __forceinline__ void batch4(int i, double2 (&p)[4], double2 (&q)[4]) const {
        const cdbl pb = pw + 2 * (size_t)i;
        const cdbl qb = aq + 2 * (size_t)i;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            p[b] = make_double2(pb[2 * b], pb[2 * b + 1]);
            q[b] = make_double2(qb[2 * b], qb[2 * b + 1]);
        }
    }
    __device__ __forceinline__ void batch(int i, double2 (&p)[kBatch], double2 (&q)[kBatch]) const {
        const cdbl pb = pw + 2 * (size_t)i;
        const cdbl qb = aq + 2 * (size_t)i;
#pragma unroll
        for (int b = 0; b < kBatch; b++) {
            p[b] = make_double2(pb[2 * b], pb[2 * b + 1]);
            q[b] = make_double2(qb[2 * b], qb[2 * b + 1]);
        }
    }
    __device__ __forceinline__ double2 p(int i) const {
        return make_double2(pw[2 * (size_t)i], pw[2 * (size_t)i + 1]);
    }
    __device__ __forceinline__ double2 q(int i) const {
        return make_double2(aq[2 * (size_t)i], aq[2 * (size_t)i + 1]);
    }
};
struct PolesLds {
    const double2* pw;
    const double2* aq;
    static constexpr bool kLazyQ = false;
    static constexpr bool kPipe = false;
    __device__ __forceinline__ void batch4(int, double2 (&)[4], double2 (&)[4]) const {}
    __device__ __forceinline__ void batch(int i, double2 (&p)[kBatch], double2 (&q)[kBatch]) const {
#pragma unroll
        for (int b = 0; b < kBatch; b++) {
            p[b] = pw[i + b];
            q[b] = aq[i + b];
        }
    }
    __device__ __forceinline__ double2 p(int i) const { return pw[i]; }
    __device__ __forceinline__ double2 q(int i) const { return aq[i]; }
};

// Merge step 5: the two boundary rows of the merged eigenvector block
// (dlaed9_br.f:272-331) at root j (group of roots from g0), over the poles
// [i0, i1) of a K-pole merge, K >= 3: x_i = w_i / delta_j(i) summed into
// temp = sum x_i^2, acc_r = sum_i q_r(i) x_i (row_r(j) = acc_r / sqrt(temp)).
template <class P>
__device__ __forceinline__ void rows_share(const P& poles, int j, double2 r, double nr1,
                                           double nr2, int g0, int i0, int i1, int K,
                                           double& temp, double& acc1, double& acc2) {
    const double dorg = r.x, tau = r.y;
    const int a = (j == K - 1) ? K - 2 : j;
    auto delta = [&](double2 p, int i, bool check) {
        double dl = (p.x - dorg) - tau;
        if (check) {
            if (i == a) dl = nr1;
            else if (i == a + 1) dl = nr2;
        }
        return dl;
    };
    auto accumulate = [&](double x, double2 q) {
        temp = fma(x, x, temp);
        acc1 = fma(q.x, x, acc1);
        acc2 = fma(q.y, x, acc2);
    };
    // one group of four poles (one shared reciprocal) from i
    auto group4 = [&](const double2 (&p)[4], const double2 (&q)[4], int i, bool check) {
        double dl[4];
#pragma unroll
        for (int u = 0; u < 4; u++) dl[u] = delta(p[u], i + u, check);
        const double x01 = dl[0] * dl[1], x23 = dl[2] * dl[3];
        const double rq = rcp_stream(x01 * x23);
        const double r01 = x23 * rq, r23 = x01 * rq;
        accumulate(p[0].y * (dl[1] * r01), q[0]);
        accumulate(p[1].y * (dl[0] * r01), q[1]);
        accumulate(p[2].y * (dl[3] * r23), q[2]);
        accumulate(p[3].y * (dl[2] * r23), q[3]);
    };
    auto range = [&](int ia, int ib, bool check) {
        int i = ia;
        if constexpr (P::kPipe) {
            if (i + kBatch <= ib) {
                // two buffers of four poles in turn, as in wprod_share: each
                // group's arithmetic overlaps the next group's scalar loads
                double2 pA[4], qA[4], pB[4], qB[4];
                poles.batch4(i, pA, qA);
                for (; i + kBatch <= ib; i += kBatch) {
                    __builtin_amdgcn_sched_barrier(0);
                    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): A is in
                    poles.batch4(i + 4, pB, qB);
                    __builtin_amdgcn_sched_barrier(0);
                    group4(pA, qA, i, check);
                    __builtin_amdgcn_sched_barrier(0);
                    __builtin_amdgcn_s_waitcnt(0xC07F);  // B is in
                    poles.batch4(min(i + kBatch, ib - 4), pA, qA);  // used if i + 16 <= ib
                    __builtin_amdgcn_sched_barrier(0);
                    group4(pB, qB, i + 4, check);
                }
            }
        }
        for (; i + kBatch <= ib; i += kBatch) {
            double2 p[kBatch], q[kBatch];
            poles.batch(i, p, q);
#pragma unroll
            for (int b = 0; b < kBatch; b += 4) {
                // one reciprocal for four poles, as in the streamed secular
                // kernel (deltas of undeflated poles keep the product in range)
                double dl[4];
#pragma unroll
                for (int u = 0; u < 4; u++) dl[u] = delta(p[b + u], i + b + u, check);
                const double x01 = dl[0] * dl[1], x23 = dl[2] * dl[3];
                const double rq = rcp_stream(x01 * x23);
                const double r01 = x23 * rq, r23 = x01 * rq;
                auto qv = [&](int u) {
                    if constexpr (P::kLazyQ) return poles.q(i + b + u);
                    else return q[b + u];
                };
                accumulate(p[b].y * (dl[1] * r01), qv(0));
                accumulate(p[b + 1].y * (dl[0] * r01), qv(1));
                accumulate(p[b + 2].y * (dl[3] * r23), qv(2));
                accumulate(p[b + 3].y * (dl[2] * r23), qv(3));
            }
        }
        for (; i < ib; i++) accumulate(poles.p(i).y * rcp_stream(delta(poles.p(i), i, check)), poles.q(i));
    };
    // poles that may be pole-adjacent to a root of this group: [g0, g0+64]
    // and the last root's K-2, K-1
    const int wa = max(i0, g0), wb = min(i1, g0 + kGroup + 1);
    const int tail = max(i0, min(i1, K - 2));  // from here on always checked
    if (wa < wb) {
        range(i0, min(wa, tail), false);
        range(wa, min(wb, tail), true);
        range(max(wb, i0), tail, false);
    } else {
        range(i0, tail, false);
    }
    range(tail, i1, true);
}

// Boundary rows of a merge of K <= 2 poles: the secular eigenvector (near1,
// near2 of each root) applied to the merged rows (dlaed9_br.f:146-172).
__device__ __forceinline__ void rows_direct(int K, double v0, double v1, double2 q0, double2 q1,
                                            double* o1, double* o2) {
    double a1 = q0.x * v0, a2 = q0.y * v0;
    if (K == 2) {
        a1 += q1.x * v1;
        a2 += q1.y * v1;
    }
    *o1 = a1;
    *o2 = a2;
}

// LDS of a small merge solved whole by one block: doubles per pole.
constexpr int kSmallMergeDoubles = 12;  // ad, az, rn1, rn2, rs (2), aq (2), pw (2), rows (2)
__host__ __device__ constexpr size_t small_merge_lds(int maxsz) {
    return sizeof(double) * kSmallMergeDoubles * (size_t)(maxsz < kSmallMaxK ? maxsz : kSmallMaxK);
}
// A merge of K <= kSmallMaxK poles at rows from s, its product form and
// boundary rows computed by the whole block (nt threads) into o1[j], o2[j]
// (LDS, j < K): poles and root states staged in LDS (sh), then w and the
// rows with L lanes per pole (root), L from K and nt: lane c of an item
// takes the c-th of L consecutive root (pole) ranges, with the arithmetic of
// the wide kernels' shares, and the L shares meet in an xor butterfly (w:
// products of the shares; rows: sums).  K <= 2 takes the secular
// eigenvector directly.  Every thread of the block calls it.
__device__ __forceinline__ void small_merge_rows(int s, int K, int nt, double* sh,
                                                 const double* __restrict__ ad,
                                                 const double* __restrict__ az,
                                                 const double2* __restrict__ rs,
                                                 const double* __restrict__ rn1,
                                                 const double* __restrict__ rn2,
                                                 const double2* __restrict__ aq, double* o1,
                                                 double* o2) {
    const int tid = threadIdx.x;
    if (K <= 2) {
        if (tid < K)
            rows_direct(K, rn1[s + tid], rn2[s + tid], aq[s], K == 2 ? aq[s + 1] : aq[s],
                        &o1[tid], &o2[tid]);
        __syncthreads();
        return;
    }
    double* const sd = sh;
    double* const sz = sd + K;
    double* const s1 = sz + K;
    double* const s2 = s1 + K;
    double2* const srs = reinterpret_cast<double2*>(s2 + K);
    double2* const saq = srs + K;
    double2* const spw = saq + K;
    for (int x = tid; x < K; x += nt) {
        sd[x] = ad[s + x];
        sz[x] = az[s + x];
        s1[x] = rn1[s + x];
        s2[x] = rn2[s + x];
        srs[x] = rs[s + x];
        saq[x] = aq[s + x];
    }
    __syncthreads();
    // lanes per item: the most (up to 16) that leave each item a lane group
    int L = 1;
    while (L < 16 && 2 * L * K <= nt) L <<= 1;
    const int c = tid & (L - 1), per = nt / L;
    const int ra = (int)(((long long)K * c) / L), rb = (int)(((long long)K * (c + 1)) / L);
    const RootsLds roots{srs, sd};
    for (int base = 0; base < K; base += per) {  // the same trip count in every lane
        const int i = min(base + tid / L, K - 1);
        double share = wprod_share(roots, s1, s2, i, sd[i], i & ~(kGroup - 1), ra, rb, K);
        for (int o = 1; o < L; o <<= 1) share *= __shfl_xor(share, o);
        if (c == 0 && base + tid / L < K) spw[i] = make_double2(sd[i], copysign(sqrt(-share), sz[i]));
    }
    __syncthreads();
    const PolesLds poles{spw, saq};
    for (int base = 0; base < K; base += per) {
        const int j = min(base + tid / L, K - 1);
        double temp = 0.0, acc1 = 0.0, acc2 = 0.0;
        rows_share(poles, j, srs[j], s1[j], s2[j], j & ~(kGroup - 1), ra, rb, K, temp, acc1, acc2);
        for (int o = 1; o < L; o <<= 1) {
            temp += __shfl_xor(temp, o);
            acc1 += __shfl_xor(acc1, o);
            acc2 += __shfl_xor(acc2, o);
        }
        if (c == 0 && base + tid / L < K) {
            const double nrm = sqrt(temp);
            o1[j] = acc1 / nrm;
            o2[j] = acc2 / nrm;
        }
    }
    __syncthreads();
}

// Merge step 4 launch (merges of more than ksmall poles): one lane per
// pole i streaming the roots j, WL waves splitting the root stream (LDS
// reduce).
template <int WL>
__global__ __launch_bounds__(kGroup * WL) void k_wprod(
    Geo geo, int bpg, int ksmall, const int* __restrict__ counts, const double* __restrict__ ad,
    const double* __restrict__ az, const double2* __restrict__ rs,
    const double* __restrict__ rn1, const double* __restrict__ rn2, double2* __restrict__ pw) {
    __shared__ double red[WL > 1 ? WL : 1][kGroup];
    const int merge = blockIdx.x / bpg, g = blockIdx.x % bpg;
    const int K = uniform(counts[kCountStride * merge]);
    const int g0 = g * kGroup;
    if (K <= ksmall || g0 >= K) return;
    const Item it = geo.item(merge);
    if (it.flags & kItemRoot) return;  // a root merge keeps no boundary rows
    const int s = it.s;
    const int lane = threadIdx.x & (kGroup - 1);
    const int c = uniform(threadIdx.x / kGroup);
    const int W = waves_for(K, WL);
    if (W == 1 && c > 0) return;
    const int j0 = (c < W) ? (int)(((long long)K * c) / W) : K;
    const int j1 = (c < W) ? (int)(((long long)K * (c + 1)) / W) : K;
    const int i = min(g0 + lane, K - 1);
    const RootsSmem roots{(cdbl) reinterpret_cast<const double*>(rs + s), (cdbl)(ad + s)};
    const double di = ad[s + i];
    double prod = wprod_share(roots, rn1 + s, rn2 + s, i, di, g0, j0, j1, K);  // this wave's share
    if (W > 1) {
        red[c][lane] = prod;
        __syncthreads();
        prod = 1.0;
        for (int cc = 0; cc < W; cc++) prod *= red[cc][lane];
    }
    if (c == 0 && g0 + lane < K) pw[s + i] = make_double2(di, copysign(sqrt(-prod), az[s + i]));
}

// Merge step 5 launch (merges of more than ksmall poles): one lane per root j
// streaming the poles i, x_i = w_i / delta_j(i), row_r(j) = sum_i q_r(i) x_i
// / ||x||, WL waves splitting the pole stream (LDS reduce).
template <int WL>
__global__ __launch_bounds__(kGroup * WL) void k_rows(
    Geo geo, int bpg, int ksmall, const int* __restrict__ counts, const double2* __restrict__ aq,
    const double2* __restrict__ pw, const double2* __restrict__ rs,
    const double* __restrict__ rn1, const double* __restrict__ rn2, double* __restrict__ qo1,
    double* __restrict__ qo2) {
    __shared__ double red[WL > 1 ? WL : 1][3][kGroup];
    const int merge = blockIdx.x / bpg, g = blockIdx.x % bpg;
    const int K = uniform(counts[kCountStride * merge]);
    const int g0 = g * kGroup;
    if (K <= ksmall || g0 >= K) return;
    const Item it = geo.item(merge);
    if (it.flags & kItemRoot) return;  // a root merge keeps no boundary rows
    const int s = it.s;
    const int lane = threadIdx.x & (kGroup - 1);
    const int c = uniform(threadIdx.x / kGroup);
    const int W = waves_for(K, WL);
    if (W == 1 && c > 0) return;
    const int i0 = (c < W) ? (int)(((long long)K * c) / W) : K;
    const int i1 = (c < W) ? (int)(((long long)K * (c + 1)) / W) : K;
    const int j = min(g0 + lane, K - 1);
    const PolesSmem poles{(cdbl) reinterpret_cast<const double*>(pw + s),
                          (cdbl) reinterpret_cast<const double*>(aq + s)};
    double temp = 0.0, acc1 = 0.0, acc2 = 0.0;
    rows_share(poles, j, rs[s + j], rn1[s + j], rn2[s + j], g0, i0, i1, K, temp, acc1, acc2);
    if (W > 1) {
        red[c][0][lane] = temp;
        red[c][1][lane] = acc1;
        red[c][2][lane] = acc2;
        __syncthreads();
        temp = acc1 = acc2 = 0.0;
        for (int cc = 0; cc < W; cc++) {
            temp += red[cc][0][lane];
            acc1 += red[cc][1][lane];
            acc2 += red[cc][2][lane];
        }
    }
    if (c == 0 && g0 + lane < K) {
        const double nrm = sqrt(temp);
        qo1[s + j] = acc1 / nrm;
        qo2[s + j] = acc2 / nrm;
    }
}

// Element idx of a merge at rows from s with K roots, G Givens-deflated and
// T tiny-deflated values: its rank in the merged order and its entries in the
// node's state (or, for the root merge of a matrix, the unscaled eigenvalue).
// qo1/qo2: the roots' boundary rows, root i at [i].
__device__ __forceinline__ void finalize_element(
    int s, int idx, int K, int G, int T, bool root, double unscale,
    const double* __restrict__ rlam, const double* qo1, const double* qo2, const double* __restrict__ gd, const double* __restrict__ gq1,
    const double* __restrict__ gq2, const double* __restrict__ td, const double* __restrict__ tq1,
    const double* __restrict__ tq2, double* __restrict__ out_final, double* __restrict__ od,
    double* __restrict__ olo, double* __restrict__ ohi) {
    double x, q1, q2;
    int rank;
    if (idx < K) {
        x = rlam[s + idx];
        rank = idx + count_less(gd + s, G, x) + count_less(td + s, T, x);
        q1 = root ? 0.0 : qo1[idx];
        q2 = root ? 0.0 : qo2[idx];
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
        out_final[s + rank] = x * unscale;
    } else {
        od[s + rank] = x;
        olo[s + rank] = q1;
        ohi[s + rank] = q2;
    }
}

// Merge step 6: the node's ascending order (DLAMRG of roots with deflated
// values, dlaed7_br.f:146, and the boundary-row persist, 155-160) as a 3-way
// merge of three sorted lists: roots, Givens-deflated, tiny-deflated.  With
// kSmall, a non-root merge of at most ksmall poles has its product form and
// boundary rows (merge steps 4 and 5) computed first by its first block,
// whose elements include every root (dynamic LDS: small_merge_lds of the
// level's largest such merge).
template <bool kSmall>
__global__ __launch_bounds__(kThreads) void k_finalize(
    Geo g, int bpm, int ksmall, const int* __restrict__ counts, const double* __restrict__ rlam,
    const double* __restrict__ qo1, const double* __restrict__ qo2, const double* __restrict__ gd,
    const double* __restrict__ gq1, const double* __restrict__ gq2, const double* __restrict__ td,
    const double* __restrict__ tq1, const double* __restrict__ tq2, double* __restrict__ out_final,
    double* __restrict__ od, double* __restrict__ olo, double* __restrict__ ohi,
    const double* __restrict__ scal, const int* __restrict__ info, int* info_user,
    const double* __restrict__ ad, const double* __restrict__ az, const double2* __restrict__ rs,
    const double* __restrict__ rn1, const double* __restrict__ rn2,
    const double2* __restrict__ aq) {
    extern __shared__ double fin_small[];
    const int merge = blockIdx.x / bpm, part = blockIdx.x % bpm;
    const Item it = g.item(merge);
    const int s = it.s, sz = it.sz;
    const bool root = (it.flags & kItemRoot) != 0;
    // one matrix: the root merge's last kernel hands INFO to the caller
    if (root && blockIdx.x == 0 && threadIdx.x == 0 && info_user) *info_user = *info;
    const int* rec = counts + kCountStride * merge;
    const int K = uniform(rec[0]);
    const double* q1r = qo1 + s;
    const double* q2r = qo2 + s;
    if constexpr (kSmall) {
        if (!root && part == 0 && K > 0 && K <= ksmall) {  // the same in every thread of the block
            double* const o1 = fin_small + (size_t)(kSmallMergeDoubles - 2) * K;
            double* const o2 = o1 + K;
            small_merge_rows(s, K, kThreads, fin_small, ad, az, rs, rn1, rn2, aq, o1, o2);
            q1r = o1;
            q2r = o2;
        }
    }
    const int idx = part * kThreads + threadIdx.x;
    if (idx >= sz) return;
    finalize_element(s, idx, K, rec[1], rec[2], root, root ? scal[2 * it.mat + 1] : 1.0,
                     rlam, q1r, q2r, gd, gq1, gq2, td, tq1, tq2, out_final, od, olo, ohi);
}

// The same product form and boundary rows for levels whose finalize grid is
// large (many merges: the separate launch keeps the finalize kernel's
// registers few): one block of small_merge_threads per merge of 0 < K <=
// kmax poles, rows to HBM (dynamic LDS: small_merge_lds of the level's
// largest such merge).
__host__ __device__ inline int small_merge_threads(int maxsz, int merges) {
    const int k = maxsz < kSmallMaxK ? maxsz : kSmallMaxK;
    int p = 64;
    while (p < k) p <<= 1;
    const int t = (merges >= 512) ? p : 4 * p;
    return t > 1024 ? 1024 : t;
}
__global__ __launch_bounds__(1024) void k_small_merge(
    Geo geo, int kmax, const int* __restrict__ counts, const double* __restrict__ ad,
    const double* __restrict__ az, const double2* __restrict__ rs,
    const double* __restrict__ rn1, const double* __restrict__ rn2,
    const double2* __restrict__ aq, double* __restrict__ qo1, double* __restrict__ qo2) {
    extern __shared__ double sm_small[];
    const int merge = blockIdx.x;
    const int K = uniform(counts[kCountStride * merge]);
    if (K > kmax || K == 0) return;
    const Item it = geo.item(merge);
    if (it.flags & kItemRoot) return;  // a root merge keeps no boundary rows
    double* const o1 = sm_small + (size_t)(kSmallMergeDoubles - 2) * K;
    double* const o2 = o1 + K;
    small_merge_rows(it.s, K, blockDim.x, sm_small, ad, az, rs, rn1, rn2, aq, o1, o2);
    for (int j = threadIdx.x; j < K; j += blockDim.x) {
        qo1[it.s + j] = o1[j];
        qo2[it.s + j] = o2[j];
    }
}

}  // namespace brdc
