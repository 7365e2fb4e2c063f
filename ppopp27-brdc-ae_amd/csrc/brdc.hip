// brdc.hip -- host driver and C ABI of the MI355X BR values-only D&C solver.
//
// One solve is a fixed, host-sync-free launch sequence on the caller's stream:
//   k_norm, k_prep_e, k_leaf, then per tree level (deepest first)
//   prepare -> deflate -> secular -> [wprod -> rows] -> finalize
// (wprod/rows are skipped at the root merge, whose boundary rows nobody
// reads -- WANTQ = .FALSE. at dlaed0_br.f:182).  The tree shape depends only
// on n and the leaf size, so the sequence is fixed per n and can be captured
// into a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <climits>
#include <cstdlib>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/brdc.h"
#include "brdc_deflate_wide.h"
#include "brdc_grad.h"
#include "brdc_kernels.h"
#include "brdc_subtree.h"

namespace brdc {
namespace {

// Leaf bound when brdc_options.leaf_max is 0: small leaves when the subtree
// kernel solves the bottom of the tree (its extra merge levels cost little
// there, and a leaf's QL chain is quadratic in its rows), else 16.
constexpr int kDefaultLeaf = 16;
constexpr int kSubtreeLeaf = 8;
// merges up to this many poles are staged in LDS (4 waves x 640 x 16 B =
// 40 KB per block, so the LDS admits the 3 blocks per CU that the secular
// kernel's register budget allows)
constexpr int kSecStageK = 640;
constexpr int kSecWaves = 4;       // waves per block of the secular kernel (independent waves)
// levels whose merges are all at most this large, and at least this many,
// run the packed secular mode (roots of several merges per wave); below
// that count the per-merge groups with LDS-staged poles are faster
constexpr int kPackMaxSize = 256;
constexpr int kPackMinMerges = 4096;
constexpr int kSecMaxWaves = 3072;  // secular grid cap: about 3 resident waves per SIMD
constexpr int kStreamWaves = 8;     // waves per workgroup of the streamed secular kernel
constexpr int kStreamMinK = 2048;   // merges with at least this many poles are streamed

// Launch-shape knobs (brdc_options; 0 there selects these defaults): the
// streaming threshold, the secular lanes in flight, and the widest group of
// the product-form and boundary-row kernels.
constexpr int kSubMaxLog2 = 10;      // subtree launches of at most 2^10 blocks by default
struct Tuning {
    int stream_min_k = kStreamMinK;
    int target_lanes = kTargetLanes;
    int wide_waves = kMaxWaves;
    long long sub_max = 1LL << kSubMaxLog2;  // most subtree blocks (0: never)
};
constexpr int kMidMaxGroups = 1024; // midpoint-pass grid cap (tiles of 64 roots)
// finalize grids of at most this many blocks compute the small merges' rows
// themselves (about one block per CU)
constexpr long long kFinFuseMaxBlocks = 256;

// Workgroups of the streamed secular kernel that fit the device at once
// (CUs x resident groups per CU), cached per device.
int stream_groups() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (cached[dev] > 0) return cached[dev];
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_secular_stream<kStreamWaves>,
                                                     kWave * kStreamWaves, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    cached[dev] = cus * per_cu;
    return cached[dev];
}
constexpr size_t kAlign = 256;
constexpr int kMergeStatFields = BRDC_MERGE_STAT_FIELDS;

inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Workspace carve-up; identical for the query and the run.  Per-row slots
// cover `rows` rows (one matrix, or a batch's concatenated rows); the
// per-matrix, plan and count records follow.
struct Layout {
    size_t scal, npart, ew;
    size_t sd[2], slo[2], shi[2];
    size_t md, mz, mq1, mq2;
    size_t ad, az, adz, aq;
    size_t gd, gq1, gq2;
    size_t td, tq1, tq2;
    size_t rlam, rs, rn1, rn2;
    size_t pw, qo1, qo2;
    size_t P, rend, fl;
    size_t wrec, wtol;  // chunk records and tolerances of the wide deflation
    size_t counts, info;
    size_t plan, moff, tpre, poff, cpof, gitm, srec;  // batch only
    size_t total;
};

Layout make_layout(int rows, int nmat, size_t merge_records, size_t plan_items,
                   size_t grad_items = 0, size_t subtrees = 0) {
    Layout L{};
    const size_t nd = (size_t)std::max(rows, 1) * sizeof(double);
    const size_t ni = (size_t)std::max(rows, 1) * sizeof(int);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = align_up(off + bytes);
        return at;
    };
    L.scal = take(2 * (size_t)std::max(nmat, 1) * sizeof(double));
    L.npart = take(2 * (size_t)kNormMaxParts * sizeof(double));
    L.ew = take(nd);
    for (int b = 0; b < 2; b++) {
        L.sd[b] = take(nd);
        L.slo[b] = take(nd);
        L.shi[b] = take(nd);
    }
    L.md = take(nd);
    L.mz = take(nd);
    L.mq1 = take(nd);
    L.mq2 = take(nd);
    L.ad = take(nd);
    L.az = take(nd);
    L.adz = take(2 * nd);
    L.aq = take(2 * nd);
    L.gd = take(nd);
    L.gq1 = take(nd);
    L.gq2 = take(nd);
    L.td = take(nd);
    L.tq1 = take(nd);
    L.tq2 = take(nd);
    L.rlam = take(nd);
    L.rs = take(2 * nd);
    L.rn1 = take(nd);
    L.rn2 = take(nd);
    L.pw = take(2 * nd);
    L.qo1 = take(nd);
    L.qo2 = take(nd);
    L.P = take(ni);
    L.rend = take(ni);
    L.fl = take(ni);
    // a wide level has at most kWideMaxMerges merges of at most rows rows
    L.wrec = take((size_t)kWideFields * kWideMaxMerges *
                  ((size_t)std::max(rows, 1) / kWideChunk + 2) * sizeof(int));
    L.wtol = take((size_t)kWideMaxMerges * sizeof(double));
    L.counts = take((size_t)kCountStride * std::max<size_t>(merge_records, 1) * sizeof(int));
    L.info = take(sizeof(int));
    L.plan = plan_items ? take(plan_items * sizeof(int4)) : 0;
    L.moff = plan_items ? take(((size_t)nmat + 1) * sizeof(int)) : 0;
    L.tpre = plan_items ? take(((size_t)nmat + 1) * sizeof(int)) : 0;
    L.poff = plan_items ? take(((size_t)nmat + 1) * sizeof(long long)) : 0;
    L.cpof = plan_items ? take(((size_t)nmat + 1) * sizeof(long long)) : 0;
    L.gitm = plan_items ? take(std::max<size_t>(grad_items, 1) * sizeof(int4)) : 0;
    L.srec = plan_items ? take(std::max<size_t>(subtrees, 1) * kSubMaxLevels * sizeof(int)) : 0;
    L.total = off;
    return L;
}

// Depth of the subtree roots of a matrix of order n with leaves at depth
// `levels`: the shallowest depth whose nodes have at most kSubRows rows, or
// -1 when no merge fits (a subtree holds at least one merge and at most
// kSubMaxLevels merge levels, which leaves of two or more rows guarantee).
inline int subtree_depth(int n, int levels) {
    for (int depth = 0; depth < levels; depth++)
        if (max_node_size(n, depth) <= kSubRows)
            return (levels - depth <= kSubMaxLevels) ? depth : -1;
    return -1;
}

// One matrix of order n: merge records at index (2^depth - 1) + i.
Layout tree_layout(int n, int leaf) {
    const int levels = (n > 0) ? tree_levels(n, leaf) : 0;
    return make_layout(n, 1, (size_t)1 << levels, 0);
}

Tuning tuning_of(const brdc_options* o) {
    Tuning t;
    if (!o) return t;
    if (o->stream_min_k) t.stream_min_k = o->stream_min_k;
    if (o->secular_lanes_log2) t.target_lanes = 1 << o->secular_lanes_log2;
    if (o->wide_max_waves) t.wide_waves = o->wide_max_waves;
    if (o->subtree_max_log2) t.sub_max = (o->subtree_max_log2 < 0) ? 0 : 1LL << o->subtree_max_log2;
    return t;
}

// Subtree roots of an order-n tree with leaves of at most `leaf` rows (0
// when no merge fits a subtree).
long long subtree_roots(int n, int leaf) {
    if (n <= 1) return 0;
    const int depth = subtree_depth(n, tree_levels(n, leaf));
    return depth >= 0 ? 1LL << depth : 0;
}

// The leaf bound of a solve over matrices of the given orders (one matrix,
// or a batch): the option if set, else kSubtreeLeaf when the subtree kernel
// takes the bottom of every tree at that bound, else kDefaultLeaf.
int leaf_of(const brdc_options* o, const int* sizes, int nmat) {
    if (o && o->leaf_max != 0) return o->leaf_max;
    long long roots = 0;
    for (int b = 0; b < nmat; b++) roots += subtree_roots(sizes[b], kSubtreeLeaf);
    return (roots > 0 && roots <= tuning_of(o).sub_max) ? kSubtreeLeaf : kDefaultLeaf;
}
int leaf_of(const brdc_options* o, int n) { return leaf_of(o, &n, 1); }

bool options_ok(const brdc_options* o) {
    if (!o) return true;
    if (o->leaf_max != 0 && (o->leaf_max < 2 || o->leaf_max > kMaxLeaf)) return false;
    if (o->stream_min_k != 0 && (o->stream_min_k < 256 || o->stream_min_k > (1 << 28)))
        return false;
    if (o->secular_lanes_log2 != 0 && (o->secular_lanes_log2 < 12 || o->secular_lanes_log2 > 24))
        return false;
    const int w = o->wide_max_waves;
    if (w != 0 && w != 1 && w != 2 && w != 4 && w != 8) return false;
    if (o->subtree_max_log2 < -1 || o->subtree_max_log2 > 30) return false;
    for (int r : o->reserved)
        if (r != 0) return false;
    return true;
}

template <class T>
T* at(void* ws, size_t off) {
    return reinterpret_cast<T*>(static_cast<char*>(ws) + off);
}

#define BRDC_LAUNCH_CHECK()                                        \
    do {                                                           \
        if (hipGetLastError() != hipSuccess) return BRDC_STATUS_HIP_ERROR; \
    } while (0)

// Kernel-class timing (diagnostics, the role of the reference's
// MergeLevelTimer, gpu/src/tridiag_solver.cu:158-359): when enabled on the
// calling host thread, every launch is bracketed by two events on its stream.
enum KernelClass { kNorm, kPrepE, kLeaf, kPrepare, kDeflate, kSecular, kWprod, kRows, kFinalize,
                   kSecularStream, kSecularMid, kSubtree, kNumClasses };
// Each mark also carries the merge level it ran for (tree depth, 0 = root
// merge, or forest height for a batch; -1 before the merges).
struct ProfMark {
    int cls, level;
    hipEvent_t a, b;
};
struct Profiler {
    bool on = false;
    int level = -1;  // level of the launches being enqueued
    std::vector<ProfMark> marks;
};
thread_local Profiler g_prof;

struct ProfScope {
    int cls;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(int c, hipStream_t s) : cls(c), st(s) {
        if (!g_prof.on) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
        (void)hipEventRecord(a, st);
    }
    ~ProfScope() {
        if (!g_prof.on || !a || !b) return;
        (void)hipEventRecord(b, st);
        g_prof.marks.push_back({cls, g_prof.level, a, b});
    }
};

// the O(K^2) kernels come in widths of 1..16 waves per 64-item group
template <class K1, class K2, class K4, class K8, class... Args>
void launch_wide(K1 k1, K2 k2, K4 k4, K8 k8, int WL, int blocks, size_t lds, hipStream_t st,
                 Args... args) {
    const dim3 grid(blocks);
    switch (WL) {
        case 1: hipLaunchKernelGGL(k1, grid, dim3(kGroup), lds, st, args...); break;
        case 2: hipLaunchKernelGGL(k2, grid, dim3(2 * kGroup), lds, st, args...); break;
        case 4: hipLaunchKernelGGL(k4, grid, dim3(4 * kGroup), lds, st, args...); break;
        default: hipLaunchKernelGGL(k8, grid, dim3(8 * kGroup), lds, st, args...); break;
    }
}

// deflation block size follows the level's node size
template <class... Args>
void launch_deflate(int maxsz, int merges, hipStream_t st, Args... args) {
    // each wave walks its share of a merge's positions in 64-wide tiles, one
    // after another: blocks wide enough for about two tiles a wave, as far
    // as the level's merges leave room on the device (2^19 threads)
    int T = (maxsz <= 128) ? 64 : (maxsz <= 512) ? 256 : 1024;
    while (T > 64 && (long long)merges * T > (1LL << 19)) T /= 4;
    if (T == 64)
        hipLaunchKernelGGL(k_deflate<64>, dim3(merges), dim3(64), 0, st, args...);
    else if (T == 256)
        hipLaunchKernelGGL(k_deflate<256>, dim3(merges), dim3(256), 0, st, args...);
    else
        hipLaunchKernelGGL(k_deflate<1024>, dim3(merges), dim3(1024), 0, st, args...);
}

// One merge level of a launch sequence: all merges of one tree depth (one
// matrix) or of one height above the leaves (a batch's merge forest).
struct LevelRun {
    Geo geo;
    int merges;
    int maxsz;
    size_t counts_rec;  // first merge record of the level
    bool any_nonroot;   // some merge keeps boundary rows (wprod/rows run)
    int report;         // level in diagnostics: tree depth, or forest height
};

// A whole solve: rows, matrices, leaves and levels (deepest first).
struct Schedule {
    int rows = 0, nmat = 1;
    Tuning tune{};
    const int* moff = nullptr;  // device: first row of each matrix (batch), else nullptr
    Geo leaves{};
    int nleaves = 0, leaf_rows = 0;
    // the bottom of the tree in k_subtree (nsub > 0): the leaves and every
    // merge below the subtree roots, which `levels` then leaves out
    SubGeo sub{};
    int nsub = 0;
    std::vector<LevelRun> levels;
};

Schedule tree_schedule(int n, int leaf, const Tuning& tune) {
    Schedule S;
    S.rows = n;
    S.tune = tune;
    const int levels = tree_levels(n, leaf);
    S.leaves = Geo{n, levels, nullptr};
    S.nleaves = 1 << levels;
    S.leaf_rows = max_node_size(n, levels);
    const int sub = subtree_depth(n, levels);
    int top = levels;  // merges at depths below this run level by level
    if (sub >= 0 && (1LL << sub) <= tune.sub_max) {
        S.sub = SubGeo{Geo{n, sub, nullptr}, levels, nullptr};
        S.nsub = 1 << sub;
        S.nleaves = 0;
        top = sub;
    }
    for (int depth = top - 1; depth >= 0; depth--) {
        LevelRun r;
        r.geo = Geo{n, depth, nullptr};
        r.merges = 1 << depth;
        r.maxsz = max_node_size(n, depth);
        r.counts_rec = (size_t)r.merges - 1;
        r.any_nonroot = (depth > 0);
        r.report = depth;
        S.levels.push_back(r);
    }
    return S;
}

// Enqueues a solve on st.  info: the INFO word(s) written by the kernels (one
// matrix: the workspace word, copied to info_user by the last kernel; batch:
// the caller's per-matrix array).
int enqueue(const Schedule& S, double* d, const double* e, void* ws, const Layout& L,
            hipStream_t st, int* info, int* info_user) {
    const int n = S.rows;
    double* scal = at<double>(ws, L.scal);
    double* ew = at<double>(ws, L.ew);
    int* counts_all = at<int>(ws, L.counts);

    // one large matrix: its norm over several blocks, finished in k_prep_e
    const int nparts = S.moff ? 0 : std::min(kNormMaxParts, n / kNormRowsPerPart);
    double* part = at<double>(ws, L.npart);
    {
        ProfScope ps(kNorm, st);
        hipLaunchKernelGGL(k_norm, dim3(nparts > 1 ? nparts : S.nmat), dim3(1024), 0, st, n,
                           S.moff, d, e, scal, info, info_user, part);
    }
    BRDC_LAUNCH_CHECK();
    if (n <= 1) return BRDC_STATUS_SUCCESS;
    {
        ProfScope ps(kPrepE, st);
        hipLaunchKernelGGL(k_prep_e, dim3((n + 255) / 256), dim3(256), 0, st, n, S.moff, S.nmat,
                           d, e, scal, ew, nparts > 1 ? nparts : 0, (const double*)part, info);
    }
    BRDC_LAUNCH_CHECK();

    int cur = 0;
    double* sd[2] = {at<double>(ws, L.sd[0]), at<double>(ws, L.sd[1])};
    double* slo[2] = {at<double>(ws, L.slo[0]), at<double>(ws, L.slo[1])};
    double* shi[2] = {at<double>(ws, L.shi[0]), at<double>(ws, L.shi[1])};
    if (S.nsub > 0) {
        ProfScope ps(kSubtree, st);
        hipLaunchKernelGGL(k_subtree, dim3(S.nsub), dim3(kSubThreads), 0, st, S.sub, S.moff,
                           (const double*)d, (const double*)ew, (const double*)scal, d, sd[cur],
                           slo[cur], shi[cur], counts_all, info,
                           S.levels.empty() ? info_user : nullptr);
        BRDC_LAUNCH_CHECK();
    }
    if (S.nleaves > 0) {
        ProfScope ps(kLeaf, st);
        // one lane per leaf, LDS sized by the largest leaf
        const int maxm = std::max(S.leaf_rows, 1);
        const int lanes = leaf_lanes(maxm);
        hipLaunchKernelGGL(lanes == kWave ? k_leaf<kWave> : k_leaf<kWave / 2>,
                           dim3((S.nleaves + lanes - 1) / lanes), dim3(lanes),
                           leaf_lds_bytes(maxm), st, S.leaves, S.nleaves, maxm, S.moff,
                           (const double*)d, (const double*)ew, (const double*)scal, d, sd[cur],
                           slo[cur], shi[cur], info, S.levels.empty() ? info_user : nullptr);
    }
    BRDC_LAUNCH_CHECK();
    if (S.levels.empty()) return BRDC_STATUS_SUCCESS;

    double* md = at<double>(ws, L.md);
    double* mz = at<double>(ws, L.mz);
    double* mq1 = at<double>(ws, L.mq1);
    double* mq2 = at<double>(ws, L.mq2);
    double* ad = at<double>(ws, L.ad);
    double* az = at<double>(ws, L.az);
    double2* adz = at<double2>(ws, L.adz);
    double2* aq = at<double2>(ws, L.aq);
    double* gd = at<double>(ws, L.gd);
    double* gq1 = at<double>(ws, L.gq1);
    double* gq2 = at<double>(ws, L.gq2);
    double* td = at<double>(ws, L.td);
    double* tq1 = at<double>(ws, L.tq1);
    double* tq2 = at<double>(ws, L.tq2);
    double* rlam = at<double>(ws, L.rlam);
    double2* rs = at<double2>(ws, L.rs);
    double* rn1 = at<double>(ws, L.rn1);
    double* rn2 = at<double>(ws, L.rn2);
    double2* pw = at<double2>(ws, L.pw);
    double* qo1 = at<double>(ws, L.qo1);
    double* qo2 = at<double>(ws, L.qo2);
    int* P = at<int>(ws, L.P);
    int* rend = at<int>(ws, L.rend);
    int* fl = at<int>(ws, L.fl);

    struct LevelReset {
        ~LevelReset() { g_prof.level = -1; }
    } level_reset;
    for (const LevelRun& lv : S.levels) {
        g_prof.level = lv.report;
        const Geo g = lv.geo;
        const int merges = lv.merges;
        const int maxsz = lv.maxsz;
        const int bpm = (maxsz + kThreads - 1) / kThreads;
        const dim3 grid(merges * bpm);
        // the O(K^2) kernels: one block per 64 roots/poles, WL waves wide
        const int bpg = (maxsz + kGroup - 1) / kGroup;
        const int WL = waves_for(maxsz, S.tune.wide_waves);
        int* counts = counts_all + kCountStride * lv.counts_rec;
        const int nxt = cur ^ 1;

        {
            ProfScope ps(kPrepare, st);
            hipLaunchKernelGGL(k_merge_prepare, grid, dim3(kThreads), 0, st, g, bpm, sd[cur],
                               slo[cur], shi[cur], ew, md, mz, mq1, mq2);
        }
        BRDC_LAUNCH_CHECK();
        if (merges <= kWideMaxMerges && maxsz >= kWideMinSize) {
            // few huge merges: the deflation runs chunked over many blocks
            ProfScope ps(kDeflate, st);
            const int chunks = (maxsz + kWideChunk - 1) / kWideChunk;
            const WideGeo wg{g, chunks, at<int>(ws, L.wrec), at<double>(ws, L.wtol)};
            const dim3 grid(merges * chunks), blk(kWideThreads);
            hipLaunchKernelGGL(k_wide_flags, grid, blk, 0, st, wg, (const double*)ew,
                               (const double*)md, (const double*)mz, fl);
            hipLaunchKernelGGL(k_wide_compact, grid, blk, 0, st, wg, (const int*)fl, P);
            hipLaunchKernelGGL(k_wide_spec, grid, blk, 0, st, wg, (const int*)P, (const double*)md,
                               (const double*)mz, rend);
            hipLaunchKernelGGL(k_wide_resolve, dim3(merges), blk, 0, st, wg, (const int*)P,
                               (const double*)md, (const double*)mz, rend, counts);
            hipLaunchKernelGGL(k_wide_replay, grid, blk, 0, st, wg, (const int*)P,
                               (const int*)rend, md, mz, mq1, mq2, fl, gd, gq1, gq2);
            hipLaunchKernelGGL(k_wide_count, grid, blk, 0, st, wg, (const int*)fl);
            hipLaunchKernelGGL(k_wide_scatter, grid, blk, 0, st, wg, (const int*)fl,
                               (const double*)md, (const double*)mz, (const double*)mq1,
                               (const double*)mq2, ad, az, adz, aq, td, tq1, tq2, counts);
        } else {
            ProfScope ps(kDeflate, st);
            launch_deflate(maxsz, merges, st, g, ew, md, mz, mq1, mq2, P, rend, fl, ad, az, adz,
                           aq, gd, gq1, gq2, td, tq1, tq2, counts);
        }
        BRDC_LAUNCH_CHECK();
        // merges of K >= kstream poles go to the streamed kernel, the
        // rest to the lanes-per-root kernel
        const int kstream = S.tune.stream_min_k, target = S.tune.target_lanes;
        const bool streamed = (maxsz >= kstream);
        if (maxsz <= kPackMaxSize && merges >= kPackMinMerges) {
            // many small merges: roots packed across merges (rend is free
            // between deflations and holds the root offsets)
            ProfScope ps(kSecular, st);
            hipLaunchKernelGGL(k_root_scan, dim3(1), dim3(kScanThreads), 0, st, merges,
                               streamed ? kstream : INT_MAX, (const int*)counts, rend);
            BRDC_LAUNCH_CHECK();
            const int blocks = kSecMaxWaves / kSecWaves;
            hipLaunchKernelGGL(k_secular_packed<kSecWaves>, dim3(blocks), dim3(kWave * kSecWaves),
                               0, st, g, merges, target, (const int*)rend, (const double*)ew,
                               (const int*)counts, (const double*)ad, (const double*)az,
                               (const double2*)adz, rlam, rs, rn1, rn2, counts, S.moff, info);
        } else {
            ProfScope ps(kSecular, st);
            // wave slots per merge: the most any K it handles can need; the
            // grid is capped and walks (merge, slot) tasks
            const int slots = secular_slots(std::min(maxsz, kstream - 1), merges, target);
            const int waves = std::min(merges * slots, kSecMaxWaves);
            const int blocks = (waves + kSecWaves - 1) / kSecWaves;
            // LDS pole stage per wave: the level's largest merge, capped
            const int stage_k = std::min(maxsz, kSecStageK);
            const size_t lds = sizeof(double2) * kSecWaves * stage_k;
            hipLaunchKernelGGL(k_secular<kSecWaves>, dim3(blocks), dim3(kWave * kSecWaves), lds,
                               st, g, merges, slots, streamed ? kstream : INT_MAX, stage_k, target,
                               (const double*)ew, (const int*)counts, (const double*)ad,
                               (const double*)az, (const double2*)adz, rlam, rs, rn1, rn2, counts,
                               S.moff, info);
        }
        BRDC_LAUNCH_CHECK();
        if (streamed) {
            {
                ProfScope ps(kSecularMid, st);
                hipLaunchKernelGGL(k_secular_mid<kStreamWaves>,
                                   dim3(std::min(kMidMaxGroups, (n + kWave - 1) / kWave)),
                                   dim3(kWave * kStreamWaves), 0, st, g, merges, kstream,
                                   (const double*)ew, (const int*)counts, (const double2*)adz,
                                   rlam);
            }
            BRDC_LAUNCH_CHECK();
            ProfScope ps(kSecularStream, st);
            const int groups = std::min(stream_groups(), (n + kWave - 1) / kWave);
            hipLaunchKernelGGL(k_secular_stream<kStreamWaves>, dim3(groups),
                               dim3(kWave * kStreamWaves), 0, st, g, merges, kstream,
                               (const double*)ew, (const int*)counts, (const double2*)adz, rlam,
                               rs, rn1, rn2, counts, S.moff, info);
            BRDC_LAUNCH_CHECK();
        }
        // merges of at most ksmall poles get their product form and rows in
        // the finalize launch; the larger ones here
        const int ksmall = small_max_k(maxsz);
        if (lv.any_nonroot && maxsz > kSmallMaxK) {
            {
                ProfScope ps(kWprod, st);
                launch_wide(k_wprod<1>, k_wprod<2>, k_wprod<4>, k_wprod<8>, WL, merges * bpg, 0,
                            st, g, bpg, ksmall, (const int*)counts, (const double*)ad,
                            (const double*)az, (const double2*)rs, (const double*)rn1,
                            (const double*)rn2, pw);
            }
            BRDC_LAUNCH_CHECK();
            {
                ProfScope ps(kRows, st);
                launch_wide(k_rows<1>, k_rows<2>, k_rows<4>, k_rows<8>, WL, merges * bpg, 0, st,
                            g, bpg, ksmall, (const int*)counts, (const double2*)aq,
                            (const double2*)pw, (const double2*)rs, (const double*)rn1,
                            (const double*)rn2, qo1, qo2);
            }
            BRDC_LAUNCH_CHECK();
        }
        // the small merges' rows inside the finalize launch while its grid is
        // small (a single matrix); a launch of their own before a many-block
        // finalize (a batch), where the fused kernel's registers would cost
        // the finalize blocks their occupancy
        const bool fused = lv.any_nonroot && (long long)merges * bpm <= kFinFuseMaxBlocks;
        const size_t small_lds = small_merge_lds(std::min(maxsz, ksmall));
        if (lv.any_nonroot && !fused) {
            ProfScope ps(kFinalize, st);
            hipLaunchKernelGGL(k_small_merge, dim3(merges),
                               dim3(small_merge_threads(maxsz, merges)), small_lds, st, g,
                               ksmall, (const int*)counts, (const double*)ad, (const double*)az,
                               (const double2*)rs, (const double*)rn1, (const double*)rn2,
                               (const double2*)aq, qo1, qo2);
            BRDC_LAUNCH_CHECK();
        }
        {
            ProfScope ps(kFinalize, st);
            hipLaunchKernelGGL(fused ? k_finalize<true> : k_finalize<false>, grid, dim3(kThreads),
                               fused ? small_lds : 0, st, g, bpm, fused ? ksmall : 0, counts, rlam,
                               qo1, qo2, gd, gq1, gq2, td, tq1, tq2, d, sd[nxt], slo[nxt],
                               shi[nxt], (const double*)scal, (const int*)info, info_user,
                               (const double*)ad, (const double*)az, (const double2*)rs,
                               (const double*)rn1, (const double*)rn2, (const double2*)aq);
        }
        BRDC_LAUNCH_CHECK();
        cur = nxt;
    }
    return BRDC_STATUS_SUCCESS;
}

int enqueue_solve(int n, double* d, const double* e, void* ws, const Layout& L, hipStream_t st,
                  int* info_user, int leaf, const Tuning& tune) {
    return enqueue(tree_schedule(n, leaf, tune), d, e, ws, L, st, at<int>(ws, L.info), info_user);
}

// ---------------------------------------------------------------------------
// Batch plan: the flattened merge forest of a batch of matrices.  Each
// matrix has its own tree (same leaf bound); level h of the forest holds
// every matrix's merges h levels above its leaves, so one launch sequence
// per height serves all matrices.  Built once on the host per size list.
struct HostLevel {
    int merges = 0;             // all merges of the height (one record each)
    int launch = 0, maxsz = 0;  // the first `launch` of them run as level launches
    size_t plan_at = 0, counts_rec = 0;
    bool any_nonroot = false;   // among the launched merges
};

}  // namespace
}  // namespace brdc

struct brdc_batch_plan {
    int nmat = 0, rows = 0, leaf = 0;
    brdc::Tuning tune{};
    std::vector<int> moff;           // nmat + 1
    std::vector<int> tprefix;        // gradient tasks (64 eigenvalues) before each matrix
    std::vector<long long> poff;     // gradient partial-sum offset of each matrix
    std::vector<long long> cpoff;    // gradient checkpoint offset of each matrix
    std::vector<int4> gitems;        // gradient items of the four launches, back to back
    int gcount[4] = {0, 0, 0, 0};    // items per launch
    int rows_max = 0;
    std::vector<int4> items;         // leaves, every level's merges, subtree roots
    int nleaves = 0, leaf_rows = 0;  // leaves of the forest; the largest
    int leaf_items = 0;              // leaves solved by the leaf kernel (first items)
    size_t sub_at = 0;               // first subtree root in items
    int nsub = 0;
    std::vector<int> srec;           // per subtree root kSubMaxLevels record bases
    std::vector<brdc::HostLevel> levels;  // height 1 upward
    size_t merge_records = 0;
    brdc::Layout layout;
};

namespace brdc {
namespace {

// Gradient scratch: per task the checkpoint arrays of (segments + 1) x 64
// doubles (grad_cp_slot); per task 64 twist indices and two |z|^2 halves; per gamma item 64
// best |gamma| values and rows; then the per-task partial sums.
struct GradShape {
    size_t cp, rarr, ndarr, nuarr, garr, rgarr, partial, gm, tflag, clus, total;
};

// checkpoint doubles of one task of an n-row matrix
inline long long grad_slot(int n) { return grad_cp_slot(n); }

// The backward's scratch grows as n^2 / 8 doubles (checkpoints of every
// task's pivot chains); above this many bytes (128 GiB, about n = 170000 for
// one matrix) the queries and entries refuse with BRDC_STATUS_INVALID_VALUE.
constexpr size_t kGradScratchMax = (size_t)1 << 37;

// rows: all rows of the call; nmax: its largest matrix (the cluster kernel's
// per-wave arrays)
GradShape grad_shape(int ntask, long long cp_doubles, long long partial_doubles, int ngamma,
                     int rows, int nmax) {
    GradShape G{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = align_up(off + bytes);
        return at;
    };
    const size_t lanes = (size_t)std::max(ntask, 1) * kWave;
    const size_t glanes = (size_t)std::max(ngamma, 1) * kWave;
    G.cp = take((size_t)std::max<long long>(cp_doubles, 1) * sizeof(double));
    G.rarr = take(lanes * sizeof(int));
    G.ndarr = take(lanes * sizeof(double));
    G.nuarr = take(lanes * sizeof(double));
    G.garr = take(glanes * sizeof(double));
    G.rgarr = take(glanes * sizeof(int));
    G.partial = take((size_t)std::max<long long>(partial_doubles, 1) * sizeof(double));
    G.gm = take((size_t)std::max(rows, 1) * sizeof(double));
    G.tflag = take((size_t)std::max(ntask, 1) * sizeof(int));
    G.clus = take((size_t)kClusterWaves * kClusterArrays * (size_t)std::max(nmax, 1) *
                  sizeof(double));
    G.total = off;
    return G;
}

// The four launches of the eigenvector kernel (chains and short matrices;
// twist search; z; contributions), then the per-matrix sums.  items: the
// four item lists back to back (count[l] each), or nullptr for one matrix,
// whose items follow from the task count and parts.
// Bound on a segment transfer's partial products of z (kScanBig);
// BRDC_GRAD_SCAN_LIMIT lowers it, and 0 sends every long-matrix z scan
// through the row walk, so the two forms can be checked against each other.
double grad_scan_limit() {
    const char* s = std::getenv("BRDC_GRAD_SCAN_LIMIT");
    if (!s || !*s) return kScanBig;
    char* end = nullptr;
    const double v = std::strtod(s, &end);
    return (end != s && v >= 0.0 && v <= kScanBig) ? v : kScanBig;
}

int enqueue_grad(int rows, int nmat, int nmax, const int* moff, const int* tprefix,
                 const long long* poff, const long long* cpoff, const int4* items,
                 const int count[4], int parts, int ntask, long long cp_doubles,
                 long long partial_doubles, const double* d, const double* e, const double* w,
                 const double* g, double* dd, double* de, void* scratch, hipStream_t st) {
    const GradShape G = grad_shape(ntask, cp_doubles, partial_doubles, count[1], rows, nmax);
    double* cp = at<double>(scratch, G.cp);
    int* rarr = at<int>(scratch, G.rarr);
    double* ndarr = at<double>(scratch, G.ndarr);
    double* nuarr = at<double>(scratch, G.nuarr);
    double* garr = at<double>(scratch, G.garr);
    int* rgarr = at<int>(scratch, G.rgarr);
    double* part = at<double>(scratch, G.partial);
    double* gm = at<double>(scratch, G.gm);
    const int n_one = moff ? 0 : rows;
    if (ntask <= 0) return BRDC_STATUS_SUCCESS;
    const double scan_big = grad_scan_limit();
    // members of numerically degenerate clusters leave the twisted
    // factorization (weight 0 there) for the cluster kernel
    int* tflag = at<int>(scratch, G.tflag);
    hipLaunchKernelGGL(k_grad_mask, dim3((ntask + 3) / 4), dim3(4 * kWave), 0, st, ntask, nmat,
                       moff, tprefix, n_one, w, g, gm, tflag);
    BRDC_LAUNCH_CHECK();
    size_t base = 0;
    for (int l = 0; l < 4; l++) {
        const int cnt = count[l];
        if (cnt > 0) {  // a wave per item, up to a few per SIMD
            const int waves = std::min(cnt, 4096);
            hipLaunchKernelGGL(k_grad_vectors, dim3((waves + kGradWaves - 1) / kGradWaves),
                               dim3(kWave * kGradWaves), 0, st, l + 1,
                               items ? items + base : (const int4*)nullptr, cnt, parts, nmat, moff,
                               tprefix, poff, cpoff, n_one, d, e, w, (const double*)gm, cp, rarr,
                               ndarr, nuarr, garr, rgarr, part, scan_big);
            BRDC_LAUNCH_CHECK();
        }
        base += (size_t)std::max(cnt, 0);
    }
    hipLaunchKernelGGL(k_grad_clusters,
                       dim3(std::min((ntask + kWave - 1) / kWave, kClusterWaves)), dim3(kWave), 0,
                       st, ntask, nmat, moff, tprefix, poff, n_one, nmax, d, e, w, g,
                       (const int*)tflag, at<double>(scratch, G.clus), part);
    BRDC_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_grad_reduce, dim3((rows + kReduceRows - 1) / kReduceRows),
                       dim3(kReduceRows * kReduceSlices), 0, st, rows, nmat, moff, tprefix, poff,
                       n_one, (const double*)part, dd, de);
    BRDC_LAUNCH_CHECK();
    return BRDC_STATUS_SUCCESS;
}

// item counts of one matrix of order n (tasks of 64 eigenvalues)
inline void grad_counts_one(int n, int count[4]) {
    const int tasks = (n + kWave - 1) / kWave, parts = grad_parts(n);
    const bool split = parts > 1;
    count[0] = split ? 2 * tasks : tasks;
    count[1] = split ? tasks * parts : 0;
    count[2] = split ? 2 * tasks : 0;
    count[3] = split ? tasks * parts : 0;
}

Schedule batch_schedule(const brdc_batch_plan& p, void* ws) {
    Schedule S;
    S.rows = p.rows;
    S.nmat = p.nmat;
    S.tune = p.tune;
    S.moff = at<int>(ws, p.layout.moff);
    const int4* plan = at<int4>(ws, p.layout.plan);
    S.leaves = Geo{p.rows, -1, plan};
    S.nleaves = p.leaf_items;
    S.leaf_rows = p.leaf_rows;
    if (p.nsub > 0) {
        S.sub = SubGeo{Geo{p.rows, -1, plan + p.sub_at}, 0, at<int>(ws, p.layout.srec)};
        S.nsub = p.nsub;
    }
    for (size_t h = 0; h < p.levels.size(); h++) {
        const HostLevel& hl = p.levels[h];
        if (hl.launch == 0) continue;  // every merge of the height is inside a subtree
        LevelRun r;
        r.geo = Geo{p.rows, -1, plan + hl.plan_at};
        r.merges = hl.launch;
        r.maxsz = hl.maxsz;
        r.counts_rec = hl.counts_rec;
        r.any_nonroot = hl.any_nonroot;
        r.report = (int)h + 1;
        S.levels.push_back(r);
    }
    return S;
}

}  // namespace
}  // namespace brdc

using namespace brdc;

extern "C" {

const char* brdc_status_string(int status) {
    switch (status) {
        case BRDC_STATUS_SUCCESS: return "success";
        case BRDC_STATUS_INVALID_POINTER: return "invalid pointer";
        case BRDC_STATUS_INVALID_VALUE: return "invalid value";
        case BRDC_STATUS_INVALID_WORKSPACE: return "workspace too small or misaligned";
        case BRDC_STATUS_ALLOC_FAILED: return "device allocation failed";
        case BRDC_STATUS_HIP_ERROR: return "HIP runtime error";
        case BRDC_STATUS_INTERNAL_ERROR: return "internal error";
        default: return "unknown status";
    }
}

int brdc_dstebr_device_workspace_ex(int n, const brdc_options* opts, size_t* bytes) {
    if (!bytes) return BRDC_STATUS_INVALID_POINTER;
    if (n < 0 || !options_ok(opts)) return BRDC_STATUS_INVALID_VALUE;
    *bytes = tree_layout(n, leaf_of(opts, n)).total;
    return BRDC_STATUS_SUCCESS;
}

int brdc_dstebr_device_workspace(int n, size_t* bytes) {
    return brdc_dstebr_device_workspace_ex(n, nullptr, bytes);
}

int brdc_dstebr_device_ex(int n, double* d_dev, const double* e_dev, void* workspace,
                          size_t workspace_bytes, void* stream, int* info_dev,
                          const brdc_options* opts) {
    if (n < 0 || !options_ok(opts)) return BRDC_STATUS_INVALID_VALUE;
    if (n == 0) return BRDC_STATUS_SUCCESS;
    if (!d_dev || (n > 1 && !e_dev) || !workspace) return BRDC_STATUS_INVALID_POINTER;
    const int leaf = leaf_of(opts, n);
    const Layout L = tree_layout(n, leaf);
    if (workspace_bytes < L.total || (reinterpret_cast<uintptr_t>(workspace) % kAlign) != 0)
        return BRDC_STATUS_INVALID_WORKSPACE;
    return enqueue_solve(n, d_dev, e_dev, workspace, L, static_cast<hipStream_t>(stream), info_dev,
                         leaf, tuning_of(opts));
}

int brdc_dstebr_device(int n, double* d_dev, const double* e_dev, void* workspace,
                       size_t workspace_bytes, void* stream, int* info_dev) {
    return brdc_dstebr_device_ex(n, d_dev, e_dev, workspace, workspace_bytes, stream, info_dev,
                                 nullptr);
}

int brdc_merge_stats(int n, const void* workspace, size_t workspace_bytes, void* stream,
                     const brdc_options* opts, int* out_host, int max_merges) {
    if (n < 0 || !options_ok(opts)) return -BRDC_STATUS_INVALID_VALUE;
    if (!workspace || (!out_host && max_merges > 0)) return -BRDC_STATUS_INVALID_POINTER;
    const int leaf = leaf_of(opts, n);
    const Layout L = tree_layout(n, leaf);
    if (workspace_bytes < L.total) return -BRDC_STATUS_INVALID_WORKSPACE;
    const int levels = (n > 0) ? tree_levels(n, leaf) : 0;
    const int total = (n > 1) ? (1 << levels) - 1 : 0;
    if (total == 0) return 0;
    std::vector<int> counts((size_t)kCountStride * total);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (hipMemcpyAsync(counts.data(),
                       static_cast<const char*>(workspace) + L.counts,
                       counts.size() * sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return -BRDC_STATUS_HIP_ERROR;
    int rec = 0;
    for (int depth = 0; depth < levels; depth++) {
        for (int i = 0; i < (1 << depth); i++) {
            if (rec >= max_merges) return total;
            int s, sz;
            node_geom(n, depth, i, &s, &sz);
            const int* c = &counts[(size_t)kCountStride * ((1 << depth) - 1 + i)];
            int* o = out_host + kMergeStatFields * rec;
            o[0] = depth;
            o[1] = s;
            o[2] = sz;
            o[3] = c[0];
            o[4] = c[1];
            o[5] = c[3];
            o[6] = c[2];
            o[7] = c[5];
            o[8] = c[6];
            rec++;
        }
    }
    return total;
}

int brdc_profile_begin(void) {
    for (auto& m : g_prof.marks) {
        (void)hipEventDestroy(m.a);
        (void)hipEventDestroy(m.b);
    }
    g_prof.marks.clear();
    g_prof.on = true;
    return BRDC_STATUS_SUCCESS;
}

int brdc_profile_end_levels(double* ms_out, int* launches_out, int nclasses, int nlevels) {
    g_prof.on = false;
    int status = BRDC_STATUS_SUCCESS;
    if (nclasses < 0 || nlevels < 0) status = BRDC_STATUS_INVALID_VALUE;
    const int nc = std::min(std::max(nclasses, 0), (int)kNumClasses);
    const size_t rows = (size_t)std::max(nlevels, 0) + 1;
    if (ms_out)
        for (size_t x = 0; x < rows * (size_t)std::max(nclasses, 0); x++) ms_out[x] = 0.0;
    if (launches_out)
        for (size_t x = 0; x < rows * (size_t)std::max(nclasses, 0); x++) launches_out[x] = 0;
    for (auto& m : g_prof.marks) {
        float t = 0.0f;
        if (hipEventSynchronize(m.b) != hipSuccess || hipEventElapsedTime(&t, m.a, m.b) != hipSuccess)
            status = BRDC_STATUS_HIP_ERROR;
        // row 0: launches before the merges; row 1 + l: merge level l
        const long row = (long)m.level + 1;
        if (m.cls < nc && row >= 0 && row < (long)rows) {
            const size_t at = (size_t)row * (size_t)nclasses + (size_t)m.cls;
            if (ms_out) ms_out[at] += t;
            if (launches_out) launches_out[at] += 1;
        }
        (void)hipEventDestroy(m.a);
        (void)hipEventDestroy(m.b);
    }
    g_prof.marks.clear();
    return status;
}

int brdc_profile_end(double* ms_out, int* launches_out, int nclasses) {
    // every level folded into one row per class
    const int nc = std::max(nclasses, 0);
    int top = -1;
    for (const auto& m : g_prof.marks) top = std::max(top, m.level);
    const size_t rows = (size_t)(top + 2);
    std::vector<double> ms(rows * (size_t)nc, 0.0);
    std::vector<int> cnt(rows * (size_t)nc, 0);
    const int status = brdc_profile_end_levels(ms.data(), cnt.data(), nclasses, top + 1);
    for (int c = 0; c < nc; c++) {
        double t = 0.0;
        int k = 0;
        for (size_t r = 0; r < rows; r++) {
            t += ms[r * (size_t)nc + (size_t)c];
            k += cnt[r * (size_t)nc + (size_t)c];
        }
        if (ms_out) ms_out[c] = t;
        if (launches_out) launches_out[c] = k;
    }
    return status;
}

int brdc_dstebr(int n, double* d, double* e, double* work, int lwork, int* iwork, int liwork) {
    // argument checks and workspace query as dstedc_dc_br.f:44-75 (NTHR = 1)
    const int smlsiz = 25;
    const bool lquery = (lwork == -1 || liwork == -1);
    if (n < 0) return -1;
    const long lwmin = (n <= smlsiz) ? 1 : std::max<long>(1, 18L * n);
    const long liwmin = (n <= smlsiz) ? 1 : std::max<long>(1, 7L * n);
    if (work) work[0] = (double)lwmin;
    if (iwork) iwork[0] = (int)liwmin;
    if (!lquery) {
        if (lwork < lwmin) return -5;
        if (liwork < liwmin) return -7;
    } else {
        return 0;
    }
    if (n <= 1) return 0;

    const int leaf = leaf_of(nullptr, n);
    const Layout L = tree_layout(n, leaf);
    double *dd = nullptr, *de = nullptr;
    void* ws = nullptr;
    int* dinfo = nullptr;
    int info = 0;
    int result = 0;
    if (hipMalloc(&dd, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&de, sizeof(double) * std::max(n - 1, 1)) != hipSuccess ||
        hipMalloc(&ws, L.total) != hipSuccess || hipMalloc(&dinfo, sizeof(int)) != hipSuccess) {
        result = -1000 - BRDC_STATUS_ALLOC_FAILED;
    } else if (hipMemcpy(dd, d, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(de, e, sizeof(double) * (n - 1), hipMemcpyHostToDevice) != hipSuccess) {
        result = -1000 - BRDC_STATUS_HIP_ERROR;
    } else {
        const int st = enqueue_solve(n, dd, de, ws, L, nullptr, dinfo, leaf, Tuning{});
        if (st != BRDC_STATUS_SUCCESS ||
            hipMemcpy(d, dd, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            result = -1000 - BRDC_STATUS_HIP_ERROR;
        } else {
            result = info;
        }
    }
    (void)hipFree(dd);
    (void)hipFree(de);
    (void)hipFree(ws);
    (void)hipFree(dinfo);
    if (work) work[0] = (double)lwmin;
    if (iwork) iwork[0] = (int)liwmin;
    return result;
}

void brdc_dstebr_(const int* n, double* d, double* e, double* work, const int* lwork, int* iwork,
                  const int* liwork, int* info) {
    *info = brdc_dstebr(*n, d, e, work, *lwork, iwork, *liwork);
}

// The reference's own Fortran symbol, weak so that a strong definition linked
// into the same program (the reference's objects) takes precedence.
__attribute__((weak)) void dstedc_dc_br_(const int* n, double* d, double* e, double* work,
                                         const int* lwork, int* iwork, const int* liwork,
                                         int* info) {
    brdc_dstebr_(n, d, e, work, lwork, iwork, liwork, info);
}

int brdc_batch_plan_create(int nmat, const int* sizes, const brdc_options* opts,
                           brdc_batch_plan** plan) {
    if (!plan) return BRDC_STATUS_INVALID_POINTER;
    *plan = nullptr;
    if (nmat > 0 && !sizes) return BRDC_STATUS_INVALID_POINTER;
    if (nmat < 0 || !options_ok(opts)) return BRDC_STATUS_INVALID_VALUE;
    long long rows = 0;
    for (int b = 0; b < nmat; b++) {
        if (sizes[b] < 0) return BRDC_STATUS_INVALID_VALUE;
        rows += sizes[b];
    }
    // row indices and plan entries are int32; per-row pair slots hold 2*rows doubles
    if (rows > (long long)(INT_MAX / 4)) return BRDC_STATUS_INVALID_VALUE;
    brdc_batch_plan* p = new (std::nothrow) brdc_batch_plan;
    if (!p) return BRDC_STATUS_ALLOC_FAILED;
    p->nmat = nmat;
    p->rows = (int)rows;
    p->leaf = leaf_of(opts, sizes, nmat);
    p->tune = tuning_of(opts);
    p->moff.resize((size_t)nmat + 1);
    std::vector<int> lev((size_t)nmat, 0);
    int H = 0;
    p->moff[0] = 0;
    p->tprefix.assign((size_t)nmat + 1, 0);
    p->poff.assign((size_t)nmat + 1, 0);
    for (int b = 0; b < nmat; b++) {
        p->moff[b + 1] = p->moff[b] + sizes[b];
        lev[b] = (sizes[b] > 0) ? tree_levels(sizes[b], p->leaf) : -1;
        H = std::max(H, lev[b]);
        const int tasks = (sizes[b] + kWave - 1) / kWave;
        p->tprefix[b + 1] = p->tprefix[b] + tasks;
        p->poff[b + 1] = p->poff[b] + (long long)tasks * 2 * sizes[b];
        p->rows_max = std::max(p->rows_max, sizes[b]);
    }
    // each matrix's subtrees (the same choice as a single solve): hs[b]
    // merge levels at the bottom of its tree run in k_subtree, 0 for none
    // (all of them, or none when the batch holds more subtree roots than
    // the tuning admits)
    std::vector<int> hs((size_t)nmat, 0);
    long long nsub = 0;
    for (int b = 0; b < nmat; b++) {
        if (lev[b] < 1) continue;
        const int sub = subtree_depth(sizes[b], lev[b]);
        if (sub >= 0) {
            hs[b] = lev[b] - sub;
            nsub += 1LL << sub;
        }
    }
    if (nsub > p->tune.sub_max) std::fill(hs.begin(), hs.end(), 0);
    // leaves of the matrices without subtrees (a matrix that fits one leaf
    // is solved there)
    for (int b = 0; b < nmat; b++) {
        if (lev[b] < 0) continue;
        for (int i = 0; i < (1 << lev[b]); i++) {
            int s, sz;
            node_geom(sizes[b], lev[b], i, &s, &sz);
            p->leaf_rows = std::max(p->leaf_rows, sz);
            p->nleaves++;
            if (hs[b] == 0)
                p->items.push_back(make_int4(p->moff[b] + s, sz, b, lev[b] == 0 ? kItemRoot : 0));
        }
    }
    p->leaf_items = (int)p->items.size();
    // merges h levels above the leaves, h = 1 .. H: first those that run as
    // level launches, then the subtrees' merges (records only), by matrix;
    // inside[h][b]: the position of matrix b's first subtree merge of height h
    std::vector<std::vector<size_t>> inside((size_t)H + 1, std::vector<size_t>((size_t)nmat, 0));
    for (int h = 1; h <= H; h++) {
        HostLevel L;
        L.plan_at = p->items.size();
        L.counts_rec = p->merge_records;
        for (int pass = 0; pass < 2; pass++) {
            for (int b = 0; b < nmat; b++) {
                if (lev[b] < h || (pass == 0) != (h > hs[b])) continue;
                const int depth = lev[b] - h;
                if (pass == 1) inside[(size_t)h][(size_t)b] = p->items.size() - L.plan_at;
                for (int i = 0; i < (1 << depth); i++) {
                    int s, sz;
                    node_geom(sizes[b], depth, i, &s, &sz);
                    p->items.push_back(
                        make_int4(p->moff[b] + s, sz, b, depth == 0 ? kItemRoot : 0));
                    if (pass == 0) {
                        L.maxsz = std::max(L.maxsz, sz);
                        L.any_nonroot = L.any_nonroot || depth > 0;
                    }
                }
            }
            if (pass == 0) L.launch = (int)(p->items.size() - L.plan_at);
        }
        L.merges = (int)(p->items.size() - L.plan_at);
        p->merge_records += (size_t)L.merges;
        p->levels.push_back(L);
    }
    // subtree roots: matrix b's nodes at depth lev[b] - hs[b]; the merges of
    // root q at relative depth r (height hs[b] - r) are records
    // inside[h][b] + q 2^r onward of height h
    p->sub_at = p->items.size();
    for (int b = 0; b < nmat; b++) {
        if (hs[b] == 0) continue;
        const int depth = lev[b] - hs[b];
        for (int q = 0; q < (1 << depth); q++) {
            int s, sz;
            node_geom(sizes[b], depth, q, &s, &sz);
            p->items.push_back(make_int4(p->moff[b] + s, sz, b,
                                         (depth == 0 ? kItemRoot : 0) | (hs[b] << kSubLevelShift)));
            for (int r = 0; r < kSubMaxLevels; r++) {
                const int h = hs[b] - r;
                p->srec.push_back(r < hs[b] ? (int)(p->levels[(size_t)h - 1].counts_rec +
                                                    inside[(size_t)h][(size_t)b]) +
                                                  (q << r)
                                            : 0);
            }
        }
    }
    p->nsub = (int)(p->items.size() - p->sub_at);
    // gradient items: matrices by descending order (ties by index) in every
    // launch, so the longest work comes first; a long matrix's tasks split
    // into chains, twist-search ranges, z directions and contribution ranges
    {
        std::vector<int> mats((size_t)nmat);
        for (int b = 0; b < nmat; b++) mats[(size_t)b] = b;
        std::stable_sort(mats.begin(), mats.end(),
                         [&](int a, int b) { return sizes[a] > sizes[b]; });
        std::vector<int4> L1, L2, L3, L4;
        std::vector<int4>* lists[4] = {&L1, &L2, &L3, &L4};
        int shorts = 0;
        for (int b : mats) {
            const int segs = (sizes[b] + kSeg - 1) / kSeg, parts = grad_parts(sizes[b]);
            for (int t = p->tprefix[b]; t < p->tprefix[b + 1]; t++) {
                if (parts == 1) {
                    // a short matrix's task runs whole, dealt round-robin over
                    // the four launches beside the long matrices' items
                    lists[shorts++ & 3]->push_back(make_int4(t, kGAll, 0, 0));
                    continue;
                }
                L1.push_back(make_int4(t, kGChainF, 0, 0));
                L1.push_back(make_int4(t, kGChainB, 0, 0));
                const int g0 = (int)L2.size();
                for (int c = 0; c < parts; c++) {
                    const int sa = (int)((long long)segs * c / parts);
                    const int sb = (int)((long long)segs * (c + 1) / parts);
                    L2.push_back(make_int4(t, kGGamma, sa, sb));
                    L4.push_back(make_int4(t, kGContrib, sa, sb));
                }
                L3.push_back(make_int4(t, kGZDown, g0, parts));
                L3.push_back(make_int4(t, kGZUp, g0, parts));
            }
        }
        p->gitems.clear();
        for (int l = 0; l < 4; l++) {
            p->gcount[l] = (int)lists[l]->size();
            p->gitems.insert(p->gitems.end(), lists[l]->begin(), lists[l]->end());
        }
        p->cpoff.assign((size_t)nmat + 1, 0);
        for (int b = 0; b < nmat; b++)
            p->cpoff[b + 1] = p->cpoff[b] +
                              (long long)(p->tprefix[b + 1] - p->tprefix[b]) * grad_slot(sizes[b]);
    }
    p->layout = make_layout(p->rows, nmat, p->merge_records, std::max<size_t>(p->items.size(), 1),
                            p->gitems.size(), (size_t)p->nsub);
    *plan = p;
    return BRDC_STATUS_SUCCESS;
}

void brdc_batch_plan_destroy(brdc_batch_plan* plan) { delete plan; }

int brdc_batch_workspace(const brdc_batch_plan* plan, size_t* bytes) {
    if (!plan || !bytes) return BRDC_STATUS_INVALID_POINTER;
    *bytes = plan->layout.total;
    return BRDC_STATUS_SUCCESS;
}

int brdc_batch_merge_stats(const brdc_batch_plan* plan, const void* workspace,
                           size_t workspace_bytes, void* stream, int* out_host, int max_merges) {
    if (!plan || !workspace || (!out_host && max_merges > 0)) return -BRDC_STATUS_INVALID_POINTER;
    const Layout& L = plan->layout;
    if (workspace_bytes < L.total) return -BRDC_STATUS_INVALID_WORKSPACE;
    const int total = (int)plan->merge_records;
    if (total == 0) return 0;
    std::vector<int> counts((size_t)kCountStride * total);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (hipMemcpyAsync(counts.data(), static_cast<const char*>(workspace) + L.counts,
                       counts.size() * sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return -BRDC_STATUS_HIP_ERROR;
    int rec = 0;
    for (size_t h = 0; h < plan->levels.size(); h++) {
        const HostLevel& lv = plan->levels[h];
        for (int m = 0; m < lv.merges; m++) {
            if (rec >= max_merges) return total;
            const int4 it = plan->items[lv.plan_at + (size_t)m];
            const int* c = &counts[(size_t)kCountStride * (lv.counts_rec + (size_t)m)];
            int* o = out_host + kMergeStatFields * rec;
            o[0] = (int)h + 1;
            o[1] = it.x;
            o[2] = it.y;
            o[3] = c[0];
            o[4] = c[1];
            o[5] = c[3];
            o[6] = c[2];
            o[7] = c[5];
            o[8] = c[6];
            rec++;
        }
    }
    return total;
}

int brdc_batch_plan_stats(const brdc_batch_plan* plan, int* out, int nout) {
    if (!plan || (!out && nout > 0)) return -BRDC_STATUS_INVALID_POINTER;
    const int v[7] = {plan->rows,       plan->nmat, plan->nleaves, (int)plan->levels.size(),
                      (int)plan->merge_records, plan->leaf_rows, plan->leaf};
    const int k = std::min(nout, 7);
    for (int i = 0; i < k; i++) out[i] = v[i];
    return k;
}

int brdc_batch_upload(const brdc_batch_plan* plan, void* workspace, size_t workspace_bytes,
                      void* stream) {
    if (!plan || !workspace) return BRDC_STATUS_INVALID_POINTER;
    const Layout& L = plan->layout;
    if (workspace_bytes < L.total || (reinterpret_cast<uintptr_t>(workspace) % kAlign) != 0)
        return BRDC_STATUS_INVALID_WORKSPACE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (!plan->items.empty() &&
        hipMemcpyAsync(at<int4>(workspace, L.plan), plan->items.data(),
                       plan->items.size() * sizeof(int4), hipMemcpyHostToDevice, st) != hipSuccess)
        return BRDC_STATUS_HIP_ERROR;
    if (hipMemcpyAsync(at<int>(workspace, L.moff), plan->moff.data(),
                       plan->moff.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(at<int>(workspace, L.tpre), plan->tprefix.data(),
                       plan->tprefix.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(at<long long>(workspace, L.poff), plan->poff.data(),
                       plan->poff.size() * sizeof(long long), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        (!plan->gitems.empty() &&
         hipMemcpyAsync(at<int4>(workspace, L.gitm), plan->gitems.data(),
                        plan->gitems.size() * sizeof(int4), hipMemcpyHostToDevice, st) !=
             hipSuccess) ||
        hipMemcpyAsync(at<long long>(workspace, L.cpof), plan->cpoff.data(),
                       plan->cpoff.size() * sizeof(long long), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        (!plan->srec.empty() &&
         hipMemcpyAsync(at<int>(workspace, L.srec), plan->srec.data(),
                        plan->srec.size() * sizeof(int), hipMemcpyHostToDevice, st) !=
             hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess)
        return BRDC_STATUS_HIP_ERROR;
    return BRDC_STATUS_SUCCESS;
}

int brdc_batch_solve(const brdc_batch_plan* plan, double* d_dev, const double* e_dev,
                     void* workspace, size_t workspace_bytes, void* stream, int* info_dev) {
    if (!plan) return BRDC_STATUS_INVALID_POINTER;
    if (plan->nmat == 0) return BRDC_STATUS_SUCCESS;
    if (!d_dev || !e_dev || !workspace || !info_dev) return BRDC_STATUS_INVALID_POINTER;
    const Layout& L = plan->layout;
    if (workspace_bytes < L.total || (reinterpret_cast<uintptr_t>(workspace) % kAlign) != 0)
        return BRDC_STATUS_INVALID_WORKSPACE;
    return enqueue(batch_schedule(*plan, workspace), d_dev, e_dev, workspace, L,
                   static_cast<hipStream_t>(stream), info_dev, nullptr);
}

int brdc_dstebr_grad_workspace(int n, size_t* bytes) {
    if (!bytes) return BRDC_STATUS_INVALID_POINTER;
    if (n < 0) return BRDC_STATUS_INVALID_VALUE;
    const int tasks = (n + kWave - 1) / kWave;
    int count[4];
    grad_counts_one(n, count);
    *bytes = grad_shape(tasks, tasks * grad_slot(n), (long long)tasks * 2 * n, count[1], n, n)
                 .total;
    return (*bytes > kGradScratchMax) ? BRDC_STATUS_INVALID_VALUE : BRDC_STATUS_SUCCESS;
}

int brdc_dstebr_grad_device(int n, const double* d_dev, const double* e_dev,
                            const double* w_dev, const double* g_dev, double* dd_dev,
                            double* de_dev, void* scratch, size_t scratch_bytes, void* stream) {
    if (n < 0) return BRDC_STATUS_INVALID_VALUE;
    if (n == 0) return BRDC_STATUS_SUCCESS;
    if (!d_dev || !e_dev || !w_dev || !g_dev || !dd_dev || !de_dev || !scratch)
        return BRDC_STATUS_INVALID_POINTER;
    const int tasks = (n + kWave - 1) / kWave;
    const long long partial = (long long)tasks * 2 * n, cpd = tasks * grad_slot(n);
    int count[4];
    grad_counts_one(n, count);
    if (grad_shape(tasks, cpd, partial, count[1], n, n).total > kGradScratchMax)
        return BRDC_STATUS_INVALID_VALUE;
    if (scratch_bytes < grad_shape(tasks, cpd, partial, count[1], n, n).total ||
        (reinterpret_cast<uintptr_t>(scratch) % kAlign) != 0)
        return BRDC_STATUS_INVALID_WORKSPACE;
    return enqueue_grad(n, 1, n, nullptr, nullptr, nullptr, nullptr, nullptr, count, grad_parts(n),
                        tasks, cpd, partial, d_dev, e_dev, w_dev, g_dev, dd_dev, de_dev,
                        scratch, static_cast<hipStream_t>(stream));
}

int brdc_batch_grad_workspace(const brdc_batch_plan* plan, size_t* bytes) {
    if (!plan || !bytes) return BRDC_STATUS_INVALID_POINTER;
    *bytes = grad_shape(plan->tprefix.back(), plan->cpoff.back(), plan->poff.back(),
                        plan->gcount[1], plan->rows, plan->rows_max).total;
    return (*bytes > kGradScratchMax) ? BRDC_STATUS_INVALID_VALUE : BRDC_STATUS_SUCCESS;
}

int brdc_batch_grad(const brdc_batch_plan* plan, const double* d_dev, const double* e_dev,
                    const double* w_dev, const double* g_dev, double* dd_dev, double* de_dev,
                    const void* workspace, size_t workspace_bytes, void* scratch,
                    size_t scratch_bytes, void* stream) {
    if (!plan) return BRDC_STATUS_INVALID_POINTER;
    if (plan->rows == 0) return BRDC_STATUS_SUCCESS;
    if (!d_dev || !e_dev || !w_dev || !g_dev || !dd_dev || !de_dev || !workspace || !scratch)
        return BRDC_STATUS_INVALID_POINTER;
    const Layout& L = plan->layout;
    const int ntask = plan->tprefix.back();
    const long long partial = plan->poff.back();
    const size_t need = grad_shape(ntask, plan->cpoff.back(), partial, plan->gcount[1],
                                   plan->rows, plan->rows_max).total;
    if (need > kGradScratchMax) return BRDC_STATUS_INVALID_VALUE;
    if (workspace_bytes < L.total || scratch_bytes < need ||
        (reinterpret_cast<uintptr_t>(scratch) % kAlign) != 0)
        return BRDC_STATUS_INVALID_WORKSPACE;
    void* ws = const_cast<void*>(workspace);
    return enqueue_grad(plan->rows, plan->nmat, plan->rows_max, at<int>(ws, L.moff),
                        at<int>(ws, L.tpre),
                        at<long long>(ws, L.poff), at<long long>(ws, L.cpof),
                        at<int4>(ws, L.gitm), plan->gcount, 1, ntask, plan->cpoff.back(), partial,
                        d_dev, e_dev,
                        w_dev, g_dev, dd_dev, de_dev, scratch, static_cast<hipStream_t>(stream));
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free device entry.  The workspace of brdc_dstebr_device lives in
// a static device array, one copy per GPU, held until the process ends.  In
// host code `scratch` names a host-side stand-in, not device memory: it is
// used only to look up each device's address, and every launch gets that
// address.  All calls on one GPU share its copy, so calls that overlap on a
// GPU race; the kernels write every word they later read (counts, records,
// INFO), so nothing needs clearing between calls.
static __device__ __align__(256) unsigned char scratch[BRDC_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of the current device's scratch, looked up on the first call
// on that device (with it current) and kept.
int scratch_address(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return BRDC_STATUS_HIP_ERROR;
    if (dev < 0 || dev >= kMaxDevices) return BRDC_STATUS_INTERNAL_ERROR;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return BRDC_STATUS_HIP_ERROR;
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return BRDC_STATUS_SUCCESS;
}
}  // namespace

extern "C" int brdc_dstebr_device_noalloc(int n, double* d_dev, const double* e_dev,
                                          void* stream, int* info_dev,
                                          const brdc_options* opts) {
    if (n < 0 || n > BRDC_NOALLOC_MAX_N || !options_ok(opts)) return BRDC_STATUS_INVALID_VALUE;
    if (n == 0) return BRDC_STATUS_SUCCESS;
    if (!d_dev || (n > 1 && !e_dev)) return BRDC_STATUS_INVALID_POINTER;
    const int leaf = leaf_of(opts, n);
    const Layout L = tree_layout(n, leaf);
    if (L.total > sizeof(scratch)) return BRDC_STATUS_INVALID_VALUE;
    void* ws = nullptr;
    const int st = scratch_address(&ws);
    if (st != BRDC_STATUS_SUCCESS) return st;
    return enqueue_solve(n, d_dev, e_dev, ws, L, static_cast<hipStream_t>(stream), info_dev,
                         leaf, tuning_of(opts));
}
