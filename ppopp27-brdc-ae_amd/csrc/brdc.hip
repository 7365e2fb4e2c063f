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
#include <vector>

#include "../../include/brdc.h"
#include "brdc_kernels.h"

namespace brdc {
namespace {

constexpr int kDefaultLeaf = 16;
constexpr int kSecWaves = 4;       // waves per block of the secular kernel (independent waves)
constexpr int kSecMaxWaves = 3072;  // secular grid cap: about 3 resident waves per SIMD
constexpr int kStreamWaves = 8;     // waves per workgroup of the streamed secular kernel
constexpr int kStreamMinK = 2048;   // merges with at least this many poles are streamed

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

inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Workspace carve-up; identical for the query and the run.
struct Layout {
    size_t scal, ew;
    size_t sd[2], slo[2], shi[2];
    size_t md, mz, mq1, mq2;
    size_t ad, az, adz, aq;
    size_t gd, gq1, gq2;
    size_t td, tq1, tq2;
    size_t rlam, rs, rn1, rn2;
    size_t pw, qo1, qo2;
    size_t P, rend, fl;
    size_t counts, info;
    size_t total;
    int levels;
    int merges;  // 2^levels (merge records, index (2^k - 1) + i at depth k)
};

Layout make_layout(int n, int leaf) {
    Layout L{};
    const size_t nd = (size_t)std::max(n, 1) * sizeof(double);
    const size_t ni = (size_t)std::max(n, 1) * sizeof(int);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = align_up(off + bytes);
        return at;
    };
    L.levels = (n > 0) ? tree_levels(n, leaf) : 0;
    L.merges = 1 << L.levels;
    L.scal = take(4 * sizeof(double));
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
    L.counts = take((size_t)kCountStride * L.merges * sizeof(int));
    L.info = take(sizeof(int));
    L.total = off;
    return L;
}

int leaf_of(const brdc_options* o) {
    if (!o || o->leaf_max == 0) return kDefaultLeaf;
    return o->leaf_max;
}

bool options_ok(const brdc_options* o) {
    if (!o) return true;
    if (o->leaf_max != 0 && (o->leaf_max < 2 || o->leaf_max > kMaxLeaf)) return false;
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
                   kNumClasses };
struct Profiler {
    bool on = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
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
        g_prof.marks.push_back({cls, {a, b}});
    }
};

// the O(K^2) kernels come in widths of 1..16 waves per 64-item group
template <class K1, class K2, class K4, class K8, class... Args>
void launch_wide(K1 k1, K2 k2, K4 k4, K8 k8, int WL, int blocks, hipStream_t st, Args... args) {
    const dim3 grid(blocks);
    switch (WL) {
        case 1: hipLaunchKernelGGL(k1, grid, dim3(kGroup), 0, st, args...); break;
        case 2: hipLaunchKernelGGL(k2, grid, dim3(2 * kGroup), 0, st, args...); break;
        case 4: hipLaunchKernelGGL(k4, grid, dim3(4 * kGroup), 0, st, args...); break;
        default: hipLaunchKernelGGL(k8, grid, dim3(8 * kGroup), 0, st, args...); break;
    }
}

// deflation block size follows the level's node size
template <class... Args>
void launch_deflate(int maxsz, int merges, hipStream_t st, Args... args) {
    if (maxsz <= 512)
        hipLaunchKernelGGL(k_deflate<64>, dim3(merges), dim3(64), 0, st, args...);
    else if (maxsz <= 8192)
        hipLaunchKernelGGL(k_deflate<256>, dim3(merges), dim3(256), 0, st, args...);
    else
        hipLaunchKernelGGL(k_deflate<1024>, dim3(merges), dim3(1024), 0, st, args...);
}

int enqueue_solve(int n, double* d, const double* e, void* ws, const Layout& L, hipStream_t st,
                  int* info_user, int leaf) {
    double* scal = at<double>(ws, L.scal);
    double* ew = at<double>(ws, L.ew);
    int* info = at<int>(ws, L.info);
    int* counts_all = at<int>(ws, L.counts);

    {
        ProfScope ps(kNorm, st);
        hipLaunchKernelGGL(k_norm, dim3(1), dim3(1024), 0, st, n, d, e, scal, info, info_user);
    }
    BRDC_LAUNCH_CHECK();
    if (n <= 1) return BRDC_STATUS_SUCCESS;
    {
        ProfScope ps(kPrepE, st);
        hipLaunchKernelGGL(k_prep_e, dim3((n - 1 + 255) / 256), dim3(256), 0, st, n, d, e, scal,
                           ew);
    }
    BRDC_LAUNCH_CHECK();

    const int levels = L.levels;
    int cur = 0;
    double* sd[2] = {at<double>(ws, L.sd[0]), at<double>(ws, L.sd[1])};
    double* slo[2] = {at<double>(ws, L.slo[0]), at<double>(ws, L.slo[1])};
    double* shi[2] = {at<double>(ws, L.shi[0]), at<double>(ws, L.shi[1])};
    const int single = (levels == 0);
    {
        ProfScope ps(kLeaf, st);
        const int leaf_rows = max_node_size(n, levels);
        auto kern = (leaf_rows <= 16) ? k_leaf<16> : (leaf_rows <= 32) ? k_leaf<32> : k_leaf<64>;
        hipLaunchKernelGGL(kern, dim3(1 << levels), dim3(kWave), 0, st, n, levels, d, ew, scal,
                           single ? d : sd[cur], slo[cur], shi[cur], info, single, info_user);
    }
    BRDC_LAUNCH_CHECK();
    if (single) return BRDC_STATUS_SUCCESS;

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

    for (int depth = levels - 1; depth >= 0; depth--) {
        const int merges = 1 << depth;
        const int maxsz = max_node_size(n, depth);
        const int bpm = (maxsz + kThreads - 1) / kThreads;
        const dim3 grid(merges * bpm);
        // the O(K^2) kernels: one block per 64 roots/poles, WL waves wide
        const int bpg = (maxsz + kGroup - 1) / kGroup;
        const int WL = waves_for(maxsz, kMaxWaves);
        int* counts = counts_all + kCountStride * (merges - 1);
        const int nxt = cur ^ 1;
        const int root = (depth == 0);

        {
            ProfScope ps(kPrepare, st);
            hipLaunchKernelGGL(k_merge_prepare, grid, dim3(kThreads), 0, st, n, depth, bpm,
                               sd[cur], slo[cur], shi[cur], ew, md, mz, mq1, mq2);
        }
        BRDC_LAUNCH_CHECK();
        {
            ProfScope ps(kDeflate, st);
            launch_deflate(maxsz, merges, st, n, depth, ew, md, mz, mq1, mq2, P, rend, fl, ad, az,
                           adz, aq, gd, gq1, gq2, td, tq1, tq2, counts);
        }
        BRDC_LAUNCH_CHECK();
        // merges of K >= kStreamMinK poles go to the streamed kernel, the
        // rest to the lanes-per-root kernel
        const bool streamed = (maxsz >= kStreamMinK);
        {
            ProfScope ps(kSecular, st);
            // wave slots per merge: the most any K it handles can need; the
            // grid is capped and walks (merge, slot) tasks
            const int slots = secular_slots(std::min(maxsz, kStreamMinK - 1), merges);
            const int waves = std::min(merges * slots, kSecMaxWaves);
            const int blocks = (waves + kSecWaves - 1) / kSecWaves;
            hipLaunchKernelGGL(k_secular<kSecWaves>, dim3(blocks), dim3(kWave * kSecWaves), 0, st,
                               n, depth, slots, streamed ? kStreamMinK : INT_MAX,
                               (const double*)ew, (const int*)counts, (const double*)ad,
                               (const double*)az, (const double2*)adz, rlam, rs, rn1, rn2, counts,
                               info);
        }
        BRDC_LAUNCH_CHECK();
        if (streamed) {
            ProfScope ps(kSecular, st);
            const int groups = std::min(stream_groups(), (n + kWave - 1) / kWave);
            hipLaunchKernelGGL(k_secular_stream<kStreamWaves>, dim3(groups),
                               dim3(kWave * kStreamWaves), 0, st, n, depth, kStreamMinK,
                               (const double*)ew, (const int*)counts, (const double2*)adz, rlam,
                               rs, rn1, rn2, counts, info);
            BRDC_LAUNCH_CHECK();
        }
        if (!root) {
            {
                ProfScope ps(kWprod, st);
                launch_wide(k_wprod<1>, k_wprod<2>, k_wprod<4>, k_wprod<8>, WL,
                            merges * bpg, st, n, depth, bpg, (const int*)counts,
                            (const double*)ad, (const double*)az, (const double2*)rs,
                            (const double*)rn1, (const double*)rn2, pw);
            }
            BRDC_LAUNCH_CHECK();
            {
                ProfScope ps(kRows, st);
                launch_wide(k_rows<1>, k_rows<2>, k_rows<4>, k_rows<8>, WL,
                            merges * bpg, st, n, depth, bpg, (const int*)counts,
                            (const double2*)aq, (const double2*)pw, (const double2*)rs,
                            (const double*)rn1, (const double*)rn2, qo1, qo2);
            }
            BRDC_LAUNCH_CHECK();
        }
        {
            ProfScope ps(kFinalize, st);
            hipLaunchKernelGGL(k_finalize, grid, dim3(kThreads), 0, st, n, depth, bpm, counts,
                               rlam, qo1, qo2, gd, gq1, gq2, td, tq1, tq2, root ? d : sd[nxt],
                               slo[nxt], shi[nxt], root, scal, info, info_user);
        }
        BRDC_LAUNCH_CHECK();
        cur = nxt;
    }
    return BRDC_STATUS_SUCCESS;
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
    *bytes = make_layout(n, leaf_of(opts)).total;
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
    const int leaf = leaf_of(opts);
    const Layout L = make_layout(n, leaf);
    if (workspace_bytes < L.total || (reinterpret_cast<uintptr_t>(workspace) % kAlign) != 0)
        return BRDC_STATUS_INVALID_WORKSPACE;
    return enqueue_solve(n, d_dev, e_dev, workspace, L, static_cast<hipStream_t>(stream), info_dev,
                         leaf);
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
    const Layout L = make_layout(n, leaf_of(opts));
    if (workspace_bytes < L.total) return -BRDC_STATUS_INVALID_WORKSPACE;
    const int total = (n > 1) ? (1 << L.levels) - 1 : 0;
    if (total == 0) return 0;
    std::vector<int> counts((size_t)kCountStride * total);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (hipMemcpyAsync(counts.data(),
                       static_cast<const char*>(workspace) + L.counts,
                       counts.size() * sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return -BRDC_STATUS_HIP_ERROR;
    int rec = 0;
    for (int depth = 0; depth < L.levels; depth++) {
        for (int i = 0; i < (1 << depth); i++) {
            if (rec >= max_merges) return total;
            int s, sz;
            node_geom(n, depth, i, &s, &sz);
            const int* c = &counts[(size_t)kCountStride * ((1 << depth) - 1 + i)];
            int* o = out_host + 6 * rec;
            o[0] = depth;
            o[1] = s;
            o[2] = sz;
            o[3] = c[0];
            o[4] = c[1];
            o[5] = c[3];
            rec++;
        }
    }
    return total;
}

int brdc_profile_begin(void) {
    for (auto& m : g_prof.marks) {
        (void)hipEventDestroy(m.second.first);
        (void)hipEventDestroy(m.second.second);
    }
    g_prof.marks.clear();
    g_prof.on = true;
    return BRDC_STATUS_SUCCESS;
}

int brdc_profile_end(double* ms_out, int* launches_out, int nclasses) {
    g_prof.on = false;
    int status = BRDC_STATUS_SUCCESS;
    std::vector<double> ms(kNumClasses, 0.0);
    std::vector<int> cnt(kNumClasses, 0);
    for (auto& m : g_prof.marks) {
        float t = 0.0f;
        if (hipEventSynchronize(m.second.second) != hipSuccess ||
            hipEventElapsedTime(&t, m.second.first, m.second.second) != hipSuccess)
            status = BRDC_STATUS_HIP_ERROR;
        ms[m.first] += t;
        cnt[m.first] += 1;
        (void)hipEventDestroy(m.second.first);
        (void)hipEventDestroy(m.second.second);
    }
    g_prof.marks.clear();
    for (int c = 0; c < nclasses && c < kNumClasses; c++) {
        if (ms_out) ms_out[c] = ms[c];
        if (launches_out) launches_out[c] = cnt[c];
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

    const Layout L = make_layout(n, kDefaultLeaf);
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
        const int st = enqueue_solve(n, dd, de, ws, L, nullptr, dinfo, kDefaultLeaf);
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

}  // extern "C"
