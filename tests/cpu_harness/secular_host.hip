// Host build of the device secular root finder (brdc_secular.h), so the CPU
// test suite can compare it root by root with the oracle's DLAED4 without a
// GPU.  Test infrastructure only.
#include "../../ppopp27-brdc-ae_amd/csrc/brdc_secular.h"

namespace {
// Sequential evaluator: the same sums the device computes with L lanes per root.
struct HostEval {
    const double* dp;
    const double* zp;
    int K;
    double d(int j) const { return dp[j]; }
    double zsq(int j) const { return zp[j] * zp[j]; }
    brdc::SecSums sums(double dorg, double tau, int ex_lo, int nex) const {
        brdc::SecSums s = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int j = 0; j < K; j++) {
            if ((unsigned)(j - ex_lo) < (unsigned)nex) continue;
            const double delta = (dp[j] - dorg) - tau;
            const double r = brdc::rcp_nr(delta);
            const double term = zsq(j) * r;
            s.S += term;
            s.A += fabs(term);
            const double x = term * fabs(r);
            if (x < 0.0) s.dpsi -= x;
            else s.dphi += x;
            s.J += (double)j * term;
        }
        return s;
    }
};
}  // namespace

extern "C" int secular_host(int K, int i, const double* d, const double* z, double rho,
                            double* out /* lambda, tau, near1, near2, org, iters */) {
    if (K <= 2) {
        double lam, v0, v1, tau;
        if (K == 1) {
            out[0] = d[0] + rho * z[0] * z[0];
            out[1] = out[2] = out[3] = out[4] = 0.0;
            out[5] = 1;
            return 0;
        }
        brdc::secular_two(i, d, z, rho, &lam, &v0, &v1, &tau);
        out[0] = lam;
        out[1] = tau;
        out[2] = v0;
        out[3] = v1;
        out[4] = 0;
        out[5] = 1;
        return 0;
    }
    HostEval E{d, z, K};
    const brdc::RootState r = brdc::secular_root(E, K, i, rho);
    out[0] = r.lambda;
    out[1] = r.tau;
    out[2] = r.near1;
    out[3] = r.near2;
    out[4] = r.org;
    out[5] = r.iters;
    return r.info;
}

// All K roots of one merge; returns the number of roots that did not converge
// and writes each root's (lambda, tau, near1, near2, org, iters) to out[6*i].
extern "C" int secular_host_all(int K, const double* d, const double* z, double rho, double* out) {
    int bad = 0;
    for (int i = 0; i < K; i++) bad += (secular_host(K, i, d, z, rho, out + 6 * i) != 0);
    return bad;
}

// Wave-slot bound of the secular launch: for every depth of an order-n tree
// with the given leaf size, count the K <= max node size whose wave need
// exceeds the slots the launch provides.  The kernel stays correct either
// way (a task walks all root groups of its merge); this checks that the
// bound is tight so no wave does two groups' work.
#include "../../ppopp27-brdc-ae_amd/csrc/brdc_launch.h"
#include "../../ppopp27-brdc-ae_amd/csrc/brdc_tree.h"

extern "C" int secular_slot_violations(int n, int leaf) {
    int bad = 0;
    const int levels = brdc::tree_levels(n, leaf);
    for (int depth = 0; depth < levels; depth++) {
        const int merges = 1 << depth;
        const int maxsz = brdc::max_node_size(n, depth);
        const int slots = brdc::secular_slots(maxsz, merges);
        for (int K = 1; K <= maxsz; K++) bad += (brdc::secular_waves(K, merges) > slots);
    }
    return bad;
}
