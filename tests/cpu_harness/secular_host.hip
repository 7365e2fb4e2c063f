// Host build of the device secular root finder (brdc_secular.h), so the CPU
// test suite can compare it root by root with the oracle's DLAED4 without a
// GPU.  Test infrastructure only.
#include "../../ppopp27-brdc-ae_amd/csrc/brdc_secular.h"

namespace {
struct HostPoles {
    const double* dp;
    const double* zp;
    double d(int j) const { return dp[j]; }
    double z(int j) const { return zp[j]; }
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
    HostPoles P{d, z};
    const brdc::RootState r = brdc::secular_root(P, K, i, rho);
    out[0] = r.lambda;
    out[1] = r.tau;
    out[2] = r.near1;
    out[3] = r.near2;
    out[4] = r.org;
    out[5] = r.iters;
    return r.info;
}
