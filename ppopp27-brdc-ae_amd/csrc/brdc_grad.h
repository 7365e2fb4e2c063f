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

constexpr int kGradWaves = 4;  // waves per block of the vector kernel

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

// slot: doubles per scratch array per wave (rows_max * 64).  partial: per
// task 2 * n_b doubles at poff[b] + (t - tprefix[b]) * 2 * n_b.
__global__ __launch_bounds__(kWave * kGradWaves) void k_grad_vectors(
    int nmat, const int* __restrict__ moff, const int* __restrict__ tprefix,
    const long long* __restrict__ poff, int ntask, int n_one, const double* __restrict__ d,
    const double* __restrict__ e, const double* __restrict__ w, const double* __restrict__ g,
    double* __restrict__ bufA, double* __restrict__ bufB, long long slot,
    double* __restrict__ partial) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = uniform((int)(blockIdx.x * kGradWaves + threadIdx.x / kWave));
    const int nwave = (int)gridDim.x * kGradWaves;
    double* __restrict__ A = bufA + (long long)wave * slot;
    double* __restrict__ B = bufB + (long long)wave * slot;
    const double safmin = 0x1p-1022;
    for (int t = wave; t < ntask; t += nwave) {
        const int b = uniform(grad_task_matrix(tprefix, nmat, t));
        const int off = moff ? uniform(moff[b]) : 0;
        const int n = moff ? uniform(moff[b + 1]) - off : n_one;
        const int t0 = tprefix ? uniform(tprefix[b]) : 0;
        const long long po = poff ? poff[b] : 0;
        double* __restrict__ part = partial + po + (long long)(t - t0) * 2 * n;
        const int i = (t - t0) * kWave + lane;
        const bool active = i < n;
        const CDoubles D = as_const(d + off), Ev = as_const(e + off);
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
        // forward stationary factorization: A[k] = D+_k
        double dp = pivot_guard(D[0] - mu, pivmin);
        A[lane] = dp;
        for (int k = 0; k < n - 1; k++) {
            const double ek = Ev[k];
            const double l = ek / dp;
            dp = pivot_guard((D[k + 1] - mu) - l * ek, pivmin);
            A[(long long)(k + 1) * kWave + lane] = dp;
        }
        // backward factorization: B[k] = u_k; twist index r = argmin |gamma_k|
        double dm = pivot_guard(D[n - 1] - mu, pivmin);
        double gbest = fabs(A[(long long)(n - 1) * kWave + lane] + dm - (D[n - 1] - mu));
        int r = n - 1;
        for (int k = n - 2; k >= 0; k--) {
            const double ek = Ev[k];
            const double u = ek / dm;
            B[(long long)k * kWave + lane] = u;
            const double dk = D[k] - mu;
            dm = pivot_guard(dk - u * ek, pivmin);
            const double gam = fabs(A[(long long)k * kWave + lane] + dm - dk);
            if (gam < gbest) {
                gbest = gam;
                r = k;
            }
        }
        // eigenvector: below r into B[k] (holding z_{k+1}), above r into A[k]
        double nrm = 1.0, z = 1.0;
        for (int k = 0; k < n - 1; k++) {
            if (k >= r) {
                const long long at = (long long)k * kWave + lane;
                z = -B[at] * z;
                B[at] = z;
                nrm += z * z;
            }
        }
        z = 1.0;
        for (int k = n - 2; k >= 0; k--) {
            if (k < r) {
                const long long at = (long long)k * kWave + lane;
                z = -(Ev[k] / A[at]) * z;
                A[at] = z;
                nrm += z * z;
            }
        }
        // contributions g_i z_k^2 / |z|^2 and 2 g_i z_k z_{k+1} / |z|^2, summed
        // over the wave's eigenvalues
        const double s = active ? gi / nrm : 0.0;
        auto zk_at = [&](int k) -> double {
            if (k < r) return A[(long long)k * kWave + lane];
            if (k == r) return 1.0;
            return B[(long long)(k - 1) * kWave + lane];
        };
        double zc = zk_at(0);
        for (int k = 0; k < n; k++) {
            const double zn = (k + 1 < n) ? zk_at(k + 1) : 0.0;
            double cd = s * zc * zc;
            double ce = 2.0 * s * zc * zn;
            for (int o = 32; o > 0; o >>= 1) {
                cd += __shfl_xor(cd, o, kWave);
                ce += __shfl_xor(ce, o, kWave);
            }
            if (lane == 0) {
                part[k] = cd;
                part[n + k] = ce;
            }
            zc = zn;
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
