// brdc_secular.h -- device secular-equation root finder for one root per lane.
//
// Solves 1 + rho * sum_j z_j^2 / (d_j - x) = 0 for root i in (d_i, d_{i+1})
// (or (d_{K-1}, d_{K-1} + rho) for the last root), d strictly increasing,
// rho > 0.  This is the middle-way method of R.-C. Li (LAWN 89), with the
// Gragg three-pole step and the two-pole closed form -- the algorithm of
// LAPACK DLAED4/DLAED6/DLAED5 that the reference calls at dlaed4_br.f:59 and
// dlaed9_br.f:87-127.  The step logic (initial guess, bracketing, switch
// between two- and three-pole models, convergence test |w| <= eps*erretm)
// follows that algorithm; the data layout does not: the K-vector DELTA is
// never stored.  Each evaluation recomputes delta_j = (d_j - d_org) - tau
// from the compact root state (org, tau) that DLAED4_BR exports
// (dlaed4_br.f:79-103), so a lane's working set is a few registers and the
// poles stream through the cache shared by all lanes of the merge.
#pragma once

#include <hip/hip_runtime.h>

namespace brdc {

constexpr double kEps = 0x1p-53;  // DLAMCH('Epsilon'), round-to-nearest
constexpr int kSecularMaxIt = 30;
constexpr int kLaed6MaxIt = 40;

// Result of one root solve: lambda = d[org] + tau; near1/near2 are the
// pole-adjacent deltas (delta[i], delta[i+1], or delta[K-2], delta[K-1] for
// the last root).  For K <= 2, v0/v1 hold the normalised secular eigenvector.
struct RootState {
    double lambda;
    double tau;
    double near1;
    double near2;
    int org;
    int iters;  // secular function evaluations over all K poles
    int info;   // 0 converged, 1 not converged
};

__host__ __device__ __forceinline__ double dev_max3(double a, double b, double c) {
    return fmax(fabs(a), fmax(fabs(b), fabs(c)));
}

// Two-pole closed form (the DLAED5 role).  vec gets the normalised eigenvector.
__host__ __device__ inline void secular_two(int i, const double* d, const double* z, double rho,
                                   double* lam, double* v0, double* v1, double* tau_out) {
    const double del = d[1] - d[0];
    const double zz0 = z[0] * z[0], zz1 = z[1] * z[1];
    double b, c, tau, x0, x1;
    if (i == 0) {
        const double w = 1.0 + 2.0 * rho * (zz1 - zz0) / del;
        if (w > 0.0) {
            b = del + rho * (zz0 + zz1);
            c = rho * zz0 * del;
            tau = 2.0 * c / (b + sqrt(fabs(b * b - 4.0 * c)));
            *lam = d[0] + tau;
            x0 = -z[0] / tau;
            x1 = z[1] / (del - tau);
            *tau_out = tau;
        } else {
            b = -del + rho * (zz0 + zz1);
            c = rho * zz1 * del;
            if (b > 0.0)
                tau = -2.0 * c / (b + sqrt(b * b + 4.0 * c));
            else
                tau = (b - sqrt(b * b + 4.0 * c)) / 2.0;
            *lam = d[1] + tau;
            x0 = -z[0] / (del + tau);
            x1 = -z[1] / tau;
            *tau_out = *lam - d[0];
        }
    } else {
        b = -del + rho * (zz0 + zz1);
        c = rho * zz1 * del;
        if (b > 0.0)
            tau = (b + sqrt(b * b + 4.0 * c)) / 2.0;
        else
            tau = 2.0 * c / (-b + sqrt(b * b + 4.0 * c));
        *lam = d[1] + tau;
        x0 = -z[0] / (del + tau);
        x1 = -z[1] / tau;
        *tau_out = *lam - d[0];
    }
    const double nrm = sqrt(x0 * x0 + x1 * x1);
    *v0 = x0 / nrm;
    *v1 = x1 / nrm;
}

// Three-pole step (the DLAED6 role): root of c + sum_k zz_k / (dd_k - t)
// between dd[1], dd[2] (orgati) or dd[0], dd[1], starting from f(0) = finit.
__host__ __device__ inline int secular_three(int kniter, bool orgati, double rho, const double dd[3],
                                    const double zz[3], double finit, double* tau_out) {
    double lbd, ubd, tau = 0.0, a, b, c, eta, f, df, ddf, fc, erretm, temp;
    double ds[3], zs[3];
    int info = 0;
    bool scale = false;
    double sclinv = 1.0;
    if (orgati) {
        lbd = dd[1];
        ubd = dd[2];
    } else {
        lbd = dd[0];
        ubd = dd[1];
    }
    if (finit < 0.0)
        lbd = 0.0;
    else
        ubd = 0.0;
    if (kniter == 2) {
        if (orgati) {
            temp = (dd[2] - dd[1]) / 2.0;
            c = rho + zz[0] / ((dd[0] - dd[1]) - temp);
            a = c * (dd[1] + dd[2]) + zz[1] + zz[2];
            b = c * dd[1] * dd[2] + zz[1] * dd[2] + zz[2] * dd[1];
        } else {
            temp = (dd[0] - dd[1]) / 2.0;
            c = rho + zz[2] / ((dd[2] - dd[1]) - temp);
            a = c * (dd[0] + dd[1]) + zz[0] + zz[1];
            b = c * dd[0] * dd[1] + zz[0] * dd[1] + zz[1] * dd[0];
        }
        temp = dev_max3(a, b, c);
        a /= temp;
        b /= temp;
        c /= temp;
        if (c == 0.0)
            tau = b / a;
        else if (a <= 0.0)
            tau = (a - sqrt(fabs(a * a - 4.0 * b * c))) / (2.0 * c);
        else
            tau = 2.0 * b / (a + sqrt(fabs(a * a - 4.0 * b * c)));
        if (tau < lbd || tau > ubd)
            tau = (lbd + ubd) / 2.0;
        if (dd[0] == tau || dd[1] == tau || dd[2] == tau) {
            tau = 0.0;
        } else {
            temp = finit + tau * zz[0] / (dd[0] * (dd[0] - tau)) +
                   tau * zz[1] / (dd[1] * (dd[1] - tau)) + tau * zz[2] / (dd[2] * (dd[2] - tau));
            if (temp <= 0.0)
                lbd = tau;
            else
                ubd = tau;
            if (fabs(finit) <= fabs(temp))
                tau = 0.0;
        }
    }
    // rescale when a pole is within 2^-340 of tau, so 1/t^3 cannot overflow
    temp = orgati ? fmin(fabs(dd[1] - tau), fabs(dd[2] - tau))
                  : fmin(fabs(dd[0] - tau), fabs(dd[1] - tau));
    if (temp <= 0x1p-340) {
        double sclfac;
        scale = true;
        if (temp <= 0x1p-680) {
            sclfac = 0x1p680;
            sclinv = 0x1p-680;
        } else {
            sclfac = 0x1p340;
            sclinv = 0x1p-340;
        }
        for (int k = 0; k < 3; k++) {
            ds[k] = dd[k] * sclfac;
            zs[k] = zz[k] * sclfac;
        }
        tau *= sclfac;
        lbd *= sclfac;
        ubd *= sclfac;
    } else {
        for (int k = 0; k < 3; k++) {
            ds[k] = dd[k];
            zs[k] = zz[k];
        }
    }
    fc = df = ddf = 0.0;
    for (int k = 0; k < 3; k++) {
        const double r = 1.0 / (ds[k] - tau);
        const double t1 = zs[k] * r, t2 = t1 * r, t3 = t2 * r;
        fc += t1 / ds[k];
        df += t2;
        ddf += t3;
    }
    f = finit + tau * fc;
    if (fabs(f) <= 0.0)
        goto done;
    if (f <= 0.0)
        lbd = tau;
    else
        ubd = tau;
    for (int niter = 2; niter <= kLaed6MaxIt; niter++) {
        double t1, t2;
        if (orgati) {
            t1 = ds[1] - tau;
            t2 = ds[2] - tau;
        } else {
            t1 = ds[0] - tau;
            t2 = ds[1] - tau;
        }
        a = (t1 + t2) * f - t1 * t2 * df;
        b = t1 * t2 * f;
        c = f - (t1 + t2) * df + t1 * t2 * ddf;
        temp = dev_max3(a, b, c);
        a /= temp;
        b /= temp;
        c /= temp;
        if (c == 0.0)
            eta = b / a;
        else if (a <= 0.0)
            eta = (a - sqrt(fabs(a * a - 4.0 * b * c))) / (2.0 * c);
        else
            eta = 2.0 * b / (a + sqrt(fabs(a * a - 4.0 * b * c)));
        if (f * eta >= 0.0)
            eta = -f / df;
        tau += eta;
        if (tau < lbd || tau > ubd)
            tau = (lbd + ubd) / 2.0;
        fc = erretm = df = ddf = 0.0;
        for (int k = 0; k < 3; k++) {
            if (ds[k] - tau == 0.0)
                goto done;
            const double r = 1.0 / (ds[k] - tau);
            const double u1 = zs[k] * r, u2 = u1 * r, u3 = u2 * r, u4 = u1 / ds[k];
            fc += u4;
            erretm += fabs(u4);
            df += u2;
            ddf += u3;
        }
        f = finit + tau * fc;
        erretm = 8.0 * (fabs(finit) + fabs(tau) * erretm) + fabs(tau) * df;
        if (fabs(f) <= 4.0 * kEps * erretm || (ubd - lbd) <= 4.0 * kEps * fabs(tau))
            goto done;
        if (f <= 0.0)
            lbd = tau;
        else
            ubd = tau;
    }
    info = 1;
done:
    if (scale)
        tau *= sclinv;
    *tau_out = tau;
    return info;
}

// Reciprocal for the pole loops: hardware v_rcp_f64 refined by two Newton
// steps (error below one ulp); the host build of this header uses 1/x.
__host__ __device__ __forceinline__ double rcp_nr(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;
#endif
}

// One evaluation pass over the poles j outside [ex_lo, ex_lo + nex), with
// delta_j = (d_j - d_org) - tau and term_j = z_j^2 / delta_j:
//   S = sum term_j,  A = sum |term_j|,  J = sum j * term_j,
//   dpsi = sum over delta_j < 0 of (z_j/delta_j)^2, dphi likewise delta_j > 0.
// Poles below the root have delta < 0 (LAPACK's psi side) and poles above
// delta > 0 (the phi side), so psi + phi = S and phi - psi = A; and the bound
// LAPACK accumulates from its sequential partial sums,
// |sum_{j<ii} psi_{<=j}| + sum_{j>ii} phi_{>=j}, equals sum_j (j - ii) term_j
// = J - ii*S.  No per-lane index ranges: every lane runs the same loop.
struct SecSums {
    double S, A, dpsi, dphi, J;
};

// Root i (0-based) of the K-pole secular equation, K >= 3.  E provides
// d(j), zsq(j) = z_j^2 and sums(d_org, tau, ex_lo, nex).
template <class Eval>
__host__ __device__ RootState secular_root(const Eval& E, int K, int i, double rho) {
    RootState r;
    r.info = 0;
    r.iters = 0;
    const double rhoinv = 1.0 / rho;
    if (i == K - 1) {
        // ---- last root, in (d_{K-1}, d_{K-1} + rho)
        const double dk = E.d(K - 1), dk1 = E.d(K - 2);
        const double zk2 = E.zsq(K - 1), zk12 = E.zsq(K - 2);
        const double midpt = rho / 2.0;
        double tau, dltlb, dltub;
        {
            const SecSums s0 = E.sums(dk, midpt, K - 2, 2);
            r.iters++;
            const double c = rhoinv + s0.S;
            const double dl_ii = (dk1 - dk) - midpt, dl_n = -midpt;
            const double w = c + zk12 / dl_ii + zk2 / dl_n;
            const double del = dk - dk1;
            const double a = -c * del + zk12 + zk2;
            const double b = zk2 * del;
            if (w <= 0.0) {
                const double temp = zk12 / (dk - dk1 + rho) + zk2 / rho;
                if (c <= temp)
                    tau = rho;
                else if (a < 0.0)
                    tau = 2.0 * b / (sqrt(a * a + 4.0 * b * c) - a);
                else
                    tau = (a + sqrt(a * a + 4.0 * b * c)) / (2.0 * c);
                dltlb = midpt;
                dltub = rho;
            } else {
                if (a < 0.0)
                    tau = 2.0 * b / (sqrt(a * a + 4.0 * b * c) - a);
                else
                    tau = (a + sqrt(a * a + 4.0 * b * c)) / (2.0 * c);
                dltlb = 0.0;
                dltub = midpt;
            }
        }
        double psi, dpsi, phi, dphi, erretm, w;
        auto eval = [&]() {
            const SecSums s = E.sums(dk, tau, K - 1, 1);
            r.iters++;
            psi = s.S;
            dpsi = s.dpsi + s.dphi;
            const double err = s.J - (double)(K - 1) * s.S;
            const double dn = -tau;
            phi = zk2 / dn;
            dphi = phi / dn;
            erretm = 8.0 * (-phi - psi) + err - phi + rhoinv + fabs(tau) * (dpsi + dphi);
            w = rhoinv + phi + psi;
        };
        eval();
        bool conv = false;
        for (int niter = 2; niter <= kSecularMaxIt; niter++) {
            if (fabs(w) <= kEps * erretm) {
                conv = true;
                break;
            }
            if (w <= 0.0)
                dltlb = fmax(dltlb, tau);
            else
                dltub = fmin(dltub, tau);
            const double dn1 = (dk1 - dk) - tau, dn = -tau;
            double c = w - dn1 * dpsi - dn * dphi;
            const double a = (dn1 + dn) * w - dn1 * dn * (dpsi + dphi);
            const double b = dn1 * dn * w;
            if (c < 0.0)
                c = fabs(c);
            double eta;
            if (c == 0.0)
                eta = -w / (dpsi + dphi);
            else if (a >= 0.0)
                eta = (a + sqrt(fabs(a * a - 4.0 * b * c))) / (2.0 * c);
            else
                eta = 2.0 * b / (a - sqrt(fabs(a * a - 4.0 * b * c)));
            if (w * eta > 0.0)
                eta = -w / (dpsi + dphi);
            const double temp = tau + eta;
            if (temp > dltub || temp < dltlb)
                eta = (w < 0.0) ? (dltub - tau) / 2.0 : (dltlb - tau) / 2.0;
            tau += eta;
            eval();
        }
        if (!conv)
            r.info = 1;
        r.lambda = dk + tau;
        r.org = K - 1;
        r.tau = tau;
        r.near1 = (dk1 - dk) - tau;
        r.near2 = -tau;
        return r;
    }

    // ---- interior root, in (d_i, d_{i+1})
    const int ip1 = i + 1;
    const double di = E.d(i), dip1 = E.d(ip1);
    const double zi2 = E.zsq(i), zip12 = E.zsq(ip1);
    const double del = dip1 - di;
    const double midpt = del / 2.0;
    bool orgati;
    double tau, dltlb, dltub;
    {
        const SecSums s0 = E.sums(di, midpt, i, 2);
        r.iters++;
        const double c = rhoinv + s0.S;
        const double w = c + zi2 / (-midpt) + zip12 / (del - midpt);
        if (w > 0.0) {
            orgati = true;
            const double a = c * del + zi2 + zip12;
            const double b = zi2 * del;
            if (a > 0.0)
                tau = 2.0 * b / (a + sqrt(fabs(a * a - 4.0 * b * c)));
            else
                tau = (a - sqrt(fabs(a * a - 4.0 * b * c))) / (2.0 * c);
            dltlb = 0.0;
            dltub = midpt;
        } else {
            orgati = false;
            const double a = c * del - zi2 - zip12;
            const double b = zip12 * del;
            if (a < 0.0)
                tau = 2.0 * b / (a - sqrt(fabs(a * a + 4.0 * b * c)));
            else
                tau = -(a + sqrt(fabs(a * a + 4.0 * b * c))) / (2.0 * c);
            dltlb = -midpt;
            dltub = 0.0;
        }
    }
    const int ii = orgati ? i : ip1;
    const int iim1 = ii - 1, iip1 = ii + 1;
    const double dorg = orgati ? di : dip1;
    const double zii2 = E.zsq(ii);
    // neighbour poles of the three-pole model (only used when 0 < ii < K-1)
    const double dm1 = (ii > 0) ? E.d(iim1) : 0.0;
    const double zm12 = (ii > 0) ? E.zsq(iim1) : 0.0;
    const double dp1 = (iip1 < K) ? E.d(iip1) : 0.0;
    const double zp12 = (iip1 < K) ? E.zsq(iip1) : 0.0;

    double A, S, dpsi, dphi, erretm, w, w0, dw;
    // w0: secular function without the origin pole's term; w: with it
    auto eval = [&](double tau_err) {
        const SecSums s = E.sums(dorg, tau, ii, 1);
        r.iters++;
        S = s.S;
        A = s.A;
        dpsi = s.dpsi;
        dphi = s.dphi;
        const double err = s.J - (double)ii * s.S;
        const double dii = -tau;  // delta at the origin pole
        const double temp = zii2 / dii;
        dw = dpsi + dphi + temp / dii;
        w0 = rhoinv + S;
        w = w0 + temp;
        erretm = 8.0 * A + err + 2.0 * rhoinv + 3.0 * fabs(temp) + tau_err * dw;
    };
    eval(0.0);
    bool swtch3 = orgati ? (w0 < 0.0) : (w0 > 0.0);
    if (ii == 0 || ii == K - 1)
        swtch3 = false;

    auto delta_at = [&](double dj) { return (dj - dorg) - tau; };
    bool conv = false;
    if (fabs(w) <= kEps * erretm) {
        conv = true;
    } else {
        if (w <= 0.0)
            dltlb = fmax(dltlb, tau);
        else
            dltub = fmin(dltub, tau);
        bool swtch = false;
        for (int niter = 2; niter <= kSecularMaxIt; niter++) {
            if (niter > 2) {
                if (fabs(w) <= kEps * erretm) {
                    conv = true;
                    break;
                }
                if (w <= 0.0)
                    dltlb = fmax(dltlb, tau);
                else
                    dltub = fmin(dltub, tau);
            }
            const double dl_i = delta_at(di), dl_ip1 = delta_at(dip1);
            double eta;
            if (!swtch3) {
                double c;
                if (niter == 2 || !swtch) {
                    if (orgati)
                        c = w - dl_ip1 * dw - (di - dip1) * (zi2 / (dl_i * dl_i));
                    else
                        c = w - dl_i * dw - (dip1 - di) * (zip12 / (dl_ip1 * dl_ip1));
                } else {
                    const double dlo = delta_at(dorg);
                    const double temp = zii2 / (dlo * dlo);
                    if (orgati)
                        dpsi += temp;
                    else
                        dphi += temp;
                    c = w - dl_i * dpsi - dl_ip1 * dphi;
                }
                double a = (dl_i + dl_ip1) * w - dl_i * dl_ip1 * dw;
                const double b = dl_i * dl_ip1 * w;
                if (c == 0.0) {
                    if (a == 0.0) {
                        if (niter == 2 || !swtch) {
                            if (orgati)
                                a = zi2 + dl_ip1 * dl_ip1 * (dpsi + dphi);
                            else
                                a = zip12 + dl_i * dl_i * (dpsi + dphi);
                        } else {
                            a = dl_i * dl_i * dpsi + dl_ip1 * dl_ip1 * dphi;
                        }
                    }
                    eta = b / a;
                } else if (a <= 0.0) {
                    eta = (a - sqrt(fabs(a * a - 4.0 * b * c))) / (2.0 * c);
                } else {
                    eta = 2.0 * b / (a + sqrt(fabs(a * a - 4.0 * b * c)));
                }
            } else {
                const double dl_m = delta_at(dm1), dl_p = delta_at(dp1);
                const double temp = rhoinv + S;
                double c, zz[3];
                if (niter > 2 && swtch) {
                    c = temp - dl_m * dpsi - dl_p * dphi;
                    zz[0] = dl_m * dl_m * dpsi;
                    zz[2] = dl_p * dl_p * dphi;
                } else if (orgati) {
                    const double t1 = zm12 / (dl_m * dl_m);
                    c = temp - dl_p * (dpsi + dphi) - (dm1 - dp1) * t1;
                    zz[0] = zm12;
                    zz[2] = dl_p * dl_p * ((dpsi - t1) + dphi);
                } else {
                    const double t1 = zp12 / (dl_p * dl_p);
                    c = temp - dl_m * (dpsi + dphi) - (dp1 - dm1) * t1;
                    zz[0] = dl_m * dl_m * (dpsi + (dphi - t1));
                    zz[2] = zp12;
                }
                zz[1] = zii2;
                const double dd3[3] = {dl_m, delta_at(dorg), dl_p};
                if (secular_three(niter, orgati, c, dd3, zz, w, &eta)) {
                    r.info = 1;
                    break;
                }
            }
            if (w * eta >= 0.0)
                eta = -w / dw;
            const double temp = tau + eta;
            if (temp > dltub || temp < dltlb)
                eta = (w < 0.0) ? (dltub - tau) / 2.0 : (dltlb - tau) / 2.0;
            const double prew = w;
            tau += eta;
            eval(fabs(tau));
            if (niter == 2) {
                swtch = orgati ? (-w > fabs(prew) / 10.0) : (w > fabs(prew) / 10.0);
            } else if (w * prew > 0.0 && fabs(w) > fabs(prew) / 10.0) {
                swtch = !swtch;
            }
        }
    }
    if (!conv && r.info == 0)
        r.info = 1;
    r.lambda = dorg + tau;
    r.org = ii;
    r.tau = tau;
    r.near1 = (di - dorg) - tau;
    r.near2 = (dip1 - dorg) - tau;
    return r;
}

}  // namespace brdc
