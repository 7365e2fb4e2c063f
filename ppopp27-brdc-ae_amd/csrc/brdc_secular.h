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

// Division in the per-root step logic: a * rcp_nr(b) on the device (within
// one ulp of a/b, about half the latency of the IEEE division sequence).  It
// only perturbs the iteration path; the stopping test and the converged root
// keep their accuracy.  The host build divides exactly.
__host__ __device__ __forceinline__ double fdiv(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__) || defined(BRDC_HOST_DEVICE_ARITH)
    return a * rcp_nr(b);
#else
    return a / b;
#endif
}

// A denominator used several times: fdiv(a, d) == Divisor(d).div(a) in
// every build (the reciprocal formed once on the device, exact division on
// the host).
struct Divisor {
    double v;
    __host__ __device__ explicit Divisor(double d) {
#if defined(__HIP_DEVICE_COMPILE__) || defined(BRDC_HOST_DEVICE_ARITH)
        v = rcp_nr(d);
#else
        v = d;
#endif
    }
    __host__ __device__ double div(double a) const {
#if defined(__HIP_DEVICE_COMPILE__) || defined(BRDC_HOST_DEVICE_ARITH)
        return a * v;
#else
        return a / v;
#endif
    }
};

// The root-of-a-quadratic step shared by the iterations:
//   c == 0:       b / a
//   else a <= 0:  (a - sq) / (2c)
//   else:         2b / (a + sq),     sq = sqrt(|a^2 - 4bc|)
// with one square root and one division whichever case applies (lanes of a
// wave in different cases then share them).
__host__ __device__ __forceinline__ double quad_step(double a, double b, double c) {
    const double sq = sqrt(fabs(a * a - 4.0 * b * c));
    const bool cz = (c == 0.0), an = (a <= 0.0);
    const double num = cz ? b : (an ? a - sq : 2.0 * b);
    const double den = cz ? a : (an ? 2.0 * c : a + sq);
    return fdiv(num, den);
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
// between dd1, dd2 (orgati) or dd0, dd1, starting from f(0) = finit.  The
// three poles are named scalars, not an array: a choice between two array
// entries on orgati would become a computed index, which puts the array in
// scratch memory.
__host__ __device__ __forceinline__ int secular_three(int kniter, bool orgati, double rho,
                                                     double dd0, double dd1, double dd2,
                                                     double zz0, double zz1, double zz2,
                                                     double finit, double* tau_out) {
    double lbd, ubd, tau = 0.0, a, b, c, eta, f, df, ddf, fc, erretm, temp;
    int info = 0;
    bool scale = false;
    double sclinv = 1.0;
    lbd = orgati ? dd1 : dd0;
    ubd = orgati ? dd2 : dd1;
    if (finit < 0.0)
        lbd = 0.0;
    else
        ubd = 0.0;
    if (kniter == 2) {
        if (orgati) {
            temp = (dd2 - dd1) / 2.0;
            c = rho + fdiv(zz0, (dd0 - dd1) - temp);
            a = c * (dd1 + dd2) + zz1 + zz2;
            b = c * dd1 * dd2 + zz1 * dd2 + zz2 * dd1;
        } else {
            temp = (dd0 - dd1) / 2.0;
            c = rho + fdiv(zz2, (dd2 - dd1) - temp);
            a = c * (dd0 + dd1) + zz0 + zz1;
            b = c * dd0 * dd1 + zz0 * dd1 + zz1 * dd0;
        }
        temp = rcp_nr(dev_max3(a, b, c));
        a *= temp;
        b *= temp;
        c *= temp;
        if (c == 0.0)
            tau = fdiv(b, a);
        else if (a <= 0.0)
            tau = fdiv(a - sqrt(fabs(a * a - 4.0 * b * c)), 2.0 * c);
        else
            tau = fdiv(2.0 * b, a + sqrt(fabs(a * a - 4.0 * b * c)));
        if (tau < lbd || tau > ubd)
            tau = (lbd + ubd) / 2.0;
        if (dd0 == tau || dd1 == tau || dd2 == tau) {
            tau = 0.0;
        } else {
            temp = finit + fdiv(tau * zz0, dd0 * (dd0 - tau)) +
                   fdiv(tau * zz1, dd1 * (dd1 - tau)) + fdiv(tau * zz2, dd2 * (dd2 - tau));
            if (temp <= 0.0)
                lbd = tau;
            else
                ubd = tau;
            if (fabs(finit) <= fabs(temp))
                tau = 0.0;
        }
    }
    // rescale when a pole is within 2^-340 of tau, so 1/t^3 cannot overflow
    temp = orgati ? fmin(fabs(dd1 - tau), fabs(dd2 - tau)) : fmin(fabs(dd0 - tau), fabs(dd1 - tau));
    double ds0 = dd0, ds1 = dd1, ds2 = dd2, zs0 = zz0, zs1 = zz1, zs2 = zz2;
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
        ds0 = dd0 * sclfac;
        ds1 = dd1 * sclfac;
        ds2 = dd2 * sclfac;
        zs0 = zz0 * sclfac;
        zs1 = zz1 * sclfac;
        zs2 = zz2 * sclfac;
        tau *= sclfac;
        lbd *= sclfac;
        ubd *= sclfac;
    }
    // the poles are fixed through the iteration: one reciprocal each
    const Divisor rd0(ds0), rd1(ds1), rd2(ds2);
    fc = df = ddf = 0.0;
    auto term0 = [&](double ds, double zs, const Divisor& rd) {
        const double r = rcp_nr(ds - tau);
        const double t1 = zs * r, t2 = t1 * r, t3 = t2 * r;
        fc += rd.div(t1);
        df += t2;
        ddf += t3;
    };
    term0(ds0, zs0, rd0);
    term0(ds1, zs1, rd1);
    term0(ds2, zs2, rd2);
    f = finit + tau * fc;
    if (fabs(f) <= 0.0)
        goto done;
    if (f <= 0.0)
        lbd = tau;
    else
        ubd = tau;
    for (int niter = 2; niter <= kLaed6MaxIt; niter++) {
        const double t1 = (orgati ? ds1 : ds0) - tau;
        const double t2 = (orgati ? ds2 : ds1) - tau;
        a = (t1 + t2) * f - t1 * t2 * df;
        b = t1 * t2 * f;
        c = f - (t1 + t2) * df + t1 * t2 * ddf;
        temp = rcp_nr(dev_max3(a, b, c));
        a *= temp;
        b *= temp;
        c *= temp;
        eta = quad_step(a, b, c);
        if (f * eta >= 0.0)
            eta = -fdiv(f, df);
        tau += eta;
        if (tau < lbd || tau > ubd)
            tau = (lbd + ubd) / 2.0;
        fc = erretm = df = ddf = 0.0;
        if (ds0 - tau == 0.0 || ds1 - tau == 0.0 || ds2 - tau == 0.0)
            goto done;
        auto term = [&](double ds, double zs, const Divisor& rd) {
            const double r = rcp_nr(ds - tau);
            const double u1 = zs * r, u2 = u1 * r, u3 = u2 * r, u4 = rd.div(u1);
            fc += u4;
            erretm += fabs(u4);
            df += u2;
            ddf += u3;
        };
        term(ds0, zs0, rd0);
        term(ds1, zs1, rd1);
        term(ds2, zs2, rd2);
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

// Root i (0-based) of the K-pole secular equation, K >= 3, as a state
// machine around a single pole-pass call site: every unconverged root of a
// wavefront streams the poles in the same pass, whatever phase it is in.
// E provides d(j), zsq(j) = z_j^2 and sums(d_org, tau, ex_lo, nex).
//
// Pass sequence (LAPACK DLAED4): E0 at the interval midpoint gives the
// initial guess; E1 evaluates it; after each E_k (k >= 1) the next step
// number niter = k+1 is tested (|w| <= eps*erretm), failing past 30 steps,
// then the step is taken and E_{k+1} evaluates the new iterate.
template <class Eval>
struct SecularSolve {
    // problem
    int K, i;
    double rho, rhoinv;
    bool last;
    // interval / origin
    double di, dip1, zi2, zip12, del, midpt;  // interior: d_i, d_{i+1}; last: d_{K-1}, d_{K-2}
    bool orgati;
    int ii;
    double dorg, zii2, dm1, zm12, dp1, zp12;
    // iterate
    double tau, dltlb, dltub;
    double S, A, dpsi, dphi, erretm, w, w0, dw, psi, phi, prew;
    bool swtch3, swtch;
    int k;  // passes done
    // next pass
    double ev_org, ev_tau;
    int ev_exlo, ev_nex;
    RootState out;

    __host__ __device__ void init(const Eval& E, int K_, int i_, double rho_) {
        K = K_;
        i = i_;
        rho = rho_;
        rhoinv = fdiv(1.0, rho);
        last = (i == K - 1);
        out.info = 0;
        out.iters = 0;
        k = 0;
        swtch3 = swtch = false;
        if (last) {
            di = E.d(K - 1);   // d_{K-1}
            dip1 = E.d(K - 2); // d_{K-2}
            zi2 = E.zsq(K - 1);
            zip12 = E.zsq(K - 2);
            midpt = rho * 0.5;
            ev_org = di;
            ev_tau = midpt;
            ev_exlo = K - 2;
            ev_nex = 2;
        } else {
            di = E.d(i);
            dip1 = E.d(i + 1);
            zi2 = E.zsq(i);
            zip12 = E.zsq(i + 1);
            del = dip1 - di;
            midpt = del * 0.5;
            ev_org = di;
            ev_tau = midpt;
            ev_exlo = i;
            ev_nex = 2;
        }
    }

    __host__ __device__ double delta_at(double dj) const { return (dj - dorg) - tau; }

    __host__ __device__ void finish(bool converged) {
        if (!converged && out.info == 0) out.info = 1;
        out.lambda = dorg + tau;
        out.org = ii;
        out.tau = tau;
        if (last) {
            out.near1 = (dip1 - dorg) - tau;
            out.near2 = -tau;
        } else {
            out.near1 = (di - dorg) - tau;
            out.near2 = (dip1 - dorg) - tau;
        }
    }

    // Consume pass k's sums; returns true when the root is finished.
    __host__ __device__ bool advance(const Eval& E, const SecSums& s) {
        out.iters++;
        if (k == 0) {
            k = 1;
            return initial_guess(E, s);
        }
        if (last) {
            psi = s.S;
            dpsi = s.dpsi + s.dphi;
            const double err = s.J - (double)(K - 1) * s.S;
            const double dn = -tau;
            phi = fdiv(zi2, dn);
            dphi = fdiv(phi, dn);
            erretm = 8.0 * (-phi - psi) + err - phi + rhoinv + fabs(tau) * (dpsi + dphi);
            w = rhoinv + phi + psi;
        } else {
            S = s.S;
            A = s.A;
            dpsi = s.dpsi;
            dphi = s.dphi;
            const double err = s.J - (double)ii * s.S;
            const double dii = -tau;
            const Divisor rdii(dii);
            const double temp = rdii.div(zii2);
            dw = dpsi + dphi + rdii.div(temp);
            w0 = rhoinv + S;
            w = w0 + temp;
            erretm = 8.0 * A + err + 2.0 * rhoinv + 3.0 * fabs(temp) + (k >= 2 ? fabs(tau) : 0.0) * dw;
            if (k == 1) {
                swtch3 = orgati ? (w0 < 0.0) : (w0 > 0.0);
                if (ii == 0 || ii == K - 1) swtch3 = false;
            } else if (k == 2) {
                swtch = orgati ? (-w > fabs(prew) / 10.0) : (w > fabs(prew) / 10.0);
            } else if (w * prew > 0.0 && fabs(w) > fabs(prew) / 10.0) {
                swtch = !swtch;
            }
        }
        const int niter = k + 1;
        if (niter > kSecularMaxIt) {
            finish(false);
            return true;
        }
        if (fabs(w) <= kEps * erretm) {
            finish(true);
            return true;
        }
        if (w <= 0.0)
            dltlb = fmax(dltlb, tau);
        else
            dltub = fmin(dltub, tau);
        double eta;
        if (last) {
            const double dn1 = (dip1 - di) - tau, dn = -tau;
            double c = w - dn1 * dpsi - dn * dphi;
            const double a = (dn1 + dn) * w - dn1 * dn * (dpsi + dphi);
            const double b = dn1 * dn * w;
            if (c < 0.0) c = fabs(c);
            if (c == 0.0)
                eta = -fdiv(w, dpsi + dphi);
            else if (a >= 0.0)
                eta = fdiv(a + sqrt(fabs(a * a - 4.0 * b * c)), 2.0 * c);
            else
                eta = fdiv(2.0 * b, a - sqrt(fabs(a * a - 4.0 * b * c)));
            if (w * eta > 0.0) eta = -fdiv(w, dpsi + dphi);
        } else {
            const double dl_i = delta_at(di), dl_ip1 = delta_at(dip1);
            if (!swtch3) {
                double c;
                if (niter == 2 || !swtch) {
                    if (orgati)
                        c = w - dl_ip1 * dw - (di - dip1) * fdiv(zi2, dl_i * dl_i);
                    else
                        c = w - dl_i * dw - (dip1 - di) * fdiv(zip12, dl_ip1 * dl_ip1);
                } else {
                    const double dlo = delta_at(dorg);
                    const double temp = fdiv(zii2, dlo * dlo);
                    if (orgati)
                        dpsi += temp;
                    else
                        dphi += temp;
                    c = w - dl_i * dpsi - dl_ip1 * dphi;
                }
                double a = (dl_i + dl_ip1) * w - dl_i * dl_ip1 * dw;
                const double b = dl_i * dl_ip1 * w;
                if (c == 0.0 && a == 0.0) {
                    if (niter == 2 || !swtch) {
                        if (orgati)
                            a = zi2 + dl_ip1 * dl_ip1 * (dpsi + dphi);
                        else
                            a = zip12 + dl_i * dl_i * (dpsi + dphi);
                    } else {
                        a = dl_i * dl_i * dpsi + dl_ip1 * dl_ip1 * dphi;
                    }
                }
                eta = quad_step(a, b, c);
            } else {
                const double dl_m = delta_at(dm1), dl_p = delta_at(dp1);
                const double temp = rhoinv + S;
                double c, zz0, zz2;
                if (niter > 2 && swtch) {
                    c = temp - dl_m * dpsi - dl_p * dphi;
                    zz0 = dl_m * dl_m * dpsi;
                    zz2 = dl_p * dl_p * dphi;
                } else if (orgati) {
                    const double t1 = fdiv(zm12, dl_m * dl_m);
                    c = temp - dl_p * (dpsi + dphi) - (dm1 - dp1) * t1;
                    zz0 = zm12;
                    zz2 = dl_p * dl_p * ((dpsi - t1) + dphi);
                } else {
                    const double t1 = fdiv(zp12, dl_p * dl_p);
                    c = temp - dl_m * (dpsi + dphi) - (dp1 - dm1) * t1;
                    zz0 = dl_m * dl_m * (dpsi + (dphi - t1));
                    zz2 = zp12;
                }
                if (secular_three(niter, orgati, c, dl_m, delta_at(dorg), dl_p, zz0, zii2, zz2, w,
                                  &eta)) {
                    out.info = 1;
                    finish(false);
                    return true;
                }
            }
            if (w * eta >= 0.0) eta = -fdiv(w, dw);
        }
        const double temp = tau + eta;
        if (temp > dltub || temp < dltlb)
            eta = (w < 0.0) ? (dltub - tau) * 0.5 : (dltlb - tau) * 0.5;
        prew = w;
        tau += eta;
        k++;
        ev_tau = tau;
        return false;
    }

    __host__ __device__ bool initial_guess(const Eval& E, const SecSums& s0) {
        if (last) {
            const double c = rhoinv + s0.S;
            const double dl_ii = (dip1 - di) - midpt, dl_n = -midpt;
            const double w_ = c + fdiv(zip12, dl_ii) + fdiv(zi2, dl_n);
            const double dl = di - dip1;
            const double a = -c * dl + zip12 + zi2;
            const double b = zi2 * dl;
            if (w_ <= 0.0) {
                const double temp = fdiv(zip12, di - dip1 + rho) + fdiv(zi2, rho);
                if (c <= temp)
                    tau = rho;
                else if (a < 0.0)
                    tau = fdiv(2.0 * b, sqrt(a * a + 4.0 * b * c) - a);
                else
                    tau = fdiv(a + sqrt(a * a + 4.0 * b * c), 2.0 * c);
                dltlb = midpt;
                dltub = rho;
            } else {
                if (a < 0.0)
                    tau = fdiv(2.0 * b, sqrt(a * a + 4.0 * b * c) - a);
                else
                    tau = fdiv(a + sqrt(a * a + 4.0 * b * c), 2.0 * c);
                dltlb = 0.0;
                dltub = midpt;
            }
            ii = K - 1;
            dorg = di;
            // the interior roots' fields, unused by the last root: every path
            // sets the same fields (a field set on one path only can keep the
            // solver state out of registers)
            orgati = false;
            zii2 = zi2;
            dm1 = dip1;
            zm12 = zip12;
            dp1 = 0.0;
            zp12 = 0.0;
            ev_org = dorg;
            ev_tau = tau;
            ev_exlo = K - 1;
            ev_nex = 1;
            return false;
        }
        const double c = rhoinv + s0.S;
        const double w_ = c + fdiv(zi2, -midpt) + fdiv(zip12, del - midpt);
        if (w_ > 0.0) {
            orgati = true;
            const double a = c * del + zi2 + zip12;
            const double b = zi2 * del;
            if (a > 0.0)
                tau = fdiv(2.0 * b, a + sqrt(fabs(a * a - 4.0 * b * c)));
            else
                tau = fdiv(a - sqrt(fabs(a * a - 4.0 * b * c)), 2.0 * c);
            dltlb = 0.0;
            dltub = midpt;
        } else {
            orgati = false;
            const double a = c * del - zi2 - zip12;
            const double b = zip12 * del;
            if (a < 0.0)
                tau = fdiv(2.0 * b, a - sqrt(fabs(a * a + 4.0 * b * c)));
            else
                tau = -fdiv(a + sqrt(fabs(a * a + 4.0 * b * c)), 2.0 * c);
            dltlb = -midpt;
            dltub = 0.0;
        }
        ii = orgati ? i : i + 1;
        dorg = orgati ? di : dip1;
        zii2 = orgati ? zi2 : zip12;
        dm1 = (ii > 0) ? E.d(ii - 1) : 0.0;
        zm12 = (ii > 0) ? E.zsq(ii - 1) : 0.0;
        dp1 = (ii + 1 < K) ? E.d(ii + 1) : 0.0;
        zp12 = (ii + 1 < K) ? E.zsq(ii + 1) : 0.0;
        ev_org = dorg;
        ev_tau = tau;
        ev_exlo = ii;
        ev_nex = 1;
        return false;
    }
};

template <class Eval>
__host__ __device__ RootState secular_root(const Eval& E, int K, int i, double rho) {
    SecularSolve<Eval> st;
    st.init(E, K, i, rho);
    for (;;) {
        const SecSums s = E.sums(st.ev_org, st.ev_tau, st.ev_exlo, st.ev_nex);
        if (st.advance(E, s)) break;
    }
    return st.out;
}

}  // namespace brdc
