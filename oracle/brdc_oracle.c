/*
 * brdc_oracle.c -- CPU restatement of the reference BR values-only D&C path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the
 * CPU baseline ("port") timed by bench.py.  The product never links it.
 *
 * Reference routines restated here (file:line in the reference tree):
 *   DSTEDC_DC_BR         cpu/src/dstedc_dc_br.f:7-151    -> brdc_oracle_dstebr_ex
 *   BRDC_WPROD_NTHREADS  cpu/src/dstedc_dc_br.f:153-213  -> brdc_oracle_wprod_nthreads
 *   DLAED0_BR            cpu/src/dlaed0_br.f:8-308       -> laed0_br
 *   DLAED0_MERGE_PAIR    cpu/src/dlaed0_br.f:314-350     -> (inlined in laed0_br)
 *   DSTEQR_BOUNDARY_BR   cpu/src/dlaed0_br.f:354-371     -> brdc_oracle_leaf
 *   DLAED7_BR            cpu/src/dlaed7_br.f:7-167       -> laed7_br
 *   DLAED8_BR            cpu/src/dlaed7_br.f:175-381     -> laed8_br
 *   DLAED9_BR            cpu/src/dlaed9_br.f:11-393      -> laed9_br
 *   DLAED4_BR            cpu/src/dlaed4_br.f:15-109      -> brdc_oracle_laed4_br
 * Third-party LAPACK routines the reference calls (absent from the reference
 * tree; Reference-LAPACK 3.12, as bundled by OpenBLAS 0.3.28), restated from
 * their published algorithms:
 *   DLAED4  R.-C. Li, "Solving secular equations stably and efficiently"
 *           (LAWN 89, 1993): middle-way rational interpolation, with
 *   DLAED6  Gragg-Thornton-Warner cubic three-pole step, and
 *   DLAED5  the closed form for two poles.
 *   DSTEQR / DSTERF roles: implicit QL with Wilkinson shift (Bowdler, Martin,
 *           Reinsch, Wilkinson, Numer. Math. 11, 1968), eigenvalues sorted
 *           ascending; only the first and last eigenvector rows are kept.
 *   DLAMRG  two-list merge permutation.
 * Pinned by tests/test_oracle.py and tests/test_secular_host.py against scipy's OpenBLAS (dsterf,
 * dlaed4/5/6 golden vectors), dense eigh, closed-form spectra and invariants.
 */
#include "brdc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* DLAMCH('Epsilon') with rounding = 2^-53; DLAMCH('Safe minimum') = 2^-1022. */
#define ORC_EPS 0x1p-53
#define ORC_SAFMIN 0x1p-1022
#define LAED4_MAXIT 30
#define LAED6_MAXIT 40
#define REF_SMLSIZ 25 /* ILAENV(9, 'DSTEDC', ...) */

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* DLAPY2: sqrt(x^2 + y^2) without avoidable overflow. */
static double lapy2(double x, double y)
{
    double xa = fabs(x), ya = fabs(y);
    double w = dmax(xa, ya), z = dmin(xa, ya);
    if (z == 0.0 || w > 1.79e308)
        return w;
    double q = z / w;
    return w * sqrt(1.0 + q * q);
}

/* ------------------------------------------------------------------------- */
/* DLAMRG: index[0..n1+n2) lists positions of a[] in merged ascending order.
 * List 1 is a[0..n1) read forward (s1 > 0) or backward; list 2 is
 * a[n1..n1+n2) likewise.  Ties take list 1 first. */
void brdc_oracle_lamrg(int n1, int n2, const double *a, int s1, int s2,
                       int *index)
{
    int r1 = n1, r2 = n2;
    int p1 = s1 > 0 ? 0 : n1 - 1;
    int p2 = s2 > 0 ? n1 : n1 + n2 - 1;
    int out = 0;
    while (r1 > 0 && r2 > 0) {
        if (a[p1] <= a[p2]) {
            index[out++] = p1;
            p1 += s1 > 0 ? 1 : -1;
            r1--;
        } else {
            index[out++] = p2;
            p2 += s2 > 0 ? 1 : -1;
            r2--;
        }
    }
    while (r1-- > 0) {
        index[out++] = p1;
        p1 += s1 > 0 ? 1 : -1;
    }
    while (r2-- > 0) {
        index[out++] = p2;
        p2 += s2 > 0 ? 1 : -1;
    }
}

/* ------------------------------------------------------------------------- */
/* DLAED5: both roots of the 2x2 problem diag(d) + rho z z^T in closed form.
 * delta receives the normalised eigenvector of root i. */
int brdc_oracle_laed5(int i, const double *d, const double *z, double *delta,
                      double rho, double *dlam)
{
    const double del = d[1] - d[0];
    const double zz0 = z[0] * z[0], zz1 = z[1] * z[1];
    double b, c, tau;
    if (i == 0) {
        double w = 1.0 + 2.0 * rho * (zz1 - zz0) / del;
        if (w > 0.0) {
            /* root nearer d[0] */
            b = del + rho * (zz0 + zz1);
            c = rho * zz0 * del;
            tau = 2.0 * c / (b + sqrt(fabs(b * b - 4.0 * c)));
            *dlam = d[0] + tau;
            delta[0] = -z[0] / tau;
            delta[1] = z[1] / (del - tau);
        } else {
            b = -del + rho * (zz0 + zz1);
            c = rho * zz1 * del;
            if (b > 0.0)
                tau = -2.0 * c / (b + sqrt(b * b + 4.0 * c));
            else
                tau = (b - sqrt(b * b + 4.0 * c)) / 2.0;
            *dlam = d[1] + tau;
            delta[0] = -z[0] / (del + tau);
            delta[1] = -z[1] / tau;
        }
    } else {
        b = -del + rho * (zz0 + zz1);
        c = rho * zz1 * del;
        if (b > 0.0)
            tau = (b + sqrt(b * b + 4.0 * c)) / 2.0;
        else
            tau = 2.0 * c / (-b + sqrt(b * b + 4.0 * c));
        *dlam = d[1] + tau;
        delta[0] = -z[0] / (del + tau);
        delta[1] = -z[1] / tau;
    }
    double nrm = sqrt(delta[0] * delta[0] + delta[1] * delta[1]);
    delta[0] /= nrm;
    delta[1] /= nrm;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* DLAED6: the root tau of g(t) = rho + sum_k z_k/(d_k - t) (three poles) that
 * lies between the two poles on the side given by orgati, starting from
 * g(0) = finit.  Gragg's cubically convergent scheme with a bracketing
 * safeguard; inputs are rescaled when a pole is close enough to t for 1/t^3 to
 * overflow. */
int brdc_oracle_laed6(int kniter, int orgati, double rho, const double *d,
                      const double *z, double finit, double *tau_out)
{
    double lbd, ubd, tau = 0.0;
    double a, b, c, eta, f, df, ddf, fc, erretm, temp;
    double dscale[3], zscale[3];
    int niter, info = 0, scale = 0;
    double sclinv = 1.0;

    if (orgati) {
        lbd = d[1];
        ubd = d[2];
    } else {
        lbd = d[0];
        ubd = d[1];
    }
    if (finit < 0.0)
        lbd = 0.0;
    else
        ubd = 0.0;

    niter = 1;
    if (kniter == 2) {
        /* initial guess from the two nearest poles with the third frozen */
        if (orgati) {
            temp = (d[2] - d[1]) / 2.0;
            c = rho + z[0] / ((d[0] - d[1]) - temp);
            a = c * (d[1] + d[2]) + z[1] + z[2];
            b = c * d[1] * d[2] + z[1] * d[2] + z[2] * d[1];
        } else {
            temp = (d[0] - d[1]) / 2.0;
            c = rho + z[2] / ((d[2] - d[1]) - temp);
            a = c * (d[0] + d[1]) + z[0] + z[1];
            b = c * d[0] * d[1] + z[0] * d[1] + z[1] * d[0];
        }
        temp = dmax(fabs(a), dmax(fabs(b), fabs(c)));
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
        if (d[0] == tau || d[1] == tau || d[2] == tau) {
            tau = 0.0;
        } else {
            temp = finit + tau * z[0] / (d[0] * (d[0] - tau)) +
                   tau * z[1] / (d[1] * (d[1] - tau)) +
                   tau * z[2] / (d[2] * (d[2] - tau));
            if (temp <= 0.0)
                lbd = tau;
            else
                ubd = tau;
            if (fabs(finit) <= fabs(temp))
                tau = 0.0;
        }
    }

    /* scaling thresholds: 2^(int(log2(safmin)/3)) and its square */
    const double small1 = 0x1p-340, sminv1 = 0x1p340;
    const double small2 = 0x1p-680, sminv2 = 0x1p680;
    if (orgati)
        temp = dmin(fabs(d[1] - tau), fabs(d[2] - tau));
    else
        temp = dmin(fabs(d[0] - tau), fabs(d[1] - tau));
    if (temp <= small1) {
        double sclfac;
        scale = 1;
        if (temp <= small2) {
            sclfac = sminv2;
            sclinv = small2;
        } else {
            sclfac = sminv1;
            sclinv = small1;
        }
        for (int k = 0; k < 3; k++) {
            dscale[k] = d[k] * sclfac;
            zscale[k] = z[k] * sclfac;
        }
        tau *= sclfac;
        lbd *= sclfac;
        ubd *= sclfac;
    } else {
        for (int k = 0; k < 3; k++) {
            dscale[k] = d[k];
            zscale[k] = z[k];
        }
    }

    fc = df = ddf = 0.0;
    for (int k = 0; k < 3; k++) {
        double r = 1.0 / (dscale[k] - tau);
        double t1 = zscale[k] * r;
        double t2 = t1 * r;
        double t3 = t2 * r;
        fc += t1 / dscale[k];
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

    for (niter = niter + 1; niter <= LAED6_MAXIT; niter++) {
        double t1, t2;
        if (orgati) {
            t1 = dscale[1] - tau;
            t2 = dscale[2] - tau;
        } else {
            t1 = dscale[0] - tau;
            t2 = dscale[1] - tau;
        }
        a = (t1 + t2) * f - t1 * t2 * df;
        b = t1 * t2 * f;
        c = f - (t1 + t2) * df + t1 * t2 * ddf;
        temp = dmax(fabs(a), dmax(fabs(b), fabs(c)));
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
            if (dscale[k] - tau == 0.0)
                goto done;
            double r = 1.0 / (dscale[k] - tau);
            double u1 = zscale[k] * r;
            double u2 = u1 * r;
            double u3 = u2 * r;
            double u4 = u1 / dscale[k];
            fc += u4;
            erretm += fabs(u4);
            df += u2;
            ddf += u3;
        }
        f = finit + tau * fc;
        erretm = 8.0 * (fabs(finit) + fabs(tau) * erretm) + fabs(tau) * df;
        if (fabs(f) <= 4.0 * ORC_EPS * erretm ||
            (ubd - lbd) <= 4.0 * ORC_EPS * fabs(tau))
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

/* ------------------------------------------------------------------------- */
/* DLAED4, last root (i = n-1): the root in (d[n-1], d[n-1] + rho). */
static void laed4_eval_last(int n, const double *z, const double *delta,
                            double rhoinv, double tau, double *psi,
                            double *dpsi, double *phi, double *dphi,
                            double *erretm, double *w)
{
    const int ii = n - 2;
    double ps = 0.0, dps = 0.0, err = 0.0;
    for (int j = 0; j <= ii; j++) {
        double t = z[j] / delta[j];
        ps += z[j] * t;
        dps += t * t;
        err += ps;
    }
    err = fabs(err);
    double t = z[n - 1] / delta[n - 1];
    double ph = z[n - 1] * t;
    double dph = t * t;
    err = 8.0 * (-ph - ps) + err - ph + rhoinv + fabs(tau) * (dps + dph);
    *psi = ps;
    *dpsi = dps;
    *phi = ph;
    *dphi = dph;
    *erretm = err;
    *w = rhoinv + ph + ps;
}

static int laed4_last(int n, const double *d, const double *z, double *delta,
                      double rho, double *dlam)
{
    const int i = n - 1, ii = n - 2;
    const double rhoinv = 1.0 / rho;
    double midpt = rho / 2.0;
    double psi, dpsi, phi, dphi, erretm, w, c, a, b, tau, eta, temp;
    double dltlb, dltub, del;

    for (int j = 0; j < n; j++)
        delta[j] = (d[j] - d[i]) - midpt;
    psi = 0.0;
    for (int j = 0; j < n - 2; j++)
        psi += z[j] * z[j] / delta[j];
    c = rhoinv + psi;
    w = c + z[ii] * z[ii] / delta[ii] + z[n - 1] * z[n - 1] / delta[n - 1];

    if (w <= 0.0) {
        temp = z[n - 2] * z[n - 2] / (d[n - 1] - d[n - 2] + rho) +
               z[n - 1] * z[n - 1] / rho;
        if (c <= temp) {
            tau = rho;
        } else {
            del = d[n - 1] - d[n - 2];
            a = -c * del + z[n - 2] * z[n - 2] + z[n - 1] * z[n - 1];
            b = z[n - 1] * z[n - 1] * del;
            if (a < 0.0)
                tau = 2.0 * b / (sqrt(a * a + 4.0 * b * c) - a);
            else
                tau = (a + sqrt(a * a + 4.0 * b * c)) / (2.0 * c);
        }
        dltlb = midpt;
        dltub = rho;
    } else {
        del = d[n - 1] - d[n - 2];
        a = -c * del + z[n - 2] * z[n - 2] + z[n - 1] * z[n - 1];
        b = z[n - 1] * z[n - 1] * del;
        if (a < 0.0)
            tau = 2.0 * b / (sqrt(a * a + 4.0 * b * c) - a);
        else
            tau = (a + sqrt(a * a + 4.0 * b * c)) / (2.0 * c);
        dltlb = 0.0;
        dltub = midpt;
    }

    for (int j = 0; j < n; j++)
        delta[j] = (d[j] - d[i]) - tau;
    laed4_eval_last(n, z, delta, rhoinv, tau, &psi, &dpsi, &phi, &dphi,
                    &erretm, &w);

    /* iteration 1 is the initial guess; each of iterations 2..MAXIT tests
     * convergence and then takes one step */
    for (int niter = 2; niter <= LAED4_MAXIT; niter++) {
        if (fabs(w) <= ORC_EPS * erretm) {
            *dlam = d[i] + tau;
            return 0;
        }
        if (w <= 0.0)
            dltlb = dmax(dltlb, tau);
        else
            dltub = dmin(dltub, tau);

        c = w - delta[n - 2] * dpsi - delta[n - 1] * dphi;
        a = (delta[n - 2] + delta[n - 1]) * w -
            delta[n - 2] * delta[n - 1] * (dpsi + dphi);
        b = delta[n - 2] * delta[n - 1] * w;
        if (c < 0.0)
            c = fabs(c);
        if (c == 0.0)
            eta = -w / (dpsi + dphi);
        else if (a >= 0.0)
            eta = (a + sqrt(fabs(a * a - 4.0 * b * c))) / (2.0 * c);
        else
            eta = 2.0 * b / (a - sqrt(fabs(a * a - 4.0 * b * c)));
        if (w * eta > 0.0)
            eta = -w / (dpsi + dphi);
        temp = tau + eta;
        if (temp > dltub || temp < dltlb) {
            if (w < 0.0)
                eta = (dltub - tau) / 2.0;
            else
                eta = (dltlb - tau) / 2.0;
        }
        for (int j = 0; j < n; j++)
            delta[j] -= eta;
        tau += eta;
        laed4_eval_last(n, z, delta, rhoinv, tau, &psi, &dpsi, &phi, &dphi,
                        &erretm, &w);
    }
    *dlam = d[i] + tau;
    return 1;
}

/* DLAED4, interior root i < n-1: the root in (d[i], d[i+1]). */
static void laed4_eval_mid(int n, int ii, const double *z,
                           const double *delta, double *psi, double *dpsi,
                           double *phi, double *dphi, double *erretm)
{
    double ps = 0.0, dps = 0.0, err = 0.0;
    for (int j = 0; j < ii; j++) {
        double t = z[j] / delta[j];
        ps += z[j] * t;
        dps += t * t;
        err += ps;
    }
    err = fabs(err);
    double ph = 0.0, dph = 0.0;
    for (int j = n - 1; j > ii; j--) {
        double t = z[j] / delta[j];
        ph += z[j] * t;
        dph += t * t;
        err += ph;
    }
    *psi = ps;
    *dpsi = dps;
    *phi = ph;
    *dphi = dph;
    *erretm = err;
}

static int laed4_mid(int n, int i, const double *d, const double *z,
                     double *delta, double rho, double *dlam)
{
    const double rhoinv = 1.0 / rho;
    const int ip1 = i + 1;
    double del = d[ip1] - d[i];
    double midpt = del / 2.0;
    double psi, dpsi, phi, dphi, erretm, w, c, a, b, tau, eta, temp, temp1;
    double dltlb, dltub, dw, prew;
    double zz[3];
    int orgati, ii, iim1, iip1, swtch3, swtch;

    for (int j = 0; j < n; j++)
        delta[j] = (d[j] - d[i]) - midpt;
    psi = 0.0;
    for (int j = 0; j < i; j++)
        psi += z[j] * z[j] / delta[j];
    phi = 0.0;
    for (int j = n - 1; j >= i + 2; j--)
        phi += z[j] * z[j] / delta[j];
    c = rhoinv + psi + phi;
    w = c + z[i] * z[i] / delta[i] + z[ip1] * z[ip1] / delta[ip1];

    if (w > 0.0) {
        /* root in (d[i], midpoint): origin d[i] */
        orgati = 1;
        a = c * del + z[i] * z[i] + z[ip1] * z[ip1];
        b = z[i] * z[i] * del;
        if (a > 0.0)
            tau = 2.0 * b / (a + sqrt(fabs(a * a - 4.0 * b * c)));
        else
            tau = (a - sqrt(fabs(a * a - 4.0 * b * c))) / (2.0 * c);
        dltlb = 0.0;
        dltub = midpt;
    } else {
        /* root in [midpoint, d[i+1]): origin d[i+1] */
        orgati = 0;
        a = c * del - z[i] * z[i] - z[ip1] * z[ip1];
        b = z[ip1] * z[ip1] * del;
        if (a < 0.0)
            tau = 2.0 * b / (a - sqrt(fabs(a * a + 4.0 * b * c)));
        else
            tau = -(a + sqrt(fabs(a * a + 4.0 * b * c))) / (2.0 * c);
        dltlb = -midpt;
        dltub = 0.0;
    }

    if (orgati)
        for (int j = 0; j < n; j++)
            delta[j] = (d[j] - d[i]) - tau;
    else
        for (int j = 0; j < n; j++)
            delta[j] = (d[j] - d[ip1]) - tau;
    ii = orgati ? i : ip1;
    iim1 = ii - 1;
    iip1 = ii + 1;
    const double dorg = d[ii];

    laed4_eval_mid(n, ii, z, delta, &psi, &dpsi, &phi, &dphi, &erretm);
    w = rhoinv + phi + psi;

    /* three-pole interpolation when the origin pole's term has the wrong
     * sign for the two-pole model, except at the ends of the pole list */
    swtch3 = 0;
    if (orgati) {
        if (w < 0.0)
            swtch3 = 1;
    } else {
        if (w > 0.0)
            swtch3 = 1;
    }
    if (ii == 0 || ii == n - 1)
        swtch3 = 0;

    temp = z[ii] / delta[ii];
    dw = dpsi + dphi + temp * temp;
    temp = z[ii] * temp;
    w += temp;
    erretm = 8.0 * (phi - psi) + erretm + 2.0 * rhoinv + 3.0 * fabs(temp);

    if (fabs(w) <= ORC_EPS * erretm) {
        *dlam = dorg + tau;
        return 0;
    }
    if (w <= 0.0)
        dltlb = dmax(dltlb, tau);
    else
        dltub = dmin(dltub, tau);

    /* first step (niter = 2) */
    int niter = 2;
    if (!swtch3) {
        if (orgati)
            c = w - delta[ip1] * dw -
                (d[i] - d[ip1]) * (z[i] / delta[i]) * (z[i] / delta[i]);
        else
            c = w - delta[i] * dw -
                (d[ip1] - d[i]) * (z[ip1] / delta[ip1]) * (z[ip1] / delta[ip1]);
        a = (delta[i] + delta[ip1]) * w - delta[i] * delta[ip1] * dw;
        b = delta[i] * delta[ip1] * w;
        if (c == 0.0) {
            if (a == 0.0) {
                if (orgati)
                    a = z[i] * z[i] + delta[ip1] * delta[ip1] * (dpsi + dphi);
                else
                    a = z[ip1] * z[ip1] + delta[i] * delta[i] * (dpsi + dphi);
            }
            eta = b / a;
        } else if (a <= 0.0) {
            eta = (a - sqrt(fabs(a * a - 4.0 * b * c))) / (2.0 * c);
        } else {
            eta = 2.0 * b / (a + sqrt(fabs(a * a - 4.0 * b * c)));
        }
    } else {
        temp = rhoinv + psi + phi;
        if (orgati) {
            temp1 = z[iim1] / delta[iim1];
            temp1 *= temp1;
            c = temp - delta[iip1] * (dpsi + dphi) - (d[iim1] - d[iip1]) * temp1;
            zz[0] = z[iim1] * z[iim1];
            zz[2] = delta[iip1] * delta[iip1] * ((dpsi - temp1) + dphi);
        } else {
            temp1 = z[iip1] / delta[iip1];
            temp1 *= temp1;
            c = temp - delta[iim1] * (dpsi + dphi) - (d[iip1] - d[iim1]) * temp1;
            zz[0] = delta[iim1] * delta[iim1] * (dpsi + (dphi - temp1));
            zz[2] = z[iip1] * z[iip1];
        }
        zz[1] = z[ii] * z[ii];
        if (brdc_oracle_laed6(niter, orgati, c, delta + iim1, zz, w, &eta)) {
            *dlam = dorg + tau;
            return 1;
        }
    }
    if (w * eta >= 0.0)
        eta = -w / dw;
    temp = tau + eta;
    if (temp > dltub || temp < dltlb) {
        if (w < 0.0)
            eta = (dltub - tau) / 2.0;
        else
            eta = (dltlb - tau) / 2.0;
    }
    prew = w;
    for (int j = 0; j < n; j++)
        delta[j] -= eta;
    laed4_eval_mid(n, ii, z, delta, &psi, &dpsi, &phi, &dphi, &erretm);
    temp = z[ii] / delta[ii];
    dw = dpsi + dphi + temp * temp;
    temp = z[ii] * temp;
    w = rhoinv + phi + psi + temp;
    erretm = 8.0 * (phi - psi) + erretm + 2.0 * rhoinv + 3.0 * fabs(temp) +
             fabs(tau + eta) * dw;
    swtch = 0;
    if (orgati) {
        if (-w > fabs(prew) / 10.0)
            swtch = 1;
    } else {
        if (w > fabs(prew) / 10.0)
            swtch = 1;
    }
    tau += eta;

    /* main loop: niter = 3 .. MAXIT */
    for (niter = 3; niter <= LAED4_MAXIT; niter++) {
        if (fabs(w) <= ORC_EPS * erretm) {
            *dlam = dorg + tau;
            return 0;
        }
        if (w <= 0.0)
            dltlb = dmax(dltlb, tau);
        else
            dltub = dmin(dltub, tau);

        if (!swtch3) {
            if (!swtch) {
                if (orgati)
                    c = w - delta[ip1] * dw -
                        (d[i] - d[ip1]) * (z[i] / delta[i]) * (z[i] / delta[i]);
                else
                    c = w - delta[i] * dw - (d[ip1] - d[i]) *
                                                (z[ip1] / delta[ip1]) *
                                                (z[ip1] / delta[ip1]);
            } else {
                temp = z[ii] / delta[ii];
                if (orgati)
                    dpsi += temp * temp;
                else
                    dphi += temp * temp;
                c = w - delta[i] * dpsi - delta[ip1] * dphi;
            }
            a = (delta[i] + delta[ip1]) * w - delta[i] * delta[ip1] * dw;
            b = delta[i] * delta[ip1] * w;
            if (c == 0.0) {
                if (a == 0.0) {
                    if (!swtch) {
                        if (orgati)
                            a = z[i] * z[i] +
                                delta[ip1] * delta[ip1] * (dpsi + dphi);
                        else
                            a = z[ip1] * z[ip1] +
                                delta[i] * delta[i] * (dpsi + dphi);
                    } else {
                        a = delta[i] * delta[i] * dpsi +
                            delta[ip1] * delta[ip1] * dphi;
                    }
                }
                eta = b / a;
            } else if (a <= 0.0) {
                eta = (a - sqrt(fabs(a * a - 4.0 * b * c))) / (2.0 * c);
            } else {
                eta = 2.0 * b / (a + sqrt(fabs(a * a - 4.0 * b * c)));
            }
        } else {
            temp = rhoinv + psi + phi;
            if (swtch) {
                c = temp - delta[iim1] * dpsi - delta[iip1] * dphi;
                zz[0] = delta[iim1] * delta[iim1] * dpsi;
                zz[2] = delta[iip1] * delta[iip1] * dphi;
            } else {
                if (orgati) {
                    temp1 = z[iim1] / delta[iim1];
                    temp1 *= temp1;
                    c = temp - delta[iip1] * (dpsi + dphi) -
                        (d[iim1] - d[iip1]) * temp1;
                    zz[0] = z[iim1] * z[iim1];
                    zz[2] = delta[iip1] * delta[iip1] * ((dpsi - temp1) + dphi);
                } else {
                    temp1 = z[iip1] / delta[iip1];
                    temp1 *= temp1;
                    c = temp - delta[iim1] * (dpsi + dphi) -
                        (d[iip1] - d[iim1]) * temp1;
                    zz[0] = delta[iim1] * delta[iim1] * (dpsi + (dphi - temp1));
                    zz[2] = z[iip1] * z[iip1];
                }
            }
            zz[1] = z[ii] * z[ii];
            if (brdc_oracle_laed6(niter, orgati, c, delta + iim1, zz, w, &eta)) {
                *dlam = dorg + tau;
                return 1;
            }
        }
        if (w * eta >= 0.0)
            eta = -w / dw;
        temp = tau + eta;
        if (temp > dltub || temp < dltlb) {
            if (w < 0.0)
                eta = (dltub - tau) / 2.0;
            else
                eta = (dltlb - tau) / 2.0;
        }
        for (int j = 0; j < n; j++)
            delta[j] -= eta;
        tau += eta;
        prew = w;
        laed4_eval_mid(n, ii, z, delta, &psi, &dpsi, &phi, &dphi, &erretm);
        temp = z[ii] / delta[ii];
        dw = dpsi + dphi + temp * temp;
        temp = z[ii] * temp;
        w = rhoinv + phi + psi + temp;
        erretm = 8.0 * (phi - psi) + erretm + 2.0 * rhoinv + 3.0 * fabs(temp) +
                 fabs(tau) * dw;
        if (w * prew > 0.0 && fabs(w) > fabs(prew) / 10.0)
            swtch = !swtch;
    }
    *dlam = dorg + tau;
    return 1;
}

int brdc_oracle_laed4(int n, int i, const double *d, const double *z,
                      double *delta, double rho, double *dlam)
{
    if (n == 1) {
        *dlam = d[0] + rho * z[0] * z[0];
        delta[0] = 1.0;
        return 0;
    }
    if (n == 2)
        return brdc_oracle_laed5(i, d, z, delta, rho, dlam);
    if (i == n - 1)
        return laed4_last(n, d, z, delta, rho, dlam);
    return laed4_mid(n, i, d, z, delta, rho, dlam);
}

/* DLAED4_BR (dlaed4_br.f:59-103): root plus (org, tau, near1, near2) with
 * lambda = d[org] + tau and delta[j] = (d[j] - d[org]) - tau. */
int brdc_oracle_laed4_br(int n, int i, const double *d, const double *z,
                         double *delta, double rho, double *dlam, int *org,
                         double *tau, double *near1, double *near2)
{
    *org = 0;
    *tau = *near1 = *near2 = 0.0;
    int info = brdc_oracle_laed4(n, i, d, z, delta, rho, dlam);
    if (info)
        return info;
    if (n == 1) {
        *org = 0;
        *tau = *dlam - d[0];
        *near1 = delta[0];
        return 0;
    }
    if (n == 2) {
        *org = 0;
        *tau = *dlam - d[0];
        *near1 = delta[0];
        *near2 = delta[1];
        return 0;
    }
    if (i == n - 1) {
        *org = n - 1;
    } else {
        double mid = 0.5 * (d[i] + d[i + 1]);
        *org = (*dlam < mid) ? i : i + 1;
    }
    *tau = -delta[*org]; /* dlaed4_br.f:95 */
    if (i == n - 1) {
        *near1 = delta[n - 2];
        *near2 = delta[n - 1];
    } else {
        *near1 = delta[i];
        *near2 = delta[i + 1];
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Implicit QL with Wilkinson shift.  e[0..n-2] couples (k, k+1).  When r0/r1
 * are given they are rows of the accumulated eigenvector matrix (initially
 * the first and last rows of the identity).  Output sorted ascending. */
static int ql_implicit(int n, double *d, const double *e_in, double *r0,
                       double *r1)
{
    if (n <= 0)
        return 0;
    double *e = (double *)malloc(sizeof(double) * (size_t)n);
    for (int k = 0; k < n - 1; k++)
        e[k] = e_in[k];
    e[n - 1] = 0.0;
    int info = 0;
    const int maxit = 30 * n;
    int iter = 0;
    for (int l = 0; l < n && !info; l++) {
        int m;
        for (;;) {
            for (m = l; m < n - 1; m++) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= ORC_EPS * dd)
                    break;
            }
            if (m == l)
                break;
            if (++iter > maxit) {
                info = l + 1;
                break;
            }
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = lapy2(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            int k;
            int underflow = 0;
            for (k = m - 1; k >= l; k--) {
                double f = s * e[k], b = c * e[k];
                r = lapy2(f, g);
                e[k + 1] = r;
                if (r == 0.0) {
                    d[k + 1] -= p;
                    e[m] = 0.0;
                    underflow = 1;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[k + 1] - p;
                r = (d[k] - g) * s + 2.0 * c * b;
                p = s * r;
                d[k + 1] = g + p;
                g = c * r - b;
                if (r0) {
                    double t = r0[k + 1];
                    r0[k + 1] = s * r0[k] + c * t;
                    r0[k] = c * r0[k] - s * t;
                    t = r1[k + 1];
                    r1[k + 1] = s * r1[k] + c * t;
                    r1[k] = c * r1[k] - s * t;
                }
            }
            if (underflow)
                continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    free(e);
    /* selection sort ascending, carrying the rows */
    for (int a = 0; a < n - 1; a++) {
        int best = a;
        for (int b = a + 1; b < n; b++)
            if (d[b] < d[best])
                best = b;
        if (best != a) {
            double t = d[a];
            d[a] = d[best];
            d[best] = t;
            if (r0) {
                t = r0[a];
                r0[a] = r0[best];
                r0[best] = t;
                t = r1[a];
                r1[a] = r1[best];
                r1[best] = t;
            }
        }
    }
    return info;
}

int brdc_oracle_leaf(int n, double *d, double *e, double *blo, double *bhi)
{
    for (int k = 0; k < n; k++) {
        blo[k] = (k == 0) ? 1.0 : 0.0;
        bhi[k] = (k == n - 1) ? 1.0 : 0.0;
    }
    return ql_implicit(n, d, e, blo, bhi);
}

int brdc_oracle_sterf(int n, double *d, double *e)
{
    return ql_implicit(n, d, e, NULL, NULL);
}

/* ------------------------------------------------------------------------- */
/* BRDC_WPROD_NTHREADS (dstedc_dc_br.f:153-213). */
int brdc_oracle_wprod_nthreads(void)
{
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    int val = 0;
    const char *env = getenv("OPENBLAS_BRDC_WPROD_THREADS");
    if (env) {
        for (const char *p = env; *p; p++) {
            if (*p >= '0' && *p <= '9')
                val = 10 * val + (*p - '0');
            else if (*p != ' ') {
                val = 0;
                break;
            }
        }
    }
    if (val > 0)
        return imax(1, val < nt ? val : nt);
    int cap = nt <= 1 ? 1 : nt <= 8 ? 2 : nt <= 28 ? 4 : 16;
    return imax(1, cap < nt ? cap : nt);
}

/* ------------------------------------------------------------------------- */
/* Per-merge scratch, allocated by the caller of laed7_br. */
typedef struct {
    double *z, *dlambda, *w, *q, *q2, *delta, *wsgn, *tau, *nr1, *nr2, *wprod;
    double *wacc, *dwork;
    int *indx, *indxp, *perm;
    int nthr;
} merge_ws;

static int merge_ws_alloc(merge_ws *ws, int n, int nthr)
{
    memset(ws, 0, sizeof(*ws));
    ws->nthr = nthr;
    size_t nd = (size_t)n;
    double *base = (double *)malloc(sizeof(double) * nd * (14 + 2 * (size_t)nthr));
    int *ibase = (int *)malloc(sizeof(int) * nd * 3);
    if (!base || !ibase) {
        free(base);
        free(ibase);
        return -1;
    }
    ws->z = base;
    ws->dlambda = ws->z + nd;
    ws->w = ws->dlambda + nd;
    ws->q = ws->w + nd;        /* 2n, column-major 2 x n */
    ws->q2 = ws->q + 2 * nd;   /* 2n */
    ws->delta = ws->q2 + 2 * nd;
    ws->wsgn = ws->delta + nd;
    ws->tau = ws->wsgn + nd;
    ws->nr1 = ws->tau + nd;
    ws->nr2 = ws->nr1 + nd;
    ws->wprod = ws->nr2 + nd;
    ws->wacc = ws->wprod + nd;           /* nthr * n */
    ws->dwork = ws->wacc + nd * (size_t)nthr; /* nthr * n */
    ws->indx = ibase;
    ws->indxp = ibase + nd;
    ws->perm = ibase + 2 * nd;
    return 0;
}

static void merge_ws_free(merge_ws *ws)
{
    free(ws->z);
    free(ws->indx);
}

/* DLAED8_BR (dlaed7_br.f:175-381): sort the two halves together, deflate tiny
 * z and close poles (with Givens rotations applied to the nbr selected rows in
 * q, which is 2 x n column-major), and compact to K secular poles.
 * On return d[K..n) holds the deflated eigenvalues (descending) and, if nbr,
 * q[:, K..n) their rows; dlambda/w/q2 hold the K secular poles.  *rho is
 * replaced by |2 rho|. */
static int laed8_br(int nbr, int n, int cut, double *d, double *q,
                    int *indxq, double *rho, double *z, merge_ws *ws, int *kout,
                    int *ngivens)
{
    double *dlambda = ws->dlambda, *w = ws->w, *q2 = ws->q2;
    int *indx = ws->indx, *indxp = ws->indxp, *perm = ws->perm;
    const int n1 = cut, n2 = n - cut;
    int k, k2, jlam = 0, j;
    *ngivens = 0;

    if (*rho < 0.0)
        for (j = n1; j < n; j++)
            z[j] = -z[j];
    const double t0 = 1.0 / sqrt(2.0);
    for (j = 0; j < n; j++) {
        indx[j] = j;
        z[j] *= t0;
    }
    *rho = fabs(2.0 * *rho);

    for (int i = cut; i < n; i++)
        indxq[i] += cut;
    for (int i = 0; i < n; i++) {
        dlambda[i] = d[indxq[i]];
        w[i] = z[indxq[i]];
    }
    brdc_oracle_lamrg(n1, n2, dlambda, 1, 1, indx);
    for (int i = 0; i < n; i++) {
        d[i] = dlambda[indx[i]];
        z[i] = w[indx[i]];
    }

    int izmax = 0, idmax = 0;
    for (int i = 1; i < n; i++) {
        if (fabs(z[i]) > fabs(z[izmax]))
            izmax = i;
        if (fabs(d[i]) > fabs(d[idmax]))
            idmax = i;
    }
    const double tol = 8.0 * ORC_EPS * fabs(d[idmax]);

    if (*rho * fabs(z[izmax]) <= tol) {
        /* everything deflates: keep the sorted order */
        *kout = 0;
        for (j = 0; j < n; j++) {
            perm[j] = indxq[indx[j]];
            if (nbr) {
                q2[2 * j] = q[2 * perm[j]];
                q2[2 * j + 1] = q[2 * perm[j] + 1];
            }
        }
        if (nbr)
            memcpy(q, q2, sizeof(double) * 2 * (size_t)n);
        return 0;
    }

    k = 0;
    k2 = n; /* deflated entries fill indxp[k2..n) from the top */
    for (j = 0; j < n; j++) {
        if (*rho * fabs(z[j]) <= tol) {
            indxp[--k2] = j;
            if (j == n - 1)
                goto compact;
        } else {
            jlam = j;
            break;
        }
    }
    for (;;) {
        j++;
        if (j > n - 1)
            break;
        if (*rho * fabs(z[j]) <= tol) {
            indxp[--k2] = j;
            continue;
        }
        double s = z[jlam], c = z[j];
        double tau = lapy2(c, s);
        double t = d[j] - d[jlam];
        c = c / tau;
        s = -s / tau;
        if (fabs(t * c * s) <= tol) {
            /* close poles: rotate jlam into j and deflate jlam */
            z[j] = tau;
            z[jlam] = 0.0;
            (*ngivens)++;
            if (nbr) {
                int c1 = indxq[indx[jlam]], c2 = indxq[indx[j]];
                for (int r = 0; r < 2; r++) {
                    double x = q[2 * c1 + r], y = q[2 * c2 + r];
                    q[2 * c1 + r] = c * x + s * y;
                    q[2 * c2 + r] = c * y - s * x;
                }
            }
            t = d[jlam] * c * c + d[j] * s * s;
            d[j] = d[jlam] * s * s + d[j] * c * c;
            d[jlam] = t;
            k2--;
            int slot = k2;
            for (;;) {
                if (slot + 1 <= n - 1) {
                    if (d[jlam] < d[indxp[slot + 1]]) {
                        indxp[slot] = indxp[slot + 1];
                        indxp[slot + 1] = jlam;
                        slot++;
                        continue;
                    }
                    indxp[slot] = jlam;
                } else {
                    indxp[slot] = jlam;
                }
                break;
            }
            jlam = j;
        } else {
            w[k] = z[jlam];
            dlambda[k] = d[jlam];
            indxp[k] = jlam;
            k++;
            jlam = j;
        }
    }
    w[k] = z[jlam];
    dlambda[k] = d[jlam];
    indxp[k] = jlam;
    k++;

compact:
    for (j = 0; j < n; j++) {
        int jp = indxp[j];
        dlambda[j] = d[jp];
        perm[j] = indxq[indx[jp]];
        if (nbr) {
            q2[2 * j] = q[2 * perm[j]];
            q2[2 * j + 1] = q[2 * perm[j] + 1];
        }
    }
    if (k < n) {
        memcpy(d + k, dlambda + k, sizeof(double) * (size_t)(n - k));
        if (nbr)
            memcpy(q + 2 * k, q2 + 2 * k, sizeof(double) * 2 * (size_t)(n - k));
    }
    *kout = k;
    return 0;
}

/* DLAED9_BR (dlaed9_br.f:11-393): K secular roots into d[0..K); when nbr, the
 * selected rows of the merged eigenvector block into qout[:, 0..K) from
 * qin[:, 0..K) via the product-form update vector and compact deltas. */
static int laed9_br(int k, int nbr, double *d, const double *qin, double *qout,
                    double rho, const double *dlambda, double *w, merge_ws *ws,
                    int nthreads)
{
    int info = 0;
    if (k == 0)
        return 0;
    if (nbr == 0) {
        /* root merge: eigenvalues only (dlaed9_br.f:77-140) */
        int par = (k > 128 && nthreads > 1);
#pragma omp parallel num_threads(par ? nthreads : 1)
        {
            double *delta = (double *)malloc(sizeof(double) * (size_t)k);
#pragma omp for schedule(static)
            for (int j = 0; j < k; j++) {
                int infoj = brdc_oracle_laed4(k, j, dlambda, w, delta, rho, &d[j]);
                if (infoj) {
#pragma omp critical
                    if (!info)
                        info = infoj;
                }
            }
            free(delta);
        }
        return info;
    }
    if (k == 1 || k == 2) {
        /* eigenvector returned directly in delta (dlaed9_br.f:146-172) */
        double delta[2];
        for (int j = 0; j < k; j++) {
            info = brdc_oracle_laed4(k, j, dlambda, w, delta, rho, &d[j]);
            if (info)
                return info;
            double a1 = 0.0, a2 = 0.0;
            for (int i = 0; i < k; i++) {
                a1 += qin[2 * i] * delta[i];
                a2 += qin[2 * i + 1] * delta[i];
            }
            qout[2 * j] = a1;
            qout[2 * j + 1] = a2;
        }
        return 0;
    }

    double *wsgn = ws->wsgn, *tau = ws->tau, *nr1 = ws->nr1, *nr2 = ws->nr2;
    double *wprod = ws->wprod;
    for (int i = 0; i < k; i++)
        wsgn[i] = w[i];

    int nthr = (k > 128) ? ws->nthr : 1;
    if (nthr < 1)
        nthr = 1;
    for (int t = 0; t < nthr; t++)
        for (int i = 0; i < k; i++)
            ws->wacc[(size_t)t * k + i] = 1.0;

    /* roots + product-form accumulation (dlaed9_br.f:181-262) */
#pragma omp parallel for num_threads(nthr) schedule(static)
    for (int j = 0; j < k; j++) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *dl = ws->dwork + (size_t)tid * k;
        double *acc = ws->wacc + (size_t)tid * k;
        int org;
        int infoj = brdc_oracle_laed4_br(k, j, dlambda, wsgn, dl, rho, &d[j],
                                         &org, &tau[j], &nr1[j], &nr2[j]);
        if (!infoj) {
            acc[j] *= dl[j];
            for (int i = 0; i < j; i++)
                acc[i] *= dl[i] / (dlambda[i] - dlambda[j]);
            for (int i = j + 1; i < k; i++)
                acc[i] *= dl[i] / (dlambda[i] - dlambda[j]);
        } else {
#pragma omp critical
            if (!info)
                info = infoj;
        }
    }
    if (info)
        return info;
    for (int i = 0; i < k; i++) {
        double p = 1.0;
        for (int t = 0; t < nthr; t++)
            p *= ws->wacc[(size_t)t * k + i];
        wprod[i] = p;
    }
    for (int i = 0; i < k; i++)
        w[i] = copysign(sqrt(-wprod[i]), wsgn[i]);

    /* stream each secular eigenvector through the two rows (dlaed9_br.f:272-331) */
#pragma omp parallel for num_threads(nthr) schedule(static)
    for (int j = 0; j < k; j++) {
        int iorg, n1, n2;
        if (j == k - 1) {
            iorg = k - 1;
            n1 = k - 2;
            n2 = k - 1;
        } else if (tau[j] >= 0.0) {
            iorg = j;
            n1 = j;
            n2 = j + 1;
        } else {
            iorg = j + 1;
            n1 = j;
            n2 = j + 1;
        }
        double temp = 0.0, a1 = 0.0, a2 = 0.0;
        for (int i = 0; i < k; i++) {
            double dlt;
            if (i == n1)
                dlt = w[i] / nr1[j];
            else if (i == n2)
                dlt = w[i] / nr2[j];
            else
                dlt = w[i] / ((dlambda[i] - dlambda[iorg]) - tau[j]);
            temp += dlt * dlt;
            a1 += qin[2 * i] * dlt;
            a2 += qin[2 * i + 1] * dlt;
        }
        temp = sqrt(temp);
        qout[2 * j] = a1 / temp;
        qout[2 * j + 1] = a2 / temp;
    }
    return 0;
}

static brdc_oracle_merge_cb g_merge_cb = NULL;
static int g_cb_level = 0, g_cb_start = 0;
void brdc_oracle_set_merge_callback(brdc_oracle_merge_cb cb) { g_merge_cb = cb; }

/* DLAED7_BR (dlaed7_br.f:7-167): one merge of two sorted children. */
static int laed7_br(int wantq, int n, int cut, double *d, int *indxq,
                    double rho, double *blo, double *bhi, merge_ws *ws,
                    int nthreads, int *kdefl, int *ngivens)
{
    *kdefl = 0;
    *ngivens = 0;
    if (n == 0)
        return 0;
    double *z = ws->z, *q = ws->q;
    for (int j = 0; j < n; j++) {
        if (j < cut) {
            z[j] = bhi[j];
            if (wantq) {
                q[2 * j] = blo[j];
                q[2 * j + 1] = 0.0;
            }
        } else {
            z[j] = blo[j];
            if (wantq) {
                q[2 * j] = 0.0;
                q[2 * j + 1] = bhi[j];
            }
        }
    }
    int nbr = wantq ? 2 : 0;
    int k;
    int info = laed8_br(nbr, n, cut, d, q, indxq, &rho, z, ws, &k, ngivens);
    if (info)
        return info;
    *kdefl = k;
    if (k != 0 && g_merge_cb)
        g_merge_cb(g_cb_level, g_cb_start, k, rho, ws->dlambda, ws->w);
    if (k != 0) {
        info = laed9_br(k, nbr, d, ws->q2, q, rho, ws->dlambda, ws->w, ws,
                        nthreads);
        if (info)
            return info;
        brdc_oracle_lamrg(k, n - k, d, 1, -1, indxq);
    } else {
        for (int i = 0; i < n; i++)
            indxq[i] = i;
    }
    if (wantq) {
        for (int j = 0; j < n; j++) {
            blo[j] = q[2 * j];
            bhi[j] = q[2 * j + 1];
        }
    }
    return 0;
}

/* DLAED0_BR (dlaed0_br.f:8-308). */
static int laed0_br(int n, double *d, double *e, int sml, int nthreads,
                    brdc_oracle_merge_stat *stats, int max_stats, int *n_stats)
{
    if (n == 0)
        return 0;
    int cap = 1;
    while (cap < 2 * n + 2)
        cap *= 2;
    int *iw = (int *)calloc((size_t)cap, sizeof(int));
    double *blo = (double *)malloc(sizeof(double) * (size_t)n);
    double *bhi = (double *)malloc(sizeof(double) * (size_t)n);
    int *indxq = (int *)malloc(sizeof(int) * (size_t)n);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
    int info = 0, err_sub = 0, err_siz = 0;

    /* split sizes: halve (left = floor, right = ceil) until the last piece
     * fits in a leaf (dlaed0_br.f:76-93) */
    iw[0] = n;
    int subpbs = 1, tlvls = 0;
    while (iw[subpbs - 1] > sml) {
        for (int j = subpbs - 1; j >= 0; j--) {
            iw[2 * j + 1] = (iw[j] + 1) / 2;
            iw[2 * j] = iw[j] / 2;
        }
        tlvls++;
        subpbs *= 2;
    }
    for (int j = 1; j < subpbs; j++)
        iw[j] += iw[j - 1];

    /* rank-one cuts (dlaed0_br.f:95-103) */
    for (int i = 0; i < subpbs - 1; i++) {
        int sub = iw[i];
        double ae = fabs(e[sub - 1]);
        d[sub - 1] -= ae;
        d[sub] -= ae;
    }

    /* leaves (dlaed0_br.f:131-170) */
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int i = 0; i < subpbs; i++) {
        int sub = (i == 0) ? 0 : iw[i - 1];
        int m = iw[i] - sub;
        int infom = brdc_oracle_leaf(m, d + sub, e + sub, blo + sub, bhi + sub);
        if (infom) {
#pragma omp critical
            if (!info) {
                info = infom;
                err_sub = sub;
                err_siz = m;
            }
        } else {
            for (int j = 0; j < m; j++)
                indxq[sub + j] = j;
        }
    }

    /* merge levels (dlaed0_br.f:180-291) */
    int level = tlvls - 1;
    while (!info && subpbs > 1) {
        int wantq = subpbs > 2;
        int nmerge = subpbs / 2;
        int wthr = brdc_oracle_wprod_nthreads();
        /* level-parallel when the level has enough merges to occupy the
         * threads, else merges in turn with root-parallel inner loops
         * (dlaed0_br.f:189-242, without the probe merge) */
        int outer = (nmerge > 1 && 4 * nmerge >= nthreads) ? nthreads : 1;
        if (outer > nmerge)
            outer = nmerge;
#pragma omp parallel for num_threads(outer) schedule(dynamic, 1) if (outer > 1)
        for (int mg = 0; mg < nmerge; mg++) {
            int left = 2 * mg, right = left + 1;
            int sub = (mg == 0) ? 0 : iw[left - 1];
            int msz = iw[right] - sub;
            int cut = iw[left] - sub;
            merge_ws ws;
            int inner = (outer > 1) ? 1 : nthreads;
            int kd = 0, ng = 0, infom;
            if (merge_ws_alloc(&ws, msz, (outer > 1) ? 1 : wthr) != 0) {
                infom = -99;
            } else {
                g_cb_level = level;
                g_cb_start = sub;
                infom = laed7_br(wantq, msz, cut, d + sub, indxq + sub,
                                 e[sub + cut - 1], blo + sub, bhi + sub, &ws,
                                 inner, &kd, &ng);
                merge_ws_free(&ws);
            }
            if (infom) {
#pragma omp critical
                if (!info) {
                    info = infom;
                    err_sub = sub;
                    err_siz = msz;
                }
            }
            if (stats && n_stats) {
                int idx = *n_stats + ((1 << level) - 1) + mg;
                if (idx < max_stats) {
                    stats[idx].level = level;
                    stats[idx].start = sub;
                    stats[idx].size = msz;
                    stats[idx].cut = cut;
                    stats[idx].k = kd;
                    stats[idx].givens = ng;
                }
            }
        }
        for (int mg = 0; mg < nmerge; mg++)
            iw[mg] = iw[2 * mg + 1];
        subpbs /= 2;
        level--;
    }
    if (stats && n_stats && !info)
        *n_stats += (1 << tlvls) - 1;

    if (!info) {
        for (int i = 0; i < n; i++)
            tmp[i] = d[indxq[i]];
        memcpy(d, tmp, sizeof(double) * (size_t)n);
    } else {
        /* failure encoding (dlaed0_br.f:302-303), 1-based submatrix */
        int submat = err_sub + 1;
        info = submat * (n + 1) + submat + err_siz - 1;
    }
    free(iw);
    free(blo);
    free(bhi);
    free(indxq);
    free(tmp);
    return info;
}

/* DLASCL('G') for the two cases used here: scale by cto/cfrom. */
static void lascl(int m, double cfrom, double cto, double *a)
{
    double mul = cto / cfrom;
    for (int i = 0; i < m; i++)
        a[i] *= mul;
}

static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int brdc_oracle_dstebr_ex(int n, double *d, double *e, double *work, int lwork,
                          int *iwork, int liwork, int smlsiz, int nthreads,
                          brdc_oracle_merge_stat *stats, int max_stats,
                          int *n_stats)
{
    int info = 0;
    const int lquery = (lwork == -1 || liwork == -1);
    const int sml = smlsiz > 0 ? smlsiz : REF_SMLSIZ;
    int lwmin = 1, liwmin = 1;
    if (n_stats)
        *n_stats = 0;
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif

    if (n < 0)
        info = -1;
    if (info == 0) {
        if (n <= sml) {
            lwmin = 1;
            liwmin = 1;
        } else {
            int nthr = brdc_oracle_wprod_nthreads();
            lwmin = imax(1, (16 + 2 * nthr) * n);
            liwmin = imax(1, 7 * n);
        }
        if (work)
            work[0] = lwmin;
        if (iwork)
            iwork[0] = liwmin;
        if (lwork < lwmin && !lquery)
            info = -5;
        else if (liwork < liwmin && !lquery)
            info = -7;
    }
    if (info != 0 || lquery)
        return info;
    if (n <= 1)
        return 0;

    if (n <= sml) {
        info = brdc_oracle_sterf(n, d, e);
        goto out;
    }

    double orgnrm = 0.0;
    for (int i = 0; i < n; i++)
        orgnrm = dmax(orgnrm, fabs(d[i]));
    for (int i = 0; i < n - 1; i++)
        orgnrm = dmax(orgnrm, fabs(e[i]));
    if (orgnrm <= 0.0)
        goto out;

    /* split at negligible off-diagonals (dstedc_dc_br.f:94-140) */
    int start = 0;
    while (start < n) {
        int finish = start;
        while (finish < n - 1) {
            double tiny = ORC_EPS * sqrt(fabs(d[finish])) * sqrt(fabs(d[finish + 1]));
            if (fabs(e[finish]) > tiny)
                finish++;
            else
                break;
        }
        int m = finish - start + 1;
        if (m == 1) {
            start = finish + 1;
            continue;
        }
        if (m > sml) {
            double bn = 0.0;
            for (int i = 0; i < m; i++)
                bn = dmax(bn, fabs(d[start + i]));
            for (int i = 0; i < m - 1; i++)
                bn = dmax(bn, fabs(e[start + i]));
            lascl(m, bn, 1.0, d + start);
            lascl(m - 1, bn, 1.0, e + start);
            info = laed0_br(m, d + start, e + start, sml, nthreads, stats,
                            max_stats, n_stats);
            if (info) {
                info = (info / (m + 1) + start) * (n + 1) + info % (m + 1) + start;
                goto out;
            }
            lascl(m, 1.0, bn, d + start);
        } else {
            info = brdc_oracle_sterf(m, d + start, e + start);
            if (info) {
                info = (start + 1) * (n + 1) + finish + 1;
                goto out;
            }
        }
        start = finish + 1;
    }
    qsort(d, (size_t)n, sizeof(double), cmp_double); /* DLASRT('I') */
out:
    if (work)
        work[0] = lwmin;
    if (iwork)
        iwork[0] = liwmin;
    return info;
}

int brdc_oracle_dstebr(int n, double *d, double *e, double *work, int lwork,
                       int *iwork, int liwork)
{
    return brdc_oracle_dstebr_ex(n, d, e, work, lwork, iwork, liwork, 0, 0,
                                 NULL, 0, NULL);
}

/* ---- binary128 Sturm bisection (high-precision checker) ---------------- */

/* number of eigenvalues of (d, e) below x: negative pivots of T - x I */
static int sturm_count_q(int n, const double *d, const double *e, __float128 x,
                         __float128 pivmin) {
    int neg = 0;
    __float128 q = (__float128)d[0] - x;
    for (int i = 0;; i++) {
        if ((q < 0 ? -q : q) < pivmin) q = -pivmin;
        if (q < 0) neg++;
        if (i + 1 >= n) break;
        const __float128 ei = (__float128)e[i];
        q = ((__float128)d[i + 1] - x) - ei * ei / q;
    }
    return neg;
}

int brdc_oracle_eig_quad(int n, const double *d, const double *e, int k, double lo,
                         double hi, int steps, double *out) {
    if (n < 1 || k < 0 || k >= n || !(lo < hi)) return -1;
    double emax = 0.0;
    for (int i = 0; i + 1 < n; i++) emax = fmax(emax, fabs(e[i]));
    /* DLAEBZ's pivmin: safe minimum times max(1, max e^2), in binary128 */
    const __float128 pivmin = (__float128)DBL_MIN * (__float128)fmax(1.0, emax * emax);
    __float128 a = lo, b = hi;
    if (sturm_count_q(n, d, e, a, pivmin) > k || sturm_count_q(n, d, e, b, pivmin) <= k)
        return -1;
    for (int s = 0; s < steps; s++) {
        const __float128 m = (a + b) / 2;
        if (sturm_count_q(n, d, e, m, pivmin) > k) b = m;
        else a = m;
    }
    const __float128 m = (a + b) / 2;
    out[0] = (double)m;
    out[1] = (double)(m - (__float128)out[0]);
    return 0;
}
