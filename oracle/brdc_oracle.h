/*
 * brdc_oracle.h -- CPU restatement of the reference's boundary-row (BR)
 * values-only divide-and-conquer eigensolver.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path and the timed CPU baseline ("port") of bench.py.  Nothing in the
 * product library links or calls it; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it.
 *
 * Every routine restates a reference routine (cpu/src/dlaed*_br.f and
 * cpu/src/dstedc_dc_br.f) or one of the
 * LAPACK routines those call (third-party, not present in the reference tree:
 * Reference-LAPACK 3.12 / OpenBLAS 0.3.28 DLAED4, DLAED5, DLAED6, DLAMRG,
 * DSTEQR, DSTERF), written from the published algorithms.  Indices are 0-based.
 */
#ifndef BRDC_ORACLE_H
#define BRDC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* One merge of the tree: level (0 = root merge), first row, size, cut and the
 * number of non-deflated poles K (DLAED7_BR's KDEFL).  Diagnostics for the GPU
 * path, which reports the same record per merge. */
typedef struct brdc_oracle_merge_stat {
    int level;
    int start;
    int size;
    int cut;
    int k;
    int givens;   /* number of close-pole Givens deflations */
} brdc_oracle_merge_stat;

/* Workspace query exactly as DSTEDC_DC_BR (dstedc_dc_br.f:44-75) with
 * SMLSIZ = 25 and NTHR = brdc_oracle_wprod_nthreads(). */
int brdc_oracle_wprod_nthreads(void);

/* DSTEDC_DC_BR (dstedc_dc_br.f:7-151).  Same argument meaning and INFO
 * convention.  smlsiz <= 0 selects the reference's 25 (ILAENV(9));
 * nthreads <= 0 uses the OpenMP default.  stats may be NULL; otherwise it
 * receives up to max_stats merge records and *n_stats their count. */
int brdc_oracle_dstebr_ex(int n, double *d, double *e, double *work, int lwork,
                          int *iwork, int liwork, int smlsiz, int nthreads,
                          brdc_oracle_merge_stat *stats, int max_stats,
                          int *n_stats);
int brdc_oracle_dstebr(int n, double *d, double *e, double *work, int lwork,
                       int *iwork, int liwork);

/* Test hook: when set, every merge's secular problem (level, first row, K,
 * rho, poles dlambda[K], weights w[K]) is passed to the callback just before
 * its roots are solved.  Not thread-safe; call brdc_oracle_dstebr_ex with
 * nthreads = 1 while it is set. */
typedef void (*brdc_oracle_merge_cb)(int level, int start, int k, double rho,
                                     const double *dlambda, const double *w);
void brdc_oracle_set_merge_callback(brdc_oracle_merge_cb cb);

/* Secular equation roots (LAPACK DLAED4/DLAED5/DLAED6 restated).  i is
 * 0-based.  Return value is LAPACK's INFO. */
int brdc_oracle_laed4(int n, int i, const double *d, const double *z,
                      double *delta, double rho, double *dlam);
int brdc_oracle_laed5(int i, const double *d, const double *z, double *delta,
                      double rho, double *dlam);
int brdc_oracle_laed6(int kniter, int orgati, double rho, const double *d,
                      const double *z, double finit, double *tau);

/* DLAED4_BR (dlaed4_br.f:15-109): root plus compact delta (org, tau, near). */
int brdc_oracle_laed4_br(int n, int i, const double *d, const double *z,
                         double *delta, double rho, double *dlam, int *org,
                         double *tau, double *near1, double *near2);

/* Leaf eigensolve (DSTEQR('I') + DSTEQR_BOUNDARY_BR, dlaed0_br.f:354-371):
 * ascending eigenvalues in d, first/last eigenvector rows in blo/bhi.
 * Implicit QL with Wilkinson shifts accumulating only the two rows. */
int brdc_oracle_leaf(int n, double *d, double *e, double *blo, double *bhi);

/* Values-only implicit QL (the DSTERF role, dstedc_dc_br.f:83,131),
 * ascending output. */
int brdc_oracle_sterf(int n, double *d, double *e);

/* High-precision checker (the role of the reference's MPFR probe,
 * run_mpfr_main_families_probe_job.sh:340-416, and of the error column of
 * README.md:70-75): eigenvalue k (0-based, ascending) of the symmetric
 * tridiagonal (d, e) by bisection on Sturm counts evaluated in binary128
 * (the count of negative pivots of the LDL^T factorization of T - x I, a
 * pivot below pivmin in magnitude replaced by -pivmin as in DLAEBZ).  The
 * bracket [lo, hi] must hold the eigenvalue (count(lo) <= k < count(hi));
 * steps halvings follow.  Returns the midpoint of the last bracket as the
 * double-double out[0] + out[1]; 0 on success, -1 if the bracket does not
 * hold eigenvalue k. */
int brdc_oracle_eig_quad(int n, const double *d, const double *e, int k, double lo,
                         double hi, int steps, double *out);

/* DLAMRG: merge permutation of two sorted sub-lists (0-based output). */
void brdc_oracle_lamrg(int n1, int n2, const double *a, int s1, int s2,
                       int *index);

#ifdef __cplusplus
}
#endif
#endif
