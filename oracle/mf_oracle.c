/*
 * mf_oracle.c -- FP64 CPU oracle for the per-channel multifit.
 *
 * TEST INFRASTRUCTURE: the parity checker and the CPU baseline.  Nothing in
 * the product library links or calls this file.
 *
 * What it restates (reference files are cited relative to the reference
 * repository root):
 *   or_llt / or_cholesky_decomp ... linalgebra/interface/cholesky.hpp:8-38 and
 *                                   Eigen LLT (called at
 *                                   multifit_cpu/src/PulseChiSqSNNLS.cpp:192,
 *                                   nnls/src/inplace_fnnls.cu:176-177)
 *   or_forward_subst / backward ... linalgebra/interface/substitution.hpp:10-71
 *   or_swap_row_column ............ linalgebra/interface/common.hpp:11-42
 *   or_fused_rcaddition ........... linalgebra/interface/fused_cholesky_substitution.hpp:18-57
 *   or_fused_inbetween_removal .... linalgebra/interface/fused_cholesky_substitution.hpp:62-102
 *   or_transpose_multiply ......... nnls/src/inplace_fnnls.cu:11-25
 *   or_inplace_fnnls .............. nnls/src/inplace_fnnls.cu:72-233
 *   or_inplace_fnnls_chol ......... nnls/src/inplace_fnnls_test_cholesky.cu:35-227
 *   or_fnnls ...................... nnls/src/fnnls.cu:28-173
 *   or_build_pulsemat ............. multifit_cpu/src/PulseChiSqSNNLS.cpp:50-62
 *   or_dofit (+ minimize, update_cov, nnls_step, chisq, approx uncertainty)
 *                                   multifit_cpu/src/PulseChiSqSNNLS.cpp:21-207,219-274
 *
 * Arithmetic conventions (the reference leaves them to Eigen, whose order is
 * not fixed by any reference test): dot products run in ascending index order
 * with fused multiply-adds, Cholesky elements are (M_ij - sum)/L_jj as in
 * cholesky.hpp, substitutions subtract term by term as in substitution.hpp.
 * Build with -ffp-contract=off so that only the explicit fma() calls fuse.
 *
 * Pinning (checks live in tests/test_oracle.py):
 *   dense primitives ...... golden vectors written by the reference's own
 *                           linalgebra headers (tests/golden), numpy
 *                           cholesky/solve, the 3x3 closed form.
 *   FNNLS variants ........ scipy.optimize.nnls, brute force over all 2^n
 *                           passive sets, KKT conditions, FP32 golden vectors
 *                           from the reference's OpenCL inplace_fnnls compiled
 *                           as host C.
 *   multifit .............. closed forms (single pulse, noise-free recovery,
 *                           identity covariance) and recomputation of chi2.
 *   Parity unpinned: Eigen's maxCoeff tie order (ties are resolved to the
 *   first index of the permuted tail, as Eigen's visitor does) and
 *   rounding-level decisions (|dchi2| ~ 1e-3, max w ~ eps); no reference test
 *   fixes them.
 */
#include "mf_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <string.h>

#define LD OR_LD
#define IDX(i, j) ((i) * LD + (j))

/* ------------------------------------------------------------------------ */
/* dense primitives                                                          */
/* ------------------------------------------------------------------------ */

/* Cholesky element recurrence shared by Eigen's unblocked LLT and
 * linalgebra's cholesky_decomp: both compute
 *   L_kk = sqrt(M_kk - sum_{j<k} L_kj^2),
 *   L_ik = (M_ik - sum_{j<k} L_ij L_kj) / L_kk.
 * Returns -1 or the failing column (pivot <= 0, Eigen llt_inplace::unblocked). */
static int chol_core(const double *M, double *L, int ld, int n, int stop_on_fail)
{
    for (int k = 0; k < n; ++k) {
        double sq = 0.0;
        for (int j = 0; j < k; ++j) sq = fma(L[k * ld + j], L[k * ld + j], sq);
        double d = M[k * ld + k] - sq;
        if (stop_on_fail && !(d > 0.0)) {
            if (d <= 0.0) {
                /* Eigen leaves columns >= k as the input's lower triangle */
                for (int c = k; c < n; ++c)
                    for (int i = c; i < n; ++i) L[i * ld + c] = M[i * ld + c];
                for (int i = 0; i < n; ++i)
                    for (int j = i + 1; j < n; ++j) L[i * ld + j] = 0.0;
                return k;
            }
            /* NaN pivot: Eigen's "x <= 0" test is false, it continues */
        }
        double lkk = sqrt(d);
        L[k * ld + k] = lkk;
        for (int i = k + 1; i < n; ++i) {
            double t = 0.0;
            for (int j = 0; j < k; ++j) t = fma(L[i * ld + j], L[k * ld + j], t);
            L[i * ld + k] = (M[i * ld + k] - t) / lkk;
        }
        for (int j = k + 1; j < n; ++j) L[k * ld + j] = 0.0;
    }
    return -1;
}

int or_llt(const double *M, double *L, int n) { return chol_core(M, L, LD, n, 1); }

void or_cholesky_decomp(const double *M, double *L, int ld, int n)
{
    /* linalgebra cholesky_decomp writes only the lower triangle */
    for (int i = 0; i < n; ++i) {
        double sumsq = 0.0;
        for (int j = 0; j < i; ++j) {
            double s2 = 0.0;
            for (int k = 0; k < j; ++k) s2 = fma(L[i * ld + k], L[j * ld + k], s2);
            double v = (M[i * ld + j] - s2) / L[j * ld + j];
            L[i * ld + j] = v;
            sumsq = fma(v, v, sumsq);
        }
        L[i * ld + i] = sqrt(M[i * ld + i] - sumsq);
    }
}

void or_forward_subst(const double *L, const double *b, double *y, int ld, int n)
{
    for (int i = 0; i < n; ++i) {
        double t = b[i];
        for (int j = 0; j < i; ++j) t = fma(-L[i * ld + j], y[j], t);
        y[i] = t / L[i * ld + i];
    }
}

/* L^T x = y with L lower (substitution.hpp:34-50 reads L transposed) */
void or_backward_subst(const double *L, const double *y, double *x, int ld, int n)
{
    for (int i = n - 1; i >= 0; --i) {
        double t = y[i];
        for (int j = i + 1; j < n; ++j) t = fma(-L[j * ld + i], x[j], t);
        x[i] = t / L[i * ld + i];
    }
}

void or_swap_row_column(double *M, int i, int j, int ld, int n)
{
    /* symmetric swap of rows/cols i and j (precondition j > i), as
     * common.hpp:11-42: diagonal, then the elements left of i and right of i
     * (the j-th entries of those loops touch the (i,j) pair itself) */
    double t = M[i * ld + i];
    M[i * ld + i] = M[j * ld + j];
    M[j * ld + j] = t;
    for (int e = 0; e < i; ++e) {
        double tmp = M[i * ld + e];
        M[i * ld + e] = M[j * ld + e];
        M[j * ld + e] = tmp;
        M[e * ld + i] = M[e * ld + j];
        M[e * ld + j] = tmp;
    }
    for (int e = i + 1; e < n; ++e) {
        double tmp = M[i * ld + e];
        M[i * ld + e] = M[j * ld + e];
        M[j * ld + e] = tmp;
        M[e * ld + i] = M[e * ld + j];
        M[e * ld + j] = tmp;
    }
}

void or_fused_rcaddition(const double *M, double *L, const double *b, double *y,
                         int ld, int n)
{
    int row = n - 1;
    double sum = 0.0, total = b[row];
    for (int col = 0; col < row; ++col) {
        double s2 = 0.0;
        for (int j = 0; j < col; ++j) s2 = fma(L[row * ld + j], L[col * ld + j], s2);
        double v = (M[row * ld + col] - s2) / L[col * ld + col];
        total = fma(-v, y[col], total);
        sum = fma(v, v, sum);
        L[row * ld + col] = v;
    }
    double d = sqrt(M[row * ld + row] - sum);
    L[row * ld + row] = d;
    y[row] = total / d;
}

void or_fused_inbetween_removal(const double *M, double *L, const double *b,
                                double *y, int position, int ld, int n)
{
    for (int i = position; i < n; ++i) {
        double sumsq = 0.0, total = b[i];
        for (int j = 0; j < i; ++j) {
            double s2 = 0.0;
            for (int k = 0; k < j; ++k) s2 = fma(L[i * ld + k], L[j * ld + k], s2);
            double v = (M[i * ld + j] - s2) / L[j * ld + j];
            L[i * ld + j] = v;
            total = fma(-v, y[j], total);
            sumsq = fma(v, v, sumsq);
        }
        double d = sqrt(M[i * ld + i] - sumsq);
        L[i * ld + i] = d;
        y[i] = total / d;
    }
}

static void tmul_mn(const double *A, double *AtA, int m, int n)
{
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            double t = 0.0;
            for (int k = 0; k < m; ++k) t = fma(A[IDX(k, i)], A[IDX(k, j)], t);
            AtA[IDX(i, j)] = t;
            AtA[IDX(j, i)] = t;
        }
}

void or_transpose_multiply(const double *A, double *AtA, int n) { tmul_mn(A, AtA, n, n); }

static void atb_mn(const double *A, const double *b, double *Atb, int m, int n)
{
    for (int i = 0; i < n; ++i) {
        double t = 0.0;
        for (int k = 0; k < m; ++k) t = fma(A[IDX(k, i)], b[k], t);
        Atb[i] = t;
    }
}

/* Decision margins of a fit (or_result.m_*): how close each data-dependent
 * decision of the reference's algorithm came to going the other way, in
 * units of the rounding error of the quantity decided on.  That error is
 * estimated as |operands| u (n + 1/rcond): n u for the sums themselves,
 * u / rcond for the amplitudes they are formed from, rcond being the
 * reciprocal condition number of the passive block that last produced
 * them.  Another code (another operation order or precision) may decide
 * differently only where a margin is a modest number of these units. */
#define OR_U 1.1102230246251565e-16
typedef struct {
    double m_eps;    /* stop test: |max w - eps| / err(w_q)                            */
    double m_arg;    /* argmax over the free w: (w_best - w_next) / err(w_q)           */
    double m_pos;    /* feasibility of a solve: min_i |s_i| / err(s)                   */
    double m_alpha;  /* which index leaves: relative gap of the two smallest step
                        ratios / their relative error                                  */
    double rc_last;  /* rcond of the passive block solved last (1 before any)         */
    double rc_min;   /* smallest rcond of all passive blocks solved                   */
    double rc_of[LD];/* inplace_fnnls: rcond of the solve that last set amplitude i    */
    int tracked;     /* rc_of is kept (the inplace_fnnls solvers)                     */
    int on;
} margins;

static void margins_init(margins *g)
{
    g->m_eps = g->m_arg = g->m_pos = g->m_alpha = INFINITY;
    g->rc_last = 1.0;
    g->rc_min = INFINITY;
    for (int i = 0; i < LD; ++i) g->rc_of[i] = INFINITY;
    g->tracked = 0;
    g->on = 1;
}
static double rnd_units(double scale, int n, double rc)
{
    return scale * OR_U * ((double)n + 1.0 / (rc > 1e-300 ? rc : 1e-300));
}
static void mmin(double *m, double v)
{
    if (v < *m) *m = v;   /* a NaN leaves it */
}

/* The margins of the fit in progress on this thread (set by nnls_step), for
 * the solvers that record only their passive solves' conditioning and
 * signs (fnnls, legacy, Lawson-Hanson). */
static _Thread_local margins *g_mg;

/* 1 / (|M|_1 |M^-1|_1) of the leading k x k block of M (symmetric positive
 * definite; 0 when its Cholesky factorization fails).  M^-1 is formed from
 * the factor column by column: an exact-enough estimate of the conditioning
 * for a test's rounding band. */
static double rcond1(const double *M, int k)
{
    double L[LD * LD], e[LD], y[LD], z[LD];
    if (chol_core(M, L, LD, k, 1) >= 0) return 0.0;
    double nm = 0.0, ni = 0.0;
    for (int j = 0; j < k; ++j) {
        double cs = 0.0;
        for (int i = 0; i < k; ++i) cs += fabs(M[IDX(i, j)]);
        if (cs > nm) nm = cs;
        for (int i = 0; i < k; ++i) e[i] = i == j ? 1.0 : 0.0;
        or_forward_subst(L, e, y, LD, k);
        or_backward_subst(L, y, z, LD, k);
        double ci = 0.0;
        for (int i = 0; i < k; ++i) ci += fabs(z[i]);
        if (ci > ni) ni = ci;
    }
    const double r = 1.0 / (nm * ni);
    return r == r ? r : 0.0;
}

/* a passive solve of the block M (leading k x k, leading dimension LD) with
 * solution s[0..k): its conditioning, and how near an s_i came to the
 * other sign */
static void record_solve(margins *g, const double *M, const double *s, int k)
{
    if (!g || !g->on) return;
    g->rc_last = rcond1(M, k);
    mmin(&g->rc_min, g->rc_last);
    double smax = 0.0, smin = INFINITY;
    for (int i = 0; i < k; ++i) {
        if (fabs(s[i]) > smax) smax = fabs(s[i]);
        mmin(&smin, fabs(s[i]));
    }
    if (smax > 0.0) mmin(&g->m_pos, smin / rnd_units(smax, k, g->rc_last));
}

/* s = LLT(M[0:k,0:k]).solve(v[0:k]) with full refactorisation */
static int llt_solve(const double *M, const double *v, double *s, int k)
{
    double L[LD * LD], y[LD];
    int f = chol_core(M, L, LD, k, 1);
    or_forward_subst(L, v, y, LD, k);
    or_backward_subst(L, y, s, LD, k);
    if (g_mg && g_mg->on) record_solve(g_mg, M, s, k);
    return f;
}

/* ------------------------------------------------------------------------ */
/* NNLS solvers                                                              */
/* ------------------------------------------------------------------------ */

static void swap_d(double *a, int i, int j) { double t = a[i]; a[i] = a[j]; a[j] = t; }
static void swap_i(int *a, int i, int j) { int t = a[i]; a[i] = a[j]; a[j] = t; }

/* full symmetric row+column swap of a 10x10 matrix, as Eigen's
 * AtA.col(a).swap(AtA.col(b)); AtA.row(a).swap(AtA.row(b)); */
static void sym_swap(double *M, int a, int b)
{
    if (a == b) return;
    for (int r = 0; r < LD; ++r) swap_d(M + r * LD, a, b);
    for (int c = 0; c < LD; ++c) {
        double t = M[IDX(a, c)];
        M[IDX(a, c)] = M[IDX(b, c)];
        M[IDX(b, c)] = t;
    }
}

static double chol_flops(int k)
{
    return k * ((double)k * k - 1.0) / 3.0 + k * (k - 1) / 2.0 + k;
}

static void stats_solve(or_nnls_stats *st, int k)
{
    if (!st) return;
    st->aflops += chol_flops(k) + 2.0 * k * k;
    st->inner++;
    st->ksum += k;
    st->k2sum += (long long)k * k;
    st->k3sum += (long long)k * k * k;
}

/* Shared driver of inplace_fnnls and its incremental-Cholesky twin
 * (inplace_fnnls.cu:72-233 and inplace_fnnls_test_cholesky.cu:35-227): the
 * loops are identical, only the passive-set solve differs.  g (optional):
 * the fit's decision margins, updated; pm_out (optional): the final
 * passive set as a bit mask over the original indices. */
static void inplace_fnnls_mn(const double *A, const double *b, double *x, int m,
                             int n, double eps, unsigned max_iter,
                             or_nnls_stats *st, int incremental, margins *g,
                             unsigned *pm_out)
{
    double AtA[LD * LD], Atb[LD], s[LD], w[LD], xp[LD];
    double pL[LD * LD], py[LD];
    int perm[LD];
    memset(AtA, 0, sizeof(AtA));
    memset(pL, 0, sizeof(pL));
    tmul_mn(A, AtA, m, n);               /* :115 */
    atb_mn(A, b, Atb, m, n);             /* :117 */
    for (int i = 0; i < n; ++i) { perm[i] = i; xp[i] = x[i]; s[i] = 0.0; }

    int np = 0;
    unsigned iter;
    for (iter = 0; iter < max_iter; ++iter) {          /* :127 */
        int nact = n - np;
        if (!nact) break;                                /* :134 */
        /* w.tail = Atb.tail - (AtA*x).tail                 :137 */
        for (int i = np; i < n; ++i) {
            double t = 0.0;
            for (int j = 0; j < n; ++j) t = fma(AtA[IDX(i, j)], xp[j], t);
            w[i] = Atb[i] - t;
        }
        if (st) st->aflops += 2.0 * n * n + (n - np);
        /* maxCoeff: first index of the maximum            :144 */
        int q = np;
        double maxw = w[np];
        for (int i = np + 1; i < n; ++i)
            if (w[i] > maxw) { maxw = w[i]; q = i; }
        if (g && g->on) {
            double sc = fabs(Atb[q]);
            for (int j = 0; j < n; ++j) sc = fma(fabs(AtA[IDX(q, j)]), fabs(xp[j]), sc);
            if (sc > 0.0) {
                /* w_q of an index with x_q > 0 is a residual of the solve that
                 * set x: backward stability bounds its error by n u |operands|
                 * whatever the conditioning; otherwise the error of x enters */
                const double ew = rnd_units(sc, n, xp[q] > 0.0 ? 1.0 : g->rc_last);
                mmin(&g->m_eps, fabs(maxw - eps) / ew);
                double next = -INFINITY;
                for (int i = np; i < n; ++i)
                    if (i != q && w[i] > next) next = w[i];
                if (maxw >= eps && next > -INFINITY) mmin(&g->m_arg, (maxw - next) / ew);
            }
        }
        if (maxw < eps) break;                           /* :147 */
        sym_swap(AtA, np, q);                            /* :157-164 */
        swap_d(Atb, np, q);
        swap_d(xp, np, q);
        swap_i(perm, np, q);
        ++np;
        if (st) st->outer++;

        int inner_iteration = 0, position_removed = -1;
        while (np > 0) {                                 /* :175 */
            if (!incremental) {
                int f = llt_solve(AtA, Atb, s, np);      /* :176-177 */
                if (f >= 0 && st) st->llt_fail = 1;
            } else {
                /* inplace_fnnls_test_cholesky.cu:145-166 */
                if (np == 1) {
                    double l00 = sqrt(AtA[0]);
                    pL[0] = l00;
                    py[0] = Atb[0] / l00;
                    position_removed = -1;
                } else if (inner_iteration == 0) {
                    or_fused_rcaddition(AtA, pL, Atb, py, LD, np);
                } else {
                    or_fused_inbetween_removal(AtA, pL, Atb, py, position_removed,
                                               LD, np);
                }
                or_backward_subst(pL, py, s, LD, np);
            }
            stats_solve(st, np);
            if (incremental && g && g->on) record_solve(g, AtA, s, np);

            if (g && g->on) {   /* every passive amplitude is set from this solve */
                g->tracked = 1;
                for (int i = 0; i < np; ++i) g->rc_of[perm[i]] = g->rc_last;
            }
            double mins = s[0];                          /* :179 */
            for (int i = 1; i < np; ++i) if (s[i] < mins) mins = s[i];
            if (mins > 0.0) {
                for (int i = 0; i < np; ++i) xp[i] = s[i];
                break;
            }
            double alpha = DBL_MAX;                      /* :188-200 */
            int aidx = 0;
            double alpha2 = DBL_MAX;
            for (int i = 0; i < np; ++i) {
                if (s[i] <= 0.0) {
                    double ratio = xp[i] / (xp[i] - s[i]);
                    if (ratio < alpha) { alpha2 = alpha; alpha = ratio; aidx = i; }
                    else if (ratio < alpha2) alpha2 = ratio;
                    if (xp[i] == 0.0) alpha2 = DBL_MAX;   /* ties at an exact 0: no rounding */
                }
            }
            if (g && g->on && alpha2 < DBL_MAX)
                mmin(&g->m_alpha, (alpha2 - alpha) / (alpha > 0.0 ? alpha : 1.0) /
                                      rnd_units(1.0, np, g->rc_last));
            if (alpha == DBL_MAX) {                      /* :202-205 */
                for (int i = 0; i < np; ++i) xp[i] = s[i];
                break;
            }
            for (int i = 0; i < np; ++i)                 /* :215 */
                xp[i] = fma(alpha, s[i] - xp[i], xp[i]);
            xp[aidx] = 0.0;                              /* :216 */
            --np;
            sym_swap(AtA, np, aidx);                     /* :222-229 */
            swap_d(Atb, np, aidx);
            swap_d(xp, np, aidx);
            swap_i(perm, np, aidx);
            position_removed = aidx;
            inner_iteration++;
        }
    }
    if (st && iter == max_iter && max_iter > 0) st->hit_max = 1;
    for (int i = 0; i < n; ++i) x[perm[i]] = xp[i];      /* :232 */
    if (pm_out) {
        unsigned pm = 0;
        for (int i = 0; i < np; ++i) pm |= 1u << perm[i];
        *pm_out = pm;
    }
}

void or_inplace_fnnls(const double *A, const double *b, double *x, int n,
                      double eps, unsigned max_iter, or_nnls_stats *st)
{
    inplace_fnnls_mn(A, b, x, n, n, eps, max_iter, st, 0, NULL, NULL);
}

void or_inplace_fnnls_chol(const double *A, const double *b, double *x, int n,
                           double eps, unsigned max_iter, or_nnls_stats *st)
{
    inplace_fnnls_mn(A, b, x, n, n, eps, max_iter, st, 1, NULL, NULL);
}

/* fnnls.cu:28-173.  Bool-mask FNNLS working on the full-length x.
 * Deviation: the reference reads an uninitialised `s` if its passive count
 * guard (:119) fires and sizes the sub-matrix from a counter that is one short
 * when w's running maximum starts on a passive index (:82-95); here `s`
 * starts at zero and the passive block is sized by the actual mask. */
static void fnnls_mn(const double *A, const double *b, double *x, int m, int n,
                     double eps, unsigned max_iter, or_nnls_stats *st)
{
    double AtA[LD * LD], Atb[LD], s[LD], w[LD];
    int active[LD];
    memset(AtA, 0, sizeof(AtA));
    tmul_mn(A, AtA, m, n);
    atb_mn(A, b, Atb, m, n);
    for (int i = 0; i < n; ++i) { active[i] = 1; s[i] = 0.0; }
    unsigned iter;
    for (iter = 0; iter < max_iter; ++iter) {
        for (int i = 0; i < n; ++i) {                    /* :73 */
            double t = 0.0;
            for (int j = 0; j < n; ++j) t = fma(AtA[IDX(i, j)], x[j], t);
            w[i] = Atb[i] - t;
        }
        if (st) st->aflops += 2.0 * n * n + n;
        int maxi = 0, nact = -1;
        for (int i = 0; i < n; ++i)                      /* :88-95 */
            if (active[i]) {
                ++nact;
                if (w[i] > w[maxi]) maxi = i;
            }
        active[maxi] = 0;                                /* :104 */
        if (!nact || w[maxi] < eps) break;               /* :113 */
        if (st) st->outer++;
        for (;;) {                                       /* :118 */
            int idx[LD], k = 0;
            for (int i = 0; i < n; ++i) if (!active[i]) idx[k++] = i;
            if (k == 0) { for (int i = 0; i < n; ++i) x[i] = s[i]; break; }
            double Ms[LD * LD], vs[LD], ss[LD];
            for (int r = 0; r < k; ++r) {
                vs[r] = Atb[idx[r]];
                for (int c = 0; c < k; ++c) Ms[IDX(r, c)] = AtA[IDX(idx[r], idx[c])];
            }
            if (llt_solve(Ms, vs, ss, k) >= 0 && st) st->llt_fail = 1;
            stats_solve(st, k);
            for (int i = 0, r = 0; i < n; ++i) s[i] = active[i] ? 0.0 : ss[r++];
            double alpha = DBL_MAX;
            int aidx = -1;
            for (int i = 0; i < n; ++i)                  /* :141-149 */
                if (!active[i] && s[i] <= 0.0) {
                    double tmp = -x[i] / (s[i] - x[i]);
                    if (tmp < alpha) { alpha = tmp; aidx = i; }
                }
            if (alpha == DBL_MAX) { for (int i = 0; i < n; ++i) x[i] = s[i]; break; }
            for (int i = 0; i < n; ++i) x[i] = fma(alpha, s[i] - x[i], x[i]);
            x[aidx] = 0.0;
            active[aidx] = 1;
        }
    }
    if (st && iter == max_iter && max_iter > 0) st->hit_max = 1;
}

void or_fnnls(const double *A, const double *b, double *x, int n, double eps,
              unsigned max_iter, or_nnls_stats *st)
{
    fnnls_mn(A, b, x, n, n, eps, max_iter, st);
}

/* ---- Lawson-Hanson (eigen_nnls.h) ---------------------------------------
 * Householder reflector of a column segment v[0..len) (Golub & Van Loan,
 * 5.1.2, the convention of LAPACK's xLARFG): beta = -sign(v0) |v|,
 * essential part v[1..) / (v0 - beta), tau = (beta - v0) / beta; a segment
 * whose tail is zero gets tau = 0 and beta = v0 (eigen_nnls.h:276-279 builds
 * it through Eigen's makeHouseholderInPlace).  col(r) reads row r of the
 * column. */
static void lh_make_householder(double *QR, int col, int k, int m, double *tau)
{
    double tail = 0.0;
    for (int r = k + 1; r < m; ++r) tail += QR[IDX(r, col)] * QR[IDX(r, col)];
    const double c0 = QR[IDX(k, col)];
    if (tail <= DBL_MIN) {
        *tau = 0.0;
        for (int r = k + 1; r < m; ++r) QR[IDX(r, col)] = 0.0;
        return;                          /* beta = c0 stays on the diagonal */
    }
    double beta = sqrt(c0 * c0 + tail);
    if (c0 >= 0.0) beta = -beta;
    for (int r = k + 1; r < m; ++r) QR[IDX(r, col)] /= (c0 - beta);
    *tau = (beta - c0) / beta;
    QR[IDX(k, col)] = beta;
}

/* applies reflector i (essential part below the diagonal of QR's column i,
 * coefficient tau) to rows i..m-1 of the vector v (stride vs) */
static void lh_apply(const double *QR, int i, double tau, int m, double *v, int vs)
{
    if (m - i == 1) {
        v[i * vs] *= 1.0 - tau;
        return;
    }
    if (tau == 0.0) return;
    double t = 0.0;
    for (int r = i + 1; r < m; ++r) t += QR[IDX(r, i)] * v[r * vs];
    t += v[i * vs];
    v[i * vs] -= tau * t;
    for (int r = i + 1; r < m; ++r) v[r * vs] -= tau * QR[IDX(r, i)] * t;
}

/* nnls_householder_qr_inplace_update (eigen_nnls.h:251-281): column k of QR
 * becomes A's column `acol`, the k reflectors before it are applied, and
 * reflector k is formed in place */
static void lh_qr_update(double *QR, double *tau, const double *A, int acol, int k, int m)
{
    for (int r = 0; r < m; ++r) QR[IDX(r, k)] = A[IDX(r, acol)];
    for (int i = 0; i < k; ++i) lh_apply(QR, i, tau[i], m, QR + k, LD);
    lh_make_householder(QR, k, k, m, &tau[k]);
}

/* eigen_nnls.h:336-414 */
static int lh_nnls_mn(const double *A, const double *b, double *x, int m, int n, double eps,
                      int max_iter, or_nnls_stats *st)
{
    double AtA[LD * LD], Atb[LD], w[LD], y[LD], QR[LD * LD], tau[LD], c[LD];
    int perm[LD];
    memset(AtA, 0, sizeof(AtA));
    memset(QR, 0, sizeof(QR));
    memset(tau, 0, sizeof(tau));
    tmul_mn(A, AtA, m, n);                                       /* :98 */
    atb_mn(A, b, Atb, m, n);                                     /* :352 */
    for (int i = 0; i < n; ++i) perm[i] = i;                     /* :348-349 */
    int np = 0, nls = 0;
    for (;;) {
        for (int i = 0; i < n; ++i) {                            /* :214-216 */
            double t = 0.0;
            for (int j = 0; j < n; ++j) t += AtA[IDX(i, j)] * x[j];
            w[i] = Atb[i] - t;
        }
        if (st) st->aflops += 2.0 * n * n;
        if (n == np) return 1;                                   /* :361 */
        int q = np;                                              /* :170-180 */
        for (int i = np + 1; i < n; ++i)
            if (w[perm[q]] < w[perm[i]]) q = i;
        if (w[perm[q]] - eps < 0) return 1;
        swap_i(perm, q, np);                                     /* :431-440 */
        ++np;
        lh_qr_update(QR, tau, A, perm[np - 1], np - 1, m);
        if (st) st->outer++;
        for (;;) {
            if (max_iter > 0 && nls >= max_iter) {               /* :375-377 */
                if (st) st->hit_max = 1;
                return 0;
            }
            for (int r = 0; r < m; ++r) c[r] = b[r];             /* :462-472 */
            for (int i = 0; i < np; ++i) lh_apply(QR, i, tau[i], m, c, 1);
            for (int i = np - 1; i >= 0; --i) {                  /* R c = Q^T b */
                c[i] /= QR[IDX(i, i)];
                for (int r = 0; r < i; ++r) c[r] -= c[i] * QR[IDX(r, i)];
            }
            for (int i = 0; i < n; ++i) y[i] = 0.0;
            for (int i = 0; i < np; ++i) y[perm[i]] = c[i];
            ++nls;
            stats_solve(st, np);
            if (g_mg && g_mg->on) {   /* the normal matrix of the passive columns */
                double Mp[LD * LD];
                for (int a = 0; a < np; ++a)
                    for (int b2 = 0; b2 < np; ++b2) Mp[IDX(a, b2)] = AtA[IDX(perm[a], perm[b2])];
                record_solve(g_mg, Mp, c, np);
            }
            int feasible = 1, rem = 0;                           /* :384-396 */
            double alpha = DBL_MAX;
            for (int i = 0; i < np; ++i) {
                const int idx = perm[i];
                if (y[idx] <= 0) {
                    const double t = -x[idx] / (y[idx] - x[idx]);
                    if (alpha > t) {
                        alpha = t;
                        rem = i;
                    }
                    feasible = 0;
                }
            }
            if (feasible) {                                      /* :398-401 */
                for (int i = 0; i < n; ++i) x[i] = y[i];
                break;
            }
            for (int i = 0; i < np; ++i) {                       /* :403-407 */
                const int idx = perm[i];
                x[idx] += alpha * (y[idx] - x[idx]);
            }
            swap_i(perm, rem, np - 1);                           /* :443-456 */
            --np;
            for (int i = rem; i < np; ++i) lh_qr_update(QR, tau, A, perm[i], i, m);
        }
    }
}

int or_lh_nnls(const double *A, const double *b, double *x, int n, double eps, int max_iter,
               or_nnls_stats *st)
{
    return lh_nnls_mn(A, b, x, n, n, eps, max_iter, st);
}

/* ------------------------------------------------------------------------ */
/* per-channel multifit                                                      */
/* ------------------------------------------------------------------------ */

static int imax(int a, int b) { return a > b ? a : b; }

/* PulseChiSqSNNLS.cpp:50-62: column j holds fullpulse[f+o ...] in rows f.. */
void or_build_pulsemat(const or_chtype *t, double *P)
{
    memset(P, 0, sizeof(double) * LD * LD);
    for (int j = 0; j < t->npulses; ++j) {
        int bx = t->bx[j];
        int f = imax(0, bx + t->first_off);
        int o = t->pulse_off - bx;
        for (int i = f; i < t->nsamples; ++i) P[IDX(i, j)] = t->fullpulse[i + o];
    }
}

typedef struct {
    const or_chtype *t;
    const or_opts *o;
    double P[LD * LD];
    double s[LD];        /* _sampvec */
    double x[LD];        /* _ampvec  */
    double C[LD * LD];   /* _invcov (lower) */
    double L[LD * LD];   /* _covdecomp.matrixL() */
    double chisq;
    int flags;
    int has_pi;          /* fullpulsecov has a nonzero entry (flop count only) */
    or_nnls_stats st;
    /* legacy solver: the column order of _pulsemat/_ampvec/_bxs (position ->
     * pulse index) and _nP, both kept from one NNLS call to the next */
    int lperm[LD];
    int lnp;
    /* decision margins of the main fit, and its last NNLS problem (normal
     * equations and final passive set) for the rounding estimate */
    margins mg;
    double m_chi2;
    double lAtA[LD * LD], lAtb[LD];
    unsigned lpm;
    /* or_dofit_from_set: the first NNLS call ends at the set force_pm */
    int force;
    unsigned force_pm;
    double force_margin;
    int force_feasible;
} fitctx;

/* updateCov, PulseChiSqSNNLS.cpp:169-194 */
static void update_cov(fitctx *c)
{
    const or_chtype *t = c->t;
    int ns = t->nsamples;
    double p2 = t->pederr * t->pederr;
    memset(c->C, 0, sizeof(c->C));
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j <= i; ++j) c->C[IDX(i, j)] = p2 * t->noisecor[IDX(i, j)];
    for (int a = 0; a < t->npulses; ++a) {    /* column order of _pulsemat */
        const int p = c->lperm[a];
        if (c->x[p] == 0.0) continue;
        int bx = t->bx[p];
        int f = imax(0, bx + t->first_off);
        int o = t->pulse_off - bx;
        double a2 = c->x[p] * c->x[p];
        for (int i = f; i < ns; ++i)
            for (int j = f; j <= i; ++j)
                c->C[IDX(i, j)] = fma(a2, t->fullpulsecov[(i + o) * OR_FULL + (j + o)],
                                      c->C[IDX(i, j)]);
    }
    if (chol_core(c->C, c->L, LD, ns, 1) >= 0) c->flags |= OR_F_LLT_FAIL;
    if (c->o->count_flops) {
        const int n = ns, p = t->npulses;
        double f = n * (n + 1) / 2.0 + chol_flops(n) + (double)p * n * n + (double)n * n +
                   (double)n * p * (p + 1) + 2.0 * n * p + (2.0 * n * p + n + (double)n * n + 2.0 * n);
        if (c->has_pi)
            for (int q = 0; q < p; ++q) {
                if (c->x[q] == 0.0) continue;
                const int m = ns - imax(0, t->bx[q] + t->first_off);
                if (m > 0) f += (double)m * (m + 1);
            }
        c->st.aflops += f;
    }
}

/* Legacy NNLS (legacy_multifit_cpu/src/PulseChiSqSNNLS.cpp:268-363) on the
 * whitened problem (A, b) of the current Minimize iteration.  The legacy code
 * keeps its pulse columns permuted between calls, so aTamat and aTbvec are
 * formed in the persisted order; here AtA and Atb are formed in pulse order
 * (the same dot products) and permuted.  Eigen's pivoted LDLT of the leading
 * block (:6-72) is restated as the Cholesky solve of the same block: the
 * same solution, different rounding. */
static void legacy_nnls(fitctx *c, const double *A, const double *b, int m, int n,
                        or_nnls_stats *st)
{
    double AtA0[LD * LD], Atb0[LD], M[LD * LD], v[LD], xp[LD], w[LD], sv[LD];
    memset(AtA0, 0, sizeof(AtA0));
    tmul_mn(A, AtA0, m, n);
    atb_mn(A, b, Atb0, m, n);
    int *perm = c->lperm;
    memset(M, 0, sizeof(M));
    for (int a = 0; a < n; ++a) {
        for (int q = 0; q < n; ++q) M[IDX(a, q)] = AtA0[IDX(perm[a], perm[q])];
        v[a] = Atb0[perm[a]];
        xp[a] = c->x[perm[a]];
    }
    unsigned iter;
    for (iter = 0; iter < c->o->max_iter; ++iter) {              /* :285 */
        if (iter > 0 || c->lnp == 0) {                           /* :287 */
            if (c->lnp == n) break;                              /* :288 */
            int np = c->lnp;
            for (int i = np; i < n; ++i) {                       /* :293-295 */
                double t = 0.0;
                for (int j = 0; j < n; ++j) t = fma(M[IDX(i, j)], xp[j], t);
                w[i] = v[i] - t;
            }
            if (st) st->aflops += 2.0 * n * n + (n - np);
            int q = np;                                          /* :297 maxCoeff */
            double wmax = w[np];
            for (int i = np + 1; i < n; ++i)
                if (w[i] > wmax) { wmax = w[i]; q = i; }
            if (wmax < c->o->eps) break;                         /* :300 */
            sym_swap(M, np, q);                                  /* :306-311 */
            swap_d(v, np, q);
            swap_d(xp, np, q);
            swap_i(perm, np, q);
            ++c->lnp;
            if (st) st->outer++;
        }
        while (c->lnp > 0) {                                     /* :318 */
            const int np = c->lnp;
            int f = llt_solve(M, v, sv, np);                     /* :324 */
            if (f >= 0 && st) st->llt_fail = 1;
            stats_solve(st, np);
            double mins = sv[0];                                 /* :326 */
            for (int i = 1; i < np; ++i) if (sv[i] < mins) mins = sv[i];
            if (mins > 0.0) {
                for (int i = 0; i < np; ++i) xp[i] = sv[i];
                break;
            }
            int idx = 0;                                         /* :331-341 */
            double minratio = DBL_MAX;
            for (int i = 0; i < np; ++i)
                if (sv[i] <= 0.0) {
                    double ratio = xp[i] / (xp[i] - sv[i]);
                    if (ratio < minratio) { minratio = ratio; idx = i; }
                }
            for (int i = 0; i < np; ++i)                         /* :343-344 */
                xp[i] = fma(minratio, sv[i] - xp[i], xp[i]);
            xp[idx] = 0.0;                                       /* :347 */
            sym_swap(M, np - 1, idx);                            /* :353-358 */
            swap_d(v, np - 1, idx);
            swap_d(xp, np - 1, idx);
            swap_i(perm, np - 1, idx);
            --c->lnp;
        }
    }
    if (iter == c->o->max_iter && c->o->max_iter > 0 && st) st->hit_max = 1;
    for (int a = 0; a < n; ++a) c->x[perm[a]] = xp[a];
}

/* KKT test of passive set pm on normal equations (AtA, Atb): the set's
 * system solved in long double into z (0 outside the set); *feasible: every
 * amplitude of the set > 0 (0 also when the block is not positive
 * definite).  Returns the largest (w_i - eps) / unit_i over the indices
 * outside the set, w = Atb - AtA z, unit_i = u (p + 1) (|Atb_i| +
 * sum_j |AtA_ij z_j|), the FP64 rounding of w_i. */
static double kkt_of_set(const double *AtA, const double *Atb, int np, unsigned pm,
                         double eps, double *z, int *feasible)
{
    int idx[LD], k = 0;
    for (int i = 0; i < np; ++i)
        if ((pm >> i) & 1u) idx[k++] = i;
    long double L[LD][LD], y[LD], zz[LD], zl[LD];
    for (int i = 0; i < LD; ++i) { zl[i] = 0; z[i] = 0.0; }
    *feasible = 1;
    for (int a = 0; a < k; ++a)
        for (int b2 = 0; b2 <= a; ++b2) {
            long double v = AtA[IDX(idx[a], idx[b2])];
            for (int q = 0; q < b2; ++q) v -= L[a][q] * L[b2][q];
            if (a == b2) {
                if (!(v > 0)) { *feasible = 0; return INFINITY; }
                L[a][a] = sqrtl(v);
            } else {
                L[a][b2] = v / L[b2][b2];
            }
        }
    for (int a = 0; a < k; ++a) {
        long double v = Atb[idx[a]];
        for (int q = 0; q < a; ++q) v -= L[a][q] * y[q];
        y[a] = v / L[a][a];
    }
    for (int a = k - 1; a >= 0; --a) {
        long double v = y[a];
        for (int q = a + 1; q < k; ++q) v -= L[q][a] * zz[q];
        zz[a] = v / L[a][a];
    }
    for (int a = 0; a < k; ++a) {
        zl[idx[a]] = zz[a];
        z[idx[a]] = (double)zz[a];
        if (!(zz[a] > 0)) *feasible = 0;
    }
    double worst = -INFINITY;
    for (int i = 0; i < np; ++i) {
        if ((pm >> i) & 1u) continue;
        long double w = Atb[i], mag = fabs(Atb[i]);
        for (int j = 0; j < np; ++j) {
            w -= (long double)AtA[IDX(i, j)] * zl[j];
            mag += fabsl((long double)AtA[IDX(i, j)] * zl[j]);
        }
        const double unit = OR_U * (np + 1) * (double)mag;
        const double m = (double)(w - eps) / (unit > 0 ? unit : DBL_MIN);
        if (m > worst) worst = m;
    }
    return worst;
}

/* NNLS, PulseChiSqSNNLS.cpp:219-274 */
static void nnls_step(fitctx *c)
{
    const or_chtype *t = c->t;
    int ns = t->nsamples, np = t->npulses;
    double A[LD * LD], b[LD], col[LD], y[LD];
    memset(A, 0, sizeof(A));
    for (int j = 0; j < np; ++j) {                        /* :235 */
        for (int i = 0; i < ns; ++i) col[i] = c->P[IDX(i, j)];
        or_forward_subst(c->L, col, y, LD, ns);
        for (int i = 0; i < ns; ++i) A[IDX(i, j)] = y[i];
    }
    or_forward_subst(c->L, c->s, b, LD, ns);              /* :236 */
    if (c->force) {   /* or_dofit_from_set: this call ends at the given set */
        double AtA[LD * LD], Atb[LD], z[LD];
        memset(AtA, 0, sizeof(AtA));
        tmul_mn(A, AtA, ns, np);
        atb_mn(A, b, Atb, ns, np);
        c->force_margin = kkt_of_set(AtA, Atb, np, c->force_pm, c->o->eps, z,
                                     &c->force_feasible);
        for (int i = 0; i < np; ++i) c->x[i] = z[i];
        if (c->mg.on) {   /* the set's block sets its amplitudes: its conditioning */
            double B[LD * LD];
            int idx[LD], k = 0;
            memset(B, 0, sizeof(B));
            for (int i = 0; i < np; ++i)
                if ((c->force_pm >> i) & 1u) idx[k++] = i;
            for (int a = 0; a < k; ++a)
                for (int b2 = 0; b2 < k; ++b2) B[IDX(a, b2)] = AtA[IDX(idx[a], idx[b2])];
            const double rc = k ? rcond1(B, k) : INFINITY;
            for (int a = 0; a < k; ++a) c->mg.rc_of[idx[a]] = rc;
            mmin(&c->mg.rc_min, rc);
            c->mg.tracked = 1;
        }
        c->force = 0;
        return;
    }
    double xx[LD];
    for (int i = 0; i < np; ++i) xx[i] = c->x[i];         /* :258 */
    or_nnls_stats st;
    memset(&st, 0, sizeof(st));
    unsigned lpm = 0;
    g_mg = c->mg.on && c->o->solver != OR_SOLVER_INPLACE && c->o->solver != OR_SOLVER_INPLACE_CHOL
               ? &c->mg : NULL;
    if (c->o->solver == OR_SOLVER_INPLACE && c->mg.on) g_mg = &c->mg;  /* its LLT solves */
    if (c->o->solver == OR_SOLVER_LEGACY) {
        legacy_nnls(c, A, b, ns, np, &st);
        for (int i = 0; i < np; ++i) xx[i] = c->x[i];
    } else
    switch (c->o->solver) {
    case OR_SOLVER_LH:                                    /* :253-257 */
        if (!lh_nnls_mn(A, b, xx, ns, np, c->o->eps, (int)c->o->max_iter, &st)) st.hit_max = 1;
        break;
    case OR_SOLVER_INPLACE_CHOL:
        inplace_fnnls_mn(A, b, xx, ns, np, c->o->eps, c->o->max_iter, &st, 1, &c->mg, &lpm);
        break;
    case OR_SOLVER_FNNLS:
        fnnls_mn(A, b, xx, ns, np, c->o->eps, c->o->max_iter, &st);
        break;
    default:
        inplace_fnnls_mn(A, b, xx, ns, np, c->o->eps, c->o->max_iter, &st, 0, &c->mg, &lpm);
        break;
    }
    if (c->mg.on && st.inner > 0) {
        /* the last call that solved sets the final amplitudes: later calls
         * that stop at once leave them */
        memset(c->lAtA, 0, sizeof(c->lAtA));
        tmul_mn(A, c->lAtA, ns, np);
        atb_mn(A, b, c->lAtb, ns, np);
        c->lpm = lpm;
    }
    g_mg = NULL;
    for (int i = 0; i < np; ++i) c->x[i] = xx[i];         /* :273 */
    c->st.outer += st.outer;
    c->st.inner += st.inner;
    c->st.ksum += st.ksum;
    c->st.k2sum += st.k2sum;
    c->st.k3sum += st.k3sum;
    c->st.aflops += st.aflops;
    if (st.hit_max) c->flags |= OR_F_FNNLS_CAP;
    if (st.llt_fail) c->flags |= OR_F_LLT_FAIL;
}

/* ComputeChiSq, PulseChiSqSNNLS.cpp:196-200: |L^-1 (P x - s)|^2 */
static double compute_chisq(const fitctx *c)
{
    int ns = c->t->nsamples, np = c->t->npulses;
    double r[LD] = {0}, y[LD];
    for (int i = 0; i < ns; ++i) {
        double t = 0.0;
        for (int a = 0; a < np; ++a) {
            const int j = c->lperm[a];
            t = fma(c->P[IDX(i, j)], c->x[j], t);
        }
        r[i] = t - c->s[i];
    }
    or_forward_subst(c->L, r, y, LD, ns);
    double q = 0.0;
    for (int i = 0; i < ns; ++i) q = fma(y[i], y[i], q);
    return q;
}

/* Minimize, PulseChiSqSNNLS.cpp:149-166 */
static int minimize(fitctx *c, int *iters)
{
    int it;
    for (it = 0; it < c->o->max_minimize; ++it) {
        update_cov(c);
        nnls_step(c);
        double now = compute_chisq(c);
        double d = now - c->chisq;
        c->chisq = now;
        if (c->mg.on) {   /* in units of the rounding of the two chi2 values */
            const double sc = fabs(now) > 1.0 ? fabs(now) : 1.0;
            const double m = fabs(fabs(d) - c->o->chi2_tol) /
                             rnd_units(2.0 * sc, c->t->nsamples, c->mg.rc_last);
            if (m < c->m_chi2) c->m_chi2 = m;
        }
        if (fabs(d) < c->o->chi2_tol) {
            if (iters) *iters = it + 1;
            return 1;
        }
    }
    if (iters) *iters = it;
    c->flags |= OR_F_MINIMIZE_CAP;
    return 1;
}

/* ComputeApproxUncertainty, PulseChiSqSNNLS.cpp:202-207 */
static double approx_uncertainty(const fitctx *c, int ip)
{
    int ns = c->t->nsamples;
    double col[LD], y[LD];
    for (int i = 0; i < ns; ++i) col[i] = c->P[IDX(i, ip)];
    or_forward_subst(c->L, col, y, LD, ns);
    double q = 0.0;
    for (int i = 0; i < ns; ++i) q = fma(y[i], y[i], q);
    return 1.0 / sqrt(q);
}

/* Rounding estimate of a fit's amplitudes: the last NNLS problem's passive
 * system (pm) solved again in long double; the largest difference from the
 * FP64 amplitudes, relative to the largest amplitude (floor 1).  Zero when
 * no passive set was recorded (the solvers other than inplace_fnnls). */
static double rounding_estimate(const double *AtA, const double *Atb, unsigned pm,
                                const double *x, int np)
{
    int idx[LD], k = 0;
    for (int i = 0; i < np; ++i)
        if ((pm >> i) & 1u) idx[k++] = i;
    if (!k) return 0.0;
    long double L[LD][LD], y[LD], z[LD];
    for (int a = 0; a < k; ++a)
        for (int b2 = 0; b2 <= a; ++b2) {
            long double v = AtA[IDX(idx[a], idx[b2])];
            for (int q = 0; q < b2; ++q) v -= L[a][q] * L[b2][q];
            if (a == b2) {
                if (!(v > 0)) return INFINITY;
                L[a][a] = sqrtl(v);
            } else {
                L[a][b2] = v / L[b2][b2];
            }
        }
    for (int a = 0; a < k; ++a) {
        long double v = Atb[idx[a]];
        for (int q = 0; q < a; ++q) v -= L[a][q] * y[q];
        y[a] = v / L[a][a];
    }
    for (int a = k - 1; a >= 0; --a) {
        long double v = y[a];
        for (int q = a + 1; q < k; ++q) v -= L[q][a] * z[q];
        z[a] = v / L[a][a];
    }
    double sc = 1.0, e = 0.0;
    for (int i = 0; i < np; ++i)
        if (fabs(x[i]) > sc) sc = fabs(x[i]);
    for (int a = 0; a < k; ++a) {
        const double d = fabs((double)(z[a] - (long double)x[idx[a]]));
        if (d > e) e = d;
    }
    return e / sc;
}

/* DoFit, PulseChiSqSNNLS.cpp:21-147.  force_pm (or_dofit_from_set): the
 * first NNLS call ends at that passive set's exact solution instead. */
static int dofit_impl(const or_chtype *t, const double *samples, const or_opts *o,
                      or_result *r, const unsigned *force_pm, double *margin, int *feasible)
{
    fitctx c;
    memset(&c, 0, sizeof(c));
    c.t = t;
    c.o = o;
    if (force_pm) {
        c.force = 1;
        c.force_pm = *force_pm;
        c.force_margin = INFINITY;
    }
    if (o->count_flops)
        for (int i = 0; i < OR_FULL * OR_FULL; ++i)
            if (t->fullpulsecov[i] != 0.0) { c.has_pi = 1; break; }
    int ns = t->nsamples, np = t->npulses;
    memset(r, 0, sizeof(*r));
    for (int i = 0; i < ns; ++i) {
        c.s[i] = samples[i];
        if (!isfinite(samples[i])) c.flags |= OR_F_NONFINITE;
    }
    or_build_pulsemat(t, c.P);
    for (int i = 0; i < LD; ++i) c.lperm[i] = i;          /* legacy: _nP = 0, pulse order */
    c.lnp = 0;
    margins_init(&c.mg);
    c.mg.on = o->record_margins;
    c.m_chi2 = INFINITY;
    c.lpm = 0;
    int nmin = 0;
    int status = minimize(&c, &nmin);                    /* :65 */
    if (force_pm) {
        *margin = c.force_margin;
        *feasible = c.force ? 0 : c.force_feasible;   /* still set: never reached */
    }
    r->m_eps = c.mg.m_eps;
    r->m_arg = c.mg.m_arg;
    r->m_pos = c.mg.m_pos;
    r->m_alpha = c.mg.m_alpha;
    r->m_chi2 = c.m_chi2;
    /* rcond behind the final amplitudes: inplace_fnnls tracks the solve
     * that last set each amplitude (stale ones included); the other solvers
     * report the worst block they solved */
    r->pivot = INFINITY;
    if (c.mg.on) {
        if (c.mg.tracked) {
            for (int i = 0; i < np; ++i)
                if (c.x[i] != 0.0) mmin(&r->pivot, c.mg.rc_of[i]);
        } else {
            r->pivot = c.mg.rc_min;
        }
    }
    r->xerr = c.mg.on ? rounding_estimate(c.lAtA, c.lAtb, c.lpm, c.x, np) : 0.0;
    c.mg.on = 0;
    r->n_minimize = nmin;
    for (int i = 0; i < np; ++i) r->x[i] = c.x[i];       /* :66 _ampvecmin */
    r->n_outer = c.st.outer;
    r->n_inner = c.st.inner;
    r->ksum = c.st.ksum;
    r->k2sum = c.st.k2sum;
    r->k3sum = c.st.k3sum;
    r->aflops = c.st.aflops;

    if (status && o->compute_errors) {                   /* :70-146 */
        int ipit = -1;
        for (int i = 0; i < np; ++i)
            if (t->bx[i] == 0) { ipit = i; break; }
        if (ipit >= 0) {
            const int ipitmin = ipit;
            double approxerr = approx_uncertainty(&c, ipit);
            double chisq0 = c.chisq;
            double x0 = r->x[ipit];
            /* :96-102 compares against _nP, which the new NNLS never moves
             * off zero, so the column swap never happens; the legacy NNLS
             * keeps _nP, and a passive in-time pulse is swapped to the last
             * passive position and leaves the passive set
             * (legacy_multifit_cpu/src/PulseChiSqSNNLS.cpp:151-157) */
            if (o->solver == OR_SOLVER_LEGACY) {
                int pos = 0;
                while (c.lperm[pos] != ipit) ++pos;
                if (pos < c.lnp) {
                    swap_i(c.lperm, pos, c.lnp - 1);
                    --c.lnp;
                }
            }
            double pulse[LD];
            for (int i = 0; i < ns; ++i) {
                pulse[i] = c.P[IDX(i, ipit)];
                c.P[IDX(i, ipit)] = 0.0;
            }
            double xplus = x0 + approxerr;
            c.x[ipit] = xplus;
            for (int i = 0; i < ns; ++i) c.s[i] = samples[i] - c.x[ipit] * pulse[i];
            status &= minimize(&c, NULL);
            double chiplus = compute_chisq(&c);
            double sigplus = fabs(xplus - x0) / sqrt(chiplus - chisq0);
            if ((x0 / sigplus) > 0.5) {
                double xminus = x0 - approxerr;
                if (xminus < 0.0) xminus = 0.0;
                c.x[ipit] = xminus;
                for (int i = 0; i < ns; ++i) c.s[i] = samples[i] - c.x[ipit] * pulse[i];
                status &= minimize(&c, NULL);
                double chiminus = compute_chisq(&c);
                double sigminus = fabs(xminus - x0) / sqrt(chiminus - chisq0);
                r->err[ipitmin] = 0.5 * (sigplus + sigminus);
            } else {
                r->err[ipitmin] = sigplus;
            }
            c.chisq = chisq0;
        }
    }
    r->chi2 = c.chisq;
    r->status = status;
    r->flags = c.flags;
    return status;
}

int or_dofit(const or_chtype *t, const double *samples, const or_opts *o, or_result *r)
{
    return dofit_impl(t, samples, o, r, NULL, NULL, NULL);
}

double or_dofit_from_set(const or_chtype *t, const double *samples, const or_opts *o,
                         unsigned pm, or_result *r, int *feasible)
{
    double margin = INFINITY;
    dofit_impl(t, samples, o, r, &pm, &margin, feasible);
    return margin;
}

void or_kkt_batch(const or_chtype *types, const uint8_t *type_of, const double *samples,
                  const double *pederr, long long nch, const long long *idx, long long k,
                  const uint32_t *pm, const or_opts *o, or_result *res, double *margin,
                  int *feasible)
{
    for (long long a = 0; a < k; ++a) {
        const long long c = idx[a];
        or_chtype tc = types[type_of ? type_of[c] : 0];
        if (pederr) tc.pederr = pederr[c];
        double s[LD];
        for (int i = 0; i < tc.nsamples; ++i) s[i] = samples[i * nch + c];
        margin[a] = or_dofit_from_set(&tc, s, o, pm[a], res + a, feasible + a);
    }
}

/* ------------------------------------------------------------------------ */
/* threaded batch (CPU baseline)                                             */
/* ------------------------------------------------------------------------ */

typedef struct {
    const or_chtype *types;
    const uint8_t *type_of;
    const double *samples;
    const double *pederr;
    long long nch;
    const or_opts *o;
    or_result *res;
    atomic_llong next;
} batchctx;

static void *batch_worker(void *arg)
{
    batchctx *b = (batchctx *)arg;
    const long long chunk = 256;
    for (;;) {
        long long c0 = atomic_fetch_add(&b->next, chunk);
        if (c0 >= b->nch) break;
        long long c1 = c0 + chunk < b->nch ? c0 + chunk : b->nch;
        for (long long c = c0; c < c1; ++c) {
            const or_chtype *t = b->types + (b->type_of ? b->type_of[c] : 0);
            double s[LD];
            for (int i = 0; i < t->nsamples; ++i) s[i] = b->samples[i * b->nch + c];
            if (b->pederr) {   /* DoFit's pederr argument, channel by channel */
                or_chtype tc = *t;
                tc.pederr = b->pederr[c];
                or_dofit(&tc, s, b->o, b->res + c);
            } else {
                or_dofit(t, s, b->o, b->res + c);
            }
        }
    }
    return NULL;
}

void or_dofit_batch(const or_chtype *types, const uint8_t *type_of,
                    const double *samples, long long nch, const or_opts *o,
                    or_result *res, int nthreads)
{
    or_dofit_batch_pederr(types, type_of, samples, NULL, nch, o, res, nthreads);
}

void or_dofit_batch_pederr(const or_chtype *types, const uint8_t *type_of,
                           const double *samples, const double *pederr, long long nch,
                           const or_opts *o, or_result *res, int nthreads)
{
    batchctx b;
    b.types = types;
    b.type_of = type_of;
    b.samples = samples;
    b.pederr = pederr;
    b.nch = nch;
    b.o = o;
    b.res = res;
    atomic_init(&b.next, 0);
    if (nthreads <= 1) {
        batch_worker(&b);
        return;
    }
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, batch_worker, &b);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
}

