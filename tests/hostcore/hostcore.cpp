// hostcore.cpp -- TEST HARNESS: the product's per-channel fit code
// (test-regression_amd/csrc/fit_core.hpp) compiled for the host CPU, so that
// CPU-only test runs can compare the kernel's arithmetic with the oracle.  The
// product library never uses this build; on a GPU the tests call the gfx950
// kernels through the C ABI instead.
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/mf_abi.h"
#include "../../test-regression_amd/csrc/fit_core.hpp"
#include "../../test-regression_amd/csrc/types_host.hpp"

using namespace mf;

static bool use_tables = true;
// 2: the per-channel iterations built for the pulse layout D = 2 (lay_first),
// as the kernels run types of that layout; -1: the generic build
static int g_layout = -1;
extern "C" void hc_set_layout(int d) { g_layout = d == 2 ? 2 : -1; }

// How the later iterations of a live pulse covariance run: 0 in one body
// (fast_p2_rest, as k_fast_p2); 1 split as k_fast_p3 + k_fast_p4 (the lazy
// iterations, then the whole body from the first iteration that needs the
// normal equations); 2 the whole body from iteration 1 (fast_p2_resume alone).
static int g_split = 0;
static long long g_pending = 0;
extern "C" void hc_set_split(int mode) { g_split = mode; }
extern "C" long long hc_pending_count(void)
{
    const long long n = g_pending;
    g_pending = 0;
    return n;
}

template <int N, int D>
static void fit_fast_split(const DevType &dt, const float (&s_in)[10], const FitParams &prm,
                           FitOut<float> &out)
{
    int flags = 0;
    FnnlsStats st;
    float s[10], x[10];
    uint32_t pm = 0;
    for (int i = 0; i < 10; ++i) {
        s[i] = s_in[i];
        x[i] = 0.0f;
    }
    if (all_finite<float, N>(s) && rho_ok(prm.rho) && prm.max_minimize > 0)
        fast_phase1<N, N>(dt, s, prm, x, pm, st, flags);
    ActiveState<double> S;
    double chi2 = 0.0;
    int nmin = 0;
    KeepSamples keep;
    flags &= F_FP64_PIVOT;
    if (fast_p2_first<N, N, true>(dt, s, prm, x, false, st, flags, S, chi2, keep)) {
        bool pend = true;
        int it = 1;
        if (g_split == 1) {
            nmin = fast_p2_rest_lazy<N, N, D>(dt, s, prm, S.x, chi2, st, flags, pend);
            it = nmin;
        } else if (prm.use_pi && fabs(chi2) < prm.chi2_tol) {
            nmin = 1;
            pend = false;
        } else if (!prm.use_pi) {
            nmin = fixed_cov_nmin(chi2, prm, flags);
            pend = false;
        }
        if (pend) {
            ++g_pending;
            RegPark<double> park;
            nmin = fast_p2_resume<N, N, D>(dt, s, prm, S.x, chi2, it, st, flags, park);
        }
    }
    fast_p2_out(S.x, chi2, flags, nmin, st, out);
}

template <int N, bool LIVE>
static void run_fast(const DevType &dt, const double *samples, int64_t nch, int64_t c,
                     const FitParams &prm, double *x, double *chi2, int32_t *flags,
                     int32_t *iters)
{
    float s[10], x0[10];
    for (int i = 0; i < 10; ++i) {
        s[i] = i < N ? (float)samples[i * nch + c] : 0.0f;
        x0[i] = 0.0f;
    }
    RegPark<double> park;   // the live-covariance iterations park L^{-1} in FP64, as the kernel's LDS
    FitOut<float> out;
    if (LIVE && g_split && use_tables && g_layout == 2)
        fit_fast_split<N, 2>(dt, s, prm, out);
    else if (LIVE && g_split && use_tables)
        fit_fast_split<N, -1>(dt, s, prm, out);
    else if (use_tables && g_layout == 2)
        fit_fast<N, N, LIVE, true, 2>(dt, s, prm, x0, 0u, false, park, out);
    else if (use_tables)
        fit_fast<N, N, LIVE, true>(dt, s, prm, x0, 0u, false, park, out);
    else
        fit_fast<N, N, LIVE, false>(dt, s, prm, x0, 0u, false, park, out);
    for (int p = 0; p < 10; ++p) x[p * nch + c] = p < N ? (double)out.x[p] : 0.0;
    chi2[c] = (double)out.chi2;
    flags[c] = out.flags;
    iters[c] = out.n_min | (out.n_outer << 8) | (out.n_inner << 16);
}

static double *g_err = nullptr;

template <typename T, int N, bool E, int XS>
static void plain_d(const DevType &dt, const T (&s)[10], const FitParams &prm, RegPark<T> &park,
                    FitOut<T> &out, int ip)
{
    if (g_layout == 2)
        fit_plain<T, N, N, E, XS, 2>(mats<T>(dt), s, prm, park, out, ip);
    else
        fit_plain<T, N, N, E, XS, -1>(mats<T>(dt), s, prm, park, out, ip);
}

template <typename T, int N>
static void run(const DevType &dt, const double *samples, int64_t nch, int64_t c,
                const FitParams &prm, double *x, double *chi2, int32_t *flags, int32_t *iters)
{
    T s[10];
    for (int i = 0; i < 10; ++i) s[i] = i < N ? (T)samples[i * nch + c] : T(0);
    RegPark<T> park;
    FitOut<T> out;
    const int ip = g_err ? dt.h.intime : -1;
    if (g_err && prm.solver == 2)
        plain_d<T, N, true, XS_LEGACY>(dt, s, prm, park, out, ip);
    else if (g_err && prm.solver == 3)
        plain_d<T, N, true, XS_LH>(dt, s, prm, park, out, ip);
    else if (g_err && prm.solver == 4)
        plain_d<T, N, true, XS_CHOL>(dt, s, prm, park, out, ip);
    else if (g_err)
        plain_d<T, N, true, XS_NONE>(dt, s, prm, park, out, ip);
    else if (prm.solver == 2)
        plain_d<T, N, false, XS_LEGACY>(dt, s, prm, park, out, -1);
    else if (prm.solver == 3)
        plain_d<T, N, false, XS_LH>(dt, s, prm, park, out, -1);
    else if (prm.solver == 4)
        plain_d<T, N, false, XS_CHOL>(dt, s, prm, park, out, -1);
    else
        plain_d<T, N, false, XS_NONE>(dt, s, prm, park, out, -1);
    for (int p = 0; p < 10; ++p) x[p * nch + c] = p < N ? (double)out.x[p] : 0.0;
    if (g_err)
        for (int p = 0; p < 10; ++p) g_err[p * nch + c] = p < N ? (double)out.err[p] : 0.0;
    chi2[c] = (double)out.chi2;
    flags[c] = out.flags;
    iters[c] = out.n_min | (out.n_outer << 8) | (out.n_inner << 16);
}

static float eps_rel32 = 1e-6f, tau32 = 1e-3f;
static double eps_rel64 = 1e-15;
// Row m of the passive-block inverse table of the type's x = 0 normal
// matrix, as the kernels read it (MASK_PACK values: packed lower triangle,
// pivot ratio in the last slot), and the type's AtA0 (100 values,
// row-major).  Where the set also has a factor row (an ill-conditioned
// block), that row instead, and the return value is 2.  Returns 1 on a bad
// type.
extern "C" int hc_mask_inverse(const mf_chtype *type, int m, double *row, double *ata0)
{
    DevType d;
    if (expand_type(*type, d)) return 1;
    std::vector<MaskTab> t(1);
    build_mask_tab(d, t[0]);
    const bool f = (t[0].factored[m >> 5] >> (m & 31)) & 1u;
    for (int e = 0; e < MASK_PACK; ++e) row[e] = f ? t[0].fac[m][e] : t[0].inv[m][e];
    for (int i = 0; i < 100; ++i) ata0[i] = d.d.AtA0[i / 10][i % 10];
    return f ? 2 : 0;   // 2: the row is the set's factor row (fac)
}

// 1: the FP64 phase of the fast path reads the passive-block inverse tables
// (as the kernels do); 0: it forms every factor (solve_passive).
extern "C" void hc_set_use_tables(int on) { use_tables = on != 0; }

extern "C" void hc_set_tolerances(float e32, float t32, double e64)
{
    eps_rel32 = e32;
    tau32 = t32;
    eps_rel64 = e64;
}

// pederr (optional): per-channel noise sigma, as mf_batch.pederr.
extern "C" int hc_fit_batch2(const mf_chtype *types, int ntypes, const uint8_t *type_of,
                             const double *samples, const double *pederr, int64_t nch,
                             const mf_opts *o, double *x, double *chi2, int32_t *flags,
                             int32_t *iters, double *err)
{
    g_err = o->compute_errors ? err : nullptr;
    std::vector<DevType> dt((size_t)ntypes);
    std::vector<MaskTab> tabs((size_t)ntypes);
    for (int t = 0; t < ntypes; ++t) {
        if (expand_type(types[t], dt[(size_t)t])) return 1;
        if (use_tables) {
            build_mask_tab(dt[(size_t)t], tabs[(size_t)t]);
            dt[(size_t)t].tab = &tabs[(size_t)t];
        }
    }
    FitParams prm;
    prm.eps = o->eps;
    prm.chi2_tol = o->chi2_tol;
    prm.max_iter = o->max_iter;
    prm.max_minimize = o->max_minimize;
    prm.use_pi = o->use_pulse_cov;
    prm.solver = o->solver;
    prm.eps_rel32 = eps_rel32;
    prm.tau32 = tau32;
    prm.eps_rel64 = eps_rel64;
    const FitParams prm0 = prm;
    for (int64_t c = 0; c < nch; ++c) {
        const DevType &d = dt[type_of ? type_of[c] : 0];
        // the kernels' lane_params
        prm = prm0;
        prm.pmask = d.h.pmask;
        prm.use_pi = prm0.use_pi && d.h.has_pi;
        if (pederr) prm.rho = pederr[c] * pederr[c] * d.inv_pederr2;
        const bool f64 = o->precision == MF_PREC_FP64;
        const bool live = d.h.has_pi && o->use_pulse_cov;
        if (d.h.ns == 10) {
            if (f64) run<double, 10>(d, samples, nch, c, prm, x, chi2, flags, iters);
            else if (live) run_fast<10, true>(d, samples, nch, c, prm, x, chi2, flags, iters);
            else run_fast<10, false>(d, samples, nch, c, prm, x, chi2, flags, iters);
        } else {
            if (f64) run<double, 8>(d, samples, nch, c, prm, x, chi2, flags, iters);
            else if (live) run_fast<8, true>(d, samples, nch, c, prm, x, chi2, flags, iters);
            else run_fast<8, false>(d, samples, nch, c, prm, x, chi2, flags, iters);
        }
    }
    return 0;
}

extern "C" int hc_fit_batch(const mf_chtype *types, int ntypes, const uint8_t *type_of,
                            const double *samples, int64_t nch, const mf_opts *o, double *x,
                            double *chi2, int32_t *flags, int32_t *iters, double *err)
{
    return hc_fit_batch2(types, ntypes, type_of, samples, nullptr, nch, o, x, chi2, flags, iters,
                         err);
}

// The fast path from given starting amplitudes xs[10][nch] (passive set =
// entries > 0), as a warm start (warm = 1: straight to the FP64 phase) or
// cold (warm = 0: xs ignored, the FP32 phase first).
extern "C" int hc_fast_from(const mf_chtype *types, int ntypes, const uint8_t *type_of,
                            const double *samples, const double *xs, int64_t nch, int warm,
                            const mf_opts *o, double *x, double *chi2, int32_t *flags,
                            int32_t *iters)
{
    std::vector<DevType> dt((size_t)ntypes);
    std::vector<MaskTab> tabs((size_t)ntypes);
    for (int t = 0; t < ntypes; ++t) {
        if (expand_type(types[t], dt[(size_t)t])) return 1;
        build_mask_tab(dt[(size_t)t], tabs[(size_t)t]);
        dt[(size_t)t].tab = &tabs[(size_t)t];
    }
    FitParams prm0;
    prm0.eps = o->eps;
    prm0.chi2_tol = o->chi2_tol;
    prm0.max_iter = o->max_iter;
    prm0.max_minimize = o->max_minimize;
    prm0.use_pi = o->use_pulse_cov;
    prm0.solver = o->solver;
    prm0.eps_rel32 = eps_rel32;
    prm0.tau32 = tau32;
    prm0.eps_rel64 = eps_rel64;
    for (int64_t c = 0; c < nch; ++c) {
        const DevType &d = dt[type_of ? type_of[c] : 0];
        if (d.h.ns != 10 || (d.h.has_pi && o->use_pulse_cov)) return 2;
        FitParams prm = prm0;
        prm.pmask = d.h.pmask;
        prm.use_pi = 0;
        float s[10], x0[10];
        uint32_t pm = 0;
        for (int i = 0; i < 10; ++i) {
            s[i] = (float)samples[i * nch + c];
            x0[i] = warm && xs[i * nch + c] > 0 ? (float)xs[i * nch + c] : 0.0f;
            if (x0[i] > 0.0f) pm |= 1u << i;
        }
        RegPark<double> park;
        FitOut<float> out;
        fit_fast<10, 10, false, true>(d, s, prm, x0, pm, warm != 0, park, out);
        for (int p = 0; p < 10; ++p) x[p * nch + c] = (double)out.x[p];
        chi2[c] = (double)out.chi2;
        flags[c] = out.flags;
        iters[c] = out.n_min | (out.n_outer << 8) | (out.n_inner << 16);
    }
    return 0;
}

// single NNLS problems (A column-major ld 10, m x m) through fit_core's solvers
// (0 inplace_fnnls, 1 fnnls, 2 Lawson-Hanson with max_iter as a signed
// solve count, 3 inplace_fnnls_test_cholesky).  Returns 0 when Lawson-Hanson stopped at max_iter, else 1.
extern "C" int hc_nnls(const double *A, const double *b, double *x, int m, int solver,
                       double eps, unsigned max_iter)
{
    double AtA[10][10] = {{0}}, Atb[10] = {0}, xx[10] = {0};
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j <= i; ++j) {
            double t = 0;
            for (int k = 0; k < m; ++k) t = fma(A[i * 10 + k], A[j * 10 + k], t);
            AtA[i][j] = t;
        }
        double t = 0;
        for (int k = 0; k < m; ++k) t = fma(A[i * 10 + k], b[k], t);
        Atb[i] = t;
        xx[i] = x[i];
    }
    FnnlsStats st;
    bool ok = true;
#define CASE(M)                                                                   \
    case M:                                                                       \
        if (solver == 2) ok = lh_nnls<double, M>(AtA, Atb, xx, eps, (int)max_iter, st); \
        else if (solver == 3) fnnls_inplace_chol<double, M>(AtA, Atb, xx, eps, (int)max_iter, st); \
        else if (solver == 1) fnnls_mask<double, M>(AtA, Atb, xx, eps, (int)max_iter, st); \
        else fnnls_inplace<double, M>(AtA, Atb, xx, eps, 0.0, (int)max_iter, st); \
        break;
    switch (m) { CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) }
#undef CASE
    for (int i = 0; i < m; ++i) x[i] = xx[i];
    return ok ? 1 : 0;
}
