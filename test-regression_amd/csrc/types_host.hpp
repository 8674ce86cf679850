// types_host.hpp -- expands a channel type (mf_chtype) into the per-type
// tables the fit kernels read (mf::DevType).  Host code.
#pragma once

#include <math.h>
#include <string.h>

#include "../../include/mf_abi.h"
#include "fit_core.hpp"

namespace mf {

inline int imax_(int a, int b) { return a > b ? a : b; }

// Returns 0 or an error string.  A type of npulses < nsamples pulses (the
// reference's dynamic BXVector, multifit_cpu/interface/EigenMatrixTypes.h:9-10)
// occupies the first npulses of the kernel's nsamples pulse slots; the other
// slots hold zero columns: their w is 0, below any eps > 0, so they never
// enter the passive set and their amplitudes stay 0.
inline const char *expand_type(const mf_chtype &t, DevType &d)
{
    if (t.nsamples != 10 && t.nsamples != 8) return "nsamples must be 10 or 8";
    if (t.npulses < 1 || t.npulses > t.nsamples) return "npulses must be in 1..nsamples";
    if (!(t.pederr > 0.0) || !(t.pederr <= 1.7976931348623157e308))
        return "pederr must be positive and finite";
    memset(&d, 0, sizeof(d));
    const int ns = t.nsamples, np = t.npulses;
    d.h.ns = ns;
    d.h.np = ns;   // kernel pulse slots
    d.h.npv = np;
    d.h.pmask = (1u << np) - 1u;
    d.h.intime = -1;
    d.inv_pederr2 = 1.0 / (t.pederr * t.pederr);
    for (int p = np; p < 10; ++p) d.h.first[p] = ns;
    bool has_pi = false;
    for (int i = 0; i < MF_FULL * MF_FULL; ++i)
        if (t.fullpulsecov[i] != 0.0) has_pi = true;
    d.h.has_pi = has_pi ? 1 : 0;
    const double p2 = t.pederr * t.pederr;
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j < ns; ++j) d.d.C0[i][j] = p2 * t.noisecor[i * 10 + j];
    for (int p = 0; p < np; ++p) {
        const int bx = t.bx[p];
        if (bx == 0 && d.h.intime < 0) d.h.intime = p;
        const int f = imax_(0, bx + t.first_off);   // PulseChiSqSNNLS.cpp:54
        const int o = t.pulse_off - bx;             // PulseChiSqSNNLS.cpp:55
        if (f < ns && (f + o < 0 || ns - 1 + o >= MF_FULL))
            return "pulse window falls outside fullpulse (check bx, first_off, pulse_off)";
        d.h.first[p] = f;
        for (int i = f; i < ns; ++i) {
            d.d.P[i][p] = t.fullpulse[i + o];
            for (int j = f; j < ns; ++j)
                d.d.Pi[p][i][j] = t.fullpulsecov[(i + o) * MF_FULL + (j + o)];
        }
    }
    // x = 0 problem: L0 = chol(C0), Li0 = L0^{-1}, A0 = Li0 P, AtA0, G0
    {
        double L[10][10] = {{0}}, Z[10][10] = {{0}};
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j <= i; ++j) {
                double v = d.d.C0[i][j];
                for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
                if (i == j) {
                    if (!(v > 0)) return "noise covariance pederr^2 * noisecor is not positive definite";
                    L[i][i] = sqrt(v);
                } else {
                    L[i][j] = v / L[j][j];
                }
            }
        for (int i = 0; i < ns; ++i) {
            Z[i][i] = 1.0 / L[i][i];
            for (int j = 0; j < i; ++j) {
                double t = 0;
                for (int k = j; k < i; ++k) t += L[i][k] * Z[k][j];
                Z[i][j] = -t / L[i][i];
            }
        }
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j < ns; ++j) d.d.Li0[i][j] = Z[i][j];
        for (int i = 0; i < ns; ++i)
            for (int p = 0; p < np; ++p) {
                double t = 0;
                for (int k = 0; k <= i; ++k) t += Z[i][k] * d.d.P[k][p];
                d.d.A0[i][p] = t;
            }
        for (int p = 0; p < np; ++p)
            for (int q = 0; q < np; ++q) {
                double t = 0;
                for (int i = 0; i < ns; ++i) t += d.d.A0[i][p] * d.d.A0[i][q];
                d.d.AtA0[p][q] = t;
            }
        for (int p = 0; p < np; ++p)
            for (int k = 0; k < ns; ++k) {
                double t = 0;
                for (int i = k; i < ns; ++i) t += d.d.A0[i][p] * Z[i][k];
                d.d.G0[p][k] = t;
            }
    }
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j) {
            d.f.Li0[i][j] = (float)d.d.Li0[i][j];
            d.f.A0[i][j] = (float)d.d.A0[i][j];
            d.f.AtA0[i][j] = (float)d.d.AtA0[i][j];
            d.f.G0[i][j] = (float)d.d.G0[i][j];
            d.f.P[i][j] = (float)d.d.P[i][j];
            d.f.C0[i][j] = (float)d.d.C0[i][j];
            for (int p = 0; p < 10; ++p) d.f.Pi[p][i][j] = (float)d.d.Pi[p][i][j];
        }
    return 0;
}

// Fills t with the passive-block inverses of d's AtA0 (see MaskTab).  Each
// inverse is formed in extended precision (long double) from the FP64
// normal matrix the kernels use, then rounded; the pivot ratio is that of
// the FP64 masked Cholesky factorization solve_passive performs, in its
// operation order.
inline void build_mask_tab(const DevType &d, MaskTab &t)
{
    const int n = d.h.np;
    const double(&M)[10][10] = d.d.AtA0;
    memset(&t, 0, sizeof(t));
    for (int m = 1; m < (1 << n) && m < MASK_SETS; ++m) {
        if (m & ~d.h.pmask) {   // a slot past the type's pulses: never passive
            t.inv[m][MASK_PACK - 1] = NAN;
            continue;
        }
        double L[10][10] = {{0}};
        double ratio = INFINITY;
        for (int k = 0; k < n; ++k) {
            const bool ink = (m >> k) & 1;
            double sq = 0.0;
            for (int j = 0; j < k; ++j) sq = fma(L[k][j], L[k][j], sq);
            const double dk = ink ? M[k][k] - sq : 1.0;
            if (ink) {
                const double r = dk / M[k][k];
                if (!(r >= ratio)) ratio = r;   // keeps a NaN
            }
            const double rk = 1.0 / sqrt(dk);
            for (int i = k + 1; i < n; ++i) {
                const bool ini = (m >> i) & 1;
                double u = 0.0;
                for (int j = 0; j < k; ++j) u = fma(L[i][j], L[k][j], u);
                L[i][k] = (ink && ini) ? (M[i][k] - u) * rk : 0.0;
            }
        }
        int idx[10], k = 0;
        for (int i = 0; i < n; ++i)
            if ((m >> i) & 1) idx[k++] = i;
        long double C[10][10] = {{0}}, W[10][10] = {{0}};
        bool spd = ratio > 0.0;
        for (int a = 0; a < k && spd; ++a)
            for (int b = 0; b <= a; ++b) {
                long double v = M[idx[a]][idx[b]];
                for (int c = 0; c < b; ++c) v -= C[a][c] * C[b][c];
                if (a == b) {
                    if (!(v > 0)) {
                        spd = false;
                        break;
                    }
                    C[a][a] = sqrtl(v);
                } else {
                    C[a][b] = v / C[b][b];
                }
            }
        double *row = t.inv[m];
        if (spd) {
            for (int a = 0; a < k; ++a) {   // W = C^{-1}, lower
                W[a][a] = 1.0L / C[a][a];
                for (int b = 0; b < a; ++b) {
                    long double v = 0;
                    for (int c = b; c < a; ++c) v += C[a][c] * W[c][b];
                    W[a][b] = -v / C[a][a];
                }
            }
            for (int a = 0; a < k; ++a)     // inverse = W^T W
                for (int b = 0; b <= a; ++b) {
                    long double v = 0;
                    for (int c = a; c < k; ++c) v += W[c][a] * W[c][b];
                    const int i = idx[a], j = idx[b];
                    row[i * (i + 1) / 2 + j] = (double)v;
                }
        } else {
            for (int a = 0; a < k; ++a)
                for (int b = 0; b <= a; ++b) {
                    const int i = idx[a], j = idx[b];
                    row[i * (i + 1) / 2 + j] = NAN;
                }
        }
        row[MASK_PACK - 1] = ratio;
        if (spd) {
            // 1-norm condition number of the block, from the extended-precision inverse
            long double nm = 0, ni = 0;
            for (int a = 0; a < k; ++a) {
                long double cm = 0, ci = 0;
                for (int b2 = 0; b2 < k; ++b2) {
                    cm += fabsl((long double)M[idx[a]][idx[b2]]);
                    long double v = 0;
                    for (int c = (a > b2 ? a : b2); c < k; ++c) v += W[c][a] * W[c][b2];
                    ci += fabsl(v);
                }
                if (cm > nm) nm = cm;
                if (ci > ni) ni = ci;
            }
            if (nm * ni > (long double)MASK_FACTOR_COND) {
                // and the block's Cholesky factor: L_ij below the diagonal,
                // 1 / L_ii on it (solve_factor)
                t.factored[m >> 5] |= 1u << (m & 31);
                double *f = t.fac[m];
                for (int a = 0; a < k; ++a)
                    for (int b2 = 0; b2 <= a; ++b2) {
                        const int i = idx[a], j = idx[b2];
                        f[i * (i + 1) / 2 + j] = a == b2 ? (double)(1.0L / C[a][a]) : (double)C[a][b2];
                    }
                f[MASK_PACK - 1] = row[MASK_PACK - 1];
            }
        }
    }
}

}  // namespace mf
