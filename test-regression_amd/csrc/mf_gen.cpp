// mf_gen.cpp -- synthetic pulse workloads (host code).  See include/mf_gen.h.
#include "../../include/mf_gen.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <thread>
#include <vector>

namespace {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies, Weyl key steps.
struct Philox {
    uint32_t k0, k1;
    static inline void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo)
    {
        const uint64_t p = (uint64_t)a * b;
        hi = (uint32_t)(p >> 32);
        lo = (uint32_t)p;
    }
    inline void block(uint32_t c[4]) const
    {
        uint32_t a0 = k0, a1 = k1;
        for (int r = 0; r < 10; ++r) {
            uint32_t hi0, lo0, hi1, lo1;
            mulhilo(0xD2511F53u, c[0], hi0, lo0);
            mulhilo(0xCD9E8D57u, c[2], hi1, lo1);
            const uint32_t n0 = hi1 ^ c[1] ^ a0, n2 = hi0 ^ c[3] ^ a1;
            c[0] = n0;
            c[1] = lo1;
            c[2] = n2;
            c[3] = lo0;
            a0 += 0x9E3779B9u;
            a1 += 0xBB67AE85u;
        }
    }
};

// Per-channel stream: counter (channel lo, channel hi, block, stream id).
struct Stream {
    Philox ph;
    uint64_t ch;
    uint32_t sid, blk = 0;
    uint32_t buf[4];
    int used = 4;
    Stream(uint64_t seed, uint64_t c, uint32_t s) : ch(c), sid(s)
    {
        ph.k0 = (uint32_t)seed;
        ph.k1 = (uint32_t)(seed >> 32);
    }
    uint32_t next()
    {
        if (used == 4) {
            buf[0] = (uint32_t)ch;
            buf[1] = (uint32_t)(ch >> 32);
            buf[2] = blk++;
            buf[3] = sid;
            ph.block(buf);
            used = 0;
        }
        return buf[used++];
    }
    // uniform in (0, 1)
    double uniform() { return ((double)next() + 0.5) * (1.0 / 4294967296.0); }
    double normal()
    {
        const double u1 = uniform(), u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
    double expo(double mean) { return -mean * log(uniform()); }
    int poisson(double mu)
    {
        if (mu <= 0.0) return 0;
        const double u = uniform();
        double p = exp(-mu), F = p;
        int k = 0;
        while (u > F && k < 1000) {
            ++k;
            p *= mu / k;
            F += p;
        }
        return k;
    }
};

inline double tmpl(double dt, double alpha, double beta)
{
    const double ab = alpha * beta;
    if (dt <= -ab) return 0.0;
    return pow(1.0 + dt / ab, alpha) * exp(-dt / beta);
}

inline double tmpl_deriv(double dt, double alpha, double beta)
{
    const double ab = alpha * beta;
    if (dt <= -ab) return 0.0;
    return tmpl(dt, alpha, beta) * (alpha / (ab + dt) - 1.0 / beta);
}

bool spec_ok(const mf_gen_spec &s)
{
    return (s.nsamples == 10 || s.nsamples == 8) && s.peak_sample >= 2 &&
           s.peak_sample < s.nsamples && s.alpha > 0 && s.beta_ns > 0 && s.sigma >= 0 &&
           s.rho > -1 && s.rho < 1;
}

}  // namespace

extern "C" mf_status mf_gen_type(const mf_gen_spec *spec, mf_chtype *out)
{
    if (!spec || !out || !spec_ok(*spec)) return MF_ERR_INVALID;
    const mf_gen_spec &s = *spec;
    memset(out, 0, sizeof(*out));
    const int n = s.nsamples;
    out->nsamples = n;
    out->npulses = n;
    for (int p = 0; p < n; ++p) out->bx[p] = p - s.peak_sample;
    out->first_off = s.peak_sample - 2;
    out->pulse_off = 9 - s.peak_sample;
    out->pederr = s.sigma;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) out->noisecor[i * 10 + j] = pow(s.rho, abs(i - j));
    for (int m = 0; m < MF_FULL; ++m) out->fullpulse[m] = tmpl(25.0 * (m - 9), s.alpha, s.beta_ns);
    if (s.sigma_t_ns > 0) {
        double d[MF_FULL];
        for (int m = 0; m < MF_FULL; ++m) d[m] = tmpl_deriv(25.0 * (m - 9), s.alpha, s.beta_ns);
        const double st2 = s.sigma_t_ns * s.sigma_t_ns;
        for (int m = 0; m < MF_FULL; ++m)
            for (int k = 0; k < MF_FULL; ++k) out->fullpulsecov[m * MF_FULL + k] = st2 * d[m] * d[k];
    }
    return MF_OK;
}

extern "C" mf_status mf_gen_samples(const mf_gen_spec *specs, int32_t nspecs, const uint8_t *type_of,
                                    int64_t nch, uint64_t seed, double *samples, double *truth,
                                    int32_t nthreads)
{
    if (!specs || nspecs <= 0 || nspecs > 255 || nch < 0 || !samples) return MF_ERR_INVALID;
    for (int t = 0; t < nspecs; ++t)
        if (!spec_ok(specs[t])) return MF_ERR_INVALID;
    // noise Cholesky per spec
    std::vector<std::array<double, 100>> Ln(nspecs);
    for (int t = 0; t < nspecs; ++t) {
        const int n = specs[t].nsamples;
        double C[100] = {0}, L[100] = {0};
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) C[i * 10 + j] = pow(specs[t].rho, abs(i - j));
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j <= i; ++j) {
                double s = C[i * 10 + j];
                for (int k = 0; k < j; ++k) s -= L[i * 10 + k] * L[j * 10 + k];
                L[i * 10 + j] = (i == j) ? sqrt(s) : s / L[j * 10 + j];
            }
        }
        for (int k = 0; k < 100; ++k) Ln[t][k] = L[k];
    }
    std::atomic<int64_t> next(0);
    auto work = [&]() {
        const int64_t chunk = 4096;
        for (;;) {
            const int64_t c0 = next.fetch_add(chunk);
            if (c0 >= nch) break;
            const int64_t c1 = std::min(nch, c0 + chunk);
            for (int64_t c = c0; c < c1; ++c) {
                const int t = type_of ? type_of[c] : 0;
                const mf_gen_spec &s = specs[t < nspecs ? t : 0];
                const int n = s.nsamples;
                Stream rs(seed, (uint64_t)c, 0u);
                double wave[10] = {0}, tru[10] = {0};
                // in-time pulse
                const double a0 = s.a0_lo + (s.a0_hi - s.a0_lo) * rs.uniform();
                const double j0 = s.jitter_ns > 0 ? s.jitter_ns * rs.normal() : 0.0;
                for (int i = 0; i < n; ++i)
                    wave[i] += a0 * tmpl(25.0 * (i - s.peak_sample) - j0, s.alpha, s.beta_ns);
                tru[s.peak_sample] += a0;
                // pileup: per BX slot, Poisson(mean / n) pulses
                const double mu = s.pu_lo + (s.pu_hi - s.pu_lo) * rs.uniform();
                if (mu > 0) {
                    for (int p = 0; p < n; ++p) {
                        const int k = p - s.peak_sample;
                        const int np = rs.poisson(mu / n);
                        for (int q = 0; q < np; ++q) {
                            const double a = rs.expo(s.pu_amp_mean);
                            const double tj = s.jitter_ns > 0 ? s.jitter_ns * rs.normal() : 0.0;
                            for (int i = 0; i < n; ++i)
                                wave[i] += a * tmpl(25.0 * (i - s.peak_sample - k) - tj, s.alpha,
                                                    s.beta_ns);
                            tru[p] += a;
                        }
                    }
                }
                // correlated noise
                double z[10];
                for (int i = 0; i < n; ++i) z[i] = rs.normal();
                for (int i = 0; i < n; ++i) {
                    double e = 0;
                    for (int k = 0; k <= i; ++k) e += Ln[t][i * 10 + k] * z[k];
                    wave[i] += s.sigma * e;
                }
                for (int i = 0; i < 10; ++i) samples[(int64_t)i * nch + c] = i < n ? wave[i] : 0.0;
                if (truth)
                    for (int i = 0; i < 10; ++i) truth[(int64_t)i * nch + c] = tru[i];
            }
        }
    };
    int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > 64) nt = 64;
    if (nt == 1 || nch < 8192) {
        work();
    } else {
        std::vector<std::thread> th;
        for (int i = 0; i < nt; ++i) th.emplace_back(work);
        for (auto &x : th) x.join();
    }
    return MF_OK;
}

extern "C" mf_status mf_gen_mix(int64_t nch, double frac1, uint64_t seed, uint8_t *type_of)
{
    if (nch < 0 || !type_of) return MF_ERR_INVALID;
    for (int64_t c = 0; c < nch; ++c) {
        Stream rs(seed, (uint64_t)c, 0x4D495855u);
        type_of[c] = rs.uniform() < frac1 ? 1 : 0;
    }
    return MF_OK;
}
