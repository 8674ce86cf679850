/*
 * mf_gen.h -- synthetic pulse workloads for the multifit engine (host code in
 * libmf_amd.so).  Stands in for the reference's ROOT generator (gen/src/gen_data.cpp),
 * which needs ROOT and its data files; it produces the channel populations
 * of the benchmark configurations from closed-form pulse shapes instead.
 *
 * Pulse shape (alpha-beta template): A(dt) = (1 + dt/(alpha*beta))^alpha *
 * exp(-dt/beta) for dt > -alpha*beta, else 0; A(0) = 1 is the peak.
 * Sample i of a pulse of amplitude a in bunch crossing k with time jitter tau:
 * a * A(25*i - 25*(peak + k) - tau) (ns).  Noise: sigma * L z with L L^T =
 * rho^|i-j| and z ~ N(0,1).  Random numbers: Philox4x32-10 (Salmon et al.,
 * SC'11) keyed by the seed and counted by (channel, draw), so output does not
 * depend on the thread count.
 */
#ifndef MF_GEN_H
#define MF_GEN_H

#include <stdint.h>

#include "mf_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mf_gen_spec {
    int32_t nsamples;      /* 10 (ECAL-like) or 8 (HCAL-like); npulses = nsamples */
    int32_t peak_sample;   /* sample of the in-time peak (ECAL 5, HCAL 4)          */
    double  alpha;         /* template alpha                                        */
    double  beta_ns;       /* template beta in ns                                   */
    double  sigma;         /* noise sigma (ADC) = pederr                            */
    double  rho;           /* noise correlation rho^|i-j|                           */
    double  a0_lo, a0_hi;  /* in-time amplitude ~ U[a0_lo, a0_hi]                   */
    double  pu_lo, pu_hi;  /* per-channel pileup mean ~ U[pu_lo, pu_hi]; per BX the
                              pulse count is Poisson(mean / npulses)                 */
    double  pu_amp_mean;   /* pileup pulse amplitude ~ Exponential(mean)            */
    double  jitter_ns;     /* per-pulse time jitter ~ N(0, jitter_ns)               */
    double  sigma_t_ns;    /* > 0: fullpulsecov = sigma_t^2 A'(t) A'(t)^T (live);
                              0: fullpulsecov = 0 (as every reference driver)        */
} mf_gen_spec;

/* Fills the channel type implied by a spec: BX list -peak..nsamples-1-peak,
 * first_off = peak - 2, pulse_off = 9 - peak, fullpulse[m] = A(25 (m - 9)),
 * noisecor = rho^|i-j|, pederr = sigma, fullpulsecov as above. */
mf_status mf_gen_type(const mf_gen_spec *spec, mf_chtype *out);

/* Generates nch channels.  type_of[c] (host, may be NULL = all 0) selects the
 * spec.  samples: host, SoA samples[i*nch + c], 10 rows (rows >= nsamples of
 * a channel are 0).  truth (host, may be NULL): SoA [10][nch] summed
 * amplitude per BX slot (pulse order).  nthreads <= 0: all hardware threads. */
mf_status mf_gen_samples(const mf_gen_spec *specs, int32_t nspecs, const uint8_t *type_of,
                         int64_t nch, uint64_t seed, double *samples, double *truth,
                         int32_t nthreads);

/* type_of[c] = 1 with probability frac1, else 0 (Philox stream of its own). */
mf_status mf_gen_mix(int64_t nch, double frac1, uint64_t seed, uint8_t *type_of);

#ifdef __cplusplus
}
#endif
#endif
