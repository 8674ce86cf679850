// host_api_check.cpp -- TEST PROGRAM for the C++ host library (mf_host.hpp).
// Reads channels from a binary file written by tests/test_gpu_host_api.py,
// fits them through PulseChiSqSNNLS::DoFit (one by one, errors on) and
// doFitWrapper (batched), solves the same NNLS problems through
// inplace_fnnls / fnnls_wrapper, and runs the linalgebra functions on the
// reference's 3x3 literal (linalgebra/tests/test_cholesky_with_solvers.cpp:25-39);
// writes everything to an output file for the test to compare with the oracle.
//
// input : int32 n, then per channel: 10 samples, 100 samplecor (row-major),
//         pederr, 19 fullpulse, 361 fullpulsecov (row-major), 10 bx (as double)
// output: per channel: DoFit X[10], Errors[10], chi2; doFitWrapper X[10], chi2,
//         ampl; inplace_fnnls x[10] of (A = I + pulsemat, b = samples); then
//         the 3x3 L (9), y (3), x (3) in double and in float; then per
//         channel legacy::PulseChiSqSNNLS::DoFit X[10], Errors[10], chi2.
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "../../include/mf_host.hpp"
#include "../../include/mf_io.h"

// --wrapper in.bin out.bin: doFitWrapper alone on a large batch in a compact
// format.  input : int32 n, int32 ntypes; per type 100 samplecor, 19
// fullpulse, 361 fullpulsecov, npulses, 10 bx (all as double); per channel
// type id, pederr, 10 samples (as double).  output: per channel X[10] (0 past
// npulses), chi2, ampl, status.
static int wrapper_mode(const char *inp, const char *outp)
{
    FILE *f = fopen(inp, "rb");
    if (!f) return 2;
    int32_t n = 0, nt = 0;
    if (fread(&n, 4, 1, f) != 1 || fread(&nt, 4, 1, f) != 1) return 2;
    struct T {
        mf::SampleMatrix sc;
        mf::FullSampleVector fp;
        mf::FullSampleMatrix fpc;
        mf::BXVector bx;
    };
    std::vector<T> types((size_t)nt);
    for (auto &t : types) {
        std::vector<double> b(100 + 19 + 361 + 1 + 10);
        if (fread(b.data(), 8, b.size(), f) != b.size()) return 2;
        size_t k = 0;
        for (int i = 0; i < 10; ++i)
            for (int j = 0; j < 10; ++j) t.sc(i, j) = b[k++];
        for (int i = 0; i < 19; ++i) t.fp(i) = b[k++];
        for (int i = 0; i < 19; ++i)
            for (int j = 0; j < 19; ++j) t.fpc(i, j) = b[k++];
        const int np = (int)b[k++];
        t.bx = mf::BXVector(np);
        for (int i = 0; i < np; ++i) t.bx.coeffRef(i) = (char)b[k + (size_t)i];
    }
    std::vector<mf::DoFitArgs> args;
    args.reserve((size_t)n);
    for (int c = 0; c < n; ++c) {
        double b[12];
        if (fread(b, 8, 12, f) != 12) return 2;
        const T &t = types[(size_t)b[0]];
        mf::SampleVector s;
        for (int i = 0; i < 10; ++i) s(i) = b[2 + i];
        args.emplace_back(s, t.sc, b[1], t.bx, t.fp, t.fpc);
    }
    fclose(f);
    std::vector<double> out;
    try {
        auto res = mf::doFitWrapper(args);
        for (auto &r : res) {
            for (int p = 0; p < 10; ++p) out.push_back(p < r.X.rows() ? r.X[p] : 0.0);
            out.push_back(r.chi2);
            out.push_back(r.ampl);
            out.push_back(r.status);
        }
    } catch (const mf::Error &e) {
        fprintf(stderr, "mf::Error %d: %s\n", e.status, e.what());
        return 1;
    }
    FILE *o = fopen(outp, "wb");
    if (!o) return 2;
    fwrite(out.data(), 8, out.size(), o);
    fclose(o);
    printf("host_api_check --wrapper: %d channels, %d types\n", n, nt);
    return 0;
}

// --driver samples.root type.bin out.root: the CPU driver's loop
// (multifit_cpu/src/multifit.cpp:88-94, 196-227, 266-279) in C++ on the
// library: the Samples tree read with mf_io_*, each entry fitted by its own
// PulseChiSqSNNLS with errors off, and the RecoAndSim tree written with
// mf_io_write_tree: the input's samples and amplitudeTruth, chi2 as the
// driver leaves it (-99), samplesReco (X by BX + 5) and activeBXs.
// type.bin: 100 samplecor (row-major), 19 fullpulse, 361 fullpulsecov,
// pederr, npulses, 10 bx (all as double).
static int driver_mode(const char *inp, const char *typ, const char *outp)
{
    FILE *f = fopen(typ, "rb");
    if (!f) return 2;
    std::vector<double> b(100 + 19 + 361 + 1 + 1 + 10);
    if (fread(b.data(), 8, b.size(), f) != b.size()) return 2;
    fclose(f);
    size_t k = 0;
    mf::SampleMatrix sc;
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j) sc(i, j) = b[k++];
    mf::FullSampleVector fp;
    for (int i = 0; i < 19; ++i) fp(i) = b[k++];
    mf::FullSampleMatrix fpc;
    for (int i = 0; i < 19; ++i)
        for (int j = 0; j < 19; ++j) fpc(i, j) = b[k++];
    const double pederr = b[k++];
    const int np = (int)b[k++];
    mf::BXVector bx(np);
    for (int i = 0; i < np; ++i) bx.coeffRef(i) = (char)b[k + (size_t)i];

    mf_rootfile *rf = nullptr;
    if (mf_io_open(inp, &rf) != MF_OK) return 3;
    int64_t nv = 0, ne = 0;
    // first call counts, second reads
    mf_io_read_branch(rf, "Samples", "samples", MF_IO_VEC_F64, nullptr, 0, &nv, nullptr, 0, &ne);
    std::vector<double> samples((size_t)nv);
    std::vector<int64_t> counts((size_t)ne);
    if (mf_io_read_branch(rf, "Samples", "samples", MF_IO_VEC_F64, samples.data(), nv, &nv,
                          counts.data(), ne, &ne) != MF_OK)
        return 3;
    std::vector<double> truth((size_t)ne);
    int64_t nt = 0, net = 0;
    if (mf_io_read_branch(rf, "Samples", "amplitudeTruth", MF_IO_F64, truth.data(), ne, &nt,
                          nullptr, 0, &net) != MF_OK || net != ne)
        return 3;
    mf_io_close(rf);

    std::vector<double> chi2((size_t)ne, -99.0), reco((size_t)ne * 10, 0.0), bxs;
    std::vector<int64_t> reco_n((size_t)ne, 10), bx_n((size_t)ne, np);
    try {
        size_t off = 0;
        for (int64_t e = 0; e < ne; ++e) {
            if (counts[(size_t)e] != 10) return 4;
            mf::SampleVector s;
            for (int i = 0; i < 10; ++i) s(i) = samples[off + (size_t)i];
            off += 10;
            mf::PulseChiSqSNNLS pulse;
            pulse.disableErrorCalculation();
            pulse.DoFit(s, sc, pederr, bx, fp, fpc);
            for (int p = 0; p < np; ++p) reco[(size_t)e * 10 + (size_t)(bx.coeff(p) + 5)] = pulse.X()[p];
            for (int p = 0; p < np; ++p) bxs.push_back(bx.coeff(p));
        }
    } catch (const mf::Error &err) {
        fprintf(stderr, "mf::Error %d: %s\n", err.status, err.what());
        return 1;
    }
    const mf_io_column cols[] = {
        {"samples", MF_IO_VEC_F64, 0, samples.data(), counts.data()},
        {"amplitudeTruth", MF_IO_F64, 0, truth.data(), nullptr},
        {"chi2", MF_IO_F32, 0, chi2.data(), nullptr},
        {"samplesReco", MF_IO_VEC_F64, 0, reco.data(), reco_n.data()},
        {"activeBXs", MF_IO_VEC_I32, 0, bxs.data(), bx_n.data()},
    };
    return mf_io_write_tree(outp, "RecoAndSim", "", cols, 5, ne) == MF_OK ? 0 : 5;
}

int main(int argc, char **argv)
{
    if (argc == 4 && std::string(argv[1]) == "--wrapper") return wrapper_mode(argv[2], argv[3]);
    if (argc == 5 && std::string(argv[1]) == "--driver")
        return driver_mode(argv[2], argv[3], argv[4]);
    if (argc != 3) {
        fprintf(stderr, "usage: %s [--wrapper] in.bin out.bin | --driver samples.root type.bin out.root\n",
                argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t n = 0;
    if (fread(&n, 4, 1, f) != 1) return 2;
    std::vector<mf::DoFitArgs> args;
    std::vector<mf::NNLS_args> nn;
    for (int c = 0; c < n; ++c) {
        std::vector<double> buf(10 + 100 + 1 + 19 + 361 + 10);
        if (fread(buf.data(), 8, buf.size(), f) != buf.size()) return 2;
        size_t k = 0;
        mf::SampleVector s;
        for (int i = 0; i < 10; ++i) s(i) = buf[k++];
        mf::SampleMatrix sc;
        for (int i = 0; i < 10; ++i)
            for (int j = 0; j < 10; ++j) sc(i, j) = buf[k++];
        const double pederr = buf[k++];
        mf::FullSampleVector fp;
        for (int i = 0; i < 19; ++i) fp(i) = buf[k++];
        mf::FullSampleMatrix fpc;
        for (int i = 0; i < 19; ++i)
            for (int j = 0; j < 19; ++j) fpc(i, j) = buf[k++];
        mf::BXVector bx(10);
        for (int i = 0; i < 10; ++i) bx.coeffRef(i) = (char)buf[k++];
        args.emplace_back(s, sc, pederr, bx, fp, fpc);
    }
    fclose(f);
    FILE *o = fopen(argv[2], "wb");
    if (!o) return 2;
    try {
        std::vector<double> out;
        for (auto &a : args) {
            mf::PulseChiSqSNNLS fit;
            fit.DoFit(a.samples, a.samplecor, a.pederr, a.bxs, a.fullpulse, a.fullpulsecov);
            for (int p = 0; p < 10; ++p) out.push_back(fit.X()[p]);
            for (int p = 0; p < 10; ++p) out.push_back(fit.Errors()[p]);
            out.push_back(fit.ChiSq());
            mf::FixedMatrix A;
            for (int i = 0; i < 10; ++i)
                for (int j = 0; j < 10; ++j) A(i, j) = fit.pulsemat()(i, j) + (i == j ? 1.0 : 0.0);
            mf::FixedVector b;
            for (int i = 0; i < 10; ++i) b(i) = a.samples(i);
            nn.emplace_back(A, b);
        }
        auto res = mf::doFitWrapper(args);
        for (auto &r : res) {
            for (int p = 0; p < 10; ++p) out.push_back(r.X[p]);
            out.push_back(r.chi2);
            out.push_back(r.ampl);
        }
        auto xs = mf::inplace_fnnls_wrapper(nn);
        for (size_t c = 0; c < xs.size(); ++c) {
            mf::FixedVector x1;  // single-problem entry, cold start
            mf::inplace_fnnls(nn[c].A, nn[c].b, x1);
            for (int i = 0; i < 10; ++i) out.push_back(xs[c](i));
            for (int i = 0; i < 10; ++i) out.push_back(x1(i));
        }
        double M[9] = {4, -1, 2, -1, 6, 0, 2, 0, 5}, L[9] = {0}, bb[3] = {3, 3, 4}, y[3], x[3];
        mf::cholesky_decomp<double>(M, L, 3, 3);
        mf::solve_forward_substitution<double>(L, bb, y, 3, 3);
        mf::solve_backward_substitution<double>(L, y, x, 3, 3);
        for (double v : L) out.push_back(v);
        for (double v : y) out.push_back(v);
        for (double v : x) out.push_back(v);
        float Mf[9] = {4, -1, 2, -1, 6, 0, 2, 0, 5}, Lf[9] = {0}, bf[3] = {3, 3, 4}, yf[3], xf[3];
        mf::cholesky_decomp<float>(Mf, Lf, 3, 3);
        mf::solve_forward_substitution<float>(Lf, bf, yf, 3, 3);
        mf::solve_backward_substitution<float>(Lf, yf, xf, 3, 3);
        for (float v : Lf) out.push_back(v);
        for (float v : yf) out.push_back(v);
        for (float v : xf) out.push_back(v);
        for (auto &a : args) {
            mf::legacy::PulseChiSqSNNLS fit;
            fit.DoFit(a.samples, a.samplecor, a.pederr, a.bxs, a.fullpulse, a.fullpulsecov);
            for (int p = 0; p < 10; ++p) out.push_back(fit.X()[p]);
            for (int p = 0; p < 10; ++p) out.push_back(fit.Errors()[p]);
            out.push_back(fit.ChiSq());
        }
        fwrite(out.data(), 8, out.size(), o);
    } catch (const mf::Error &e) {
        fprintf(stderr, "mf::Error %d: %s\n", e.status, e.what());
        return 1;
    }
    fclose(o);
    printf("host_api_check: %d channels\n", n);
    return 0;
}
