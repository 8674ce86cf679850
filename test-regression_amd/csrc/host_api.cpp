// host_api.cpp -- the C++ host library (include/mf_host.hpp) over the C ABI.
// Every numerical call goes to the gfx950 kernels; this file only marshals
// the reference's Eigen-shaped arguments into device batches and back.
#include <hip/hip_runtime.h>
#include <string.h>

#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/mf_host.hpp"

namespace mf {

static void throw_status(mf_status st, const char *what)
{
    if (st != MF_OK) throw Error((int)st, std::string(what) + ": " + mf_last_error());
}

static void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw Error(MF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffers of one call, carved from the device's staging area (which
// grows to the largest call and is kept): parts are 256-byte aligned.
struct Staging {
    std::vector<size_t> off;
    size_t total = 0;
    size_t add(size_t n)
    {
        off.push_back(total);
        total += (n + 255) / 256 * 256;
        return off.size() - 1;
    }
};

Device::Device(int device) : device_(device)
{
    throw_status(mf_context_create(&ctx_, device), "mf_context_create");
    hip_check(hipSetDevice(device), "hipSetDevice");
    hipStream_t s;
    hip_check(hipStreamCreate(&s), "hipStreamCreate");
    stream_ = s;
}

Device::~Device()
{
    if (stage_) (void)hipFree(stage_);
    if (stream_) (void)hipStreamDestroy(static_cast<hipStream_t>(stream_));
    if (ctx_) mf_context_destroy(ctx_);
}

void *Device::staging(size_t bytes)
{
    if (bytes > stage_cap_) {
        hip_check(hipStreamSynchronize(static_cast<hipStream_t>(stream_)), "hipStreamSynchronize");
        if (stage_) hip_check(hipFree(stage_), "hipFree");
        stage_ = nullptr;
        stage_cap_ = 0;
        hip_check(hipMalloc(&stage_, bytes), "hipMalloc");
        stage_cap_ = bytes;
    }
    return stage_;
}

Device &Device::get(int device)
{
    static std::mutex mu;
    static std::map<int, std::unique_ptr<Device>> devs;
    std::lock_guard<std::mutex> lk(mu);
    auto it = devs.find(device);
    if (it == devs.end()) it = devs.emplace(device, std::unique_ptr<Device>(new Device(device))).first;
    return *it->second;
}

void Device::set_types(const std::vector<mf_chtype> &types)
{
    if (types.size() == cached_.size() &&
        memcmp(types.data(), cached_.data(), sizeof(mf_chtype) * types.size()) == 0)
        return;
    throw_status(mf_set_types(ctx_, types.data(), (int32_t)types.size()), "mf_set_types");
    cached_ = types;
}

mf_chtype make_chtype(const SampleMatrix &samplecor, double pederr, const BXVector &bxs,
                      const FullSampleVector &fullpulse, const FullSampleMatrix &fullpulsecov)
{
    if (bxs.rows() < 1 || bxs.rows() > 10)
        throw Error(MF_ERR_INVALID, "bxs must hold 1..10 pulses");
    mf_chtype t;
    memset(&t, 0, sizeof(t));
    t.nsamples = 10;
    t.npulses = bxs.rows();
    for (int p = 0; p < bxs.rows(); ++p) t.bx[p] = (int32_t)bxs.coeff(p);
    t.first_off = 3;  // firstsamplet = max(0, bx + 3)   PulseChiSqSNNLS.cpp:54
    t.pulse_off = 4;  // offset = 7 - 3 - bx             PulseChiSqSNNLS.cpp:55
    t.pederr = pederr;
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j) t.noisecor[i * 10 + j] = samplecor(i, j);
    for (int i = 0; i < MF_FULL; ++i) t.fullpulse[i] = fullpulse(i);
    for (int i = 0; i < MF_FULL; ++i)
        for (int j = 0; j < MF_FULL; ++j) t.fullpulsecov[i * MF_FULL + j] = fullpulsecov(i, j);
    return t;
}

// ---------------------------------------------------------------- fits
namespace {
struct Fitted {
    std::vector<double> x, err, chi2;
    std::vector<int32_t> flags;
};

Fitted run_fits(const std::vector<mf_chtype> &types, const std::vector<uint8_t> &type_of,
                const std::vector<double> &samples_soa, const std::vector<double> &pederr,
                int64_t nch, bool errors, int solver = MF_SOLVER_INPLACE_FNNLS)
{
    Device &dev = Device::get();
    dev.set_types(types);
    hipStream_t st = static_cast<hipStream_t>(dev.stream());
    const size_t n = (size_t)nch;
    Staging sg;
    const size_t ks = sg.add(sizeof(double) * 10 * n), kt = sg.add(type_of.size()),
                 kp = sg.add(sizeof(double) * pederr.size()), kx = sg.add(sizeof(double) * 10 * n),
                 kc = sg.add(sizeof(double) * n), kf = sg.add(sizeof(int32_t) * n),
                 ke = sg.add(errors ? sizeof(double) * 10 * n : 0);
    char *base = static_cast<char *>(dev.staging(sg.total));
    auto at = [&](size_t k) { return static_cast<void *>(base + sg.off[k]); };
    hip_check(hipMemcpyAsync(at(ks), samples_soa.data(), sizeof(double) * 10 * n,
                             hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    if (!type_of.empty())
        hip_check(hipMemcpyAsync(at(kt), type_of.data(), type_of.size(), hipMemcpyHostToDevice, st),
                  "hipMemcpyAsync");
    if (!pederr.empty())
        hip_check(hipMemcpyAsync(at(kp), pederr.data(), sizeof(double) * pederr.size(),
                                 hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    mf_batch in;
    memset(&in, 0, sizeof(in));
    in.nch = nch;
    in.sample_dtype = MF_DTYPE_F64;
    in.samples = at(ks);
    in.type_of = type_of.empty() ? nullptr : static_cast<const uint8_t *>(at(kt));
    in.pederr = pederr.empty() ? nullptr : static_cast<const double *>(at(kp));
    mf_result out;
    memset(&out, 0, sizeof(out));
    out.x = at(kx);
    out.chi2 = at(kc);
    out.flags = static_cast<int32_t *>(at(kf));
    out.err = errors ? at(ke) : nullptr;
    mf_opts o;
    mf_default_opts(&o);
    o.compute_errors = errors ? 1 : 0;
    o.solver = solver;
    throw_status(mf_fit_batch(dev.ctx(), &in, &out, &o, st), "mf_fit_batch");
    Fitted f;
    f.x.resize(10 * n);
    f.chi2.resize(n);
    f.flags.resize(n);
    hip_check(hipMemcpyAsync(f.x.data(), out.x, f.x.size() * sizeof(double), hipMemcpyDeviceToHost,
                             st), "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(f.chi2.data(), out.chi2, f.chi2.size() * sizeof(double),
                             hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(f.flags.data(), out.flags, f.flags.size() * sizeof(int32_t),
                             hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (errors) {
        f.err.resize(10 * n);
        hip_check(hipMemcpyAsync(f.err.data(), out.err, f.err.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    }
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    return f;
}
}  // namespace

PulseChiSqSNNLS::PulseChiSqSNNLS() {}
PulseChiSqSNNLS::PulseChiSqSNNLS(int solver) : _solver(solver) {}
PulseChiSqSNNLS::~PulseChiSqSNNLS() {}

bool PulseChiSqSNNLS::DoFit(const SampleVector &samples, const SampleMatrix &samplecor,
                            double pederr, const BXVector &bxs, const FullSampleVector &fullpulse,
                            const FullSampleMatrix &fullpulsecov)
{
    const int np = bxs.rows();
    // the channel's constants with pederr 1 (the type table stays cached from
    // one call to the next while they repeat) and its pederr per channel
    std::vector<mf_chtype> types{make_chtype(samplecor, 1.0, bxs, fullpulse, fullpulsecov)};
    std::vector<double> s(10);
    for (int i = 0; i < 10; ++i) s[(size_t)i] = samples(i);
    Fitted f = run_fits(types, {}, s, {pederr}, 1, _computeErrors, _solver);
    // pulse matrix as built by DoFit (PulseChiSqSNNLS.cpp:50-62)
    _pulsemat = SamplePulseMatrix::Zero();
    for (int p = 0; p < np; ++p) {
        const int bx = bxs.coeff(p), f0 = bx + 3 > 0 ? bx + 3 : 0, o = 4 - bx;
        for (int i = f0; i < 10; ++i) _pulsemat(i, p) = fullpulse(i + o);
    }
    _ampvecmin = PulseVector(np);
    _errvec = PulseVector(np);
    for (int p = 0; p < np; ++p) {
        _ampvecmin[p] = f.x[(size_t)p];
        _errvec[p] = _computeErrors ? f.err[(size_t)p] : 0.0;
    }
    _bxsmin = bxs;
    _chisq = f.chi2[0];
    _flags = f.flags[0];
    return true;  // Minimize always reports success (PulseChiSqSNNLS.cpp:165)
}

std::vector<Output> doFitWrapper(std::vector<DoFitArgs> const &vargs)
{
    const int64_t n = (int64_t)vargs.size();
    std::vector<Output> res((size_t)n);
    if (!n) return res;
    // channels sharing their constants (all but pederr, which goes per
    // channel) share a type
    std::vector<mf_chtype> types;
    std::vector<uint8_t> type_of((size_t)n);
    std::vector<double> ped((size_t)n);
    for (int64_t c = 0; c < n; ++c) {
        const DoFitArgs &a = vargs[(size_t)c];
        ped[(size_t)c] = a.pederr;
        mf_chtype t = make_chtype(a.samplecor, 1.0, a.bxs, a.fullpulse, a.fullpulsecov);
        size_t k = 0;
        for (; k < types.size(); ++k)
            if (!memcmp(&types[k], &t, sizeof(t))) break;
        if (k == types.size()) {
            if (types.size() == 256)
                throw Error(MF_ERR_INVALID, "doFitWrapper: more than 256 distinct channel types");
            types.push_back(t);
        }
        type_of[(size_t)c] = (uint8_t)k;
    }
    std::vector<double> s(10 * (size_t)n);
    for (int64_t c = 0; c < n; ++c)
        for (int i = 0; i < 10; ++i) s[(size_t)(i * n + c)] = vargs[(size_t)c].samples(i);
    Fitted f = run_fits(types, types.size() > 1 ? type_of : std::vector<uint8_t>{}, s, ped, n,
                        false);
    for (int64_t c = 0; c < n; ++c) {
        Output &o = res[(size_t)c];
        const DoFitArgs &a = vargs[(size_t)c];
        o.chi2 = f.chi2[(size_t)c];
        o.status = 1;
        o.BXs = a.bxs;
        const int np = a.bxs.rows();
        o.X = PulseVector(np);
        int it = -1;
        for (int p = 0; p < np; ++p) {
            o.X[p] = f.x[(size_t)(p * n + c)];
            if (it < 0 && a.bxs.coeff(p) == 0) it = p;
        }
        o.ampl = it >= 0 ? o.X[it] : o.X[0];
    }
    return res;
}

// ---------------------------------------------------------------- solvers
static std::vector<FixedVector> nnls_many(std::vector<NNLS_args> const &args, double eps,
                                          unsigned max_iter, bool inplace,
                                          const std::vector<FixedVector> *x0)
{
    const size_t n = args.size();
    std::vector<FixedVector> x(n);
    if (!n) return x;
    Device &dev = Device::get();
    hipStream_t st = static_cast<hipStream_t>(dev.stream());
    std::vector<double> A(100 * n), b(10 * n), xx(10 * n, 0.0);
    for (size_t c = 0; c < n; ++c) {
        memcpy(&A[100 * c], args[c].A.data(), sizeof(double) * 100);  // column-major
        memcpy(&b[10 * c], args[c].b.data(), sizeof(double) * 10);
        if (x0) memcpy(&xx[10 * c], (*x0)[c].data(), sizeof(double) * 10);
    }
    Staging sg;
    const size_t kA = sg.add(sizeof(double) * A.size()), kb = sg.add(sizeof(double) * b.size()),
                 kx = sg.add(sizeof(double) * xx.size());
    char *base = static_cast<char *>(dev.staging(sg.total));
    double *dA = reinterpret_cast<double *>(base + sg.off[kA]),
           *db = reinterpret_cast<double *>(base + sg.off[kb]),
           *dx = reinterpret_cast<double *>(base + sg.off[kx]);
    hip_check(hipMemcpyAsync(dA, A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice, st),
              "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(db, b.data(), sizeof(double) * b.size(), hipMemcpyHostToDevice, st),
              "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(dx, xx.data(), sizeof(double) * xx.size(), hipMemcpyHostToDevice,
                             st), "hipMemcpyAsync");
    auto f = inplace ? mf_inplace_fnnls_batch : mf_fnnls_batch;
    throw_status(f(dA, db, dx, (int64_t)n, 10, eps, max_iter, st),
                 inplace ? "mf_inplace_fnnls_batch" : "mf_fnnls_batch");
    hip_check(hipMemcpyAsync(xx.data(), dx, sizeof(double) * xx.size(), hipMemcpyDeviceToHost,
                             st), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (size_t c = 0; c < n; ++c) memcpy(x[c].data(), &xx[10 * c], sizeof(double) * 10);
    return x;
}

void inplace_fnnls(const FixedMatrix &A, const FixedVector &b, FixedVector &x, const double eps,
                   const unsigned int max_iterations)
{
    std::vector<FixedVector> x0{x};
    x = nnls_many({NNLS_args(A, b)}, eps, max_iterations, true, &x0)[0];
}

void fnnls(const FixedMatrix &A, const FixedVector &b, FixedVector &x, const double eps,
           const unsigned int max_iterations)
{
    std::vector<FixedVector> x0{x};
    x = nnls_many({NNLS_args(A, b)}, eps, max_iterations, false, &x0)[0];
}

std::vector<FixedVector> fnnls_wrapper(std::vector<NNLS_args> const &args, const double eps,
                                       const unsigned int max_iterations)
{
    return nnls_many(args, eps, max_iterations, false, nullptr);
}

std::vector<FixedVector> inplace_fnnls_wrapper(std::vector<NNLS_args> const &args,
                                               const double eps, const unsigned int max_iterations)
{
    return nnls_many(args, eps, max_iterations, true, nullptr);
}

// ---------------------------------------------------------------- linalgebra
template <typename T> static int dtype_of();
template <> int dtype_of<double>() { return MF_DTYPE_F64; }
template <> int dtype_of<float>() { return MF_DTYPE_F32; }

// one problem: M, L full*full, b, y full; nullptr operands are not copied
template <typename T>
static void la_call(int op, T *M, T *L, T *b, T *y, int full, int view, int i, int j)
{
    if (full < 1 || full > 10) throw Error(MF_ERR_INVALID, "linalgebra: full_size must be 1..10");
    Device &dev = Device::get();
    hipStream_t st = static_cast<hipStream_t>(dev.stream());
    const size_t mb = sizeof(T) * (size_t)(full * full), vb = sizeof(T) * (size_t)full;
    Staging sg;
    const size_t kM = sg.add(mb), kL = sg.add(mb), kb = sg.add(vb), ky = sg.add(vb);
    char *base = static_cast<char *>(dev.staging(sg.total));
    void *dM = M ? base + sg.off[kM] : nullptr, *dL = L ? base + sg.off[kL] : nullptr,
         *db = b ? base + sg.off[kb] : nullptr, *dy = y ? base + sg.off[ky] : nullptr;
    if (M) hip_check(hipMemcpyAsync(dM, M, mb, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    if (L) hip_check(hipMemcpyAsync(dL, L, mb, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    if (b) hip_check(hipMemcpyAsync(db, b, vb, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    if (y) hip_check(hipMemcpyAsync(dy, y, vb, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    throw_status(mf_linalg_batch(op, dtype_of<T>(), dM, dL, db, dy, 1, full, view, i, j, st),
                 "mf_linalg_batch");
    if (M) hip_check(hipMemcpyAsync(M, dM, mb, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (L) hip_check(hipMemcpyAsync(L, dL, mb, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (y) hip_check(hipMemcpyAsync(y, dy, vb, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

template <typename T> void cholesky_decomp(T *pM, T *pL, int full_size, int view_size)
{
    la_call<T>(MF_LA_CHOLESKY, pM, pL, nullptr, nullptr, full_size, view_size, 0, 0);
}
template <typename T>
void solve_forward_substitution(T *pM, T *pb, T *psolution, int full_size, int view_size)
{
    la_call<T>(MF_LA_FORWARD, nullptr, pM, pb, psolution, full_size, view_size, 0, 0);
}
template <typename T>
void solve_backward_substitution(T *pM, T *pb, T *psolution, int full_size, int view_size)
{
    la_call<T>(MF_LA_BACKWARD, nullptr, pM, pb, psolution, full_size, view_size, 0, 0);
}
template <typename T> void swap_row_column(T *pM, int i, int j, int full_size, int view_size)
{
    la_call<T>(MF_LA_SWAP, pM, nullptr, nullptr, nullptr, full_size, view_size, i, j);
}
template <typename T>
void fused_cholesky_forward_substitution_solver_rcaddition(T *pM, T *pL, T *pb, T *py,
                                                           int full_size, int view_size)
{
    la_call<T>(MF_LA_RCADDITION, pM, pL, pb, py, full_size, view_size, 0, 0);
}
template <typename T>
void fused_cholesky_forward_substitution_solver_inbetween_removal(T *pM, T *pL, T *pb, T *py,
                                                                  int position, int full_size,
                                                                  int view_size)
{
    la_call<T>(MF_LA_INBETWEEN, pM, pL, pb, py, full_size, view_size, position, 0);
}

#define MF_INSTANTIATE(T)                                                                   \
    template void cholesky_decomp<T>(T *, T *, int, int);                                   \
    template void solve_forward_substitution<T>(T *, T *, T *, int, int);                   \
    template void solve_backward_substitution<T>(T *, T *, T *, int, int);                  \
    template void swap_row_column<T>(T *, int, int, int, int);                              \
    template void fused_cholesky_forward_substitution_solver_rcaddition<T>(T *, T *, T *, T *, \
                                                                           int, int);      \
    template void fused_cholesky_forward_substitution_solver_inbetween_removal<T>(          \
        T *, T *, T *, T *, int, int, int);
MF_INSTANTIATE(double)
MF_INSTANTIATE(float)
#undef MF_INSTANTIATE

}  // namespace mf
