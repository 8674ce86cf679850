// mf_host.hpp -- C++ host library over the C ABI (include/mf_abi.h): the
// reference's C++ entry points for the multifit path, Eigen-free, running on
// the gfx950 kernels of libmf_amd.so.  Paths are relative to the reference root.
//
//   PulseChiSqSNNLS ........ multifit_cpu/interface/PulseChiSqSNNLS.h:9-59
//   DoFitArgs, Output ...... multifit_gpu/interface/DeviceData.h:6-35
//   doFitWrapper ........... multifit_gpu/interface/PulseChiSqSNNLSWrapper.h:11
//   inplace_fnnls, fnnls ... nnls/interface/inplace_fnnls.h:9-14, fnnls.h:9-14
//   NNLS_args, fnnls_wrapper nnls/interface/data_types.h:98-106, kernel_wrapper.h:21-24
//   cholesky_decomp, solve_{forward,backward}_substitution, swap_row_column,
//   fused_cholesky_forward_substitution_solver_{rcaddition,inbetween_removal}
//                            linalgebra/interface/*.hpp
//
// Matrix types keep the reference's storage: Eigen's default column-major
// order for SampleMatrix / FixedMatrix (element (r, c) at data()[c*rows + r]),
// row-major raw pointers for the linalgebra functions.  Every call runs on the
// GPU; errors are thrown as mf::Error (the C ABI's status and message).
#pragma once

#include <stdint.h>

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "mf_abi.h"

namespace mf {

struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string &m) : std::runtime_error(m), status(st) {}
};

// fixed-size column-major matrix / vector (Eigen::Matrix<double, R, C>)
template <int R, int Cc> struct Mat {
    std::array<double, R * Cc> a{};
    static Mat Zero() { return Mat(); }
    double &operator()(int r, int c) { return a[(size_t)(c * R + r)]; }
    double operator()(int r, int c) const { return a[(size_t)(c * R + r)]; }
    double *data() { return a.data(); }
    const double *data() const { return a.data(); }
    static constexpr int rows() { return R; }
    static constexpr int cols() { return Cc; }
};
template <int N> struct Vec {
    std::array<double, N> a{};
    static Vec Zero() { return Vec(); }
    double &operator()(int i) { return a[(size_t)i]; }
    double operator()(int i) const { return a[(size_t)i]; }
    double &operator[](int i) { return a[(size_t)i]; }
    double operator[](int i) const { return a[(size_t)i]; }
    double &coeffRef(int i) { return a[(size_t)i]; }
    double coeff(int i) const { return a[(size_t)i]; }
    double *data() { return a.data(); }
    const double *data() const { return a.data(); }
    static constexpr int rows() { return N; }
};
// dynamic length <= 10 (Eigen::Matrix<T, Dynamic, 1, 0, 10, 1>)
template <typename T> struct DynVec {
    std::array<T, 10> a{};
    int n = 0;
    DynVec() = default;
    explicit DynVec(int len) : n(len) {}
    void resize(int len) { n = len; }
    int rows() const { return n; }
    T &coeffRef(int i) { return a[(size_t)i]; }
    T coeff(int i) const { return a[(size_t)i]; }
    T &operator[](int i) { return a[(size_t)i]; }
    T operator[](int i) const { return a[(size_t)i]; }
    T &operator()(int i) { return a[(size_t)i]; }
    T operator()(int i) const { return a[(size_t)i]; }
};

using SampleVector = Vec<10>;
using FullSampleVector = Vec<19>;
using PulseVector = DynVec<double>;
using BXVector = DynVec<char>;
using SampleMatrix = Mat<10, 10>;
using FullSampleMatrix = Mat<19, 19>;
using SamplePulseMatrix = Mat<10, 10>;  // 10 x npulses, unused columns zero
using FixedMatrix = Mat<10, 10>;
using FixedVector = Vec<10>;

// Per-process device state: one mf_context per device, a stream and a small
// cache of the last channel types uploaded (host library only).
class Device {
  public:
    static Device &get(int device = 0);
    mf_context *ctx() { return ctx_; }
    void *stream() { return stream_; }
    int device() const { return device_; }
    // uploads types unless the same table is already on the device
    void set_types(const std::vector<mf_chtype> &types);
    // a device buffer of at least `bytes` for one call's inputs and outputs:
    // the device's staging area, grown to the largest call and kept (one call
    // at a time, like the reference's single-threaded drivers)
    void *staging(size_t bytes);
    ~Device();
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;

  private:
    explicit Device(int device);
    int device_;
    mf_context *ctx_ = nullptr;
    void *stream_ = nullptr;
    std::vector<mf_chtype> cached_;
    void *stage_ = nullptr;
    size_t stage_cap_ = 0;
};

// mf_chtype from DoFit's arguments: ECAL layout (firstsamplet = max(0, bx + 3),
// offset = 4 - bx, PulseChiSqSNNLS.cpp:54-55) over the given BX list.
mf_chtype make_chtype(const SampleMatrix &samplecor, double pederr, const BXVector &bxs,
                      const FullSampleVector &fullpulse, const FullSampleMatrix &fullpulsecov);

class PulseChiSqSNNLS {
  public:
    typedef int Index;
    PulseChiSqSNNLS();
    ~PulseChiSqSNNLS();

    bool DoFit(const SampleVector &samples, const SampleMatrix &samplecor, double pederr,
               const BXVector &bxs, const FullSampleVector &fullpulse,
               const FullSampleMatrix &fullpulsecov);

    const SamplePulseMatrix &pulsemat() const { return _pulsemat; }
    const PulseVector &X() const { return _ampvecmin; }
    const PulseVector &Errors() const { return _errvec; }
    const BXVector &BXs() const { return _bxsmin; }
    double ChiSq() const { return _chisq; }
    void disableErrorCalculation() { _computeErrors = false; }
    // per-channel MF_F_* flags of the last fit (no reference counterpart)
    int flags() const { return _flags; }

  protected:
    explicit PulseChiSqSNNLS(int solver);  // MF_SOLVER_*

  private:
    SamplePulseMatrix _pulsemat;
    PulseVector _ampvecmin, _errvec;
    BXVector _bxsmin;
    double _chisq = 0.0;
    bool _computeErrors = true;
    int _flags = 0;
    int _solver = MF_SOLVER_INPLACE_FNNLS;
};

namespace legacy {
// legacy_multifit_cpu's PulseChiSqSNNLS (interface/PulseChiSqSNNLS.h:8-33):
// the same DoFit, whose NNLS keeps its passive set and column order from one
// Minimize iteration to the next (src/PulseChiSqSNNLS.cpp:268-363), run on
// the GPU with MF_SOLVER_LEGACY.  The legacy class returns X(), BXs() and
// Errors() jointly permuted in its final column order; here they are in the
// order of the bxs given to DoFit.  Its driver reads them as (BX, X) pairs
// (src/multifit.cpp:204-224), which are the same.  invcov() is not provided.
class PulseChiSqSNNLS : public mf::PulseChiSqSNNLS {
  public:
    PulseChiSqSNNLS() : mf::PulseChiSqSNNLS(MF_SOLVER_LEGACY) {}
};
}  // namespace legacy

struct DoFitArgs {
    SampleVector samples;
    SampleMatrix samplecor;
    double pederr;
    BXVector bxs;
    FullSampleVector fullpulse;
    FullSampleMatrix fullpulsecov;
    DoFitArgs(const SampleVector &s, const SampleMatrix &sc, double pe, const BXVector &b,
              const FullSampleVector &fp, const FullSampleMatrix &fpc)
        : samples(s), samplecor(sc), pederr(pe), bxs(b), fullpulse(fp), fullpulsecov(fpc) {}
};

struct Output {
    double chi2;
    double ampl;  // amplitude of the first pulse with BX 0 (see doFitWrapper)
    int status;
    BXVector BXs;
    PulseVector X;
};

// Batched DoFit (errors disabled, as kernel_multifit runs it): channels that
// share their constants but pederr share a type (at most 256 types), pederr
// goes per channel (mf_batch.pederr); one mf_fit_batch, plain FP64 path,
// FNNLS max_iter 1000 (multifit_gpu/src/PulseChiSqSNNLS.cu:215).  BX lists
// of 1..10 pulses; Output::X has bxs.rows() entries.
// Output::ampl is X at the first BX == 0, as multifit_cpu computes it
// (multifit_cpu/src/multifit.cpp:203-211); the reference GPU kernel's test
// `ip < BXs(ip) == 0` (PulseChiSqSNNLS.cu:245) picks index 0 instead.
std::vector<Output> doFitWrapper(std::vector<DoFitArgs> const &vargs);

void inplace_fnnls(const FixedMatrix &A, const FixedVector &b, FixedVector &x,
                   const double eps = 1e-11, const unsigned int max_iterations = 1000);
void fnnls(const FixedMatrix &A, const FixedVector &b, FixedVector &x, const double eps = 1e-11,
           const unsigned int max_iterations = 1000);

struct NNLS_args {
    FixedMatrix A;
    FixedVector b;
    NNLS_args(FixedMatrix const &a, FixedVector const &bb) : A(a), b(bb) {}
};
// x starts at zero for every problem (the reference reads unset device memory)
std::vector<FixedVector> fnnls_wrapper(std::vector<NNLS_args> const &args, const double eps = 1e-11,
                                       const unsigned int max_iterations = 1000);
std::vector<FixedVector> inplace_fnnls_wrapper(std::vector<NNLS_args> const &args,
                                               const double eps = 1e-11,
                                               const unsigned int max_iterations = 1000);

// linalgebra on one problem (row-major raw pointers, full_size <= 10), run on
// the GPU through mf_linalg_batch; T = double or float.
template <typename T> void cholesky_decomp(T *pM, T *pL, int full_size, int view_size);
template <typename T>
void solve_forward_substitution(T *pM, T *pb, T *psolution, int full_size, int view_size);
template <typename T>
void solve_backward_substitution(T *pM, T *pb, T *psolution, int full_size, int view_size);
template <typename T> void swap_row_column(T *pM, int i, int j, int full_size, int view_size);
template <typename T>
void fused_cholesky_forward_substitution_solver_rcaddition(T *pM, T *pL, T *pb, T *py,
                                                           int full_size, int view_size);
template <typename T>
void fused_cholesky_forward_substitution_solver_inbetween_removal(T *pM, T *pL, T *pb, T *py,
                                                                  int position, int full_size,
                                                                  int view_size);

}  // namespace mf
