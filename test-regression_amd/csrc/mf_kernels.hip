// mf_kernels.hip -- gfx950 kernels of the multifit engine and the C ABI
// (include/mf_abi.h).
//
// Fit kernel: one channel per lane.  Samples are read structure-of-arrays
// (samples[i*nch + c]: 64 lanes read 64 consecutive values per sample row),
// the per-type constants (pulse matrix, noise covariance, pulse covariance
// blocks) are wave-uniform and are read through the scalar cache, and the
// whole fit -- covariance, its Cholesky and inverse, the whitened normal
// equations, the FNNLS active-set loop and chi2 -- runs in registers.  The
// one value parked in LDS is L^{-1} of the covariance, between the normal
// equations and chi2, laid out [element][lane] so that 64 lanes touch 64
// consecutive words.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <type_traits>
#include <vector>

#include "../../include/mf_abi.h"
#include "fit_core.hpp"
#include "mlp.hpp"
#include "mlp_device.hpp"
#include "types_host.hpp"

using namespace mf;

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";

static mf_status fail(mf_status st, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return st;
}

#define HIP_TRY(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(MF_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));     \
    } while (0)

// the same message slot for the library's other translation units
namespace mf {
mf_status set_error(mf_status st, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return st;
}
}  // namespace mf

extern "C" const char *mf_last_error(void) { return g_err; }
extern "C" int mf_abi_version(void) { return MF_ABI_VERSION; }

// ---------------------------------------------------------------------------
// fit kernels
// ---------------------------------------------------------------------------
template <typename T> struct LdsPark {
    T *base;
    int stride;
    __device__ __forceinline__ void store(int k, T v) { base[k * stride] = v; }
    __device__ __forceinline__ T load(int k) const { return base[k * stride]; }
};

// The channels one launch fits: positions 0..count-1 of a list (idx; null:
// channel = position).  A list built on the device (the type partition)
// gives its length and start through device memory, so that launching on it
// needs no host round trip; n is then an upper bound that sizes the grid.
struct Seg {
    const int32_t *idx;
    const int64_t *dcount;   // device: number of items (null: n)
    const int64_t *doffset;  // device: first item's position in idx (null: 0)
    int64_t n;
};
// The pointers of a Seg are global memory; read through the global address
// space (pointers held in a struct argument would otherwise be read with
// flat loads).
__device__ __forceinline__ int64_t seg_count(const Seg &s)
{
    return s.dcount ? *(GlobalPtr<int64_t>)s.dcount : s.n;
}
__device__ __forceinline__ GlobalPtr<int32_t> seg_list(const Seg &s)
{
    return s.idx ? (GlobalPtr<int32_t>)s.idx + (s.doffset ? *(GlobalPtr<int64_t>)s.doffset : 0)
                 : nullptr;
}

// What one launch of a fit kernel covers: either one list of one type
// (nrange == 0), or up to MAX_RANGES channel ranges, each of its own type,
// laid out wave by wave: range r takes waves wave0[r] .. wave0[r+1]-1, so
// that every wavefront holds channels of one type (its tables stay
// wave-uniform) and one launch keeps the whole chip busy across types.
constexpr int MAX_RANGES = 8;
struct SegArg {
    int nrange;
    int type_id;   // single list: its type
    Seg list;      // single list
    int rtype[MAX_RANGES];
    int64_t rfirst[MAX_RANGES], rcount[MAX_RANGES], wave0[MAX_RANGES + 1];
};
// This thread's channel c and type t; false: no channel.  t is wave-uniform.
// *range (optional): the range's index within the launch (0 for a list).
__device__ __forceinline__ bool seg_item(const SegArg &a, int &t, int64_t &c,
                                         int *range = nullptr)
{
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (range) *range = 0;
    if (a.nrange == 0) {
        t = a.type_id;
        if (gt >= seg_count(a.list)) return false;
        const GlobalPtr<int32_t> idx = seg_list(a.list);
        c = idx ? (int64_t)idx[gt] : gt;
        return true;
    }
    const int64_t w = gt >> 6;
    int r = 0;
#pragma unroll
    for (int k = 1; k < MAX_RANGES; ++k)
        if (k < a.nrange && w >= a.wave0[k]) r = k;
    r = __builtin_amdgcn_readfirstlane(r);
    if (range) *range = r;
    t = __builtin_amdgcn_readfirstlane(a.rtype[r]);
    const int64_t g = (w - a.wave0[r]) * 64 + (gt & 63);
    if (w >= a.wave0[a.nrange] || g >= a.rcount[r]) return false;
    c = a.rfirst[r] + g;
    return true;
}

// The lane's copy of the launch's parameters: its channel's noise scale
// (ped: per-channel pederr, null for the type's) and its type's pulses.
// PED: the kernel is built for a per-channel pederr; without it rho stays
// the constant 1 and costs nothing in the fit loops.
template <bool PED = true>
__device__ __forceinline__ FitParams lane_params(const FitParams &prm, const DevType &ty,
                                                 const double *__restrict__ ped, int64_t c)
{
    FitParams p = prm;
    p.pmask = ty.h.pmask;
    p.use_pi = prm.use_pi && ty.h.has_pi;   // no pulse covariance: nothing to add
    p.rho = 1.0;   // a compile-time constant unless PED
    if (PED && ped) {
        const double e = ped[c];
        p.rho = e * e * ty.inv_pederr2;
    }
    return p;
}

// D: the pulse layout the kernel is specialised for (lay_first, fit_core.hpp).
template <typename T, typename TS, int N, bool ERRORS, int XS, int D>
__global__ __launch_bounds__(256)
void k_fit_plain(const DevType *__restrict__ types, SegArg seg,
                 const TS *__restrict__ samples, const double *__restrict__ ped, int64_t nch,
                 T *__restrict__ xo, T *__restrict__ chi2o, int32_t *__restrict__ flo,
                 int32_t *__restrict__ ito, T *__restrict__ erro, FitParams prm0)
{
    extern __shared__ unsigned char lds_raw[];
    int type_id;
    int64_t c;
    if (!seg_item(seg, type_id, c)) return;
    const FitParams prm = lane_params(prm0, types[type_id], ped, c);
    const TypeMats<T> &tm = mats<T>(types[type_id]);
    T s[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) s[i] = i < N ? (T)samples[(int64_t)i * nch + c] : T(0);
    LdsPark<T> park{reinterpret_cast<T *>(lds_raw) + threadIdx.x, (int)blockDim.x};
    FitOut<T> out;
    fit_plain<T, N, N, ERRORS, XS, D>(tm, s, prm, park, out, types[type_id].h.intime);
#pragma unroll
    for (int p = 0; p < 10; ++p) xo[(int64_t)p * nch + c] = out.x[p];
    if (ERRORS) {
#pragma unroll
        for (int p = 0; p < 10; ++p) erro[(int64_t)p * nch + c] = out.err[p];
    }
    chi2o[c] = out.chi2;
    flo[c] = out.flags;
    if (ito) ito[c] = out.n_min | (out.n_outer << 8) | (out.n_inner << 16);
}

// Fast path (MF_PREC_FP32), phases fast_phase1 (FP32 active-set
// exploration) and fast_phase2 (FP64 finish).  A cold start runs both; a warm
// start goes straight to the FP64 phase from the prediction: the prediction
// is exact for most channels, and an FP64 step with the passive set's
// inverse table costs fewer instructions than an FP32 factorization (the
// FP64 vector rate equals the unpacked FP32 rate), so the FP32 phase only
// repeated work.  Types without a live pulse covariance run in one launch
// (k_fast_fused, or k_fast_warm with the MLP inside).  With a live covariance
// phase 2 runs the reference's later Minimize iterations in FP64 and needs
// about twice the registers of phase 1, so it is a launch of its own:
// k_fast_p1 leaves its amplitudes in the output x (FP32), its flags and
// counters in flags / iters; k_fast_p2 reads them back, finishes the fit and
// writes the final outputs.
// Warm start (WarmIn): the MLP's predicted passive set pm[c] and amplitudes
// x[i][c]; a predicted passive value starts at max(x, 1e-6) (phase 1 needs a
// feasible point).  Null pointers: cold start.
struct WarmIn {
    const float *x;     // [10][nch]
    const uint32_t *pm; // [nch]
};
template <int N>
__device__ __forceinline__ uint32_t warm_load(const WarmIn &w, int64_t c, int64_t nch,
                                              float (&x)[10], uint32_t pmask)
{
    uint32_t pm = 0;
#pragma unroll
    for (int i = 0; i < 10; ++i) x[i] = 0.0f;
    if (w.pm) {
        pm = w.pm[c] & pmask & ((1u << N) - 1u);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const float v = w.x[(int64_t)i * nch + c];
            x[i] = ((pm >> i) & 1u) ? fmaxf(v, 1e-6f) : 0.0f;
        }
    }
    return pm;
}

template <typename TS, int N, bool PED>
__global__ __launch_bounds__(256)
void k_fast_p1(const DevType *__restrict__ types, SegArg seg,
               const TS *__restrict__ samples, const double *__restrict__ ped, int64_t nch,
               float *__restrict__ xo, int32_t *__restrict__ flo, int32_t *__restrict__ ito,
               WarmIn warm, FitParams prm0)
{
    int type_id;
    int64_t c;
    if (!seg_item(seg, type_id, c)) return;
    const FitParams prm = lane_params<PED>(prm0, types[type_id], ped, c);
    float s[10], x[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) s[i] = i < N ? (float)samples[(int64_t)i * nch + c] : 0.0f;
    uint32_t pm = warm_load<N>(warm, c, nch, x, prm.pmask);
    int flags = 0;
    FnnlsStats st;
    if (!warm.pm && all_finite<float, N>(s) && rho_ok(prm.rho) && prm.max_minimize > 0)
        fast_phase1<N, N>(types[type_id], s, prm, x, pm, st, flags);
#pragma unroll
    for (int p = 0; p < N; ++p) xo[(int64_t)p * nch + c] = ((pm >> p) & 1u) ? x[p] : 0.0f;
    flo[c] = flags | (st.hit_max ? F_FNNLS_CAP : 0) | (st.llt_fail ? F_LLT_FAIL : 0);
    ito[c] = (st.outer << 8) | (st.inner << 16);
}

// Both phases in one launch, for types whose pulse covariance is not live:
// the FP64 phase of one wavefront overlaps the FP32 phase of others, and the
// hand-over through x / flags / iters stays in registers.  Results are
// identical to k_fast_p1 + k_fast_p2.
struct NoPark {
    __device__ void store(int, double) {}
    __device__ double load(int) const { return 0.0; }
};
// Reads a channel's samples again where phase 2 uses them, so that they need
// not stay in registers (or spill) through the solves.  The pointer is
// laundered so the compiler does not reuse the values it loaded before, then
// cast back to the global address space: a generic pointer would make these
// flat loads, which the scalar loads' waits also wait for.
template <typename TS, int N> struct SampleReload {
    const TS *s;
    int64_t nch, c;
    __device__ void operator()(float (&v)[10]) const
    {
        const TS *q = s;
        asm volatile("" : "+s"(q));
        const GlobalPtr<TS> g = (GlobalPtr<TS>)q;
#pragma unroll
        for (int i = 0; i < 10; ++i) v[i] = i < N ? (float)g[(int64_t)i * nch + c] : 0.0f;
    }
};
// Pending lists (after the hand-over's FP64 values): one count per range of
// the launch (MAX_RANGES), cleared by the first-iteration launch, then the
// items (the index seg_item reads).  Range r files its
// items from slot wave0[r] * 64 on (a single list from slot 0), so that the
// lists need no shared count and k_fast_p4 takes one range per grid row.
constexpr int F_PENDING = 1 << 30;   // between k_fast_p3 and k_fast_p4 only
__host__ __device__ inline size_t hand_fp64_bytes(int64_t nch)
{
    return ((size_t)nch * 11 * sizeof(double) + 255) / 256 * 256;
}
// slots: at most nch items plus each range's last partial wavefront
__host__ __device__ inline size_t hand_bytes(int64_t nch)
{
    return hand_fp64_bytes(nch) + 256 + sizeof(int32_t) * ((size_t)nch + 64 * MAX_RANGES);
}
__host__ __device__ inline unsigned *hand_pending(double *hand, int64_t nch)
{
    return reinterpret_cast<unsigned *>(reinterpret_cast<unsigned char *>(hand) +
                                        hand_fp64_bytes(nch));
}

// Hand-over of a channel after phase 2's first iteration, for types whose
// pulse covariance is live: that iteration runs in a light kernel (four
// waves per SIMD), the later iterations, which need about twice the
// registers, in k_fast_p3.  `hand` holds the FP64 amplitudes [10][nch] and
// chi2 [nch] (after them); flags and the FNNLS counters go to the outputs, as
// between k_fast_p1 and k_fast_p2.  A channel that is not fitted gets its
// final outputs here (its flags say so: F_NONFINITE or F_MINIMIZE_CAP, which
// no fitted channel carries at this point).
__device__ __forceinline__ void hand_over(bool begun, const ActiveState<double> &S, double chi2,
                                          int flags, const FnnlsStats &st, int64_t c,
                                          int64_t nch, double *__restrict__ hand,
                                          float *__restrict__ xo, float *__restrict__ chi2o,
                                          int32_t *__restrict__ flo, int32_t *__restrict__ ito)
{
    if (begun) {
#pragma unroll
        for (int p = 0; p < 10; ++p) hand[(int64_t)p * nch + c] = S.x[p];
        hand[10 * nch + c] = chi2;
    } else {
#pragma unroll
        for (int p = 0; p < 10; ++p) xo[(int64_t)p * nch + c] = 0.0f;
        chi2o[c] = 0.0f;
    }
    flo[c] = flags | (st.hit_max ? F_FNNLS_CAP : 0) | (st.llt_fail ? F_LLT_FAIL : 0);
    ito[c] = (st.outer << 8) | (st.inner << 16);
}

// Four waves per SIMD: at 128 VGPRs a few values spill (about 56 bytes per
// lane, outside the solve loops); measured faster than 137 VGPRs at three.
// HO: live pulse covariance, the first iteration only (hand_over).
template <typename TS, int N, bool PED, bool HO = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void k_fast_fused(const DevType *__restrict__ types, SegArg seg,
                  const TS *__restrict__ samples, const double *__restrict__ ped, int64_t nch,
                  float *__restrict__ xo, float *__restrict__ chi2o, int32_t *__restrict__ flo,
                  int32_t *__restrict__ ito, WarmIn warm, FitParams prm0,
                  double *__restrict__ hand = nullptr)
{
    // HO: the first block clears the pending counts of k_fast_p3 (this
    // launch precedes it on the stream)
    if (HO && blockIdx.x == 0 && threadIdx.x < MAX_RANGES) hand_pending(hand, nch)[threadIdx.x] = 0u;
    int type_id;
    int64_t c;
    if (!seg_item(seg, type_id, c)) return;
    const FitParams prm = lane_params<PED>(prm0, types[type_id], ped, c);
    float s[10], x[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) s[i] = i < N ? (float)samples[(int64_t)i * nch + c] : 0.0f;
    uint32_t pm = warm_load<N>(warm, c, nch, x, prm.pmask);
    int flags = 0;
    FnnlsStats st;
    const DevType &ty = types[type_id];
    if (!warm.pm && all_finite<float, N>(s) && rho_ok(prm.rho) && prm.max_minimize > 0)
        fast_phase1<N, N>(ty, s, prm, x, pm, st, flags);
#pragma unroll
    for (int p = 0; p < 10; ++p) x[p] = (p < N && ((pm >> p) & 1u)) ? x[p] : 0.0f;
    if (HO) {
        ActiveState<double> S;
        double chi2;
        int fl = flags & F_FP64_PIVOT;
        SampleReload<TS, N> rl{samples, nch, c};
        const bool begun = fast_p2_first<N, N, true>(ty, s, prm, x, warm.pm != nullptr, st, fl,
                                                     S, chi2, rl);
        hand_over(begun, S, chi2, fl, st, c, nch, hand, xo, chi2o, flo, ito);
        return;
    }
    NoPark park;
    FitOut<float> out;
    fast_phase2<N, N, false>(ty, s, prm, x, warm.pm != nullptr, st, flags & F_FP64_PIVOT, park,
                             out, SampleReload<TS, N>{samples, nch, c});
#pragma unroll
    for (int p = 0; p < 10; ++p) xo[(int64_t)p * nch + c] = out.x[p];
    chi2o[c] = out.chi2;
    flo[c] = out.flags;
    ito[c] = out.n_min | (out.n_outer << 8) | (out.n_inner << 16);
}

// The warm-started fast path in one launch: the MLP's prediction for the
// block's channels (the network staged in LDS once per block, the forward
// pass on MFMA, as k_mlp_warm), then the FP64 phase from that prediction.
// The prediction never leaves the chip.  Blocks of FAST_WARM_BLOCK threads.
constexpr int FAST_WARM_BLOCK = 256;
template <typename TS, int N, bool PED, bool HO = false>
__global__ __launch_bounds__(FAST_WARM_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
void k_fast_warm(const DevType *__restrict__ types, SegArg seg,
                 const TS *__restrict__ samples, const double *__restrict__ ped, int64_t nch,
                 float *__restrict__ xo, float *__restrict__ chi2o, int32_t *__restrict__ flo,
                 int32_t *__restrict__ ito, const uint16_t *__restrict__ frag,
                 const float *__restrict__ pq, FitParams prm0, double *__restrict__ hand = nullptr)
{
    __shared__ mlp::MlpLds L;
    __shared__ __bf16 xs[FAST_WARM_BLOCK / 64][64][16];
    __shared__ float pz[FAST_WARM_BLOCK / 64][10][64];
    __shared__ uint32_t pmw[FAST_WARM_BLOCK / 64][64];
    // HO: the first block clears the pending counts of k_fast_p3 (this
    // launch precedes it on the stream)
    if (HO && blockIdx.x == 0 && threadIdx.x < MAX_RANGES) hand_pending(hand, nch)[threadIdx.x] = 0u;
    // a block past the end of its type's list (the grid of a type's launch is
    // sized for the whole batch) stages nothing
    if (seg.nrange == 0 && (int64_t)blockIdx.x * FAST_WARM_BLOCK >= seg_count(seg.list)) return;
    mlp::mlp_stage(L, frag, pq, threadIdx.x, FAST_WARM_BLOCK);
    __syncthreads();
    int type_id = 0;
    int64_t c = 0;
    const bool live = seg_item(seg, type_id, c);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float s[10], xp[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) s[i] = (live && i < N) ? (float)samples[(int64_t)i * nch + c] : 0.0f;
    // every lane of the wavefront takes part in the forward pass
    uint32_t pm = mlp::mlp_warm_wave(L, xs[wv], pz[wv], pmw[wv], lane, s, N, xp);
    if (!live) return;
    const DevType &ty = types[type_id];
    const FitParams prm = lane_params<PED>(prm0, ty, ped, c);
    pm &= prm.pmask & ((1u << N) - 1u);
    float x[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) x[i] = (i < N && ((pm >> i) & 1u)) ? fmaxf(xp[i], 1e-6f) : 0.0f;
    // a warm start goes straight to the FP64 phase (see k_fast_p1)
    FnnlsStats st;
    if (HO) {
        ActiveState<double> S;
        double chi2;
        int fl = 0;
        SampleReload<TS, N> rl{samples, nch, c};
        const bool begun = fast_p2_first<N, N, true>(ty, s, prm, x, true, st, fl, S, chi2, rl);
        hand_over(begun, S, chi2, fl, st, c, nch, hand, xo, chi2o, flo, ito);
        return;
    }
    NoPark park;
    FitOut<float> out;
    fast_phase2<N, N, false>(ty, s, prm, x, true, st, 0, park, out,
                             SampleReload<TS, N>{samples, nch, c});
#pragma unroll
    for (int p = 0; p < 10; ++p) xo[(int64_t)p * nch + c] = out.x[p];
    chi2o[c] = out.chi2;
    flo[c] = out.flags;
    ito[c] = out.n_min | (out.n_outer << 8) | (out.n_inner << 16);
}

template <typename TS, int N, bool LIVE, bool PED, int D>
__global__ __launch_bounds__(256)
void k_fast_p2(const DevType *__restrict__ types, SegArg seg,
               const TS *__restrict__ samples, const double *__restrict__ ped, int64_t nch,
               float *__restrict__ xo, float *__restrict__ chi2o, int32_t *__restrict__ flo,
               int32_t *__restrict__ ito, int warm, FitParams prm0)
{
    extern __shared__ unsigned char lds_raw[];
    int type_id;
    int64_t c;
    if (!seg_item(seg, type_id, c)) return;
    const FitParams prm = lane_params<PED>(prm0, types[type_id], ped, c);
    float s[10], x[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        s[i] = i < N ? (float)samples[(int64_t)i * nch + c] : 0.0f;
        x[i] = i < N ? xo[(int64_t)i * nch + c] : 0.0f;
    }
    const int f1 = flo[c], it1 = ito[c];
    FnnlsStats st;
    st.outer = (it1 >> 8) & 255;
    st.inner = (it1 >> 16) & 0xFFFF;
    st.hit_max = (f1 & F_FNNLS_CAP) != 0;
    st.llt_fail = (f1 & F_LLT_FAIL) != 0;
    LdsPark<double> park{reinterpret_cast<double *>(lds_raw) + threadIdx.x, (int)blockDim.x};
    FitOut<float> out;
    fast_phase2<N, N, LIVE, true, D>(types[type_id], s, prm, x, warm != 0, st, f1 & F_FP64_PIVOT,
                                     park, out);
#pragma unroll
    for (int p = 0; p < 10; ++p) xo[(int64_t)p * nch + c] = out.x[p];
    chi2o[c] = out.chi2;
    flo[c] = out.flags;
    ito[c] = out.n_min | (out.n_outer << 8) | (out.n_inner << 16);
}

// The later Minimize iterations of a live pulse covariance after a hand-over
// (k_fast_warm / k_fast_fused with HO), in two kernels.  k_fast_p3 runs them
// while each one stops at its start or adds a single index
// (fast_p2_rest_lazy): no normal equations, so it needs half the registers
// of the whole iteration body and no park for L^{-1}.  A channel that reaches
// an iteration needing the normal equations leaves its state in `hand`, its
// flags with F_PENDING and its iteration index in the outputs, and its item
// in the pending list; k_fast_p4 finishes those channels with the whole body
// (fast_p2_resume).  Results are those of fast_p2_rest bit for bit.
//
// Two waves per SIMD: the compiler then spills a few values (84 bytes per
// lane); measured faster than one wave at 276 registers (0.316 against
// 0.374 ms for the warm step of 1M live channels).
template <typename TS, int N, bool PED, int D>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_fast_p3(const DevType *__restrict__ types, SegArg seg,
               const TS *__restrict__ samples, const double *__restrict__ ped, int64_t nch,
               float *__restrict__ xo, float *__restrict__ chi2o, int32_t *__restrict__ flo,
               int32_t *__restrict__ ito, double *__restrict__ hand, FitParams prm0)
{
    int type_id, r;
    int64_t c;
    if (!seg_item(seg, type_id, c, &r)) return;
    const int f1 = flo[c], it1 = ito[c];
    if (f1 & (F_NONFINITE | F_MINIMIZE_CAP)) return;   // not fitted: outputs written
    const FitParams prm = lane_params<PED>(prm0, types[type_id], ped, c);
    float s[10];
    double x[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        s[i] = i < N ? (float)samples[(int64_t)i * nch + c] : 0.0f;
        x[i] = hand[(int64_t)i * nch + c];
    }
    double chi2 = hand[10 * nch + c];
    FnnlsStats st;
    st.outer = (it1 >> 8) & 255;
    st.inner = (it1 >> 16) & 0xFFFF;
    st.hit_max = (f1 & F_FNNLS_CAP) != 0;
    st.llt_fail = (f1 & F_LLT_FAIL) != 0;
    int flags = f1 & ~(F_FNNLS_CAP | F_LLT_FAIL);
    bool pend;
    const int nmin = fast_p2_rest_lazy<N, N, D>(types[type_id], s, prm, x, chi2, st, flags, pend);
    // the stores' addresses are formed here again from laundered bases, not
    // kept in registers through the iterations; the bases are cast back to
    // the global address space (generic pointers would make flat stores)
    asm volatile("" : "+s"(hand), "+s"(xo), "+s"(chi2o), "+s"(flo), "+s"(ito));
    const GlobalOut<double> hg = (GlobalOut<double>)hand;
    const GlobalOut<int32_t> fg = (GlobalOut<int32_t>)flo, ig = (GlobalOut<int32_t>)ito;
    if (pend) {
#pragma unroll
        for (int p = 0; p < 10; ++p) hg[(int64_t)p * nch + c] = x[p];
        hg[10 * nch + c] = chi2;
        fg[c] = flags | F_PENDING | (st.hit_max ? F_FNNLS_CAP : 0) |
                (st.llt_fail ? F_LLT_FAIL : 0);
        ig[c] = nmin | (st.outer << 8) | (st.inner << 16);
        const GlobalOut<unsigned> pq = (GlobalOut<unsigned>)hand_pending(hand, nch);
        const unsigned k =
            __hip_atomic_fetch_add(pq + r, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int64_t base = seg.nrange ? seg.wave0[r] * 64 : 0;
        ((GlobalOut<int32_t>)(pq + 64))[base + k] =
            (int32_t)((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
        return;
    }
    FitOut<float> out;
    fast_p2_out(x, chi2, flags, nmin, st, out);
    const GlobalOut<float> xg = (GlobalOut<float>)xo, cg = (GlobalOut<float>)chi2o;
#pragma unroll
    for (int p = 0; p < 10; ++p) xg[(int64_t)p * nch + c] = out.x[p];
    cg[c] = out.chi2;
    fg[c] = out.flags;
    ig[c] = out.n_min | (out.n_outer << 8) | (out.n_inner << 16);
}

// The pending channels of k_fast_p3: grid row blockIdx.y takes range
// blockIdx.y's list (one type: its tables are read with uniform addresses),
// one wavefront per block, grid-stride over the list.  L^{-1} is parked in
// LDS (55 doubles per lane).
constexpr int FAST_P4_BLOCK = 64;
template <typename TS, int N, bool PED, int D>
__global__ __launch_bounds__(FAST_P4_BLOCK)
void k_fast_p4(const DevType *__restrict__ types, SegArg seg,
               const TS *__restrict__ samples, const double *__restrict__ ped, int64_t nch,
               float *__restrict__ xo, float *__restrict__ chi2o, int32_t *__restrict__ flo,
               int32_t *__restrict__ ito, double *__restrict__ hand, FitParams prm0)
{
    extern __shared__ unsigned char lds_raw[];
    const int r = blockIdx.y;
    const unsigned *pq = hand_pending(hand, nch);
    const int64_t np = (int64_t)pq[r];
    const int32_t *list =
        reinterpret_cast<const int32_t *>(pq + 64) + (seg.nrange ? seg.wave0[r] * 64 : 0);
    const int type_id = seg.nrange ? seg.rtype[r] : seg.type_id;
    const DevType &ty = types[type_id];
    for (int64_t k0 = (int64_t)blockIdx.x * FAST_P4_BLOCK; k0 < np;
         k0 += (int64_t)gridDim.x * FAST_P4_BLOCK) {
        const int64_t k = k0 + threadIdx.x;
        if (k >= np) continue;
        // the item's channel: as seg_item maps the item's thread index
        const int64_t gt = list[k];
        int64_t c;
        if (seg.nrange == 0) {
            const GlobalPtr<int32_t> idx = seg_list(seg.list);
            c = idx ? (int64_t)idx[gt] : gt;
        } else {
            c = seg.rfirst[r] + (gt - seg.wave0[r] * 64);
        }
        const int f1 = flo[c], it1 = ito[c];
        const FitParams prm = lane_params<PED>(prm0, ty, ped, c);
        float s[10];
        double x[10];
#pragma unroll
        for (int i = 0; i < 10; ++i) {
            s[i] = i < N ? (float)samples[(int64_t)i * nch + c] : 0.0f;
            x[i] = hand[(int64_t)i * nch + c];
        }
        double chi2 = hand[10 * nch + c];
        FnnlsStats st;
        st.outer = (it1 >> 8) & 255;
        st.inner = (it1 >> 16) & 0xFFFF;
        st.hit_max = (f1 & F_FNNLS_CAP) != 0;
        st.llt_fail = (f1 & F_LLT_FAIL) != 0;
        int flags = f1 & ~(F_PENDING | F_FNNLS_CAP | F_LLT_FAIL);
        LdsPark<double> park{reinterpret_cast<double *>(lds_raw) + threadIdx.x, FAST_P4_BLOCK};
        const int nmin = fast_p2_resume<N, N, D>(ty, s, prm, x, chi2, it1 & 255, st, flags, park);
        FitOut<float> out;
        fast_p2_out(x, chi2, flags, nmin, st, out);
#pragma unroll
        for (int p = 0; p < 10; ++p) xo[(int64_t)p * nch + c] = out.x[p];
        chi2o[c] = out.chi2;
        flo[c] = out.flags;
        ito[c] = out.n_min | (out.n_outer << 8) | (out.n_inner << 16);
    }
}

// ---------------------------------------------------------------------------
// persistent fast-path kernels: lanes refill from a device work queue
// ---------------------------------------------------------------------------
// Active-set iteration counts vary from channel to channel (pileup), and a
// wavefront runs as long as its slowest lane.  In these kernels a lane that
// has finished its channel takes the next one from a work queue while the
// other lanes go on: the grid is sized to the machine (persistent waves), and
// each pass of a wavefront's loop runs one active-set step (fnnls_step) on
// every lane that holds a channel.
//
// Queue claims of a wavefront: items are claimed 64 at a time (one atomic
// add by lane 0), and the next range is claimed one range ahead.  Lane 0
// keeps that claim's result unread until the range is needed, so the atomic
// completes in the background.  rb, ru: wave-uniform.
struct QueueCursor {
    unsigned rb, ru;  // current range base, items used from it
    unsigned nb0;     // lane 0: base of the next range (claimed ahead)
    __device__ void start(unsigned *head)
    {
        unsigned a = 0;
        nb0 = 0;
        if (__lane_id() == 0) {
            a = atomicAdd(head, 64u);
            nb0 = atomicAdd(head, 64u);
        }
        rb = __shfl(a, 0);
        ru = 0;
    }
    // Items for the k lanes of `empty` (all 64 lanes call): this lane's item,
    // INT64_MAX for a lane outside `empty`.
    __device__ int64_t take(unsigned *head, uint64_t empty)
    {
        const int k = __popcll(empty);
        const int r = __popcll(empty & ((1ull << __lane_id()) - 1ull));
        const unsigned t1 = min((unsigned)k, 64u - ru);
        int64_t it = (int64_t)rb + ru + r;
        if (ru + (unsigned)k >= 64u) {
            // the current range runs out: move to the claimed-ahead one and
            // claim the next
            const unsigned nb = __shfl(nb0, 0);
            if ((unsigned)r >= t1) it = (int64_t)nb + (r - t1);
            if (__lane_id() == 0) nb0 = atomicAdd(head, 64u);
            rb = nb;
            ru = (unsigned)k - t1;
        } else {
            ru += (unsigned)k;
        }
        return ((empty >> __lane_id()) & 1ull) ? it : INT64_MAX;
    }
    __device__ bool exhausted(int64_t nitems) const { return (int64_t)rb + ru >= nitems; }
};

// Raw warm-start prediction of one channel (loaded unconditionally, so the
// loads need not complete until the channel is started); warm_start_state
// turns it into the starting point.
template <int N>
__device__ __forceinline__ void warm_fetch(const WarmIn &w, int64_t c, int64_t nch,
                                           uint32_t &pm, float (&x)[10])
{
    pm = 0;
#pragma unroll
    for (int i = 0; i < 10; ++i) x[i] = 0.0f;
    if (w.pm) {
        pm = w.pm[c];
#pragma unroll
        for (int i = 0; i < N; ++i) x[i] = w.x[(int64_t)i * nch + c];
    }
}
template <int N>
__device__ __forceinline__ uint32_t warm_start_state(uint32_t pm, const float (&xr)[10],
                                                     float (&x)[10], uint32_t pmask)
{
    pm &= pmask & ((1u << N) - 1u);
#pragma unroll
    for (int i = 0; i < 10; ++i) x[i] = (i < N && ((pm >> i) & 1u)) ? fmaxf(xr[i], 1e-6f) : 0.0f;
    return pm;
}

// Phase 1, persistent.  Each lane holds its current channel and a prefetched
// next one (samples and warm-start prediction, loaded while the current one
// is worked on); a lane that finishes switches to its next channel at once,
// and empty prefetch slots are refilled from the queue once at least `need`
// of them are empty.
template <typename TS, int N, bool PED>
__global__ __launch_bounds__(256)
void k_fast_p1q(const DevType *__restrict__ types, int type_id, Seg seg,
                const TS *__restrict__ samples, const double *__restrict__ ped, int64_t nch,
                float *__restrict__ xo, int32_t *__restrict__ flo, int32_t *__restrict__ ito,
                WarmIn warm, FitParams prm, unsigned *__restrict__ head, int need)
{
    const DevType &ty = types[type_id];
    const int64_t nitems = seg_count(seg);
    const GlobalPtr<int32_t> idx = seg_list(seg);
    int64_t c = -1, cn = -1;
    float Atb[10], svn[10], xn[10];
    float eps1 = 0.0f;   // the lane's threshold: eps times its channel's noise scale
    double rhon = 1.0;
    uint32_t pmn = 0;
    ActiveState<float> S;
    FnnlsStats st;
    bool ill = false;
    QueueCursor q;
    q.start(head);
    for (;;) {
        // refill empty prefetch slots
        const uint64_t em = __ballot(cn < 0);
        if (!q.exhausted(nitems) && __popcll(em) >= min(need, 64)) {
            const int64_t it = q.take(head, em);
            if (it < nitems) {
                cn = idx ? (int64_t)idx[it] : it;
#pragma unroll
                for (int i = 0; i < 10; ++i) svn[i] = i < N ? (float)samples[(int64_t)i * nch + cn] : 0.0f;
                warm_fetch<N>(warm, cn, nch, pmn, xn);
                if (PED && ped) {
                    const double e = ped[cn];
                    rhon = e * e * ty.inv_pederr2;
                }
            }
        }
        // a lane without a channel takes its prefetched one
        if (c < 0 && cn >= 0) {
            c = cn;
            cn = -1;
            S.pm = warm_start_state<N>(pmn, xn, S.x, ty.h.pmask);
            eps1 = (float)(prm.eps * rhon);
            S.solve = S.pm != 0;
            S.adds = 0;
            st = FnnlsStats();
            ill = false;
            if (warm.pm) {
                // a warm start goes straight to the FP64 phase (k_fast_p2q)
#pragma unroll
                for (int p = 0; p < N; ++p)
                    xo[(int64_t)p * nch + c] = ((S.pm >> p) & 1u) ? S.x[p] : 0.0f;
                flo[c] = 0;
                ito[c] = 0;
                c = -1;
            } else if (all_finite<float, N>(svn) && rho_ok(rhon) && prm.max_minimize > 0) {
                const TypeMats<float> &tf = *opaque(&ty.f);
#pragma unroll
                for (int p = 0; p < 10; ++p) {
                    float t = 0.0f;
#pragma unroll
                    for (int k = 0; k < N; ++k) t = mfma(tf.G0[p][k], svn[k], t);
                    Atb[p] = p < N ? t : 0.0f;
                }
            } else {
#pragma unroll
                for (int p = 0; p < N; ++p) xo[(int64_t)p * nch + c] = 0.0f;
                flo[c] = 0;
                ito[c] = 0;
                c = -1;
            }
        }
        if (c >= 0 &&
            fnnls_step<float, N, true>(ty.f.AtA0, Atb, S, eps1, prm.eps_rel32,
                                       prm.max_iter, st, prm.tau32, &ill)) {
#pragma unroll
            for (int p = 0; p < N; ++p)
                xo[(int64_t)p * nch + c] = ((S.pm >> p) & 1u) ? S.x[p] : 0.0f;
            flo[c] = (ill ? F_FP64_PIVOT : 0) | (st.hit_max ? F_FNNLS_CAP : 0) |
                     (st.llt_fail ? F_LLT_FAIL : 0);
            ito[c] = (st.outer << 8) | (st.inner << 16);
            c = -1;
        }
        if (q.exhausted(nitems) && __ballot(c >= 0 || cn >= 0) == 0) break;
    }
}

// Phase 2 (zero or unused pulse covariance), persistent, with the same
// prefetch and refill scheme; the next channel's samples and phase-1 result
// are loaded while the current one is worked on.
template <typename TS, int N, bool PED>
__global__ __launch_bounds__(256)
void k_fast_p2q(const DevType *__restrict__ types, int type_id, Seg seg,
                const TS *__restrict__ samples, const double *__restrict__ ped, int64_t nch,
                float *__restrict__ xo, float *__restrict__ chi2o, int32_t *__restrict__ flo,
                int32_t *__restrict__ ito, int warm, FitParams prm0, unsigned *__restrict__ head,
                int need)
{
    FitParams prm = prm0;   // rho: the lane's current channel
    prm.rho = 1.0;
    double rhon = 1.0;
    const DevType &ty = types[type_id];
    const int64_t nitems = seg_count(seg);
    const GlobalPtr<int32_t> idx = seg_list(seg);
    int64_t c = -1, cn = -1;
    double Atb[10];
    ActiveState<double> S;
    FnnlsStats st;
    int flags = 0, outer0 = 0;
    // 0: the start is a prediction, 1: it fell back to x = 0, 2: it was the
    // empty set (the cold start: nothing to fall back to)
    int restart = 0;
    uint32_t polished = 0;  // the set last solved by its factor (polish_end)
    float svn[10], xn[10];
    int fn = 0, itn = 0;
    QueueCursor q;
    q.start(head);
    for (;;) {
        const uint64_t em = __ballot(cn < 0);
        if (!q.exhausted(nitems) && __popcll(em) >= min(need, 64)) {
            const int64_t it = q.take(head, em);
            if (it < nitems) {
                cn = idx ? (int64_t)idx[it] : it;
#pragma unroll
                for (int i = 0; i < 10; ++i) {
                    svn[i] = i < N ? (float)samples[(int64_t)i * nch + cn] : 0.0f;
                    xn[i] = i < N ? xo[(int64_t)i * nch + cn] : 0.0f;
                }
                fn = flo[cn];
                itn = ito[cn];
                if (PED && ped) {
                    const double e = ped[cn];
                    rhon = e * e * ty.inv_pederr2;
                }
            }
        }
        if (c < 0 && cn >= 0) {
            c = cn;
            cn = -1;
            prm.rho = rhon;
            st = FnnlsStats();
            st.outer = (itn >> 8) & 255;
            st.inner = (itn >> 16) & 0xFFFF;
            st.hit_max = (fn & F_FNNLS_CAP) != 0;
            st.llt_fail = (fn & F_LLT_FAIL) != 0;
            outer0 = st.outer;
            flags = fn & F_FP64_PIVOT;
            const bool begun = fast_p2_begin<N, N>(ty, svn, prm, xn, Atb, S, flags);
            restart = S.pm == 0 ? 2 : 0;
            polished = 0;
            if (!begun) {
#pragma unroll
                for (int p = 0; p < 10; ++p) xo[(int64_t)p * nch + c] = 0.0f;
                chi2o[c] = 0.0f;
                flo[c] = flags | (st.hit_max ? F_FNNLS_CAP : 0) | (st.llt_fail ? F_LLT_FAIL : 0);
                ito[c] = (st.outer << 8) | (st.inner << 16);
                c = -1;
            }
        }
        if (c >= 0 &&
            fnnls_step<double, N, true, true>(ty.d.AtA0, Atb, S, prm.eps * prm.rho, prm.eps_rel64,
                                              prm.max_iter, st, 0.0, nullptr, *opaque(&ty.tab)) &&
            !(warm && st.hit_max && restart == 0) &&
            !polish_end<N>(*opaque(&ty.tab), Atb, S, st, polished)) {
            flags |= F_REFINED;
            if (warm && st.outer != outer0) flags |= F_WARM_REJECT;
            if (restart == 1) flags |= F_WARM_FALLBACK;
            // the samples are read again here rather than held through the steps
            float sv[10];
#pragma unroll
            for (int i = 0; i < 10; ++i) sv[i] = i < N ? (float)samples[(int64_t)i * nch + c] : 0.0f;
            const double chi2 = fast_p2_chi2<N, N>(ty, sv, S.x) / prm.rho;
            const int nmin = fixed_cov_nmin(chi2, prm, flags);
            if (st.hit_max) flags |= F_FNNLS_CAP;
            if (st.llt_fail) flags |= F_LLT_FAIL;
#pragma unroll
            for (int p = 0; p < 10; ++p) xo[(int64_t)p * nch + c] = (float)S.x[p];
            chi2o[c] = (float)chi2;
            flo[c] = flags;
            ito[c] = nmin | (st.outer << 8) | (st.inner << 16);
            c = -1;
        }
        // a warm start whose iterations reached max_iter starts over cold
        if (c >= 0 && warm && st.hit_max && restart == 0) {
            warm_fallback_reset(S, st);
            restart = 1;
            polished = 0;
        }
        if (q.exhausted(nitems) && __ballot(c >= 0 || cn >= 0) == 0) break;
    }
}

__global__ void k_zero_u32(unsigned *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Partition of a mixed batch by channel type, on the device only (no host
// round trip, so a fit call can be captured in a graph), without atomics on
// shared addresses and stable: each type's list holds its channels in
// ascending order.  The batch is cut into at most PART_BLOCKS contiguous
// chunks, one per block:
//   k_part_count    the chunk's count of each type (wavefront ballots);
//   k_part_scan     per type, the exclusive scan of the chunk counts (the
//                   chunk's first slot within the type) and the type's total;
//   k_type_scan     the exclusive scan of the totals (each type's first slot);
//   k_part_scatter  each chunk again, 256 channels at a time: a channel's slot
//                   is its type's next slot in the chunk plus its rank among
//                   the earlier channels of its type in those 256.
// part layout: counts[256], (unused)[256], offsets[256] (int64), then
// bcount[256][PART_BLOCKS] and boff[256][PART_BLOCKS] (int32).
constexpr int PART_BLOCKS = 512;
constexpr size_t PART_BYTES = sizeof(int64_t) * 768 + 2 * sizeof(int32_t) * 256 * PART_BLOCKS;

__global__ void k_zero_u64(unsigned long long *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// The type ids of a wavefront's lanes (t < 0: no channel), one type at a
// time: fn(type, lanes of that type, this lane's rank among them) on every
// lane, with this lane's own rank only where the type is its own.
template <typename Fn>
__device__ __forceinline__ void wave_types(int t, Fn &&fn)
{
    const int lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    bool pending = t >= 0;
    for (uint64_t pm = __ballot(pending); pm; pm = __ballot(pending)) {
        const int lead = __ffsll((unsigned long long)pm) - 1;
        const int tt = __shfl(t, lead);
        const bool me = pending && t == tt;
        const uint64_t mine = __ballot(me);
        fn(tt, lane == lead, mine, __popcll(mine & below));
        if (me) pending = false;
    }
}

__global__ __launch_bounds__(256) void k_part_count(const uint8_t *__restrict__ type_of,
                                                    int64_t nch, int ntypes, int64_t chunk,
                                                    int32_t *__restrict__ bcount)
{
    __shared__ int cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t c0 = (int64_t)blockIdx.x * chunk, c1 = c0 + chunk < nch ? c0 + chunk : nch;
    for (int64_t base = c0; base < c1; base += 256) {
        const int64_t c = base + threadIdx.x;
        int t = c < c1 ? (int)type_of[c] : -1;
        if (t >= ntypes) t = -1;
        wave_types(t, [&](int tt, bool lead, uint64_t mine, int) {
            if (lead) atomicAdd(&cnt[tt], __popcll(mine));
        });
    }
    __syncthreads();
    if ((int)threadIdx.x < ntypes) bcount[threadIdx.x * PART_BLOCKS + blockIdx.x] = cnt[threadIdx.x];
}

// one block per type, nblk <= 2 * 256 chunks
__global__ __launch_bounds__(256) void k_part_scan(int nblk, const int32_t *__restrict__ bcount,
                                                   int32_t *__restrict__ boff,
                                                   long long *__restrict__ counts)
{
    __shared__ int v[256];
    const int t = blockIdx.x, i = threadIdx.x;
    const int32_t *bc = bcount + (size_t)t * PART_BLOCKS;
    const int a0 = 2 * i < nblk ? bc[2 * i] : 0, a1 = 2 * i + 1 < nblk ? bc[2 * i + 1] : 0;
    v[i] = a0 + a1;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const int x = i >= d ? v[i - d] : 0;
        __syncthreads();
        v[i] += x;
        __syncthreads();
    }
    const int ex = v[i] - (a0 + a1);
    if (2 * i < nblk) boff[(size_t)t * PART_BLOCKS + 2 * i] = ex;
    if (2 * i + 1 < nblk) boff[(size_t)t * PART_BLOCKS + 2 * i + 1] = ex + a0;
    if (i == 255) counts[t] = v[255];
}

// one block of 256 threads: offsets[t] = sum of counts[u], u < t
__global__ void k_type_scan(const long long *__restrict__ counts, long long *__restrict__ offsets)
{
    __shared__ long long v[256];
    const int t = threadIdx.x;
    v[t] = counts[t];
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const long long a = t >= d ? v[t - d] : 0;
        __syncthreads();
        v[t] += a;
        __syncthreads();
    }
    offsets[t] = v[t] - counts[t];
}

// The index lists of the types; a channel whose type id is not set
// (>= ntypes) is not fitted: x = 0, chi2 = 0, flags = F_BAD_TYPE.
template <typename T>
__global__ __launch_bounds__(256) void k_part_scatter(
    const uint8_t *__restrict__ type_of, int64_t nch, int ntypes, int64_t chunk,
    const int32_t *__restrict__ boff, const long long *__restrict__ offsets,
    int32_t *__restrict__ idx, T *__restrict__ xo, T *__restrict__ chi2o,
    int32_t *__restrict__ flo, int32_t *__restrict__ ito)
{
    __shared__ int run[256];      // the type's next slot in this chunk
    __shared__ int wc[4][256];    // the type's channels in each wavefront of the 256
    const int tid = threadIdx.x, wv = tid >> 6;
    if (tid < ntypes) run[tid] = (int)(offsets[tid] + boff[(size_t)tid * PART_BLOCKS + blockIdx.x]);
#pragma unroll
    for (int w = 0; w < 4; ++w) wc[w][tid] = 0;
    __syncthreads();
    const int64_t c0 = (int64_t)blockIdx.x * chunk, c1 = c0 + chunk < nch ? c0 + chunk : nch;
    for (int64_t base = c0; base < c1; base += 256) {
        const int64_t c = base + tid;
        int t = c < c1 ? (int)type_of[c] : -1;
        if (t >= ntypes) {
#pragma unroll
            for (int i = 0; i < 10; ++i) xo[(int64_t)i * nch + c] = T(0);
            chi2o[c] = T(0);
            flo[c] = F_BAD_TYPE;
            if (ito) ito[c] = 0;
            t = -1;
        }
        int rank = 0;
        wave_types(t, [&](int tt, bool lead, uint64_t mine, int r) {
            if (lead) wc[wv][tt] = __popcll(mine);
            if (t == tt) rank = r;
        });
        __syncthreads();
        if (t >= 0) {
            int off = run[t] + rank;
            for (int w = 0; w < wv; ++w) off += wc[w][t];
            idx[off] = (int32_t)c;
        }
        __syncthreads();
        if (tid < ntypes) {
            run[tid] += wc[0][tid] + wc[1][tid] + wc[2][tid] + wc[3][tid];
#pragma unroll
            for (int w = 0; w < 4; ++w) wc[w][tid] = 0;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// solver-only kernels (inplace_fnnls / fnnls on given A, b)
// ---------------------------------------------------------------------------
template <int M, bool CHOL = false>
__device__ void solve_one(const double *__restrict__ A, const double *__restrict__ b,
                          double *__restrict__ x, int solver, double eps, unsigned max_iter,
                          int32_t *__restrict__ solved)
{
    // A column-major, leading dimension 10: A[col*10 + row]
    double AtA[10][10], Atb[10], xx[10];
#pragma unroll
    for (int i = 0; i < M; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < M; ++k) t = fma(A[i * 10 + k], A[j * 10 + k], t);
            AtA[i][j] = t;
        }
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < M; ++k) t = fma(A[i * 10 + k], b[k], t);
        Atb[i] = t;
        xx[i] = x[i];
    }
    FnnlsStats st;
    if (solver == 2) {
        const bool ok = lh_nnls<double, M>(AtA, Atb, xx, eps, (int)max_iter, st);
        if (solved) *solved = ok ? 1 : 0;
    } else if (solver == 3) {
        if constexpr (CHOL) fnnls_inplace_chol<double, M>(AtA, Atb, xx, eps, (int)max_iter, st);
    } else if (solver == 1) {
        fnnls_mask<double, M>(AtA, Atb, xx, eps, (int)max_iter, st);
    } else {
        fnnls_inplace<double, M>(AtA, Atb, xx, eps, 0.0, (int)max_iter, st);
    }
#pragma unroll
    for (int i = 0; i < M; ++i) x[i] = xx[i];
}

__global__ __launch_bounds__(256) void k_nnls(const double *__restrict__ A,
                                              const double *__restrict__ b, double *__restrict__ x,
                                              int64_t n, int m, int solver, double eps,
                                              unsigned max_iter, int32_t *__restrict__ solved)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double *Ac = A + c * 100;
    const double *bc = b + c * 10;
    double *xc = x + c * 10;
    int32_t *ok = solved ? solved + c : nullptr;
    switch (m) {
    case 1: solve_one<1>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 2: solve_one<2>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 3: solve_one<3>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 4: solve_one<4>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 5: solve_one<5>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 6: solve_one<6>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 7: solve_one<7>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 8: solve_one<8>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 9: solve_one<9>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    default: solve_one<10>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    }
}

// inplace_fnnls_test_cholesky batches: a kernel of their own, so that the
// persistent factor's registers do not weigh on the other solvers
__global__ __launch_bounds__(256) void k_nnls_chol(const double *__restrict__ A,
                                              const double *__restrict__ b, double *__restrict__ x,
                                              int64_t n, int m, int solver, double eps,
                                              unsigned max_iter, int32_t *__restrict__ solved)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double *Ac = A + c * 100;
    const double *bc = b + c * 10;
    double *xc = x + c * 10;
    int32_t *ok = solved ? solved + c : nullptr;
    switch (m) {
    case 1: solve_one<1, true>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 2: solve_one<2, true>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 3: solve_one<3, true>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 4: solve_one<4, true>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 5: solve_one<5, true>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 6: solve_one<6, true>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 7: solve_one<7, true>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 8: solve_one<8, true>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    case 9: solve_one<9, true>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    default: solve_one<10, true>(Ac, bc, xc, solver, eps, max_iter, ok); break;
    }
}

// ---------------------------------------------------------------------------
// linalgebra primitives: one problem per lane, 32 problems per block
// (operation order as linalgebra/interface/*.hpp)
// ---------------------------------------------------------------------------
// The caller's layout is a row-major matrix per problem, problem after
// problem (the reference's raw pointers).  A lane reading its own problem
// straight from there would stride by full^2 elements and no load would
// coalesce, so a block first copies its 64 problems' operands into LDS with
// consecutive threads on consecutive addresses, transposed to
// [element][problem]; each of the first 32 lanes then works on its column
// (32 lanes touch 32 consecutive words: no bank conflicts), and the outputs go
// back the same way.  32 problems per block keep the FP64 tiles (M, L, b, y of
// 10 x 10) at 56 KB of LDS.
constexpr int LA_PROBLEMS = 32;
constexpr int LA_THREADS = 64;

template <typename T>
__device__ __forceinline__ void la_load(T *__restrict__ tile, const T *__restrict__ g, int elems,
                                        int64_t c0, int64_t n)
{
    const int64_t np = (n - c0) < LA_PROBLEMS ? (n - c0) : LA_PROBLEMS;
    const int total = (int)np * elems;
    const T *src = g + c0 * elems;
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
        const int p = k / elems, e = k - p * elems;
        tile[e * LA_PROBLEMS + p] = src[k];
    }
}
template <typename T>
__device__ __forceinline__ void la_store(T *__restrict__ g, const T *__restrict__ tile, int elems,
                                         int64_t c0, int64_t n)
{
    const int64_t np = (n - c0) < LA_PROBLEMS ? (n - c0) : LA_PROBLEMS;
    const int total = (int)np * elems;
    T *dst = g + c0 * elems;
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
        const int p = k / elems, e = k - p * elems;
        dst[k] = tile[e * LA_PROBLEMS + p];
    }
}

// needM/needL/needb/needy: operands the op reads; outM/outL/outy: the ones it writes
template <typename T>
__global__ __launch_bounds__(LA_THREADS)
void k_linalg(int op, T *M, T *L, T *b, T *y, int64_t n, int full, int view, int pi, int pj)
{
    extern __shared__ unsigned char la_raw[];
    const int ff = full * full;
    T *tM = reinterpret_cast<T *>(la_raw);
    T *tL = tM + ff * LA_PROBLEMS;
    T *tb = tL + ff * LA_PROBLEMS;
    T *ty = tb + full * LA_PROBLEMS;
    const int64_t c0 = (int64_t)blockIdx.x * LA_PROBLEMS;
    const bool rM = op == MF_LA_CHOLESKY || op == MF_LA_SWAP || op == MF_LA_RCADDITION ||
                    op == MF_LA_INBETWEEN;
    const bool rL = op != MF_LA_CHOLESKY && op != MF_LA_SWAP;
    const bool rb = op != MF_LA_CHOLESKY && op != MF_LA_SWAP;
    const bool ry = op == MF_LA_RCADDITION || op == MF_LA_INBETWEEN;
    if (rM) la_load(tM, M, ff, c0, n);
    if (rL) la_load(tL, L, ff, c0, n);
    if (op == MF_LA_CHOLESKY)   // the output's entries the op leaves alone come back as they were
        la_load(tL, L, ff, c0, n);
    if (rb) la_load(tb, b, full, c0, n);
    if (ry) la_load(ty, y, full, c0, n);
    if (op == MF_LA_FORWARD || op == MF_LA_BACKWARD) la_load(ty, y, full, c0, n);
    __syncthreads();
    const int lane = threadIdx.x;
    if (lane < LA_PROBLEMS && c0 + lane < n) {
#define AT(P, r, q) P[((r) * full + (q)) * LA_PROBLEMS + lane]
#define V(P, i) P[(i) * LA_PROBLEMS + lane]
        switch (op) {
        case MF_LA_CHOLESKY:  // cholesky.hpp:8-38
            for (int i = 0; i < view; ++i) {
                T sumsq = 0;
                for (int j = 0; j < i; ++j) {
                    T s2 = 0;
                    for (int k = 0; k < j; ++k) s2 += AT(tL, i, k) * AT(tL, j, k);
                    AT(tL, i, j) = (AT(tM, i, j) - s2) / AT(tL, j, j);
                    sumsq += AT(tL, i, j) * AT(tL, i, j);
                }
                AT(tL, i, i) = sqrt(AT(tM, i, i) - sumsq);
            }
            break;
        case MF_LA_FORWARD:  // substitution.hpp:10-26
            V(ty, 0) = V(tb, 0) / AT(tL, 0, 0);
            for (int i = 1; i < view; ++i) {
                T t = V(tb, i);
                for (int j = 0; j < i; ++j) t -= AT(tL, i, j) * V(ty, j);
                V(ty, i) = t / AT(tL, i, i);
            }
            break;
        case MF_LA_BACKWARD:  // substitution.hpp:34-50
            V(ty, view - 1) = V(tb, view - 1) / AT(tL, view - 1, view - 1);
            for (int i = view - 2; i >= 0; --i) {
                T t = V(tb, i);
                for (int j = i + 1; j < view; ++j) t -= AT(tL, j, i) * V(ty, j);
                V(ty, i) = t / AT(tL, i, i);
            }
            break;
        case MF_LA_SWAP: {  // common.hpp:11-42 (j > i)
            const int i = pi, j = pj;
            T t = AT(tM, i, i);
            AT(tM, i, i) = AT(tM, j, j);
            AT(tM, j, j) = t;
            for (int e = 0; e < i; ++e) {
                T tmp = AT(tM, i, e);
                AT(tM, i, e) = AT(tM, j, e);
                AT(tM, j, e) = tmp;
                AT(tM, e, i) = AT(tM, e, j);
                AT(tM, e, j) = tmp;
            }
            for (int e = i + 1; e < view; ++e) {
                T tmp = AT(tM, i, e);
                AT(tM, i, e) = AT(tM, j, e);
                AT(tM, j, e) = tmp;
                AT(tM, e, i) = AT(tM, e, j);
                AT(tM, e, j) = tmp;
            }
            break;
        }
        case MF_LA_RCADDITION: {  // fused_cholesky_substitution.hpp:18-57
            const int row = view - 1;
            T sum = 0, total = V(tb, row);
            for (int col = 0; col < row; ++col) {
                T s2 = 0;
                for (int j = 0; j < col; ++j) s2 += AT(tL, row, j) * AT(tL, col, j);
                const T v = (AT(tM, row, col) - s2) / AT(tL, col, col);
                total -= v * V(ty, col);
                sum += v * v;
                AT(tL, row, col) = v;
            }
            const T d = sqrt(AT(tM, row, row) - sum);
            AT(tL, row, row) = d;
            V(ty, row) = total / d;
            break;
        }
        case MF_LA_INBETWEEN:  // fused_cholesky_substitution.hpp:62-102
            for (int i = pi; i < view; ++i) {
                T sumsq = 0, total = V(tb, i);
                for (int j = 0; j < i; ++j) {
                    T s2 = 0;
                    for (int k = 0; k < j; ++k) s2 += AT(tL, i, k) * AT(tL, j, k);
                    const T v = (AT(tM, i, j) - s2) / AT(tL, j, j);
                    AT(tL, i, j) = v;
                    total -= v * V(ty, j);
                    sumsq += v * v;
                }
                const T d = sqrt(AT(tM, i, i) - sumsq);
                AT(tL, i, i) = d;
                V(ty, i) = total / d;
            }
            break;
        }
#undef V
#undef AT
    }
    __syncthreads();
    if (op == MF_LA_SWAP) la_store(M, tM, ff, c0, n);
    if (op == MF_LA_CHOLESKY || op == MF_LA_RCADDITION || op == MF_LA_INBETWEEN)
        la_store(L, tL, ff, c0, n);
    if (op != MF_LA_CHOLESKY && op != MF_LA_SWAP) la_store(y, ty, full, c0, n);
}

// ---------------------------------------------------------------------------
// context and host entry points
// ---------------------------------------------------------------------------
// Launch shape of the persistent fast-path kernels (mf_autotune chooses it):
// threads per block, resident wavefronts per compute unit for each phase,
// and the number of idle lanes that triggers a refill from the queue.
struct FastTune {
    int block = 256;
    int p1_wpc = 16, p2_wpc = 12;
    int p1_need = 16, p2_need = 16;
    int fit_block = 64;        // per-channel kernels, opts.block == 0
    bool use_persistent = false;
};
static bool tune_valid(const mf_tune &t)
{
    auto blk = [](int b) { return b >= 64 && b <= 256 && b % 64 == 0; };
    return blk(t.block) && blk(t.fit_block) && t.p1_wpc >= 1 && t.p1_wpc <= 32 &&
           t.p2_wpc >= 1 && t.p2_wpc <= 32 && t.p1_refill >= 1 && t.p1_refill <= 64 &&
           t.p2_refill >= 1 && t.p2_refill <= 64 && (t.use_persistent == 0 || t.use_persistent == 1);
}
constexpr int QUEUE_SLOTS = 64;

struct mf_context {
    int device = 0;
    std::vector<mf_chtype> types;
    DevType *d_types = nullptr;
    MaskTab *d_tabs = nullptr;     // one per type, pointed to by DevType::tab
    // partition scratch
    void *d_part = nullptr;  // PART_BYTES: type counts, offsets, per-chunk counts and slots
    int32_t *d_idx = nullptr;
    int64_t idx_cap = 0;
    // warm-start MLP: masters, bfloat16 copies (as floats and as bits),
    // gradient, Adam moments; activations of the largest batch seen
    float *mlp_p = nullptr, *mlp_pq = nullptr, *mlp_g = nullptr, *mlp_m = nullptr,
          *mlp_v = nullptr;
    uint16_t *mlp_bf = nullptr, *mlp_frag = nullptr;
    int mlp_t = 0;
    bool mlp_ready = false;
    mf::mlp::Scratch mlp_w;
    float *pred = nullptr;  // warm-start prediction: x [10][nch], then pm [nch] (uint32)
    int64_t pred_cap = 0;
    double *hand = nullptr;  // live-covariance hand-over: x [10][nch] and chi2 [nch] in FP64,
                             // then the pending list (hand_bytes)
    int64_t hand_cap = 0;
    // live mf_graph objects whose kernels point into the prediction / the
    // partition / the hand-over scratch (that scratch must not move while
    // they exist)
    int graphs_pred = 0, graphs_part = 0, graphs_hand = 0;
    // every live graph of the context: their kernels also point into the
    // type tables, so mf_set_types is refused while any exists, and
    // mf_context_destroy detaches them (a detached graph refuses to launch)
    std::vector<mf_graph *> graphs;
    // the device address of this GPU's scratch of mf_fit_batch_noalloc
    void *noalloc = nullptr;
    // persistent kernels: compute units, queue heads, launch shape
    int cus = 0;
    unsigned *d_queue = nullptr;  // QUEUE_SLOTS heads
    FastTune tune;
    // opt-in kernel timing: events around the MLP and the fit launches
    bool timing = false;
    hipEvent_t ev[2 + 2 * 256] = {};
    int n_fit_ev = 0;
    bool mlp_timed = false;
};

// A captured fit call (mf_graph_capture); ctx is null once its context is gone.
struct mf_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
    mf_context *ctx = nullptr;
    bool pred = false, part = false, hand = false;  // scratch its kernels use
};

static void time_mark(mf_context *ctx, int slot, hipStream_t st)
{
    if (ctx->timing) (void)hipEventRecord(ctx->ev[slot], st);
}

extern "C" mf_status mf_default_opts(mf_opts *o)
{
    if (!o) return fail(MF_ERR_INVALID, "opts is NULL");
    memset(o, 0, sizeof(*o));
    o->precision = MF_PREC_FP64;
    o->solver = MF_SOLVER_INPLACE_FNNLS;
    o->max_iter = 1000;      // multifit_gpu/src/PulseChiSqSNNLS.cu:215
    o->max_minimize = 50;    // multifit_cpu/src/PulseChiSqSNNLS.cpp:153
    o->eps = 1e-11;          // PulseChiSqSNNLS.cpp:242
    o->chi2_tol = 1e-3;      // PulseChiSqSNNLS.cpp:162
    o->use_pulse_cov = 1;
    return MF_OK;
}

// Scratch of mf_fit_batch_noalloc: a file-scope device array, one copy on
// each GPU the library runs on, held until the process ends.  On the host
// the name is only a stand-in; the device address of a GPU's copy comes from
// hipGetSymbolAddress with that GPU current, once per GPU (at the first
// mf_context_create on it), and only that address is passed to kernels.
static __device__ __align__(256) unsigned char scratch[MF_NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[MAX_DEVICES] = {};

static mf_status noalloc_scratch(int device, void **ptr)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= MAX_DEVICES)
        return fail(MF_ERR_INVALID, "device %d: the library keeps scratch for devices 0..%d",
                    device, MAX_DEVICES - 1);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MF_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return MF_OK;
}

extern "C" mf_status mf_context_create(mf_context **ctx, int device)
{
    if (!ctx) return fail(MF_ERR_INVALID, "ctx is NULL");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(MF_ERR_INVALID, "device %d out of range (%d devices)", device, ndev);
    HIP_TRY(hipSetDevice(device));
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    void *na = nullptr;
    const mf_status sr = noalloc_scratch(device, &na);
    if (sr != MF_OK) return sr;
    unsigned *q = nullptr;
    HIP_TRY(hipMalloc(&q, sizeof(unsigned) * QUEUE_SLOTS));
    mf_context *c = new mf_context;
    c->device = device;
    c->cus = cus > 0 ? cus : 1;
    c->d_queue = q;
    c->noalloc = na;
    *ctx = c;
    return MF_OK;
}

extern "C" mf_status mf_get_tune(mf_context *ctx, mf_tune *out)
{
    if (!ctx || !out) return fail(MF_ERR_INVALID, "NULL argument");
    memset(out, 0, sizeof(*out));
    out->block = ctx->tune.block;
    out->p1_wpc = ctx->tune.p1_wpc;
    out->p2_wpc = ctx->tune.p2_wpc;
    out->p1_refill = ctx->tune.p1_need;
    out->p2_refill = ctx->tune.p2_need;
    out->fit_block = ctx->tune.fit_block;
    out->use_persistent = ctx->tune.use_persistent ? 1 : 0;
    return MF_OK;
}

extern "C" mf_status mf_set_tune(mf_context *ctx, const mf_tune *t)
{
    if (!ctx || !t) return fail(MF_ERR_INVALID, "NULL argument");
    if (!tune_valid(*t))
        return fail(MF_ERR_INVALID, "tune out of range (block 64..256 step 64, wpc 1..32, "
                                    "refill 1..64)");
    ctx->tune.block = t->block;
    ctx->tune.p1_wpc = t->p1_wpc;
    ctx->tune.p2_wpc = t->p2_wpc;
    ctx->tune.p1_need = t->p1_refill;
    ctx->tune.p2_need = t->p2_refill;
    ctx->tune.fit_block = t->fit_block;
    ctx->tune.use_persistent = t->use_persistent != 0;
    return MF_OK;
}

extern "C" mf_status mf_autotune(mf_context *ctx, const mf_batch *in, const mf_result *out,
                                 const mf_opts *opts, void *stream, mf_tune *best, float *ms)
{
    if (!ctx || !in || !out || !opts) return fail(MF_ERR_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    struct Cand {
        FastTune t;
        bool pers;
    };
    std::vector<Cand> cands;
    const FastTune base = ctx->tune;
    for (int b : {64, 128, 256}) {
        FastTune t = base;
        t.fit_block = b;
        t.use_persistent = false;
        cands.push_back({t, false});
    }
    if (opts->precision == MF_PREC_FP32)   // the persistent kernels exist for the fast path
        for (int wpc : {8, 16})
            for (int need : {16, 64}) {
                FastTune t = base;
                t.block = 256;
                t.p1_wpc = wpc;
                t.p2_wpc = wpc;
                t.p1_need = need;
                t.p2_need = need;
                t.use_persistent = true;
                cands.push_back({t, true});
            }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const bool timing = ctx->timing;
    ctx->timing = false;
    float best_ms = 3.0e38f;
    FastTune chosen = base;
    mf_status status = MF_OK;
    for (const Cand &c : cands) {
        ctx->tune = c.t;
        mf_opts o = *opts;
        o.block = 0;
        o.persistent = c.pers ? 1 : 0;
        float t_min = 3.0e38f;
        for (int rep = 0; rep < 4 && status == MF_OK; ++rep) {
            if (hipEventRecord(e0, st) != hipSuccess) status = fail(MF_ERR_HIP, "hipEventRecord");
            if (status == MF_OK) status = mf_fit_batch(ctx, in, out, &o, st);
            if (status == MF_OK && hipEventRecord(e1, st) != hipSuccess)
                status = fail(MF_ERR_HIP, "hipEventRecord");
            if (status == MF_OK && hipEventSynchronize(e1) != hipSuccess)
                status = fail(MF_ERR_HIP, "hipEventSynchronize");
            float t = 0.0f;
            if (status == MF_OK && rep > 0 && hipEventElapsedTime(&t, e0, e1) == hipSuccess)
                t_min = std::min(t_min, t);
        }
        if (status != MF_OK) break;
        if (t_min < best_ms) {
            best_ms = t_min;
            chosen = c.t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->timing = timing;
    ctx->tune = status == MF_OK ? chosen : base;
    if (status != MF_OK) return status;
    if (best) (void)mf_get_tune(ctx, best);
    if (ms) *ms = best_ms;
    return MF_OK;
}

extern "C" mf_status mf_context_destroy(mf_context *ctx)
{
    if (!ctx) return MF_OK;
    (void)hipSetDevice(ctx->device);
    for (mf_graph *g : ctx->graphs) g->ctx = nullptr;
    if (ctx->d_types) (void)hipFree(ctx->d_types);
    if (ctx->d_tabs) (void)hipFree(ctx->d_tabs);
    if (ctx->d_queue) (void)hipFree(ctx->d_queue);
    if (ctx->d_part) (void)hipFree(ctx->d_part);
    if (ctx->d_idx) (void)hipFree(ctx->d_idx);
    for (hipEvent_t &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (void *p : {(void *)ctx->mlp_frag, (void *)ctx->mlp_p, (void *)ctx->mlp_pq, (void *)ctx->mlp_g,
                    (void *)ctx->mlp_m, (void *)ctx->mlp_v, (void *)ctx->mlp_bf,
                    (void *)ctx->mlp_w.X, (void *)ctx->mlp_w.m, (void *)ctx->mlp_w.Z1,
                    (void *)ctx->mlp_w.H1, (void *)ctx->mlp_w.Z2, (void *)ctx->mlp_w.H2,
                    (void *)ctx->mlp_w.Z3, (void *)ctx->mlp_w.lossdummy, (void *)ctx->pred,
                    (void *)ctx->hand})
        if (p) (void)hipFree(p);
    delete ctx;
    return MF_OK;
}

extern "C" mf_status mf_set_types(mf_context *ctx, const mf_chtype *types, int32_t ntypes)
{
    if (!ctx || !types) return fail(MF_ERR_INVALID, "ctx or types is NULL");
    if (ntypes <= 0 || ntypes > 256) return fail(MF_ERR_INVALID, "ntypes must be in 1..256");
    if (!ctx->graphs.empty())
        return fail(MF_ERR_INVALID, "%d captured graph(s) of this context read its type tables; "
                                    "destroy them before mf_set_types", (int)ctx->graphs.size());
    std::vector<DevType> h((size_t)ntypes);
    for (int t = 0; t < ntypes; ++t) {
        const char *e = expand_type(types[t], h[(size_t)t]);
        if (e) return fail(MF_ERR_INVALID, "type %d: %s", t, e);
    }
    std::vector<MaskTab> tabs((size_t)ntypes);
    for (int t = 0; t < ntypes; ++t) build_mask_tab(h[(size_t)t], tabs[(size_t)t]);
    HIP_TRY(hipSetDevice(ctx->device));
    if (ctx->d_types) {
        HIP_TRY(hipFree(ctx->d_types));
        ctx->d_types = nullptr;
    }
    if (ctx->d_tabs) {
        HIP_TRY(hipFree(ctx->d_tabs));
        ctx->d_tabs = nullptr;
    }
    HIP_TRY(hipMalloc(&ctx->d_tabs, sizeof(MaskTab) * (size_t)ntypes));
    HIP_TRY(hipMemcpy(ctx->d_tabs, tabs.data(), sizeof(MaskTab) * (size_t)ntypes,
                      hipMemcpyHostToDevice));
    for (int t = 0; t < ntypes; ++t) h[(size_t)t].tab = ctx->d_tabs + t;
    HIP_TRY(hipMalloc(&ctx->d_types, sizeof(DevType) * (size_t)ntypes));
    HIP_TRY(hipMemcpy(ctx->d_types, h.data(), sizeof(DevType) * (size_t)ntypes,
                      hipMemcpyHostToDevice));
    ctx->types.assign(types, types + ntypes);
    return MF_OK;
}

static int64_t seg_threads(const SegArg &a) { return a.nrange ? a.wave0[a.nrange] * 64 : a.list.n; }
static int seg_type0(const SegArg &a) { return a.nrange ? a.rtype[0] : a.type_id; }

// Whether the kernels' layout specialisation D = 2 holds for a type (see
// lay_first): every pulse window starts at or after sample max(0, p - 2).
static bool layout2(const mf_chtype &t)
{
    for (int p = 0; p < t.npulses; ++p)
        if (std::max(0, t.bx[p] + t.first_off) < std::max(0, p - 2)) return false;
    return true;
}
// ... for every type of a launch
static bool seg_layout2(const mf_context *ctx, const SegArg &seg)
{
    if (seg.nrange == 0) return layout2(ctx->types[(size_t)seg.type_id]);
    for (int r = 0; r < seg.nrange; ++r)
        if (!layout2(ctx->types[(size_t)seg.rtype[r]])) return false;
    return true;
}

template <typename T, typename TS, int N>
static mf_status launch_fit(mf_context *ctx, const SegArg &seg, const void *samples,
                            const double *ped, int64_t nch, const mf_result *out,
                            const FitParams &prm, int block, bool errors, hipStream_t st)
{
    const int64_t nt = seg_threads(seg);
    if (nt <= 0) return MF_OK;
    const int64_t grid = (nt + block - 1) / block;
    const size_t lds = sizeof(T) * 55 * (size_t)block;
    const bool l2 = seg_layout2(ctx, seg);
    auto pick = [&](auto e) {
        constexpr bool E = decltype(e)::value;
        return prm.solver == MF_SOLVER_LEGACY ? k_fit_plain<T, TS, N, E, XS_LEGACY, -1>
               : prm.solver == MF_SOLVER_LAWSON_HANSON ? k_fit_plain<T, TS, N, E, XS_LH, -1>
               : prm.solver == MF_SOLVER_INPLACE_FNNLS_CHOL
                   ? (l2 ? k_fit_plain<T, TS, N, E, XS_CHOL, 2> : k_fit_plain<T, TS, N, E, XS_CHOL, -1>)
                   : (l2 ? k_fit_plain<T, TS, N, E, XS_NONE, 2> : k_fit_plain<T, TS, N, E, XS_NONE, -1>);
    };
    auto kern = errors ? pick(std::true_type{}) : pick(std::false_type{});
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(block), lds, st, ctx->d_types, seg,
                       static_cast<const TS *>(samples), ped, nch, static_cast<T *>(out->x),
                       static_cast<T *>(out->chi2), out->flags, out->iters,
                       static_cast<T *>(out->err), prm);
    HIP_TRY(hipGetLastError());
    return MF_OK;
}

// mlp_here: a warm start whose prediction the fit kernel computes itself
// (k_fast_warm); warm then holds no prediction.
template <typename TS, int N, bool LIVE, bool PED>
static mf_status launch_fast(mf_context *ctx, const SegArg &seg, const void *samples,
                             const double *ped, int64_t nch, const mf_result *out,
                             const FitParams &prm, int block, WarmIn warm, bool persistent,
                             unsigned *queue, hipStream_t st, bool mlp_here = false,
                             double *hand = nullptr)
{
    const int64_t nt = seg_threads(seg);
    if (nt <= 0) return MF_OK;
    if (!out->iters) return fail(MF_ERR_INVALID, "the FP32 path needs mf_result.iters");
    if (LIVE && hand) {
        // first iteration in a light kernel, the later ones in k_fast_p3 and,
        // for the channels that need the normal equations, k_fast_p4
        float *xo = static_cast<float *>(out->x), *c2 = static_cast<float *>(out->chi2);
        if (mlp_here) {
            const int64_t g = (nt + FAST_WARM_BLOCK - 1) / FAST_WARM_BLOCK;
            hipLaunchKernelGGL((k_fast_warm<TS, N, PED, true>), dim3((unsigned)g),
                               dim3(FAST_WARM_BLOCK), 0, st, ctx->d_types, seg,
                               static_cast<const TS *>(samples), ped, nch, xo, c2, out->flags,
                               out->iters, ctx->mlp_frag, ctx->mlp_pq, prm, hand);
        } else {
            const int64_t g = (nt + block - 1) / block;
            hipLaunchKernelGGL((k_fast_fused<TS, N, PED, true>), dim3((unsigned)g), dim3(block), 0,
                               st, ctx->d_types, seg, static_cast<const TS *>(samples), ped, nch,
                               xo, c2, out->flags, out->iters, warm, prm, hand);
        }
        HIP_TRY(hipGetLastError());
        const int64_t g = (nt + block - 1) / block;
        const bool l2 = seg_layout2(ctx, seg);
        auto p3 = l2 ? k_fast_p3<TS, N, PED, 2> : k_fast_p3<TS, N, PED, -1>;
        hipLaunchKernelGGL(p3, dim3((unsigned)g), dim3(block), 0, st, ctx->d_types, seg,
                           static_cast<const TS *>(samples), ped, nch, xo, c2, out->flags,
                           out->iters, hand, prm);
        HIP_TRY(hipGetLastError());
        // the pending channels: one grid row per range, a few wavefronts per
        // CU in all
        const int nr = seg.nrange ? seg.nrange : 1;
        const int64_t g4 = std::max<int64_t>(
            1, std::min<int64_t>((nt + FAST_P4_BLOCK - 1) / FAST_P4_BLOCK,
                                 2 * (int64_t)ctx->cus / nr));
        auto p4 = l2 ? k_fast_p4<TS, N, PED, 2> : k_fast_p4<TS, N, PED, -1>;
        hipLaunchKernelGGL(p4, dim3((unsigned)g4, (unsigned)nr), dim3(FAST_P4_BLOCK),
                           sizeof(double) * 55 * FAST_P4_BLOCK, st, ctx->d_types, seg,
                           static_cast<const TS *>(samples), ped, nch, xo, c2, out->flags,
                           out->iters, hand, prm);
        HIP_TRY(hipGetLastError());
        return MF_OK;
    }
    if (mlp_here) {
        if (LIVE || persistent) return fail(MF_ERR_INVALID, "in-kernel warm start: not for this launch");
        const int64_t g = (nt + FAST_WARM_BLOCK - 1) / FAST_WARM_BLOCK;
        hipLaunchKernelGGL((k_fast_warm<TS, N, PED>), dim3((unsigned)g), dim3(FAST_WARM_BLOCK), 0, st,
                           ctx->d_types, seg, static_cast<const TS *>(samples), ped, nch,
                           static_cast<float *>(out->x), static_cast<float *>(out->chi2),
                           out->flags, out->iters, ctx->mlp_frag, ctx->mlp_pq, prm);
        HIP_TRY(hipGetLastError());
        return MF_OK;
    }
    if (persistent && !LIVE && seg.nrange == 0) {
        const FastTune &tu = ctx->tune;
        const int wpb = tu.block / 64;
        const int64_t need_w = (nt + 63) / 64;
        const int64_t g1 = std::max<int64_t>(1, std::min<int64_t>(need_w, (int64_t)ctx->cus * tu.p1_wpc) / wpb);
        const int64_t g2 = std::max<int64_t>(1, std::min<int64_t>(need_w, (int64_t)ctx->cus * tu.p2_wpc) / wpb);
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, queue, 2);
        hipLaunchKernelGGL((k_fast_p1q<TS, N, PED>), dim3((unsigned)g1), dim3(tu.block), 0, st,
                           ctx->d_types, seg.type_id, seg.list, static_cast<const TS *>(samples),
                           ped, nch, static_cast<float *>(out->x), out->flags, out->iters, warm,
                           prm, queue, tu.p1_need);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL((k_fast_p2q<TS, N, PED>), dim3((unsigned)g2), dim3(tu.block), 0, st,
                           ctx->d_types, seg.type_id, seg.list, static_cast<const TS *>(samples),
                           ped, nch, static_cast<float *>(out->x), static_cast<float *>(out->chi2),
                           out->flags, out->iters, warm.pm ? 1 : 0, prm, queue + 1,
                           tu.p2_need);
        HIP_TRY(hipGetLastError());
        return MF_OK;
    }
    const int64_t grid = (nt + block - 1) / block;
    if (!LIVE) {
        hipLaunchKernelGGL((k_fast_fused<TS, N, PED>), dim3((unsigned)grid), dim3(block), 0, st,
                           ctx->d_types, seg, static_cast<const TS *>(samples), ped, nch,
                           static_cast<float *>(out->x), static_cast<float *>(out->chi2),
                           out->flags, out->iters, warm, prm);
        HIP_TRY(hipGetLastError());
        return MF_OK;
    }
    hipLaunchKernelGGL((k_fast_p1<TS, N, PED>), dim3((unsigned)grid), dim3(block), 0, st,
                       ctx->d_types, seg, static_cast<const TS *>(samples), ped, nch,
                       static_cast<float *>(out->x), out->flags, out->iters, warm, prm);
    HIP_TRY(hipGetLastError());
    const size_t lds = LIVE ? sizeof(double) * 55 * (size_t)block : 0;
    auto p2 = LIVE && seg_layout2(ctx, seg) ? k_fast_p2<TS, N, LIVE, PED, 2>
                                            : k_fast_p2<TS, N, LIVE, PED, -1>;
    hipLaunchKernelGGL(p2, dim3((unsigned)grid), dim3(block), lds, st,
                       ctx->d_types, seg, static_cast<const TS *>(samples), ped, nch,
                       static_cast<float *>(out->x), static_cast<float *>(out->chi2), out->flags,
                       out->iters, warm.pm ? 1 : 0, prm);
    HIP_TRY(hipGetLastError());
    return MF_OK;
}

// Launch shape of a type: samples per channel and whether its pulse
// covariance is live (one launch covers types of the same shape).
static bool type_live(const mf_chtype &t, const FitParams &prm)
{
    bool has_pi = false;
    for (int i = 0; i < MF_FULL * MF_FULL && !has_pi; ++i) has_pi = t.fullpulsecov[i] != 0.0;
    return has_pi && prm.use_pi;
}

template <typename TS>
static mf_status dispatch_fast(mf_context *ctx, const SegArg &seg, const mf_batch *in,
                               const mf_result *out, const FitParams &prm, int block,
                               WarmIn warm, bool persistent, unsigned *queue, hipStream_t st,
                               bool mlp_here, double *hand)
{
    const mf_chtype &t = ctx->types[(size_t)seg_type0(seg)];
    const bool live = type_live(t, prm);
    auto go = [&](auto ped) {
        constexpr bool P = decltype(ped)::value;
        if (t.nsamples == 10)
            return live ? launch_fast<TS, 10, true, P>(ctx, seg, in->samples, in->pederr, in->nch,
                                                       out, prm, block, warm, persistent, queue, st,
                                                       mlp_here && hand, hand)
                        : launch_fast<TS, 10, false, P>(ctx, seg, in->samples, in->pederr,
                                                        in->nch, out, prm, block, warm,
                                                        persistent, queue, st, mlp_here);
        return live ? launch_fast<TS, 8, true, P>(ctx, seg, in->samples, in->pederr, in->nch, out,
                                                  prm, block, warm, persistent, queue, st,
                                                  mlp_here && hand, hand)
                    : launch_fast<TS, 8, false, P>(ctx, seg, in->samples, in->pederr, in->nch, out,
                                                   prm, block, warm, persistent, queue, st,
                                                   mlp_here);
    };
    return in->pederr ? go(std::true_type{}) : go(std::false_type{});
}

template <typename T>
static mf_status dispatch_fit(mf_context *ctx, const SegArg &seg, const mf_batch *in,
                              const mf_result *out, const FitParams &prm, int block, bool errors,
                              hipStream_t st)
{
    const int ns = ctx->types[(size_t)seg_type0(seg)].nsamples;
    const bool f64in = in->sample_dtype == MF_DTYPE_F64;
    if (ns == 10)
        return f64in ? launch_fit<T, double, 10>(ctx, seg, in->samples, in->pederr, in->nch, out,
                                                 prm, block, errors, st)
                     : launch_fit<T, float, 10>(ctx, seg, in->samples, in->pederr, in->nch, out,
                                                prm, block, errors, st);
    return f64in ? launch_fit<T, double, 8>(ctx, seg, in->samples, in->pederr, in->nch, out, prm,
                                            block, errors, st)
                 : launch_fit<T, float, 8>(ctx, seg, in->samples, in->pederr, in->nch, out, prm,
                                           block, errors, st);
}

static SegArg seg_single(int type_id, const Seg &list)
{
    SegArg a;
    memset(&a, 0, sizeof(a));
    a.nrange = 0;
    a.type_id = type_id;
    a.list = list;
    return a;
}

// Device memory a fit call uses beyond its inputs and outputs.
struct CallScratch {
    void *part = nullptr;       // PART_BYTES: type-partition counts and slots
    int32_t *idx = nullptr;     // nch entries: the types' index lists
    float *pred = nullptr;      // 11 nch floats: warm-start prediction, x [10][nch] then pm [nch]
    unsigned *queue = nullptr;  // 2 heads: the persistent kernels' work queues
    double *hand = nullptr;     // hand_bytes(nch): live-covariance hand-over (hand_over), pending list
};
// What a call needs of it.
struct ScratchNeed {
    bool part = false, pred = false, queue = false, hand = false;
    int64_t nch = 0;
};
// Offsets of the parts in one contiguous region (256-byte aligned): queue
// heads, partition counts, index lists, prediction.
static size_t align256(size_t v) { return (v + 255) / 256 * 256; }
static size_t scratch_layout(const ScratchNeed &n, size_t off[5])
{
    size_t k = 0;
    off[0] = k;
    if (n.queue) k += 256;
    off[1] = k;
    if (n.part) k += align256(PART_BYTES);
    off[2] = k;
    if (n.part) k += align256(sizeof(int32_t) * (size_t)n.nch);
    off[3] = k;
    if (n.pred) k += align256(sizeof(float) * 11 * (size_t)n.nch);
    off[4] = k;
    if (n.hand) k += align256(hand_bytes(n.nch));
    return k;
}

// Grows the context's own scratch to what the call needs (never under a
// graph that uses it) and returns it.
static mf_status ctx_scratch(mf_context *ctx, const ScratchNeed &n, CallScratch &sc)
{
    if (n.pred && ctx->pred_cap < n.nch) {
        if (ctx->graphs_pred)
            return fail(MF_ERR_INVALID, "the scratch would grow under a captured graph; "
                                        "destroy the context's graphs first");
        if (ctx->pred) HIP_TRY(hipFree(ctx->pred));
        ctx->pred = nullptr;
        ctx->pred_cap = 0;
        HIP_TRY(hipMalloc(&ctx->pred, sizeof(float) * 11 * (size_t)n.nch));
        ctx->pred_cap = n.nch;
    }
    if (n.part) {
        if ((!ctx->d_part || ctx->idx_cap < n.nch) && ctx->graphs_part)
            return fail(MF_ERR_INVALID, "the scratch would grow under a captured graph; "
                                        "destroy the context's graphs first");
        if (!ctx->d_part) HIP_TRY(hipMalloc(&ctx->d_part, PART_BYTES));
        if (ctx->idx_cap < n.nch) {
            if (ctx->d_idx) HIP_TRY(hipFree(ctx->d_idx));
            ctx->d_idx = nullptr;
            ctx->idx_cap = 0;
            HIP_TRY(hipMalloc(&ctx->d_idx, sizeof(int32_t) * (size_t)n.nch));
            ctx->idx_cap = n.nch;
        }
    }
    if (n.hand && ctx->hand_cap < n.nch) {
        if (ctx->graphs_hand)
            return fail(MF_ERR_INVALID, "the scratch would grow under a captured graph; "
                                        "destroy the context's graphs first");
        if (ctx->hand) HIP_TRY(hipFree(ctx->hand));
        ctx->hand = nullptr;
        ctx->hand_cap = 0;
        HIP_TRY(hipMalloc(&ctx->hand, hand_bytes(n.nch)));
        ctx->hand_cap = n.nch;
    }
    sc.part = ctx->d_part;
    sc.idx = ctx->d_idx;
    sc.pred = ctx->pred;
    sc.queue = ctx->d_queue;
    sc.hand = n.hand ? ctx->hand : nullptr;
    return MF_OK;
}

// Validates a call and fills the per-call constants; shared by the entry points.
struct CallPlan {
    FitParams prm;
    int block = 0;
    bool pers = false, mlp_here = false;
    ScratchNeed need;
};
// split: types with a live pulse covariance run the fast path's first
// iteration and the later ones in two kernels (hand_over), with FP64 state
// between them in scratch; without it (the allocation-free entry point,
// whose fixed scratch would not hold 88 B per channel of a large batch) both
// run in k_fast_p2, with the same results.
static mf_status plan_call(mf_context *ctx, const mf_batch *in, const mf_result *out,
                           const mf_opts *opts, CallPlan &pl, bool split = true)
{
    if (!ctx || !in || !out || !opts) return fail(MF_ERR_INVALID, "NULL argument");
    if (ctx->types.empty()) return fail(MF_ERR_INVALID, "no channel types set (mf_set_types)");
    if (in->nch < 0 || in->nch > (int64_t)INT32_MAX)
        return fail(MF_ERR_INVALID, "nch must be in 0..2^31-1");
    if (in->nch == 0) return MF_OK;
    if (!in->samples || !out->x || !out->chi2 || !out->flags)
        return fail(MF_ERR_INVALID, "samples, x, chi2 and flags are required");
    if (in->sample_dtype != MF_DTYPE_F64 && in->sample_dtype != MF_DTYPE_F32)
        return fail(MF_ERR_INVALID, "sample_dtype must be MF_DTYPE_F64 or MF_DTYPE_F32");
    if (in->nranges < 0 || in->nranges > (int32_t)ctx->types.size())
        return fail(MF_ERR_INVALID, "nranges must be in 0..ntypes");
    if (in->nranges > 0) {
        if (in->type_of) return fail(MF_ERR_INVALID, "give type_of or ranges, not both");
        if (!in->ranges) return fail(MF_ERR_INVALID, "ranges is NULL");
        if (in->ranges[0] != 0 || in->ranges[in->nranges] != in->nch)
            return fail(MF_ERR_INVALID, "ranges must start at 0 and end at nch");
        for (int r = 0; r < in->nranges; ++r)
            if (in->ranges[r + 1] < in->ranges[r])
                return fail(MF_ERR_INVALID, "ranges must not decrease");
    }
    if (opts->precision != MF_PREC_FP64 && opts->precision != MF_PREC_FP32)
        return fail(MF_ERR_INVALID, "precision must be MF_PREC_FP64 or MF_PREC_FP32");
    if (opts->max_minimize < 0 || opts->max_iter < 0)
        return fail(MF_ERR_INVALID, "max_minimize and max_iter must be >= 0");
    if (opts->solver != MF_SOLVER_INPLACE_FNNLS && opts->solver != MF_SOLVER_FNNLS &&
        opts->solver != MF_SOLVER_LEGACY && opts->solver != MF_SOLVER_LAWSON_HANSON &&
        opts->solver != MF_SOLVER_INPLACE_FNNLS_CHOL)
        return fail(MF_ERR_INVALID, "unknown solver");
    if (opts->warm_start && opts->precision != MF_PREC_FP32)
        return fail(MF_ERR_UNSUPPORTED, "warm_start needs MF_PREC_FP32");
    if (opts->warm_start && !ctx->mlp_ready)
        return fail(MF_ERR_INVALID, "warm_start needs MLP parameters (mf_mlp_init / mf_mlp_set_params)");
    pl.block = opts->block > 0 ? opts->block : ctx->tune.fit_block;
    if (pl.block % 64 || pl.block > 256)
        return fail(MF_ERR_INVALID, "block must be 64, 128, 192 or 256");
    FitParams &prm = pl.prm;
    prm.eps = opts->eps;
    prm.chi2_tol = opts->chi2_tol;
    prm.max_iter = opts->max_iter;
    prm.max_minimize = opts->max_minimize;
    prm.use_pi = opts->use_pulse_cov;
    prm.solver = opts->solver;
    prm.eps_rel32 = 1e-6f;  // FP32 phase: allowance on w (about 16 ulp of its terms)
    prm.tau32 = 1e-3f;      // FP32 phase: pivot keeping < 0.1% of its diagonal -> FP64
    prm.eps_rel64 = 1e-15;  // FP64 phase: allowance on w (about 4 ulp of its terms)
    if (opts->precision == MF_PREC_FP32 && opts->solver != MF_SOLVER_INPLACE_FNNLS)
        return fail(MF_ERR_UNSUPPORTED, "the FP32 path implements inplace_fnnls only");
    if (opts->compute_errors && opts->precision != MF_PREC_FP64)
        return fail(MF_ERR_UNSUPPORTED, "compute_errors needs MF_PREC_FP64");
    if (opts->compute_errors && !out->err)
        return fail(MF_ERR_INVALID, "compute_errors needs mf_result.err");
    pl.pers = opts->persistent == 1 ||
              (opts->persistent == MF_PERSISTENT_AUTO && ctx->tune.use_persistent);
    // A warm start runs the MLP inside the fit kernel (k_fast_warm) unless
    // the persistent kernels run, or a type's live pulse covariance runs in
    // k_fast_p2 (no split): then the MLP kernel writes the prediction for
    // all channels first.
    bool any_live = false;
    for (const mf_chtype &t : ctx->types) any_live = any_live || type_live(t, prm);
    const bool fast = opts->precision == MF_PREC_FP32;
    pl.mlp_here = opts->warm_start && !pl.pers && (!any_live || split);
    pl.need.nch = in->nch;
    pl.need.hand = fast && any_live && split;
    pl.need.pred = opts->warm_start && !pl.mlp_here;
    pl.need.part = in->type_of != nullptr && in->nranges == 0;
    pl.need.queue = opts->precision == MF_PREC_FP32 && pl.pers;
    return MF_OK;
}

// The fit itself, once the call is planned and its scratch is in place.
static mf_status run_call(mf_context *ctx, const mf_batch *in, const mf_result *out,
                          const mf_opts *opts, const CallPlan &pl, const CallScratch &sc,
                          hipStream_t st)
{
    const FitParams &prm = pl.prm;
    const int block = pl.block;
    const bool pers = pl.pers, mlp_here = pl.mlp_here;
    const int ntypes = (int)ctx->types.size();
    WarmIn warm{nullptr, nullptr};
    if (pl.need.pred) {
        float *px = sc.pred;
        uint32_t *ppm = reinterpret_cast<uint32_t *>(sc.pred + 10 * (size_t)in->nch);
        // the prediction sees each channel's own sample count
        mf::mlp::NsOfType nst;
        memset(&nst, 0, sizeof(nst));
        for (int t = 0; t < ntypes; ++t)
            if (ctx->types[(size_t)t].nsamples <= 8) nst.is8[t >> 5] |= 1u << (t & 31);
        auto warm_run = [&](int ns, int64_t c0, int64_t n, const uint8_t *tof) {
            if (in->sample_dtype == MF_DTYPE_F64)
                mf::mlp::launch_warm<double>(in->samples, in->nch, ns, c0, n, tof, nst,
                                             ctx->mlp_frag, ctx->mlp_pq, px, ppm, st);
            else
                mf::mlp::launch_warm<float>(in->samples, in->nch, ns, c0, n, tof, nst,
                                            ctx->mlp_frag, ctx->mlp_pq, px, ppm, st);
        };
        time_mark(ctx, 0, st);
        ctx->mlp_timed = ctx->timing;
        if (in->type_of && in->nranges == 0) {
            warm_run(10, 0, in->nch, in->type_of);
        } else if (in->nranges > 0) {
            // runs of ranges with the same sample count
            int r = 0;
            while (r < in->nranges) {
                const int ns = nst.samples_of((unsigned)r);
                int e = r + 1;
                while (e < in->nranges && nst.samples_of((unsigned)e) == ns) ++e;
                warm_run(ns, in->ranges[r], in->ranges[e] - in->ranges[r], nullptr);
                r = e;
            }
        } else {
            warm_run(nst.samples_of(0), 0, in->nch, nullptr);
        }
        HIP_TRY(hipGetLastError());
        time_mark(ctx, 1, st);
        warm = WarmIn{px, ppm};
    } else {
        ctx->mlp_timed = false;
    }
    auto go_fit = [&](const SegArg &seg) {
        if (opts->precision == MF_PREC_FP64)
            return dispatch_fit<double>(ctx, seg, in, out, prm, block, opts->compute_errors != 0,
                                        st);
        return in->sample_dtype == MF_DTYPE_F64
                   ? dispatch_fast<double>(ctx, seg, in, out, prm, block, warm, pers, sc.queue, st,
                                           mlp_here, sc.hand)
                   : dispatch_fast<float>(ctx, seg, in, out, prm, block, warm, pers, sc.queue, st,
                                          mlp_here, sc.hand);
    };
    ctx->n_fit_ev = 0;
    auto go = [&](const SegArg &seg) {
        const int k = ctx->n_fit_ev < 256 ? ctx->n_fit_ev++ : 255;
        time_mark(ctx, 2 + 2 * k, st);
        const mf_status r = go_fit(seg);
        time_mark(ctx, 3 + 2 * k, st);
        return r;
    };
    if (in->nranges > 0) {
        // type ranges: one launch per shape (samples, live pulse covariance),
        // up to MAX_RANGES ranges each, every wavefront within one range
        // (fixed arrays: no host allocation on this path)
        int pending[256], rest[256], np = 0;
        for (int r = 0; r < in->nranges; ++r)
            if (in->ranges[r + 1] > in->ranges[r]) pending[np++] = r;
        while (np > 0) {
            const mf_chtype &t0 = ctx->types[(size_t)pending[0]];
            const bool live0 = type_live(t0, prm);
            SegArg a;
            memset(&a, 0, sizeof(a));
            int nr = 0;
            for (int k = 0; k < np; ++k) {
                const int r = pending[k];
                const mf_chtype &t = ctx->types[(size_t)r];
                if (a.nrange < MAX_RANGES && t.nsamples == t0.nsamples && type_live(t, prm) == live0) {
                    a.rtype[a.nrange] = r;
                    a.rfirst[a.nrange] = in->ranges[r];
                    a.rcount[a.nrange] = in->ranges[r + 1] - in->ranges[r];
                    a.wave0[a.nrange + 1] = a.wave0[a.nrange] + (a.rcount[a.nrange] + 63) / 64;
                    a.nrange++;
                } else {
                    rest[nr++] = r;
                }
            }
            const mf_status r = go(a);
            if (r != MF_OK) return r;
            for (int k = 0; k < nr; ++k) pending[k] = rest[k];
            np = nr;
        }
        return MF_OK;
    }
    if (!in->type_of) return go(seg_single(0, Seg{nullptr, nullptr, nullptr, in->nch}));
    // mixed batch: partition channel indices by type on the device; each
    // type's launch reads its count and offset from device memory
    int64_t *counts = static_cast<int64_t *>(sc.part);
    int64_t *offsets = counts + 512;
    int32_t *bcount = reinterpret_cast<int32_t *>(counts + 768);
    int32_t *boff = bcount + 256 * PART_BLOCKS;
    // chunks: multiples of 256 channels, at most PART_BLOCKS of them
    const int64_t nblk0 = std::min<int64_t>(PART_BLOCKS, (in->nch + 255) / 256);
    const int64_t chunk = ((in->nch + nblk0 - 1) / nblk0 + 255) / 256 * 256;
    const int nblk = (int)((in->nch + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_zero_u64, dim3(3), dim3(256), 0, st,
                       reinterpret_cast<unsigned long long *>(counts), 768);
    hipLaunchKernelGGL(k_part_count, dim3(nblk), dim3(256), 0, st, in->type_of, in->nch, ntypes,
                       chunk, bcount);
    hipLaunchKernelGGL(k_part_scan, dim3(ntypes), dim3(256), 0, st, nblk, bcount, boff,
                       reinterpret_cast<long long *>(counts));
    hipLaunchKernelGGL(k_type_scan, dim3(1), dim3(256), 0, st,
                       reinterpret_cast<const long long *>(counts),
                       reinterpret_cast<long long *>(offsets));
    if (opts->precision == MF_PREC_FP64)
        hipLaunchKernelGGL(k_part_scatter<double>, dim3(nblk), dim3(256), 0, st, in->type_of,
                           in->nch, ntypes, chunk, boff,
                           reinterpret_cast<const long long *>(offsets), sc.idx,
                           static_cast<double *>(out->x), static_cast<double *>(out->chi2),
                           out->flags, out->iters);
    else
        hipLaunchKernelGGL(k_part_scatter<float>, dim3(nblk), dim3(256), 0, st, in->type_of,
                           in->nch, ntypes, chunk, boff,
                           reinterpret_cast<const long long *>(offsets), sc.idx,
                           static_cast<float *>(out->x), static_cast<float *>(out->chi2),
                           out->flags, out->iters);
    HIP_TRY(hipGetLastError());
    for (int t = 0; t < ntypes; ++t) {
        mf_status r = go(seg_single(t, Seg{sc.idx, counts + t, offsets + t, in->nch}));
        if (r != MF_OK) return r;
    }
    return MF_OK;
}

extern "C" mf_status mf_fit_batch(mf_context *ctx, const mf_batch *in, const mf_result *out,
                                  const mf_opts *opts, void *stream)
{
    CallPlan pl;
    mf_status r = plan_call(ctx, in, out, opts, pl);
    if (r != MF_OK || in->nch == 0) return r;
    HIP_TRY(hipSetDevice(ctx->device));
    CallScratch sc;
    r = ctx_scratch(ctx, pl.need, sc);
    if (r != MF_OK) return r;
    return run_call(ctx, in, out, opts, pl, sc, static_cast<hipStream_t>(stream));
}

extern "C" mf_status mf_fit_batch_noalloc(mf_context *ctx, const mf_batch *in,
                                          const mf_result *out, const mf_opts *opts, void *stream)
{
    CallPlan pl;
    mf_status r = plan_call(ctx, in, out, opts, pl, false);
    if (r != MF_OK || in->nch == 0) return r;
    size_t off[5];
    const size_t need = scratch_layout(pl.need, off);
    if (need > (size_t)MF_NOALLOC_SCRATCH_BYTES)
        return fail(MF_ERR_INVALID,
                    "mf_fit_batch_noalloc: this call needs %zu bytes of scratch (%lld channels%s%s), "
                    "more than its %zu; use mf_fit_batch or fewer channels per call",
                    need, (long long)in->nch, pl.need.part ? ", type partition" : "",
                    pl.need.pred ? ", warm-start prediction" : "",
                    (size_t)MF_NOALLOC_SCRATCH_BYTES);
    if (!ctx->noalloc) return fail(MF_ERR_INVALID, "the context has no scratch address");
    HIP_TRY(hipSetDevice(ctx->device));
    char *base = static_cast<char *>(ctx->noalloc);
    CallScratch sc;
    sc.queue = pl.need.queue ? reinterpret_cast<unsigned *>(base + off[0]) : nullptr;
    sc.part = pl.need.part ? static_cast<void *>(base + off[1]) : nullptr;
    sc.idx = pl.need.part ? reinterpret_cast<int32_t *>(base + off[2]) : nullptr;
    sc.pred = pl.need.pred ? reinterpret_cast<float *>(base + off[3]) : nullptr;
    return run_call(ctx, in, out, opts, pl, sc, static_cast<hipStream_t>(stream));
}

// ---------------------------------------------------------------------------
// graph capture of a fit call
// ---------------------------------------------------------------------------

extern "C" mf_status mf_graph_capture(mf_context *ctx, const mf_batch *in, const mf_result *out,
                                      const mf_opts *opts, void *stream, mf_graph **graph)
{
    if (!ctx || !graph) return fail(MF_ERR_INVALID, "NULL argument");
    *graph = nullptr;
    // eager run: checks the arguments and grows the scratch, so that nothing
    // is allocated while capturing
    mf_status r = mf_fit_batch(ctx, in, out, opts, stream);
    if (r != MF_OK) return r;
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream), own = nullptr;
    if (!st) {  // the null stream cannot be captured
        HIP_TRY(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
        HIP_TRY(hipDeviceSynchronize());
        st = own;
    }
    const bool timing = ctx->timing;
    ctx->timing = false;  // no event records inside the graph
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        r = mf_fit_batch(ctx, in, out, opts, st);
        e = hipStreamEndCapture(st, &g);
    }
    ctx->timing = timing;
    if (own) (void)hipStreamDestroy(own);
    if (e != hipSuccess || r != MF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r != MF_OK ? r : fail(MF_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(MF_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    mf_graph *h = new mf_graph;
    h->graph = g;
    h->exec = x;
    h->device = ctx->device;
    h->ctx = ctx;
    CallPlan pl;
    (void)plan_call(ctx, in, out, opts, pl);   // as the captured call: checked by its eager run
    h->pred = opts->warm_start != 0;
    h->part = in->type_of != nullptr && in->nranges == 0;
    h->hand = pl.need.hand;
    ctx->graphs_pred += h->pred;
    ctx->graphs_part += h->part;
    ctx->graphs_hand += h->hand;
    ctx->graphs.push_back(h);
    *graph = h;
    return MF_OK;
}

extern "C" mf_status mf_graph_launch(mf_graph *graph, void *stream)
{
    if (!graph) return fail(MF_ERR_INVALID, "graph is NULL");
    if (!graph->ctx)
        return fail(MF_ERR_INVALID, "the graph's context was destroyed: its kernels would read "
                                    "freed tables");
    HIP_TRY(hipSetDevice(graph->device));
    HIP_TRY(hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream)));
    return MF_OK;
}

extern "C" mf_status mf_graph_destroy(mf_graph *graph)
{
    if (!graph) return MF_OK;
    (void)hipSetDevice(graph->device);
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    if (graph->ctx) {
        mf_context *c = graph->ctx;
        c->graphs_pred -= graph->pred;
        c->graphs_part -= graph->part;
        c->graphs_hand -= graph->hand;
        for (size_t k = 0; k < c->graphs.size(); ++k)
            if (c->graphs[k] == graph) {
                c->graphs.erase(c->graphs.begin() + (long)k);
                break;
            }
    }
    delete graph;
    return MF_OK;
}

// ---------------------------------------------------------------------------
// warm-start MLP entry points
// ---------------------------------------------------------------------------
static mf_status mlp_alloc(mf_context *ctx)
{
    if (ctx->mlp_p) return MF_OK;
    const size_t n = sizeof(float) * mf::mlp::NPARAM;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMalloc(&ctx->mlp_p, n));
    HIP_TRY(hipMalloc(&ctx->mlp_pq, n));
    HIP_TRY(hipMalloc(&ctx->mlp_g, n));
    HIP_TRY(hipMalloc(&ctx->mlp_m, n));
    HIP_TRY(hipMalloc(&ctx->mlp_v, n));
    HIP_TRY(hipMalloc(&ctx->mlp_bf, sizeof(uint16_t) * mf::mlp::NPARAM));
    HIP_TRY(hipMalloc(&ctx->mlp_frag, sizeof(uint16_t) * mf::mlp::FRAG_ELEMS));
    return MF_OK;
}

static mf_status mlp_upload(mf_context *ctx, const float *params)
{
    mf_status s = mlp_alloc(ctx);
    if (s != MF_OK) return s;
    const size_t n = sizeof(float) * mf::mlp::NPARAM;
    HIP_TRY(hipMemcpy(ctx->mlp_p, params, n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(ctx->mlp_m, 0, n));
    HIP_TRY(hipMemset(ctx->mlp_v, 0, n));
    HIP_TRY(mf::mlp::quantize(ctx->mlp_p, ctx->mlp_pq, ctx->mlp_bf, nullptr));
    HIP_TRY(mf::mlp::pack(ctx->mlp_bf, ctx->mlp_frag, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    ctx->mlp_t = 0;
    ctx->mlp_ready = true;
    return MF_OK;
}

extern "C" mf_status mf_mlp_init(mf_context *ctx, uint64_t seed)
{
    if (!ctx) return fail(MF_ERR_INVALID, "ctx is NULL");
    // Xavier-uniform weights from a 64-bit LCG (Knuth MMIX constants), zero biases
    std::vector<float> p((size_t)mf::mlp::NPARAM, 0.0f);
    uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto uni = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    const int shapes[3][2] = {{10, 64}, {64, 64}, {64, 20}};
    size_t k = 0;
    for (auto &sh : shapes) {
        const double lim = sqrt(6.0 / (sh[0] + sh[1]));
        for (int i = 0; i < sh[0] * sh[1]; ++i) p[k++] = (float)((2.0 * uni() - 1.0) * lim);
        k += (size_t)sh[1];  // bias stays 0
    }
    return mlp_upload(ctx, p.data());
}

extern "C" mf_status mf_mlp_set_params(mf_context *ctx, const float *params, int32_t n)
{
    if (!ctx || !params) return fail(MF_ERR_INVALID, "NULL argument");
    if (n != mf::mlp::NPARAM) return fail(MF_ERR_INVALID, "expected %d parameters", mf::mlp::NPARAM);
    return mlp_upload(ctx, params);
}

extern "C" mf_status mf_mlp_get_params(mf_context *ctx, float *params, int32_t n)
{
    if (!ctx || !params) return fail(MF_ERR_INVALID, "NULL argument");
    if (n != mf::mlp::NPARAM) return fail(MF_ERR_INVALID, "expected %d parameters", mf::mlp::NPARAM);
    if (!ctx->mlp_ready) return fail(MF_ERR_INVALID, "MLP parameters not set");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(params, ctx->mlp_p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
    return MF_OK;
}

static mf_status mlp_scratch(mf_context *ctx, int64_t B)
{
    mf::mlp::Scratch &w = ctx->mlp_w;
    if (w.cap >= B) return MF_OK;
    for (float **p : {&w.X, &w.m, &w.Z1, &w.H1, &w.Z2, &w.H2, &w.Z3, &w.lossdummy})
        if (*p) {
            HIP_TRY(hipFree(*p));
            *p = nullptr;
        }
    const size_t b = (size_t)B;
    HIP_TRY(hipMalloc(&w.X, sizeof(float) * 16 * b));
    HIP_TRY(hipMalloc(&w.m, sizeof(float) * b));
    HIP_TRY(hipMalloc(&w.Z1, sizeof(float) * 64 * b));
    HIP_TRY(hipMalloc(&w.H1, sizeof(float) * 64 * b));
    HIP_TRY(hipMalloc(&w.Z2, sizeof(float) * 64 * b));
    HIP_TRY(hipMalloc(&w.H2, sizeof(float) * 64 * b));
    HIP_TRY(hipMalloc(&w.Z3, sizeof(float) * 20 * b));
    HIP_TRY(hipMalloc(&w.lossdummy, sizeof(float)));
    w.cap = B;
    return MF_OK;
}

extern "C" mf_status mf_mlp_grad(mf_context *ctx, const float *samples, int64_t ld,
                                 const float *xref, int64_t ldx, int64_t B, float *grad,
                                 float *loss, void *stream)
{
    if (!ctx || !samples || !xref || !grad) return fail(MF_ERR_INVALID, "NULL argument");
    if (B <= 0 || B > (1 << 22) || ld < B || ldx < B)
        return fail(MF_ERR_INVALID, "need 0 < B <= 4194304 and ld, ldx >= B");
    if (!ctx->mlp_ready) return fail(MF_ERR_INVALID, "MLP parameters not set");
    HIP_TRY(hipSetDevice(ctx->device));
    mf_status s = mlp_scratch(ctx, B);
    if (s != MF_OK) return s;
    HIP_TRY(mf::mlp::grad_step(ctx->mlp_w, samples, ld, xref, ldx, B, ctx->mlp_pq, grad, loss,
                               static_cast<hipStream_t>(stream)));
    return MF_OK;
}

extern "C" mf_status mf_mlp_train_step(mf_context *ctx, const float *samples, int64_t ld,
                                       const float *xref, int64_t ldx, int64_t B, float lr,
                                       float *loss, void *stream)
{
    if (!ctx) return fail(MF_ERR_INVALID, "ctx is NULL");
    mf_status s = mf_mlp_grad(ctx, samples, ld, xref, ldx, B, ctx ? ctx->mlp_g : nullptr, loss,
                              stream);
    if (s != MF_OK) return s;
    ctx->mlp_t += 1;
    HIP_TRY(mf::mlp::adam_step(ctx->mlp_p, ctx->mlp_g, ctx->mlp_m, ctx->mlp_v, ctx->mlp_t, lr,
                               ctx->mlp_pq, ctx->mlp_bf, static_cast<hipStream_t>(stream)));
    HIP_TRY(mf::mlp::pack(ctx->mlp_bf, ctx->mlp_frag, static_cast<hipStream_t>(stream)));
    return MF_OK;
}

extern "C" mf_status mf_mlp_predict(mf_context *ctx, const mf_batch *in, float *out,
                                    void *stream)
{
    if (!ctx || !in || !out) return fail(MF_ERR_INVALID, "NULL argument");
    if (!ctx->mlp_ready) return fail(MF_ERR_INVALID, "MLP parameters not set");
    if (in->nch <= 0) return MF_OK;
    if (!in->samples) return fail(MF_ERR_INVALID, "samples is NULL");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (in->sample_dtype == MF_DTYPE_F64)
        mf::mlp::launch_infer<double>(in->samples, in->nch, 10, nullptr, in->nch, ctx->mlp_frag,
                                      ctx->mlp_pq, out, in->nch, st);
    else
        mf::mlp::launch_infer<float>(in->samples, in->nch, 10, nullptr, in->nch, ctx->mlp_frag,
                                     ctx->mlp_pq, out, in->nch, st);
    HIP_TRY(hipGetLastError());
    return MF_OK;
}

extern "C" mf_status mf_set_timing(mf_context *ctx, int32_t enable)
{
    if (!ctx) return fail(MF_ERR_INVALID, "ctx is NULL");
    HIP_TRY(hipSetDevice(ctx->device));
    if (enable && !ctx->ev[0])
        for (hipEvent_t &e : ctx->ev) HIP_TRY(hipEventCreate(&e));
    ctx->timing = enable != 0;
    return MF_OK;
}

extern "C" mf_status mf_get_timings(mf_context *ctx, mf_timings *out)
{
    if (!ctx || !out) return fail(MF_ERR_INVALID, "NULL argument");
    memset(out, 0, sizeof(*out));
    if (!ctx->timing) return fail(MF_ERR_INVALID, "timing is not enabled (mf_set_timing)");
    HIP_TRY(hipSetDevice(ctx->device));
    if (ctx->mlp_timed) {
        HIP_TRY(hipEventSynchronize(ctx->ev[1]));
        HIP_TRY(hipEventElapsedTime(&out->mlp_ms, ctx->ev[0], ctx->ev[1]));
    }
    for (int k = 0; k < ctx->n_fit_ev; ++k) {
        float ms = 0.0f;
        HIP_TRY(hipEventSynchronize(ctx->ev[3 + 2 * k]));
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[2 + 2 * k], ctx->ev[3 + 2 * k]));
        out->fit_ms += ms;
    }
    out->fit_launches = ctx->n_fit_ev;
    return MF_OK;
}

static mf_status nnls_batch(const double *A, const double *b, double *x, int64_t n, int32_t m,
                            double eps, uint32_t max_iter, int solver, void *stream,
                            int32_t *solved = nullptr)
{
    if (n < 0) return fail(MF_ERR_INVALID, "n < 0");
    if (n == 0) return MF_OK;
    if (!A || !b || !x) return fail(MF_ERR_INVALID, "NULL argument");
    if (m < 1 || m > 10) return fail(MF_ERR_INVALID, "m must be in 1..10");
    const int block = 64;
    hipLaunchKernelGGL(solver == 3 ? k_nnls_chol : k_nnls,
                       dim3((unsigned)((n + block - 1) / block)), dim3(block), 0,
                       static_cast<hipStream_t>(stream), A, b, x, n, m, solver, eps, max_iter,
                       solved);
    HIP_TRY(hipGetLastError());
    return MF_OK;
}

extern "C" mf_status mf_inplace_fnnls_batch(const double *A, const double *b, double *x,
                                            int64_t n, int32_t m, double eps, uint32_t max_iter,
                                            void *stream)
{
    return nnls_batch(A, b, x, n, m, eps, max_iter, 0, stream);
}

extern "C" mf_status mf_fnnls_batch(const double *A, const double *b, double *x, int64_t n,
                                    int32_t m, double eps, uint32_t max_iter, void *stream)
{
    return nnls_batch(A, b, x, n, m, eps, max_iter, 1, stream);
}

extern "C" mf_status mf_inplace_fnnls_chol_batch(const double *A, const double *b, double *x,
                                                 int64_t n, int32_t m, double eps,
                                                 uint32_t max_iter, void *stream)
{
    return nnls_batch(A, b, x, n, m, eps, max_iter, 3, stream);
}

extern "C" mf_status mf_lh_nnls_batch(const double *A, const double *b, double *x, int64_t n,
                                      int32_t m, double eps, int32_t max_iter, int32_t *solved,
                                      void *stream)
{
    return nnls_batch(A, b, x, n, m, eps, (uint32_t)(max_iter > 0 ? max_iter : 0), 2, stream,
                      solved);
}

extern "C" mf_status mf_linalg_batch(int32_t op, int32_t dtype, void *M, void *L, void *b, void *y,
                                     int64_t n, int32_t full, int32_t view, int32_t i, int32_t j,
                                     void *stream)
{
    if (n < 0) return fail(MF_ERR_INVALID, "n < 0");
    if (n == 0) return MF_OK;
    if (full < 1 || full > 10 || view < 1 || view > full)
        return fail(MF_ERR_INVALID, "need 1 <= view <= full <= 10");
    if (op < MF_LA_CHOLESKY || op > MF_LA_INBETWEEN) return fail(MF_ERR_INVALID, "unknown op");
    if (op == MF_LA_SWAP && !(0 <= i && i < j && j < view))
        return fail(MF_ERR_INVALID, "swap needs 0 <= i < j < view");
    if (op == MF_LA_INBETWEEN && !(0 <= i && i < view))
        return fail(MF_ERR_INVALID, "position must be in 0..view-1");
    const bool needM = op == MF_LA_CHOLESKY || op == MF_LA_SWAP || op == MF_LA_RCADDITION ||
                       op == MF_LA_INBETWEEN;
    const bool needL = op != MF_LA_SWAP;
    const bool needv = op != MF_LA_CHOLESKY && op != MF_LA_SWAP;
    if ((needM && !M) || (needL && !L) || (needv && (!b || !y)))
        return fail(MF_ERR_INVALID, "missing operand for op %d", op);
    const dim3 grid((unsigned)((n + LA_PROBLEMS - 1) / LA_PROBLEMS));
    const dim3 block(LA_THREADS);
    hipStream_t st = static_cast<hipStream_t>(stream);
    // LDS tiles: M and L (full^2 each) and b, y (full each) for 32 problems
    const size_t elems = (size_t)(2 * full * full + 2 * full) * LA_PROBLEMS;
    if (dtype == MF_DTYPE_F64)
        hipLaunchKernelGGL(k_linalg<double>, grid, block, elems * sizeof(double), st, op,
                           static_cast<double *>(M), static_cast<double *>(L),
                           static_cast<double *>(b), static_cast<double *>(y), n, full, view, i, j);
    else if (dtype == MF_DTYPE_F32)
        hipLaunchKernelGGL(k_linalg<float>, grid, block, elems * sizeof(float), st, op,
                           static_cast<float *>(M),
                           static_cast<float *>(L), static_cast<float *>(b),
                           static_cast<float *>(y), n, full, view, i, j);
    else
        return fail(MF_ERR_INVALID, "dtype must be MF_DTYPE_F64 or MF_DTYPE_F32");
    HIP_TRY(hipGetLastError());
    return MF_OK;
}
