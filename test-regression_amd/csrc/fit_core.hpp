// fit_core.hpp -- per-channel multifit, one channel per lane.
//
// Everything here is written for registers: matrix sizes are template
// constants, every loop is fully unrolled, every array index is a
// compile-time constant.  Data-dependent choices (which pulses are in the
// passive set, which index enters or leaves) become predicates and selects,
// never dynamic register indexing, so a wavefront runs the linear algebra of
// 64 channels in lock step with no scratch traffic.
//
// The same header compiles for gfx950 (hipcc) and for the host (g++), so that
// the kernel's arithmetic can be checked against the oracle on a machine
// without a GPU.  The product library only ever runs the device build.
//
// Reference semantics restated (paths relative to the reference root):
//   fit_channel ........... PulseChiSqSNNLS::DoFit/Minimize/updateCov/NNLS/
//                           ComputeChiSq, multifit_cpu/src/PulseChiSqSNNLS.cpp:21-71,149-200,219-274
//   fnnls_inplace ......... inplace_fnnls, nnls/src/inplace_fnnls.cu:72-233
//   fnnls_mask ............ fnnls, nnls/src/fnnls.cu:28-173
//
// How the reference's physical permutation is handled.  inplace_fnnls swaps
// rows/columns of AtA so that the passive set occupies the leading block, and
// solves that block with a fresh LLT.  Here AtA stays in original index order
// and the passive block is solved by a *masked* Cholesky: rows/columns of
// free indices are replaced by the identity on the fly, which leaves exactly
// the Cholesky factor of AtA[P,P] (in ascending index order) embedded in a
// full-size factor.  The reference's permutation is still tracked, as 4-bit
// fields packed in a 64-bit word, because it fixes the tie-breaking of its
// argmax over w (first maximum in permuted order, :144) and of its argmin
// over the step ratios (first minimum in passive order, :192-199).  The
// solution of the passive system is the same; only the rounding of the solve
// differs from an LLT in permuted order.
#pragma once

#include <stdint.h>

#include <type_traits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MF_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define MF_HD inline
#endif

namespace mf {

// ---------------------------------------------------------------------------
// scalar helpers
// ---------------------------------------------------------------------------
// Inside a function body: no contraction of a * b + c into an FMA there, so
// that every kernel that inlines the function rounds it the same way (device
// code contracts by default; the host builds of the tests do not).
#if defined(__clang__)
#define MF_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define MF_NO_CONTRACT
#endif

MF_HD double mfma(double a, double b, double c) { return fma(a, b, c); }
MF_HD float mfma(float a, float b, float c) { return fmaf(a, b, c); }
MF_HD double mabs(double a) { return fabs(a); }
MF_HD float mabs(float a) { return fabsf(a); }

// 1/sqrt(d).  FP64 on the device: the hardware estimate (v_rsq_f64) and two
// Newton-Raphson steps r <- r (3 - d r^2) / 2, each of which squares the
// relative error of the estimate, to full FP64 accuracy (about 1 ulp) at a
// fraction of the cost of a correctly rounded sqrt and divide.  FP32 on the
// device: the hardware reciprocal square root (1 ulp).  Host builds: 1/sqrt.
MF_HD double mrsqrt(double d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rsq(d);
    const double h = 0.5 * d;
    r = r * fma(-h * r, r, 1.5);
    r = r * fma(-h * r, r, 1.5);
    return r;
#else
    return 1.0 / sqrt(d);
#endif
}
MF_HD float mrsqrt(float d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(d);
#else
    return 1.0f / sqrtf(d);
#endif
}
// x / y: FP64 correctly rounded; FP32 on the device via the hardware
// reciprocal (1 ulp), which is all the step-length ratio needs.
MF_HD double mdiv(double x, double y) { return x / y; }
MF_HD float mdiv(float x, float y)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return x * __builtin_amdgcn_rcpf(y);
#else
    return x / y;
#endif
}

template <typename T> struct Lim;
template <> struct Lim<double> {
    MF_HD static double max() { return 1.7976931348623157e308; }
    MF_HD static double eps() { return 2.220446049250313e-16; }
};
template <> struct Lim<float> {
    MF_HD static float max() { return 3.40282347e38f; }
    MF_HD static float eps() { return 1.1920929e-7f; }
};

// Opaque copy of a pointer: stops the compiler from hoisting loads of the
// wave-uniform type tables (hundreds of values) out of the Minimize loop and
// keeping them all live in scalar registers.
// The copy is made in the constant address space, so that the loads through
// it stay scalar (s_load) rather than per-lane flat loads.
template <typename P> MF_HD const P *opaque(const P *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    auto q = (const __attribute__((address_space(4))) P *)p;
    asm volatile("" : "+s"(q));
    return (const P *)q;
#else
    return p;
#endif
}

// 4-bit fields packed in a 64-bit word (permutation and its inverse)
MF_HD int nib(uint64_t w, int k) { return (int)((w >> (4 * k)) & 15u); }
MF_HD uint64_t setnib(uint64_t w, int k, int v)
{
    const int sh = 4 * k;
    return (w & ~((uint64_t)15u << sh)) | ((uint64_t)(unsigned)v << sh);
}
MF_HD uint64_t ident_nibbles()
{
    uint64_t w = 0;
    for (int k = 0; k < 16; ++k) w |= (uint64_t)k << (4 * k);
    return w;
}

// ---------------------------------------------------------------------------
// per-type constants, already expanded to the form the kernels read
// ---------------------------------------------------------------------------
template <typename T> struct TypeMats {
    T P[10][10];       // pulse matrix, P[sample][pulse]
    T C0[10][10];      // pederr^2 * noisecor (lower triangle used)
    T Pi[10][10][10];  // Pi[p][i][j] = fullpulsecov[i+o_p][j+o_p] if i,j >= f_p else 0
    // The whitened problem at x = 0, where the covariance is C0 for every
    // channel of the type (computed once on the host in FP64):
    T Li0[10][10];     // L0^{-1}, L0 L0^T = C0 (lower)
    T A0[10][10];      // L0^{-1} P
    T AtA0[10][10];    // A0^T A0 (both triangles)
    T G0[10][10];      // A0^T L0^{-1}: Atb = G0 s
};
struct TypeHdr {
    int32_t ns, np, has_pi, intime;  // np: the kernel's pulse slots (= ns)
    int32_t first[10];
    uint32_t pmask;  // the type's pulses: bits 0 .. npulses-1; slots past them hold
                     // zero columns, whose w is 0 and which never become passive
    int32_t npv;     // npulses of the type (<= ns)
};
// Inverses of the passive blocks of the x = 0 normal matrix AtA0, one per
// passive set m (bit mask over the pulses, 1 <= m < 2^np).  At x = 0 the
// normal matrix is the same for every channel of a type, so the solve of
// the passive system depends on the passive set alone: s_P = inv(AtA0[P,P])
// Atb_P.  inv[m] holds that inverse embedded in n x n (zeros outside m),
// its lower triangle packed by rows (entry (i, j), j <= i, at i(i+1)/2 + j),
// computed on the host in extended precision and rounded to FP64; slot
// MASK_PACK - 1 holds the smallest pivot ratio (M_kk - sum_j L_kj^2) / M_kk
// over k in m of the FP64 Cholesky of the block (<= 0 or NaN: that
// factorization fails; the inverse is then NaN).  A product with a rounded
// inverse is not backward stable: its residual grows with the block's
// condition number.  Where that (1-norm, from the extended-precision
// inverse) exceeds MASK_FACTOR_COND, bit m of factored is set and fac[m]
// holds the block's Cholesky factor, packed the same way with the
// reciprocals of its diagonal and the same last slot, for two triangular
// solves (solve_factor, at the end point of the steps: polish_end).
constexpr int MASK_PACK = 56;
constexpr int MASK_SETS = 1024;
constexpr double MASK_FACTOR_COND = 1e6;
struct MaskTab {
    alignas(16) double inv[MASK_SETS][MASK_PACK];
    alignas(16) double fac[MASK_SETS][MASK_PACK];
    uint32_t factored[MASK_SETS / 32];
};

struct DevType {
    TypeHdr h;
    TypeMats<double> d;
    TypeMats<float> f;
    const MaskTab *tab;  // passive-set factors of AtA0 (same memory space as the DevType)
    double inv_pederr2;  // 1 / pederr^2 of the type (a per-channel pederr p scales C0 by
                         // rho = p^2 / pederr^2)
};
template <typename T> MF_HD const TypeMats<T> &mats(const DevType &t);
template <> MF_HD const TypeMats<double> &mats<double>(const DevType &t) { return t.d; }
template <> MF_HD const TypeMats<float> &mats<float>(const DevType &t) { return t.f; }

struct FitParams {
    double eps;          // FNNLS max-w threshold
    double chi2_tol;     // Minimize |dchi2| threshold
    int32_t max_iter;    // FNNLS outer cap
    int32_t max_minimize;
    int32_t use_pi;      // amplitude-weighted pulse covariance live
    int32_t solver;      // 0 inplace_fnnls, 1 fnnls, 2 legacy (persisted passive set),
                         // 3 Lawson-Hanson (Eigen::NNLS)
    float eps_rel32;     // FP32 path: rounding allowance of w (see fnnls_inplace)
    float tau32;         // FP32 path: pivot ratio below which a solve goes to FP64
    double eps_rel64;    // FP32 path: rounding allowance of w in the FP64 phase
    // Per channel (a kernel sets them on its lane's copy):
    double rho = 1.0;          // (channel pederr / type pederr)^2: the noise covariance is
                               // rho C0, so the x = 0 normal equations are AtA0 / rho and
                               // G0 s / rho and chi2 of the type's whitening is rho chi2
    uint32_t pmask = 0x3FFu;   // the type's pulses (TypeHdr::pmask)
    int32_t pad_ = 0;
};

// A per-channel noise sigma must be positive and finite; a channel whose
// rho is not is not fitted (F_NONFINITE), like one with non-finite samples.
MF_HD bool rho_ok(double r) { return r > 0.0 && r <= 1.7976931348623157e308; }

enum {
    F_MINIMIZE_CAP = 1, F_LLT_FAIL = 2, F_FNNLS_CAP = 4, F_NONFINITE = 8,
    F_FP64_PIVOT = 16, F_WARM_REJECT = 32, F_REFINED = 64, F_BAD_TYPE = 128,
    F_WARM_FALLBACK = 256
};

struct FnnlsStats {
    int outer = 0, inner = 0;
    bool hit_max = false;
    bool llt_fail = false;
};

// ---------------------------------------------------------------------------
// dense kernels on register arrays (lower triangles of symmetric matrices)
// ---------------------------------------------------------------------------

// In-place Cholesky of the lower triangle of C (row-oriented, as
// linalgebra/interface/cholesky.hpp:8-38); ri[k] = 1/L_kk.  Returns false if a
// pivot was <= 0 or NaN.
template <typename T, int N>
MF_HD bool cholesky_inplace(T (&C)[10][10], T (&ri)[10])
{
    bool ok = true;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        T sumsq = T(0);
#pragma unroll
        for (int j = 0; j < i; ++j) {
            T s2 = T(0);
#pragma unroll
            for (int k = 0; k < j; ++k) s2 = mfma(C[i][k], C[j][k], s2);
            const T v = (C[i][j] - s2) * ri[j];
            C[i][j] = v;
            sumsq = mfma(v, v, sumsq);
        }
        const T d = C[i][i] - sumsq;
        ok = ok && (d > T(0));
        const T r = mrsqrt(d);
        ri[i] = r;
        C[i][i] = d * r;
    }
    return ok;
}

// Lower-triangular L (diagonal reciprocals ri) -> L^{-1}, in place.  Row i of
// the inverse needs row i of the factor and rows < i of the inverse, so row i
// of the factor is copied out before it is overwritten:
//   Z_ii = 1/L_ii,  Z_ij = -(sum_{k=j}^{i-1} L_ik Z_kj) / L_ii  (j < i).
template <typename T, int N>
MF_HD void invert_lower(T (&L)[10][10], const T (&ri)[10])
{
#pragma unroll
    for (int i = 0; i < N; ++i) {
        T row[10];
#pragma unroll
        for (int k = 0; k < i; ++k) row[k] = L[i][k];
#pragma unroll
        for (int j = 0; j < i; ++j) {
            T t = T(0);
#pragma unroll
            for (int k = j; k < i; ++k) t = mfma(row[k], L[k][j], t);
            L[i][j] = -t * ri[i];
        }
        L[i][i] = ri[i];
    }
}

// A pulse layout known at compile time.  D >= 0: pulse p's window starts at
// sample first_p >= lay_first<D>(p) = max(0, p - D) (consecutive BXs, as in
// every reference configuration: BX -5..4 with 10 samples and BX -4..3 with
// 8 give D = 2), so the rows of P, Pi_p and A = L^-1 P above lay_first<D>(p)
// are zero and their terms are left out.  Each such term adds an exact zero
// (an accumulator that starts at +0 stays +0, fma(0, y, a) == a), so the
// results are identical bit for bit.  D < 0: nothing known, nothing left out.
template <int D> MF_HD constexpr int lay_first(int p) { return D < 0 ? 0 : (p > D ? p - D : 0); }

// updateCov (PulseChiSqSNNLS.cpp:169-194): lower(C) = pederr^2 R + sum_p x_p^2 Pi_p.
// A pulse with x_p == 0 adds fma(0, Pi, C) == C exactly, so the reference's
// `continue` (:178-179) needs no branch.
// rho: the channel's scale of the type's noise covariance (1: the type's
// pederr; the type's pederr 1 and rho = pederr^2 give pederr^2 * noisecor
// exactly as the reference forms it).
template <typename T, int NS, int NP, int D = -1>
MF_HD void build_cov(const TypeMats<T> &tm, bool use_pi, const T (&x)[10], T (&C)[10][10],
                     T rho = T(1))
{
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) C[i][j] = rho * tm.C0[i][j];
    if (use_pi) {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            // a pulse of amplitude 0 adds exactly 0: skipped (a wavefront
            // skips the block when none of its lanes needs it)
            if (x[p] == T(0)) continue;
            const T a2 = x[p] * x[p];
#pragma unroll
            for (int i = 0; i < NS; ++i)
#pragma unroll
                for (int j = 0; j <= i; ++j)
                    if (j >= lay_first<D>(p)) C[i][j] = mfma(a2, tm.Pi[p][i][j], C[i][j]);
        }
    }
}

// Normal equations of the whitened problem A = L^{-1} P, b = L^{-1} s
// (PulseChiSqSNNLS.cpp:235-236, inplace_fnnls.cu:115-117), streamed by rows
// of A so that A is never held whole: row i of A needs only row i of L^{-1}.
//
// zc >= 0: column zc of the pulse matrix is taken as zero (DoFit's error
// computation zeroes the in-time column, PulseChiSqSNNLS.cpp:104-105).
template <typename T, int NS, int NP, int D = -1>
MF_HD void normal_equations(const TypeMats<T> &tm, const T (&Li)[10][10], const T (&s)[10],
                            T (&AtA)[10][10], T (&Atb)[10], int zc = -1)
{
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        Atb[i] = T(0);
#pragma unroll
        for (int j = 0; j <= i; ++j) AtA[i][j] = T(0);
    }
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        T b = T(0);
#pragma unroll
        for (int k = 0; k <= r; ++k) b = mfma(Li[r][k], s[k], b);
        // a[p] = row r of A: zero above column p's first sample
        T a[10];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            T t = T(0);
#pragma unroll
            for (int k = 0; k <= r; ++k)
                if (k >= lay_first<D>(p)) t = mfma(Li[r][k], tm.P[k][p], t);
            a[p] = (p == zc) ? T(0) : t;
        }
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            if (r < lay_first<D>(i)) continue;   // a[i] = 0: adds exactly 0
            Atb[i] = mfma(a[i], b, Atb[i]);
#pragma unroll
            for (int j = 0; j <= i; ++j) AtA[i][j] = mfma(a[i], a[j], AtA[i][j]);
        }
    }
}

// chi2 = |L^{-1} (P x - s)|^2 (ComputeChiSq, PulseChiSqSNNLS.cpp:196-200)
template <typename T, int NS, int NP, int D = -1>
MF_HD T chisq(const TypeMats<T> &tm, const T (&Li)[10][10], const T (&s)[10], const T (&x)[10],
              int zc = -1)
{
    T r[10];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        T t = T(0);
#pragma unroll
        for (int p = 0; p < NP; ++p)
            if (i >= lay_first<D>(p)) t = mfma(p == zc ? T(0) : tm.P[i][p], x[p], t);
        r[i] = t - s[i];
    }
    T q = T(0);
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        T y = T(0);
#pragma unroll
        for (int k = 0; k <= i; ++k) y = mfma(Li[i][k], r[k], y);
        q = mfma(y, y, q);
    }
    return q;
}

// Solve AtA[P,P] s_P = Atb_P for the passive set P (bit mask pm) with a
// masked full-size Cholesky; s is 0 outside P.  Returns false on a
// non-positive pivot inside P.
// tau > 0: *ill is set when a pivot inside P keeps less than a fraction tau
// of its diagonal entry (M_kk - sum_j L_kj^2 < tau M_kk): the cancellation has
// eaten the digits T carries and the solve should be redone in FP64.
template <typename T, int N>
MF_HD bool solve_passive(const T (&M)[10][10], const T (&b)[10], uint32_t pm, T (&s)[10],
                         T tau = T(0), bool *ill = nullptr)
{
    T L[10][10], ri[10], y[10];
    bool ok = true, bad = false;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const bool ink = (pm >> k) & 1u;
        T sq = T(0);
#pragma unroll
        for (int j = 0; j < k; ++j) sq = mfma(L[k][j], L[k][j], sq);
        const T d = ink ? M[k][k] - sq : T(1);
        ok = ok && (d > T(0));
        if (tau > T(0)) bad = bad || (ink && d < tau * M[k][k]);
        const T r = mrsqrt(d);
        ri[k] = r;
        T t = ink ? b[k] : T(0);
#pragma unroll
        for (int j = 0; j < k; ++j) t = mfma(-L[k][j], y[j], t);
        y[k] = t * r;
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            const bool ini = (pm >> i) & 1u;
            T u = T(0);
#pragma unroll
            for (int j = 0; j < k; ++j) u = mfma(L[i][j], L[k][j], u);
            L[i][k] = (ink && ini) ? (M[i][k] - u) * r : T(0);
        }
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        T t = y[i];
#pragma unroll
        for (int j = i + 1; j < N; ++j) t = mfma(-L[j][i], s[j], t);
        s[i] = t * ri[i];
    }
    if (ill) *ill = bad || !ok;
    return ok;
}

// The solve of solve_passive with the factor read from a MaskFactor: the
// same two triangular solves, in the same order, with L streamed from memory
// instead of being formed (and held) in registers.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T> using GlobalPtr = const __attribute__((address_space(1))) T *;
template <typename T> using GlobalOut = __attribute__((address_space(1))) T *;
typedef double Dbl2 __attribute__((ext_vector_type(2)));
#else
template <typename T> using GlobalPtr = const T *;
template <typename T> using GlobalOut = T *;
struct Dbl2 {
    double v[2];
    double operator[](int e) const { return v[e]; }
};
#endif

// (row, column) of packed lower-triangle entry e
struct TriIndex {
    int8_t i[MASK_PACK], j[MASK_PACK];
    constexpr TriIndex() : i(), j()
    {
        int e = 0;
        for (int r = 0; r < 10; ++r)
            for (int c = 0; c <= r; ++c, ++e) {
                i[e] = (int8_t)r;
                j[e] = (int8_t)c;
            }
        i[e] = j[e] = -1;
    }
};
constexpr TriIndex TRI{};

// s = inv b for the passive set pm, inv read from a MaskTab row (see there):
// s_i = sum_j inv_ij b_j over the packed lower triangle, each entry used for
// both (i, j) and (j, i).  A row is zero outside its passive block, so b
// needs no masking and s is zero outside P.  The row is streamed in 16-byte
// chunks, two stages of S chunks in flight.  Returns the block's pivot ratio
// (the row's last slot).
template <int N>
MF_HD double solve_inverse(const double *row, const double (&b)[10], uint32_t pm,
                           double (&s)[10])
{
    constexpr int NE = N * (N + 1) / 2;  // packed entries
    constexpr int NC = MASK_PACK / 2;    // chunks in a row (the last one carries the ratio)
    constexpr int NCU = (NE + 1) / 2;    // chunks holding entries (for N = 10 also the ratio)
    constexpr int S = 4;                 // chunks per stage
    constexpr int NST = (NCU + S - 1) / S;
    static_assert(NE <= MASK_PACK - 1 && NCU <= NC, "packing");
    (void)pm;
    const GlobalPtr<Dbl2> p = (GlobalPtr<Dbl2>)row;
#pragma unroll
    for (int k = 0; k < 10; ++k) s[k] = 0.0;
    Dbl2 buf[2][S];
    double ratio = 0.0;
#pragma unroll
    for (int st = 0; st < 2 && st < NST; ++st)
#pragma unroll
        for (int c = 0; c < S; ++c)
            if (st * S + c < NCU) buf[st][c] = p[st * S + c];
#pragma unroll
    for (int st = 0; st < NST; ++st) {
#pragma unroll
        for (int c = 0; c < S; ++c) {
            const int ch = st * S + c;
            if (ch >= NCU) continue;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int e = 2 * ch + h;
                const double v = buf[st & 1][c][h];
                if (e < NE) {
                    const int i = TRI.i[e], j = TRI.j[e];
                    s[i] = fma(v, b[j], s[i]);
                    if (i != j) s[j] = fma(v, b[i], s[j]);
                } else if (e == MASK_PACK - 1) {
                    ratio = v;
                }
            }
        }
        if (st + 2 < NST) {
            // the address of stage st + 2 depends on the last chunk of this
            // stage, so its loads issue once this stage has arrived: at most
            // two stages (2 S chunks) are in flight per lane
            GlobalPtr<Dbl2> q = p;
#if defined(__HIP_DEVICE_COMPILE__)
            asm volatile("" : "+v"(q) : "v"(buf[st & 1][S - 1]));
#endif
#pragma unroll
            for (int c = 0; c < S; ++c)
                if ((st + 2) * S + c < NCU) buf[st & 1][c] = q[(st + 2) * S + c];
        }
    }
    if (NCU < NC) ratio = ((GlobalPtr<double>)row)[MASK_PACK - 1];
    return ratio;
}

// s = (L L^T)^{-1} b for a passive set whose MaskTab row holds the block's
// Cholesky factor L (see there): the entries (i, j), j < i, of L by rows and
// 1 / L_ii on the diagonal, zero outside the block.  Forward substitution
// streams the rows in order, y_i = (b_i - sum_j L_ij y_j) / L_ii; backward
// substitution streams them again from the last, s_k = y_k / L_kk and then
// y_j -= L_kj s_k for j < k (s and y share the array).  A zero reciprocal
// outside the block leaves s zero there.  Returns the pivot ratio (the row's
// last slot), as solve_inverse does.
template <int N>
MF_HD double solve_factor(const double *row, const double (&b)[10], double (&s)[10])
{
    constexpr int NE = N * (N + 1) / 2;
    const GlobalPtr<double> p = (GlobalPtr<double>)row;
#pragma unroll
    for (int k = 0; k < 10; ++k) s[k] = k < N ? b[k] : 0.0;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const int i = TRI.i[e], j = TRI.j[e];
        const double v = p[e];
        if (j < i) s[i] = fma(-v, s[j], s[i]);
        else s[i] *= v;
    }
    // the same row again for the transposed factor, last row first
    GlobalPtr<double> q = p;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(q) : "v"(s[N - 1]));
#endif
#pragma unroll
    for (int e = NE - 1; e >= 0; --e) {
        const int i = TRI.i[e], j = TRI.j[e];
        const double v = q[e];
        if (j == i) s[i] *= v;
        else s[j] = fma(-v, s[i], s[j]);
    }
    return p[MASK_PACK - 1];
}

template <typename T>
MF_HD T sym(const T (&M)[10][10], int i, int j) { return i >= j ? M[i][j] : M[j][i]; }

// inplace_fnnls (nnls/src/inplace_fnnls.cu:72-233) on the normal equations.
// x: warm start in, solution out (original index order).  eps_rel > 0 adds
// eps_rel * (|Atb_i| + sum_j |AtA_ij x_j|), a bound on the rounding error of
// w_i, to the convergence threshold (used by the FP32 path only; the FP64
// path passes 0 and keeps the reference's test max_w < eps).
template <typename T, int N>
MF_HD void fnnls_inplace(const T (&AtA)[10][10], const T (&Atb)[10], T (&x)[10], T eps,
                         T eps_rel, int max_iter, FnnlsStats &st)
{
    uint32_t pm = 0;
    uint64_t perm = ident_nibbles();  // perm[pos] = original index
    uint64_t ipos = perm;             // ipos[idx] = position
    int np = 0;
    int iter = 0;
    for (; iter < max_iter; ++iter) {
        if (np == N) break;                                     // :134
        // w = Atb - AtA x on the free set; argmax, first in permuted order
        T wmax = T(0), wtol = T(0);
        int q = -1, qpos = 16;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T t = T(0), ta = T(0);
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const T a = sym(AtA, i, j);
                t = mfma(a, x[j], t);
                if (eps_rel > T(0)) ta = mfma(mabs(a), mabs(x[j]), ta);
            }
            const T w = Atb[i] - t;
            const int pos = nib(ipos, i);
            const bool isfree = !((pm >> i) & 1u);
            const bool better = (q < 0) || (w > wmax) || (w == wmax && pos < qpos);
            if (isfree && better) {
                wmax = w;
                q = i;
                qpos = pos;
                wtol = eps_rel > T(0) ? eps_rel * (mabs(Atb[i]) + ta) : T(0);
            }
        }
        if (wmax < eps + wtol) break;                           // :147
        {   // swap positions np <-> qpos                          :157-164
            const int inp = nib(perm, np);
            perm = setnib(setnib(perm, np, q), qpos, inp);
            ipos = setnib(setnib(ipos, q, np), inp, qpos);
        }
        pm |= 1u << q;
        ++np;
        st.outer++;

        for (;;) {                                              // :175
            T s[10];
            if (!solve_passive<T, N>(AtA, Atb, pm, s)) st.llt_fail = true;
            st.inner++;
            bool allpos = true;
#pragma unroll
            for (int i = 0; i < N; ++i)
                if ((pm >> i) & 1u) allpos = allpos && (s[i] > T(0));
            if (allpos) {                                       // :179-182
#pragma unroll
                for (int i = 0; i < N; ++i)
                    if ((pm >> i) & 1u) x[i] = s[i];
                break;
            }
            T alpha = Lim<T>::max();                            // :188-200
            int a = -1, apos = 16;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                if (((pm >> i) & 1u) && s[i] <= T(0)) {
                    const T ratio = mdiv(x[i], x[i] - s[i]);
                    const int pos = nib(ipos, i);
                    if (ratio < alpha || (a >= 0 && ratio == alpha && pos < apos)) {
                        alpha = ratio;
                        a = i;
                        apos = pos;
                    }
                }
            }
            if (a < 0) {                                        // :202-205
#pragma unroll
                for (int i = 0; i < N; ++i)
                    if ((pm >> i) & 1u) x[i] = s[i];
                break;
            }
#pragma unroll
            for (int i = 0; i < N; ++i)                         // :215-216
                if ((pm >> i) & 1u) x[i] = (i == a) ? T(0) : mfma(alpha, s[i] - x[i], x[i]);
            --np;
            {   // swap positions apos <-> np                      :222-229
                const int il = nib(perm, np);
                perm = setnib(setnib(perm, apos, il), np, a);
                ipos = setnib(setnib(ipos, il, apos), a, np);
            }
            pm &= ~(1u << a);
            if (np == 0) break;
        }
    }
    if (iter == max_iter && max_iter > 0) st.hit_max = true;
}

// inplace_fnnls_test_cholesky (nnls/src/inplace_fnnls_test_cholesky.cu:35-227):
// the loop of fnnls_inplace with the Cholesky factor of the passive block
// kept from one solve to the next.  The reference keeps the factor in its
// permuted (position) order: a 1 x 1 factor for one passive index
// (:145-155), the new last row appended after an add (:156-159, through
// fused_cholesky_substitution.hpp:18-57) and, after a removal (the removed
// position swapped with the last passive one), the rows from that position
// on recomputed (:160-165, :62-102).  Here the factor of AtA[P,P] is kept
// in ascending index order inside a masked 10 x 10 factor (rows and columns
// of free indices the identity), which keeps every register index static:
// adding index q changes rows >= q of the factor, removing index a rows >=
// a; those rows of the factor and of the forward solution y = L^{-1} Atb_P
// are recomputed (row-oriented, as linalgebra/interface/cholesky.hpp:8-38
// and the fused forward step), the rows before them are kept.  A row whose
// index no lane of a wavefront needs is skipped.  The permutation is kept as
// in fnnls_inplace and decides the ties the same way.
template <typename T> struct CholState {
    T L[10][10];  // masked factor, lower triangle (L[i][i] unused: ri holds 1/L_ii)
    T ri[10];
    T y[10];      // L^{-1} Atb_P
};

// Rows from .. N-1 of the masked factor of M[P,P] and of y; returns false on
// a non-positive pivot inside P among them.
template <typename T, int N>
MF_HD bool chol_rows_from(const T (&M)[10][10], const T (&b)[10], uint32_t pm, int from,
                          CholState<T> &C)
{
    bool ok = true;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (i < from) continue;   // kept from the solve before
        const bool ini = (pm >> i) & 1u;
        T sumsq = T(0), t = ini ? b[i] : T(0);
#pragma unroll
        for (int j = 0; j < i; ++j) {
            const bool inj = (pm >> j) & 1u;
            T s2 = T(0);
#pragma unroll
            for (int k = 0; k < j; ++k) s2 = mfma(C.L[i][k], C.L[j][k], s2);
            const T v = (ini && inj) ? (M[i][j] - s2) * C.ri[j] : T(0);
            C.L[i][j] = v;
            sumsq = mfma(v, v, sumsq);
            t = mfma(-v, C.y[j], t);
        }
        const T d = ini ? M[i][i] - sumsq : T(1);
        ok = ok && (d > T(0));
        const T r = mrsqrt(d);
        C.ri[i] = r;
        C.y[i] = t * r;
    }
    return ok;
}

template <typename T, int N>
MF_HD void fnnls_inplace_chol(const T (&AtA)[10][10], const T (&Atb)[10], T (&x)[10], T eps,
                              int max_iter, FnnlsStats &st)
{
    uint32_t pm = 0;
    uint64_t perm = ident_nibbles();
    uint64_t ipos = perm;
    int np = 0;
    int from = 0;   // first row of the factor that no longer holds
    CholState<T> C;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        C.ri[i] = T(1);
        C.y[i] = T(0);
#pragma unroll
        for (int j = 0; j < 10; ++j) C.L[i][j] = T(0);
    }
    int iter = 0;
    for (; iter < max_iter; ++iter) {
        if (np == N) break;                                     // :101-102
        T wmax = T(0);
        int q = -1, qpos = 16;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T t = T(0);
#pragma unroll
            for (int j = 0; j < N; ++j) t = mfma(sym(AtA, i, j), x[j], t);
            const T w = Atb[i] - t;
            const int pos = nib(ipos, i);
            const bool isfree = !((pm >> i) & 1u);
            const bool better = (q < 0) || (w > wmax) || (w == wmax && pos < qpos);
            if (isfree && better) {
                wmax = w;
                q = i;
                qpos = pos;
            }
        }
        if (wmax < eps) break;                                  // :114
        {   // swap positions np <-> qpos                          :124-133
            const int inp = nib(perm, np);
            perm = setnib(setnib(perm, np, q), qpos, inp);
            ipos = setnib(setnib(ipos, q, np), inp, qpos);
        }
        pm |= 1u << q;
        ++np;
        from = q < from ? q : from;                             // rows >= q change
        st.outer++;
        for (;;) {                                              // :144
            if (!chol_rows_from<T, N>(AtA, Atb, pm, from, C)) st.llt_fail = true;
            from = N;
            T s[10];
#pragma unroll
            for (int i = N - 1; i >= 0; --i) {                  // :166
                T t = C.y[i];
#pragma unroll
                for (int j = i + 1; j < N; ++j) t = mfma(-C.L[j][i], s[j], t);
                s[i] = t * C.ri[i];
            }
            st.inner++;
            bool allpos = true;
#pragma unroll
            for (int i = 0; i < N; ++i)
                if ((pm >> i) & 1u) allpos = allpos && (s[i] > T(0));
            if (allpos) {                                       // :170-173
#pragma unroll
                for (int i = 0; i < N; ++i)
                    if ((pm >> i) & 1u) x[i] = s[i];
                break;
            }
            T alpha = Lim<T>::max();                            // :179-191
            int a = -1, apos = 16;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                if (((pm >> i) & 1u) && s[i] <= T(0)) {
                    const T ratio = mdiv(x[i], x[i] - s[i]);
                    const int pos = nib(ipos, i);
                    if (ratio < alpha || (a >= 0 && ratio == alpha && pos < apos)) {
                        alpha = ratio;
                        a = i;
                        apos = pos;
                    }
                }
            }
            if (a < 0) {                                        // :193-196
#pragma unroll
                for (int i = 0; i < N; ++i)
                    if ((pm >> i) & 1u) x[i] = s[i];
                break;
            }
#pragma unroll
            for (int i = 0; i < N; ++i)                         // :206-207
                if ((pm >> i) & 1u) x[i] = (i == a) ? T(0) : mfma(alpha, s[i] - x[i], x[i]);
            --np;
            {   // swap positions apos <-> np                      :213-220
                const int il = nib(perm, np);
                perm = setnib(setnib(perm, apos, il), np, a);
                ipos = setnib(setnib(ipos, il, apos), a, np);
            }
            pm &= ~(1u << a);
            from = a;                                           // rows >= a change
            if (np == 0) break;
        }
    }
    if (iter == max_iter && max_iter > 0) st.hit_max = true;
}

// State of the legacy NNLS that lives across the Minimize iterations of one
// fit: the passive set, the permutation (position -> index and its inverse)
// and the number of passive positions (PulseChiSqSNNLS members _nP, and the
// column order of _pulsemat / _ampvec / _bxs, which the legacy code permutes
// in place: legacy_multifit_cpu/src/PulseChiSqSNNLS.cpp:304-311, 353-360).
struct LegacyState {
    uint32_t pm = 0;
    uint64_t perm = 0, ipos = 0;
    int np = 0;
    MF_HD void reset()
    {
        pm = 0;
        perm = ident_nibbles();
        ipos = perm;
        np = 0;
    }
    // takes index i out of the passive set: its position is swapped with the
    // last passive one (PulseChiSqSNNLS.cpp:151-157, the MINOS set-up)
    MF_HD void drop(int i)
    {
        const int pos = nib(ipos, i);
        if (pos >= np) return;
        --np;
        const int il = nib(perm, np);
        perm = setnib(setnib(perm, pos, il), np, i);
        ipos = setnib(setnib(ipos, il, pos), i, np);
        pm &= ~(1u << i);
    }
};

// Legacy NNLS (legacy_multifit_cpu/src/PulseChiSqSNNLS.cpp:268-363): FNNLS
// whose passive set and order persist from one Minimize iteration to the
// next.  The first pass of a call with passive positions already set skips
// the KKT test and solves first (:287).  w is taken on the free tail with the
// first maximum in permuted order (:296-297), the passive system is solved
// (the reference uses a fixed-size LDLT of the permuted leading block,
// :6-72; here the masked Cholesky of the same block), and the step back
// takes the first minimum ratio in permuted order, position 0 when no ratio
// is below the maximum (:331-341).  Amplitudes stay in original index order;
// the permutation only decides ties and which position a removal swaps with.
template <typename T, int N>
MF_HD void fnnls_legacy(const T (&AtA)[10][10], const T (&Atb)[10], T (&x)[10], T eps,
                        int max_iter, FnnlsStats &st, LegacyState &L)
{
    int iter = 0;
    for (; iter < max_iter; ++iter) {                           // :285
        if (iter > 0 || L.np == 0) {                            // :287
            if (L.np == N) break;                               // :288
            T wmax = T(0);
            int q = -1, qpos = 16;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                T t = T(0);
#pragma unroll
                for (int j = 0; j < N; ++j) t = mfma(sym(AtA, i, j), x[j], t);
                const T w = Atb[i] - t;
                const int pos = nib(L.ipos, i);
                const bool isfree = !((L.pm >> i) & 1u);
                const bool better = (q < 0) || (w > wmax) || (w == wmax && pos < qpos);
                if (isfree && better) {
                    wmax = w;
                    q = i;
                    qpos = pos;
                }
            }
            if (wmax < eps) break;                              // :300
            {   // swap positions np <-> qpos                      :304-311
                const int inp = nib(L.perm, L.np);
                L.perm = setnib(setnib(L.perm, L.np, q), qpos, inp);
                L.ipos = setnib(setnib(L.ipos, q, L.np), inp, qpos);
            }
            L.pm |= 1u << q;
            ++L.np;
            st.outer++;
        }
        while (L.np > 0) {                                      // :318
            T sv[10];
            if (!solve_passive<T, N>(AtA, Atb, L.pm, sv)) st.llt_fail = true;
            st.inner++;
            bool allpos = true;
#pragma unroll
            for (int i = 0; i < N; ++i)
                if ((L.pm >> i) & 1u) allpos = allpos && (sv[i] > T(0));
            if (allpos) {                                       // :326-329
#pragma unroll
                for (int i = 0; i < N; ++i)
                    if ((L.pm >> i) & 1u) x[i] = sv[i];
                break;
            }
            T alpha = Lim<T>::max();                            // :331-341
            int apos = 0;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                if (((L.pm >> i) & 1u) && sv[i] <= T(0)) {
                    const T ratio = mdiv(x[i], x[i] - sv[i]);
                    const int pos = nib(L.ipos, i);
                    if (ratio < alpha || (ratio == alpha && alpha != Lim<T>::max() && pos < apos)) {
                        alpha = ratio;
                        apos = pos;
                    }
                }
            }
            const int a = nib(L.perm, apos);
#pragma unroll
            for (int i = 0; i < N; ++i)                         // :343-349
                if ((L.pm >> i) & 1u) x[i] = mfma(alpha, sv[i] - x[i], x[i]);
            x[a] = T(0);
            L.drop(a);                                          // :353-360
        }
    }
    if (iter == max_iter && max_iter > 0) st.hit_max = true;
}

// Bound on the least-squares solves of lh_nnls when the caller sets none
// (max_iter <= 0): the reference then loops until convergence; a GPU lane
// needs an end.  Reaching it is reported like max_iter.
constexpr int LH_SOLVE_CAP = 1000;

// Lawson-Hanson NNLS with the semantics of the reference's Eigen::NNLS
// class (nnls/interface/eigen_nnls.h:336-414).  The passive set starts
// empty whatever x holds (x is the setX starting point, :124); the gradient
// is taken on the whole x; the argmax over the free set takes the first
// maximum in permuted order (:170-180) and stops when it is below eps
// (:361); a feasible solve replaces all of x, free entries becoming 0
// (:398-401); an infeasible one interpolates the passive entries and removes
// one index, the first minimum ratio in permuted order (:384-396, :443-456),
// without zeroing it.  max_iter bounds the least-squares solves (:375-377).
// The reference solves the passive system by a Householder QR of A's passive
// columns; here the masked Cholesky of AtA's passive block: the same
// solution, different rounding.  Returns false when max_iter stopped it.
template <typename T, int N>
MF_HD bool lh_nnls(const T (&AtA)[10][10], const T (&Atb)[10], T (&x)[10], T eps, int max_iter,
                   FnnlsStats &st)
{
    LegacyState L;   // passive set and permutation
    L.reset();
    const int cap = max_iter > 0 ? max_iter : LH_SOLVE_CAP;
    int nls = 0;
    for (;;) {
        if (L.np == N) return true;                             // :361
        T wmax = T(0);
        int q = -1, qpos = 16;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T t = T(0);
#pragma unroll
            for (int j = 0; j < N; ++j) t = mfma(sym(AtA, i, j), x[j], t);
            const T w = Atb[i] - t;
            const int pos = nib(L.ipos, i);
            const bool better = (q < 0) || (w > wmax) || (w == wmax && pos < qpos);
            if (!((L.pm >> i) & 1u) && better) {
                wmax = w;
                q = i;
                qpos = pos;
            }
        }
        if (wmax - eps < T(0)) return true;
        {   // swap positions np <-> qpos                         :431-440
            const int inp = nib(L.perm, L.np);
            L.perm = setnib(setnib(L.perm, L.np, q), qpos, inp);
            L.ipos = setnib(setnib(L.ipos, q, L.np), inp, qpos);
        }
        L.pm |= 1u << q;
        ++L.np;
        st.outer++;
        for (;;) {
            if (nls >= cap) {                                   // :375-377
                st.hit_max = true;
                return false;
            }
            T y[10];
            if (!solve_passive<T, N>(AtA, Atb, L.pm, y)) st.llt_fail = true;
            ++nls;
            st.inner++;
            bool feasible = true;
            T alpha = Lim<T>::max();
            int rpos = 16;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                if (((L.pm >> i) & 1u) && y[i] <= T(0)) {
                    feasible = false;
                    const T t = mdiv(-x[i], y[i] - x[i]);
                    const int pos = nib(L.ipos, i);
                    if (t < alpha || (t == alpha && pos < rpos)) {
                        alpha = t;
                        rpos = pos;
                    }
                }
            }
            if (feasible) {
#pragma unroll
                for (int i = 0; i < N; ++i) x[i] = y[i];
                break;
            }
#pragma unroll
            for (int i = 0; i < N; ++i)
                if ((L.pm >> i) & 1u) x[i] = x[i] + alpha * (y[i] - x[i]);
            L.drop(nib(L.perm, rpos));
        }
    }
}

// fnnls (nnls/src/fnnls.cu:28-173): boolean-mask FNNLS on the full x.
// Its outer test `!n_active || w < eps` (:113) counts the active indices
// minus one, so it stops as soon as a single index is still free, without
// adding it: a full-support optimum is never reached.  Kept as is.
// The reference starts its running maximum at index 0 whether or not index 0
// is still active (:82-95) and then sizes the passive block from a counter
// that is one short in that case, reading an unset row; here the passive
// block is the mask itself, as in the oracle.
// pmask: the columns that exist (a problem of fewer than N columns embedded
// in N; the "one index short" test counts only these).
template <typename T, int N>
MF_HD void fnnls_mask(const T (&AtA)[10][10], const T (&Atb)[10], T (&x)[10], T eps,
                      int max_iter, FnnlsStats &st, uint32_t pmask = 0x3FFu)
{
    uint32_t act = pmask & ((1u << N) - 1u);  // active (= not passive) set
    T s[10];
#pragma unroll
    for (int i = 0; i < N; ++i) s[i] = T(0);
    int iter = 0;
    for (; iter < max_iter; ++iter) {
        T w[10];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            T t = T(0);
#pragma unroll
            for (int j = 0; j < N; ++j) t = mfma(sym(AtA, i, j), x[j], t);
            w[i] = Atb[i] - t;
        }
        int maxi = 0, nact = -1;
        T wm = w[0];
#pragma unroll
        for (int i = 0; i < N; ++i)
            if ((act >> i) & 1u) {
                ++nact;
                if (w[i] > wm) { wm = w[i]; maxi = i; }
            }
        act &= ~(1u << maxi);                                   // :104
        if (!nact || wm < eps) break;                           // :113
        st.outer++;
        for (;;) {                                              // :118
            const uint32_t pm = pmask & ((1u << N) - 1u) & ~act;
            if (!pm) {
#pragma unroll
                for (int i = 0; i < N; ++i) x[i] = s[i];
                break;
            }
            if (!solve_passive<T, N>(AtA, Atb, pm, s)) st.llt_fail = true;
            st.inner++;
            T alpha = Lim<T>::max();
            int a = -1;
#pragma unroll
            for (int i = 0; i < N; ++i)                         // :141-149
                if (((pm >> i) & 1u) && s[i] <= T(0)) {
                    const T tmp = mdiv(-x[i], s[i] - x[i]);
                    if (tmp < alpha) { alpha = tmp; a = i; }
                }
            if (a < 0) {
#pragma unroll
                for (int i = 0; i < N; ++i) x[i] = s[i];
                break;
            }
#pragma unroll
            for (int i = 0; i < N; ++i)                         // :163-165
                x[i] = (i == a) ? T(0) : mfma(alpha, s[i] - x[i], x[i]);
            act |= 1u << a;
        }
    }
    if (iter == max_iter && max_iter > 0) st.hit_max = true;
}

// ---------------------------------------------------------------------------
// the per-channel fit (DoFit with errors disabled)
// ---------------------------------------------------------------------------
template <typename T> struct FitOut {
    T x[10];
    T err[10];
    T chi2;
    int flags;
    int n_min, n_outer, n_inner;
};

// One Minimize iteration of the reference (PulseChiSqSNNLS.cpp:154-164):
// updateCov + LLT, NNLS on the whitened problem (warm x), chi2.  Li (L^{-1} of
// the covariance) is parked in `park` between the normal equations and chi2;
// Park is any type with store(k, v) / load(k) for k < 55.  Returns chi2.
// XS: solver compiled in besides inplace_fnnls / fnnls (chosen at run time
// by prm.solver): XS_LEGACY (with ls), XS_LH or XS_CHOL.
enum { XS_NONE = 0, XS_LEGACY = 2, XS_LH = 3, XS_CHOL = 4 };

// The whitened normal equations of one Minimize iteration (updateCov + LLT,
// PulseChiSqSNNLS.cpp:155-157, 235-236), with L^{-1} parked in `park`.
template <typename T, int NS, int NP, int D = -1, typename Park>
MF_HD void iteration_problem(const TypeMats<T> &tm, const T (&s)[10], const FitParams &prm,
                             Park &park, const T (&x)[10], T (&AtA)[10][10], T (&Atb)[10],
                             int &flags, int zc)
{
    T Li[10][10], ri[10];
    build_cov<T, NS, NP, D>(tm, prm.use_pi != 0, x, Li, T(prm.rho));
    if (!cholesky_inplace<T, NS>(Li, ri)) flags |= F_LLT_FAIL;
    invert_lower<T, NS>(Li, ri);
    normal_equations<T, NS, NP, D>(tm, Li, s, AtA, Atb, zc);
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) park.store(i * (i + 1) / 2 + j, Li[i][j]);
}

// The NNLS call and chi2 of one Minimize iteration on its normal equations
// (PulseChiSqSNNLS.cpp:158-160); L^{-1} is read back from `park`.
template <typename T, int NS, int NP, int XS = XS_NONE, int D = -1, typename Park>
MF_HD T iteration_solve(const TypeMats<T> &tm, const T (&s)[10], const FitParams &prm,
                        Park &park, const T (&AtA)[10][10], const T (&Atb)[10], T (&x)[10],
                        FnnlsStats &st, int zc, LegacyState *ls)
{
    if (XS == XS_LEGACY && ls)
        fnnls_legacy<T, NP>(AtA, Atb, x, T(prm.eps), prm.max_iter, st, *ls);
    else if (XS == XS_LH)   // PulseChiSqSNNLS.cpp:253-257 (setX, solve, x)
        lh_nnls<T, NP>(AtA, Atb, x, T(prm.eps), prm.max_iter, st);
    else if (XS == XS_CHOL)  // the incremental-Cholesky twin (PulseChiSqSNNLS.cpp:267)
        fnnls_inplace_chol<T, NP>(AtA, Atb, x, T(prm.eps), prm.max_iter, st);
    else if (prm.solver == 1)
        fnnls_mask<T, NP>(AtA, Atb, x, T(prm.eps), prm.max_iter, st, prm.pmask);
    else
        fnnls_inplace<T, NP>(AtA, Atb, x, T(prm.eps), T(0), prm.max_iter, st);
    T Li[10][10];
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) Li[i][j] = park.load(i * (i + 1) / 2 + j);
    return chisq<T, NS, NP, D>(tm, Li, s, x, zc);
}

// The stop test that opens an inplace_fnnls call whose passive set starts
// empty (nnls/src/inplace_fnnls.cu:140-147: max_i w_i < eps), at x, with
// w = A^T (b - A x) formed from the residual instead of the normal
// equations: w = P^T C^-1 (s - P x) = -P^T L^-T u, u = L^-1 (P x - s).  Also
// gives |u|^2, the chi2 at x (ComputeChiSq).  It is the reference's
// Atb - AtA x in exact arithmetic; in rounding it differs, which can move
// only a decision taken within rounding of eps.  L: the covariance's
// Cholesky factor (cholesky_inplace: ri the reciprocal diagonal).
template <typename T, int NS, int NP, int D>
MF_HD void forward_subst(const T (&L)[10][10], const T (&ri)[10], T (&y)[10])
{
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        T v = y[i];
#pragma unroll
        for (int k = 0; k < i; ++k) v = mfma(-L[i][k], y[k], v);
        y[i] = v * ri[i];
    }
}
template <typename T, int NS>
MF_HD void backward_subst(const T (&L)[10][10], const T (&ri)[10], T (&y)[10])
{
#pragma unroll
    for (int i = NS - 1; i >= 0; --i) {
        T v = y[i];
#pragma unroll
        for (int k = i + 1; k < NS; ++k) v = mfma(-L[k][i], y[k], v);
        y[i] = v * ri[i];
    }
}
// y_p = (P^T v)_p
template <typename T, int NS, int NP, int D>
MF_HD void pt_times(const TypeMats<T> &tm, const T (&v)[10], T (&y)[10])
{
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        T t = T(0);
#pragma unroll
        for (int i = 0; i < NS; ++i)
            if (i >= lay_first<D>(p)) t = mfma(tm.P[i][p], v[i], t);
        y[p] = t;
    }
}

// u = L^-1 (P x - s), w = -P^T L^-T u (see above), chi2 = |u|^2; returns
// whether max_p w_p < eps.
template <typename T, int NS, int NP, int D, typename TS>
MF_HD bool stop_at_start(const TypeMats<T> &tm, const TS (&s)[10], const T (&x)[10],
                         const T (&L)[10][10], const T (&ri)[10], T eps, T &chi2, T (&w)[10])
{
    MF_NO_CONTRACT
    T u[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) u[i] = T(0);
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        T t = T(0);
#pragma unroll
        for (int p = 0; p < NP; ++p)
            if (i >= lay_first<D>(p)) t = mfma(tm.P[i][p], x[p], t);
        u[i] = t - T(s[i]);
    }
    forward_subst<T, NS, NP, D>(L, ri, u);
    chi2 = T(0);
#pragma unroll
    for (int i = 0; i < NS; ++i) chi2 = mfma(u[i], u[i], chi2);
    T v[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) v[i] = u[i];
    backward_subst<T, NS>(L, ri, v);
    pt_times<T, NS, NP, D>(tm, v, w);
    bool stop = true;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        w[p] = -w[p];
        stop = stop && (w[p] < eps);
    }
    return stop;
}

// One add and one solve, what the later iterations of the fast path do
// when the stop test fails: the reference's inplace_fnnls from the warm x
// with its passive set reset adds the index q of the largest w (the first
// of equal ones), solves the passive system of {q} alone (x_q = Atb_q /
// AtA_qq by the masked factorisation's steps; the other amplitudes kept)
// and stops when no free w reaches eps (nnls/src/inplace_fnnls.cu:140-233).
// Formed from the factor L and the w, chi2 of stop_at_start: a_q = L^-1 P_q,
// AtA_qq = a_q . a_q, g = L^-T a_q = C^-1 P_q, Atb_q = a_q . L^-1 s = g . s,
// column q of AtA = P^T g, the new w = w - (P^T g) dx_q, the new chi2 =
// |u + dx_q a_q|^2 = chi2 + dx_q (dx_q AtA_qq - 2 w_q) (a_q . u = -w_q), so
// that neither u nor a_q is held past the solve.  Returns false when the
// channel leaves that case (the solve is not positive, a second index is
// added, max_iter < 2); x is then unchanged and the caller runs the
// iteration in full.
template <typename T, int NS, int NP, int D, typename TS>
MF_HD bool single_add(const TypeMats<T> &tm, const TS (&s)[10], T (&x)[10], const T (&L)[10][10],
                      const T (&ri)[10], const T (&w)[10], T eps, int max_iter, T &chi2)
{
    MF_NO_CONTRACT
    if (max_iter < 2) return false;
    int q = 0;
    T wq = w[0];
#pragma unroll
    for (int p = 1; p < NP; ++p)
        if (w[p] > wq) {
            wq = w[p];
            q = p;
        }
    T a[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        T c = T(0);
#pragma unroll
        for (int p = 0; p < NP; ++p)
            if (i < NS && i >= lay_first<D>(p)) c = (p == q) ? tm.P[i][p] : c;
        a[i] = c;
    }
    forward_subst<T, NS, NP, D>(L, ri, a);
    T att = T(0);
#pragma unroll
    for (int i = 0; i < NS; ++i) att = mfma(a[i], a[i], att);
    if (!(att > T(0))) return false;
    T g[10], c[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) g[i] = a[i];
    backward_subst<T, NS>(L, ri, g);
    T atb = T(0);
#pragma unroll
    for (int i = 0; i < NS; ++i) atb = mfma(g[i], T(s[i]), atb);
    const T r = mrsqrt(att);
    const T sq = (atb * r) * r;
    if (!(sq > T(0))) return false;
    const T dx = sq - x[q];
    pt_times<T, NS, NP, D>(tm, g, c);   // column q of AtA
    bool stop = true;
#pragma unroll
    for (int p = 0; p < NP; ++p) stop = stop && (p == q || mfma(-c[p], dx, w[p]) < eps);
    if (!stop) return false;
    x[q] = sq;
    chi2 = mfma(dx, mfma(dx, att, T(-2) * wq), chi2);
    return true;
}

// LAZY (the fast path's later iterations, inplace_fnnls): the stop test is
// taken first from the residual (stop_at_start); a channel that passes it
// keeps x and gets its chi2 without the normal equations, and one that adds
// a single index and stops (single_add) needs only that index's column.
// Only the channels left form the normal equations.  *kept: set when the
// stop test passed at the start (x unchanged).
// ONLY (with LAZY): the iteration is not run in full; a channel that needs
// the normal equations sets *pend and keeps x, for a kernel that holds them
// (so that this one needs neither their registers nor L^{-1}'s park).
template <typename T, int NS, int NP, int XS = XS_NONE, int D = -1, bool LAZY = false,
          bool ONLY = false, typename Park, typename TS>
MF_HD T minimize_step(const TypeMats<T> &tm0, const TS (&s)[10], const FitParams &prm,
                      Park &park, T (&x)[10], FnnlsStats &st, int &flags, int zc = -1,
                      LegacyState *ls = nullptr, bool *kept = nullptr, bool *pend = nullptr)
{
    const TypeMats<T> &tm = *opaque(&tm0);
    // TS: the samples' type; other than T only for ONLY (the lazy steps
    // convert each sample where they use it)
    static_assert(std::is_same<TS, T>::value || ONLY, "samples in T");
    if (LAZY && XS == XS_NONE && zc < 0 && prm.solver == 0) {
        T L[10][10], ri[10], chi2, w[10];
        build_cov<T, NS, NP, D>(tm, prm.use_pi != 0, x, L, T(prm.rho));
        if (!cholesky_inplace<T, NS>(L, ri)) flags |= F_LLT_FAIL;
        if (stop_at_start<T, NS, NP, D>(tm, s, x, L, ri, T(prm.eps), chi2, w)) {
            if (kept) *kept = true;
            return chi2;
        }
        if (single_add<T, NS, NP, D>(tm, s, x, L, ri, w, T(prm.eps), prm.max_iter, chi2)) {
            st.outer++;
            st.inner++;
            return chi2;
        }
        if constexpr (ONLY) {
            *pend = true;
            return chi2;
        } else {
            T AtA[10][10], Atb[10];
            invert_lower<T, NS>(L, ri);
            normal_equations<T, NS, NP, D>(tm, L, s, AtA, Atb, zc);
#pragma unroll
            for (int i = 0; i < NS; ++i)
#pragma unroll
                for (int j = 0; j <= i; ++j) park.store(i * (i + 1) / 2 + j, L[i][j]);
            return iteration_solve<T, NS, NP, XS, D>(tm, s, prm, park, AtA, Atb, x, st, zc, ls);
        }
    }
    if constexpr (ONLY) {
        *pend = true;
        return T(0);
    } else {
        T AtA[10][10], Atb[10];
        iteration_problem<T, NS, NP, D>(tm, s, prm, park, x, AtA, Atb, flags, zc);
        return iteration_solve<T, NS, NP, XS, D>(tm, s, prm, park, AtA, Atb, x, st, zc, ls);
    }
}

// Minimize (PulseChiSqSNNLS.cpp:149-166) from iteration `it0` on, with the
// previous chi2 `chi2` and amplitudes x.  Returns the iteration count.
// KEEP: form the normal equations once when the covariance does not depend
// on x (off for the error refits, whose extra register state costs more
// than the repeated set-up there).
// ONLY (with LAZY, see minimize_step): when an iteration needs the normal
// equations, *pend is set and the index of that iteration returned, with x
// and chi2 as they were at its start.
template <typename T, int NS, int NP, int XS = XS_NONE, int D = -1, bool KEEP = true,
          bool LAZY = false, bool ONLY = false, typename Park, typename TS>
MF_HD int minimize_from(const TypeMats<T> &tm, const TS (&s)[10], const FitParams &prm,
                        Park &park, T (&x)[10], T &chi2, int it0, FnnlsStats &st, int &flags,
                        int zc = -1, LegacyState *ls = nullptr, bool *pend = nullptr)
{
    int it = it0;
    if constexpr (KEEP) if (!prm.use_pi && it < prm.max_minimize) {
        // The covariance does not depend on x: every iteration forms the
        // same normal equations and L^{-1}, bit for bit, so they are formed
        // once and kept (the solves and chi2 are the reference's).
        const TypeMats<T> &tmo = *opaque(&tm);
        T AtA[10][10], Atb[10];
        iteration_problem<T, NS, NP, D>(tmo, s, prm, park, x, AtA, Atb, flags, zc);
        for (; it < prm.max_minimize; ++it) {                   // :154
            const T now =
                iteration_solve<T, NS, NP, XS, D>(tmo, s, prm, park, AtA, Atb, x, st, zc, ls);
            const T d = now - chi2;
            chi2 = now;
            if (mabs(d) < T(prm.chi2_tol)) return it + 1;       // :162
        }
        flags |= F_MINIMIZE_CAP;
        return it;
    }
    if constexpr (!LAZY) {
        for (; it < prm.max_minimize; ++it) {                   // :154
            const T now =
                minimize_step<T, NS, NP, XS, D>(tm, s, prm, park, x, st, flags, zc, ls);
            const T d = now - chi2;
            chi2 = now;
            if (mabs(d) < T(prm.chi2_tol)) return it + 1;       // :162
        }
        flags |= F_MINIMIZE_CAP;
        return it;
    }
    // LAZY: one exit from the loop (the iteration count of the end).  With
    // a return inside the loop, k_fast_p2 (the whole later-iteration body in
    // one kernel) kept a lane that had returned iterating with the rest of
    // its wavefront: only its count and the cap flag showed it (the extra
    // iterations kept x), and only in wavefronts whose other lanes ran on.
    int done = -1;
    for (; it < prm.max_minimize && done < 0; ++it) {           // :154
        bool kept = false, pd = false;
        const T now = minimize_step<T, NS, NP, XS, D, LAZY, ONLY>(tm, s, prm, park, x, st, flags,
                                                                  zc, ls, &kept, &pd);
        if (ONLY && pd) {
            *pend = true;
            done = it;
        } else {
            const T d = now - chi2;
            chi2 = now;
            if (mabs(d) < T(prm.chi2_tol)) {                    // :162
                done = it + 1;
            } else if (LAZY && kept && now - now == T(0) && T(prm.chi2_tol) > T(0) &&
                       it + 1 < prm.max_minimize) {
                // LAZY: an iteration that kept x is followed by one that
                // rebuilds the same covariance from the same x, passes the
                // same stop test and returns the same chi2 bit for bit; its
                // difference is 0, so Minimize ends after it (a finite chi2
                // and a positive tolerance).  It is counted, not run.
                done = it + 2;
            }
        }
    }
    if (done >= 0) return done;
    flags |= F_MINIMIZE_CAP;
    return it;
}

template <typename T, int NS>
MF_HD bool all_finite(const T (&s)[10])
{
    bool finite = true;
#pragma unroll
    for (int i = 0; i < NS; ++i) finite = finite && (s[i] - s[i] == T(0));
    return finite;
}

// MINOS-like uncertainty of the in-time amplitude (PulseChiSqSNNLS.cpp:73-146)
// after a fit that ended with amplitudes x, chi2 and L^{-1} of its last
// covariance in `park`.  ip: the first pulse with BX 0 (type constant).
// approxerr = 1 / |L^{-1} p_ip| (ComputeApproxUncertainty, :202-207); the
// in-time column is zeroed, its amplitude fixed at x0 + approxerr and the
// samples reduced by it, Minimize continues from the current state, and
// sigma+ = approxerr / sqrt(chi2+ - chi2_0); if x0 / sigma+ > 0.5 the same is
// done at max(0, x0 - approxerr) and the two are averaged (:120-142).
//
// ls (legacy solver): the fit's persisted passive state.  When the in-time
// pulse is passive it leaves the passive set first (:151-157 -- a branch the
// current reference never takes, its _nP staying 0), and the two Minimize
// calls continue from that state.
template <typename T, int NS, int NP, int XS = XS_NONE, int D = -1, typename Park>
MF_HD T minos_error(const TypeMats<T> &tm, const T (&s)[10], const FitParams &prm, Park &park,
                    const T (&xfit)[10], T chi2_0, int ip, FnnlsStats &st, int &flags,
                    LegacyState *ls = nullptr)
{
    if (ls) ls->drop(ip);
    T y = T(0), q = T(0);
    {
        T col[10];
#pragma unroll
        for (int i = 0; i < NS; ++i) col[i] = tm.P[i][ip];
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            T t = T(0);
#pragma unroll
            for (int k = 0; k <= i; ++k) t = mfma(park.load(i * (i + 1) / 2 + k), col[k], t);
            q = mfma(t, t, q);
        }
        y = q;
    }
    const T approxerr = T(1) / sqrt(y);
    const T x0 = xfit[ip];
    T x[10], sr[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) x[i] = xfit[i];
    T chi2 = chi2_0;
    // upper point
    const T xplus = x0 + approxerr;
    x[ip] = xplus;
#pragma unroll
    for (int i = 0; i < 10; ++i) sr[i] = i < NS ? s[i] - xplus * tm.P[i][ip] : T(0);
    minimize_from<T, NS, NP, XS, D, false>(tm, sr, prm, park, x, chi2, 0, st, flags, ip, ls);
    T Li[10][10];
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) Li[i][j] = park.load(i * (i + 1) / 2 + j);
    const T chiplus = chisq<T, NS, NP, D>(tm, Li, sr, x, ip);      // :117
    const T sigplus = mabs(xplus - x0) / sqrt(chiplus - chi2_0);
    if (!(x0 / sigplus > T(0.5))) return sigplus;              // :122, :141-142
    T xminus = x0 - approxerr;                                  // :130
    if (xminus < T(0)) xminus = T(0);
    x[ip] = xminus;
#pragma unroll
    for (int i = 0; i < 10; ++i) sr[i] = i < NS ? s[i] - xminus * tm.P[i][ip] : T(0);
    minimize_from<T, NS, NP, XS, D, false>(tm, sr, prm, park, x, chi2, 0, st, flags, ip, ls);
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) Li[i][j] = park.load(i * (i + 1) / 2 + j);
    const T chiminus = chisq<T, NS, NP, D>(tm, Li, sr, x, ip);
    const T sigminus = mabs(xminus - x0) / sqrt(chiminus - chi2_0);
    return T(0.5) * (sigplus + sigminus);
}

// Plain fit: the reference's algorithm in precision T, cold start, and with
// ERRORS the in-time uncertainty (ip >= 0: index of the in-time pulse).
// XS: XS_LEGACY for the legacy NNLS (prm.solver 2), whose passive state
// lives through the whole fit, XS_LH for Lawson-Hanson (prm.solver 3);
// separate instantiations, so that the default solvers do not carry their
// state and code.
// Non-finite samples are not fitted (x = 0, F_NONFINITE).
template <typename T, int NS, int NP, bool ERRORS = false, int XS = XS_NONE, int D = -1,
          typename Park>
MF_HD void fit_plain(const TypeMats<T> &tm, const T (&s)[10], const FitParams &prm,
                     Park &park, FitOut<T> &out, int ip = -1)
{
    T x[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        x[i] = T(0);
        out.err[i] = T(0);
    }
    T chi2 = T(0);
    int flags = 0, nmin = 0;
    FnnlsStats st;
    LegacyState ls;   // legacy solver: passive state across Minimize iterations
    ls.reset();
    LegacyState *lsp = XS == XS_LEGACY ? &ls : nullptr;
    if (all_finite<T, NS>(s) && rho_ok(prm.rho))
        nmin = minimize_from<T, NS, NP, XS, D>(tm, s, prm, park, x, chi2, 0, st, flags, -1, lsp);
    else
        flags |= F_NONFINITE;
    if (ERRORS && ip >= 0 && !(flags & F_NONFINITE) && prm.max_minimize > 0) {
        FnnlsStats st2;
        int fl2 = 0;
        const T e = minos_error<T, NS, NP, XS, D>(tm, s, prm, park, x, chi2, ip, st2, fl2, lsp);
#pragma unroll
        for (int i = 0; i < 10; ++i)
            if (i == ip) out.err[i] = e;
    }
    if (st.hit_max) flags |= F_FNNLS_CAP;
    if (st.llt_fail) flags |= F_LLT_FAIL;
#pragma unroll
    for (int i = 0; i < 10; ++i) out.x[i] = x[i];
    out.chi2 = chi2;
    out.flags = flags;
    out.n_min = nmin;
    out.n_outer = st.outer;
    out.n_inner = st.inner;
}

// ---------------------------------------------------------------------------
// the fast path: FP32 bulk, exact NNLS optimum in the first Minimize
// iteration (FP64 refinement + FP64 KKT verification), the reference's FNNLS
// semantics in later iterations
// ---------------------------------------------------------------------------

// Active-set NNLS (Lawson-Hanson / FNNLS steps) from a feasible start: x > 0
// on the passive mask pm and x = 0 off it.  x = 0 with any pm is accepted
// too: the first solve then drops infeasible indices one at a time (alpha =
// 0).  Free entries stay exactly 0, so the fixed point is the NNLS optimum of
// (M, Atb).  Ties go to the lowest index.
//
// tau > 0 turns on the pivot monitor of solve_passive: the first solve that
// reports an ill-conditioned passive block returns with *ill set and (x, pm)
// left at the last feasible state, for the caller to continue in FP64.
//
// UNI: M lives in memory and is the same for every lane (a type's table): it
// is re-read from the scalar cache on every pass instead of being held in
// scalar registers across the loops.
//
// TAB (FP64): M0 is the type's AtA0 and tab holds the inverses of its
// passive blocks: each solve is a product with the set's inverse instead of
// a Cholesky factorization, and the factorization's success is read from
// the set's FP64 pivot ratio.
template <typename T> struct ActiveState {
    T x[10];
    uint32_t pm;  // passive set
    bool solve;   // a solve of the passive system is due
    int adds;     // indices added so far
};

// One step of fnnls_active.  If a solve is due: solve the passive system,
// then either accept it (all positive) or step from x towards it until the
// first passive value reaches zero and drop that index (a solve stays due).
// If no solve is due afterwards: test the KKT conditions and add the free
// index of largest w (a solve becomes due).  Returns true when the iteration
// has ended: KKT met, max_iter adds, or (tau > 0) an ill-conditioned block
// (*ill set, state left at the last feasible point).  A lane of a persistent
// kernel runs one step per pass and takes a new channel when this returns
// true; fnnls_active runs the steps to the end.
template <typename T, int N, bool UNI = false, bool TAB = false>
MF_HD bool fnnls_step(const T (&M0)[10][10], const T (&Atb)[10], ActiveState<T> &S, T eps,
                      T eps_rel, int max_iter, FnnlsStats &st, T tau = T(0),
                      bool *ill = nullptr, const MaskTab *tab = nullptr)
{
    if (S.solve && !S.pm) S.solve = false;
    if (S.solve) {
        const uint32_t pm = S.pm;
        T s[10];
        bool bad = false, ok;
        if constexpr (TAB) {
            static_assert(sizeof(T) == 8, "the inverse tables are FP64");
            const double r = solve_inverse<N>(tab->inv[pm], Atb, pm, s);
            ok = r > 0.0;
            bad = !(r >= (double)tau);
        } else {
            const T(&M)[10][10] = UNI ? *opaque(&M0) : M0;
            ok = solve_passive<T, N>(M, Atb, pm, s, tau, &bad);
        }
        st.inner++;
        if (tau > T(0) && bad) {
            *ill = true;
            return true;
        }
        if (!ok) st.llt_fail = true;
        bool allpos = true;
#pragma unroll
        for (int i = 0; i < N; ++i)
            if ((pm >> i) & 1u) allpos = allpos && (s[i] > T(0));
        if (allpos) {
#pragma unroll
            for (int i = 0; i < N; ++i)
                if ((pm >> i) & 1u) S.x[i] = s[i];
            S.solve = false;
        } else {
            T alpha = Lim<T>::max();
            int a = -1, afirst = -1;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                if (((pm >> i) & 1u) && !(s[i] > T(0))) {
                    if (afirst < 0) afirst = i;
                    const T ratio = mdiv(S.x[i], S.x[i] - s[i]);
                    if (ratio < alpha) {
                        alpha = ratio;
                        a = i;
                    }
                }
            }
            if (a < 0) {  // only 0/0 ratios (x_i = s_i = 0) or NaN: drop without a step
                a = afirst;
                alpha = T(0);
            }
#pragma unroll
            for (int i = 0; i < N; ++i)
                if ((pm >> i) & 1u)
                    S.x[i] = (i == a) ? T(0) : mfma(alpha, s[i] - S.x[i], S.x[i]);
            S.pm = pm & ~(1u << a);
            return false;
        }
    }
    if (S.adds >= max_iter) {
        st.hit_max = true;
        return true;
    }
    int q = -1;
    T best = T(0);
    const T(&M)[10][10] = UNI ? *opaque(&M0) : M0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        T t = T(0), ta = T(0);
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const T a = sym(M, i, j);
            t = mfma(a, S.x[j], t);
            ta = mfma(mabs(a), mabs(S.x[j]), ta);
        }
        const T w = Atb[i] - t;
        const T thr = eps + eps_rel * (mabs(Atb[i]) + ta);
        if (!((S.pm >> i) & 1u) && w > thr && (q < 0 || w > best)) {
            best = w;
            q = i;
        }
    }
    if (q < 0) return true;
    S.pm |= 1u << q;
    ++S.adds;
    st.outer++;
    S.solve = true;
    return false;
}

// The KKT point of the table steps, solved again where its set has a MaskTab
// factor row (an ill-conditioned block, where a
// product with the rounded inverse is not backward stable): when that solve
// is all positive it replaces x, and the caller repeats the KKT test by
// running the steps again (returns true).  The steps before use the inverse
// on every set: only the end point's accuracy rests on the solve, and the
// factor is not on the path of the common case.  polished: the last set so
// solved (0 for none), so that each set is solved this way once.
template <int N>
MF_HD bool polish_end(const MaskTab *tab, const double (&Atb)[10], ActiveState<double> &S,
                      FnnlsStats &st, uint32_t &polished)
{
    const uint32_t pm = S.pm;
    if (!pm || pm == polished || st.hit_max ||
        !((((GlobalPtr<uint32_t>)tab->factored)[pm >> 5] >> (pm & 31u)) & 1u))
        return false;
    polished = pm;
    double s[10];
    bool allpos = solve_factor<N>(tab->fac[pm], Atb, s) > 0.0;
    st.inner++;
#pragma unroll
    for (int i = 0; i < N; ++i)
        if ((pm >> i) & 1u) allpos = allpos && (s[i] > 0.0);
    if (!allpos) return false;
#pragma unroll
    for (int i = 0; i < N; ++i)
        if ((pm >> i) & 1u) S.x[i] = s[i];
    S.solve = false;
    return true;
}

template <typename T, int N, bool UNI = false, bool TAB = false>
MF_HD void fnnls_active(const T (&M0)[10][10], const T (&Atb)[10], T (&x)[10], uint32_t &pm,
                        T eps, T eps_rel, int max_iter, FnnlsStats &st, T tau = T(0),
                        bool *ill = nullptr, const MaskTab *tab = nullptr)
{
    ActiveState<T> S;
#pragma unroll
    for (int i = 0; i < 10; ++i) S.x[i] = x[i];
    S.pm = pm;
    S.solve = pm != 0;
    S.adds = 0;
    while (!fnnls_step<T, N, UNI, TAB>(M0, Atb, S, eps, eps_rel, max_iter, st, tau, ill, tab)) {
    }
#pragma unroll
    for (int i = 0; i < 10; ++i) x[i] = S.x[i];
    pm = S.pm;
}

// chi2 with an FP64 residual: r = s - P x in FP64, y = Li r with the given
// (FP32 or FP64) inverse factor, chi2 = |y|^2.
template <typename TL, typename TX, int NS, int NP>
MF_HD double chisq64(const TypeMats<double> &td, const TL (&Li)[10][10], const double (&s)[10],
                     const TX (&x)[10])
{
    double r[10];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        double t = s[i];
#pragma unroll
        for (int p = 0; p < NP; ++p) t = fma(-td.P[i][p], (double)x[p], t);
        r[i] = t;
    }
    double q = 0.0;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        double y = 0.0;
#pragma unroll
        for (int k = 0; k <= i; ++k) y = fma((double)Li[i][k], r[k], y);
        q = fma(y, y, q);
    }
    return q;
}

// First Minimize iteration of the fast fit.  The covariance is C0 for every
// channel of the type at x = 0, so the normal equations are the type's AtA0
// and Atb = G0 s (no per-channel covariance, Cholesky or whitening).
//   phase 1 (FP32, fast_phase1, cold starts): active-set iterations from x = 0
//     until the FP32 KKT test passes or a pivot of the passive block loses
//     more than a fraction tau32 of its diagonal (F_FP64_PIVOT);
//   phase 2 (FP64, fast_phase2): the same iterations continue from that state
//     (a warm start: from the predicted passive set and amplitudes) on the
//     FP64 normal equations with the reference's threshold eps: the first
//     solve refines the amplitudes to FP64, the KKT test verifies the active
//     set (a warm start that needed more steps is flagged F_WARM_REJECT).
// The result is the NNLS optimum of the x = 0 problem, which is what the
// reference's first (cold) FNNLS computes.  The two phases are separate
// functions so that a kernel can run them as separate launches (with a live
// pulse covariance phase 2 needs about twice the registers of phase 1).
// Phase 2 takes the passive set as the entries with x > 0.
template <int NS, int NP>
MF_HD void fast_phase1(const DevType &ty, const float (&s32)[10], const FitParams &prm,
                       float (&x)[10], uint32_t &pm, FnnlsStats &st, int &flags)
{
    const TypeMats<float> &tf = ty.f;
    float Atb[10];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        float t = 0.0f;
#pragma unroll
        for (int k = 0; k < NS; ++k) t = mfma(tf.G0[p][k], s32[k], t);
        Atb[p] = t;
    }
    bool ill = false;
    // AtA0 and G0 s are rho times the channel's normal equations: the same
    // solutions, w scaled by rho, so the threshold is rho eps
    fnnls_active<float, NP, true>(tf.AtA0, Atb, x, pm, (float)(prm.eps * prm.rho),
                                  prm.eps_rel32, prm.max_iter, st, prm.tau32, &ill);
    if (ill) flags |= F_FP64_PIVOT;
}

// The fallback of a warm start: the state of a cold start (x = 0, nothing
// passive), with the adds counted afresh.
MF_HD void warm_fallback_reset(ActiveState<double> &S, FnnlsStats &st)
{
#pragma unroll
    for (int i = 0; i < 10; ++i) S.x[i] = 0.0;
    S.pm = 0;
    S.solve = false;
    S.adds = 0;
    st.hit_max = false;
}

// Start of phase 2 for one channel: the FP64 right side Atb = G0 s and the
// FP32 phase's amplitudes as the starting state (passive set = entries > 0).
// Returns false for a channel that is not fitted (non-finite samples, or
// max_minimize <= 0; the flag is set and x stays 0).
template <int NS, int NP>
MF_HD bool fast_p2_begin(const DevType &ty, const float (&s32)[10], const FitParams &prm,
                         const float (&x)[10], double (&Atb64)[10], ActiveState<double> &S,
                         int &flags)
{
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        S.x[i] = 0.0;
        Atb64[i] = 0.0;
    }
    S.pm = 0;
    S.solve = false;
    S.adds = 0;
    if (!all_finite<float, NS>(s32) || !rho_ok(prm.rho)) {
        flags |= F_NONFINITE;
        return false;
    }
    if (prm.max_minimize <= 0) {
        flags |= F_MINIMIZE_CAP;
        return false;
    }
    const TypeMats<double> &td = *opaque(&ty.d);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        double t64 = 0.0;
#pragma unroll
        for (int k = 0; k < NS; ++k) t64 = fma(td.G0[p][k], (double)s32[k], t64);
        Atb64[p] = t64;
        S.x[p] = x[p] > 0.0f ? (double)x[p] : 0.0;
        if (x[p] > 0.0f) S.pm |= 1u << p;
    }
    S.solve = S.pm != 0;
    return true;
}

// chi2 of the phase-2 result, with the count of Minimize iterations when
// the type's pulse covariance is zero or not used: the covariance then never
// depends on x, the reference's iteration 2 rebuilds the same normal
// equations, its FNNLS starts at their optimum and stops, chi2 repeats and
// Minimize ends after 2 iterations with the same amplitudes -- so the fit
// ends here.
template <int NS, int NP>
MF_HD double fast_p2_chi2(const DevType &ty, const float (&s32)[10], const double (&x64)[10])
{
    const TypeMats<double> &td = *opaque(&ty.d);
    double s64[10];   // formed here, not held through the solves
#pragma unroll
    for (int i = 0; i < 10; ++i) s64[i] = (double)s32[i];
    return chisq64<double, double, NS, NP>(td, td.Li0, s64, x64);
}
MF_HD int fixed_cov_nmin(double chi2, const FitParams &prm, int &flags)
{
    if (fabs(chi2) < prm.chi2_tol) return 1;
    const int nmin = prm.max_minimize >= 2 ? 2 : 1;
    if (nmin < 2) flags |= F_MINIMIZE_CAP;
    return nmin;
}

// Phase 2's first Minimize iteration: the NNLS of the x = 0 normal equations
// from the start x (phase 1's amplitudes or the prediction), with the cold
// fallback of a warm start, and its chi2.  Returns false for a channel that
// is not fitted (fast_p2_begin; S.x stays 0, chi2 0).
template <int NS, int NP, bool TAB, typename Reload>
MF_HD bool fast_p2_first(const DevType &ty, float (&s32)[10], const FitParams &prm,
                         const float (&x)[10], bool warm, FnnlsStats &st, int &flags,
                         ActiveState<double> &S, double &chi2, Reload &reload)
{
    double Atb64[10];
    chi2 = 0.0;
    reload(s32);
    if (!fast_p2_begin<NS, NP>(ty, s32, prm, x, Atb64, S, flags)) return false;
    const int outer0 = st.outer;
    const bool from_empty = S.pm == 0;   // a start that is the cold start
    const MaskTab *tab = *opaque(&ty.tab);   // wave-uniform: one scalar load
    const double eps = prm.eps * prm.rho;
    uint32_t polished = 0;
    do {
        while (!fnnls_step<double, NP, true, TAB>(ty.d.AtA0, Atb64, S, eps, prm.eps_rel64,
                                                  prm.max_iter, st, 0.0, nullptr, tab)) {
        }
    } while (TAB && polish_end<NP>(tab, Atb64, S, st, polished));
    flags |= F_REFINED;
    if (warm && st.outer != outer0) flags |= F_WARM_REJECT;
    if (warm && st.hit_max && !from_empty) {
        // The cold start is the fallback: a prediction whose iterations
        // reach max_iter is dropped and the channel fitted from x = 0.
        warm_fallback_reset(S, st);
        flags |= F_WARM_FALLBACK;
        polished = 0;
        do {
            while (!fnnls_step<double, NP, true, TAB>(ty.d.AtA0, Atb64, S, eps, prm.eps_rel64,
                                                      prm.max_iter, st, 0.0, nullptr, tab)) {
            }
        } while (TAB && polish_end<NP>(tab, Atb64, S, st, polished));
    }
    reload(s32);
    chi2 = fast_p2_chi2<NS, NP>(ty, s32, S.x) / prm.rho;
    return true;
}

// The fit after its first iteration: the Minimize count; with a live pulse
// covariance the later iterations run the reference's FP64 iteration body
// (per-channel covariance, inplace_fnnls with warm x and reset passive set).
// They are FP64 because the reference decides them on gradients near its
// 1e-11 threshold.
template <int NS, int NP, bool LIVE, int D, typename Park>
MF_HD int fast_p2_rest(const DevType &ty, const float (&s32)[10], const FitParams &prm,
                       double (&x)[10], double &chi2, FnnlsStats &st, int &flags, Park &park)
{
    if (LIVE && prm.use_pi) {
        if (fabs(chi2) < prm.chi2_tol) return 1;
        double s64[10];
#pragma unroll
        for (int i = 0; i < 10; ++i) s64[i] = (double)s32[i];
        return minimize_from<double, NS, NP, XS_NONE, D, false, true>(ty.d, s64, prm, park, x,
                                                                      chi2, 1, st, flags);
    }
    return fixed_cov_nmin(chi2, prm, flags);
}

// fast_p2_rest in two parts, for kernels that cannot afford the registers of
// the normal equations: fast_p2_rest_lazy runs the later iterations while
// each one stops at its start or adds a single index (minimize_step, LAZY);
// at the first that needs the normal equations it sets pend and returns that
// iteration's index, x and chi2 as at its start, and fast_p2_resume goes on
// from there with the whole iteration body.  Together they give
// fast_p2_rest's results bit for bit.
struct NullPark {
    MF_HD void store(int, double) {}
    MF_HD double load(int) const { return 0.0; }
};
template <int NS, int NP, int D>
MF_HD int fast_p2_rest_lazy(const DevType &ty, const float (&s32)[10], const FitParams &prm,
                            double (&x)[10], double &chi2, FnnlsStats &st, int &flags, bool &pend)
{
    pend = false;
    if (!prm.use_pi) return fixed_cov_nmin(chi2, prm, flags);
    if (fabs(chi2) < prm.chi2_tol) return 1;
    NullPark park;
    return minimize_from<double, NS, NP, XS_NONE, D, false, true, true>(
        ty.d, s32, prm, park, x, chi2, 1, st, flags, -1, nullptr, &pend);
}
template <int NS, int NP, int D, typename Park>
MF_HD int fast_p2_resume(const DevType &ty, const float (&s32)[10], const FitParams &prm,
                         double (&x)[10], double &chi2, int it, FnnlsStats &st, int &flags,
                         Park &park)
{
    double s64[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) s64[i] = (double)s32[i];
    return minimize_from<double, NS, NP, XS_NONE, D, false, true>(ty.d, s64, prm, park, x, chi2,
                                                                  it, st, flags);
}

MF_HD void fast_p2_out(const double (&x)[10], double chi2, int flags, int nmin,
                       const FnnlsStats &st, FitOut<float> &out)
{
    if (st.hit_max) flags |= F_FNNLS_CAP;
    if (st.llt_fail) flags |= F_LLT_FAIL;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        out.x[i] = (float)x[i];
        out.err[i] = 0.0f;
    }
    out.chi2 = (float)chi2;
    out.flags = flags;
    out.n_min = nmin;
    out.n_outer = st.outer;
    out.n_inner = st.inner;
}

// Phase 2 and the rest of the fit (see fit_fast).  x: phase 1's amplitudes
// (passive set = entries > 0).  Fills out.  reload(s32) is called before
// each use of the samples (the start and chi2): a kernel can read them again
// from memory instead of holding them in registers through the solves.
struct KeepSamples {
    MF_HD void operator()(float (&)[10]) const {}
};
template <int NS, int NP, bool LIVE, bool TAB = true, int D = -1, typename Park,
          typename Reload = KeepSamples>
MF_HD void fast_phase2(const DevType &ty, float (&s32)[10], const FitParams &prm,
                       const float (&x)[10], bool warm, FnnlsStats &st, int flags, Park &park,
                       FitOut<float> &out, Reload &&reload = KeepSamples{})
{
    ActiveState<double> S;
    double chi2 = 0.0;
    int nmin = 0;
    if (fast_p2_first<NS, NP, TAB>(ty, s32, prm, x, warm, st, flags, S, chi2, reload))
        nmin = fast_p2_rest<NS, NP, LIVE, D>(ty, s32, prm, S.x, chi2, st, flags, park);
    fast_p2_out(S.x, chi2, flags, nmin, st, out);
}

// The fast fit of one channel in one call: phase 1 from a cold start, then
// phase 2 (a warm start begins at phase 2).
template <int NS, int NP, bool LIVE, bool TAB = true, int D = -1, typename Park>
MF_HD void fit_fast(const DevType &ty, const float (&s32_in)[10], const FitParams &prm,
                    float (&x)[10], uint32_t pm, bool warm, Park &park, FitOut<float> &out)
{
    int flags = 0;
    FnnlsStats st;
    float s32[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) s32[i] = s32_in[i];
    if (!warm && all_finite<float, NS>(s32) && rho_ok(prm.rho) && prm.max_minimize > 0)
        fast_phase1<NS, NP>(ty, s32, prm, x, pm, st, flags);
    fast_phase2<NS, NP, LIVE, TAB, D>(ty, s32, prm, x, warm, st, flags, park, out);
}

// Register-array park for host builds and for kernels that keep L^{-1} live.
template <typename T> struct RegPark {
    T v[55];
    MF_HD void store(int k, T x) { v[k] = x; }
    MF_HD T load(int k) const { return v[k]; }
};

}  // namespace mf
