/*
 * mf_abi.h -- C ABI of the MI355X multifit engine (libmf_amd.so).
 *
 * Plain C, plain pointers and sizes.  Device buffers are HIP device pointers,
 * streams are hipStream_t passed as void* (NULL = the null stream).  Every
 * entry point returns an mf_status; mf_last_error() gives the message of the
 * last failure on the calling thread.  Device memory is allocated only into
 * the context, kept there and freed by mf_context_destroy: the queue heads
 * (mf_context_create), the type tables (mf_set_types), the scratch of
 * mf_fit_batch (type partition, warm-start prediction; grown on demand, never
 * while a graph that uses it exists) and the MLP's parameters and training
 * activations (mf_mlp_*).
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference repository root):
 *   mf_fit_batch ............. doFitWrapper (multifit_gpu/interface/PulseChiSqSNNLSWrapper.h:11,
 *                              multifit_gpu/src/PulseChiSqSNNLSWrapper.cu:8-52) and the
 *                              per-channel PulseChiSqSNNLS::DoFit it runs
 *                              (multifit_cpu/interface/PulseChiSqSNNLS.h:13-18)
 *   mf_inplace_fnnls_batch ... inplace_fnnls / inplace_fnnls_kernel
 *                              (nnls/interface/inplace_fnnls.h:9-14, nnls/interface/kernel_wrapper.h:7-12)
 *   mf_fnnls_batch ........... fnnls / fnnls_kernel / fnnls_wrapper
 *                              (nnls/interface/fnnls.h:9-14, nnls/interface/kernel_wrapper.h:14-24)
 *   mf_inplace_fnnls_chol_batch  inplace_fnnls_test_cholesky
 *                              (nnls/interface/inplace_fnnls_test_cholesky.h:9-14)
 *   mf_lh_nnls_batch ......... Eigen::NNLS<FixedMatrix>::solve (nnls/interface/eigen_nnls.h:51-414)
 *   mf_linalg_batch .......... cholesky_decomp, solve_forward_substitution,
 *                              solve_backward_substitution, swap_row_column,
 *                              fused_cholesky_forward_substitution_solver_{rcaddition,inbetween_removal}
 *                              (linalgebra/interface/{cholesky,substitution,common,fused_cholesky_substitution}.hpp)
 */
#ifndef MF_ABI_H
#define MF_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MF_ABI_VERSION 3
#define MF_MAX_N 10   /* samples and pulses per channel */
#define MF_FULL 19    /* FullSampleVector length (EigenMatrixTypes.h:7) */

typedef enum {
    MF_OK = 0,
    MF_ERR_INVALID = 1,     /* bad argument (sizes, NULL, unsupported shape) */
    MF_ERR_HIP = 2,         /* a HIP runtime call failed */
    MF_ERR_NOMEM = 3,       /* device allocation failed */
    MF_ERR_UNSUPPORTED = 4  /* option combination not built */
} mf_status;

/* One channel type: the per-channel constants of DoFit
 * (PulseChiSqSNNLS.h:13-18: samplecor, pederr, bxs, fullpulse, fullpulsecov).
 * The pulse matrix column of pulse j holds fullpulse[f_j + o_j ...] in rows
 * f_j .. nsamples-1 with f_j = max(0, bx_j + first_off), o_j = pulse_off - bx_j
 * (PulseChiSqSNNLS.cpp:50-62: first_off = 3, pulse_off = 4 for ECAL).
 * npulses is bxs.rows() of the reference (its BXVector is dynamic, at most
 * 10: multifit_cpu/interface/EigenMatrixTypes.h:9-10); amplitudes of the
 * slots npulses .. nsamples-1 come back 0.
 * Matrices are row-major: noisecor[i*10+j], fullpulsecov[i*19+j].
 * Layout is identical to the oracle's or_chtype. */
typedef struct mf_chtype {
    int32_t nsamples;          /* 8 or 10                                  */
    int32_t npulses;           /* 1 .. nsamples                            */
    int32_t bx[MF_MAX_N];      /* BX offset of each pulse                  */
    int32_t first_off;
    int32_t pulse_off;
    double  pederr;            /* noise sigma                              */
    double  noisecor[MF_MAX_N * MF_MAX_N];
    double  fullpulse[MF_FULL];
    double  fullpulsecov[MF_FULL * MF_FULL];
} mf_chtype;

/* precision / algorithm of mf_fit_batch */
enum {
    MF_PREC_FP64 = 0,  /* plain cold-start FP64 fit with the reference's
                          semantics (multifit_gpu: FNNLS max_iter 1000)     */
    MF_PREC_FP32 = 1   /* fast path: active-set exploration in FP32 on the
                          type's precomputed normal equations, finished in
                          FP64 (refinement + KKT verification); later
                          Minimize iterations of a live pulse covariance
                          in FP64                                           */
};
enum {
    MF_SOLVER_INPLACE_FNNLS = 0,  /* nnls/interface/inplace_fnnls.h:9-14 (default)  */
    MF_SOLVER_FNNLS = 1,          /* nnls/interface/fnnls.h:9-14                    */
    MF_SOLVER_LEGACY = 2,         /* legacy_multifit_cpu/src/PulseChiSqSNNLS.cpp:
                                     268-363: the passive set and its order carry
                                     over from one Minimize iteration to the next
                                     (and into the MINOS refits); FP64 only; use
                                     max_iter = 1000 as the legacy code does      */
    MF_SOLVER_LAWSON_HANSON = 3,  /* Eigen::NNLS started at the current amplitudes
                                     (setX), the alternative kept in
                                     multifit_cpu/src/PulseChiSqSNNLS.cpp:253-257;
                                     max_iter bounds least-squares solves; FP64 */
    MF_SOLVER_INPLACE_FNNLS_CHOL = 4  /* inplace_fnnls_test_cholesky, the alternative
                                     kept in PulseChiSqSNNLS.cpp:267: the same
                                     loop with the passive factor updated from
                                     one solve to the next (nnls/src/
                                     inplace_fnnls_test_cholesky.cu:35-227); FP64 */
};
enum { MF_DTYPE_F64 = 0, MF_DTYPE_F32 = 1 };

typedef struct mf_opts {
    int32_t precision;      /* MF_PREC_*                                      */
    int32_t solver;         /* MF_SOLVER_* (FP32 path: inplace_fnnls only)    */
    int32_t max_iter;       /* FNNLS outer cap (reference GPU: 1000, CPU: 10) */
    int32_t max_minimize;   /* Minimize cap (reference: 50)                   */
    double  eps;            /* FNNLS max-w threshold (reference: 1e-11)       */
    double  chi2_tol;       /* Minimize |dchi2| threshold (reference: 1e-3)   */
    int32_t use_pulse_cov;  /* 1: add x_j^2 * fullpulsecov blocks (updateCov) */
    int32_t warm_start;     /* FP32 only: 1 = start from the MLP's prediction
                               (mf_mlp_*; every result is still KKT-verified) */
    int32_t persistent;     /* 1: persistent kernel + device work queue;
                               MF_PERSISTENT_AUTO: as mf_autotune chose       */
    int32_t block;          /* threads per block (0 = tuned default)          */
    int32_t compute_errors; /* FP64 only: MINOS-like in-time uncertainty
                               (PulseChiSqSNNLS.cpp:73-146) into mf_result.err */
    int32_t pad;
} mf_opts;

/* Per-channel flags (mf_result.flags). */
enum {
    MF_F_MINIMIZE_CAP = 1,   /* Minimize ran max_minimize iterations          */
    MF_F_LLT_FAIL     = 2,   /* a Cholesky pivot was <= 0 or NaN              */
    MF_F_FNNLS_CAP    = 4,   /* an FNNLS call hit max_iter                    */
    MF_F_NONFINITE    = 8,   /* non-finite input sample, or a per-channel pederr
                                whose square is not positive and finite: not
                                fitted, x = 0                                 */
    MF_F_FP64_PIVOT   = 16,  /* FP32 path: a pivot was recomputed in FP64     */
    MF_F_WARM_REJECT  = 32,  /* warm start failed its KKT check; the active-set
                                iterations went on from it                    */
    MF_F_REFINED      = 64,  /* FP32 path: FP64 refinement step applied       */
    MF_F_BAD_TYPE     = 128, /* type_of[c] >= ntypes: not fitted, x = 0       */
    MF_F_WARM_FALLBACK = 256 /* the iterations from a non-empty prediction
                                reached max_iter: the prediction was dropped
                                and the channel fitted from a cold start
                                (x = 0; an empty prediction is that start)    */
};

/* Device-resident batch, structure of arrays.  samples[i*nch + c] for sample
 * i of channel c (dtype MF_DTYPE_*).  With MF_PREC_FP32 the samples are
 * used rounded to float32 (also FP64 samples, in both of the fast path's
 * phases): its accuracy figures are those against the FP64 fit of the
 * rounded samples.  pederr (device, optional): the noise sigma of each
 * channel, the pederr argument of DoFit, replacing the type's (NULL: the
 * type's pederr for every channel); the noise covariance is then
 * pederr[c]^2 / type.pederr^2 times the type's, formed exactly as
 * pederr^2 * noisecor when the type's pederr is 1.  The channel types, one of:
 *   type_of (device, one byte per channel): index into the table passed to
 *     mf_set_types; the batch is partitioned by type on the device;
 *   ranges (host, nranges + 1 ascending channel numbers): channels
 *     [ranges[r], ranges[r+1]) are of type r -- the layout of a readout
 *     ordered by subdetector; no partition is needed and one launch covers
 *     every range of the same shape (type_of must then be NULL);
 *   neither: every channel is type 0. */
typedef struct mf_batch {
    int64_t        nch;
    int32_t        sample_dtype;
    int32_t        nranges;
    const void    *samples;
    const uint8_t *type_of;
    const int64_t *ranges;
    const double  *pederr;
} mf_batch;

/* Device-resident results, structure of arrays, in the compute precision
 * (double for MF_PREC_FP64, float for MF_PREC_FP32): x[p*nch + c] amplitude
 * of pulse p (pulse order of the channel's type, the BXs() order of the
 * reference; all 10 rows are written, 0 past the type's npulses),
 * chi2[c], flags[c] (MF_F_*), and optionally iters[c] =
 * n_minimize | n_fnnls_outer << 8 | n_fnnls_inner << 16 (NULL: skipped) and
 * err[p*nch + c] (Errors(): only the in-time pulse's entry is nonzero; needed
 * when compute_errors is set, else ignored). */
typedef struct mf_result {
    void     *x;
    void     *chi2;
    int32_t  *flags;
    int32_t  *iters;
    void     *err;
} mf_result;

typedef struct mf_context mf_context;

const char *mf_last_error(void);
int mf_abi_version(void);

mf_status mf_context_create(mf_context **ctx, int device);
mf_status mf_context_destroy(mf_context *ctx);
/* Copies ntypes channel types (host memory) to the device.  nsamples must be
 * 10 or 8, npulses 1..nsamples.  Refused while a graph of the context exists
 * (its kernels read the tables). */
mf_status mf_set_types(mf_context *ctx, const mf_chtype *types, int32_t ntypes);
mf_status mf_default_opts(mf_opts *opts);

/* Fits every channel of the batch: DoFit, by default with error computation
 * disabled as every reference driver runs it (multifit_cpu/src/multifit.cpp:200,
 * multifit_gpu/src/PulseChiSqSNNLS.cu:234).  Asynchronous on `stream`, with no
 * host round trip (a batch with type_of is partitioned by type on the device),
 * so that a call can be captured in a graph once its buffers exist (one call
 * of the same size beforehand allocates them). */
mf_status mf_fit_batch(mf_context *ctx, const mf_batch *in, const mf_result *out,
                       const mf_opts *opts, void *stream);

/* mf_fit_batch that allocates nothing (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  The same fit and the
 * same results as mf_fit_batch; its scratch (type-partition lists and counts,
 * warm-start prediction, persistent-queue heads) lives in a fixed device
 * array of MF_NOALLOC_SCRATCH_BYTES that each GPU holds from the first
 * mf_context_create on it until the process ends.  Calls whose scratch would
 * not fit are refused with MF_ERR_INVALID: 4 bytes per channel with type_of
 * (plus 1 MiB), 44 per channel for a warm start that runs its prediction as
 * a separate kernel (live pulse covariance or persistent kernels); about
 * 16.5M channels with type_of alone, 1.37M with both.  Every such call on a
 * GPU shares that GPU's array, whatever its context: calls that overlap on
 * one GPU (other streams, other threads, other contexts) race. */
#define MF_NOALLOC_SCRATCH_BYTES (64u << 20)
mf_status mf_fit_batch_noalloc(mf_context *ctx, const mf_batch *in, const mf_result *out,
                               const mf_opts *opts, void *stream);

/* A fit call captured in a HIP graph, for launch-bound repeats such as one
 * event after another (no counterpart in the reference).  mf_graph_capture
 * runs the call once eagerly (which sizes the context's scratch and checks
 * the arguments), then records it; each mf_graph_launch repeats it on the
 * same buffers, reading their current contents.  The buffers must outlive
 * the graph, and the graphs must be destroyed before their context.  Graphs
 * of one context share its scratch: do not run two of them, or a graph and
 * an eager call, at the same time; while a graph exists, a call that would
 * grow the scratch that graph uses (more channels than any call before) is
 * refused. */
typedef struct mf_graph mf_graph;
mf_status mf_graph_capture(mf_context *ctx, const mf_batch *in, const mf_result *out,
                           const mf_opts *opts, void *stream, mf_graph **graph);
mf_status mf_graph_launch(mf_graph *graph, void *stream);
mf_status mf_graph_destroy(mf_graph *graph);

/* Opt-in timing of mf_fit_batch's kernels with HIP events recorded on the
 * launch stream (no counterpart in the reference; used by the benchmark).
 * mf_get_timings waits for the events of the last timed mf_fit_batch. */
typedef struct mf_timings {
    float mlp_ms;     /* warm-start inference kernel                       */
    float fit_ms;     /* fit kernel(s), summed over the type partitions    */
    int32_t fit_launches;
    int32_t pad;
} mf_timings;
mf_status mf_set_timing(mf_context *ctx, int32_t enable);
mf_status mf_get_timings(mf_context *ctx, mf_timings *out);

/* Launch shape of the persistent fast-path kernels (opts.persistent = 1; no
 * counterpart in the reference).  A lane that finishes its channel takes the
 * next one from a device work queue; the grid holds a fixed number of
 * resident wavefronts per compute unit. */
typedef struct mf_tune {
    int32_t block;      /* threads per block: 64, 128, 192 or 256             */
    int32_t p1_wpc;     /* resident wavefronts per CU, FP32 phase (1..32)     */
    int32_t p2_wpc;     /* resident wavefronts per CU, FP64 phase (1..32)     */
    int32_t p1_refill;  /* idle lanes that trigger a refill, FP32 phase (1..64) */
    int32_t p2_refill;  /* idle lanes that trigger a refill, FP64 phase (1..64) */
    int32_t fit_block;  /* threads per block of the per-channel kernels when
                           opts.block == 0 (64, 128, 192 or 256)            */
    int32_t use_persistent; /* 1: opts.persistent == MF_PERSISTENT_AUTO runs
                               the persistent kernels                      */
    int32_t pad;
} mf_tune;
#define MF_PERSISTENT_AUTO 2   /* opts.persistent: as the context's tune says */
mf_status mf_get_tune(mf_context *ctx, mf_tune *out);
mf_status mf_set_tune(mf_context *ctx, const mf_tune *tune);

/* Autotuner: fits `in` (into `out`) under each candidate launch shape --
 * per-channel kernels at 64, 128 and 256 threads per block, persistent
 * kernels at several resident-wave counts and refill thresholds -- timing
 * each with HIP events on `stream` (best of 3 after a warm-up), and keeps the
 * fastest in the context (mf_get_tune).  *best and *ms (optional) receive
 * the choice and its time per call.  Synchronizes the stream. */
mf_status mf_autotune(mf_context *ctx, const mf_batch *in, const mf_result *out,
                      const mf_opts *opts, void *stream, mf_tune *best, float *ms);

/* Batched NNLS solvers, one problem per channel, FP64.
 * A: n matrices of m x m, column-major as the reference's FixedMatrix
 * (nnls/interface/data_types.h:25-26), A[c*100 + col*10 + row] (leading
 * dimension 10 whatever m is); b[c*10 + i]; x[c*10 + i] in/out (warm start
 * in, solution out).  m <= 10; rows/cols >= m are ignored. */
mf_status mf_inplace_fnnls_batch(const double *A, const double *b, double *x,
                                 int64_t n, int32_t m, double eps,
                                 uint32_t max_iter, void *stream);
mf_status mf_fnnls_batch(const double *A, const double *b, double *x,
                         int64_t n, int32_t m, double eps, uint32_t max_iter,
                         void *stream);
mf_status mf_inplace_fnnls_chol_batch(const double *A, const double *b, double *x,
                                      int64_t n, int32_t m, double eps, uint32_t max_iter,
                                      void *stream);
/* Lawson-Hanson NNLS of the reference's Eigen::NNLS class
 * (nnls/interface/eigen_nnls.h:84-98 constructor, :124 setX, :336-414
 * solve, as tests/main.cpp:53-58 calls it).  Same layout as above; x in is
 * the setX starting point.  max_iter bounds the least-squares solves; <= 0:
 * none (the solver then stops after 1000 solves, reported as not solved).
 * solved (optional, n entries): 1 where solve() returned true. */
mf_status mf_lh_nnls_batch(const double *A, const double *b, double *x,
                           int64_t n, int32_t m, double eps, int32_t max_iter,
                           int32_t *solved, void *stream);

/* Batched linalgebra primitives on n row-major matrices of full x full
 * (full <= 10), operating on the leading view x view block, FP64 or FP32
 * (dtype MF_DTYPE_*).  Arguments by op (per problem c; M, L: full*full
 * elements at c*full*full; vectors: full elements at c*full):
 *   MF_LA_CHOLESKY        M -> L                 (cholesky_decomp)
 *   MF_LA_FORWARD         L, b -> y              (solve_forward_substitution)
 *   MF_LA_BACKWARD        L, b(=y) -> y(=x)      (solve_backward_substitution)
 *   MF_LA_SWAP            M in place, i, j       (swap_row_column, j > i)
 *   MF_LA_RCADDITION      M, L in/out, b, y in/out
 *   MF_LA_INBETWEEN       M, L in/out, b, y in/out, i = position */
enum {
    MF_LA_CHOLESKY = 0, MF_LA_FORWARD = 1, MF_LA_BACKWARD = 2, MF_LA_SWAP = 3,
    MF_LA_RCADDITION = 4, MF_LA_INBETWEEN = 5
};
mf_status mf_linalg_batch(int32_t op, int32_t dtype, void *M, void *L, void *b,
                          void *y, int64_t n, int32_t full, int32_t view,
                          int32_t i, int32_t j, void *stream);

/* ---- learned warm start ---------------------------------------------------
 * No counterpart in the reference: a 10 -> 64 -> 64 -> 20 MLP (GELU) predicts
 * a channel's passive set (logits 0..9 > 0) and amplitudes (10..19) from its
 * samples scaled by the largest sample; mf_fit_batch with opts.warm_start = 1
 * starts the fast path's active-set solve there.  Parameters: float32
 * [W1 10x64, b1 64, W2 64x64, b2 64, W3 64x20, b3 20] row-major [in][out],
 * used rounded to bfloat16 with float32 accumulation.  Device pointers
 * unless stated; `samples` / `xref` are [10][ld] float32 (row i at i*ld). */
#define MF_MLP_NPARAM 6164
/* Xavier-uniform weights, zero biases, Adam state reset (seeded). */
mf_status mf_mlp_init(mf_context *ctx, uint64_t seed);
/* host arrays of MF_MLP_NPARAM floats; set also resets the Adam state */
mf_status mf_mlp_set_params(mf_context *ctx, const float *params, int32_t n);
mf_status mf_mlp_get_params(mf_context *ctx, float *params, int32_t n);
/* gradient of the mean loss over B channels (grad: MF_MLP_NPARAM floats,
 * loss: 1 float, may be NULL); loss per channel = mean_i BCE(logit_i,
 * [xref_i > 0]) + mean_i (amp_i / m - xref_i / m)^2 */
mf_status mf_mlp_grad(mf_context *ctx, const float *samples, int64_t ld, const float *xref,
                      int64_t ldx, int64_t B, float *grad, float *loss, void *stream);
/* one Adam step (lr) on that gradient; loss may be NULL */
mf_status mf_mlp_train_step(mf_context *ctx, const float *samples, int64_t ld,
                            const float *xref, int64_t ldx, int64_t B, float lr, float *loss,
                            void *stream);
/* inference for a batch: out [20][nch] float32 (logits, then amplitudes) */
mf_status mf_mlp_predict(mf_context *ctx, const mf_batch *in, float *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* MF_ABI_H */
