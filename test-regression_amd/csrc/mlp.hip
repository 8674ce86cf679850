// mlp.hip -- the learned warm start: a 10 -> 64 -> 64 -> 20 MLP (GELU) that
// predicts a channel's passive set (10 logits) and amplitudes (10 values,
// softplus, in units of the channel's largest sample) from its samples.
//
// Parameters: flat float32 [W1 10x64, b1 64, W2 64x64, b2 64, W3 64x20, b3 20],
// row-major [in][out], 6164 values (FP32 masters for Adam).  The forward pass
// uses them rounded to bfloat16 and accumulates in float32.
//
// Training kernels (FP32 VALU, batch-major activations [B][N]):
//   k_mlp_features  samples -> X [B][16] (s_i / max(max_i s_i, 1), zero-padded), m [B]
//   k_gemm          C (+)= op(A) op(B), 64x64 tiles through LDS, split-K with
//                   atomics for the weight gradients (K = batch)
//   k_bias_gelu     Z += b; H = gelu(Z)   (tanh form)
//   k_head_loss     logits / softplus amplitudes, BCE + MSE loss, dZ3
//   k_gelu_bwd      dZ = dH * gelu'(Z)
//   k_colsum        bias gradients
//   k_adam          Adam on the masters, then the bfloat16 copies
// Inference kernel (MFMA): k_mlp_infer, one wavefront per 64 channels, the
// three layers as v_mfma_f32_16x16x16_bf16 products in transposed form
// (Z^T = W^T H^T): the output fragment of one layer is, after bias and GELU,
// exactly the B-operand fragment of the next, so activations never leave
// registers.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp.hpp"
#include "mlp_device.hpp"

namespace mf {
namespace mlp {

// d/dx [x sigma(2u)] = sigma + x sigma (1 - sigma) 2u'(x)
__device__ __forceinline__ float gelu_grad(float x)
{
    const float c2 = 1.5957691216057308f;
    const float sg = sig2u(x);
    return sg + x * sg * (1.0f - sg) * c2 * fmaf(3.0f * 0.044715f, x * x, 1.0f);
}

__device__ __forceinline__ float softplusf(float x)
{
    return fmaxf(x, 0.0f) + __logf(1.0f + __expf(-fabsf(x)));
}

// bfloat16 round to nearest even (NaN not expected: weights are finite)
__host__ __device__ __forceinline__ uint16_t f2bf(float f)
{
    union {
        float f;
        uint32_t u;
    } v;
    v.f = f;
    const uint32_t r = v.u + 0x7FFFu + ((v.u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t b)
{
    union {
        float f;
        uint32_t u;
    } v;
    v.u = (uint32_t)b << 16;
    return v.f;
}

template <typename TS>
__global__ void k_mlp_features(const TS *__restrict__ s, int64_t ld, int64_t B, int ns,
                               float *__restrict__ X, float *__restrict__ m)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float v[10], mx = -3.0e38f;
    for (int i = 0; i < 10; ++i) {
        v[i] = i < ns ? (float)s[(int64_t)i * ld + b] : 0.0f;
        if (i < ns) mx = fmaxf(mx, v[i]);
    }
    const float sc = fmaxf(mx, 1.0f);
    for (int i = 0; i < 16; ++i) X[b * 16 + i] = i < 10 ? v[i] / sc : 0.0f;
    m[b] = sc;
}

// C[M x N] = op(A) op(B) (or += atomically when atomic_acc).  op(A)(i, k) =
// TA ? A[k*lda + i] : A[i*lda + k]; op(B)(k, j) = TB ? B[j*ldb + k] : B[k*ldb + j].
// Grid (ceil(N/64), ceil(M/64), splits); split z covers K in [z*kc, (z+1)*kc).
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float *__restrict__ A,
                                              int lda, const float *__restrict__ Bm, int ldb,
                                              float *__restrict__ C, int ldc, int kc,
                                              int atomic_acc)
{
    __shared__ float As[16][65], Bs[16][65];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int kb = blockIdx.z * kc, ke = min(K, kb + kc);
    float acc[4][4] = {{0}};
    for (int k0 = kb; k0 < ke; k0 += 16) {
        for (int e = tid; e < 1024; e += 256) {
            const int mm = e / 16, kk = e % 16;
            const int gi = m0 + mm, gk = k0 + kk;
            float a = 0.0f;
            if (gi < M && gk < ke) a = TA ? A[(int64_t)gk * lda + gi] : A[(int64_t)gi * lda + gk];
            As[kk][mm] = a;
            const int nn = e / 16, kk2 = e % 16;
            const int gj = n0 + nn, gk2 = k0 + kk2;
            float bv = 0.0f;
            if (gj < N && gk2 < ke)
                bv = TB ? Bm[(int64_t)gj * ldb + gk2] : Bm[(int64_t)gk2 * ldb + gj];
            Bs[kk2][nn] = bv;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], bb[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = As[kk][ty * 4 + r];
#pragma unroll
            for (int q = 0; q < 4; ++q) bb[q] = Bs[kk][tx * 4 + q];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[r][q] = fmaf(a[r], bb[q], acc[r][q]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int gi = m0 + ty * 4 + r, gj = n0 + tx * 4 + q;
            if (gi < M && gj < N) {
                if (atomic_acc)
                    atomicAdd(&C[(int64_t)gi * ldc + gj], acc[r][q]);
                else
                    C[(int64_t)gi * ldc + gj] = acc[r][q];
            }
        }
}

__global__ void k_bias_gelu(float *__restrict__ Z, float *__restrict__ H,
                            const float *__restrict__ bias, int64_t B, int N)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= B * N) return;
    const float z = Z[e] + bias[e % N];
    Z[e] = z;
    H[e] = gelu(z);
}

__global__ void k_gelu_bwd(float *__restrict__ dH, const float *__restrict__ Z, int64_t n)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    dH[e] *= gelu_grad(Z[e]);
}

// z3 += b3; loss and dZ3 (mean over the batch).  xref: [10][ldx] float.
__global__ void k_head_loss(float *__restrict__ Z3, const float *__restrict__ b3,
                            const float *__restrict__ xref, int64_t ldx,
                            const float *__restrict__ m, int64_t B, float *__restrict__ loss)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (b < B) {
        const float invB = 1.0f / (float)B, sc = m[b];
        for (int i = 0; i < 10; ++i) {
            const float lg = Z3[b * 20 + i] + b3[i];
            const float ar = Z3[b * 20 + 10 + i] + b3[10 + i];
            const float xr = xref[(int64_t)i * ldx + b];
            const float y = xr > 0.0f ? 1.0f : 0.0f, t = xr / sc;
            const float a = softplusf(ar);
            const float bce = fmaxf(lg, 0.0f) - lg * y + log1pf(expf(-fabsf(lg)));
            l += 0.1f * (bce + (a - t) * (a - t));
            const float sig = 1.0f / (1.0f + expf(-lg));
            Z3[b * 20 + i] = 0.1f * (sig - y) * invB;
            Z3[b * 20 + 10 + i] = 0.2f * (a - t) / (1.0f + expf(-ar)) * invB;
        }
        l *= invB;
    }
    // wave-level reduction, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

__global__ void k_colsum(const float *__restrict__ D, int64_t B, int N, float *__restrict__ out)
{
    const int j = blockIdx.y;
    float s = 0.0f;
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < B;
         b += (int64_t)gridDim.x * blockDim.x)
        s += D[b * N + j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(&out[j], s);
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ mo,
                       float *__restrict__ vo, int n, float lr, float c1, float c2,
                       float *__restrict__ pq, uint16_t *__restrict__ pbf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gi = g[i];
    const float m = b1 * mo[i] + (1.0f - b1) * gi;
    const float v = b2 * vo[i] + (1.0f - b2) * gi * gi;
    mo[i] = m;
    vo[i] = v;
    const float w = p[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    p[i] = w;
    const uint16_t bq = f2bf(w);
    pq[i] = bf2f(bq);
    pbf[i] = bq;
}

__global__ void k_zero(float *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_quantize(const float *__restrict__ p, int n, float *__restrict__ pq,
                           uint16_t *__restrict__ pbf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint16_t bq = f2bf(p[i]);
    pq[i] = bf2f(bq);
    pbf[i] = bq;
}

__global__ void k_mlp_pack(const uint16_t *__restrict__ pbf, uint16_t *__restrict__ frag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 16 * 64) return;
    const int f = t / 64, lane = t % 64, g = lane >> 4, n = lane & 15;
    if (f < 4) {
        const int o = f * 16 + n;
        for (int j = 0; j < 4; ++j) {
            const int k = 4 * g + j;
            frag[FRAG_W1 + f * 256 + lane * 4 + j] = k < 10 ? pbf[k * 64 + o] : (uint16_t)0;
        }
        return;
    }
    const bool l2 = f < 12;
    const int q = l2 ? f - 4 : f - 12, mt = q / 2, kb = q % 2;
    const uint16_t *W = l2 ? pbf + 704 : pbf + 704 + 4160;
    const int nout = l2 ? 64 : 20, o = mt * 16 + n;
    uint16_t *dst = frag + (l2 ? FRAG_W2 : FRAG_W3) + q * 512 + lane * 8;
    for (int i = 0; i < 8; ++i) {
        const int h = 32 * kb + 16 * (i / 4) + 4 * g + i % 4;
        dst[i] = o < nout ? W[h * nout + o] : (uint16_t)0;
    }
}

// out: [20][ldo] float: logits 0..9, then amplitudes (softplus * m) 10..19,
// column = position in the channel list.  Grid-stride over 64-channel
// wavefront tiles; the network is staged in LDS once per block.
constexpr int INFER_BLOCK = 256;
template <typename TS>
__global__ __launch_bounds__(INFER_BLOCK) void k_mlp_infer(
    const TS *__restrict__ s, int64_t ld, int ns, const int32_t *__restrict__ idx, int64_t n,
    const uint16_t *__restrict__ frag, const float *__restrict__ pq, float *__restrict__ out,
    int64_t ldo)
{
    __shared__ MlpLds L;
    __shared__ __bf16 xs[INFER_BLOCK / 64][64][16];
    __shared__ float ms[INFER_BLOCK / 64][64];
    mlp_stage(L, frag, pq, threadIdx.x, INFER_BLOCK);
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t base = ((int64_t)blockIdx.x * (INFER_BLOCK / 64) + wv) * 64; base < n;
         base += (int64_t)gridDim.x * INFER_BLOCK) {
        const int64_t g = base + lane;
        const bool live = g < n;
        const int64_t c = live ? (idx ? (int64_t)idx[g] : g) : 0;
        float v[10];
#pragma unroll
        for (int i = 0; i < 10; ++i) v[i] = (live && i < ns) ? (float)s[(int64_t)i * ld + c] : 0.0f;
        ms[wv][lane] = mlp_features(v, ns, xs[wv], lane);
        // the tiles read other lanes' features: LDS operations of a wavefront
        // complete in order, the fence keeps the compiler from reordering
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        mlp_tiles(L, xs[wv], lane, [&](int t, int o, float z) {
            const int64_t gt = base + t;
            if (gt < n) out[(int64_t)o * ldo + gt] = o < 10 ? z : softplusf(z) * ms[wv][t];
        });
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Warm-start form of the prediction, for the fit kernels: pm[c] = the
// passive-set bits {i: logit_i > 0}, px[i][c] = the predicted amplitude
// (softplus * m) of pulse i, for channels c = 0..n-1.
// The tiles leave each output with the lane that computed it (4 lanes per
// channel); the outputs go through LDS (pz: softplus by pulse and channel,
// pmw: logit bits by channel) so that each lane stores its own channel's 10
// amplitudes as 10 coalesced rows.  The per-output work is branch-free.
// Channels c0 .. c0+n-1; a channel's sample count is ns, or, with type_of,
// that of its type (nst).
template <typename TS>
__global__ __launch_bounds__(INFER_BLOCK) void k_mlp_warm(
    const TS *__restrict__ s, int64_t ld, int ns, int64_t c0, int64_t n,
    const uint8_t *__restrict__ type_of, NsOfType nst, const uint16_t *__restrict__ frag,
    const float *__restrict__ pq, float *__restrict__ px, uint32_t *__restrict__ ppm)
{
    __shared__ MlpLds L;
    __shared__ __bf16 xs[INFER_BLOCK / 64][64][16];
    __shared__ float pz[INFER_BLOCK / 64][10][64];
    __shared__ uint32_t pmw[INFER_BLOCK / 64][64];
    mlp_stage(L, frag, pq, threadIdx.x, INFER_BLOCK);
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t base = ((int64_t)blockIdx.x * (INFER_BLOCK / 64) + wv) * 64; base < n;
         base += (int64_t)gridDim.x * INFER_BLOCK) {
        const bool live = base + lane < n;
        const int64_t c = c0 + base + lane;
        int nsl = ns;
        if (type_of && live) nsl = nst.samples_of(type_of[c]);
        float v[10], x[10];
#pragma unroll
        for (int i = 0; i < 10; ++i) v[i] = (live && i < nsl) ? (float)s[(int64_t)i * ld + c] : 0.0f;
        const uint32_t pm = mlp_warm_wave(L, xs[wv], pz[wv], pmw[wv], lane, v, nsl, x);
        if (live) {
            ppm[c] = pm;
#pragma unroll
            for (int i = 0; i < 10; ++i) px[(int64_t)i * ld + c] = x[i];
        }
    }
}

// ------------------------------------------------------------------ host side
static int infer_grid(int64_t n)
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
    }
    const int64_t need = (n + INFER_BLOCK - 1) / INFER_BLOCK;
    return (int)std::min<int64_t>(need, (int64_t)cus * 8);
}

template <typename TS>
void launch_infer(const void *s, int64_t ld, int ns, const int32_t *idx, int64_t n,
                  const uint16_t *pbf, const float *pq, float *out, int64_t ldo, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_infer<TS>, dim3((unsigned)infer_grid(n)), dim3(INFER_BLOCK), 0, st,
                       static_cast<const TS *>(s), ld, ns, idx, n, pbf, pq, out, ldo);
}
template void launch_infer<float>(const void *, int64_t, int, const int32_t *, int64_t,
                                  const uint16_t *, const float *, float *, int64_t, hipStream_t);
template void launch_infer<double>(const void *, int64_t, int, const int32_t *, int64_t,
                                   const uint16_t *, const float *, float *, int64_t, hipStream_t);

template <typename TS>
void launch_warm(const void *s, int64_t ld, int ns, int64_t c0, int64_t n,
                 const uint8_t *type_of, const NsOfType &nst, const uint16_t *frag,
                 const float *pq, float *px, uint32_t *ppm, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_warm<TS>, dim3((unsigned)infer_grid(n)), dim3(INFER_BLOCK), 0, st,
                       static_cast<const TS *>(s), ld, ns, c0, n, type_of, nst, frag, pq, px, ppm);
}
template void launch_warm<float>(const void *, int64_t, int, int64_t, int64_t, const uint8_t *,
                                 const NsOfType &, const uint16_t *, const float *, float *,
                                 uint32_t *, hipStream_t);
template void launch_warm<double>(const void *, int64_t, int, int64_t, int64_t, const uint8_t *,
                                  const NsOfType &, const uint16_t *, const float *, float *,
                                  uint32_t *, hipStream_t);

static void gemm(bool ta, bool tb, int M, int N, int K, const float *A, int lda, const float *B,
                 int ldb, float *C, int ldc, int splits, hipStream_t st)
{
    const int kc = ((K + splits - 1) / splits + 15) / 16 * 16;
    splits = (K + kc - 1) / kc;
    const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64), (unsigned)splits);
    const int at = splits > 1;
    if (!ta && !tb)
        hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb,
                           C, ldc, kc, at);
    else if (ta && !tb)
        hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb,
                           C, ldc, kc, at);
    else
        hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb,
                           C, ldc, kc, at);
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t grad_step(Scratch &w, const float *samples, int64_t lds, const float *xref,
                     int64_t ldx, int64_t B, const float *pq, float *grad, float *loss,
                     hipStream_t st)
{
    const float *W1 = pq, *b1 = pq + 640, *W2 = pq + 704, *b2 = pq + 704 + 4096,
                *W3 = pq + 704 + 4160, *b3 = pq + 704 + 4160 + 1280;
    float *gW1 = grad, *gb1 = grad + 640, *gW2 = grad + 704, *gb2 = grad + 704 + 4096,
          *gW3 = grad + 704 + 4160, *gb3 = grad + 704 + 4160 + 1280;
    // zero with kernels, not memset nodes, so that the step can be captured
    // in a graph and replayed
    (void)zero(grad, NPARAM, st);
    if (loss) (void)zero(loss, 1, st);
    hipLaunchKernelGGL(k_mlp_features<float>, dim3(nblk(B, 256)), dim3(256), 0, st, samples, lds,
                       B, 10, w.X, w.m);
    const int splits = (int)std::min<int64_t>(64, std::max<int64_t>(1, B / 256));
    // forward
    gemm(false, false, (int)B, 64, 10, w.X, 16, W1, 64, w.Z1, 64, 1, st);
    hipLaunchKernelGGL(k_bias_gelu, dim3(nblk(B * 64, 256)), dim3(256), 0, st, w.Z1, w.H1, b1, B,
                       64);
    gemm(false, false, (int)B, 64, 64, w.H1, 64, W2, 64, w.Z2, 64, 1, st);
    hipLaunchKernelGGL(k_bias_gelu, dim3(nblk(B * 64, 256)), dim3(256), 0, st, w.Z2, w.H2, b2, B,
                       64);
    gemm(false, false, (int)B, 20, 64, w.H2, 64, W3, 20, w.Z3, 20, 1, st);
    hipLaunchKernelGGL(k_head_loss, dim3(nblk(B, 256)), dim3(256), 0, st, w.Z3, b3, xref, ldx,
                       w.m, B, loss ? loss : w.lossdummy);
    // backward (w.Z3 now holds dZ3)
    gemm(true, false, 64, 20, (int)B, w.H2, 64, w.Z3, 20, gW3, 20, splits, st);
    hipLaunchKernelGGL(k_colsum, dim3(64, 20), dim3(256), 0, st, w.Z3, B, 20, gb3);
    gemm(false, true, (int)B, 64, 20, w.Z3, 20, W3, 20, w.H2, 64, 1, st);  // dH2 over H2
    hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(B * 64, 256)), dim3(256), 0, st, w.H2, w.Z2, B * 64);
    gemm(true, false, 64, 64, (int)B, w.H1, 64, w.H2, 64, gW2, 64, splits, st);
    hipLaunchKernelGGL(k_colsum, dim3(64, 64), dim3(256), 0, st, w.H2, B, 64, gb2);
    gemm(false, true, (int)B, 64, 64, w.H2, 64, W2, 64, w.H1, 64, 1, st);  // dH1 over H1
    hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(B * 64, 256)), dim3(256), 0, st, w.H1, w.Z1, B * 64);
    gemm(true, false, 10, 64, (int)B, w.X, 16, w.H1, 64, gW1, 64, splits, st);
    hipLaunchKernelGGL(k_colsum, dim3(64, 64), dim3(256), 0, st, w.H1, B, 64, gb1);
    return hipGetLastError();
}

hipError_t adam_step(float *p, const float *g, float *m, float *v, int t, float lr, float *pq,
                     uint16_t *pbf, hipStream_t st)
{
    const float c1 = 1.0f - powf(0.9f, (float)t), c2 = 1.0f - powf(0.999f, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(nblk(NPARAM, 256)), dim3(256), 0, st, p, g, m, v, NPARAM, lr,
                       c1, c2, pq, pbf);
    return hipGetLastError();
}

hipError_t pack(const uint16_t *pbf, uint16_t *frag, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_pack, dim3(nblk(16 * 64, 256)), dim3(256), 0, st, pbf, frag);
    return hipGetLastError();
}

hipError_t zero(float *p, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero, dim3(nblk(n, 256)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t quantize(const float *p, float *pq, uint16_t *pbf, hipStream_t st)
{
    hipLaunchKernelGGL(k_quantize, dim3(nblk(NPARAM, 256)), dim3(256), 0, st, p, NPARAM, pq, pbf);
    return hipGetLastError();
}

}  // namespace mlp
}  // namespace mf
