// mlp_device.hpp -- device side of the warm-start MLP's inference, shared by
// the MLP kernels (mlp.hip) and the fit kernel that runs the prediction in
// place (mf_kernels.hip): activation functions, the weight-fragment layout,
// the LDS image of the network and the per-wavefront forward pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp.hpp"

namespace mf {
namespace mlp {

// GELU, tanh form: 0.5 x (1 + tanh(u)), u = sqrt(2/pi) (x + 0.044715 x^3).
// Since (1 + tanh(u)) / 2 = sigmoid(2u) this is x / (1 + exp(-2u)), one
// hardware exp and one hardware reciprocal.
__device__ __forceinline__ float sig2u(float x)
{
    // 1 / (1 + exp(-2u)) = 1 / (1 + 2^(x (A + B x^2))), A = -2 sqrt(2/pi) log2(e),
    // B = 0.044715 A: one v_exp_f32 and one v_rcp_f32
    const float A = -1.5957691216057308f * 1.4426950408889634f;
    const float B = A * 0.044715f;
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(x * fmaf(B, x * x, A)));
}

__device__ __forceinline__ float gelu(float x) { return x * sig2u(x); }

// ------------------------------------------------------------------ inference
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even, in hardware (v_cvt_pk_bf16_f32)
__device__ __forceinline__ bf16x4 cvt4(f32x4 v) { return __builtin_convertvector(v, bf16x4); }

// Weight fragments, packed once per parameter update (k_mlp_pack) and read by
// the inference kernels from LDS.  Lane l = 16 g + n.
//   layer 1 (v_mfma_f32_16x16x16_bf16), A = W1^T tile mt: lane l holds
//     A[n][4g + j] = W1[4g + j][16 mt + n], j < 4 (0 for inputs >= 10);
//   layers 2, 3 (v_mfma_f32_16x16x32_bf16), A = W^T tile (mt, kb): lane l
//     holds A[n][8g + i] = W[h][16 mt + n], i < 8, with the hidden units of K
//     block kb taken in the order h(kb, g, i) = 32 kb + 16 (i / 4) + 4 g + i % 4.
// That order is the one in which a lane already holds them: the D fragment
// of output tile mt gives lane l rows 16 mt + 4 g + j, so the D fragments of
// tiles 2 kb and 2 kb + 1, after bias and GELU, are the B operand of K block
// kb as they lie (a contraction does not depend on the order of its index).
// Offsets (bfloat16 elements): W1 tile mt at 256 mt; W2 (mt, kb) at 1024 +
// 512 (2 mt + kb); W3 (mt, kb) at 5120 + 512 (2 mt + kb); 7168 in all.
constexpr int FRAG_W1 = 0, FRAG_W2 = 1024, FRAG_W3 = 5120;
static_assert(FRAG_W3 + 4 * 512 == FRAG_ELEMS, "fragment size");

// LDS image of the network, shared by a block: fragments and biases
// (b1 at 0, b2 at 64, b3 at 128, zero to 160).
struct MlpLds {
    alignas(16) uint16_t w[FRAG_ELEMS];
    alignas(16) float b[160];
};

__device__ __forceinline__ void mlp_stage(MlpLds &L, const uint16_t *__restrict__ frag,
                                          const float *__restrict__ pq, int tid, int nthr)
{
    const uint4 *src = reinterpret_cast<const uint4 *>(frag);
    uint4 *dst = reinterpret_cast<uint4 *>(L.w);
    for (int e = tid; e < FRAG_ELEMS / 8; e += nthr) dst[e] = src[e];
    for (int e = tid; e < 160; e += nthr)
        L.b[e] = e < 64 ? pq[640 + e] : e < 128 ? pq[704 + 4096 + e - 64]
               : e < 148 ? pq[704 + 4160 + 1280 + e - 128] : 0.0f;
}

// Features of one channel (this lane's), written where the tiles read them:
// xs[lane][0..15] = s_i / m (bfloat16, zero past ns), m = max(max_i s_i, 1).
__device__ __forceinline__ float mlp_features(const float (&v)[10], int ns, __bf16 (*xs)[16],
                                              int lane)
{
    float mx = -3.0e38f;
#pragma unroll
    for (int i = 0; i < 10; ++i)
        if (i < ns) mx = fmaxf(mx, v[i]);
    const float sc = fmaxf(mx, 1.0f);
    // the hardware reciprocal (1 ulp): the features are rounded to bfloat16
    // (2^-9 relative) next
    const float rs = __builtin_amdgcn_rcpf(sc);
    f32x4 q[4];
#pragma unroll
    for (int i = 0; i < 16; ++i) q[i / 4][i % 4] = (i < 10 && i < ns) ? v[i] * rs : 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) *reinterpret_cast<bf16x4 *>(&xs[lane][4 * k]) = cvt4(q[k]);
    return sc;
}

struct NoTileEnd {
    __device__ void operator()(int) const {}
};

// The network on the 64 channels of a wavefront, as 4 tiles of 16 channels.
// xs: the wavefront's features (mlp_features), L: the staged network.
// tile(t, g, d0, d1) is called once per tile with the output pre-activations
// the lane holds for channel t (0..63), g = lane / 16: d0[j] is output
// 4 g + j, d1[j] is output 16 + 4 g + j (outputs past 19 are padding).
template <typename TileFn>
__device__ __forceinline__ void mlp_tiles_d(const MlpLds &L, const __bf16 (*xs)[16], int lane,
                                            TileFn &&tile)
{
    const int g = lane >> 4, n = lane & 15;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        // an opaque zero offset per tile: the fragments are read from LDS
        // where they are used, not hoisted into registers for all tiles
        int w0 = 0;
        asm volatile("" : "+s"(w0));
        const bf16x4 bx = *reinterpret_cast<const bf16x4 *>(&xs[nt * 16 + n][4 * g]);
        // The bias enters as the accumulator operand: in the transposed form
        // the lane's accumulator j is output unit 16 mt + 4 g + j.
        f32x4 d[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const bf16x4 a = *reinterpret_cast<const bf16x4 *>(&L.w[w0 + FRAG_W1 + mt * 256 + lane * 4]);
            const f32x4 bb = *reinterpret_cast<const f32x4 *>(&L.b[mt * 16 + 4 * g]);
            d[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                __builtin_bit_cast(short __attribute__((ext_vector_type(4))), a),
                __builtin_bit_cast(short __attribute__((ext_vector_type(4))), bx), bb, 0, 0, 0);
        }
        bf16x8 h[2];
#pragma unroll
        for (int layer = 0; layer < 2; ++layer) {
            // GELU of the 4 tiles, regrouped as the 2 K blocks of the next layer
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                bf16x4 lo, hi;
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    const int mt = 2 * kb + half;
                    f32x4 v;
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] = gelu(d[mt][j]);
                    (half ? hi : lo) = cvt4(v);
                }
                h[kb] = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
            }
            const int base = layer == 0 ? FRAG_W2 : FRAG_W3;
            const int nmt = layer == 0 ? 4 : 2;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                if (mt >= nmt) break;
                f32x4 acc = *reinterpret_cast<const f32x4 *>(&L.b[64 * (layer + 1) + mt * 16 + 4 * g]);
#pragma unroll
                for (int kb = 0; kb < 2; ++kb) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(
                        &L.w[w0 + base + (2 * mt + kb) * 512 + lane * 8]);
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, h[kb], acc, 0, 0, 0);
                }
                d[mt] = acc;
            }
        }
        tile(nt * 16 + n, g, d[0], d[1]);
    }
}

// The same, one output at a time: emit(t, o, z) for the lane's channel t and
// outputs o = 16 mt + 4 g + j < 20 with z = the output pre-activation (logits
// for o < 10; the amplitude head's softplus is left to the caller);
// tile_end(t) after the last output of the lane's channel t of a tile.
template <typename Emit, typename TileEnd = NoTileEnd>
__device__ __forceinline__ void mlp_tiles(const MlpLds &L, const __bf16 (*xs)[16], int lane,
                                          Emit &&emit, TileEnd &&tile_end = NoTileEnd{})
{
    mlp_tiles_d(L, xs, lane, [&](int t, int g, const f32x4 &d0, const f32x4 &d1) {
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int o = mt * 16 + 4 * g + j;
                if (o < 20) emit(t, o, mt ? d1[j] : d0[j]);
            }
        }
        tile_end(t);
    });
}

// softplus(z) = max(z, 0) + ln 2 log2(1 + 2^(-|z| log2 e)) with the hardware
// exp2 and log2 (the log's argument lies in (1, 2])
__device__ __forceinline__ float softplus_hw(float z)
{
    const float e = __builtin_amdgcn_exp2f(-fabsf(z) * 1.4426950408889634f);
    return fmaf(0.6931471805599453f, __builtin_amdgcn_logf(1.0f + e), fmaxf(z, 0.0f));
}

// The warm-start prediction of the 64 channels of a wavefront (one per lane,
// samples v, ns of them): the lane receives its channel's predicted passive
// set (logit > 0, bits below ns) and amplitudes (softplus * m).  xs, pz,
// pmw: the wavefront's LDS areas.  All 64 lanes must call (the tiles read
// other lanes' features); a lane without a channel passes zero samples.
__device__ __forceinline__ uint32_t mlp_warm_wave(const MlpLds &L, __bf16 (*xs)[16],
                                                  float (*pz)[64], uint32_t *pmw, int lane,
                                                  const float (&v)[10], int ns, float (&px)[10])
{
    const uint32_t valid = (1u << (ns < 10 ? ns : 10)) - 1u;
    const float sc = mlp_features(v, ns, xs, lane);
    pmw[lane] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    mlp_tiles_d(L, xs, lane, [&](int t, int g, const f32x4 &d0, const f32x4 &d1) {
        // logits, o = 4 g + j < 10: d0 of lanes g = 0, 1, 2
        uint32_t bits = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = 4 * g + j;
            bits |= (o < 10 && d0[j] > 0.0f) ? 1u << o : 0u;
        }
        if (bits) atomicOr(&pmw[t], bits);
        // amplitude head, o = 10..19: d0 of lanes g = 2 (o = 8 + j, j >= 2) and
        // g = 3 (o = 12 + j), d1 of lanes g = 0 (o = 16 + j); one softplus per j
        // on the value this lane holds
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float sp = softplus_hw(g == 0 ? d1[j] : d0[j]);
            const int r = g == 0 ? 6 + j : g == 3 ? 2 + j : j - 2;   // o - 10
            if (g == 0 || g == 3 || (g == 2 && j >= 2)) pz[r][t] = sp;
        }
    });
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // masked by this lane's own channel's sample count (the bits of channel
    // t were set by the lanes that hold its outputs, whose ns may differ)
    const uint32_t pm = pmw[lane] & valid;
#pragma unroll
    for (int i = 0; i < 10; ++i) px[i] = pz[i][lane] * sc;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return pm;
}

}  // namespace mlp
}  // namespace mf
