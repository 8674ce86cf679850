// mlp.hpp -- host-side interface of the warm-start MLP kernels (mlp.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace mf {
namespace mlp {

constexpr int NPARAM = 6164;  // [W1 10x64, b1 64, W2 64x64, b2 64, W3 64x20, b3 20]

// activations of one training batch (device memory, batch-major)
struct Scratch {
    int64_t cap = 0;
    float *X = nullptr, *m = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr,
          *H2 = nullptr, *Z3 = nullptr, *lossdummy = nullptr;
};

// loss (device scalar, may be null) and gradient of one batch
hipError_t grad_step(Scratch &w, const float *samples, int64_t lds, const float *xref,
                     int64_t ldx, int64_t B, const float *pq, float *grad, float *loss,
                     hipStream_t st);
hipError_t adam_step(float *p, const float *g, float *m, float *v, int t, float lr, float *pq,
                     uint16_t *pbf, hipStream_t st);
hipError_t quantize(const float *p, float *pq, uint16_t *pbf, hipStream_t st);
hipError_t zero(float *p, int64_t n, hipStream_t st);
// packs the bfloat16 weights into the inference kernels' fragment order
// (frag: FRAG_ELEMS uint16)
hipError_t pack(const uint16_t *pbf, uint16_t *frag, hipStream_t st);
constexpr int FRAG_ELEMS = 7168;
// frag: packed fragments (pack), pq: bfloat16-rounded parameters as floats
template <typename TS>
void launch_infer(const void *s, int64_t ld, int ns, const int32_t *idx, int64_t n,
                  const uint16_t *frag, const float *pq, float *out, int64_t ldo, hipStream_t st);

// Sample counts of up to 256 channel types (8 or 10): bit t of is8 set when
// type t has 8 samples.
struct NsOfType {
    uint32_t is8[8];
    __host__ __device__ int samples_of(unsigned t) const
    {
        return ((is8[(t >> 5) & 7u] >> (t & 31u)) & 1u) ? 8 : 10;
    }
};

// warm-start form over channels c0..c0+n-1 of s ([10][ld]): passive-set bits
// ppm[c] and predicted amplitudes px[i*ld + c]; each channel has ns samples,
// or, when type_of is given (device, one byte per channel), its type's count
// in nst
template <typename TS>
void launch_warm(const void *s, int64_t ld, int ns, int64_t c0, int64_t n,
                 const uint8_t *type_of, const NsOfType &nst, const uint16_t *frag,
                 const float *pq, float *px, uint32_t *ppm, hipStream_t st);

}  // namespace mlp
}  // namespace mf
