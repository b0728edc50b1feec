// mlp.h -- BF16 matrix-core building blocks of the troubled-interface MLP
// (24 -> 32 -> 32 -> 1, ReLU, logistic output): operand rounding and the
// 16x16x16 MFMA tile with FP32 accumulation.  No counterpart in the
// reference.
//
// Lane layout of v_mfma_f32_16x16x16_bf16 (wave64), used everywhere below:
//   A (16 x 16, M x K): lane l holds A[l % 16][4*(l/16) + v], v = 0..3
//   B (16 x 16, K x N): lane l holds B[4*(l/16) + v][l % 16]
//   D (16 x 16, M x N): lane l holds D[4*(l/16) + v][l % 16]
// so a D tile is, lane for lane, the B operand of a following product whose
// K index is D's row index.  The inference inside the stage kernel computes
// the transposed network H^T = W^T X^T for that reason: each layer's
// accumulators feed the next layer's MFMA without moving data between lanes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace finess {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs).
__host__ __device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Two FP32 values to a packed BF16 pair (low half = a) with the gfx950
// conversion instruction, which rounds to nearest even like f2bf.
__device__ __forceinline__ uint32_t bf16_pair(float a, float b)
{
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    bf16x2_t v;
    v[0] = (__bf16)a;
    v[1] = (__bf16)b;
    uint32_t u;
    __builtin_memcpy(&u, &v, 4);
    return u;
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d)
{
    const uint32_t u[2] = {bf16_pair(a, b), bf16_pair(c, d)};
    bf16x4 r;
    __builtin_memcpy(&r, u, 8);
    return r;
}

// ReLU of a packed BF16 pair as a signed 16-bit max with zero: a negative
// value (sign bit set) becomes +0, others are unchanged, so it equals
// rounding max(x, 0) for every finite x (-0 gives +0 instead of -0).
__device__ __forceinline__ uint32_t relu_bf16_pair(uint32_t u)
{
    typedef short s16x2 __attribute__((ext_vector_type(2)));
    s16x2 v;
    __builtin_memcpy(&v, &u, 4);
    const s16x2 z = {0, 0};
    v = __builtin_elementwise_max(v, z);
    uint32_t r;
    __builtin_memcpy(&r, &v, 4);
    return r;
}

// max(., 0) of four values, rounded to BF16 and packed (the hidden layers'
// activation ahead of the next product).
__device__ __forceinline__ bf16x4 pack4_relu(float a, float b, float c, float d)
{
    const uint32_t u[2] = {relu_bf16_pair(bf16_pair(a, b)), relu_bf16_pair(bf16_pair(c, d))};
    bf16x4 r;
    __builtin_memcpy(&r, u, 8);
    return r;
}

// Weight fragments of the transposed network for the in-kernel inference,
// prepared on the host from the FP32 parameter vector (finess_mlp layout):
//   w1[mb][kb][lane][4]  W1^T rows 16mb.., cols (inputs) 16kb..  (2 x 2)
//   w2[mb][kb][lane][4]  W2^T                                      (2 x 2)
//   w3[kb][lane][4]      w3^T in row 0 of a 16-row tile            (2)
//   b1[32], b2[32], b3, threshold on the logit (troubled if above)
struct __attribute__((aligned(16))) MlpFrags {
    uint16_t w1[2][2][64][4];
    uint16_t w2[2][2][64][4];
    uint16_t w3[2][64][4];
    float b1[32], b2[32];
    float b3;
    float threshold;
    float pad[2];  // whole 16-byte chunks, so a workgroup can stage it in LDS with vector copies
};
static_assert(sizeof(MlpFrags) % 16 == 0, "MlpFrags is copied in 16-byte chunks");

void mlp_make_frags(const float* params, float threshold, MlpFrags* f);

} // namespace finess
