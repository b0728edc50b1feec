// mlp_infer.h -- one wave evaluates the troubled-interface MLP for its 64
// interfaces (one per lane) on the matrix cores.
//
// The network is evaluated transposed, H^T = W^T X^T: with the lane layout of
// v_mfma_f32_16x16x16_bf16 (see mlp.h) the accumulator tile of one layer is
// already the B operand of the next, so activations stay in registers.  Only
// the inputs make one trip through LDS (each lane writes its 24 features as
// BF16, then reads the X^T fragments) and the logits one trip back.
//
// Every thread of the workgroup must call this routine (it synchronises the
// workgroup); lanes without an interface pass zero features.
#pragma once
#include "indicator.h"
#include "mlp.h"

namespace finess {

constexpr int MLP_FEAT_LD = 36;  // BF16 row pitch of the feature tile (8-byte aligned rows)

__device__ __forceinline__ float mlp_wave_logit(const float f[MLP_IN], const MlpFrags& fr,
                                                uint16_t (*feat)[MLP_FEAT_LD], float* lo, int lane)
{
    for (int k = 0; k < MLP_IN; k++) feat[lane][k] = f2bf(f[k]);
    for (int k = MLP_IN; k < 32; k++) feat[lane][k] = 0;
    __syncthreads();
    const int r = lane & 15, q = lane >> 4;
    // layer 1: H1^T (32 x 64) = W1^T (32 x 32) X^T (32 x 64)
    f32x4 h1[2][4];
    for (int nb = 0; nb < 4; nb++) {
        bf16x4 xb[2];
        for (int kb = 0; kb < 2; kb++) xb[kb] = *reinterpret_cast<const bf16x4*>(&feat[16 * nb + r][16 * kb + 4 * q]);
        for (int mb = 0; mb < 2; mb++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int kb = 0; kb < 2; kb++)
                acc = mfma16(*reinterpret_cast<const bf16x4*>(fr.w1[mb][kb][lane]), xb[kb], acc);
            for (int v = 0; v < 4; v++) {
                const float x = acc[v] + fr.b1[16 * mb + 4 * q + v];
                h1[mb][nb][v] = x > 0.f ? x : 0.f;
            }
        }
    }
    // layer 2: H2^T = W2^T H1^T
    f32x4 h2[2][4];
    for (int nb = 0; nb < 4; nb++) {
        bf16x4 hb[2];
        for (int kb = 0; kb < 2; kb++) hb[kb] = pack4(h1[kb][nb][0], h1[kb][nb][1], h1[kb][nb][2], h1[kb][nb][3]);
        for (int mb = 0; mb < 2; mb++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int kb = 0; kb < 2; kb++)
                acc = mfma16(*reinterpret_cast<const bf16x4*>(fr.w2[mb][kb][lane]), hb[kb], acc);
            for (int v = 0; v < 4; v++) {
                const float x = acc[v] + fr.b2[16 * mb + 4 * q + v];
                h2[mb][nb][v] = x > 0.f ? x : 0.f;
            }
        }
    }
    // layer 3: z^T (row 0 of a 16-row tile) = w3^T H2^T
    for (int nb = 0; nb < 4; nb++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < 2; kb++)
            acc = mfma16(*reinterpret_cast<const bf16x4*>(fr.w3[kb][lane]),
                         pack4(h2[kb][nb][0], h2[kb][nb][1], h2[kb][nb][2], h2[kb][nb][3]), acc);
        if (q == 0) lo[16 * nb + r] = acc[0] + fr.b3;
    }
    __syncthreads();
    return lo[lane];
}

} // namespace finess
