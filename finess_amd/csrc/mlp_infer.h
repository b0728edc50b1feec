// mlp_infer.h -- one wave evaluates the troubled-interface MLP for its 64
// interfaces (one per lane) on the matrix cores.
//
// The network is evaluated transposed, H^T = W^T X^T: with the lane layout of
// v_mfma_f32_16x16x16_bf16 (see mlp.h) the accumulator tile of one layer is
// already the B operand of the next, so activations stay in registers.  Only
// the inputs make one trip through LDS (each lane writes its 24 features as
// BF16, then reads the X^T fragments) and the logits one trip back.
//
// All 64 lanes of the wave must call this routine; lanes without an interface
// pass zero features.  The feature tile and the logit row belong to the
// calling wave alone, so only the wave synchronises.
#pragma once
#include "indicator.h"
#include "mlp.h"

namespace finess {

constexpr int MLP_FEAT_LD = 36;  // BF16 row pitch of the feature tile (8-byte aligned rows)

// Orders the wave's own LDS writes before its following LDS reads.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float mlp_wave_logit(const float f[MLP_IN], const MlpFrags& fr,
                                                uint16_t (*feat)[MLP_FEAT_LD], float* lo, int lane)
{
    // each lane's features as BF16 pairs; columns 24..31 pad K with zeros
    uint32_t* row = reinterpret_cast<uint32_t*>(feat[lane]);
    for (int k = 0; k < MLP_IN / 2; k++) row[k] = bf16_pair(f[2 * k], f[2 * k + 1]);
    for (int k = MLP_IN / 2; k < 16; k++) row[k] = 0u;
    wave_lds_sync();
    const int r = lane & 15, q = lane >> 4;
    // biases enter as the accumulators' initial values: row 16mb + 4q + v
    f32x4 b1[2], b2[2];
    for (int mb = 0; mb < 2; mb++)
        for (int v = 0; v < 4; v++) {
            b1[mb][v] = fr.b1[16 * mb + 4 * q + v];
            b2[mb][v] = fr.b2[16 * mb + 4 * q + v];
        }
    // layer 1: H1^T (32 x 64) = W1^T (32 x 32) X^T (32 x 64)
    f32x4 h1[2][4];
    for (int nb = 0; nb < 4; nb++) {
        bf16x4 xb[2];
        for (int kb = 0; kb < 2; kb++) xb[kb] = *reinterpret_cast<const bf16x4*>(&feat[16 * nb + r][16 * kb + 4 * q]);
        for (int mb = 0; mb < 2; mb++) {
            f32x4 acc = b1[mb];
            for (int kb = 0; kb < 2; kb++)
                acc = mfma16(*reinterpret_cast<const bf16x4*>(fr.w1[mb][kb][lane]), xb[kb], acc);
            h1[mb][nb] = acc;  // the ReLU is applied when packing for layer 2
        }
    }
    // layer 2: H2^T = W2^T H1^T
    f32x4 h2[2][4];
    for (int nb = 0; nb < 4; nb++) {
        bf16x4 hb[2];
        for (int kb = 0; kb < 2; kb++) hb[kb] = pack4_relu(h1[kb][nb][0], h1[kb][nb][1], h1[kb][nb][2], h1[kb][nb][3]);
        for (int mb = 0; mb < 2; mb++) {
            f32x4 acc = b2[mb];
            for (int kb = 0; kb < 2; kb++)
                acc = mfma16(*reinterpret_cast<const bf16x4*>(fr.w2[mb][kb][lane]), hb[kb], acc);
            h2[mb][nb] = acc;
        }
    }
    // layer 3: z^T (row 0 of a 16-row tile) = w3^T H2^T
    for (int nb = 0; nb < 4; nb++) {
        f32x4 acc = {fr.b3, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < 2; kb++)
            acc = mfma16(*reinterpret_cast<const bf16x4*>(fr.w3[kb][lane]),
                         pack4_relu(h2[kb][nb][0], h2[kb][nb][1], h2[kb][nb][2], h2[kb][nb][3]), acc);
        if (q == 0) lo[16 * nb + r] = acc[0];
    }
    wave_lds_sync();
    const float z = lo[lane];
    wave_lds_sync();  // the next call of this wave overwrites feat and lo
    return z;
}

// Stages the network fragments in LDS (all threads of the workgroup take
// part; the caller synchronises before the first mlp_wave_logit).
__device__ __forceinline__ void mlp_stage_frags(const MlpFrags* src, MlpFrags* dst, int t, int nt)
{
    const int4* s = reinterpret_cast<const int4*>(src);
    int4* d = reinterpret_cast<int4*>(dst);
    for (int k = t; k < (int)(sizeof(MlpFrags) / 16); k += nt) d[k] = s[k];
}

} // namespace finess
