// mlp.hip -- training, evaluation and data harvesting of the troubled-
// interface MLP on gfx950 matrix cores (BF16 operands, FP32 accumulation).
// No counterpart in the reference.
//
// Training is a sequence of small GEMMs on a batch of B interfaces:
//   H1 = relu(X W1 + b1), H2 = relu(H1 W2 + b2), z = H2 w3 + b3,
//   loss = mean BCE(sigmoid(z), y),  dz = (sigmoid(z) - y) / B,
//   gW3 = H2^T dz,  dH2 = dz w3^T . [H2 > 0],  gW2 = H1^T dH2,
//   dH1 = dH2 W2^T . [H1 > 0],  gW1 = X^T dH1,  gb = column sums,
// then Adam on the FP32 master parameters.  Every GEMM runs on
// v_mfma_f32_16x16x16_bf16 with both operands rounded to BF16 on load.  The
// reductions over the batch (weight and bias gradients) run in a fixed
// order, with no atomics, so a training run gives the same network every
// time.
#include <hip/hip_runtime.h>

#include "indicator.h"
#include "kernels.h"
#include "mlp.h"
#include "stage2d_body.h"

namespace finess {

// ---------------------------------------------------------------------------
// generic 16x16-tile GEMM: C (M x N) = op(A) (M x K) * op(B) (K x N)
//   A(m,k) = ta ? A[k*lda + m] : A[m*lda + k]
//   B(k,n) = tb ? B[n*ldb + k] : B[k*ldb + n]
//   epilogue: + bias[n]; relu; * [mask(m,n) > 0]; store or atomic add
// One wave per 16x16 output tile and K-chunk; 4 waves per workgroup.
// ---------------------------------------------------------------------------
struct GemmArgs {
    int M, N, K;
    const float* A; int lda; int ta;
    const float* B; int ldb; int tb;
    float* C; int ldc;
    const float* bias;     // length N or null
    const float* mask; int ldm;  // mask(m, n) = mask[m*ldm + n] or null
    int relu;
    int atomic;            // C += the product, K split over the 4 waves of a workgroup (k_gemm_bf16_splitk)
    int kchunk;            // K range per wave (multiple of 16)
    int tiles_m, tiles_n, chunks;
};

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g)
{
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int total = g.tiles_m * g.tiles_n * g.chunks;
    if (wave >= total) return;
    const int chunk = wave % g.chunks;
    const int tile = wave / g.chunks;
    const int tm = tile / g.tiles_n, tn = tile - tm * g.tiles_n;
    const int m0 = tm * 16, n0 = tn * 16;
    const int k_begin = chunk * g.kchunk;
    const int k_end = min(g.K, k_begin + g.kchunk);
    const int r = lane & 15, q = lane >> 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = k_begin; k0 < k_end; k0 += 16) {
        float av[4], bv[4];
        for (int v = 0; v < 4; v++) {
            const int k = k0 + 4 * q + v;
            const int m = m0 + r, n = n0 + r;
            const bool kin = k < k_end;
            av[v] = (kin && m < g.M) ? (g.ta ? g.A[(long long)k * g.lda + m] : g.A[(long long)m * g.lda + k]) : 0.f;
            bv[v] = (kin && n < g.N) ? (g.tb ? g.B[(long long)n * g.ldb + k] : g.B[(long long)k * g.ldb + n]) : 0.f;
        }
        acc = mfma16(pack4(av[0], av[1], av[2], av[3]), pack4(bv[0], bv[1], bv[2], bv[3]), acc);
    }
    const int n = n0 + r;
    if (n >= g.N) return;
    for (int v = 0; v < 4; v++) {
        const int m = m0 + 4 * q + v;
        if (m >= g.M) continue;
        float x = acc[v];
        if (g.bias) x += g.bias[n];
        if (g.relu) x = x > 0.f ? x : 0.f;
        if (g.mask && !(g.mask[(long long)m * g.ldm + n] > 0.f)) x = 0.f;
        g.C[(long long)m * g.ldc + n] = x;
    }
}

// C (M x N) += op(A) op(B) for the weight gradients, whose K is the batch:
// one workgroup per 16x16 output tile, its 4 waves taking interleaved
// 16-wide slices of K, the four partial tiles summed in a fixed order by
// wave 0 (no bias, ReLU or mask).
__global__ __launch_bounds__(256) void k_gemm_bf16_splitk(GemmArgs g)
{
    __shared__ float part[4][64][4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tm = blockIdx.x / g.tiles_n, tn = blockIdx.x - tm * g.tiles_n;
    const int m0 = tm * 16, n0 = tn * 16;
    const int r = lane & 15, q = lane >> 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 16 * w; k0 < g.K; k0 += 64) {
        float av[4], bv[4];
        for (int v = 0; v < 4; v++) {
            const int k = k0 + 4 * q + v;
            const int m = m0 + r, n = n0 + r;
            const bool kin = k < g.K;
            av[v] = (kin && m < g.M) ? (g.ta ? g.A[(long long)k * g.lda + m] : g.A[(long long)m * g.lda + k]) : 0.f;
            bv[v] = (kin && n < g.N) ? (g.tb ? g.B[(long long)n * g.ldb + k] : g.B[(long long)k * g.ldb + n]) : 0.f;
        }
        acc = mfma16(pack4(av[0], av[1], av[2], av[3]), pack4(bv[0], bv[1], bv[2], bv[3]), acc);
    }
    for (int v = 0; v < 4; v++) part[w][lane][v] = acc[v];
    __syncthreads();
    if (w != 0) return;
    const int n = n0 + r;
    if (n >= g.N) return;
    for (int v = 0; v < 4; v++) {
        const int m = m0 + 4 * q + v;
        if (m >= g.M) continue;
        g.C[(long long)m * g.ldc + n] += ((part[0][lane][v] + part[1][lane][v]) + part[2][lane][v]) + part[3][lane][v];
    }
}

static int gemm(GemmArgs g, hipStream_t st)
{
    g.tiles_m = (g.M + 15) / 16;
    g.tiles_n = (g.N + 15) / 16;
    if (g.atomic) {
        if (g.bias || g.relu || g.mask) return -1;
        hipLaunchKernelGGL(k_gemm_bf16_splitk, dim3(g.tiles_m * g.tiles_n), dim3(256), 0, st, g);
        return hipGetLastError() == hipSuccess ? 0 : -4;
    }
    g.kchunk = ((g.K + 15) / 16) * 16;
    g.chunks = 1;
    const int waves = g.tiles_m * g.tiles_n * g.chunks;
    hipLaunchKernelGGL(k_gemm_bf16, dim3((waves + 3) / 4), dim3(256), 0, st, g);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// dz = (sigmoid(z) - y) / B and the summed BCE loss (stable form).
__global__ void k_bce(const float* z, const float* y, float* dz, int n, float invB, float* loss)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        const float zi = z[i], yi = y[i];
        const float s = 1.f / (1.f + __expf(-zi));
        dz[i] = (s - yi) * invB;
        l = fmaxf(zi, 0.f) - zi * yi + __logf(1.f + __expf(-fabsf(zi)));
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
    if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, l);
}

// out[col] += the column sum of a (n x c) matrix, one workgroup of 256
// threads per column, summed in a fixed order.
__global__ __launch_bounds__(256) void k_colsum(const float* a, int n, int c, float* out)
{
    __shared__ float red[4];
    const int col = blockIdx.y;
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) s += a[(long long)i * c + col];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[col] += ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ void k_adam(float* p, float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    g[i] = 0.f;
}

__global__ void k_zero(float* p, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// Forward pass of a batch (rows of X), writing H1, H2 and the logits z.
int mlp_forward(const float* params, const float* X, int B, float* H1, float* H2, float* z, hipStream_t st)
{
    GemmArgs g = {};
    int rc;
    g.M = B; g.N = MLP_H; g.K = MLP_IN;
    g.A = X; g.lda = MLP_IN; g.B = params + MLP_OFF_W1; g.ldb = MLP_H; g.C = H1; g.ldc = MLP_H;
    g.bias = params + MLP_OFF_B1; g.relu = 1;
    if ((rc = gemm(g, st))) return rc;
    g = GemmArgs{};
    g.M = B; g.N = MLP_H; g.K = MLP_H;
    g.A = H1; g.lda = MLP_H; g.B = params + MLP_OFF_W2; g.ldb = MLP_H; g.C = H2; g.ldc = MLP_H;
    g.bias = params + MLP_OFF_B2; g.relu = 1;
    if ((rc = gemm(g, st))) return rc;
    g = GemmArgs{};
    g.M = B; g.N = 1; g.K = MLP_H;
    g.A = H2; g.lda = MLP_H; g.B = params + MLP_OFF_W3; g.ldb = 1; g.C = z; g.ldc = 1;
    g.bias = params + MLP_OFF_B3;
    return gemm(g, st);
}

// Gradient of the mean BCE loss of a batch, accumulated into grad (which the
// caller zeroes); loss_sum receives the summed loss (may be null).
int mlp_backward(const float* params, const float* X, const float* y, int B, float* H1, float* H2, float* z,
                 float* dz, float* dH2, float* dH1, float* grad, float* loss_sum, hipStream_t st)
{
    int rc;
    if ((rc = mlp_forward(params, X, B, H1, H2, z, st))) return rc;
    hipLaunchKernelGGL(k_bce, dim3((B + 255) / 256), dim3(256), 0, st, z, y, dz, B, 1.f / B, loss_sum);
    GemmArgs g = {};
    // gW3 = H2^T dz
    g.M = MLP_H; g.N = 1; g.K = B;
    g.A = H2; g.lda = MLP_H; g.ta = 1; g.B = dz; g.ldb = 1; g.C = grad + MLP_OFF_W3; g.ldc = 1; g.atomic = 1;
    if ((rc = gemm(g, st))) return rc;
    // dH2 = dz w3^T . [H2 > 0]
    g = GemmArgs{};
    g.M = B; g.N = MLP_H; g.K = 1;
    g.A = dz; g.lda = 1; g.B = params + MLP_OFF_W3; g.ldb = MLP_H; g.C = dH2; g.ldc = MLP_H; g.mask = H2; g.ldm = MLP_H;
    if ((rc = gemm(g, st))) return rc;
    // gW2 = H1^T dH2
    g = GemmArgs{};
    g.M = MLP_H; g.N = MLP_H; g.K = B;
    g.A = H1; g.lda = MLP_H; g.ta = 1; g.B = dH2; g.ldb = MLP_H; g.C = grad + MLP_OFF_W2; g.ldc = MLP_H; g.atomic = 1;
    if ((rc = gemm(g, st))) return rc;
    // dH1 = dH2 W2^T . [H1 > 0]
    g = GemmArgs{};
    g.M = B; g.N = MLP_H; g.K = MLP_H;
    g.A = dH2; g.lda = MLP_H; g.B = params + MLP_OFF_W2; g.ldb = MLP_H; g.tb = 1; g.C = dH1; g.ldc = MLP_H;
    g.mask = H1; g.ldm = MLP_H;
    if ((rc = gemm(g, st))) return rc;
    // gW1 = X^T dH1
    g = GemmArgs{};
    g.M = MLP_IN; g.N = MLP_H; g.K = B;
    g.A = X; g.lda = MLP_IN; g.ta = 1; g.B = dH1; g.ldb = MLP_H; g.C = grad + MLP_OFF_W1; g.ldc = MLP_H; g.atomic = 1;
    if ((rc = gemm(g, st))) return rc;
    // bias gradients
    hipLaunchKernelGGL(k_colsum, dim3(1, 1), dim3(256), 0, st, dz, B, 1, grad + MLP_OFF_B3);
    hipLaunchKernelGGL(k_colsum, dim3(1, MLP_H), dim3(256), 0, st, dH2, B, MLP_H, grad + MLP_OFF_B2);
    hipLaunchKernelGGL(k_colsum, dim3(1, MLP_H), dim3(256), 0, st, dH1, B, MLP_H, grad + MLP_OFF_B1);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int mlp_adam(float* p, float* g, float* m, float* v, float lr, int t, hipStream_t st)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, p, g, m, v, MLP_NPARAM, lr, b1, b2,
                       1e-8f, c1, c2);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int mlp_zero(float* p, long long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// ---------------------------------------------------------------------------
// Harvest: features and labels of every face of a 2D Euler state (one
// thread per face, the state read through the boundary map).  Faces k <
// (mx+1)*my are x-faces (i = k / my, j = k % my), the rest y-faces.  Every
// `stride`-th face (with offset `phase`) is written at X[out0 + slot].
// ---------------------------------------------------------------------------

__device__ void load_rec(const HarvestArgs& a, int gi, int gj, double rec[8])
{
    int fi, fj;
    const int si = bc_map(a.bc[0], a.bc[1], gi, a.mx, fi);
    const int sj = bc_map(a.bc[2], a.bc[3], gj, a.my, fj);
    double q[5];
    for (int m = 0; m < 5; m++) q[m] = a.q[m * a.plane + (long long)si * a.my + sj];
    if (fi) q[1] = -q[1];
    if (fj) q[2] = -q[2];
    Euler2D::cell_record(q, a.pp, rec);
}

__global__ void k_harvest(HarvestArgs a)
{
    const long long nfx = (long long)(a.mx + 1) * a.my, nf = nfx + (long long)a.mx * (a.my + 1);
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nf) return;
    if ((k % a.stride) != a.phase) return;
    const long long slot = k / a.stride;
    if (a.out0 + slot >= a.cap) return;
    int dir, gi, gj;
    if (k < nfx) { dir = 0; gi = (int)(k / a.my); gj = (int)(k - (long long)gi * a.my); }
    else { const long long h = k - nfx; dir = 1; gi = (int)(h / (a.my + 1)); gj = (int)(h - (long long)gi * (a.my + 1)); }
    double rec[6][8];
    for (int s = 0; s < 6; s++) {
        if (dir == 0) load_rec(a, gi - 3 + s, gj, rec[s]);
        else load_rec(a, gi, gj - 3 + s, rec[s]);
    }
    float f[MLP_IN];
    face_features<Euler2D, double>(rec, dir, a.pp, f);
    const double d = face_departure<Euler2D, double>(rec, dir, a.pp, a.eps, a.alpha_scaling);
    const long long o = a.out0 + slot;
    for (int i = 0; i < MLP_IN; i++) a.X[o * MLP_IN + i] = f[i];
    a.y[o] = d > a.tau ? 1.f : 0.f;
    if (a.dep) a.dep[o] = (float)d;
}

int launch_harvest(const HarvestArgs& a, hipStream_t st)
{
    const long long nf = (long long)(a.mx + 1) * a.my + (long long)a.mx * (a.my + 1);
    hipLaunchKernelGGL(k_harvest, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

// ---------------------------------------------------------------------------
// Inference of the transposed network on one wave's 64 interfaces, the same
// routine the routed stage kernel uses (declared in mlp_infer.h).
// ---------------------------------------------------------------------------
} // namespace finess

#include "mlp_infer.h"

namespace finess {

__global__ __launch_bounds__(256) void k_mlp_eval(const float* X, long long n, const MlpFrags* fr, float* logits)
{
    __shared__ uint16_t feat[4][64][MLP_FEAT_LD];
    __shared__ float lo[4][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    float f[MLP_IN];
    for (int k = 0; k < MLP_IN; k++) f[k] = i < n ? X[i * MLP_IN + k] : 0.f;
    const float z = mlp_wave_logit(f, *fr, feat[w], lo[w], lane);
    if (i < n) logits[i] = z;
}

int launch_mlp_eval(const float* X, long long n, const MlpFrags* fr, float* logits, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, X, n, fr, logits);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

} // namespace finess
