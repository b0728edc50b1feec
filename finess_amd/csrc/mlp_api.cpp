// mlp_api.cpp -- C-ABI of the troubled-interface network (include/finess_hip.h):
// device training set harvested from the solver's own states, Adam training
// with shuffled mini-batches on the matrix-core kernels of mlp.hip, gradient
// and evaluation entry points for the tests.  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/finess_hip.h"
#include "indicator.h"
#include "kernels.h"
#include "mlp.h"
#include "stage2d_body.h"

extern "C" {
hipStream_t finess_ctx_stream(finess_ctx* c);
const double* finess_ctx_state(finess_ctx* c);
const finess_params* finess_ctx_params(finess_ctx* c);
}

namespace finess {


void mlp_make_frags(const float* P, float threshold, MlpFrags* f)
{
    std::memset(f, 0, sizeof *f);
    for (int mb = 0; mb < 2; mb++)
        for (int kb = 0; kb < 2; kb++)
            for (int lane = 0; lane < 64; lane++)
                for (int v = 0; v < 4; v++) {
                    const int r = lane & 15, q = lane >> 4;
                    const int n = 16 * mb + r, k = 16 * kb + 4 * q + v;
                    f->w1[mb][kb][lane][v] = k < MLP_IN ? f2bf(P[MLP_OFF_W1 + k * MLP_H + n]) : 0;
                    f->w2[mb][kb][lane][v] = f2bf(P[MLP_OFF_W2 + k * MLP_H + n]);
                }
    for (int kb = 0; kb < 2; kb++)
        for (int lane = 0; lane < 64; lane++)
            for (int v = 0; v < 4; v++) {
                const int r = lane & 15, q = lane >> 4, k = 16 * kb + 4 * q + v;
                f->w3[kb][lane][v] = r == 0 ? f2bf(P[MLP_OFF_W3 + k]) : 0;
            }
    for (int n = 0; n < MLP_H; n++) {
        f->b1[n] = P[MLP_OFF_B1 + n];
        f->b2[n] = P[MLP_OFF_B2 + n];
    }
    f->b3 = P[MLP_OFF_B3];
    f->threshold = threshold;
}

__global__ void k_gather(const float* X, const float* y, const int* perm, long long start, int B, float* Xb, float* yb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const long long src = perm[start + i];
    for (int k = 0; k < MLP_IN; k++) Xb[(long long)i * MLP_IN + k] = X[src * MLP_IN + k];
    yb[i] = y[src];
}

} // namespace finess

using namespace finess;

struct finess_mlp {
    hipStream_t st = nullptr;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    float *X = nullptr, *y = nullptr, *dep = nullptr;
    long long cap = 0, n = 0;
    int *perm = nullptr;
    float *Xb = nullptr, *yb = nullptr, *H1 = nullptr, *H2 = nullptr, *z = nullptr, *dz = nullptr, *dH2 = nullptr,
          *dH1 = nullptr, *loss = nullptr;
    MlpFrags* frags = nullptr;
    int adam_t = 0;
};

namespace {
constexpr int BMAX = 65536;

int mk(void** p, size_t bytes) { return hipMalloc(p, bytes) == hipSuccess ? 0 : FINESS_ERR_HIP; }

// He-normal initial weights from splitmix64 + Box-Muller; zero biases.
void init_params(unsigned long long seed, std::vector<float>& P)
{
    P.assign(MLP_NPARAM, 0.f);
    unsigned long long s = seed;
    auto uni = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        unsigned long long z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto normal = [&]() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni()); };
    for (int i = 0; i < MLP_IN * MLP_H; i++) P[MLP_OFF_W1 + i] = (float)(normal() * std::sqrt(2.0 / MLP_IN));
    for (int i = 0; i < MLP_H * MLP_H; i++) P[MLP_OFF_W2 + i] = (float)(normal() * std::sqrt(2.0 / MLP_H));
    for (int i = 0; i < MLP_H; i++) P[MLP_OFF_W3 + i] = (float)(normal() * std::sqrt(1.0 / MLP_H));
}
} // namespace

extern "C" {

void finess_mlp_destroy(finess_mlp* t)
{
    if (!t) return;
    if (t->st) hipStreamSynchronize(t->st);
    void* ptrs[] = {t->p, t->g, t->m, t->v, t->X, t->y, t->dep, t->perm, t->Xb, t->yb, t->H1, t->H2, t->z,
                    t->dz, t->dH2, t->dH1, t->loss, t->frags};
    for (void* q : ptrs) hipFree(q);
    if (t->st) hipStreamDestroy(t->st);
    delete t;
}

int finess_mlp_create(unsigned long long seed, long long capacity, finess_mlp** out)
{
    if (!out || capacity < 1) return FINESS_ERR_ARG;
    *out = nullptr;
    if (finess_device_count() < 1) return FINESS_ERR_NODEVICE;
    finess_mlp* t = new finess_mlp();
    t->cap = capacity;
    int rc = hipStreamCreate(&t->st) == hipSuccess ? 0 : FINESS_ERR_HIP;
    const size_t pb = MLP_NPARAM * sizeof(float);
    if (!rc) rc = mk((void**)&t->p, pb);
    if (!rc) rc = mk((void**)&t->g, pb);
    if (!rc) rc = mk((void**)&t->m, pb);
    if (!rc) rc = mk((void**)&t->v, pb);
    if (!rc) rc = mk((void**)&t->X, (size_t)capacity * MLP_IN * sizeof(float));
    if (!rc) rc = mk((void**)&t->y, (size_t)capacity * sizeof(float));
    if (!rc) rc = mk((void**)&t->dep, (size_t)capacity * sizeof(float));
    if (!rc) rc = mk((void**)&t->perm, (size_t)capacity * sizeof(int));
    if (!rc) rc = mk((void**)&t->Xb, (size_t)BMAX * MLP_IN * sizeof(float));
    if (!rc) rc = mk((void**)&t->yb, (size_t)BMAX * sizeof(float));
    if (!rc) rc = mk((void**)&t->H1, (size_t)BMAX * MLP_H * sizeof(float));
    if (!rc) rc = mk((void**)&t->H2, (size_t)BMAX * MLP_H * sizeof(float));
    if (!rc) rc = mk((void**)&t->dH1, (size_t)BMAX * MLP_H * sizeof(float));
    if (!rc) rc = mk((void**)&t->dH2, (size_t)BMAX * MLP_H * sizeof(float));
    if (!rc) rc = mk((void**)&t->z, (size_t)BMAX * sizeof(float));
    if (!rc) rc = mk((void**)&t->dz, (size_t)BMAX * sizeof(float));
    if (!rc) rc = mk((void**)&t->loss, sizeof(float));
    if (!rc) rc = mk((void**)&t->frags, sizeof(MlpFrags));
    if (!rc) {
        std::vector<float> P;
        init_params(seed, P);
        if (hipMemcpy(t->p, P.data(), pb, hipMemcpyHostToDevice) != hipSuccess) rc = FINESS_ERR_HIP;
        if (!rc && (mlp_zero(t->g, MLP_NPARAM, t->st) || mlp_zero(t->m, MLP_NPARAM, t->st) ||
                    mlp_zero(t->v, MLP_NPARAM, t->st)))
            rc = FINESS_ERR_HIP;
    }
    if (rc) {
        finess_mlp_destroy(t);
        return rc;
    }
    *out = t;
    return FINESS_OK;
}

int finess_mlp_harvest(finess_mlp* t, finess_ctx* c, int stride, double tau, long long* count)
{
    if (!t || !c || stride < 1) return FINESS_ERR_ARG;
    const finess_params* p = finess_ctx_params(c);
    const double* q = finess_ctx_state(c);
    if (!q) return FINESS_ERR_ARG;
    if (p->ndims != 2 || p->physics != FINESS_PHYS_EULER) return FINESS_ERR_CONFIG;
    HarvestArgs a;
    std::memset(&a, 0, sizeof a);
    a.q = q;
    a.mx = p->mx;
    a.my = p->my;
    a.plane = (long long)p->mx * p->my;
    for (int k = 0; k < 4; k++) a.bc[k] = p->bc[k];
    a.pp.gamma = p->gamma;
    a.eps = p->epsilon;
    a.alpha_scaling = p->alpha_scaling;
    a.tau = tau;
    a.stride = stride;
    a.phase = (int)(t->n % stride);
    a.X = t->X;
    a.y = t->y;
    a.dep = t->dep;
    a.out0 = t->n;
    a.cap = t->cap;
    const hipStream_t st = finess_ctx_stream(c);
    if (launch_harvest(a, st)) return FINESS_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return FINESS_ERR_HIP;
    const long long nf = (long long)(p->mx + 1) * p->my + (long long)p->mx * (p->my + 1);
    const long long got = nf > a.phase ? (nf - a.phase + stride - 1) / stride : 0;
    t->n = std::min(t->cap, t->n + got);
    if (count) *count = t->n;
    return FINESS_OK;
}

int finess_mlp_get_data(finess_mlp* t, long long first, long long n, float* X, float* y, float* dep)
{
    if (!t || first < 0 || first + n > t->n) return FINESS_ERR_ARG;
    hipError_t e = hipSuccess;
    if (X) e = hipMemcpy(X, t->X + first * MLP_IN, (size_t)n * MLP_IN * sizeof(float), hipMemcpyDeviceToHost);
    if (y && e == hipSuccess) e = hipMemcpy(y, t->y + first, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
    if (dep && e == hipSuccess) e = hipMemcpy(dep, t->dep + first, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
    return e == hipSuccess ? FINESS_OK : FINESS_ERR_HIP;
}

int finess_mlp_set_data(finess_mlp* t, const float* X, const float* y, long long n)
{
    if (!t || !X || !y || n < 0 || n > t->cap) return FINESS_ERR_ARG;
    hipError_t e = hipMemcpy(t->X, X, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t->y, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return FINESS_ERR_HIP;
    t->n = n;
    return FINESS_OK;
}

int finess_mlp_get_params(finess_mlp* t, float* P)
{
    if (!t || !P) return FINESS_ERR_ARG;
    return hipMemcpy(P, t->p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess ? FINESS_OK
                                                                                             : FINESS_ERR_HIP;
}

int finess_mlp_set_params(finess_mlp* t, const float* P)
{
    if (!t || !P) return FINESS_ERR_ARG;
    return hipMemcpy(t->p, P, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice) == hipSuccess ? FINESS_OK
                                                                                             : FINESS_ERR_HIP;
}

int finess_mlp_train(finess_mlp* t, int epochs, int batch, float lr, float* loss)
{
    if (!t || epochs < 1 || batch < 16 || batch > BMAX) return FINESS_ERR_ARG;
    if (t->n < batch) return FINESS_ERR_ARG;
    std::mt19937_64 rng(12345 + t->adam_t);
    std::vector<int> perm((size_t)t->n);
    const long long nb = t->n / batch;
    for (int e = 0; e < epochs; e++) {
        std::iota(perm.begin(), perm.end(), 0);
        std::shuffle(perm.begin(), perm.end(), rng);
        if (hipMemcpyAsync(t->perm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, t->st) !=
            hipSuccess)
            return FINESS_ERR_HIP;
        if (mlp_zero(t->loss, 1, t->st)) return FINESS_ERR_HIP;
        for (long long b = 0; b < nb; b++) {
            hipLaunchKernelGGL(k_gather, dim3((batch + 255) / 256), dim3(256), 0, t->st, t->X, t->y, t->perm,
                               b * batch, batch, t->Xb, t->yb);
            if (mlp_backward(t->p, t->Xb, t->yb, batch, t->H1, t->H2, t->z, t->dz, t->dH2, t->dH1, t->g, t->loss,
                             t->st))
                return FINESS_ERR_HIP;
            if (mlp_adam(t->p, t->g, t->m, t->v, lr, ++t->adam_t, t->st)) return FINESS_ERR_HIP;
        }
        float l = 0.f;
        if (hipMemcpyAsync(&l, t->loss, sizeof l, hipMemcpyDeviceToHost, t->st) != hipSuccess ||
            hipStreamSynchronize(t->st) != hipSuccess)
            return FINESS_ERR_HIP;
        if (loss) loss[e] = l / (float)(nb * batch);
    }
    return FINESS_OK;
}

int finess_mlp_gradient(finess_mlp* t, const float* X, const float* y, int n, float* grad, float* loss)
{
    if (!t || !X || !y || !grad || n < 1 || n > BMAX) return FINESS_ERR_ARG;
    hipError_t e = hipMemcpyAsync(t->Xb, X, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, t->st);
    if (e == hipSuccess) e = hipMemcpyAsync(t->yb, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice, t->st);
    if (e != hipSuccess || mlp_zero(t->g, MLP_NPARAM, t->st) || mlp_zero(t->loss, 1, t->st)) return FINESS_ERR_HIP;
    if (mlp_backward(t->p, t->Xb, t->yb, n, t->H1, t->H2, t->z, t->dz, t->dH2, t->dH1, t->g, t->loss, t->st))
        return FINESS_ERR_HIP;
    float l = 0.f;
    e = hipMemcpyAsync(grad, t->g, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, t->st);
    if (e == hipSuccess) e = hipMemcpyAsync(&l, t->loss, sizeof l, hipMemcpyDeviceToHost, t->st);
    if (e == hipSuccess) e = hipStreamSynchronize(t->st);
    if (e != hipSuccess || mlp_zero(t->g, MLP_NPARAM, t->st)) return FINESS_ERR_HIP;
    if (loss) *loss = l / (float)n;
    return FINESS_OK;
}

int finess_mlp_eval(finess_mlp* t, const float* X, long long n, int path, float* logits)
{
    if (!t || !X || !logits || n < 0) return FINESS_ERR_ARG;
    if (path == 1) {
        std::vector<float> P(MLP_NPARAM);
        MlpFrags f;
        if (finess_mlp_get_params(t, P.data())) return FINESS_ERR_HIP;
        mlp_make_frags(P.data(), 0.f, &f);
        if (hipMemcpyAsync(t->frags, &f, sizeof f, hipMemcpyHostToDevice, t->st) != hipSuccess) return FINESS_ERR_HIP;
    }
    for (long long s = 0; s < n; s += BMAX) {
        const int b = (int)std::min<long long>(BMAX, n - s);
        if (hipMemcpyAsync(t->Xb, X + s * MLP_IN, (size_t)b * MLP_IN * sizeof(float), hipMemcpyHostToDevice,
                           t->st) != hipSuccess)
            return FINESS_ERR_HIP;
        const int rc = path == 1 ? launch_mlp_eval(t->Xb, b, t->frags, t->z, t->st)
                                 : mlp_forward(t->p, t->Xb, b, t->H1, t->H2, t->z, t->st);
        if (rc) return FINESS_ERR_HIP;
        if (hipMemcpyAsync(logits + s, t->z, (size_t)b * sizeof(float), hipMemcpyDeviceToHost, t->st) != hipSuccess ||
            hipStreamSynchronize(t->st) != hipSuccess)
            return FINESS_ERR_HIP;
    }
    return FINESS_OK;
}

} // extern "C"
