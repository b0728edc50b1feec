// batch_api.cpp -- C-ABI of the batched mode (include/finess_hip.h; no
// counterpart in the reference).  Host side: the problem table, the tile map
// of the chosen tile shape, the tile-shape autotuner and the HIP graph that
// holds a run of steps (three persistent stage launches and one controller
// launch per step).
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>

#include "../../include/finess_hip.h"
#include "batch.h"
#include "kernels.h"

using namespace finess;

struct finess_batch {
    finess_params p;
    int n = 0, M = 0, device = 0;
    int physics = 0, fp32 = 0;
    hipStream_t st = nullptr;
    std::vector<BatchProb> host;            // mirror of the device table
    BatchProb* probs = nullptr;
    int* tile_prob = nullptr;
    int ntiles = 0, tile_cap = 0;
    unsigned long long* smax = nullptr;     // [n][2][BATCH_SPREAD]
    std::vector<double*> bufs;
    double* staging = nullptr;              // reference layout of the largest problem
    size_t staging_n = 0;
    int shape = 0, grid = 0;
    hipGraphExec_t exec = nullptr;
    int exec_steps = 0, exec_shape = -1;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
};

namespace {

int bfail(finess_batch* b, int code, const std::string& msg)
{
    if (b) b->err = msg;
    return code;
}

BatchArgs args_of(finess_batch* b)
{
    BatchArgs a;
    memset(&a, 0, sizeof a);
    a.probs = b->probs;
    a.tile_prob = b->tile_prob;
    a.ntiles = b->ntiles;
    a.nprob = b->n;
    for (int k = 0; k < 4; k++) a.bc[k] = b->p.bc[k];
    a.pp.gamma = b->p.gamma;
    a.pp.sw_compat_hl = b->p.sw_compat_hl;
    a.eps = b->p.epsilon;
    a.power = b->p.power_param;
    a.alpha_scaling = b->p.alpha_scaling;
    a.desired_cfl = b->p.desired_cfl;
    a.max_cfl = b->p.max_cfl;
    a.max_dt = b->p.max_dt;
    return a;
}

int upload_table(finess_batch* b)
{
    if (hipMemcpyAsync(b->probs, b->host.data(), b->n * sizeof(BatchProb), hipMemcpyHostToDevice, b->st) !=
            hipSuccess ||
        hipStreamSynchronize(b->st) != hipSuccess)
        return bfail(b, FINESS_ERR_HIP, "table upload failed");
    return 0;
}

int download_table(finess_batch* b)
{
    if (hipMemcpyAsync(b->host.data(), b->probs, b->n * sizeof(BatchProb), hipMemcpyDeviceToHost, b->st) !=
            hipSuccess ||
        hipStreamSynchronize(b->st) != hipSuccess)
        return bfail(b, FINESS_ERR_HIP, "table download failed");
    return 0;
}

// Tile map of a shape: problem k's tiles are entries tile0 .. tile0 + nti*ntj - 1.
int set_shape(finess_batch* b, int shape)
{
    const int TI = BATCH_SHAPES[shape][0], TJ = BATCH_SHAPES[shape][1];
    std::vector<int> map;
    for (int k = 0; k < b->n; k++) {
        BatchProb& pr = b->host[k];
        const int nti = (pr.mx + TI - 1) / TI;
        pr.ntj = (pr.my + TJ - 1) / TJ;
        pr.tile0 = (int)map.size();
        map.insert(map.end(), (size_t)nti * pr.ntj, k);
    }
    if ((int)map.size() > b->tile_cap) {
        hipFree(b->tile_prob);
        b->tile_prob = nullptr;
        if (hipMalloc((void**)&b->tile_prob, map.size() * sizeof(int)) != hipSuccess)
            return bfail(b, FINESS_ERR_HIP, "tile map allocation failed");
        b->tile_cap = (int)map.size();
    }
    b->ntiles = (int)map.size();
    if (hipMemcpy(b->tile_prob, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(b, FINESS_ERR_HIP, "tile map upload failed");
    b->shape = shape;
    b->grid = batch_grid(b->physics, b->fp32, shape, b->device);
    if (b->grid < 1) return bfail(b, FINESS_ERR_HIP, "occupancy query failed");
    if (b->grid > b->ntiles) b->grid = b->ntiles;
    return upload_table(b);
}

int enqueue_steps(finess_batch* b, int steps)
{
    const BatchArgs a = args_of(b);
    for (int s = 0; s < steps; s++) {
        for (int stage = 1; stage <= 3; stage++)
            if (launch_batch_stage(b->physics, b->fp32, b->shape, stage, a, b->grid, b->st))
                return bfail(b, FINESS_ERR_HIP, "stage launch failed");
        if (launch_batch_control(a, b->st)) return bfail(b, FINESS_ERR_HIP, "controller launch failed");
    }
    return 0;
}

// The graph of `steps` steps for the current tile shape.
int ensure_graph(finess_batch* b, int steps)
{
    if (b->exec && b->exec_shape == b->shape && b->exec_steps == steps) return 0;
    if (b->exec) hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return bfail(b, FINESS_ERR_HIP, "graph capture failed");
    const int rc = enqueue_steps(b, steps);
    const hipError_t e = hipStreamEndCapture(b->st, &g);
    if (rc) return rc;
    if (e != hipSuccess || hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return bfail(b, FINESS_ERR_HIP, "graph instantiation failed");
    }
    hipGraphDestroy(g);
    b->exec_shape = b->shape;
    b->exec_steps = steps;
    return 0;
}

} // namespace

extern "C" {

int finess_batch_create(const finess_params* base, int nprob, const int* mx, const int* my, finess_batch** out)
{
    if (!base || !mx || !my || !out || nprob < 1) return FINESS_ERR_ARG;
    *out = nullptr;
    finess_params p = *base;
    int rc = finess_params_finalize(&p);
    if (rc) return rc;
    if (p.ndims != 2 || p.time_stepping_method != FINESS_TS_RK || p.time_order != 3 ||
        p.weno_version != FINESS_WENO_JS || p.power_param != 2.0 || p.routing != FINESS_ROUTE_NONE ||
        p.space_order != 5 || p.global_alpha || p.source_term || p.bc[0] == FINESS_BC_HOST ||
        p.bc[1] == FINESS_BC_HOST || p.bc[2] == FINESS_BC_HOST || p.bc[3] == FINESS_BC_HOST ||
        (p.physics != FINESS_PHYS_SHALLOW_WATER && p.physics != FINESS_PHYS_EULER))
        return FINESS_ERR_CONFIG;
    if (finess_device_count() < 1) return FINESS_ERR_NODEVICE;
    finess_batch* b = new finess_batch();
    b->p = p;
    b->n = nprob;
    b->M = p.meqn;
    b->physics = p.physics == FINESS_PHYS_SHALLOW_WATER ? PHYS_SHALLOW_WATER : PHYS_EULER;
    b->fp32 = p.precision == FINESS_PREC_FP32;
    rc = 0;
    if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreate(&b->st) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
        rc = FINESS_ERR_HIP;
    if (!rc && (hipMalloc((void**)&b->probs, nprob * sizeof(BatchProb)) != hipSuccess ||
                hipMalloc((void**)&b->smax, (size_t)nprob * 2 * BATCH_SPREAD * sizeof(unsigned long long)) !=
                    hipSuccess))
        rc = FINESS_ERR_HIP;
    b->host.assign(nprob, BatchProb());
    for (int k = 0; k < nprob && !rc; k++) {
        if (mx[k] < 1 || my[k] < 1) { rc = FINESS_ERR_ARG; break; }
        BatchProb& pr = b->host[k];
        memset(&pr, 0, sizeof pr);
        pr.mx = mx[k];
        pr.my = my[k];
        pr.plane = (long long)mx[k] * my[k];
        pr.dx = (p.xhigh - p.xlow) / mx[k];
        pr.dy = (p.yhigh - p.ylow) / my[k];
        pr.dt = p.initial_dt;
        pr.smax = b->smax + (size_t)k * 2 * BATCH_SPREAD;
        pr.done = 1;  // idle until a run sets its target
        for (int s = 0; s < 3 && !rc; s++) {
            double* d = nullptr;
            if (hipMalloc((void**)&d, (size_t)b->M * pr.plane * sizeof(double)) != hipSuccess) rc = FINESS_ERR_HIP;
            else if (hipMemset(d, 0, (size_t)b->M * pr.plane * sizeof(double)) != hipSuccess) rc = FINESS_ERR_HIP;
            pr.buf[s] = d;
            b->bufs.push_back(d);
        }
        const size_t ref = (size_t)(mx[k] + 2 * p.mbc) * (my[k] + 2 * p.mbc) * b->M;
        if (ref > b->staging_n) b->staging_n = ref;
    }
    if (!rc && hipMalloc((void**)&b->staging, b->staging_n * sizeof(double)) != hipSuccess) rc = FINESS_ERR_HIP;
    if (!rc && launch_fill_u64(b->smax, nprob * 2 * BATCH_SPREAD, 0ull, b->st)) rc = FINESS_ERR_HIP;
    if (!rc) rc = set_shape(b, 0);
    if (rc) {
        finess_batch_destroy(b);
        return rc;
    }
    *out = b;
    return FINESS_OK;
}

void finess_batch_destroy(finess_batch* b)
{
    if (!b) return;
    if (b->st) hipStreamSynchronize(b->st);
    if (b->exec) hipGraphExecDestroy(b->exec);
    for (double* d : b->bufs) hipFree(d);
    hipFree(b->probs);
    hipFree(b->tile_prob);
    hipFree(b->smax);
    hipFree(b->staging);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->st) hipStreamDestroy(b->st);
    delete b;
}

const char* finess_batch_last_error(finess_batch* b) { return b ? b->err.c_str() : "null batch"; }

int finess_batch_set_state(finess_batch* b, int k, const double* q, double t, double dt)
{
    if (!b || !q || k < 0 || k >= b->n || !(dt > 0.0)) return FINESS_ERR_ARG;
    BatchProb& pr = b->host[k];
    const int g = b->p.mbc, M = b->M;
    const long long ld = pr.my + 2 * g;
    std::vector<double> planes((size_t)M * pr.plane);
    for (int i = 0; i < pr.mx; i++)
        for (int j = 0; j < pr.my; j++)
            for (int m = 0; m < M; m++) {
                const double v = q[((i + g) * ld + (j + g)) * M + m];
                if (!isfinite(v)) return bfail(b, FINESS_ERR_NONFINITE, "non-finite initial state");
                planes[(size_t)m * pr.plane + (size_t)i * pr.my + j] = v;
            }
    int rc = download_table(b);
    if (rc) return rc;
    pr.cur = 0;
    pr.t = t;
    pr.dt = dt;
    pr.accepted = pr.rejected = pr.target = 0;
    pr.err = 0;
    pr.done = 1;
    if (hipMemcpy(pr.buf[0], planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(b, FINESS_ERR_HIP, "state upload failed");
    return upload_table(b);
}

int finess_batch_get_state(finess_batch* b, int k, double* q, double* t, double* dt, long long* accepted,
                           long long* rejected)
{
    if (!b || k < 0 || k >= b->n) return FINESS_ERR_ARG;
    int rc = download_table(b);
    if (rc) return rc;
    const BatchProb& pr = b->host[k];
    if (q) {
        // momentum components normal to x and y walls (both physics): 1 and 2
        if (launch_planes_to_ref(pr.buf[pr.cur], b->staging, pr.mx, pr.my, b->M, b->p.mbc, b->p.bc, 1, 2, b->st))
            return bfail(b, FINESS_ERR_HIP, "layout conversion failed");
        const size_t n = (size_t)(pr.mx + 2 * b->p.mbc) * (pr.my + 2 * b->p.mbc) * b->M;
        if (hipMemcpyAsync(q, b->staging, n * sizeof(double), hipMemcpyDeviceToHost, b->st) != hipSuccess ||
            hipStreamSynchronize(b->st) != hipSuccess)
            return bfail(b, FINESS_ERR_HIP, "state download failed");
    }
    if (t) *t = pr.t;
    if (dt) *dt = pr.dt;
    if (accepted) *accepted = pr.accepted;
    if (rejected) *rejected = pr.rejected;
    return pr.err;
}

int finess_batch_autotune(finess_batch* b, int trial_launches, int* tile_i, int* tile_j)
{
    if (!b || trial_launches < 1) return FINESS_ERR_ARG;
    int rc = download_table(b);
    if (rc) return rc;
    const std::vector<BatchProb> saved = b->host;
    for (BatchProb& pr : b->host) pr.done = pr.err ? 1 : 0;
    int best = -1;
    float best_ms = 0.f;
    for (int s = 0; s < BATCH_NSHAPES; s++) {
        if ((rc = set_shape(b, s))) return rc;
        const BatchArgs a = args_of(b);
        // stage 1 writes only the scratch plane set and the wave-speed words
        if (launch_batch_stage(b->physics, b->fp32, s, 1, a, b->grid, b->st))
            return bfail(b, FINESS_ERR_HIP, "stage launch failed");
        hipEventRecord(b->ev0, b->st);
        for (int r = 0; r < trial_launches; r++)
            if (launch_batch_stage(b->physics, b->fp32, s, 1, a, b->grid, b->st))
                return bfail(b, FINESS_ERR_HIP, "stage launch failed");
        hipEventRecord(b->ev1, b->st);
        if (hipEventSynchronize(b->ev1) != hipSuccess) return bfail(b, FINESS_ERR_HIP, "autotune timing failed");
        float ms = 0.f;
        hipEventElapsedTime(&ms, b->ev0, b->ev1);
        if (b->p.verbosity) printf("finess-hip batch autotune: tile %dx%d  %.3f ms per stage\n", BATCH_SHAPES[s][0],
                                   BATCH_SHAPES[s][1], ms / trial_launches);
        if (best < 0 || ms < best_ms) { best = s; best_ms = ms; }
    }
    if (launch_fill_u64(b->smax, b->n * 2 * BATCH_SPREAD, 0ull, b->st)) return bfail(b, FINESS_ERR_HIP, "reset failed");
    for (int k = 0; k < b->n; k++) {  // controller state as it was, tile geometry of the winner
        b->host[k] = saved[k];
    }
    if ((rc = set_shape(b, best))) return rc;
    if (tile_i) *tile_i = BATCH_SHAPES[best][0];
    if (tile_j) *tile_j = BATCH_SHAPES[best][1];
    return FINESS_OK;
}

int finess_batch_run(finess_batch* b, int steps, double* seconds)
{
    if (!b || steps < 0) return FINESS_ERR_ARG;
    int rc = download_table(b);
    if (rc) return rc;
    for (BatchProb& pr : b->host) {
        if (pr.err) continue;
        pr.target = pr.accepted + steps;
        pr.done = steps > 0 ? 0 : 1;
    }
    if ((rc = upload_table(b))) return rc;
    if (steps == 0) {
        if (seconds) *seconds = 0.0;
        return FINESS_OK;
    }
    const int chunk = steps < 50 ? steps : 50;
    if ((rc = ensure_graph(b, chunk))) return rc;
    hipEventRecord(b->ev0, b->st);
    // every replay advances each unfinished problem by up to `chunk` attempts;
    // rejected attempts need further replays (up to three per accepted step)
    const long long max_replays = 4LL * steps / chunk + 2;
    bool all_done = false;
    for (long long r = 0; r < max_replays && !all_done; r++) {
        if (hipGraphLaunch(b->exec, b->st) != hipSuccess) return bfail(b, FINESS_ERR_HIP, "graph launch failed");
        if ((rc = download_table(b))) return rc;
        all_done = true;
        for (const BatchProb& pr : b->host) all_done = all_done && pr.done;
    }
    hipEventRecord(b->ev1, b->st);
    hipEventSynchronize(b->ev1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, b->ev0, b->ev1);
    if (seconds) *seconds = ms * 1e-3;
    for (int k = 0; k < b->n; k++)
        if (b->host[k].err) {
            char msg[96];
            snprintf(msg, sizeof msg, "problem %d stopped with status %d", k, b->host[k].err);
            return bfail(b, b->host[k].err, msg);
        }
    if (!all_done) return bfail(b, FINESS_ERR_NV, "exceeded allowed # of time steps");
    return FINESS_OK;
}

} // extern "C"
