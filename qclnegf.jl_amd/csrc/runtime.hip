// Context, grid and shift-plan management of the qclhip C-ABI.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "common.hpp"

namespace qclhip {

int fail(qclhip_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

int hip_fail(qclhip_ctx* ctx, hipError_t err, const char* what) {
    std::string msg = std::string(what) + ": " + hipGetErrorString(err);
    return fail(ctx, QCLHIP_ERR_RUNTIME, msg);
}

void* workspace(qclhip_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->workspace_bytes) return ctx->workspace;
    if (ctx->capturing) return nullptr;  // no allocation under graph capture
    if (ctx->workspace) {
        // Work already enqueued may still read the old scratch.
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) return nullptr;
        hipFree(ctx->workspace);
        ctx->workspace = nullptr;
        ctx->workspace_bytes = 0;
    }
    size_t want = std::max(bytes, (size_t)1 << 20);
    ++ctx->workspace_gen;
    if (hipMalloc(&ctx->workspace, want) != hipSuccess) {
        ctx->workspace = nullptr;
        return nullptr;
    }
    ctx->workspace_bytes = want;
    return ctx->workspace;
}

int persistent_grid(qclhip_ctx* ctx, const void* kernel, int threads, size_t lds, long long ntiles, unsigned* grid) {
    int per_cu = 0;
    QCLHIP_TRY_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
    long long g = (long long)std::max(per_cu, 1) * std::max(ctx->sm_count, 1);
    g = std::max<long long>(8, g / 8 * 8);
    *grid = (unsigned)std::max<long long>(1, std::min(g, ntiles));
    return QCLHIP_OK;
}

__global__ void k_copy_doubles(const double* src, int n, double* dst) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

__global__ void k_copy_coefficients(const PairCoefficients* c, double* dst) {
    if (threadIdx.x == 0) {
        dst[0] = c->lambda;
        dst[1] = c->number;
        dst[2] = c->holes;
        dst[3] = c->a;
        dst[4] = c->c;
    }
}

int defer_result(qclhip_ctx* ctx, const double* src, int n, bool* deferred) {
    *deferred = false;
    if (!ctx->defer_dev) return QCLHIP_OK;
    if (ctx->defer_next + n > ctx->defer_cap) return fail(ctx, QCLHIP_ERR_ARGUMENT, "deferred-result buffer is full");
    hipLaunchKernelGGL(k_copy_doubles, dim3(1), dim3(256), 0, ctx->stream, src, n, ctx->defer_dev + ctx->defer_next);
    ctx->defer_next += n;
    *deferred = true;
    return launch_check(ctx, "k_copy_doubles");
}

int defer_coefficients(qclhip_ctx* ctx, const PairCoefficients* src, bool* deferred) {
    *deferred = false;
    if (!ctx->defer_dev) return QCLHIP_OK;
    if (ctx->defer_next + 5 > ctx->defer_cap) return fail(ctx, QCLHIP_ERR_ARGUMENT, "deferred-result buffer is full");
    hipLaunchKernelGGL(k_copy_coefficients, dim3(1), dim3(64), 0, ctx->stream, src, ctx->defer_dev + ctx->defer_next);
    ctx->defer_next += 5;
    *deferred = true;
    return launch_check(ctx, "k_copy_coefficients");
}

int launch_check(qclhip_ctx* ctx, const char* what) {
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return hip_fail(ctx, err, what);
    return QCLHIP_OK;
}

}  // namespace qclhip

using namespace qclhip;

extern "C" {

int qclhip_ctx_create(int device, void* stream, qclhip_ctx** out) {
    if (!out) return QCLHIP_ERR_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return QCLHIP_ERR_RUNTIME;
    if (device < 0 || device >= count) return QCLHIP_ERR_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return QCLHIP_ERR_RUNTIME;
    // the fixed scratch of the allocation-free entry points, once per device
    int src = static_scratch_setup(device);
    if (src != QCLHIP_OK) return src;
    auto* ctx = new qclhip_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete ctx;
        return QCLHIP_ERR_RUNTIME;
    }
    ctx->sm_count = prop.multiProcessorCount;
    ctx->max_lds_per_block = prop.sharedMemPerBlock;
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, device) == hipSuccess &&
        (size_t)optin > ctx->max_lds_per_block)
        ctx->max_lds_per_block = (size_t)optin;
    // The caller's stream; NULL is HIP's default stream (torch's default
    // stream), so work stays ordered with the caller's own kernels and copies.
    ctx->stream = (hipStream_t)stream;
    ctx->owns_stream = false;
    bool ok = hipMalloc(&ctx->status, sizeof(DeviceStatus)) == hipSuccess;
    ok = ok && hipMalloc(&ctx->coef, sizeof(PairCoefficients)) == hipSuccess;
    ok = ok && hipHostMalloc(&ctx->status_host, sizeof(DeviceStatus), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc(&ctx->host_scalars, 64 * sizeof(double), hipHostMallocDefault) == hipSuccess;
    if (ok) {
        DeviceStatus zero{0, 0, ~0ull};
        ok = hipMemcpy(ctx->status, &zero, sizeof(zero), hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
        qclhip_ctx_destroy(ctx);
        return QCLHIP_ERR_RUNTIME;
    }
    *out = ctx;
    return QCLHIP_OK;
}

int qclhip_ctx_destroy(qclhip_ctx* ctx) {
    if (!ctx) return QCLHIP_OK;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    if (ctx->workspace) hipFree(ctx->workspace);
    if (ctx->status) hipFree(ctx->status);
    if (ctx->coef) hipFree(ctx->coef);
    if (ctx->status_host) hipHostFree(ctx->status_host);
    if (ctx->host_scalars) hipHostFree(ctx->host_scalars);
    if (ctx->owns_stream && ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return QCLHIP_OK;
}

// The latched device status is read back and cleared here. Singular blocks
// win over other codes, like the reference's SingularException.
int qclhip_ctx_set_stream(qclhip_ctx* ctx, void* stream) {
    if (!ctx) return QCLHIP_ERR_ARGUMENT;
    if (ctx->capturing) return fail(ctx, QCLHIP_ERR_ARGUMENT, "cannot change the stream during graph capture");
    ctx->stream = (hipStream_t)stream;
    return QCLHIP_OK;
}

int qclhip_ctx_defer_results(qclhip_ctx* ctx, double* dev_buffer, int capacity) {
    if (!ctx || (dev_buffer && capacity <= 0)) return QCLHIP_ERR_ARGUMENT;
    ctx->defer_dev = dev_buffer;
    ctx->defer_cap = dev_buffer ? capacity : 0;
    ctx->defer_next = 0;
    return QCLHIP_OK;
}

int qclhip_ctx_result_slot(const qclhip_ctx* ctx) { return ctx ? ctx->defer_next : QCLHIP_ERR_ARGUMENT; }

struct qclhip_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // The capturing context and its scratch generation at capture: the graph's
    // kernels hold that scratch address.
    const qclhip_ctx* ctx = nullptr;
    unsigned long long workspace_gen = 0;
};

int qclhip_graph_begin(qclhip_ctx* ctx) {
    if (!ctx) return QCLHIP_ERR_ARGUMENT;
    if (!ctx->stream) return fail(ctx, QCLHIP_ERR_ARGUMENT, "the default stream cannot be captured; set a stream");
    if (ctx->capturing) return fail(ctx, QCLHIP_ERR_ARGUMENT, "capture already in progress");
    QCLHIP_TRY_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    return QCLHIP_OK;
}

int qclhip_graph_end(qclhip_ctx* ctx, qclhip_graph** out) {
    if (!ctx || !out) return QCLHIP_ERR_ARGUMENT;
    *out = nullptr;
    if (!ctx->capturing) return fail(ctx, QCLHIP_ERR_ARGUMENT, "no capture in progress");
    ctx->capturing = false;
    auto* g = new qclhip_graph();
    hipError_t err = hipStreamEndCapture(ctx->stream, &g->graph);
    if (err == hipSuccess) err = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (err != hipSuccess) {
        if (g->graph) hipGraphDestroy(g->graph);
        delete g;
        return hip_fail(ctx, err, "graph capture");
    }
    g->ctx = ctx;
    g->workspace_gen = ctx->workspace_gen;
    *out = g;
    return QCLHIP_OK;
}

int qclhip_graph_launch(qclhip_ctx* ctx, const qclhip_graph* g) {
    if (!ctx || !g) return QCLHIP_ERR_ARGUMENT;
    if (g->ctx != ctx) return fail(ctx, QCLHIP_ERR_ARGUMENT, "the graph was captured on another context");
    if (g->workspace_gen != ctx->workspace_gen)
        return fail(ctx, QCLHIP_ERR_ARGUMENT,
                    "the context's scratch was reallocated after the graph was captured (an eager call needed "
                    "more); capture the graph again");
    QCLHIP_TRY_HIP(ctx, hipGraphLaunch(g->exec, ctx->stream));
    return QCLHIP_OK;
}

int qclhip_graph_destroy(qclhip_graph* g) {
    if (!g) return QCLHIP_OK;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
    return QCLHIP_OK;
}

int qclhip_ctx_synchronize(qclhip_ctx* ctx) {
    if (!ctx) return QCLHIP_ERR_ARGUMENT;
    if (ctx->capturing) return fail(ctx, QCLHIP_ERR_ARGUMENT, "synchronizing entry called during graph capture");
    hipError_t err = hipMemcpyAsync(ctx->status_host, ctx->status, sizeof(DeviceStatus),
                                    hipMemcpyDeviceToHost, ctx->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
    if (err != hipSuccess) return hip_fail(ctx, err, "qclhip_ctx_synchronize");
    DeviceStatus st = *ctx->status_host;
    if (st.code == 0) return QCLHIP_OK;
    DeviceStatus zero{0, 0, ~0ull};
    err = hipMemcpyAsync(ctx->status, &zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
    if (err != hipSuccess) return hip_fail(ctx, err, "qclhip_ctx_synchronize(clear)");
    if (st.code > 0) {
        unsigned long long b = st.block;
        ctx->last_error = "singular Dyson block " + std::to_string(b);
        return b > 0x7fffffffull ? 0x7fffffff : (int)b;
    }
    switch (st.code) {
        case QCLHIP_ERR_DOMAIN:
            ctx->last_error = "domain error (number functional, lambda, Pauli capacity or a singular Poisson system)";
            break;
        default: ctx->last_error = "device error"; break;
    }
    return st.code;
}

const char* qclhip_ctx_last_error(const qclhip_ctx* ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

int qclhip_grid_create(qclhip_ctx* ctx, int NE, int Nk, int Nb, int gs, const double* eps_host,
                       const double* wE_host, const double* wk_host, qclhip_grid** out) {
    if (!ctx || !out || !eps_host || !wE_host || !wk_host) return QCLHIP_ERR_ARGUMENT;
    *out = nullptr;
    if (NE < 1 || Nk < 1 || Nb < 1) return fail(ctx, QCLHIP_ERR_DIMENSION, "grid sizes must be positive");
    if (gs <= 0) return fail(ctx, QCLHIP_ERR_ARGUMENT, "g_s must be positive");
    auto* g = new qclhip_grid();
    g->NE = NE;
    g->Nk = Nk;
    g->Nb = Nb;
    g->gs = gs;
    g->eps_host.assign(eps_host, eps_host + NE);
    g->wE_host.assign(wE_host, wE_host + NE);
    g->wk_host.assign(wk_host, wk_host + Nk);
    // Uniform spacing test of _uniform_energy_spacing (production_operators.jl:724-733).
    g->de = 0.0;
    if (NE >= 2) {
        double d = eps_host[1] - eps_host[0], dev = 0.0, emax = 0.0;
        for (int e = 1; e < NE; ++e) dev = std::max(dev, std::fabs((eps_host[e] - eps_host[e - 1]) - d));
        for (int e = 0; e < NE; ++e) emax = std::max(emax, std::fabs(eps_host[e]));
        double tol = 256.0 * 2.220446049250313e-16 * std::max(std::max(emax, std::fabs(d)), 1.0);
        if (d > 0 && dev <= tol) g->de = d;
    }
    // Cancellation-aware order of the current (production_iteration.jl:191-194).
    std::vector<int> order(NE);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        double ax = std::fabs(eps_host[x]), ay = std::fabs(eps_host[y]);
        if (ax != ay) return ax < ay;
        return eps_host[x] < eps_host[y];
    });
    bool ok = hipMalloc(&g->eps, NE * sizeof(double)) == hipSuccess &&
              hipMalloc(&g->wE, NE * sizeof(double)) == hipSuccess &&
              hipMalloc(&g->wk, Nk * sizeof(double)) == hipSuccess &&
              hipMalloc(&g->flux_order, NE * sizeof(int)) == hipSuccess;
    ok = ok && hipMemcpy(g->eps, eps_host, NE * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(g->wE, wE_host, NE * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(g->wk, wk_host, Nk * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(g->flux_order, order.data(), NE * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        qclhip_grid_destroy(g);
        return fail(ctx, QCLHIP_ERR_RUNTIME, "grid allocation failed");
    }
    *out = g;
    return QCLHIP_OK;
}

int qclhip_grid_destroy(qclhip_grid* g) {
    if (!g) return QCLHIP_OK;
    hipDeviceSynchronize();
    if (g->eps) hipFree(g->eps);
    if (g->wE) hipFree(g->wE);
    if (g->wk) hipFree(g->wk);
    if (g->flux_order) hipFree(g->flux_order);
    if (g->block_flags) hipFree(g->block_flags);
    delete g;
    return QCLHIP_OK;
}

int qclhip_grid_set_precision(qclhip_ctx* ctx, qclhip_grid* g, int precision, int refine_steps, double refine_tol) {
    if (!ctx || !g) return QCLHIP_ERR_ARGUMENT;
    if (precision != QCLHIP_PRECISION_FP64 && precision != QCLHIP_PRECISION_FP32_REFINE &&
        precision != QCLHIP_PRECISION_FP32_PRODUCTS)
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "precision must be QCLHIP_PRECISION_FP64, _FP32_REFINE or _FP32_PRODUCTS");
    if (refine_steps < 0 || refine_steps > 4) return fail(ctx, QCLHIP_ERR_ARGUMENT, "refine_steps must be in [0, 4]");
    if (!(refine_tol <= 0.0 || std::isfinite(refine_tol)))
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "refine_tol must be finite");
    if (precision != QCLHIP_PRECISION_FP64 && !g->block_flags)
        QCLHIP_TRY_HIP(ctx, hipMalloc(&g->block_flags, (size_t)g->NE * g->Nk * sizeof(unsigned)));
    g->precision = precision;
    g->refine_steps = refine_steps;
    g->refine_tol = refine_tol > 0.0 ? refine_tol : 1e-10;
    return QCLHIP_OK;
}

int qclhip_grid_set_engine(qclhip_ctx* ctx, qclhip_grid* g, int keldysh_engine) {
    if (!ctx || !g) return QCLHIP_ERR_ARGUMENT;
    if (keldysh_engine != QCLHIP_ENGINE_VECTOR && keldysh_engine != QCLHIP_ENGINE_MATRIX &&
        keldysh_engine != QCLHIP_ENGINE_FUSED)
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "engine must be QCLHIP_ENGINE_VECTOR, _MATRIX or _FUSED");
    g->keldysh_engine = keldysh_engine;
    return QCLHIP_OK;
}

// build_shift_plan (production_cache.jl:258-288), 0-based indices.
int qclhip_shift_create(qclhip_ctx* ctx, const qclhip_grid* grid, double delta, qclhip_shift** out) {
    if (!ctx || !grid || !out) return QCLHIP_ERR_ARGUMENT;
    *out = nullptr;
    const int N = grid->NE;
    const std::vector<double>& eps = grid->eps_host;
    if (N < 2) return fail(ctx, QCLHIP_ERR_ARGUMENT, "at least two energy nodes are required");
    for (int e = 0; e < N; ++e)
        if (!std::isfinite(eps[e])) return fail(ctx, QCLHIP_ERR_ARGUMENT, "energy nodes must be finite");
    for (int e = 1; e < N; ++e)
        if (!(eps[e] - eps[e - 1] > 0))
            return fail(ctx, QCLHIP_ERR_ARGUMENT, "energy nodes must be strictly increasing");
    if (!std::isfinite(delta)) return fail(ctx, QCLHIP_ERR_ARGUMENT, "energy displacement must be finite");
    std::vector<int> left(N, 0), right(N, 0);
    std::vector<double> wr(N, 0.0);
    std::vector<unsigned char> valid(N, 0);
    int j = 0;
    for (int e = 0; e < N; ++e) {
        double target = eps[e] + delta;
        if (target < eps[0] || target > eps[N - 1]) continue;
        if (target == eps[N - 1]) {
            left[e] = right[e] = N - 1;
            valid[e] = 1;
            continue;
        }
        while (j < N - 2 && eps[j + 1] <= target) ++j;
        left[e] = j;
        right[e] = j + 1;
        wr[e] = (target - eps[j]) / (eps[j + 1] - eps[j]);
        valid[e] = 1;
    }
    // Transpose (adjoint) of the interpolation as a gather per source node j:
    // X[j] += sum over valid e with left_e = j of (1 - wr_e) Y[e] and with
    // right_e = j of wr_e Y[e], in ascending e (deterministic).
    std::vector<std::vector<std::pair<int, double>>> rows(N);
    for (int e = 0; e < N; ++e) {
        if (!valid[e]) continue;
        if (1.0 - wr[e] != 0.0) rows[left[e]].push_back({e, 1.0 - wr[e]});
        if (wr[e] != 0.0) rows[right[e]].push_back({e, wr[e]});
    }
    std::vector<int> tptr(N + 1, 0), tsrc;
    std::vector<double> tw;
    for (int jj = 0; jj < N; ++jj) {
        for (auto& pr : rows[jj]) {
            tsrc.push_back(pr.first);
            tw.push_back(pr.second);
        }
        tptr[jj + 1] = (int)tsrc.size();
    }
    if (tsrc.empty()) {  // keep device allocations non-empty
        tsrc.push_back(0);
        tw.push_back(0.0);
    }
    auto* s = new qclhip_shift();
    s->NE = N;
    s->delta = delta;
    bool ok = hipMalloc(&s->tptr, (N + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&s->tsrc, tsrc.size() * sizeof(int)) == hipSuccess &&
              hipMalloc(&s->tw, tw.size() * sizeof(double)) == hipSuccess &&
              hipMemcpy(s->tptr, tptr.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s->tsrc, tsrc.data(), tsrc.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(s->tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMalloc(&s->left, N * sizeof(int)) == hipSuccess &&
              hipMalloc(&s->right, N * sizeof(int)) == hipSuccess &&
              hipMalloc(&s->wr, N * sizeof(double)) == hipSuccess &&
              hipMalloc(&s->valid, N) == hipSuccess;
    ok = ok && hipMemcpy(s->left, left.data(), N * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(s->right, right.data(), N * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(s->wr, wr.data(), N * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(s->valid, valid.data(), N, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        qclhip_shift_destroy(s);
        return fail(ctx, QCLHIP_ERR_RUNTIME, "shift allocation failed");
    }
    *out = s;
    return QCLHIP_OK;
}

int qclhip_shift_destroy(qclhip_shift* s) {
    if (!s) return QCLHIP_OK;
    hipDeviceSynchronize();
    if (s->left) hipFree(s->left);
    if (s->right) hipFree(s->right);
    if (s->wr) hipFree(s->wr);
    if (s->valid) hipFree(s->valid);
    if (s->tptr) hipFree(s->tptr);
    if (s->tsrc) hipFree(s->tsrc);
    if (s->tw) hipFree(s->tw);
    delete s;
    return QCLHIP_OK;
}

}  // extern "C"
