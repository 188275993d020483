// Batched GF solve over independent instances with uneven sizes (bias points
// or structures, each with its own grid: N_E, N_k and N_b may all differ).
//
// The reference solves one problem per call (solve_production,
// src/QCLNEGF.jl:982-983); a sweep calls it once per instance. Here the
// instances are grouped by tile class (N_b <= 8B) and each kernel of
// current_green runs ONE launch per class over the tiles of all instances
// of that class, so small instances share the device with large ones
// instead of each leaving it partly idle. Instance tables live in device
// memory; the number functionals and paired maps of all instances are one
// launch each.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dk_common.hpp"

using namespace qclhip;

struct qclhip_batch {
    int ninst = 0;
    std::vector<const qclhip_grid*> grids;
    // Per tile class B = 1..6: instances, and tile prefix sums for the Dyson
    // and the Keldysh tile widths.
    std::vector<int> cls_inst[7];
    std::vector<long long> cls_begin[2][7];
    int* d_inst[7] = {};
    long long* d_begin[2][7] = {};
    DKArgs* d_args = nullptr;
    DKArgs* h_args = nullptr;  // pinned staging
    BatchScalars* d_scal = nullptr;
    BatchScalars* h_scal = nullptr;  // pinned staging
    PairCoefficients* d_coef = nullptr;
    c64* traces = nullptr;
    std::vector<long long> trace_off;
    long long maxn = 0;
    // The instance tables last staged to the device: a call with the same
    // fields and targets launches without restaging (no host wait), which is
    // what lets a captured graph contain the batched solve.
    std::vector<DKArgs> last_args;
    std::vector<BatchScalars> last_scal;
    bool staged = false;
    // QCLHIP_BATCH_PERSISTENT: one launch per kernel and tile class over all
    // instances; QCLHIP_BATCH_PER_INSTANCE: the single-instance launches
    // (register prefetch) per instance, then the batched scalars.
    int mode = QCLHIP_BATCH_PERSISTENT;
};

namespace {
// (lambda, number, holes, a, c) of every instance into the deferred buffer.
__global__ void k_copy_coefficients_batch(const PairCoefficients* c, int n, double* dst) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        dst[5 * i + 0] = c[i].lambda;
        dst[5 * i + 1] = c[i].number;
        dst[5 * i + 2] = c[i].holes;
        dst[5 * i + 3] = c[i].a;
        dst[5 * i + 4] = c[i].c;
    }
}
}  // namespace

static int tile_class(int N) { return (N + 7) / 8; }

extern "C" {

int qclhip_batch_destroy(qclhip_batch* b) {
    if (!b) return QCLHIP_OK;
    hipDeviceSynchronize();
    for (int c = 0; c < 7; ++c) {
        if (b->d_inst[c]) hipFree(b->d_inst[c]);
        for (int k = 0; k < 2; ++k)
            if (b->d_begin[k][c]) hipFree(b->d_begin[k][c]);
    }
    if (b->d_args) hipFree(b->d_args);
    if (b->h_args) hipHostFree(b->h_args);
    if (b->d_scal) hipFree(b->d_scal);
    if (b->h_scal) hipHostFree(b->h_scal);
    if (b->d_coef) hipFree(b->d_coef);
    if (b->traces) hipFree(b->traces);
    delete b;
    return QCLHIP_OK;
}

int qclhip_batch_create(qclhip_ctx* ctx, int ninst, const qclhip_grid* const* grids, qclhip_batch** out) {
    if (!ctx || !grids || !out) return QCLHIP_ERR_ARGUMENT;
    *out = nullptr;
    if (ninst < 1) return fail(ctx, QCLHIP_ERR_ARGUMENT, "a batch needs at least one instance");
    auto* b = new qclhip_batch();
    b->ninst = ninst;
    long long trace_total = 0;
    for (int i = 0; i < ninst; ++i) {
        const qclhip_grid* g = grids[i];
        if (!g) {
            qclhip_batch_destroy(b);
            return fail(ctx, QCLHIP_ERR_ARGUMENT, "null grid in batch");
        }
        if (g->Nb > 48) {
            qclhip_batch_destroy(b);
            return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");
        }
        b->grids.push_back(g);
        const int B = tile_class(g->Nb);
        b->cls_inst[B].push_back(i);
        b->trace_off.push_back(trace_total);
        trace_total += 2LL * g->NE * g->Nk;
        b->maxn = std::max(b->maxn, (long long)g->NE * g->Nk * g->Nb * g->Nb);
    }
    bool ok = true;
    for (int B = 1; B <= 6 && ok; ++B) {
        if (b->cls_inst[B].empty()) continue;
        for (int k = 0; k < 2; ++k) {
            const int W = batch_tile_waves(B, k == 1);
            std::vector<long long>& beg = b->cls_begin[k][B];
            beg.push_back(0);
            for (int i : b->cls_inst[B]) {
                const qclhip_grid* g = b->grids[i];
                beg.push_back(beg.back() + (long long)((g->NE + W - 1) / W) * g->Nk);
            }
            ok = ok && hipMalloc(&b->d_begin[k][B], beg.size() * sizeof(long long)) == hipSuccess &&
                 hipMemcpy(b->d_begin[k][B], beg.data(), beg.size() * sizeof(long long), hipMemcpyHostToDevice) ==
                     hipSuccess;
        }
        ok = ok && hipMalloc(&b->d_inst[B], b->cls_inst[B].size() * sizeof(int)) == hipSuccess &&
             hipMemcpy(b->d_inst[B], b->cls_inst[B].data(), b->cls_inst[B].size() * sizeof(int),
                       hipMemcpyHostToDevice) == hipSuccess;
    }
    ok = ok && hipMalloc(&b->d_args, ninst * sizeof(DKArgs)) == hipSuccess &&
         hipHostMalloc(&b->h_args, ninst * sizeof(DKArgs)) == hipSuccess &&
         hipMalloc(&b->d_scal, ninst * sizeof(BatchScalars)) == hipSuccess &&
         hipHostMalloc(&b->h_scal, ninst * sizeof(BatchScalars)) == hipSuccess &&
         hipMalloc(&b->d_coef, ninst * sizeof(PairCoefficients)) == hipSuccess &&
         hipMalloc(&b->traces, trace_total * sizeof(c64)) == hipSuccess;
    if (!ok) {
        qclhip_batch_destroy(b);
        return fail(ctx, QCLHIP_ERR_RUNTIME, "batch allocation failed");
    }
    *out = b;
    return QCLHIP_OK;
}

int qclhip_batch_current_green(qclhip_ctx* ctx, qclhip_batch* b, const double* const* h, int nfam,
                               const double* const* famR, const double* const* famL, const double* const* famG,
                               const double* target_host, double* const* GR, double* const* GL, double* const* GG,
                               double* out_host) {
    if (!ctx || !b || !h || !famR || !famL || !famG || !target_host || !GR || !GL || !GG)
        return QCLHIP_ERR_ARGUMENT;
    if (nfam < 1 || nfam > kMaxFamilies)
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "the batched solve takes 1..8 self-energy families per instance");
    std::vector<DKArgs> args(b->ninst);
    std::vector<BatchScalars> scal(b->ninst);
    for (int i = 0; i < b->ninst; ++i) {
        const qclhip_grid* g = b->grids[i];
        if (!h[i] || !GR[i] || !GL[i] || !GG[i]) return fail(ctx, QCLHIP_ERR_ARGUMENT, "null field in batch");
        if (!(std::isfinite(target_host[i]) && target_host[i] > 0))
            return fail(ctx, QCLHIP_ERR_ARGUMENT, "target charge must be finite and positive");
        DKArgs a{};
        a.NE = g->NE;
        a.Nk = g->Nk;
        a.N = g->Nb;
        a.eps = g->eps;
        a.h = (const c64*)h[i];
        for (int f = 0; f < nfam; ++f) {
            const size_t k = (size_t)i * nfam + f;
            if (!famR[k] || !famL[k] || !famG[k]) return fail(ctx, QCLHIP_ERR_ARGUMENT, "null family in batch");
            a.famR[f] = (const c64*)famR[k];
            a.famL[f] = (const c64*)famL[k];
            a.famG[f] = (const c64*)famG[k];
        }
        a.nfamR = a.nfamL = a.nfamG = nfam;
        a.eta = 0.0;
        a.GR = (c64*)GR[i];
        a.GL = (c64*)GL[i];
        a.GG = (c64*)GG[i];
        a.scale = nullptr;
        a.traces = b->traces + b->trace_off[i];
        a.status = ctx->status;
        a.keldysh = 1;
        args[i] = a;
        BatchScalars sc{};
        sc.traces = a.traces;
        sc.NE = g->NE;
        sc.Nk = g->Nk;
        sc.gs = g->gs;
        sc.wE = g->wE;
        sc.wk = g->wk;
        sc.target = target_host[i];
        sc.limit = 1.0e-10;  // the reference's realness limit (numerical_invariants.jl:5)
        sc.coef = b->d_coef + i;
        sc.GL = a.GL;
        sc.GG = a.GG;
        sc.n = (long long)g->NE * g->Nk * g->Nb * g->Nb;
        scal[i] = sc;
    }
    const bool same = b->staged && std::memcmp(args.data(), b->last_args.data(), b->ninst * sizeof(DKArgs)) == 0 &&
                      std::memcmp(scal.data(), b->last_scal.data(), b->ninst * sizeof(BatchScalars)) == 0;
    if (!same) {
        // The staging buffers may still feed the previous call's copies; the
        // wait below cannot be captured.
        if (ctx->capturing)
            return fail(ctx, QCLHIP_ERR_ARGUMENT,
                        "the batched solve restages its instance tables (new fields or targets) and cannot do so "
                        "during graph capture; run it once with these arguments before capturing");
        QCLHIP_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        std::memcpy(b->h_args, args.data(), b->ninst * sizeof(DKArgs));
        std::memcpy(b->h_scal, scal.data(), b->ninst * sizeof(BatchScalars));
        QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(b->d_args, b->h_args, b->ninst * sizeof(DKArgs), hipMemcpyHostToDevice,
                                           ctx->stream));
        QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(b->d_scal, b->h_scal, b->ninst * sizeof(BatchScalars),
                                           hipMemcpyHostToDevice, ctx->stream));
        b->last_args = args;
        b->last_scal = scal;
        b->staged = true;
    }
    if (b->mode == QCLHIP_BATCH_PER_INSTANCE) {
        for (int i = 0; i < b->ninst; ++i) {
            int rc = launch_dyson_fp64(ctx, args[i]);
            if (rc) return rc;
            rc = launch_keldysh_valu(ctx, args[i]);
            if (rc) return rc;
        }
    } else {
        for (int k = 0; k < 2; ++k) {
            for (int B = 1; B <= 6; ++B) {
                if (b->cls_inst[B].empty()) continue;
                BatchClass c;
                c.insts = b->d_args;
                c.inst = b->d_inst[B];
                c.begin = b->d_begin[k][B];
                c.n = (int)b->cls_inst[B].size();
                int rc = launch_dk_batch(ctx, B, k == 1, c, b->cls_begin[k][B].back());
                if (rc) return rc;
            }
        }
    }
    int rc = launch_pair_batch(ctx, b->d_scal, b->ninst, b->maxn);
    if (rc) return rc;
    if (out_host && ctx->defer_dev) {
        if (ctx->defer_next + 5 * b->ninst > ctx->defer_cap)
            return fail(ctx, QCLHIP_ERR_ARGUMENT, "deferred-result buffer is full");
        hipLaunchKernelGGL(k_copy_coefficients_batch, dim3(1), dim3(256), 0, ctx->stream, b->d_coef, b->ninst,
                           ctx->defer_dev + ctx->defer_next);
        ctx->defer_next += 5 * b->ninst;
        return launch_check(ctx, "k_copy_coefficients_batch");
    }
    if (out_host) {
        std::vector<PairCoefficients> pc(b->ninst);
        QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(pc.data(), b->d_coef, b->ninst * sizeof(PairCoefficients),
                                           hipMemcpyDeviceToHost, ctx->stream));
        rc = qclhip_ctx_synchronize(ctx);
        if (rc) return rc;
        for (int i = 0; i < b->ninst; ++i) {
            out_host[5 * i + 0] = pc[i].lambda;
            out_host[5 * i + 1] = pc[i].number;
            out_host[5 * i + 2] = pc[i].holes;
            out_host[5 * i + 3] = pc[i].a;
            out_host[5 * i + 4] = pc[i].c;
        }
    }
    return QCLHIP_OK;
}

int qclhip_batch_set_mode(qclhip_ctx* ctx, qclhip_batch* b, int mode) {
    if (!ctx || !b) return QCLHIP_ERR_ARGUMENT;
    if (mode != QCLHIP_BATCH_PERSISTENT && mode != QCLHIP_BATCH_PER_INSTANCE)
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "mode must be QCLHIP_BATCH_PERSISTENT or QCLHIP_BATCH_PER_INSTANCE");
    b->mode = mode;
    return QCLHIP_OK;
}

int qclhip_batch_get_mode(const qclhip_batch* b) { return b ? b->mode : QCLHIP_ERR_ARGUMENT; }

int qclhip_batch_autotune(qclhip_ctx* ctx, qclhip_batch* b, const double* const* h, int nfam,
                          const double* const* famR, const double* const* famL, const double* const* famG,
                          const double* target_host, double* const* GR, double* const* GL, double* const* GG,
                          double* ms_host) {
    if (!ctx || !b) return QCLHIP_ERR_ARGUMENT;
    if (ctx->capturing) return fail(ctx, QCLHIP_ERR_ARGUMENT, "the autotuner times on the stream; not during capture");
    double* saved_defer = ctx->defer_dev;
    if (saved_defer) return fail(ctx, QCLHIP_ERR_ARGUMENT, "the autotuner does not run in deferred-results mode");
    hipEvent_t e0, e1;
    QCLHIP_TRY_HIP(ctx, hipEventCreate(&e0));
    QCLHIP_TRY_HIP(ctx, hipEventCreate(&e1));
    const int modes[2] = {QCLHIP_BATCH_PERSISTENT, QCLHIP_BATCH_PER_INSTANCE};
    float ms[2] = {0.f, 0.f};
    int rc = QCLHIP_OK;
    for (int k = 0; k < 2 && rc == QCLHIP_OK; ++k) {
        b->mode = modes[k];
        rc = qclhip_batch_current_green(ctx, b, h, nfam, famR, famL, famG, target_host, GR, GL, GG, nullptr);
        if (rc) break;
        hipEventRecord(e0, ctx->stream);
        for (int r = 0; r < 2 && rc == QCLHIP_OK; ++r)
            rc = qclhip_batch_current_green(ctx, b, h, nfam, famR, famL, famG, target_host, GR, GL, GG, nullptr);
        hipEventRecord(e1, ctx->stream);
        if (rc) break;
        rc = qclhip_ctx_synchronize(ctx);
        if (rc) break;
        hipEventElapsedTime(&ms[k], e0, e1);
        ms[k] /= 2.f;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    b->mode = ms[1] < ms[0] ? QCLHIP_BATCH_PER_INSTANCE : QCLHIP_BATCH_PERSISTENT;
    if (ms_host) {
        ms_host[0] = ms[0];
        ms_host[1] = ms[1];
    }
    return QCLHIP_OK;
}

}  // extern "C"
