// Autotuner of the GF-solve launch configuration. For a grid shape it times
// the FP64 Keldysh engines (vector ALUs, matrix cores), the tile shapes of
// the vector kernels and the fused Dyson + Keldysh kernel on synthetic blocks
// of that shape and keeps the fastest launch set for the grid, caching the
// choice per (Nb, NE, Nk) in the context. Timing uses a sub-grid of at most 2048 energies
// at the grid's Nk and Nb: the kernels' throughput per block depends on the
// block size and occupancy, not on the number of tiles once the device is
// full.
#include <algorithm>
#include <cmath>

#include "dk_common.hpp"

namespace qclhip {

// Synthetic, well-conditioned inputs: h diagonal, Sigma^R = -i(0.05 + u) I,
// Sigma< = i f g I, Sigma> = -i (1-f) g I, with u, f, g from a hash.
__device__ inline double hash01(unsigned long long x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void k_tune_fields(long long nblk, int N, c64* SR, c64* SL, c64* SG) {
    const long long n = nblk * N * N;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long blk = i % nblk;
        const long long ab = i / nblk;
        const bool diag = (ab % N) == (ab / N);
        const double u = hash01(2 * i + 1), f = hash01(2 * i + 2);
        SR[i] = diag ? cmk(0.0, -(0.05 + 0.1 * u)) : cmk(1e-3 * (u - 0.5), 0.0);
        SL[i] = diag ? cmk(0.0, f * 0.1) : cmk(0.0, 0.0);
        SG[i] = diag ? cmk(0.0, -(1.0 - f) * 0.1) : cmk(0.0, 0.0);
        (void)blk;
    }
}

__global__ void k_tune_h(int Nk, int N, c64* h) {
    const int n = Nk * N * N;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int ab = i / Nk;
        h[i] = ((ab % N) == (ab / N)) ? cmk(0.2 * (ab % N) / N, 0.0) : cmk(0.0, 0.0);
    }
}

}  // namespace qclhip

using namespace qclhip;

extern "C" {

int qclhip_grid_autotune(qclhip_ctx* ctx, qclhip_grid* grid, double* ms_host) {
    if (!ctx || !grid) return QCLHIP_ERR_ARGUMENT;
    if (ctx->capturing) return fail(ctx, QCLHIP_ERR_ARGUMENT, "autotuning is not possible during graph capture");
    for (const auto& t : ctx->tune_cache)
        if (t.Nb == grid->Nb && t.NE == grid->NE && t.Nk == grid->Nk) {
            grid->keldysh_engine = t.engine;
            grid->dyson_waves = t.dyson_waves;
            grid->keldysh_waves = t.keldysh_waves;
            grid->keldysh_split = t.keldysh_split;
            if (ms_host) {
                ms_host[0] = t.ms_vector;
                ms_host[1] = t.ms_matrix;
            }
            return QCLHIP_OK;
        }
    const int NE = std::min(grid->NE, 2048), Nk = grid->Nk, N = grid->Nb;
    const long long nblk = (long long)NE * Nk, n = nblk * N * N;
    c64 *h = nullptr, *f = nullptr;
    double* eps = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (h) hipFree(h);
        if (f) hipFree(f);
        if (eps) hipFree(eps);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
    };
    if (hipMalloc(&h, (size_t)Nk * N * N * sizeof(c64)) != hipSuccess ||
        hipMalloc(&f, (size_t)6 * n * sizeof(c64)) != hipSuccess || hipMalloc(&eps, NE * sizeof(double)) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return fail(ctx, QCLHIP_ERR_RUNTIME, "autotune allocation failed");
    }
    // eps: the first NE energies of the grid.
    if (hipMemcpyAsync(eps, grid->eps, NE * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
        cleanup();
        return fail(ctx, QCLHIP_ERR_RUNTIME, "autotune copy failed");
    }
    hipLaunchKernelGGL(k_tune_h, dim3(64), dim3(256), 0, ctx->stream, Nk, N, h);
    hipLaunchKernelGGL(k_tune_fields, dim3(grid_blocks(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream, nblk, N, f,
                       f + n, f + 2 * n);
    int rc = launch_check(ctx, "autotune fields");
    if (rc) {
        cleanup();
        return rc;
    }
    DKArgs a{};
    a.NE = NE;
    a.Nk = Nk;
    a.N = N;
    a.eps = eps;
    a.h = h;
    a.famR[0] = f;
    a.famL[0] = f + n;
    a.famG[0] = f + 2 * n;
    a.nfamR = a.nfamL = a.nfamG = 1;
    a.GR = f + 3 * n;
    a.GL = f + 4 * n;
    a.GG = f + 5 * n;
    a.status = ctx->status;
    a.keldysh = 1;
    rc = launch_dyson_fp64(ctx, a);
    float ms[2] = {0.f, 0.f};
    for (int engine = 0; engine < 2 && !rc; ++engine) {
        rc = engine ? launch_keldysh_mfma(ctx, a, false) : launch_keldysh_valu(ctx, a);  // warm-up
        if (rc) break;
        const int reps = 3;
        hipEventRecord(e0, ctx->stream);
        for (int r = 0; r < reps && !rc; ++r)
            rc = engine ? launch_keldysh_mfma(ctx, a, false) : launch_keldysh_valu(ctx, a);
        hipEventRecord(e1, ctx->stream);
        if (!rc && hipEventSynchronize(e1) == hipSuccess) hipEventElapsedTime(&ms[engine], e0, e1);
        ms[engine] /= reps;
    }
    // Tile shapes of the FP64 vector kernels: energies per workgroup of the
    // Dyson kernel (4 or 8) and of the Keldysh kernel (2 or 4), each timed
    // where it is compiled and fits the LDS of a CU (0 ms: not possible).
    auto timed = [&](int which, int W, float* out) {
        DKArgs b = a;
        b.wd = which == 0 ? W : 0;
        b.wk = which == 1 ? W : 0;
        auto run = [&]() { return which == 0 ? launch_dyson_fp64(ctx, b) : launch_keldysh_valu(ctx, b); };
        int r = run();
        const int reps = 3;
        hipEventRecord(e0, ctx->stream);
        for (int k = 0; k < reps && !r; ++k) r = run();
        hipEventRecord(e1, ctx->stream);
        if (!r && hipEventSynchronize(e1) == hipSuccess) hipEventElapsedTime(out, e0, e1);
        *out /= reps;
        return r;
    };
    float msd[2] = {0.f, 0.f}, msk[2] = {0.f, 0.f};
    const int dw[2] = {kDysonTileWaves, 8}, kw[2] = {2, 4};
    for (int k = 0; k < 2 && !rc; ++k)
        if (dyson_waves_possible(N, dw[k])) rc = timed(0, dw[k], &msd[k]);
    for (int k = 0; k < 2 && !rc; ++k)
        if (keldysh_waves_possible(N, kw[k])) rc = timed(1, kw[k], &msk[k]);
    // Waves per block of the vector Keldysh kernel (rows of a block split
    // over 1..6 waves where compiled), at the chosen energies per workgroup.
    int best_ks = 1, best_ks_w = 0;
    float ms_ks = 0.f;
    {
        const int kw_best = [&]() {
            if (msk[0] > 0.f && msk[1] > 0.f) return msk[1] < msk[0] ? kw[1] : kw[0];
            return msk[1] > 0.f ? kw[1] : (msk[0] > 0.f ? kw[0] : 0);
        }();
        for (int ks = 1; ks <= 6 && !rc; ++ks) {
            if (!keldysh_split_possible(N, ks)) continue;
            // split kernels also at twice the energies per workgroup
            const int widths[2] = {kw_best, ks > 1 ? 2 * (kw_best ? kw_best : 1) : 0};
            for (int iw = 0; iw < 2 && !rc; ++iw) {
                const int wk = widths[iw];
                if (iw == 1 && (wk == 0 || !keldysh_waves_possible(N, wk))) continue;
                DKArgs b = a;
                b.wk = wk;
                b.ks = ks;
                rc = launch_keldysh_valu(ctx, b);  // warm-up
                const int reps = 3;
                hipEventRecord(e0, ctx->stream);
                for (int k = 0; k < reps && !rc; ++k) rc = launch_keldysh_valu(ctx, b);
                hipEventRecord(e1, ctx->stream);
                float t = 0.f;
                if (!rc && hipEventSynchronize(e1) == hipSuccess) hipEventElapsedTime(&t, e0, e1);
                t /= reps;
                if (ms_ks == 0.f || t < ms_ks) {
                    ms_ks = t;
                    best_ks = ks;
                    best_ks_w = wk;
                }
            }
        }
    }
    // The fused Dyson + Keldysh kernel, against the two launches.
    float msf = 0.f;
    if (!rc) {
        rc = launch_fused_fp64(ctx, a);  // warm-up
        const int reps = 3;
        hipEventRecord(e0, ctx->stream);
        for (int k = 0; k < reps && !rc; ++k) rc = launch_fused_fp64(ctx, a);
        hipEventRecord(e1, ctx->stream);
        if (!rc && hipEventSynchronize(e1) == hipSuccess) hipEventElapsedTime(&msf, e0, e1);
        msf /= reps;
    }
    if (!rc) rc = qclhip_ctx_synchronize(ctx);
    cleanup();
    if (rc) return rc;
    auto pick = [](const float* t, const int* w) {
        if (t[0] > 0.f && t[1] > 0.f) return t[1] < t[0] ? w[1] : w[0];
        if (t[0] <= 0.f && t[1] <= 0.f) return 0;  // neither shape exists: the built-in rule
        return t[1] > 0.f ? w[1] : w[0];
    };
    auto best = [](const float* t, float fallback) {
        float b = fallback;
        for (int k = 0; k < 2; ++k)
            if (t[k] > 0.f && t[k] < b) b = t[k];
        return b;
    };
    grid->dyson_waves = pick(msd, dw);
    grid->keldysh_waves = pick(msk, kw);
    // Whole launch sets: the best Dyson tile with either Keldysh engine, or
    // the fused kernel.
    const float dy = best(msd, 3.0e38f);
    const float t_vec = dy + std::min(best(msk, ms[0]), ms_ks > 0.f ? ms_ks : 3.0e38f), t_mat = dy + ms[1];
    int engine = t_mat < t_vec ? QCLHIP_ENGINE_MATRIX : QCLHIP_ENGINE_VECTOR;
    if (msf > 0.f && msf < std::min(t_vec, t_mat)) engine = QCLHIP_ENGINE_FUSED;
    grid->keldysh_engine = engine;
    grid->keldysh_split = best_ks;
    if (best_ks > 1) grid->keldysh_waves = best_ks_w;  // the split kernel's energies per workgroup
    qclhip_ctx::TuneEntry te{grid->Nb,
                             grid->NE,
                             grid->Nk,
                             engine,
                             grid->dyson_waves,
                             grid->keldysh_waves,
                             grid->keldysh_split,
                             ms[0],
                             ms[1],
                             {msd[0], msd[1]},
                             {msk[0], msk[1]},
                             msf,
                             ms_ks};
    ctx->tune_cache.push_back(te);
    if (ms_host) {
        ms_host[0] = ms[0];
        ms_host[1] = ms[1];
    }
    return QCLHIP_OK;
}

int qclhip_grid_set_tiles(qclhip_ctx* ctx, qclhip_grid* grid, int dyson_waves, int keldysh_waves) {
    if (!ctx || !grid) return QCLHIP_ERR_ARGUMENT;
    if ((dyson_waves && !dyson_waves_possible(grid->Nb, dyson_waves)) ||
        (keldysh_waves && !keldysh_waves_possible(grid->Nb, keldysh_waves)))
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "tile shape not available for this basis size");
    grid->dyson_waves = dyson_waves;
    grid->keldysh_waves = keldysh_waves;
    return QCLHIP_OK;
}

int qclhip_grid_set_keldysh_split(qclhip_ctx* ctx, qclhip_grid* grid, int waves_per_block) {
    if (!ctx || !grid) return QCLHIP_ERR_ARGUMENT;
    if (waves_per_block && !keldysh_split_possible(grid->Nb, waves_per_block))
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "waves per block not available for this basis size");
    grid->keldysh_split = waves_per_block;
    return QCLHIP_OK;
}

int qclhip_grid_tuning(qclhip_ctx* ctx, const qclhip_grid* grid, double* ms_host, int* choice_host) {
    if (!ctx || !grid || !ms_host || !choice_host) return QCLHIP_ERR_ARGUMENT;
    for (const auto& t : ctx->tune_cache)
        if (t.Nb == grid->Nb && t.NE == grid->NE && t.Nk == grid->Nk) {
            const double v[8] = {t.ms_vector, t.ms_matrix, t.ms_dyson[0], t.ms_dyson[1], t.ms_keldysh[0],
                                 t.ms_keldysh[1], t.ms_fused, t.ms_split};
            for (int k = 0; k < 8; ++k) ms_host[k] = v[k];
            choice_host[0] = t.engine;
            choice_host[1] = t.dyson_waves;
            choice_host[2] = t.keldysh_waves;
            choice_host[3] = t.keldysh_split;
            return QCLHIP_OK;
        }
    return fail(ctx, QCLHIP_ERR_ARGUMENT, "this grid shape has not been autotuned");
}

}  // extern "C"
