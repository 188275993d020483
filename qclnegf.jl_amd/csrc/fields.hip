// Energy shifts, Gamma, and the observables of one SCBA iteration.
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace qclhip {

// apply_energy_shift! (production_cache.jl:300-325): column-wise two-point
// interpolation along energy. Thread = energy (its interpolation entry read
// once), block rows stride over the columns; consecutive threads touch
// consecutive energies.
__global__ void k_apply_shift(int NE, long long ncols, const int* __restrict__ left, const int* __restrict__ right,
                              const double* __restrict__ wr, const unsigned char* __restrict__ valid,
                              const c64* __restrict__ X, c64* __restrict__ Y) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NE) return;
    const bool ok = valid[e];
    const int l = left[e], r = right[e];
    const double t = wr[e];
    for (long long col = blockIdx.y; col < ncols; col += gridDim.y) {
        const long long base = col * NE;
        c64 y = cmk(0.0, 0.0);
        if (ok) {
            const c64 xl = X[base + l];
            if (l == r || t == 0.0) {
                y = xl;
            } else {
                const c64 xr = X[base + r];
                y = cmk((1.0 - t) * xl.x + t * xr.x, (1.0 - t) * xl.y + t * xr.y);
            }
        }
        Y[base + e] = y;
    }
}

// Thermal LO combination with the shifts applied after the contraction:
// out = ca * (S_a Y) + cb * (S_b Y), exact since the shift acts on energy
// and the kernel on (m,a,b) (production_operators.jl:1107-1162). Layout of
// k_apply_shift.
__global__ void k_lo_combine(int NE, long long ncols, const c64* __restrict__ Y, double ca, const int* la,
                             const int* ra, const double* wa, const unsigned char* va, double cb, const int* lb,
                             const int* rb, const double* wb, const unsigned char* vb, c64* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NE) return;
    const bool oka = va[e], okb = vb[e];
    const int l1 = la[e], r1 = ra[e], l2 = lb[e], r2 = rb[e];
    const double t1 = wa[e], t2 = wb[e];
    for (long long col = blockIdx.y; col < ncols; col += gridDim.y) {
        const long long base = col * NE;
        c64 sa = cmk(0.0, 0.0), sb = cmk(0.0, 0.0);
        if (oka) {
            const c64 yl = Y[base + l1];
            if (l1 == r1 || t1 == 0.0) sa = yl;
            else {
                const c64 yr = Y[base + r1];
                sa = cmk((1.0 - t1) * yl.x + t1 * yr.x, (1.0 - t1) * yl.y + t1 * yr.y);
            }
        }
        if (okb) {
            const c64 yl = Y[base + l2];
            if (l2 == r2 || t2 == 0.0) sb = yl;
            else {
                const c64 yr = Y[base + r2];
                sb = cmk((1.0 - t2) * yl.x + t2 * yr.x, (1.0 - t2) * yl.y + t2 * yr.y);
            }
        }
        out[base + e] = cmk(ca * sa.x + cb * sb.x, ca * sa.y + cb * sb.y);
    }
}

__global__ void k_gamma(long long n, const c64* __restrict__ SL, const c64* __restrict__ SG, c64* __restrict__ Gm) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        c64 l = SL[i], g = SG[i];
        Gm[i] = cmk(-(g.y - l.y), g.x - l.x);
    }
}

// Per block: wE wk tr(S>+ G< - S<+ G>) (production_iteration.jl:174-187).
__global__ void k_flux_terms(int NE, int Nk, int N, const double* wE, const double* wk, const c64* __restrict__ SL,
                             const c64* __restrict__ SG, const c64* __restrict__ GL, const c64* __restrict__ GG,
                             c64* __restrict__ terms) {
    const long long nblk = (long long)NE * Nk;
    const long long blk = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= nblk) return;
    const int e = (int)(blk % NE), m = (int)(blk / NE);
    c64 value = cmk(0.0, 0.0);
    for (int a = 0; a < N; ++a) {
        c64 outgoing = cmk(0.0, 0.0), incoming = cmk(0.0, 0.0);
        for (int b = 0; b < N; ++b) {
            const size_t ab = blk + nblk * (a + (size_t)N * b), ba = blk + nblk * (b + (size_t)N * a);
            outgoing = cfma(outgoing, SG[ab], GL[ba]);
            incoming = cfma(incoming, SL[ab], GG[ba]);
        }
        value = cadd(value, csub(outgoing, incoming));
    }
    const double w = wE[e] * wk[m];
    terms[blk] = cmk(w * value.x, w * value.y);
}

// Sum of the flux terms in the order (e sorted by (|e|,e), m inner); fixed
// contiguous partition plus a fixed tree, then Re[g_s value / (2 pi)].
__global__ __launch_bounds__(1024) void k_flux_reduce(int NE, int Nk, const int* order, const c64* terms, int gs,
                                                      double* out) {
    __shared__ double red[2][1024];
    const int t = threadIdx.x;
    const long long n = (long long)NE * Nk;
    const long long per = (n + 1023) / 1024;
    double sx = 0.0, sy = 0.0;
    for (long long s = t * per; s < min(n, (t + 1) * per); ++s) {
        const int e = order[s / Nk], m = (int)(s % Nk);
        c64 v = terms[e + (long long)NE * m];
        sx += v.x;
        sy += v.y;
    }
    red[0][t] = sx;
    red[1][t] = sy;
    __syncthreads();
    for (int width = 512; width > 0; width >>= 1) {
        if (t < width) {
            red[0][t] += red[0][t + width];
            red[1][t] += red[1][t + width];
        }
        __syncthreads();
    }
    if (t == 0) out[0] = ((double)gs * red[0][0]) / (2.0 * M_PI);
}

// p_a = Re[-i g_s/(2 pi) sum wE wk G<_aa] / target, one workgroup per a
// (production_iteration.jl:204-215).
__global__ __launch_bounds__(1024) void k_populations(int NE, int Nk, int N, const double* wE, const double* wk,
                                                      int gs, double target, const c64* __restrict__ GL,
                                                      double* out) {
    __shared__ double red[1024];
    const int a = blockIdx.x, t = threadIdx.x;
    const long long nblk = (long long)NE * Nk;
    const long long per = (nblk + 1023) / 1024;
    double s = 0.0;
    // Re[(-i k) w z] = k w Im z with k = g_s/(2 pi); order (e outer, m inner).
    for (long long lin = t * per; lin < min(nblk, (t + 1) * per); ++lin) {
        const int e = (int)(lin / Nk), m = (int)(lin % Nk);
        c64 z = GL[e + (long long)NE * m + nblk * (a + (long long)N * a)];
        s += (wE[e] * wk[m]) * z.y;
    }
    red[t] = s;
    __syncthreads();
    for (int width = 512; width > 0; width >>= 1) {
        if (t < width) red[t] += red[t + width];
        __syncthreads();
    }
    if (t == 0) out[a] = ((double)gs / (2.0 * M_PI)) * red[0] / target;
}

int gamma_field(qclhip_ctx* ctx, long long n, const double* SL, const double* SG, double* Gamma) {
    hipLaunchKernelGGL(k_gamma, dim3(grid_blocks(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream, n,
                       (const c64*)SL, (const c64*)SG, (c64*)Gamma);
    return launch_check(ctx, "k_gamma");
}

int lo_combine(qclhip_ctx* ctx, int NE, long long n, const c64* Y, double ca, const qclhip_shift* a, double cb,
               const qclhip_shift* b, c64* out) {
    if (NE < 1 || n % NE) return fail(ctx, QCLHIP_ERR_DIMENSION, "LO combination: field size not a multiple of NE");
    const long long ncols = n / NE;
    if (ncols == 0) return QCLHIP_OK;
    hipLaunchKernelGGL(k_lo_combine, ecol_grid(NE, ncols, ctx->sm_count), dim3(256), 0, ctx->stream, NE, ncols, Y,
                       ca, a->left, a->right, a->wr, a->valid, cb, b->left, b->right, b->wr, b->valid, out);
    return launch_check(ctx, "k_lo_combine");
}

// X_em <- (X_em - X_em^dagger) / 2 in place; one thread per pair a <= b.
__global__ void k_antihermitian_part(long long nblk, int N, c64* X) {
    const long long npair = (long long)N * (N + 1) / 2;
    const long long n = nblk * npair;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long blk = idx % nblk;
        long long pr = idx / nblk;
        int b = 0;
        while (pr >= b + 1) {
            pr -= b + 1;
            ++b;
        }
        const int a = (int)pr;  // a <= b
        const size_t ab = blk + nblk * (a + (size_t)N * b), ba = blk + nblk * (b + (size_t)N * a);
        const c64 x = X[ab], y = X[ba];
        X[ab] = cmk(0.5 * (x.x - y.x), 0.5 * (x.y + y.y));
        if (a != b) X[ba] = cmk(0.5 * (y.x - x.x), 0.5 * (y.y + x.y));
    }
}

}  // namespace qclhip

using namespace qclhip;

extern "C" {

int qclhip_antihermitian_part(qclhip_ctx* ctx, const qclhip_grid* grid, double* X) {
    if (!ctx || !grid || !X) return QCLHIP_ERR_ARGUMENT;
    const long long nblk = (long long)grid->NE * grid->Nk;
    const long long n = nblk * grid->Nb * (grid->Nb + 1) / 2;
    hipLaunchKernelGGL(k_antihermitian_part, dim3(grid_blocks(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream,
                       nblk, grid->Nb, (c64*)X);
    return launch_check(ctx, "k_antihermitian_part");
}

int qclhip_apply_shift(qclhip_ctx* ctx, const qclhip_shift* shift, int64_t ncols, const double* X, double* Y) {
    if (!ctx || !shift || !X || !Y || ncols < 0) return QCLHIP_ERR_ARGUMENT;
    if (X == Y) return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    if (ncols == 0 || shift->NE == 0) return QCLHIP_OK;
    hipLaunchKernelGGL(k_apply_shift, ecol_grid(shift->NE, ncols, ctx->sm_count), dim3(256), 0, ctx->stream,
                       shift->NE, (long long)ncols, shift->left, shift->right, shift->wr, shift->valid,
                       (const c64*)X, (c64*)Y);
    return launch_check(ctx, "k_apply_shift");
}

int qclhip_boundary_flux(qclhip_ctx* ctx, const qclhip_grid* grid, const double* plusL, const double* plusG,
                         const double* GL, const double* GG, double* Jbar_host) {
    if (!ctx || !grid || !plusL || !plusG || !GL || !GG || !Jbar_host) return QCLHIP_ERR_ARGUMENT;
    const long long nblk = (long long)grid->NE * grid->Nk;
    char* ws = (char*)workspace(ctx, (size_t)nblk * sizeof(c64) + 64);
    if (!ws) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    c64* terms = (c64*)ws;
    double* res = (double*)(ws + (size_t)nblk * sizeof(c64));
    hipLaunchKernelGGL(k_flux_terms, dim3((unsigned)((nblk + 127) / 128)), dim3(128), 0, ctx->stream, grid->NE,
                       grid->Nk, grid->Nb, grid->wE, grid->wk, (const c64*)plusL, (const c64*)plusG, (const c64*)GL,
                       (const c64*)GG, terms);
    hipLaunchKernelGGL(k_flux_reduce, dim3(1), dim3(1024), 0, ctx->stream, grid->NE, grid->Nk, grid->flux_order,
                       terms, grid->gs, res);
    int rc = launch_check(ctx, "boundary_flux");
    if (rc) return rc;
    bool deferred = false;
    rc = defer_result(ctx, res, 1, &deferred);
    if (rc || deferred) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    rc = qclhip_ctx_synchronize(ctx);
    if (rc) return rc;
    *Jbar_host = ctx->host_scalars[0];
    return QCLHIP_OK;
}

int qclhip_state_populations(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GL, double target,
                             double* out_host) {
    if (!ctx || !grid || !GL || !out_host) return QCLHIP_ERR_ARGUMENT;
    if (!(target > 0)) return fail(ctx, QCLHIP_ERR_ARGUMENT, "target charge must be positive");
    double* res = (double*)workspace(ctx, (size_t)grid->Nb * sizeof(double));
    if (!res) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    hipLaunchKernelGGL(k_populations, dim3(grid->Nb), dim3(1024), 0, ctx->stream, grid->NE, grid->Nk, grid->Nb,
                       grid->wE, grid->wk, grid->gs, target, (const c64*)GL, res);
    int rc = launch_check(ctx, "k_populations");
    if (rc) return rc;
    bool deferred = false;
    rc = defer_result(ctx, res, grid->Nb, &deferred);
    if (rc || deferred) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(out_host, res, (size_t)grid->Nb * sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->stream));
    return qclhip_ctx_synchronize(ctx);
}

}  // extern "C"
