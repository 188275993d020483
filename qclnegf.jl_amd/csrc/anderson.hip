// Anderson mixing of the self-energy families on the device
// (production_mixing.jl:135-317): the history append with its Gram row,
// the linear combination of the history, and the causality guard of the
// mixed families. The coefficient solve of the small Gram matrix and the
// history bookkeeping stay on the host.
//
// A history entry holds every field of the ordered families (mechanisms,
// plus, minus; R, <, > each), so the passes below stream nfields fields
// of n complex values; the guard runs one wavefront per (family, e, m)
// block.
#include <vector>

#include "wave_eig.hpp"

namespace qclhip {

constexpr int kAndBlocks = 512;
constexpr int kAndMaxHistory = 16;

struct AppendArgs {
    const c64* old_resid[kAndMaxHistory];
    int nold;
};

// saved <- state, resid <- cand - state, and the partial sums of
// Re <old_h, resid> (h < nold) and Re <resid, resid> on a fixed grid.
__global__ __launch_bounds__(256) void k_anderson_append(long long n, const c64* __restrict__ state,
                                                         const c64* __restrict__ cand, c64* __restrict__ saved,
                                                         c64* __restrict__ resid, AppendArgs a, double* part) {
    __shared__ double red[256];
    double acc[kAndMaxHistory + 1];
#pragma unroll
    for (int h = 0; h <= kAndMaxHistory; ++h) acc[h] = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)kAndBlocks * 256) {
        const c64 s = state[i], c = cand[i];
        const c64 r = cmk(c.x - s.x, c.y - s.y);
        saved[i] = s;
        resid[i] = r;
#pragma unroll
        for (int h = 0; h < kAndMaxHistory; ++h)
            if (h < a.nold) {
                const c64 o = a.old_resid[h][i];
                acc[h] += o.x * r.x + o.y * r.y;
            }
        acc[kAndMaxHistory] += r.x * r.x + r.y * r.y;
    }
    // one fixed tree per sum; row h of part holds the block partials
#pragma unroll
    for (int h = 0; h <= kAndMaxHistory; ++h) {
        if (h < a.nold || h == kAndMaxHistory) {
            red[threadIdx.x] = acc[h];
            __syncthreads();
            for (int w = 128; w > 0; w >>= 1) {
                if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
                __syncthreads();
            }
            if (threadIdx.x == 0) part[(h == kAndMaxHistory ? a.nold : h) * kAndBlocks + blockIdx.x] = red[0];
            __syncthreads();
        }
    }
}

// dots[k] = fixed-tree sum of part[k][0..kAndBlocks).
__global__ __launch_bounds__(kAndBlocks) void k_anderson_dots(const double* part, double* dots) {
    __shared__ double red[kAndBlocks];
    red[threadIdx.x] = part[blockIdx.x * kAndBlocks + threadIdx.x];
    __syncthreads();
    for (int w = kAndBlocks / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) dots[blockIdx.x] = red[0];
}

struct CombineArgs {
    const c64* X[kAndMaxHistory];
    const c64* R[kAndMaxHistory];
    double c[kAndMaxHistory];
    int count;
};

// out = sum_h c_h (X_h + damping R_h), history order.
__global__ __launch_bounds__(256) void k_anderson_combine(long long n, CombineArgs a, double damping,
                                                          c64* __restrict__ out) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        c64 v = cmk(0.0, 0.0);
#pragma unroll
        for (int h = 0; h < kAndMaxHistory; ++h)
            if (h < a.count) {
                const c64 x = a.X[h][i], r = a.R[h][i];
                v = cmk(v.x + a.c[h] * (x.x + damping * r.x), v.y + a.c[h] * (x.y + damping * r.y));
            }
        out[i] = v;
    }
}

// ---- guard (production_mixing.jl:36-80) ----

struct GuardArgs {
    int NE, Nk, N, nfam;
    const c64* R[kMaxFamilies];
    const c64* L[kMaxFamilies];
    const c64* G[kMaxFamilies];
    unsigned long long* first;  // min of (index + 1) * 8 + code over failing blocks
};

// sigma_max of the N x N block in tile X: sqrt(lambda_max(X X^dagger)).
template <int B, int LD>
__device__ inline double sigma_max(const c64* X, int N, int lane, EigWork<B>* w, c64 (&T)[B][B]) {
    rprod<B>(X, X, N, lane, T, true);
    const BlockScale s = block_scale<B>(T, N, lane);
    if (s.bad) return INFINITY;
    if (s.amax == 0.0) return 0.0;
    int ex;
    frexp(s.amax, &ex);
    scale_regs<B>(T, ldexp(1.0, -ex));
    double lmin, lmax;
    eig_core<B>(T, N, lane, w, lmin, lmax);
    return sqrt(fmax(0.0, ldexp(lmax, ex)));
}

// Extreme eigenvalue of the Hermitian part of f X (lowest when want_min).
template <int B, int LD>
__device__ inline double herm_extreme(const c64* X, int N, int lane, EigWork<B>* w, double fr, double fi,
                                      bool want_min) {
    c64 a[B][B];
    load_herm<B, LD>(X, N, lane, fr, fi, a);
    const BlockScale s = block_scale<B>(a, N, lane);
    if (s.bad) return NAN;
    if (s.amax == 0.0) return 0.0;
    int ex;
    frexp(s.amax, &ex);
    scale_regs<B>(a, ldexp(1.0, -ex));
    double lmin, lmax;
    eig_core<B>(a, N, lane, w, lmin, lmax);
    return ldexp(want_min ? lmin : lmax, ex);
}

// True when the Hermitian part H of f X certainly satisfies
// lambda_min(H) >= -delta (Cholesky of H + delta I; its backward error is
// rounding-level next to delta).
template <int B, int LD>
__device__ inline bool herm_screen(const c64* X, int N, int lane, EigWork<B>* w, double fr, double fi, double delta) {
    c64 a[B][B];
    load_herm<B, LD>(X, N, lane, fr, fi, a);
    const BlockScale s = block_scale<B>(a, N, lane);
    if (s.bad) return false;
    if (s.amax == 0.0) return delta >= 0.0;
    int ex;
    frexp(s.amax, &ex);
    const double sc = ldexp(1.0, -ex);
    scale_regs<B>(a, sc);
    return chol_screen<B>(a, N, lane, w, delta * sc);
}

template <int LD>
__device__ inline bool tile_all_finite(const c64* T, int N, int lane) {
    double bad = 0.0;
    for (int f = lane; f < N * N; f += 64) {
        const c64 v = T[(f / N) * LD + f % N];
        if (!(isfinite(v.x) && isfinite(v.y))) bad = 1.0;
    }
    return wave_max(bad) == 0.0;
}

// The failure code of one block (0 = admissible), in the reference's order:
// 1 non-finite, 2 Im S^R above 64 eps ||S^R||_2, 3 anti-Hermitian part of
// f S (f = -i for <, i for >) above the PSD budget, 4 lambda_min of its
// Hermitian part below -budget, 5 S^R - S^R^dagger != S> - S< beyond the
// budget. Exact norms and eigenvalues are computed only where Frobenius
// bounds and a Cholesky screen cannot settle the comparison.
template <int B>
__device__ inline int guard_block(const c64* SR, const c64* SL, const c64* SG, int N, int lane, EigWork<B>* w,
                                  c64 (&T)[B][B]) {
    constexpr int LD = 8 * B + 1;
    constexpr double kEps = 2.220446049250313e-16, kTiny = 2.2250738585072014e-308;
    const double sqN = sqrt((double)N);
    if (!tile_all_finite<LD>(SR, N, lane)) return 1;
    const double fR = sqrt(tile_frob2<LD>(SR, N, lane, 0.0));
    double sR = -1.0;  // exact ||S^R||_2 once computed
    // code 2: lambda_max(Im S^R) <= 64 eps max(||S^R||_2, floatmin)
    if (!herm_screen<B, LD>(SR, N, lane, w, 0.0, 1.0, 32.0 * kEps * fR / sqN)) {
        sR = sigma_max<B, LD>(SR, N, lane, w, T);
        const double lmax = herm_extreme<B, LD>(SR, N, lane, w, 0.0, -1.0, false);
        if (!(lmax <= 64.0 * kEps * fmax(sR, kTiny))) return 2;
    }
    if (!tile_all_finite<LD>(SL, N, lane) || !tile_all_finite<LD>(SG, N, lane)) return 1;
    double sig[2] = {-1.0, -1.0};
    double fro[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const c64* X = k == 0 ? SL : SG;
        const double fi = k == 0 ? -1.0 : 1.0;
        fro[k] = sqrt(tile_frob2<LD>(X, N, lane, 0.0));
        // ||f X - (f X)^dagger||_F / 2 = ||X + X^dagger||_F / 2 for f = -+i
        double an = 0.0;
        for (int f = lane; f < N * N; f += 64) {
            const int i = f / N, j = f % N;
            const c64 a = X[i * LD + j], b = X[j * LD + i];
            const double x = a.x + b.x, y = a.y - b.y;
            an += x * x + y * y;
        }
        an = 0.5 * sqrt(wave_sum(an));
        const double b_lo = 64.0 * N * kEps * fmax(fro[k] / sqN, 1.0);
        const double b_hi = 64.0 * N * kEps * fmax(fro[k], 1.0);
        if (an > b_hi) return 3;
        if (an > b_lo) {
            sig[k] = sigma_max<B, LD>(X, N, lane, w, T);
            if (!(an <= 64.0 * N * kEps * fmax(sig[k], 1.0))) return 3;
        }
        if (!herm_screen<B, LD>(X, N, lane, w, 0.0, fi, 0.5 * b_lo)) {
            if (sig[k] < 0.0) sig[k] = sigma_max<B, LD>(X, N, lane, w, T);
            const double lmin = herm_extreme<B, LD>(X, N, lane, w, 0.0, fi, true);
            if (!(lmin >= -64.0 * N * kEps * fmax(sig[k], 1.0))) return 4;
        }
    }
    // code 5: ||S^R - S^R^dagger - (S> - S<)||_F <= budget(max norms)
    double d5 = 0.0;
    for (int f = lane; f < N * N; f += 64) {
        const int i = f / N, j = f % N;
        const c64 r = SR[i * LD + j], rt = SR[j * LD + i], g = SG[i * LD + j], l = SL[i * LD + j];
        const double x = (r.x - rt.x) - (g.x - l.x), y = (r.y + rt.y) - (g.y - l.y);
        d5 += x * x + y * y;
    }
    d5 = sqrt(wave_sum(d5));
    const double lo = 64.0 * N * kEps * fmax(fmax(fR, fmax(fro[0], fro[1])) / sqN, 1.0);
    const double hi = 64.0 * N * kEps * fmax(fmax(fR, fmax(fro[0], fro[1])), 1.0);
    if (d5 > hi) return 5;
    if (d5 > lo) {
        if (sR < 0.0) sR = sigma_max<B, LD>(SR, N, lane, w, T);
        if (sig[0] < 0.0) sig[0] = sigma_max<B, LD>(SL, N, lane, w, T);
        if (sig[1] < 0.0) sig[1] = sigma_max<B, LD>(SG, N, lane, w, T);
        if (!(d5 <= 64.0 * N * kEps * fmax(fmax(sR, fmax(sig[0], sig[1])), 1.0))) return 5;
    }
    return 0;
}

template <int B>
__global__ __launch_bounds__(64) void k_anderson_guard(GuardArgs p) {
    constexpr int LD = 8 * B + 1;
    const int wstride = TileStride<B>::value;
    extern __shared__ __align__(16) unsigned char smem[];
    c64* SR = reinterpret_cast<c64*>(smem);
    c64* SL = SR + wstride;
    c64* SG = SL + wstride;
    EigWork<B>* w = reinterpret_cast<EigWork<B>*>(SG + wstride);
    const int lane = threadIdx.x;
    const long long nblk = (long long)p.NE * p.Nk;
    const long long b = xcd_tile(blockIdx.x, nblk * p.nfam);
    const int f = (int)(b / nblk);
    const long long blk = b % nblk;  // e + NE m
    const int e = (int)(blk % p.NE), m = (int)(blk / p.NE);
    const int N = p.N;
    const size_t S = (size_t)nblk;
    const c64* one[1];
#pragma unroll
    for (int k = 0; k < kMaxFamilies; ++k)
        if (k == f) {
            one[0] = p.R[k];
            tile_load<1, LD>(SR, wstride, one, 1, blk, 1, N, S);
            one[0] = p.L[k];
            tile_load<1, LD>(SL, wstride, one, 1, blk, 1, N, S);
            one[0] = p.G[k];
            tile_load<1, LD>(SG, wstride, one, 1, blk, 1, N, S);
        }
    wave_sync();
    c64 T[B][B];
    const int code = guard_block<B>(SR, SL, SG, N, lane, w, T);
    if (code && lane == 0) {
        // the reference's linear block order: family, then e, then m fastest
        const unsigned long long index = (unsigned long long)f * nblk + (unsigned long long)e * p.Nk + m;
        atomicMin(p.first, (index + 1) * 8ull + (unsigned long long)code);
    }
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) {
    if (threadIdx.x == 0) *p = v;
}

template <int B>
int launch_guard(qclhip_ctx* ctx, const GuardArgs& a) {
    const size_t lds = 3 * (size_t)TileStride<B>::value * sizeof(c64) + sizeof(EigWork<B>);
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_anderson_guard<B>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long nb = (long long)a.NE * a.Nk * a.nfam;
    hipLaunchKernelGGL((k_anderson_guard<B>), dim3((unsigned)nb), dim3(64), lds, ctx->stream, a);
    return launch_check(ctx, "k_anderson_guard");
}

}  // namespace qclhip

using namespace qclhip;

extern "C" {

int qclhip_anderson_append(qclhip_ctx* ctx, int64_t n, int nfields, const double* const* state,
                           const double* const* cand, double* const* saved, double* const* resid, int nold,
                           const double* const* old_resid, double* dots_host) {
    if (!ctx || !state || !cand || !saved || !resid || !dots_host || n < 0 || nfields < 1)
        return QCLHIP_ERR_ARGUMENT;
    if (nold < 0 || nold >= kAndMaxHistory || (nold > 0 && !old_resid))
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "Anderson history holds at most 15 older residuals");
    const int per = nold + 1;
    double* part = (double*)workspace(ctx, ((size_t)per * kAndBlocks + (size_t)nfields * per) * sizeof(double));
    if (!part) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    double* dots = part + (size_t)per * kAndBlocks;
    for (int f = 0; f < nfields; ++f) {
        if (!state[f] || !cand[f] || !saved[f] || !resid[f]) return QCLHIP_ERR_ARGUMENT;
        AppendArgs a{};
        a.nold = nold;
        for (int h = 0; h < nold; ++h) {
            a.old_resid[h] = (const c64*)old_resid[(size_t)h * nfields + f];
            if (!a.old_resid[h]) return QCLHIP_ERR_ARGUMENT;
        }
        hipLaunchKernelGGL(k_anderson_append, dim3(kAndBlocks), dim3(256), 0, ctx->stream, (long long)n,
                           (const c64*)state[f], (const c64*)cand[f], (c64*)saved[f], (c64*)resid[f], a, part);
        hipLaunchKernelGGL(k_anderson_dots, dim3(per), dim3(kAndBlocks), 0, ctx->stream, part,
                           dots + (size_t)f * per);
    }
    int rc = launch_check(ctx, "k_anderson_append");
    if (rc) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(dots_host, dots, (size_t)nfields * per * sizeof(double),
                                       hipMemcpyDeviceToHost, ctx->stream));
    return qclhip_ctx_synchronize(ctx);
}

int qclhip_anderson_combine(qclhip_ctx* ctx, int64_t n, int nfields, int count, const double* coef_host,
                            double damping, const double* const* X, const double* const* R, double* const* out) {
    if (!ctx || !coef_host || !X || !R || !out || n < 0 || nfields < 1) return QCLHIP_ERR_ARGUMENT;
    if (count < 1 || count > kAndMaxHistory) return fail(ctx, QCLHIP_ERR_ARGUMENT, "1..16 history entries");
    for (int f = 0; f < nfields; ++f) {
        CombineArgs a{};
        a.count = count;
        for (int h = 0; h < count; ++h) {
            a.X[h] = (const c64*)X[(size_t)h * nfields + f];
            a.R[h] = (const c64*)R[(size_t)h * nfields + f];
            a.c[h] = coef_host[h];
            if (!a.X[h] || !a.R[h]) return QCLHIP_ERR_ARGUMENT;
        }
        if (!out[f]) return QCLHIP_ERR_ARGUMENT;
        long long blocks = std::min<long long>(((long long)n + 255) / 256, (long long)ctx->sm_count * 32);
        hipLaunchKernelGGL(k_anderson_combine, dim3((unsigned)std::max<long long>(1, blocks)), dim3(256), 0,
                           ctx->stream, (long long)n, a, damping, (c64*)out[f]);
    }
    return launch_check(ctx, "k_anderson_combine");
}

int qclhip_anderson_guard(qclhip_ctx* ctx, const qclhip_grid* grid, int nfam, const double* const* famR,
                          const double* const* famL, const double* const* famG, int64_t* first_failure,
                          int* code) {
    if (!ctx || !grid || !famR || !famL || !famG || !first_failure || !code) return QCLHIP_ERR_ARGUMENT;
    if (nfam < 1 || nfam > kMaxFamilies) return fail(ctx, QCLHIP_ERR_ARGUMENT, "1..8 families");
    GuardArgs a{};
    a.NE = grid->NE;
    a.Nk = grid->Nk;
    a.N = grid->Nb;
    a.nfam = nfam;
    for (int f = 0; f < nfam; ++f) {
        if (!famR[f] || !famL[f] || !famG[f]) return QCLHIP_ERR_ARGUMENT;
        a.R[f] = (const c64*)famR[f];
        a.L[f] = (const c64*)famL[f];
        a.G[f] = (const c64*)famG[f];
    }
    a.first = (unsigned long long*)workspace(ctx, 64);
    if (!a.first) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, ctx->stream, a.first, ~0ull);
    int rc;
    const int N = grid->Nb;
    if (N <= 8)
        rc = launch_guard<1>(ctx, a);
    else if (N <= 16)
        rc = launch_guard<2>(ctx, a);
    else if (N <= 24)
        rc = launch_guard<3>(ctx, a);
    else if (N <= 32)
        rc = launch_guard<4>(ctx, a);
    else if (N <= 40)
        rc = launch_guard<5>(ctx, a);
    else if (N <= 48)
        rc = launch_guard<6>(ctx, a);
    else
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");
    if (rc) return rc;
    unsigned long long key = 0;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(&key, a.first, sizeof(key), hipMemcpyDeviceToHost, ctx->stream));
    rc = qclhip_ctx_synchronize(ctx);
    if (rc) return rc;
    if (key == ~0ull) {
        *first_failure = 0;
        *code = 0;
    } else {
        *first_failure = (int64_t)(key / 8);  // one-based, as the reference reports it
        *code = (int)(key % 8);
    }
    return QCLHIP_OK;
}

}  // extern "C"
