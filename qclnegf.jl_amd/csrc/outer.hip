// The Poisson step that closes the SCBA loop on the device
// (production_solver.jl:1044-1047): electron density from G<
// (_electron_density_bar, src/physics/observables.jl:1-26), the bordered
// periodic Poisson solve (solve_periodic_poisson,
// src/numerics/reference/poisson.jl:12-65), the outer mixing of the Hartree
// energy, and the projected Hamiltonians of the next SCBA run
// (project_hamiltonians, src/numerics/reference/basis.jl:288-308).
//
// Sizes are small (N_z ~ 512, N_b^2 coherences) next to the SCBA fields; the
// only field-sized pass is the energy/momentum integral of G<, one block per
// coherence reading a contiguous run of N_E N_k values.
#include <cmath>

#include "common.hpp"

namespace qclhip {

constexpr int kDensityThreads = 256;

// I_ab = sum_{m,e} (wE_e wk_m) G<_{e,m,a,b}: block ab, fixed strided partials
// and a fixed tree.
__global__ __launch_bounds__(kDensityThreads) void k_density_integrate(int NE, long long nem, const c64* GL,
                                                                    const double* wE, const double* wk, c64* I) {
    __shared__ double re[kDensityThreads], im[kDensityThreads];
    const long long ab = blockIdx.x;
    const c64* col = GL + ab * nem;
    double sr = 0.0, si = 0.0;
    for (long long l = threadIdx.x; l < nem; l += kDensityThreads) {
        const int e = (int)(l % NE), m = (int)(l / NE);
        const double w = wE[e] * wk[m];
        const c64 g = col[l];
        sr += w * g.x;
        si += w * g.y;
    }
    re[threadIdx.x] = sr;
    im[threadIdx.x] = si;
    __syncthreads();
    for (int s = kDensityThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            re[threadIdx.x] += re[threadIdx.x + s];
            im[threadIdx.x] += im[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) I[ab] = cmk(re[0], im[0]);
}

// nbar_g = gs/(2 pi) Re(-i sum_{b,a} chi_{g,a} I_ab conj(chi_{g,b})).
__global__ void k_density_project(int Nz, int Nb, const c64* chi, const c64* I, double pref, double* nbar) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= Nz) return;
    c64 v = cmk(0.0, 0.0);
    for (int b = 0; b < Nb; ++b) {
        const c64 cb = cconj(chi[g + (size_t)Nz * b]);
        for (int a = 0; a < Nb; ++a) v = cadd(v, cmul(cmul(chi[g + (size_t)Nz * a], I[a + (size_t)Nb * b]), cb));
    }
    nbar[g] = pref * v.y;  // Re(-i v) = Im v
}

// One workgroup: assemble the bordered system in M ((N+1)^2, column-major,
// global scratch), LU with partial pivoting (largest magnitude, first index
// on ties), the two triangular solves, and the residuals.
constexpr int kPoissonThreads = 1024;
constexpr int kMaxPoisson = 4096;

__device__ inline double block_sum(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kPoissonThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__device__ inline double block_max(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kPoissonThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// Row g of L = -d/dx(eps_r d/dx) (harmonic-mean faces), assembled as the
// dense reference does (diagonal, then the left and the right neighbour, a
// later write replacing an earlier one when they coincide).
struct LRow {
    int c[3];
    double v[3];
};

__device__ inline LRow l_row(int N, int g, const double* er, double dx2) {
    auto face = [&](int i) {
        const int ip = i == N - 1 ? 0 : i + 1;
        return 2.0 * er[i] * er[ip] / (er[i] + er[ip]);
    };
    const int gm = g == 0 ? N - 1 : g - 1, gp = g == N - 1 ? 0 : g + 1;
    LRow r;
    r.c[0] = g;
    r.v[0] = (face(gm) + face(g)) / dx2;
    r.c[1] = gm;
    r.v[1] = -face(gm) / dx2;
    r.c[2] = gp;
    r.v[2] = -face(g) / dx2;
    // later writes win on coinciding columns (N <= 2)
    if (r.c[2] == r.c[1]) r.v[1] = r.v[2], r.c[2] = -1;
    if (r.c[1] == r.c[0]) r.v[0] = r.v[1], r.c[1] = -1;
    if (r.c[2] == r.c[0]) r.v[0] = r.v[2], r.c[2] = -1;
    // ascending columns (dropped entries last): the dense row order
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b) {
            const unsigned ca = (unsigned)r.c[a], cb = (unsigned)r.c[b];  // -1 sorts last
            if (cb < ca) {
                const int tc = r.c[a];
                const double tv = r.v[a];
                r.c[a] = r.c[b], r.v[a] = r.v[b];
                r.c[b] = tc, r.v[b] = tv;
            }
        }
    return r;
}

__global__ __launch_bounds__(kPoissonThreads) void k_periodic_poisson(int N, const double* er, const double* wx,
                                                                   const double* ND, double lambdaP,
                                                                   const double* nbar, double* M, double* x,
                                                                   int* piv, double* U, double* out,
                                                                   DeviceStatus* status) {
    __shared__ double red[kPoissonThreads];
    __shared__ int redi[kPoissonThreads];
    __shared__ int rows[kMaxPoisson + 1], cols[kMaxPoisson + 1];
    __shared__ int nrow, ncol;
    const int n = N + 1, t = threadIdx.x;
    const double dx2 = wx[0] * wx[0];
    // sP: the largest absolute row sum of L
    double rs = 0.0;
    for (int g = t; g < N; g += kPoissonThreads) {
        const LRow r = l_row(N, g, er, dx2);
        double s = 0.0;
        for (int k = 0; k < 3; ++k)
            if (r.c[k] >= 0) s += fabs(r.v[k]);
        rs = fmax(rs, s);
    }
    const double sP = block_max(rs, red);
    const double v = 1.0 / sqrt((double)N);
    for (long long f = t; f < (long long)n * n; f += kPoissonThreads) M[f] = 0.0;
    __syncthreads();
    for (int g = t; g < N; g += kPoissonThreads) {
        const LRow r = l_row(N, g, er, dx2);
        for (int k = 0; k < 3; ++k)
            if (r.c[k] >= 0) M[g + (size_t)n * r.c[k]] = r.v[k] / sP;
        M[g + (size_t)n * N] = v;
        M[N + (size_t)n * g] = v;
        x[g] = -lambdaP * (ND[g] - nbar[g]) / sP;
    }
    if (t == 0) x[N] = 0.0;
    __syncthreads();
    bool singular = false;
    for (int k = 0; k < n; ++k) {
        // pivot: largest |M_ik|, i >= k, first on ties
        double best = -1.0;
        int bi = n;
        for (int i = k + t; i < n; i += kPoissonThreads) {
            const double a = fabs(M[i + (size_t)n * k]);
            if (a > best) best = a, bi = i;
        }
        red[t] = best;
        redi[t] = bi;
        __syncthreads();
        for (int s = kPoissonThreads / 2; s > 0; s >>= 1) {
            if (t < s) {
                const double o = red[t + s];
                const int oi = redi[t + s];
                if (o > red[t] || (o == red[t] && oi < redi[t])) red[t] = o, redi[t] = oi;
            }
            __syncthreads();
        }
        const int p = redi[0];
        const double pv = red[0];
        __syncthreads();
        if (!(pv > 0.0)) {
            singular = true;
            break;
        }
        if (t == 0) piv[k] = p;
        if (p != k)
            for (int j = t; j < n; j += kPoissonThreads) {
                const double a = M[k + (size_t)n * j];
                M[k + (size_t)n * j] = M[p + (size_t)n * j];
                M[p + (size_t)n * j] = a;
            }
        __syncthreads();
        const double d = M[k + (size_t)n * k];
        if (t == 0) nrow = ncol = 0;
        __syncthreads();
        // The update touches only rows with a nonzero multiplier and columns
        // with a nonzero pivot-row entry (skipped terms are exact zeros):
        // the system is tridiagonal-cyclic with a border, so both lists stay
        // a few entries long.
        for (int i = k + 1 + t; i < n; i += kPoissonThreads) {
            const double mult = M[i + (size_t)n * k] / d;
            M[i + (size_t)n * k] = mult;
            if (mult != 0.0) rows[atomicAdd(&nrow, 1)] = i;
        }
        for (int j = k + 1 + t; j < n; j += kPoissonThreads)
            if (M[k + (size_t)n * j] != 0.0) cols[atomicAdd(&ncol, 1)] = j;
        __syncthreads();
        const int nr = nrow, nc = ncol;
        for (long long f = t; f < (long long)nr * nc; f += kPoissonThreads) {
            const int i = rows[f % nr], j = cols[f / nr];
            M[i + (size_t)n * j] -= M[i + (size_t)n * k] * M[k + (size_t)n * j];
        }
        __syncthreads();
    }
    if (singular) {
        if (t == 0) report_status(status, QCLHIP_ERR_DOMAIN, 0);
        return;
    }
    if (t == 0)
        for (int k = 0; k < n; ++k) {
            const int p = piv[k];
            if (p != k) {
                const double a = x[k];
                x[k] = x[p];
                x[p] = a;
            }
        }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        const double xj = x[j];
        for (int i = j + 1 + t; i < n; i += kPoissonThreads) x[i] -= M[i + (size_t)n * j] * xj;
        __syncthreads();
    }
    for (int j = n - 1; j >= 0; --j) {
        if (t == 0) x[j] /= M[j + (size_t)n * j];
        __syncthreads();
        const double xj = x[j];
        for (int i = t; i < j; i += kPoissonThreads) x[i] -= M[i + (size_t)n * j] * xj;
        __syncthreads();
    }
    // residuals: ||L U - b|| / (||b|| + lambda_P 1e-14), neutrality
    double rn = 0.0, bn = 0.0, nd = 0.0, ne = 0.0;
    for (int g = t; g < N; g += kPoissonThreads) {
        U[g] = x[g];
        const LRow r = l_row(N, g, er, dx2);
        double s = 0.0;
        for (int a = 0; a < 3; ++a)
            if (r.c[a] >= 0) s += r.v[a] * x[r.c[a]];
        const double b = -lambdaP * (ND[g] - nbar[g]);
        rn += (s - b) * (s - b);
        bn += b * b;
        nd += wx[g] * ND[g];
        ne += wx[g] * (ND[g] - nbar[g]);
    }
    rn = block_sum(rn, red);
    bn = block_sum(bn, red);
    nd = block_sum(nd, red);
    ne = block_sum(ne, red);
    if (t == 0) {
        out[0] = x[N];
        out[1] = sqrt(rn) / (sqrt(bn) + lambdaP * 1e-14);
        out[2] = fabs(ne) / nd;
    }
}

// h_m = (S + Phi^dagger diag(U) Phi) + kin_m Minv: one wavefront per (a,b).
__global__ __launch_bounds__(64) void k_project(int Nz, int Nb, int Nk, const c64* Phi, const double* U,
                                              const c64* S, const double* kin, const c64* Minv, c64* h) {
    const int ab = blockIdx.x, a = ab % Nb, b = ab / Nb, lane = threadIdx.x;
    double ur = 0.0, ui = 0.0;
    for (int g = lane; g < Nz; g += 64) {
        const c64 pa = Phi[g + (size_t)Nz * a], pb = Phi[g + (size_t)Nz * b];
        const c64 t = cscale(U[g], cconj(pa));
        const c64 v = cmul(t, pb);
        ur += v.x;
        ui += v.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        ur += shfl_xor_d(ur, off);
        ui += shfl_xor_d(ui, off);
    }
    const c64 st = cadd(S[ab], cmk(ur, ui));
    const c64 mi = Minv[ab];
    for (int m = lane; m < Nk; m += 64) h[m + (size_t)Nk * ab] = cadd(st, cscale(kin[m], mi));
}

// U <- (1 - alpha) U + alpha cand; out = max |cand - U_old|.
__global__ __launch_bounds__(kPoissonThreads) void k_mix_potential(int N, double* U, const double* cand,
                                                                double alpha, double* out) {
    __shared__ double red[kPoissonThreads];
    double r = 0.0;
    for (int g = threadIdx.x; g < N; g += kPoissonThreads) {
        const double u = U[g], c = cand[g];
        r = fmax(r, fabs(c - u));
        U[g] = (1.0 - alpha) * u + alpha * c;
    }
    r = block_max(r, red);
    if (threadIdx.x == 0) out[0] = r;
}

}  // namespace qclhip

using namespace qclhip;

extern "C" {

int qclhip_electron_density(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GL, int Nz, const double* chi,
                            double gs, double* nbar) {
    if (!ctx || !grid || !GL || !chi || !nbar || Nz < 1) return QCLHIP_ERR_ARGUMENT;
    const int Nb = grid->Nb;
    c64* I = (c64*)workspace(ctx, (size_t)Nb * Nb * sizeof(c64));
    if (!I) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    hipLaunchKernelGGL(k_density_integrate, dim3(Nb * Nb), dim3(kDensityThreads), 0, ctx->stream, grid->NE,
                       (long long)grid->NE * grid->Nk, (const c64*)GL, grid->wE, grid->wk, I);
    hipLaunchKernelGGL(k_density_project, dim3((Nz + 255) / 256), dim3(256), 0, ctx->stream, Nz, Nb,
                       (const c64*)chi, (const c64*)I, gs / (2.0 * M_PI), nbar);
    return launch_check(ctx, "electron density");
}

int qclhip_periodic_poisson(qclhip_ctx* ctx, int N, const double* eps_r, const double* wx, const double* ND,
                            double lambdaP, const double* nbar, double* U, double* out_host) {
    if (!ctx || !eps_r || !wx || !ND || !nbar || !U || !out_host) return QCLHIP_ERR_ARGUMENT;
    if (N < 1 || N > kMaxPoisson) return fail(ctx, QCLHIP_ERR_ARGUMENT, "Poisson grid size must lie in 1..4096");
    if (!std::isfinite(lambdaP)) return fail(ctx, QCLHIP_ERR_ARGUMENT, "lambda_P must be finite");
    const size_t n = (size_t)N + 1;
    const size_t bytes = (n * n + n + 8) * sizeof(double) + n * sizeof(int);
    double* M = (double*)workspace(ctx, bytes);
    if (!M) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    double* x = M + n * n;
    double* out = x + n;
    int* piv = (int*)(out + 8);
    hipLaunchKernelGGL(k_periodic_poisson, dim3(1), dim3(kPoissonThreads), 0, ctx->stream, N, eps_r, wx, ND,
                       lambdaP, nbar, M, x, piv, U, out, ctx->status);
    int rc = launch_check(ctx, "k_periodic_poisson");
    if (rc) return rc;
    bool deferred = false;
    rc = defer_result(ctx, out, 3, &deferred);
    if (rc || deferred) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(out_host, out, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return qclhip_ctx_synchronize(ctx);
}

int qclhip_project_hamiltonians(qclhip_ctx* ctx, int Nz, int Nb, int Nk, const double* Phi, const double* U,
                                const double* S, const double* kin, const double* Minv, double* h) {
    if (!ctx || !Phi || !U || !S || !kin || !Minv || !h) return QCLHIP_ERR_ARGUMENT;
    if (Nz < 1 || Nb < 1 || Nk < 1) return fail(ctx, QCLHIP_ERR_ARGUMENT, "sizes must be positive");
    hipLaunchKernelGGL(k_project, dim3(Nb * Nb), dim3(64), 0, ctx->stream, Nz, Nb, Nk, (const c64*)Phi, U,
                       (const c64*)S, kin, (const c64*)Minv, (c64*)h);
    return launch_check(ctx, "k_project");
}

int qclhip_mix_potential(qclhip_ctx* ctx, int N, double* U, const double* cand, double alpha, double* rU_host) {
    if (!ctx || !U || !cand || !rU_host || N < 1) return QCLHIP_ERR_ARGUMENT;
    if (!(alpha > 0 && alpha <= 1)) return fail(ctx, QCLHIP_ERR_ARGUMENT, "mixing factor must lie in (0,1]");
    double* out = (double*)workspace(ctx, 8 * sizeof(double));
    if (!out) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    hipLaunchKernelGGL(k_mix_potential, dim3(1), dim3(kPoissonThreads), 0, ctx->stream, N, U, cand, alpha, out);
    int rc = launch_check(ctx, "k_mix_potential");
    if (rc) return rc;
    bool deferred = false;
    rc = defer_result(ctx, out, 1, &deferred);
    if (rc || deferred) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(rU_host, out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return qclhip_ctx_synchronize(ctx);
}

}  // extern "C"
