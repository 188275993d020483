// Scattering-kernel contractions (production_operators.jl:402-658) and the
// thermal LO map (:1089-1162).
//
// With G stored energy-fastest, G viewed as the column-major matrix
// G2 (NE x Nc), Nc = Nk*Nb^2 (column q = m + Nk*(a + Nb*b)), every
// representation is a short chain of plain GEMMs, which run on the FP64
// matrix cores (k_zgemm_nt, below):
//   dense:      S2 = f * G2 * (K diag(w))^T                 (NE x Nc)
//   low rank:   S2 = f * (G2 * (R^+ diag(w))^T) * left^T
//   momentum independent: X(e,(c,d)) = sum_m' wk_m' G(e,m',c,d) (kernel),
//               Y = X * block^T, S(e,m,(a,b)) = f*Y(e,(a,b)) for every m
//   separable IFR: C1 = Gm * right^T (Gm = G as (NE*Nk) x Nb^2),
//               C2_j = C1_j * (radial diag(wk))^T per interface j (batched),
//               S = f * C2 * left^T
// with w(q) = wk(q mod Nk) folded into the device copy of the kernel.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

struct qclhip_kernel {
    int kind = 0, rank = 0, NE = 0, Nk = 0, Nb = 0;
    // Momentum-independent form: 0 = the quadrature weights enter the
    // momentum sum (the forward map), 1 = they scale the broadcast output
    // (its adjoint).
    int weights_out = 0;
    const qclhip_grid* grid = nullptr;
    qclhip::c64* a = nullptr;
    qclhip::c64* b = nullptr;
    qclhip::c64* c = nullptr;
};

namespace qclhip {

int lo_combine(qclhip_ctx* ctx, int NE, long long n, const c64* Y, double ca, const qclhip_shift* a, double cb,
               const qclhip_shift* b, c64* out);

namespace {

__global__ void k_mi_reduce(int NE, int Nk, int N2, const double* __restrict__ wk, const c64* __restrict__ G,
                            c64* __restrict__ X) {
    const long long n = (long long)NE * N2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const int e = (int)(idx % NE);
        const long long q = idx / NE;
        c64 v = cmk(0.0, 0.0);
        for (int mp = 0; mp < Nk; ++mp) {
            c64 g = G[e + (long long)NE * (mp + (long long)Nk * q)];
            const double w = wk ? wk[mp] : 1.0;
            v = cmk(v.x + w * g.x, v.y + w * g.y);
        }
        X[idx] = v;
    }
}

// S[e + NE (m + Nk q)] = f w_m Y[e + NE q] + beta S; thread = energy, block
// row y takes a contiguous range of the columns (m, q), so that Y is read
// once per q (Nk consecutive columns) instead of once per column.
__global__ void k_mi_broadcast(int NE, int Nk, int N2, const c64* __restrict__ Y, double f, double beta,
                               const double* __restrict__ wout, c64* __restrict__ S) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NE) return;
    const long long ncols = (long long)Nk * N2;
    const long long cchunk = (ncols + gridDim.y - 1) / gridDim.y;
    const long long c0 = (long long)blockIdx.y * cchunk, c1 = min(ncols, c0 + cchunk);
    long long qcur = -1;
    c64 y = cmk(0.0, 0.0);
    for (long long col = c0; col < c1; ++col) {
        const long long q = col / Nk;  // uniform over the block
        const int m = (int)(col - q * Nk);
        const long long idx = col * NE + e;
        if (q != qcur) {
            y = Y[e + (long long)NE * q];
            qcur = q;
        }
        const double fw = wout ? f * wout[m] : f;
        c64 v = cmk(fw * y.x, fw * y.y);
        if (beta != 0.0) {
            const c64 o = S[idx];
            v = cmk(beta * o.x + v.x, beta * o.y + v.y);
        }
        S[idx] = v;
    }
}

// ---- complex FP64 GEMM on the matrix cores -------------------------------
// C (m x n, column-major, ld ldc) = alpha A (m x k, ld lda) B^T (B n x k, ld
// ldb) + beta C. A 256-thread workgroup owns a 64 x 64 tile of C; its four
// waves own 32 x 32 quadrants, i.e. 2 x 2 MFMA tiles of 16 x 16 with a real
// and an imaginary accumulator each. K advances in stages of 16: the A and B
// panels (64 x 16 complex each) are loaded coalesced (column-major: 64
// consecutive rows of one column per pass) into LDS, then each wave issues,
// per 4-deep step and MFMA tile, the four real 16x16x4 FP64 MFMAs of the
// complex product (re += ar br - ai bi, im += ar bi + ai br). The 16x16x4
// FP64 operand layout: lane l holds A[l%16][l/16] and B[l/16][l%16]; result
// element v of lane l is row (l/16) + 4v, column l%16.
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int kGemmTile = 64;
constexpr int kGemmStage = 16;

__global__ __launch_bounds__(256) void k_zgemm_nt(long long m, long long n, long long k, double alpha,
                                                  const c64* __restrict__ A, long long lda,
                                                  const c64* __restrict__ Bm, long long ldb, double beta,
                                                  c64* __restrict__ C, long long ldc) {
    __shared__ c64 As[kGemmStage][kGemmTile + 1];
    __shared__ c64 Bs[kGemmStage][kGemmTile + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const long long i0 = (long long)blockIdx.x * kGemmTile, j0 = (long long)blockIdx.y * kGemmTile;
    const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
    f64x4 accR[2][2], accI[2][2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) accR[t][u][v] = accI[t][u][v] = 0.0;
    for (long long p0 = 0; p0 < k; p0 += kGemmStage) {
        // 64 x 16 complex per panel: four elements per thread, 64 rows of one
        // column per group of 64 threads (1 KiB contiguous)
#pragma unroll
        for (int r = 0; r < (kGemmTile * kGemmStage) / 256; ++r) {
            const int e = tid + 256 * r;
            const int ii = e % kGemmTile, pp = e / kGemmTile;
            const long long gi = i0 + ii, gj = j0 + ii, gp = p0 + pp;
            As[pp][ii] = (gi < m && gp < k) ? A[gi + lda * gp] : cmk(0.0, 0.0);
            Bs[pp][ii] = (gj < n && gp < k) ? Bm[gj + ldb * gp] : cmk(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll
        for (int ps = 0; ps < kGemmStage; ps += 4) {
            c64 a[2], b[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) a[t] = As[ps + lk][wi + 16 * t + li];
#pragma unroll
            for (int u = 0; u < 2; ++u) b[u] = Bs[ps + lk][wj + 16 * u + li];
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    accR[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t].x, b[u].x, accR[t][u], 0, 0, 0);
                    accR[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[t].y, b[u].y, accR[t][u], 0, 0, 0);
                    accI[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t].x, b[u].y, accI[t][u], 0, 0, 0);
                    accI[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t].y, b[u].x, accI[t][u], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long i = i0 + wi + 16 * t + lk + 4 * v, j = j0 + wj + 16 * u + li;
                if (i < m && j < n) {
                    c64 out = cmk(alpha * accR[t][u][v], alpha * accI[t][u][v]);
                    if (beta != 0.0) {
                        const c64 o = C[i + ldc * j];
                        out = cmk(fma(beta, o.x, out.x), fma(beta, o.y, out.y));
                    }
                    C[i + ldc * j] = out;
                }
            }
}

// C (m x n, ld m) = alpha * A (m x k, ld lda) * B^T (B is n x k, ld ldb) + beta * C
int gemm_nt(qclhip_ctx* ctx, long long m, long long n, long long k, double alpha, const c64* A, long long lda,
            const c64* Bm, long long ldb, double beta, c64* C, long long ldc) {
    if (m <= 0 || n <= 0) return QCLHIP_OK;
    const long long gx = (m + kGemmTile - 1) / kGemmTile, gy = (n + kGemmTile - 1) / kGemmTile;
    if (gy > 65535) return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "contraction too wide for one launch");
    // blockIdx.x over rows: the workgroups that share a B panel run together
    hipLaunchKernelGGL(k_zgemm_nt, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, ctx->stream, m, n, k, alpha, A, lda,
                       Bm, ldb, beta, C, ldc);
    return launch_check(ctx, "k_zgemm_nt");
}

// Streaming kernels for the thin factors (rank J or r of a few), where a
// plain GEMM library tiles poorly: each thread owns one row of the tall
// operand and walks its long dimension, so every load and store is
// coalesced across threads and the field is read or written exactly once.
constexpr int kMaxRank = 16;

// C[row, j] = sum_q A[row + M q] * Bm[j + J q]  (q < K), J <= kMaxRank.
__global__ __launch_bounds__(1024) void k_thin_reduce(long long M, int K, int J, const c64* __restrict__ A, const c64* __restrict__ Bm,
                              c64* __restrict__ C) {
    extern __shared__ c64 sb[];
    for (int idx = threadIdx.x; idx < J * K; idx += blockDim.x) sb[idx] = Bm[idx];
    __syncthreads();
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= M) return;
    c64 acc[kMaxRank];
#pragma unroll
    for (int j = 0; j < kMaxRank; ++j) acc[j] = cmk(0.0, 0.0);
    for (int q = 0; q < K; ++q) {
        const c64 a = A[row + M * q];
#pragma unroll
        for (int j = 0; j < kMaxRank; ++j)
            if (j < J) acc[j] = cfma(acc[j], a, sb[j + J * q]);
    }
    // Unrolled with a guard: a runtime-bounded loop over acc would index the
    // register array dynamically and move it to scratch.
#pragma unroll
    for (int j = 0; j < kMaxRank; ++j)
        if (j < J) C[row + M * j] = acc[j];
}

// S[row + M q] = f * sum_j C[row + M j] * Bm[q + K j] + beta * S  (q < K).
__global__ __launch_bounds__(1024) void k_thin_expand(long long M, int K, int J, const c64* __restrict__ C, const c64* __restrict__ Bm,
                              double f, double beta, c64* __restrict__ S) {
    extern __shared__ c64 sb[];
    for (int idx = threadIdx.x; idx < J * K; idx += blockDim.x) sb[idx] = Bm[idx];
    __syncthreads();
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= M) return;
    c64 c[kMaxRank];
#pragma unroll
    for (int j = 0; j < kMaxRank; ++j) c[j] = (j < J) ? C[row + M * j] : cmk(0.0, 0.0);
    for (int q = 0; q < K; ++q) {
        c64 v = cmk(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < kMaxRank; ++j)
            if (j < J) v = cfma(v, c[j], sb[q + K * j]);
        v = cmk(f * v.x, f * v.y);
        if (beta != 0.0) {
            const c64 o = S[row + M * q];
            v = cmk(beta * o.x + v.x, beta * o.y + v.y);
        }
        S[row + M * q] = v;
    }
}

// Long reduction of the low-rank factor: partial[chunk][e][i] over a chunk
// of the compound index q (rows e < NE, ranks i < r).
constexpr int kQChunk = 256;
__global__ void k_lowrank_partial(int NE, long long Nc, int r, const c64* __restrict__ G,
                                  const c64* __restrict__ RAw, c64* __restrict__ partial) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const long long q0 = (long long)blockIdx.y * kQChunk;
    if (e >= NE) return;
    c64 acc[kMaxRank];
#pragma unroll
    for (int i = 0; i < kMaxRank; ++i) acc[i] = cmk(0.0, 0.0);
    const long long q1 = min(Nc, q0 + kQChunk);
    for (long long q = q0; q < q1; ++q) {
        const c64 g = G[e + (long long)NE * q];
#pragma unroll
        for (int i = 0; i < kMaxRank; ++i)
            if (i < r) acc[i] = cfma(acc[i], g, RAw[i + (long long)r * q]);
    }
#pragma unroll
    for (int i = 0; i < kMaxRank; ++i)
        if (i < r) partial[((long long)blockIdx.y * r + i) * NE + e] = acc[i];
}

// red[e + NE i] = sum over chunks, in chunk order.
__global__ void k_lowrank_sum(int NE, int r, int nchunk, const c64* __restrict__ partial, c64* __restrict__ red) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)NE * r) return;
    const int e = (int)(idx % NE), i = (int)(idx / NE);
    c64 v = cmk(0.0, 0.0);
    for (int ch = 0; ch < nchunk; ++ch) v = cadd(v, partial[((long long)ch * r + i) * NE + e]);
    red[e + (long long)NE * i] = v;
}

// S2[e + NE q] = f sum_i red[e + NE i] left[q + Nc i] + beta S2; thread =
// energy (its r reduced values held in registers); block row y takes a
// contiguous chunk of q, whose factor entries are staged in LDS first (one
// load pass instead of a dependent load per q).
constexpr int kExpandChunk = 128;
__global__ void k_lowrank_expand(int NE, long long Nc, int r, const c64* __restrict__ red,
                                 const c64* __restrict__ left, double f, double beta, c64* __restrict__ S) {
    __shared__ c64 lf[kExpandChunk * kMaxRank];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    c64 rv[kMaxRank];
#pragma unroll
    for (int i = 0; i < kMaxRank; ++i) rv[i] = (e < NE && i < r) ? red[e + (long long)NE * i] : cmk(0.0, 0.0);
    for (long long qb = (long long)blockIdx.y * kExpandChunk; qb < Nc; qb += (long long)gridDim.y * kExpandChunk) {
        const int nq = (int)min((long long)kExpandChunk, Nc - qb);
        __syncthreads();  // the previous chunk's factor entries are no longer read
        for (int t = threadIdx.x; t < nq * r; t += blockDim.x) {
            const int qq = t % nq, i = t / nq;
            lf[qq * kMaxRank + i] = left[qb + qq + Nc * i];
        }
        __syncthreads();
        if (e >= NE) continue;
        for (int qq = 0; qq < nq; ++qq) {
            const long long q = qb + qq;
            c64 v = cmk(0.0, 0.0);
#pragma unroll
            for (int i = 0; i < kMaxRank; ++i)
                if (i < r) v = cfma(v, rv[i], lf[qq * kMaxRank + i]);
            v = cmk(f * v.x, f * v.y);
            const long long idx = q * NE + e;
            if (beta != 0.0) {
                const c64 o = S[idx];
                v = cmk(beta * o.x + v.x, beta * o.y + v.y);
            }
            S[idx] = v;
        }
    }
}

// IFR radial step: C2[(e,m), j] = sum_m' Rw[m, m'] C1[(e,m'), j] for every (e, j).
__global__ void k_ifr_radial(int NE, int Nk, int J, const c64* __restrict__ C1, const c64* __restrict__ Rw,
                             c64* __restrict__ C2) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long M = (long long)NE * Nk;
    if (idx >= M * J) return;
    const int e = (int)(idx % NE);
    const long long rest = idx / NE;
    const int m = (int)(rest % Nk), j = (int)(rest / Nk);
    c64 v = cmk(0.0, 0.0);
    for (int mp = 0; mp < Nk; ++mp) v = cfma(v, Rw[m + (long long)Nk * mp], C1[e + (long long)NE * mp + M * j]);
    C2[e + (long long)NE * m + M * j] = v;
}

}  // namespace

// Thermal LO step of a low-rank kernel (rank <= kMaxRank): both energy
// shifts act on the NE x r reduced coefficients red = K_right[w G] before
// the expansion, S = qK left (ca S_a red + cb S_b red)^T, exact since the
// shifts act on energy and the expansion on (m,a,b); the full-size
// intermediate K[G] and its two shifted reads are not formed.
static int lo_lowrank(qclhip_ctx* ctx, const qclhip_kernel* K, const c64* G, double qK, double ca,
                      const qclhip_shift* a, double cb, const qclhip_shift* b, c64* S) {
    const qclhip_grid* g = K->grid;
    const long long NE = g->NE, Nc = (long long)g->Nk * g->Nb * g->Nb, r = K->rank;
    const long long nchunk = (Nc + kQChunk - 1) / kQChunk;
    c64* red = (c64*)workspace(ctx, (size_t)(2 * NE * r + nchunk * r * NE) * sizeof(c64));
    if (!red) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    c64* coef = red + NE * r;
    c64* partial = coef + NE * r;
    hipLaunchKernelGGL(k_lowrank_partial, dim3((unsigned)((NE + 255) / 256), (unsigned)nchunk), dim3(256), 0,
                       ctx->stream, (int)NE, Nc, (int)r, G, K->b, partial);
    hipLaunchKernelGGL(k_lowrank_sum, dim3((unsigned)((NE * r + 255) / 256)), dim3(256), 0, ctx->stream, (int)NE,
                       (int)r, (int)nchunk, partial, red);
    int st = lo_combine(ctx, (int)NE, NE * r, red, ca, a, cb, b, coef);
    if (st) return st;
    hipLaunchKernelGGL(k_lowrank_expand, ecol_grid((int)NE, (Nc + kExpandChunk - 1) / kExpandChunk, ctx->sm_count),
                       dim3(256), 0, ctx->stream, (int)NE, Nc, (int)r, coef, K->a, qK, 0.0, S);
    return launch_check(ctx, "low-rank LO contraction");
}

// Device scratch that contract_core takes from the context workspace.
size_t core_scratch_bytes(const qclhip_kernel* K) {
    const qclhip_grid* g = K->grid;
    const size_t NE = g->NE, Nk = g->Nk, N2 = (size_t)g->Nb * g->Nb, r = (size_t)std::max(1, K->rank);
    switch (K->kind) {
        case QCLHIP_KERNEL_LOW_RANK: return (NE * r + ((Nk * N2 + 255) / 256) * r * NE) * sizeof(c64);
        case QCLHIP_KERNEL_MOMENTUM_INDEPENDENT: return 2 * NE * N2 * sizeof(c64);
        case QCLHIP_KERNEL_SEPARABLE_IFR: return 2 * NE * Nk * r * sizeof(c64);
        default: return 0;
    }
}

// The momentum-independent representation: X = sum_m' w_m' G(., m', .),
// Y = X block^T, then S(e, m, q) = f Y(e, q) + beta S for nk_out momenta:
// nk_out = Nk writes the field, nk_out = 1 the single column set that every
// momentum holds in the forward form (the values of each momentum's columns
// bit for bit).
int mi_contract(qclhip_ctx* ctx, const qclhip_kernel* K, const c64* G, double f, double beta, c64* S, int nk_out) {
    const qclhip_grid* g = K->grid;
    const long long NE = g->NE, Nk = g->Nk, N2 = (long long)g->Nb * g->Nb;
    c64* X = (c64*)workspace(ctx, (size_t)2 * NE * N2 * sizeof(c64));
    if (!X) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    c64* Y = X + NE * N2;
    hipLaunchKernelGGL(k_mi_reduce, dim3(grid_blocks(NE * N2, 256, ctx->sm_count)), dim3(256), 0, ctx->stream,
                       (int)NE, (int)Nk, (int)N2, K->weights_out ? (const double*)nullptr : g->wk, G, X);
    int st = launch_check(ctx, "k_mi_reduce");
    if (st) return st;
    st = gemm_nt(ctx, NE, N2, N2, 1.0, X, NE, K->a, N2, 0.0, Y, NE);
    if (st) return st;
    hipLaunchKernelGGL(k_mi_broadcast, ecol_grid((int)NE, (long long)nk_out * N2, ctx->sm_count), dim3(256), 0,
                       ctx->stream, (int)NE, nk_out, (int)N2, Y, f, beta,
                       K->weights_out ? (const double*)g->wk : nullptr, S);
    return launch_check(ctx, "k_mi_broadcast");
}

// S <- f K[w G] + beta S for the given representation.
int contract_core(qclhip_ctx* ctx, const qclhip_kernel* K, const c64* G, double f, double beta, c64* S) {
    const qclhip_grid* g = K->grid;
    const long long NE = g->NE, Nk = g->Nk, N2 = (long long)g->Nb * g->Nb, Nc = Nk * N2;
    switch (K->kind) {
        case QCLHIP_KERNEL_DENSE:
            return gemm_nt(ctx, NE, Nc, Nc, f, G, NE, K->a, Nc, beta, S, NE);
        case QCLHIP_KERNEL_LOW_RANK: {
            const long long r = K->rank;
            if (r > kMaxRank) {
                c64* red = (c64*)workspace(ctx, (size_t)NE * r * sizeof(c64));
                if (!red) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
                int st = gemm_nt(ctx, NE, r, Nc, 1.0, G, NE, K->b, r, 0.0, red, NE);
                if (st) return st;
                return gemm_nt(ctx, NE, Nc, r, f, red, NE, K->a, Nc, beta, S, NE);
            }
            const long long nchunk = (Nc + kQChunk - 1) / kQChunk;
            c64* red = (c64*)workspace(ctx, (size_t)(NE * r + nchunk * r * NE) * sizeof(c64));
            if (!red) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
            c64* partial = red + NE * r;
            hipLaunchKernelGGL(k_lowrank_partial, dim3((unsigned)((NE + 255) / 256), (unsigned)nchunk), dim3(256), 0,
                               ctx->stream, (int)NE, Nc, (int)r, G, K->b, partial);
            hipLaunchKernelGGL(k_lowrank_sum, dim3((unsigned)((NE * r + 255) / 256)), dim3(256), 0, ctx->stream,
                               (int)NE, (int)r, (int)nchunk, partial, red);
            // S2 (NE x Nc) = f red left^T + beta S2, streamed over q.
            hipLaunchKernelGGL(k_lowrank_expand, ecol_grid((int)NE, (Nc + kExpandChunk - 1) / kExpandChunk, ctx->sm_count),
                               dim3(256), 0, ctx->stream, (int)NE, Nc, (int)r, red, K->a, f, beta, S);
            return launch_check(ctx, "low-rank contraction");
        }
        case QCLHIP_KERNEL_MOMENTUM_INDEPENDENT:
            return mi_contract(ctx, K, G, f, beta, S, (int)Nk);
        case QCLHIP_KERNEL_SEPARABLE_IFR: {
            const long long J = K->rank, M = NE * Nk;
            // The J x N^2 factor sits in LDS; above 64 KiB (e.g. six interfaces
            // at N = 32: 96 KiB) one workgroup fills a CU, so it takes 1024
            // threads (16 waves) to keep the strided field reads in flight.
            const size_t lds = (size_t)J * N2 * sizeof(c64);
            if (J > kMaxRank || lds > ctx->max_lds_per_block)
                return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "separable kernel rank too large for this build");
            const int threads = lds > 65536 ? 1024 : 256;
            if (lds > 65536) {
                QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_thin_reduce,
                                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_thin_expand,
                                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            }
            c64* C1 = (c64*)workspace(ctx, (size_t)2 * M * J * sizeof(c64));
            if (!C1) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
            c64* C2 = C1 + M * J;
            const unsigned blocks = (unsigned)((M + threads - 1) / threads);
            hipLaunchKernelGGL(k_thin_reduce, dim3(blocks), dim3(threads), lds, ctx->stream, M, (int)N2, (int)J, G,
                               K->b, C1);
            hipLaunchKernelGGL(k_ifr_radial, dim3((unsigned)((M * J + 255) / 256)), dim3(256), 0, ctx->stream,
                               (int)NE, (int)Nk, (int)J, C1, K->c, C2);
            hipLaunchKernelGGL(k_thin_expand, dim3(blocks), dim3(threads), lds, ctx->stream, M, (int)N2, (int)J, C2,
                               K->a, f, beta, S);
            return launch_check(ctx, "separable contraction");
        }
        default:
            return fail(ctx, QCLHIP_ERR_ARGUMENT, "unknown kernel representation");
    }
}

}  // namespace qclhip

using namespace qclhip;

extern "C" {

int qclhip_kernel_create(qclhip_ctx* ctx, const qclhip_grid* grid, int kind, int rank, const double* a_host,
                         const double* b_host, const double* c_host, qclhip_kernel** out) {
    if (!ctx || !grid || !out || !a_host) return QCLHIP_ERR_ARGUMENT;
    *out = nullptr;
    const long long Nk = grid->Nk, N2 = (long long)grid->Nb * grid->Nb, Nc = Nk * N2;
    typedef std::complex<double> cd;
    const cd* A = reinterpret_cast<const cd*>(a_host);
    const cd* Bh = reinterpret_cast<const cd*>(b_host);
    const cd* Ch = reinterpret_cast<const cd*>(c_host);
    std::vector<cd> da, db, dc;
    const std::vector<double>& wk = grid->wk_host;
    switch (kind) {
        case QCLHIP_KERNEL_DENSE:
            da.resize((size_t)Nc * Nc);
            for (long long qp = 0; qp < Nc; ++qp)
                for (long long q = 0; q < Nc; ++q) da[q + Nc * qp] = A[q + Nc * qp] * wk[qp % Nk];
            break;
        case QCLHIP_KERNEL_MOMENTUM_INDEPENDENT:
            da.assign(A, A + N2 * N2);
            break;
        case QCLHIP_KERNEL_SEPARABLE_IFR:
            if (rank < 1 || !Bh || !Ch) return fail(ctx, QCLHIP_ERR_ARGUMENT, "IFR kernel needs left, right, radial");
            da.assign(A, A + N2 * rank);
            db.assign(Bh, Bh + (long long)rank * N2);
            dc.resize((size_t)Nk * Nk);
            for (long long mp = 0; mp < Nk; ++mp)
                for (long long m = 0; m < Nk; ++m) dc[m + Nk * mp] = Ch[m + Nk * mp] * wk[mp];
            break;
        case QCLHIP_KERNEL_LOW_RANK:
            if (rank < 1 || !Bh) return fail(ctx, QCLHIP_ERR_ARGUMENT, "low-rank kernel needs left and right_adjoint");
            da.assign(A, A + Nc * rank);
            db.resize((size_t)rank * Nc);
            for (long long q = 0; q < Nc; ++q)
                for (long long i = 0; i < rank; ++i) db[i + rank * q] = Bh[i + rank * q] * wk[q % Nk];
            break;
        default:
            return fail(ctx, QCLHIP_ERR_ARGUMENT, "unknown kernel representation");
    }
    auto* K = new qclhip_kernel();
    K->kind = kind;
    K->rank = rank;
    K->grid = grid;
    K->Nk = (int)Nk;
    K->Nb = grid->Nb;
    bool ok = true;
    auto upload = [&](std::vector<cd>& h, c64** d) {
        if (h.empty() || !ok) return;
        ok = hipMalloc(d, h.size() * sizeof(c64)) == hipSuccess &&
             hipMemcpy(*d, h.data(), h.size() * sizeof(c64), hipMemcpyHostToDevice) == hipSuccess;
    };
    upload(da, &K->a);
    upload(db, &K->b);
    upload(dc, &K->c);
    if (!ok) {
        qclhip_kernel_destroy(K);
        return fail(ctx, QCLHIP_ERR_RUNTIME, "kernel allocation failed");
    }
    *out = K;
    return QCLHIP_OK;
}

namespace {
// dst[q + Nc qp] = src[q + Nc qp] * wk[qp mod Nk] (the weights folded into the
// dense kernel's columns).
__global__ void k_fold_weights(long long Nc, int Nk, const double* __restrict__ wk, const qclhip::c64* __restrict__ src,
                               qclhip::c64* __restrict__ dst) {
    const long long n = Nc * Nc;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double w = wk[(i / Nc) % Nk];
        const qclhip::c64 v = src[i];
        dst[i] = qclhip::cmk(v.x * w, v.y * w);
    }
}
}  // namespace

int qclhip_kernel_create_dense_device(qclhip_ctx* ctx, const qclhip_grid* grid, const double* a_dev,
                                      qclhip_kernel** out) {
    if (!ctx || !grid || !out || !a_dev) return QCLHIP_ERR_ARGUMENT;
    *out = nullptr;
    if (ctx->capturing) return fail(ctx, QCLHIP_ERR_ARGUMENT, "kernel creation allocates; not during capture");
    const long long Nk = grid->Nk, Nc = Nk * (long long)grid->Nb * grid->Nb;
    auto* K = new qclhip_kernel();
    K->kind = QCLHIP_KERNEL_DENSE;
    K->grid = grid;
    K->Nk = (int)Nk;
    K->Nb = grid->Nb;
    if (hipMalloc(&K->a, (size_t)Nc * Nc * sizeof(c64)) != hipSuccess) {
        qclhip_kernel_destroy(K);
        return fail(ctx, QCLHIP_ERR_RUNTIME, "kernel allocation failed");
    }
    hipLaunchKernelGGL(k_fold_weights, dim3(grid_blocks(Nc * Nc, 256, ctx->sm_count)), dim3(256), 0, ctx->stream, Nc,
                       (int)Nk, grid->wk, (const c64*)a_dev, K->a);
    int rc = launch_check(ctx, "k_fold_weights");
    if (rc) {
        qclhip_kernel_destroy(K);
        return rc;
    }
    *out = K;
    return QCLHIP_OK;
}

// Conjugate transpose of the weighted map S = K diag(w) G: G_bar =
// (K diag(w))^dagger S_bar, in the same representation (the reverse of the
// static contraction, production_operators.jl:402-658). Built once from the
// device copy of the forward kernel.
int qclhip_kernel_adjoint(qclhip_ctx* ctx, const qclhip_kernel* K, qclhip_kernel** out) {
    if (!ctx || !K || !out) return QCLHIP_ERR_ARGUMENT;
    *out = nullptr;
    typedef std::complex<double> cd;
    const long long Nk = K->Nk, N2 = (long long)K->Nb * K->Nb, Nc = Nk * N2, r = K->rank;
    auto fetch = [&](const c64* d, size_t n, std::vector<cd>& h) -> bool {
        h.resize(n);
        return hipMemcpy(h.data(), d, n * sizeof(c64), hipMemcpyDeviceToHost) == hipSuccess;
    };
    QCLHIP_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<cd> a, b, c, da, db, dc;
    bool ok = true;
    switch (K->kind) {
        case QCLHIP_KERNEL_DENSE:
            ok = fetch(K->a, (size_t)Nc * Nc, a);
            da.resize(a.size());
            for (long long qp = 0; qp < Nc; ++qp)
                for (long long q = 0; q < Nc; ++q) da[qp + Nc * q] = std::conj(a[q + Nc * qp]);
            break;
        case QCLHIP_KERNEL_MOMENTUM_INDEPENDENT:
            ok = fetch(K->a, (size_t)N2 * N2, a);
            da.resize(a.size());
            for (long long qp = 0; qp < N2; ++qp)
                for (long long q = 0; q < N2; ++q) da[qp + N2 * q] = std::conj(a[q + N2 * qp]);
            break;
        case QCLHIP_KERNEL_SEPARABLE_IFR:
            ok = fetch(K->a, (size_t)N2 * r, a) && fetch(K->b, (size_t)r * N2, b) && fetch(K->c, (size_t)Nk * Nk, c);
            da.resize(a.size());
            db.resize(b.size());
            dc.resize(c.size());
            for (long long j = 0; j < r; ++j)
                for (long long q = 0; q < N2; ++q) {
                    da[q + N2 * j] = std::conj(b[j + r * q]);  // left' = right^dagger
                    db[j + r * q] = std::conj(a[q + N2 * j]);  // right' = left^dagger
                }
            for (long long mp = 0; mp < Nk; ++mp)
                for (long long m = 0; m < Nk; ++m) dc[mp + Nk * m] = std::conj(c[m + Nk * mp]);
            break;
        case QCLHIP_KERNEL_LOW_RANK:
            ok = fetch(K->a, (size_t)Nc * r, a) && fetch(K->b, (size_t)r * Nc, b);
            da.resize(a.size());
            db.resize(b.size());
            for (long long i = 0; i < r; ++i)
                for (long long q = 0; q < Nc; ++q) {
                    da[q + Nc * i] = std::conj(b[i + r * q]);
                    db[i + r * q] = std::conj(a[q + Nc * i]);
                }
            break;
        default:
            return fail(ctx, QCLHIP_ERR_ARGUMENT, "unknown kernel representation");
    }
    if (!ok) return fail(ctx, QCLHIP_ERR_RUNTIME, "kernel download failed");
    auto* A = new qclhip_kernel();
    A->kind = K->kind;
    A->rank = K->rank;
    A->grid = K->grid;
    A->Nk = K->Nk;
    A->Nb = K->Nb;
    A->weights_out = K->kind == QCLHIP_KERNEL_MOMENTUM_INDEPENDENT ? 1 - K->weights_out : 0;
    auto upload = [&](std::vector<cd>& h, c64** d) {
        if (h.empty() || !ok) return;
        ok = hipMalloc(d, h.size() * sizeof(c64)) == hipSuccess &&
             hipMemcpy(*d, h.data(), h.size() * sizeof(c64), hipMemcpyHostToDevice) == hipSuccess;
    };
    upload(da, &A->a);
    upload(db, &A->b);
    upload(dc, &A->c);
    if (!ok) {
        qclhip_kernel_destroy(A);
        return fail(ctx, QCLHIP_ERR_RUNTIME, "kernel allocation failed");
    }
    *out = A;
    return QCLHIP_OK;
}

int qclhip_kernel_destroy(qclhip_kernel* K) {
    if (!K) return QCLHIP_OK;
    hipDeviceSynchronize();
    if (K->a) hipFree(K->a);
    if (K->b) hipFree(K->b);
    if (K->c) hipFree(K->c);
    delete K;
    return QCLHIP_OK;
}

int qclhip_contract(qclhip_ctx* ctx, const qclhip_kernel* kernel, const double* G, double qK, double alpha,
                    double beta, double* S) {
    if (!ctx || !kernel || !G || !S) return QCLHIP_ERR_ARGUMENT;
    if (G == S) return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    return contract_core(ctx, kernel, (const c64*)G, alpha * qK, beta, (c64*)S);
}

int qclhip_contract_compact(qclhip_ctx* ctx, const qclhip_kernel* kernel, const double* G, double qK, double alpha,
                            double* Y) {
    if (!ctx || !kernel || !G || !Y) return QCLHIP_ERR_ARGUMENT;
    if (G == Y) return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    if (kernel->kind != QCLHIP_KERNEL_MOMENTUM_INDEPENDENT || kernel->weights_out)
        return fail(ctx, QCLHIP_ERR_ARGUMENT,
                    "the compact contraction takes a momentum-independent kernel in its forward form");
    return mi_contract(ctx, kernel, (const c64*)G, alpha * qK, 0.0, (c64*)Y, 1);
}

int qclhip_broadcast_momenta(qclhip_ctx* ctx, const qclhip_grid* grid, const double* Y, double* S) {
    if (!ctx || !grid || !Y || !S) return QCLHIP_ERR_ARGUMENT;
    if (Y == S) return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    const long long N2 = (long long)grid->Nb * grid->Nb;
    // S = 1.0 * Y + 0: every momentum's columns are Y bit for bit.
    hipLaunchKernelGGL(k_mi_broadcast, ecol_grid(grid->NE, (long long)grid->Nk * N2, ctx->sm_count), dim3(256), 0,
                       ctx->stream, grid->NE, grid->Nk, (int)N2, (const c64*)Y, 1.0, 0.0, (const double*)nullptr,
                       (c64*)S);
    return launch_check(ctx, "k_mi_broadcast");
}

int qclhip_lo_contraction(qclhip_ctx* ctx, const qclhip_kernel* kernel, const double* GL, const double* GG,
                          double qK, double NLO, const qclhip_shift* plus, const qclhip_shift* minus, double* SL,
                          double* SG) {
    if (!ctx || !kernel || !GL || !GG || !plus || !minus || !SL || !SG) return QCLHIP_ERR_ARGUMENT;
    const qclhip_grid* g = kernel->grid;
    if (plus->NE != g->NE || minus->NE != g->NE) return fail(ctx, QCLHIP_ERR_DIMENSION, "shift plan has the wrong energy size");
    if (kernel->kind == QCLHIP_KERNEL_LOW_RANK && kernel->rank <= kMaxRank) {
        int st = lo_lowrank(ctx, kernel, (const c64*)GL, qK, NLO + 1.0, plus, NLO, minus, (c64*)SL);
        if (st) return st;
        return lo_lowrank(ctx, kernel, (const c64*)GG, qK, NLO + 1.0, minus, NLO, plus, (c64*)SG);
    }
    const long long n = (long long)g->NE * g->Nk * g->Nb * g->Nb;
    // Y = qK K[w G] once per component, then both shifts of Y: the shifts act
    // on energy and K on (m,a,b), so S K[G] = K[S G]. Y sits behind the
    // scratch that contract_core uses, in one workspace request.
    const size_t core = (core_scratch_bytes(kernel) + 255) / 256 * 256;
    char* ws = (char*)workspace(ctx, core + (size_t)n * sizeof(c64));
    if (!ws) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    c64* Y = (c64*)(ws + core);
    int st = contract_core(ctx, kernel, (const c64*)GL, qK, 0.0, Y);
    if (st) return st;
    st = lo_combine(ctx, g->NE, n, Y, NLO + 1.0, plus, NLO, minus, (c64*)SL);
    if (st) return st;
    st = contract_core(ctx, kernel, (const c64*)GG, qK, 0.0, Y);
    if (st) return st;
    return lo_combine(ctx, g->NE, n, Y, NLO + 1.0, minus, NLO, plus, (c64*)SG);
}

}  // extern "C"
