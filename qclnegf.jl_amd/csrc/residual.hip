// Per-iteration residual suite on the device (production_residual_suite,
// residuals.jl:537-772, and production_selfenergy_residual, :802-865): the
// convergence gates of the SCBA loop without downloading fields.
//
// One wavefront per (e,m) block. Block products reuse the lane-grid product
// of the Keldysh kernels; the positivity and causality checks need Hermitian
// eigenvalues, computed by a parallel-ordered two-sided Jacobi method on the
// block in LDS: each round rotates N/2 disjoint pairs (round-robin
// "circle" schedule) at once, J = diag(1, u*) R with u = H_pq/|H_pq| and R the
// real symmetric Schur rotation (Golub-Van Loan 8.5.2), H <- J^H H J; sweeps
// stop when the off-diagonal mass is below 1e-34 of the total.
// Per-block values are reduced afterwards in a fixed order.
#include "dk_common.hpp"

namespace qclhip {

template <int B>
__device__ inline void rprod(const c64* __restrict__ L, const c64* __restrict__ R, int N, int lane,
                             c64 (&out)[B][B], bool radj) {
    constexpr int LD = 8 * B + 1;
    const int lr = lane & 7, lc = lane >> 3;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) out[r][c] = cmk(0.0, 0.0);
    for (int q = 0; q < N; ++q) {
        c64 lv[B], rv[B];
#pragma unroll
        for (int r = 0; r < B; ++r) lv[r] = L[(lr * B + r) * LD + q];
#pragma unroll
        for (int c = 0; c < B; ++c) rv[c] = radj ? cconj(R[(lc * B + c) * LD + q]) : R[q * LD + lc * B + c];
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) out[r][c] = cfma(out[r][c], lv[r], rv[c]);
    }
}

template <int B>
__device__ inline void rstore(c64* T, int N, int lane, const c64 (&v)[B][B]) {
    constexpr int LD = 8 * B + 1;
    const int lr = lane & 7, lc = lane >> 3;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int i = lr * B + r, j = lc * B + c;
            if (i < N && j < N) T[i * LD + j] = v[r][c];
        }
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += shfl_xor_d(v, off);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, shfl_xor_d(v, off));
    return v;
}
__device__ inline double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, shfl_xor_d(v, off));
    return v;
}

// Squared Frobenius norm of a tile (optionally of tile - c I).
template <int LD>
__device__ inline double tile_frob2(const c64* T, int N, int lane, double cdiag) {
    double s = 0.0;
    for (int f = lane; f < N * N; f += 64) {
        const int i = f / N, j = f % N;
        c64 v = T[i * LD + j];
        if (i == j) v.x -= cdiag;
        s += v.x * v.x + v.y * v.y;
    }
    return wave_sum(s);
}

// Hermitian part of f X into H: H = (f X + (f X)^dagger) / 2; f in {1, -i, i}
// as (fr, fi).
template <int LD>
__device__ inline void herm_part(const c64* X, c64* H, int N, int lane, double fr, double fi) {
    for (int f = lane; f < N * N; f += 64) {
        const int i = f / N, j = f % N;
        const c64 a = X[i * LD + j], b = X[j * LD + i];
        const c64 fa = cmk(fr * a.x - fi * a.y, fr * a.y + fi * a.x);
        const c64 fb = cmk(fr * b.x - fi * b.y, fr * b.y + fi * b.x);
        H[i * LD + j] = cmk(0.5 * (fa.x + fb.x), 0.5 * (fa.y - fb.y));
    }
}

struct JacobiPair {
    double c, s;
    c64 u;
    int p, q;
};

// Eigenvalues of the Hermitian tile H (destroyed); returns lambda_min,
// lambda_max, max |lambda| (wave-uniform). prs: LDS scratch of 24 pairs.
template <int LD>
__device__ inline void wave_eig_extremes(c64* H, int N, int lane, JacobiPair* prs, double& lmin, double& lmax,
                                         double& lbig) {
    const int ne = (N + 1) & ~1;  // even number of slots; slot N (odd N) is idle
    const int mm = ne - 1;        // odd
    const int npairs = ne / 2;
    for (int sweep = 0; sweep < 60 && N > 1; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int f = lane; f < N * N; f += 64) {
            const int i = f / N, j = f % N;
            const c64 v = H[i * LD + j];
            const double a = v.x * v.x + v.y * v.y;
            tot += a;
            if (i != j) off += a;
        }
        off = wave_sum(off);
        tot = wave_sum(tot);
        if (off == 0.0 || off <= 1e-34 * tot) break;
        for (int r = 0; r < mm; ++r) {
            const int ks = (int)((long long)r * ((mm + 1) / 2) % mm);  // 2 ks = r (mod mm)
            if (lane < npairs) {
                int p, q;
                if (lane == 0) {
                    p = ks;
                    q = mm;
                } else {
                    p = (ks + lane) % mm;
                    q = (ks - lane + mm) % mm;
                }
                JacobiPair jp;
                jp.p = p;
                jp.q = q;
                jp.c = 1.0;
                jp.s = 0.0;
                jp.u = cmk(1.0, 0.0);
                if (p < N && q < N) {
                    const c64 b = H[p * LD + q];
                    const double ab = sqrt(b.x * b.x + b.y * b.y);
                    if (ab > 0.0) {
                        const double a = H[p * LD + p].x, d = H[q * LD + q].x;
                        const double tau = (d - a) / (2.0 * ab);
                        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        jp.c = 1.0 / sqrt(1.0 + t * t);
                        jp.s = t * jp.c;
                        jp.u = cmk(b.x / ab, b.y / ab);
                    }
                }
                prs[lane] = jp;
            }
            wave_sync();
            // columns: H <- H J  (col p = c col_p - s u* col_q, col q = s col_p + c u* col_q)
            for (int t = lane; t < npairs * N; t += 64) {
                const JacobiPair jp = prs[t / N];
                const int i = t % N;
                if (jp.q >= N || jp.p >= N || jp.s == 0.0) continue;
                const c64 hp = H[i * LD + jp.p], hq = H[i * LD + jp.q];
                const c64 uq = cmul(cconj(jp.u), hq);
                H[i * LD + jp.p] = cmk(jp.c * hp.x - jp.s * uq.x, jp.c * hp.y - jp.s * uq.y);
                H[i * LD + jp.q] = cmk(jp.s * hp.x + jp.c * uq.x, jp.s * hp.y + jp.c * uq.y);
            }
            wave_sync();
            // rows: H <- J^H H  (row p = c row_p - s u row_q, row q = s row_p + c u row_q)
            for (int t = lane; t < npairs * N; t += 64) {
                const JacobiPair jp = prs[t / N];
                const int j = t % N;
                if (jp.q >= N || jp.p >= N || jp.s == 0.0) continue;
                const c64 hp = H[jp.p * LD + j], hq = H[jp.q * LD + j];
                const c64 uq = cmul(jp.u, hq);
                c64 np = cmk(jp.c * hp.x - jp.s * uq.x, jp.c * hp.y - jp.s * uq.y);
                c64 nq = cmk(jp.s * hp.x + jp.c * uq.x, jp.s * hp.y + jp.c * uq.y);
                // The rotated pair is diagonal with a real diagonal in exact
                // arithmetic: store it so (rows p and q belong to this pair).
                if (j == jp.p) {
                    np.y = 0.0;
                    nq = cmk(0.0, 0.0);
                } else if (j == jp.q) {
                    np = cmk(0.0, 0.0);
                    nq.y = 0.0;
                }
                H[jp.p * LD + j] = np;
                H[jp.q * LD + j] = nq;
            }
            wave_sync();
        }
    }
    double mn = INFINITY, mx = -INFINITY, bg = 0.0, bad = 0.0;
    for (int i = lane; i < N; i += 64) {
        const double v = H[i * LD + i].x;
        if (!isfinite(v)) bad = 1.0;
        mn = fmin(mn, v);
        mx = fmax(mx, v);
        bg = fmax(bg, fabs(v));
    }
    lmin = wave_min(mn);
    lmax = wave_max(mx);
    lbig = wave_max(bg);
    if (wave_max(bad) != 0.0) lmin = lmax = lbig = NAN;
    wave_sync();  // the caller may overwrite H next
}

// 1 when every element of the N x N tile is finite (wave-uniform).
template <int LD>
__device__ inline bool tile_finite(const c64* T, int N, int lane) {
    double bad = 0.0;
    for (int f = lane; f < N * N; f += 64) {
        const c64 v = T[(f / N) * LD + f % N];
        if (!(isfinite(v.x) && isfinite(v.y))) bad = 1.0;
    }
    return wave_max(bad) == 0.0;
}

// The positivity ratio of the reference's witness (positivity_diagnostics.jl:
// 61-96): max(0, -lambda_min - 64 n eps max(|lambda|_max, 1)) /
// max(|lambda|_max, floatmin).
__device__ inline double psd_ratio(double lmin, double lbig, int N) {
    if (!(isfinite(lmin) && isfinite(lbig))) return INFINITY;
    const double budget = 64.0 * N * 2.220446049250313e-16 * fmax(lbig, 1.0);
    return fmax(0.0, fmax(0.0, -lmin) - budget) / fmax(lbig, 2.2250738585072014e-308);
}

// Positivity ratio of the Hermitian part of f X (X intact, E overwritten);
// Inf for a non-finite block (residuals.jl:413-452).
template <int LD>
__device__ inline double psd_term(const c64* X, c64* E, int N, int lane, JacobiPair* prs, double fr, double fi) {
    if (!tile_finite<LD>(X, N, lane)) return INFINITY;
    herm_part<LD>(X, E, N, lane, fr, fi);
    wave_sync();
    double lmin, lmax, lbig;
    wave_eig_extremes<LD>(E, N, lane, prs, lmin, lmax, lbig);
    return psd_ratio(lmin, lbig, N);
}

struct ResArgs {
    int NE, Nk, N;
    const double* eps;
    const c64* h;
    const c64* GR;
    const c64* GL;
    const c64* GG;
    const c64* A;
    const c64* famR[kMaxFamilies];
    const c64* famL[kMaxFamilies];
    const c64* famG[kMaxFamilies];
    const c64* cand[kMaxFamilies];
    int nfam, ncand;
    double* out;  // [7][NE*Nk]: rD, specN, specD, kelN, kelD, psd, caus
};

template <int B>
__global__ __launch_bounds__(64) void k_residual_suite(ResArgs p) {
    constexpr int LD = 8 * B + 1;
    constexpr int W = 1;
    const int wstride = TileStride<B>::value;
    extern __shared__ __align__(16) unsigned char smem[];
    c64* G = reinterpret_cast<c64*>(smem);
    c64* X = G + wstride;
    c64* Y = X + wstride;
    c64* E = Y + wstride;
    JacobiPair* prs = reinterpret_cast<JacobiPair*>(E + wstride);
    const int lane = threadIdx.x;
    const long long nblk = (long long)p.NE * p.Nk;
    const long long blk = blockIdx.x;  // e + NE m
    const int e = (int)(blk % p.NE), m = (int)(blk / p.NE);
    const int N = p.N;
    const size_t S = (size_t)nblk;
    c64 T[B][B];
    double lmin, lmax, lbig;
    const c64* one[1];
    one[0] = p.GR;
    tile_load<W, LD>(G, wstride, one, 1, blk, 1, N, S);
    tile_load<W, LD>(X, wstride, p.famR, p.nfam, blk, 1, N, S);
    wave_sync();
    // Causality of the total S^R (residuals.jl:455-492).
    double caus = 0.0;
    const bool sr_finite = tile_finite<LD>(X, N, lane);
    for (int f = lane; f < N * N; f += 64) {
        const int i = f / N, j = f % N;
        const c64 a = X[i * LD + j], b = X[j * LD + i];
        E[i * LD + j] = cmk(0.5 * (a.y + b.y), -0.5 * (a.x - b.x));  // (S - S^dagger) / (2i)
    }
    wave_sync();
    wave_eig_extremes<LD>(E, N, lane, prs, lmin, lmax, lbig);
    if (!sr_finite || !isfinite(lmax)) {
        caus = INFINITY;
    } else if (lmax > 0.0) {
        // ||S^R||_2 = sqrt(lambda_max(S^dagger S))
        for (int f = lane; f < N * N; f += 64) {
            const int i = f / N, j = f % N;
            c64 v = cmk(0.0, 0.0);
            for (int c = 0; c < N; ++c) v = cfma(v, cconj(X[c * LD + i]), X[c * LD + j]);
            E[i * LD + j] = v;
        }
        wave_sync();
        double a0, a1, a2;
        wave_eig_extremes<LD>(E, N, lane, prs, a0, a1, a2);
        caus = lmax / (sqrt(fmax(0.0, a1)) + 1e-14);
    }
    // r_D: D = eps - h - S^R in X, ||D G - I|| / (||D|| ||G|| + sqrt(N)).
    for (int f = lane; f < N * N; f += 64) {
        const int i = f / N, j = f % N;
        const c64 hv = p.h[(size_t)m + (size_t)p.Nk * (i + (size_t)N * j)];
        const c64 sv = X[i * LD + j];
        X[i * LD + j] = cmk((i == j ? p.eps[e] : 0.0) - hv.x - sv.x, -hv.y - sv.y);
    }
    wave_sync();
    rprod<B>(X, G, N, lane, T, false);
    wave_sync();
    rstore<B>(Y, N, lane, T);
    wave_sync();
    const double rD = sqrt(tile_frob2<LD>(Y, N, lane, 1.0)) /
                      (sqrt(tile_frob2<LD>(X, N, lane, 0.0)) * sqrt(tile_frob2<LD>(G, N, lane, 0.0)) + sqrt((double)N));
    double psd = 0.0;
    // Gamma = i (S> - S<) in X.
    for (int f = lane; f < N * N; f += 64) {
        const int i = f / N, j = f % N;
        const size_t idx = blk + S * (i + (size_t)N * j);
        const c64 g = fam_load(p.famG, p.nfam, idx), l = fam_load(p.famL, p.nfam, idx);
        X[i * LD + j] = cmk(-(g.y - l.y), g.x - l.x);
    }
    wave_sync();
    // r_A terms: ||A - G Gamma G^dagger||^2, ||A||^2.
    rprod<B>(G, X, N, lane, T, false);
    wave_sync();
    rstore<B>(Y, N, lane, T);
    wave_sync();
    rprod<B>(Y, G, N, lane, T, true);
    wave_sync();
    rstore<B>(E, N, lane, T);
    one[0] = p.A;
    tile_load<W, LD>(Y, wstride, one, 1, blk, 1, N, S);
    wave_sync();
    double sn = 0.0;
    for (int f = lane; f < N * N; f += 64) {
        const int i = f / N, j = f % N;
        const c64 d = csub(Y[i * LD + j], E[i * LD + j]);
        sn += d.x * d.x + d.y * d.y;
    }
    const double specN = wave_sum(sn);
    const double specD = tile_frob2<LD>(Y, N, lane, 0.0);
    wave_sync();
    // Positivity of Gamma and of A.
    psd = fmax(psd, psd_term<LD>(X, E, N, lane, prs, 1.0, 0.0));
    psd = fmax(psd, psd_term<LD>(Y, E, N, lane, prs, 1.0, 0.0));
    // r_K terms against the unmixed candidate S<.
    tile_load<W, LD>(X, wstride, p.cand, p.ncand, blk, 1, N, S);
    wave_sync();
    rprod<B>(G, X, N, lane, T, false);
    wave_sync();
    rstore<B>(Y, N, lane, T);
    wave_sync();
    rprod<B>(Y, G, N, lane, T, true);
    wave_sync();
    rstore<B>(E, N, lane, T);
    one[0] = p.GL;
    tile_load<W, LD>(Y, wstride, one, 1, blk, 1, N, S);
    wave_sync();
    sn = 0.0;
    for (int f = lane; f < N * N; f += 64) {
        const int i = f / N, j = f % N;
        const c64 d = csub(Y[i * LD + j], E[i * LD + j]);
        sn += d.x * d.x + d.y * d.y;
    }
    const double kelN = wave_sum(sn);
    const double kelD = tile_frob2<LD>(Y, N, lane, 0.0);
    wave_sync();
    psd = fmax(psd, psd_term<LD>(Y, E, N, lane, prs, 0.0, -1.0));  // -i G<
    one[0] = p.GG;
    tile_load<W, LD>(Y, wstride, one, 1, blk, 1, N, S);
    wave_sync();
    psd = fmax(psd, psd_term<LD>(Y, E, N, lane, prs, 0.0, 1.0));  // i G>
    // Raw Keldysh products before the paired map: -i G S< G^dagger, i G S> G^dagger.
    for (int comp = 0; comp < 2; ++comp) {
        tile_load<W, LD>(X, wstride, comp ? p.famG : p.famL, p.nfam, blk, 1, N, S);
        wave_sync();
        rprod<B>(G, X, N, lane, T, false);
        wave_sync();
        rstore<B>(Y, N, lane, T);
        wave_sync();
        rprod<B>(Y, G, N, lane, T, true);
        wave_sync();
        rstore<B>(X, N, lane, T);
        wave_sync();
        psd = fmax(psd, psd_term<LD>(X, E, N, lane, prs, 0.0, comp ? 1.0 : -1.0));
    }
    if (lane == 0) {
        p.out[0 * nblk + blk] = rD;
        p.out[1 * nblk + blk] = specN;
        p.out[2 * nblk + blk] = specD;
        p.out[3 * nblk + blk] = kelN;
        p.out[4 * nblk + blk] = kelD;
        p.out[5 * nblk + blk] = psd;
        p.out[6 * nblk + blk] = caus;
    }
}

// Fixed-order reduction of the per-block values: sums for rows 1-4, maxima
// for rows 0, 5, 6; a non-finite value makes its residual Inf.
__global__ __launch_bounds__(1024) void k_residual_reduce(long long nblk, const double* v, double* out) {
    __shared__ double red[7][1024];
    const int t = threadIdx.x;
    const long long per = (nblk + 1023) / 1024;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long long i = t * per; i < min(nblk, (t + 1) * per); ++i)
        for (int k = 0; k < 7; ++k) {
            const double x = v[k * nblk + i];
            const bool is_sum = k >= 1 && k <= 4;
            if (!isfinite(x))
                acc[k] = INFINITY;
            else
                acc[k] = is_sum ? acc[k] + x : fmax(acc[k], x);
        }
    for (int k = 0; k < 7; ++k) red[k][t] = acc[k];
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 7; ++k) {
                const bool is_sum = k >= 1 && k <= 4;
                red[k][t] = is_sum ? red[k][t] + red[k][t + w] : fmax(red[k][t], red[k][t + w]);
            }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = red[0][0];
        out[1] = sqrt(red[1][0]) / (sqrt(red[2][0]) + 1e-14);
        out[2] = sqrt(red[3][0]) / (sqrt(red[4][0]) + 1e-14);
        out[3] = red[5][0];
        out[4] = red[6][0];
        for (int k = 1; k <= 2; ++k)
            if (!isfinite(out[k])) out[k] = INFINITY;
    }
}

// Self-energy residual: per component, partial sums of |cand - old|^2 and
// |cand|^2 on a fixed grid, then a fixed tree.
constexpr int kSigBlocks = 1024;

__global__ __launch_bounds__(256) void k_sigma_partial(long long n, const c64* o0, const c64* o1, const c64* o2,
                                                       const c64* c0, const c64* c1, const c64* c2, double* part) {
    __shared__ double red[6][256];
    const c64* os[3] = {o0, o1, o2};
    const c64* cs[3] = {c0, c1, c2};
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)kSigBlocks * 256)
        for (int k = 0; k < 3; ++k) {
            const c64 c = cs[k][i], o = os[k][i];
            const double dx = c.x - o.x, dy = c.y - o.y;
            acc[2 * k] += dx * dx + dy * dy;
            acc[2 * k + 1] += c.x * c.x + c.y * c.y;
        }
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[k * kSigBlocks + blockIdx.x] = red[k][0];
}

__global__ __launch_bounds__(kSigBlocks) void k_sigma_reduce(const double* part, double* out) {
    __shared__ double red[6][kSigBlocks];
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = part[k * kSigBlocks + threadIdx.x];
    __syncthreads();
    for (int w = kSigBlocks / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double d = red[2 * k][0], c = red[2 * k + 1][0];
            const double v = (isfinite(d) && isfinite(c)) ? sqrt(d) / (sqrt(c) + 1e-14) : INFINITY;
            r = fmax(r, v);
        }
        *out = r;
    }
}

template <int B>
int launch_residual(qclhip_ctx* ctx, const ResArgs& a) {
    const size_t lds = 4 * (size_t)TileStride<B>::value * sizeof(c64) + 24 * sizeof(JacobiPair);
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_residual_suite<B>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_residual_suite<B>), dim3((unsigned)((long long)a.NE * a.Nk)), dim3(64), lds, ctx->stream, a);
    return launch_check(ctx, "k_residual_suite");
}

}  // namespace qclhip

using namespace qclhip;

extern "C" {

int qclhip_residual_suite(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, const double* GR,
                          const double* GL, const double* GG, const double* A, int nfam,
                          const double* const* famR, const double* const* famL, const double* const* famG,
                          int ncand, const double* const* candL, double* out_host) {
    if (!ctx || !grid || !h || !GR || !GL || !GG || !A || !famR || !famL || !famG || !candL || !out_host)
        return QCLHIP_ERR_ARGUMENT;
    if (nfam < 1 || nfam > kMaxFamilies || ncand < 1 || ncand > kMaxFamilies)
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "1..8 families and candidate families");
    const long long nblk = (long long)grid->NE * grid->Nk;
    double* v = (double*)workspace(ctx, (size_t)(7 * nblk + 8) * sizeof(double));
    if (!v) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    ResArgs a{};
    a.NE = grid->NE;
    a.Nk = grid->Nk;
    a.N = grid->Nb;
    a.eps = grid->eps;
    a.h = (const c64*)h;
    a.GR = (const c64*)GR;
    a.GL = (const c64*)GL;
    a.GG = (const c64*)GG;
    a.A = (const c64*)A;
    for (int f = 0; f < nfam; ++f) {
        a.famR[f] = (const c64*)famR[f];
        a.famL[f] = (const c64*)famL[f];
        a.famG[f] = (const c64*)famG[f];
    }
    for (int f = 0; f < ncand; ++f) a.cand[f] = (const c64*)candL[f];
    a.nfam = nfam;
    a.ncand = ncand;
    a.out = v;
    int rc;
    const int N = grid->Nb;
    if (N <= 8)
        rc = launch_residual<1>(ctx, a);
    else if (N <= 16)
        rc = launch_residual<2>(ctx, a);
    else if (N <= 24)
        rc = launch_residual<3>(ctx, a);
    else if (N <= 32)
        rc = launch_residual<4>(ctx, a);
    else if (N <= 40)
        rc = launch_residual<5>(ctx, a);
    else if (N <= 48)
        rc = launch_residual<6>(ctx, a);
    else
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");
    if (rc) return rc;
    double* res = v + 7 * nblk;
    hipLaunchKernelGGL(k_residual_reduce, dim3(1), dim3(1024), 0, ctx->stream, nblk, v, res);
    rc = launch_check(ctx, "k_residual_reduce");
    if (rc) return rc;
    bool deferred = false;
    rc = defer_result(ctx, res, 5, &deferred);
    if (rc || deferred) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(out_host, res, 5 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return qclhip_ctx_synchronize(ctx);
}

int qclhip_selfenergy_residual(qclhip_ctx* ctx, int64_t n, const double* const* old_, const double* const* cand,
                               double* out_host) {
    if (!ctx || !old_ || !cand || !out_host || n < 0) return QCLHIP_ERR_ARGUMENT;
    for (int k = 0; k < 3; ++k)
        if (!old_[k] || !cand[k]) return QCLHIP_ERR_ARGUMENT;
    double* part = (double*)workspace(ctx, (size_t)(6 * kSigBlocks + 8) * sizeof(double));
    if (!part) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    double* res = part + 6 * kSigBlocks;
    hipLaunchKernelGGL(k_sigma_partial, dim3(kSigBlocks), dim3(256), 0, ctx->stream, (long long)n,
                       (const c64*)old_[0], (const c64*)old_[1], (const c64*)old_[2], (const c64*)cand[0],
                       (const c64*)cand[1], (const c64*)cand[2], part);
    hipLaunchKernelGGL(k_sigma_reduce, dim3(1), dim3(kSigBlocks), 0, ctx->stream, part, res);
    int rc = launch_check(ctx, "selfenergy residual");
    if (rc) return rc;
    bool deferred = false;
    rc = defer_result(ctx, res, 1, &deferred);
    if (rc || deferred) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(out_host, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return qclhip_ctx_synchronize(ctx);
}

}  // extern "C"
