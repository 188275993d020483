// Wave-level dense linear algebra on (N x N) complex blocks for the
// residual suite and the Anderson guard: the lane-grid product (8 x 8 lanes,
// B x B register blocks), Hermitian parts, a Cholesky screen, and extreme
// Hermitian eigenvalues by Householder tridiagonalization in registers and
// Sturm multisection. One wavefront per block; LDS vectors per wave.
#pragma once

#include "dk_common.hpp"

namespace qclhip {

// out (lane block) = L R or L R^dagger (radj) of two LDS tiles, with the
// per-element arithmetic of block_product (the Keldysh kernels' products:
// FMAs in q order, cfmac for the adjoint), so a sandwich G X G^dagger formed
// here equals the Keldysh kernel's bit for bit.
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
        for (int c = 0; c < B; ++c) rv[c] = radj ? R[(lc * B + c) * LD + q] : R[q * LD + lc * B + c];
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c)
                out[r][c] = radj ? cfmac(out[r][c], lv[r], rv[c]) : cfma(out[r][c], lv[r], rv[c]);
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

// Wave reductions: DPP within each row of 16 lanes (xor 1, xor 2, then the
// half-row and row mirrors, after which every lane of a row holds the row's
// value), then the four rows read back in a fixed order. Every lane gets
// the same result; no LDS traffic.
template <typename Op>
__device__ inline double wave_reduce(double v, Op op) {
    v = op(v, dpp_d<0xB1>(v));   // quad_perm [1,0,3,2]
    v = op(v, dpp_d<0x4E>(v));   // quad_perm [2,3,0,1]
    v = op(v, dpp_d<0x141>(v));  // row_half_mirror
    v = op(v, dpp_d<0x140>(v));  // row_mirror
    return op(op(readlane_d(v, 0), readlane_d(v, 16)), op(readlane_d(v, 32), readlane_d(v, 48)));
}
__device__ inline double wave_sum(double v) {
    return wave_reduce(v, [](double a, double b) { return a + b; });
}
__device__ inline double wave_max(double v) {
    return wave_reduce(v, [](double a, double b) { return fmax(a, b); });
}
__device__ inline double wave_min(double v) {
    return wave_reduce(v, [](double a, double b) { return fmin(a, b); });
}

// Squared Frobenius norm of a tile (optionally of tile - c I); each lane
// sums the positions of its register block on the 8 x 8 lane grid.
template <int LD>
__device__ inline double tile_frob2(const c64* T, int N, int lane, double cdiag) {
    static_assert((LD - 1) % 8 == 0, "tiles of the lane grid have LD = 8B + 1");
    constexpr int B = (LD - 1) / 8;
    const int lr = lane & 7, lc = lane >> 3;
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int i = lr * B + r, j = lc * B + c;
            if (i < N && j < N) {
                c64 v = T[i * LD + j];
                if (i == j) v.x -= cdiag;
                s += v.x * v.x + v.y * v.y;
            }
        }
    return wave_sum(s);
}

// LDS vectors of one wave's tridiagonal reduction and Cholesky screen, sized
// for the tile class (N <= 8B).
template <int B>
struct EigWork {
    c64 u[8 * B];
    c64 q[8 * B];
    double d[8 * B];
    double e2[8 * B];
};

// Sums over the 8 lanes of one row group (lane bits 3..5) and of one column
// group (lane bits 0..2) of the 8 x 8 lane grid.
__device__ inline double sum_over_lc(double v) {
    v += shfl_xor_d(v, 8);
    v += shfl_xor_d(v, 16);
    return v + shfl_xor_d(v, 32);
}
__device__ inline double sum_over_lr(double v) {
    v += shfl_xor_d(v, 1);
    v += shfl_xor_d(v, 2);
    return v + shfl_xor_d(v, 4);
}

// Hermitian part of f X, f = fr + i fi, from an LDS tile into the lane-grid
// registers (rows lr*B + r, columns lc*B + c; zero outside N x N).
template <int B, int LD>
__device__ inline void load_herm(const c64* X, int N, int lane, double fr, double fi, c64 (&a)[B][B]) {
    const int lr = lane & 7, lc = lane >> 3;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int i = lr * B + r, j = lc * B + c;
            if (i < N && j < N) {
                const c64 x = X[i * LD + j], y = X[j * LD + i];
                const double ax = fr * x.x - fi * x.y, ay = fr * x.y + fi * x.x;
                const double bx = fr * y.x - fi * y.y, by = fr * y.y + fi * y.x;
                a[r][c] = cmk(0.5 * (ax + bx), 0.5 * (ay - by));
            } else {
                a[r][c] = cmk(0.0, 0.0);
            }
        }
}

struct BlockScale {
    bool bad;     // a non-finite entry
    double amax;  // largest |Re|, |Im|
    double fro2;  // squared Frobenius norm
};

// Wave-uniform scale figures of the N x N part of a register block.
template <int B>
__device__ inline BlockScale block_scale(const c64 (&a)[B][B], int N, int lane) {
    const int lr = lane & 7, lc = lane >> 3;
    double amax = 0.0, fro2 = 0.0, bad = 0.0;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c)
            if (lr * B + r < N && lc * B + c < N) {
                const c64 v = a[r][c];
                if (!(isfinite(v.x) && isfinite(v.y))) bad = 1.0;
                amax = fmax(amax, fmax(fabs(v.x), fabs(v.y)));
                fro2 += v.x * v.x + v.y * v.y;
            }
    BlockScale s;
    s.bad = wave_max(bad) != 0.0;
    s.amax = wave_max(amax);
    s.fro2 = wave_sum(fro2);
    return s;
}

template <int B>
__device__ inline void scale_regs(c64 (&a)[B][B], double s) {
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) a[r][c] = cmk(a[r][c].x * s, a[r][c].y * s);
}

// True when the Hermitian register block a + shift I (N x N part) has a
// Cholesky factorization a + shift I = L L^H, i.e. every pivot of the
// right-looking elimination is positive; a is destroyed. The column of each
// step is passed through LDS.
template <int B>
__device__ inline bool chol_screen(c64 (&a)[B][B], int N, int lane, EigWork<B>* w, double shift) {
    const int lr = lane & 7, lc = lane >> 3;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c)
            if (lr * B + r == lc * B + c) a[r][c].x += shift;
    // Steps k = ck*B + kk with kk unrolled, so the pivot column is a fixed
    // register column of the lanes lc == ck; row and column masks are 0/1
    // weights. Padding rows and columns (index >= N) hold zeros off the
    // diagonal and never change.
    c64* u = w->u;
    const int nck = (N + B - 1) / B;
    for (int ck = 0; ck < nck; ++ck) {
        const bool is_c = lc == ck;
#pragma unroll
        for (int kk = 0; kk < B; ++kk) {
            const int k = ck * B + kk;
            if (k >= N) break;
            if (is_c) {
#pragma unroll
                for (int r = 0; r < B; ++r) u[lr * B + r] = a[r][kk];
            }
            wave_sync();
            const double piv = u[k].x;
            if (!(piv > 0.0)) {
                wave_sync();
                return false;
            }
            const double rs = rsqrt_nr(piv);
            c64 li[B], lj[B];
#pragma unroll
            for (int r = 0; r < B; ++r) {
                const double wi = lr * B + r > k ? rs : 0.0;
                const double wj = lc * B + r > k ? rs : 0.0;
                const c64 ui = u[lr * B + r], uj = u[lc * B + r];
                li[r] = cmk(wi * ui.x, wi * ui.y);
                lj[r] = cmk(wj * uj.x, -(wj * uj.y));
            }
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int c = 0; c < B; ++c) a[r][c] = cfms(a[r][c], li[r], lj[c]);
            wave_sync();  // the next step rewrites u
        }
    }
    return true;
}

// Number of eigenvalues below x of the real symmetric tridiagonal matrix
// with diagonal d and squared off-diagonal e2: the negative pivots of
// LDL^T(T - x I), a zero pivot replaced by -pivmin (Kahan's count, as in
// LAPACK's bisection).
__device__ inline int sturm_count(const double* d, const double* e2, int N, double x, double pivmin) {
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    int cnt = q < 0.0;
    for (int i = 1; i < N; ++i) {
        q = (d[i] - x) - e2[i - 1] / q;
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += q < 0.0;
    }
    return cnt;
}

// Narrow [a, b] to the sub-interval that holds the transition of a
// monotone predicate sampled at x_j = a + (b - a)(j + 1)/33, j = 0..31
// (bit j of mask set where the predicate holds).
__device__ inline void multisect(double& a, double& b, unsigned mask) {
    const double w = b - a;
    if (mask == 0u) {
        a = a + w * (32 / 33.0);
        return;
    }
    const int f = __builtin_ctz(mask);
    const double nb = a + w * ((f + 1) / 33.0);
    if (f > 0) a = a + w * (f / 33.0);
    b = nb;
}

// Smallest and largest eigenvalue of the finite Hermitian register block a
// (N x N part; a is destroyed; entries outside are ignored), wave-uniform.
// a is reduced to real symmetric tridiagonal form by N-2 Householder
// reflections applied in registers on the 8 x 8 lane grid
// (P = I - beta u u^H with u = x + e^{i arg x_0} ||x|| e_1, A <- P A P
// through p = beta A u, q = p - (beta/2)(u^H p) u, A <- A - u q^H - q u^H),
// and the two extreme eigenvalues are located by 32-point Sturm
// multisection each (half-waves), to eps * ||T|| absolute. Callers scale a
// to max |a_ij| ~ 1 first.
template <int B>
__device__ inline void eig_core(c64 (&a)[B][B], int N, int lane, EigWork<B>* w, double& lmin, double& lmax) {
    constexpr double kEps = 2.220446049250313e-16, kTiny = 2.2250738585072014e-308;
    const int lr = lane & 7, lc = lane >> 3;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c)
            if (lr * B + r >= N || lc * B + c >= N) a[r][c] = cmk(0.0, 0.0);
    for (int k = 0; k + 2 < N; ++k) {
        const int kg = k / B, kc = k - kg * B;  // column k: lanes lc == kg, register column kc
        double part = 0.0;
        if (lc == kg) {
#pragma unroll
            for (int c = 0; c < B; ++c) {
                if (c != kc) continue;
#pragma unroll
                for (int r = 0; r < B; ++r) {
                    const int i = lr * B + r;
                    if (i > k && i < N) {
                        w->u[i] = a[r][c];
                        part += a[r][c].x * a[r][c].x + a[r][c].y * a[r][c].y;
                    }
                    if (i == k) w->d[k] = a[r][c].x;
                }
            }
        }
        const double nrm2 = wave_sum(part);
        wave_sync();
        const c64 x0 = w->u[k + 1];
        if (lane == 0) w->e2[k] = nrm2;
        if (nrm2 == 0.0) {  // column already reduced
            wave_sync();
            continue;
        }
        const double ax0 = sqrt(x0.x * x0.x + x0.y * x0.y), xn = sqrt(nrm2);
        const c64 ph = ax0 > 0.0 ? cmk(x0.x / ax0, x0.y / ax0) : cmk(1.0, 0.0);
        const c64 u1 = cmk(x0.x + ph.x * xn, x0.y + ph.y * xn);
        const double beta = 1.0 / (nrm2 + xn * ax0);
        c64 ur[B], uc[B], p[B];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            const int i = lr * B + r;
            ur[r] = i == k + 1 ? u1 : (i > k + 1 && i < N ? w->u[i] : cmk(0.0, 0.0));
            const int j = lc * B + r;
            uc[r] = j == k + 1 ? u1 : (j > k + 1 && j < N ? w->u[j] : cmk(0.0, 0.0));
        }
#pragma unroll
        for (int r = 0; r < B; ++r) {
            p[r] = cmk(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < B; ++c) p[r] = cfma(p[r], a[r][c], uc[c]);
        }
        double kd = 0.0;
#pragma unroll
        for (int r = 0; r < B; ++r) {
            const int i = lr * B + r;
            p[r] = cmk(beta * sum_over_lc(p[r].x), beta * sum_over_lc(p[r].y));
            if (i <= k || i >= N) p[r] = cmk(0.0, 0.0);
            kd += ur[r].x * p[r].x + ur[r].y * p[r].y;  // Re(conj(u_i) p_i)
        }
        const double K = 0.5 * beta * sum_over_lr(kd);
#pragma unroll
        for (int r = 0; r < B; ++r) p[r] = cmk(p[r].x - K * ur[r].x, p[r].y - K * ur[r].y);  // q rows
        if (lc == 0) {
#pragma unroll
            for (int r = 0; r < B; ++r)
                if (lr * B + r < N) w->q[lr * B + r] = p[r];
        }
        wave_sync();
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int j = lc * B + c;
            const c64 qc = (j > k && j < N) ? w->q[j] : cmk(0.0, 0.0);
#pragma unroll
            for (int r = 0; r < B; ++r) {
                const int i = lr * B + r;
                if (i > k && i < N && j > k && j < N) {
                    // a -= u_i conj(q_j) + q_i conj(u_j)
                    a[r][c] = cfms(a[r][c], ur[r], cconj(qc));
                    a[r][c] = cfms(a[r][c], p[r], cconj(uc[c]));
                }
            }
        }
        wave_sync();  // the next step rewrites u and q
    }
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int i = lr * B + r, j = lc * B + c;
            if (i == j && i < N && i >= N - 2) w->d[i] = a[r][c].x;
            if (i == N - 1 && j == N - 2) w->e2[N - 2] = a[r][c].x * a[r][c].x + a[r][c].y * a[r][c].y;
        }
    wave_sync();
    double e2max = 0.0, glo = INFINITY, ghi = -INFINITY;
    for (int i = lane; i < N; i += 64) {
        const double ri = (i > 0 ? sqrt(w->e2[i - 1]) : 0.0) + (i + 1 < N ? sqrt(w->e2[i]) : 0.0);
        glo = fmin(glo, w->d[i] - ri);
        ghi = fmax(ghi, w->d[i] + ri);
        if (i + 1 < N) e2max = fmax(e2max, w->e2[i]);
    }
    glo = wave_min(glo);
    ghi = wave_max(ghi);
    e2max = wave_max(e2max);
    const double pivmin = kTiny * fmax(1.0, e2max);
    const double tnorm = fmax(fabs(glo), fabs(ghi));
    const double widen = 2.0 * kEps * tnorm * N + 2.0 * pivmin;
    double a0 = glo - widen, b0 = ghi + widen, a1 = a0, b1 = b0;
    const int half = lane >> 5, jx = lane & 31;
    for (int round = 0; round < 24; ++round) {
        const double t0 = fmax(2.0 * kEps * fmax(fabs(a0), fabs(b0)), kEps * tnorm);
        const double t1 = fmax(2.0 * kEps * fmax(fabs(a1), fabs(b1)), kEps * tnorm);
        if (b0 - a0 <= t0 && b1 - a1 <= t1) break;
        const double lo = half ? a1 : a0, hi = half ? b1 : b0;
        const double x = lo + (hi - lo) * ((jx + 1) / 33.0);
        const int cnt = sturm_count(w->d, w->e2, N, x, pivmin);
        const unsigned long long m = __ballot(half ? cnt >= N : cnt >= 1);
        multisect(a0, b0, (unsigned)(m & 0xffffffffull));
        multisect(a1, b1, (unsigned)(m >> 32));
    }
    lmin = 0.5 * (a0 + b0);
    lmax = 0.5 * (a1 + b1);
    wave_sync();  // d and e2 are rewritten by the next call
}

// The positivity ratio of the reference's witness (positivity_diagnostics.jl:
// 61-96): max(0, -lambda_min - 64 n eps max(|lambda|_max, 1)) /
// max(|lambda|_max, floatmin).
__device__ inline double psd_ratio(double lmin, double lbig, int N) {
    if (!(isfinite(lmin) && isfinite(lbig))) return INFINITY;
    const double budget = 64.0 * N * 2.220446049250313e-16 * fmax(lbig, 1.0);
    return fmax(0.0, fmax(0.0, -lmin) - budget) / fmax(lbig, 2.2250738585072014e-308);
}

// Positivity ratio of the Hermitian part H of f X (tile X intact); Inf for
// a non-finite block (residuals.jl:413-452). Screen first: the ratio is 0
// whenever lambda_min(H) >= -budget, and max |lambda| >= ||H||_F / sqrt(N)
// bounds the budget from below, so a Cholesky factorization of H + delta I
// with delta = 32 N eps max(||H||_F / sqrt(N), 1) (half that bound) settles
// the block; its backward error (<= N(N+1) eps ||H||) stays within the other
// half of the budget. Otherwise the extreme eigenvalues are computed.
template <int B, int LD>
__device__ inline double psd_from_tile(const c64* X, int N, int lane, EigWork<B>* w, double fr, double fi) {
    constexpr double kEps = 2.220446049250313e-16;
    c64 a[B][B];
    load_herm<B, LD>(X, N, lane, fr, fi, a);
    const BlockScale s = block_scale<B>(a, N, lane);
    if (s.bad) return INFINITY;
    if (s.amax == 0.0) return 0.0;
    int ex;
    frexp(s.amax, &ex);
    const double sc = ldexp(1.0, -ex);
    scale_regs<B>(a, sc);
    const double delta = 32.0 * N * kEps * fmax(sqrt(s.fro2 / N), 1.0);
    if (chol_screen<B>(a, N, lane, w, delta * sc)) return 0.0;
    load_herm<B, LD>(X, N, lane, fr, fi, a);
    scale_regs<B>(a, sc);
    double lmin, lmax;
    eig_core<B>(a, N, lane, w, lmin, lmax);
    lmin = ldexp(lmin, ex);
    lmax = ldexp(lmax, ex);
    return psd_ratio(lmin, fmax(fabs(lmin), fabs(lmax)), N);
}

// The screen on the cyclic lane grid: lane (lr, lc) holds rows lr + 8r and
// columns lc + 8c (r, c < B). Column k = 8 kc + kl lives in register column
// kc of the lanes lc == kl, and once k >= 8 kc every row and column of
// register index below kc is eliminated on every lane, so step k updates
// only the (B - kc)^2 trailing registers: a third of the blocked screen's
// FMAs for B = 3. Each entry sees the same operations in the same order as
// in chol_screen (skipped updates there subtract exact zeros).
// (The entries of the matrix come through at(i, j): a tile, or A formed on
// the fly from a G^R tile.)
template <int B, typename At>
__device__ inline void load_herm_cyc_at(At at, int N, int lane, double fr, double fi, c64 (&a)[B][B]) {
    const int lr = lane & 7, lc = lane >> 3;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int i = lr + 8 * r, j = lc + 8 * c;
            if (i < N && j < N) {
                const c64 x = at(i, j), y = at(j, i);
                const double ax = fr * x.x - fi * x.y, ay = fr * x.y + fi * x.x;
                const double bx = fr * y.x - fi * y.y, by = fr * y.y + fi * y.x;
                a[r][c] = cmk(0.5 * (ax + bx), 0.5 * (ay - by));
            } else {
                a[r][c] = cmk(0.0, 0.0);
            }
        }
}

template <int B, int LD>
__device__ inline void load_herm_cyc(const c64* X, int N, int lane, double fr, double fi, c64 (&a)[B][B]) {
    load_herm_cyc_at<B>([&](int i, int j) { return X[i * LD + j]; }, N, lane, fr, fi, a);
}

// A = i (G - G^dagger) at (i, j) from a G^R tile (the arithmetic of the
// spectral kernel, so the values are bitwise those of a stored A).
template <int LD>
__device__ inline c64 spectral_at(const c64* G, int i, int j) {
    const c64 g = G[i * LD + j], gt = G[j * LD + i];
    const double dx = g.x - gt.x, dy = g.y + gt.y;
    return cmk(-dy, dx);
}

template <int B>
__device__ inline BlockScale block_scale_cyc(const c64 (&a)[B][B], int N, int lane) {
    const int lr = lane & 7, lc = lane >> 3;
    double amax = 0.0, fro2 = 0.0, bad = 0.0;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c)
            if (lr + 8 * r < N && lc + 8 * c < N) {
                const c64 v = a[r][c];
                if (!(isfinite(v.x) && isfinite(v.y))) bad = 1.0;
                amax = fmax(amax, fmax(fabs(v.x), fabs(v.y)));
                fro2 += v.x * v.x + v.y * v.y;
            }
    BlockScale s;
    s.bad = wave_max(bad) != 0.0;
    s.amax = wave_max(amax);
    s.fro2 = wave_sum(fro2);
    return s;
}

template <int B>
__device__ inline bool chol_screen_cyc(c64 (&a)[B][B], int N, int lane, EigWork<B>* w, double shift) {
    const int lr = lane & 7, lc = lane >> 3;
    if (lr == lc) {
#pragma unroll
        for (int r = 0; r < B; ++r) a[r][r].x += shift;
    }
    c64* u = w->u;
    bool ok = true;
    static_for<0, B>([&](auto kcc) {
        constexpr int kc = decltype(kcc)::value;
        for (int kl = 0; kl < 8 && ok; ++kl) {
            const int k = 8 * kc + kl;
            if (k >= N) break;
            if (lc == kl) {
#pragma unroll
                for (int r = kc; r < B; ++r) u[lr + 8 * r] = a[r][kc];
            }
            wave_sync();
            const double piv = u[k].x;
            if (!(piv > 0.0)) {
                wave_sync();
                ok = false;
                break;
            }
            const double rs = rsqrt_nr(piv);
            c64 li[B], lj[B];
#pragma unroll
            for (int r = kc; r < B; ++r) {
                const double wi = lr + 8 * r > k ? rs : 0.0;
                const double wj = lc + 8 * r > k ? rs : 0.0;
                const c64 ui = u[lr + 8 * r], uj = u[lc + 8 * r];
                li[r] = cmk(wi * ui.x, wi * ui.y);
                lj[r] = cmk(wj * uj.x, -(wj * uj.y));
            }
#pragma unroll
            for (int r = kc; r < B; ++r)
#pragma unroll
                for (int c = kc; c < B; ++c) a[r][c] = cfms(a[r][c], li[r], lj[c]);
            wave_sync();  // the next step rewrites u
        }
    });
    return ok;
}

// Checks of the residual suite, as bits of a block's screen word: the
// suite's list of unsettled blocks and the Keldysh kernels' raw-product
// screens (with kScrNonFinite for a raw product with a non-finite entry).
enum ResScreen : unsigned {
    kScrCaus = 1u << 0,   // lambda_max(Im S^R)
    kScrGamma = 1u << 1,  // Gamma
    kScrA = 1u << 2,      // A
    kScrLess = 1u << 3,   // -i G<
    kScrGrt = 1u << 4,    // i G>
    kScrRawL = 1u << 5,   // -i G^R S< G^R^dagger
    kScrRawG = 1u << 6,   // i G^R S> G^R^dagger
    kScrNonFinite = 1u << 7
};

// Screen-only forms of psd_from_tile and causality_from_tile for the first
// pass of the residual suite: the value 0 (or Inf for a non-finite block)
// wherever the finiteness check, the zero block or the Cholesky screen
// settles the block; otherwise `unsettled` is set and 0 returned, and a
// second pass computes the eigenvalues (the full functions, which repeat the
// screen on the blocked grid). A screen that passes proves the ratio 0 in
// either layout, so the outcome does not depend on which grid screened.
// Without the eigenvalue code the pass needs far fewer registers.
template <int B, typename At>
__device__ inline double psd_screen_at(At at, int N, int lane, EigWork<B>* w, double fr, double fi, bool& unsettled) {
    constexpr double kEps = 2.220446049250313e-16;
    c64 a[B][B];
    load_herm_cyc_at<B>(at, N, lane, fr, fi, a);
    const BlockScale s = block_scale_cyc<B>(a, N, lane);
    if (s.bad) return INFINITY;
    if (s.amax == 0.0) return 0.0;
    int ex;
    frexp(s.amax, &ex);
    const double sc = ldexp(1.0, -ex);
    scale_regs<B>(a, sc);
    const double delta = 32.0 * N * kEps * fmax(sqrt(s.fro2 / N), 1.0);
    if (!chol_screen_cyc<B>(a, N, lane, w, delta * sc)) unsettled = true;
    return 0.0;
}

template <int B, int LD>
__device__ inline double psd_screen_tile(const c64* X, int N, int lane, EigWork<B>* w, double fr, double fi,
                                         bool& unsettled) {
    return psd_screen_at<B>([&](int i, int j) { return X[i * LD + j]; }, N, lane, w, fr, fi, unsettled);
}

template <int B, int LD>
__device__ inline double causality_screen_tile(const c64* X, int N, int lane, EigWork<B>* w, bool& unsettled) {
    c64 a[B][B];
    load_herm_cyc<B, LD>(X, N, lane, 0.0, 1.0, a);  // Hermitian part of i S = -Im S
    const BlockScale s = block_scale_cyc<B>(a, N, lane);
    if (s.bad) return INFINITY;
    if (s.amax == 0.0) return 0.0;
    int ex;
    frexp(s.amax, &ex);
    const double sc = ldexp(1.0, -ex);
    scale_regs<B>(a, sc);
    const double delta = 32.0 * N * 2.220446049250313e-16 * sqrt(s.fro2 / N);
    if (!chol_screen_cyc<B>(a, N, lane, w, delta * sc)) unsettled = true;
    return 0.0;
}

// Causality of the retarded block S in tile X (residuals.jl:455-492):
// lambda_max(Im S) / (||S||_2 + 1e-14) when positive, else 0; Inf for a
// non-finite block. A Cholesky factorization of -Im S + delta I with
// delta = 32 N eps ||Im S||_F / sqrt(N) settles the block at 0: it proves
// lambda_max(Im S) < delta + (factorization error), i.e. a ratio below
// (32 N + N(N+1)) eps (~2e-13 at N = 24), rounding noise for a causal
// block whose Im S is singular (couplings of low rank). ||S||_2^2 is the
// largest eigenvalue of S S^dagger (that of S^dagger S). T: register
// scratch.
template <int B, int LD>
__device__ inline double causality_from_tile(const c64* X, int N, int lane, EigWork<B>* w, c64 (&T)[B][B]) {
    c64 a[B][B];
    load_herm<B, LD>(X, N, lane, 0.0, 1.0, a);  // Hermitian part of i S = -Im S
    const BlockScale s = block_scale<B>(a, N, lane);
    if (s.bad) return INFINITY;
    if (s.amax == 0.0) return 0.0;
    int ex;
    frexp(s.amax, &ex);
    const double sc = ldexp(1.0, -ex);
    scale_regs<B>(a, sc);
    const double delta = 32.0 * N * 2.220446049250313e-16 * sqrt(s.fro2 / N);
    if (chol_screen<B>(a, N, lane, w, delta * sc)) return 0.0;
    load_herm<B, LD>(X, N, lane, 0.0, -1.0, a);  // Im S = (S - S^dagger) / (2i)
    scale_regs<B>(a, sc);
    double lmin, lmax;
    eig_core<B>(a, N, lane, w, lmin, lmax);
    lmax = ldexp(lmax, ex);
    if (!isfinite(lmax)) return INFINITY;
    if (!(lmax > 0.0)) return 0.0;
    rprod<B>(X, X, N, lane, T, true);
    const BlockScale t = block_scale<B>(T, N, lane);
    if (t.bad) return INFINITY;
    int ex2;
    frexp(t.amax, &ex2);
    scale_regs<B>(T, ldexp(1.0, -ex2));
    double s0, s1;
    eig_core<B>(T, N, lane, w, s0, s1);
    const double v = lmax / (sqrt(fmax(0.0, ldexp(s1, ex2))) + 1e-14);
    return isfinite(v) ? v : INFINITY;
}

// Sum over the N x N part of |X_ij - T_ij|^2 (tile against registers).
template <int B, int LD>
__device__ inline double tile_reg_diff2(const c64* X, const c64 (&T)[B][B], int N, int lane) {
    const int lr = lane & 7, lc = lane >> 3;
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int i = lr * B + r, j = lc * B + c;
            if (i < N && j < N) {
                const c64 d = csub(X[i * LD + j], T[r][c]);
                s += d.x * d.x + d.y * d.y;
            }
        }
    return wave_sum(s);
}

}  // namespace qclhip
