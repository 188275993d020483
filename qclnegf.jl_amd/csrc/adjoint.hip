// Adjoint (vector-Jacobian products) of the GF-solve operators, for
// gradients of the current functional. The reference has no adjoint
// (algorithm_registry.jl:226-249 lists it as missing); these are the exact
// reverses of the forward operators of this library.
//
// Convention: for a real functional J of a complex field Z the gradient is
// Zbar = dJ/dRe Z + i dJ/dIm Z, so dJ = Re sum conj(Zbar) dZ.
//
// current_green (production_state.jl:1-74) reversed, per block:
//   paired map    G<r_bar = a G<_bar - (1-a) G>_bar + Nbar * i k w I
//                 G>r_bar = -c G<_bar + (1-c) G>_bar - Hbar * i k w I
//                 (k = g_s/2pi, w = wE_e wk_m; Nbar, Hbar from the chain
//                 through a, c of reference/greens.jl:155-185)
//   Keldysh       G^R_bar += Gc_bar G^R Sc^dagger + Gc_bar^dagger G^R Sc,
//                 Sc_bar = G^R^dagger Gc_bar G^R          (c = <, >)
//   Dyson         SigmaR_bar = G^R^dagger G^R_bar G^R^dagger,
//                 h_bar_m = sum_e SigmaR_bar_em  (D = eps - h - SigmaR).
#include "dk_common.hpp"

namespace qclhip {

// out (+)= op(L) op(R) over LDS tiles (row-major, stride 8B+1); op = ^dagger
// when LA / RA.
template <int B, bool LA, bool RA>
__device__ inline void bprod(const c64* __restrict__ L, const c64* __restrict__ R, int N, int lr, int lc,
                             c64 (&out)[B][B], bool accumulate) {
    constexpr int LD = 8 * B + 1;
    if (!accumulate) {
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) out[r][c] = cmk(0.0, 0.0);
    }
#pragma unroll 2
    for (int q = 0; q < N; ++q) {
        c64 lv[B], rv[B];
#pragma unroll
        for (int r = 0; r < B; ++r) {
            const int i = lr * B + r;
            lv[r] = LA ? cconj(L[q * LD + i]) : L[i * LD + q];
        }
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int j = lc * B + c;
            rv[c] = RA ? cconj(R[j * LD + q]) : R[q * LD + j];
        }
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) out[r][c] = cfma(out[r][c], lv[r], rv[c]);
    }
}

template <int B>
__device__ inline void bstore(c64* T, int lr, int lc, const c64 (&v)[B][B]) {
    constexpr int LD = 8 * B + 1;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) T[(lr * B + r) * LD + lc * B + c] = v[r][c];
}

// Rows [s 8 B/SPLIT, (s+1) 8 B/SPLIT) of bprod, the share of wave s of SPLIT
// waves on one block (each entry the same q-ordered FMA chain as bprod's).
template <int B, int SPLIT, bool LA, bool RA>
__device__ inline void bprod_rows(const c64* __restrict__ L, const c64* __restrict__ R, int N, int s, int lr, int lc,
                                  c64 (&out)[B / SPLIT][B], bool accumulate) {
    constexpr int LD = 8 * B + 1;
    constexpr int BR = B / SPLIT;
    const int row0 = s * 8 * BR + lr * BR;
    if (!accumulate) {
#pragma unroll
        for (int r = 0; r < BR; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) out[r][c] = cmk(0.0, 0.0);
    }
#pragma unroll 2
    for (int q = 0; q < N; ++q) {
        c64 lv[BR], rv[B];
#pragma unroll
        for (int r = 0; r < BR; ++r) {
            const int i = row0 + r;
            lv[r] = LA ? cconj(L[q * LD + i]) : L[i * LD + q];
        }
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int j = lc * B + c;
            rv[c] = RA ? cconj(R[j * LD + q]) : R[q * LD + j];
        }
#pragma unroll
        for (int r = 0; r < BR; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) out[r][c] = cfma(out[r][c], lv[r], rv[c]);
    }
}

template <int B, int SPLIT>
__device__ inline void bstore_rows(c64* T, int s, int lr, int lc, const c64 (&v)[B / SPLIT][B]) {
    constexpr int LD = 8 * B + 1;
    constexpr int BR = B / SPLIT;
    const int row0 = s * 8 * BR + lr * BR;
#pragma unroll
    for (int r = 0; r < BR; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) T[(row0 + r) * LD + lc * B + c] = v[r][c];
}

// Coefficients of the paired-map reverse: a, c, Nbar, Hbar, k.
struct PairAdjoint {
    double a, c, nbar, hbar, k;
};

struct GAArgs {
    int NE, Nk, N;
    const c64* GR;
    const c64* famL[kMaxFamilies];
    const c64* famG[kMaxFamilies];
    int nfam;
    const c64* GRbar;  // may be NULL
    const c64* GLbar;
    const c64* GGbar;
    const double* wE;
    const double* wk;
    const PairAdjoint* coef;  // device
    c64* SRbar;
    c64* SLbar;
    c64* SGbar;
};

// Tile of k1 A + k2 B (+ kd i k w on the diagonal) for W energies.
template <int W, int LD, int TH = 64 * W>
__device__ inline void tile_load_combo(c64* lds, int wstride, const c64* A, double k1, const c64* Bf, double k2,
                                       double kd, const double* wE, double wkm, int e0, long long blk0, int nvalid,
                                       int N, size_t S) {
    const int total = N * N * W;
    int f = threadIdx.x;
    const int w = f % W;
    int ij = f / W;
    int i = ij % N, j = ij / N;
    const double dw = (w < nvalid) ? kd * wE[e0 + w] * wkm : 0.0;
    for (; f < total; f += TH) {
        if (w < nvalid) {
            const size_t idx = blk0 + w + S * (i + (size_t)N * j);
            const c64 x = A[idx], y = Bf[idx];
            c64 v = cmk(k1 * x.x + k2 * y.x, k1 * x.y + k2 * y.y);
            if (i == j) v.y += dw;
            lds[w * wstride + i * LD + j] = v;
        }
        i += TH / W;
        while (i >= N) {
            i -= N;
            ++j;
        }
    }
}

template <int B, int W>
__global__ __launch_bounds__(64 * W) void k_green_adjoint(GAArgs p) {
    constexpr int NP = 8 * B;
    constexpr int LD = NP + 1;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nte = (p.NE + W - 1) / W;
    const long long t = xcd_tile(blockIdx.x, gridDim.x);
    const int e0 = (int)(t % nte) * W;
    const int m = (int)(t / nte);
    const int nvalid = min(W, p.NE - e0);
    const long long nblk = (long long)p.NE * p.Nk;
    const long long blk0 = e0 + (long long)p.NE * m;
    const int wstride = TileStride<B>::value;
    c64* G0 = reinterpret_cast<c64*>(smem);
    c64* X0 = G0 + (size_t)W * wstride;
    c64* Y0 = X0 + (size_t)W * wstride;
    c64* Z0 = Y0 + (size_t)W * wstride;
    c64* Gs = G0 + (size_t)wave * wstride;
    c64* Xs = X0 + (size_t)wave * wstride;
    c64* Ys = Y0 + (size_t)wave * wstride;
    c64* Zs = Z0 + (size_t)wave * wstride;
    const int N = p.N;
    const size_t S = (size_t)nblk;
    const int lr = lane & 7, lc = lane >> 3;
    const bool active = wave < nvalid;
    const unsigned mask = (nvalid >= 32) ? 0xffffffffu : ((1u << nvalid) - 1u);
    const PairAdjoint cf = *p.coef;
    const c64* gr_fam[1] = {p.GR};
    tile_load<W, LD>(G0, wstride, gr_fam, 1, blk0, nvalid, N, S);
    if (p.GRbar) {
        const c64* grb[1] = {p.GRbar};
        tile_load<W, LD>(Z0, wstride, grb, 1, blk0, nvalid, N, S);
    }
    __syncthreads();
    c64 acc[B][B];
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int i = lr * B + r, j = lc * B + c;
            acc[r][c] = (p.GRbar && i < N && j < N) ? Zs[i * LD + j] : cmk(0.0, 0.0);
        }
    c64 T[B][B];
    for (int comp = 0; comp < 2; ++comp) {
        __syncthreads();
        if (comp == 0)
            tile_load_combo<W, LD>(X0, wstride, p.GLbar, cf.a, p.GGbar, -(1.0 - cf.a), cf.nbar * cf.k, p.wE, p.wk[m],
                                   e0, blk0, nvalid, N, S);
        else
            tile_load_combo<W, LD>(X0, wstride, p.GLbar, -cf.c, p.GGbar, 1.0 - cf.c, -cf.hbar * cf.k, p.wE, p.wk[m],
                                   e0, blk0, nvalid, N, S);
        tile_load<W, LD>(Y0, wstride, comp ? p.famG : p.famL, p.nfam, blk0, nvalid, N, S);
        __syncthreads();
        if (active) {
            bprod<B, false, true>(Gs, Ys, N, lr, lc, T, false);  // G Sc^dagger
            wave_sync();
            bstore<B>(Zs, lr, lc, T);
            wave_sync();
            bprod<B, false, false>(Xs, Zs, N, lr, lc, acc, true);  // += Gc_bar G Sc^dagger
            wave_sync();
            bprod<B, false, false>(Gs, Ys, N, lr, lc, T, false);  // G Sc
            wave_sync();
            bstore<B>(Zs, lr, lc, T);
            wave_sync();
            bprod<B, true, false>(Xs, Zs, N, lr, lc, acc, true);  // += Gc_bar^dagger G Sc
            wave_sync();
            bprod<B, true, false>(Gs, Xs, N, lr, lc, T, false);  // G^dagger Gc_bar
            wave_sync();
            bstore<B>(Zs, lr, lc, T);
            wave_sync();
            bprod<B, false, false>(Zs, Gs, N, lr, lc, T, false);  // (G^dagger Gc_bar) G
            wave_sync();
            bstore<B>(Ys, lr, lc, T);
        }
        __syncthreads();
        tile_store<W, LD>(Y0, wstride, comp ? p.SGbar : p.SLbar, blk0, mask, N, S);
    }
    __syncthreads();
    if (active) {
        bstore<B>(Zs, lr, lc, acc);
        wave_sync();
        bprod<B, true, false>(Gs, Zs, N, lr, lc, T, false);  // G^dagger G^R_bar
        wave_sync();
        bstore<B>(Ys, lr, lc, T);
        wave_sync();
        bprod<B, false, true>(Ys, Gs, N, lr, lc, T, false);  // (.) G^dagger
        wave_sync();
        bstore<B>(Xs, lr, lc, T);
    }
    __syncthreads();
    tile_store<W, LD>(X0, wstride, p.SRbar, blk0, mask, N, S);
}

// k_green_adjoint with SPLIT waves per block sharing its four tiles (wave s
// computes rows s 8 B/SPLIT .. of every product, each entry the same FMA
// chain as in the one-wave kernel, so the results are bitwise equal): four
// tiles per wave would hold a CU to four waves at N <= 24.
template <int B, int W, int SPLIT>
__global__ __launch_bounds__(64 * SPLIT * W) void k_green_adjoint_split(GAArgs p) {
    constexpr int NP = 8 * B;
    constexpr int LD = NP + 1;
    constexpr int TH = 64 * SPLIT * W;
    constexpr int BR = B / SPLIT;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int b = wave / SPLIT, s = wave % SPLIT;
    const int nte = (p.NE + W - 1) / W;
    const long long t = xcd_tile(blockIdx.x, gridDim.x);
    const int e0 = (int)(t % nte) * W;
    const int m = (int)(t / nte);
    const int nvalid = min(W, p.NE - e0);
    const long long nblk = (long long)p.NE * p.Nk;
    const long long blk0 = e0 + (long long)p.NE * m;
    const int wstride = TileStride<B>::value;
    c64* G0 = reinterpret_cast<c64*>(smem);
    c64* X0 = G0 + (size_t)W * wstride;
    c64* Y0 = X0 + (size_t)W * wstride;
    c64* Z0 = Y0 + (size_t)W * wstride;
    c64* Gs = G0 + (size_t)b * wstride;
    c64* Xs = X0 + (size_t)b * wstride;
    c64* Ys = Y0 + (size_t)b * wstride;
    c64* Zs = Z0 + (size_t)b * wstride;
    const int N = p.N;
    const size_t S = (size_t)nblk;
    const int lr = lane & 7, lc = lane >> 3;
    const int row0 = s * 8 * BR + lr * BR;
    const bool active = b < nvalid;
    const unsigned mask = (nvalid >= 32) ? 0xffffffffu : ((1u << nvalid) - 1u);
    const PairAdjoint cf = *p.coef;
    const c64* gr_fam[1] = {p.GR};
    tile_load<W, LD, TH>(G0, wstride, gr_fam, 1, blk0, nvalid, N, S);
    if (p.GRbar) {
        const c64* grb[1] = {p.GRbar};
        tile_load<W, LD, TH>(Z0, wstride, grb, 1, blk0, nvalid, N, S);
    }
    __syncthreads();
    c64 acc[BR][B];
#pragma unroll
    for (int r = 0; r < BR; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int i = row0 + r, j = lc * B + c;
            acc[r][c] = (p.GRbar && i < N && j < N) ? Zs[i * LD + j] : cmk(0.0, 0.0);
        }
    c64 T[BR][B];
    for (int comp = 0; comp < 2; ++comp) {
        __syncthreads();
        if (comp == 0)
            tile_load_combo<W, LD, TH>(X0, wstride, p.GLbar, cf.a, p.GGbar, -(1.0 - cf.a), cf.nbar * cf.k, p.wE,
                                       p.wk[m], e0, blk0, nvalid, N, S);
        else
            tile_load_combo<W, LD, TH>(X0, wstride, p.GLbar, -cf.c, p.GGbar, 1.0 - cf.c, -cf.hbar * cf.k, p.wE,
                                       p.wk[m], e0, blk0, nvalid, N, S);
        tile_load<W, LD, TH>(Y0, wstride, comp ? p.famG : p.famL, p.nfam, blk0, nvalid, N, S);
        __syncthreads();
        if (active) bprod_rows<B, SPLIT, false, true>(Gs, Ys, N, s, lr, lc, T, false);  // G Sc^dagger
        if (active) bstore_rows<B, SPLIT>(Zs, s, lr, lc, T);
        __syncthreads();
        if (active) bprod_rows<B, SPLIT, false, false>(Xs, Zs, N, s, lr, lc, acc, true);  // += Gc_bar G Sc^dagger
        __syncthreads();
        if (active) bprod_rows<B, SPLIT, false, false>(Gs, Ys, N, s, lr, lc, T, false);  // G Sc
        if (active) bstore_rows<B, SPLIT>(Zs, s, lr, lc, T);
        __syncthreads();
        if (active) bprod_rows<B, SPLIT, true, false>(Xs, Zs, N, s, lr, lc, acc, true);  // += Gc_bar^dagger G Sc
        __syncthreads();
        if (active) bprod_rows<B, SPLIT, true, false>(Gs, Xs, N, s, lr, lc, T, false);  // G^dagger Gc_bar
        if (active) bstore_rows<B, SPLIT>(Zs, s, lr, lc, T);
        wave_sync();  // (G^dagger Gc_bar) G reads only this wave's rows
        if (active) bprod_rows<B, SPLIT, false, false>(Zs, Gs, N, s, lr, lc, T, false);
        if (active) bstore_rows<B, SPLIT>(Ys, s, lr, lc, T);  // every wave is past its reads of Sc
        __syncthreads();
        tile_store<W, LD, false, TH>(Y0, wstride, comp ? p.SGbar : p.SLbar, blk0, mask, N, S);
    }
    __syncthreads();
    if (active) bstore_rows<B, SPLIT>(Zs, s, lr, lc, acc);
    __syncthreads();
    if (active) bprod_rows<B, SPLIT, true, false>(Gs, Zs, N, s, lr, lc, T, false);  // G^dagger G^R_bar
    if (active) bstore_rows<B, SPLIT>(Ys, s, lr, lc, T);
    wave_sync();
    if (active) bprod_rows<B, SPLIT, false, true>(Ys, Gs, N, s, lr, lc, T, false);  // (.) G^dagger
    if (active) bstore_rows<B, SPLIT>(Xs, s, lr, lc, T);
    __syncthreads();
    tile_store<W, LD, false, TH>(X0, wstride, p.SRbar, blk0, mask, N, S);
}

// Partial inner products for the paired-map reverse: per thread of a fixed
// grid, Re sum conj(GLbar + GGbar) G<r and Re sum conj(GLbar + GGbar) G>r,
// with the raw fields recovered from the normalized ones,
//   G<r = ((1-c) G< + c G>) / (a - c),  G>r = ((1-a) G< + a G>) / (a - c).
constexpr int kPairBlocks = 1024;
constexpr int kPairThreads = 256;

__global__ __launch_bounds__(kPairThreads) void k_pair_adjoint_partial(long long n, const c64* GL, const c64* GG,
                                                                        const c64* GLbar, const c64* GGbar, double a,
                                                                        double c, double* partial) {
    __shared__ double red[2][kPairThreads];
    const double inv = 1.0 / (a - c);
    double sa = 0.0, sc = 0.0;
    for (long long i = (long long)blockIdx.x * kPairThreads + threadIdx.x; i < n;
         i += (long long)kPairBlocks * kPairThreads) {
        const c64 l = GL[i], g = GG[i], lb = GLbar[i], gb = GGbar[i];
        const double lrx = ((1.0 - c) * l.x + c * g.x) * inv, lry = ((1.0 - c) * l.y + c * g.y) * inv;
        const double grx = ((1.0 - a) * l.x + a * g.x) * inv, gry = ((1.0 - a) * l.y + a * g.y) * inv;
        const double bx = lb.x + gb.x, by = lb.y + gb.y;
        sa += bx * lrx + by * lry;
        sc += bx * grx + by * gry;
    }
    red[0][threadIdx.x] = sa;
    red[1][threadIdx.x] = sc;
    __syncthreads();
    for (int w = kPairThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = red[0][0];
        partial[kPairBlocks + blockIdx.x] = red[1][0];
    }
}

// abar = <Gbar, G<r>, cbar = -<Gbar, G>r>; then Nbar, Hbar through the
// branch of reference/greens.jl:155-185 (number >= target: a = target/N,
// c = 0; else a = 1, c = (target - N)/holes).
__global__ __launch_bounds__(kPairBlocks) void k_pair_adjoint_coef(const double* partial, double a, double c,
                                                                  double number, double holes, double target,
                                                                  double k, PairAdjoint* out) {
    __shared__ double red[2][kPairBlocks];
    red[0][threadIdx.x] = partial[threadIdx.x];
    red[1][threadIdx.x] = partial[kPairBlocks + threadIdx.x];
    __syncthreads();
    for (int w = kPairBlocks / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double abar = red[0][0], cbar = -red[1][0];
    PairAdjoint r;
    r.a = a;
    r.c = c;
    r.k = k;
    if (number >= target) {
        r.nbar = abar * (-target / (number * number));
        r.hbar = 0.0;
    } else {
        r.nbar = -cbar / holes;
        r.hbar = -cbar * (target - number) / (holes * holes);
    }
    *out = r;
}

// hbar[m + Nk*q] = sum_e X[e + NE*(m + Nk*q)]: one wavefront per output,
// fixed lane partition and shuffle tree.
__global__ void k_energy_sum(int NE, long long nout, const c64* X, c64* out) {
    const long long o = (long long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (o >= nout) return;
    double sx = 0.0, sy = 0.0;
    const c64* col = X + (size_t)o * NE;
    for (int e = lane; e < NE; e += 64) {
        sx += col[e].x;
        sy += col[e].y;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off, 64);
        sy += __shfl_xor(sy, off, 64);
    }
    if (lane == 0) out[o] = cmk(sx, sy);
}

// Reverse of the boundary flux Jbar = Re[k sum w tr(S>+ G< - S<+ G>)]
// (production_iteration.jl:189-201), accumulated:
//   GLbar_ab += s conj(S>+_ba), GGbar_ab -= s conj(S<+_ba),
//   S>+bar_ab += s conj(G<_ba), S<+bar_ab -= s conj(G>_ba),  s = Jbar k w.
__global__ void k_flux_adjoint(int NE, int Nk, int N, const double* wE, const double* wk, double sk,
                               const c64* __restrict__ SL, const c64* __restrict__ SG, const c64* __restrict__ GL,
                               const c64* __restrict__ GG, c64* SLbar, c64* SGbar, c64* GLbar, c64* GGbar) {
    const long long nblk = (long long)NE * Nk;
    const long long n = nblk * N * N;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long blk = idx % nblk;
        const long long ab = idx / nblk;
        const int a = (int)(ab % N), b = (int)(ab / N);
        const size_t t = blk + nblk * (b + (size_t)N * a);
        const int e = (int)(blk % NE), m = (int)(blk / NE);
        const double s = sk * wE[e] * wk[m];
        if (GLbar) {
            const c64 v = SG[t];
            GLbar[idx] = cmk(GLbar[idx].x + s * v.x, GLbar[idx].y - s * v.y);
        }
        if (GGbar) {
            const c64 v = SL[t];
            GGbar[idx] = cmk(GGbar[idx].x - s * v.x, GGbar[idx].y + s * v.y);
        }
        if (SGbar) {
            const c64 v = GL[t];
            SGbar[idx] = cmk(SGbar[idx].x + s * v.x, SGbar[idx].y - s * v.y);
        }
        if (SLbar) {
            const c64 v = GG[t];
            SLbar[idx] = cmk(SLbar[idx].x - s * v.x, SLbar[idx].y + s * v.y);
        }
    }
}

// out[j, col] (+)= alpha sum_t tw[t] Y[tsrc[t], col] over the CSR row j of
// the shift's transpose (apply_energy_shift!, production_cache.jl:300-325,
// reversed).
__global__ void k_shift_transpose(int NE, long long ncols, const int* __restrict__ tptr, const int* __restrict__ tsrc,
                                  const double* __restrict__ tw, const c64* __restrict__ Y, double alpha,
                                  int accumulate, c64* __restrict__ out) {
    const long long n = (long long)NE * ncols;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(idx % NE);
        const size_t base = (size_t)(idx - j);
        double sx = 0.0, sy = 0.0;
        for (int q = tptr[j]; q < tptr[j + 1]; ++q) {
            const c64 y = Y[base + tsrc[q]];
            sx += tw[q] * y.x;
            sy += tw[q] * y.y;
        }
        c64 v = cmk(alpha * sx, alpha * sy);
        if (accumulate) v = cadd(out[idx], v);
        out[idx] = v;
    }
}

int shift_transpose(qclhip_ctx* ctx, const qclhip_shift* s, long long ncols, const c64* Y, double alpha,
                    int accumulate, c64* out) {
    const long long n = (long long)s->NE * ncols;
    hipLaunchKernelGGL(k_shift_transpose, dim3(grid_blocks(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream, s->NE,
                       ncols, s->tptr, s->tsrc, s->tw, Y, alpha, accumulate, out);
    return launch_check(ctx, "k_shift_transpose");
}

template <int B, int W, int SPLIT>
int launch_green_adjoint_split(qclhip_ctx* ctx, const GAArgs& a) {
    const size_t lds = 4 * (size_t)TileStride<B>::value * sizeof(c64) * W;
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_green_adjoint_split<B, W, SPLIT>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long tiles = (long long)((a.NE + W - 1) / W) * a.Nk;
    hipLaunchKernelGGL((k_green_adjoint_split<B, W, SPLIT>), dim3((unsigned)tiles), dim3(64 * SPLIT * W), lds,
                       ctx->stream, a);
    return launch_check(ctx, "k_green_adjoint_split");
}

template <int B>
int launch_green_adjoint(qclhip_ctx* ctx, const GAArgs& a) {
    // N = 17..24: three waves per block, two blocks per workgroup (measured
    // at N = 24, NE = 2048, Nk = 64: 10.97 ms for the reverse GF solve
    // against 12.45 ms with one wave per block, 11.66 ms with four blocks)
    if constexpr (B == 3) return launch_green_adjoint_split<B, 2, 3>(ctx, a);
    constexpr size_t per_wave = 4 * (size_t)TileStride<B>::value * sizeof(c64);
    // 4 energies (64-byte segments, one workgroup per CU) where they fit, else 2
    // (two workgroups per CU) or 1: the same waves per CU either way.
    constexpr int W = per_wave * 4 <= 160 * 1024 ? 4 : (per_wave * 2 <= 80 * 1024 ? 2 : 1);
    const size_t lds = per_wave * W;
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_green_adjoint<B, W>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long tiles = (long long)((a.NE + W - 1) / W) * a.Nk;
    hipLaunchKernelGGL((k_green_adjoint<B, W>), dim3((unsigned)tiles), dim3(64 * W), lds, ctx->stream, a);
    return launch_check(ctx, "k_green_adjoint");
}

int green_adjoint_dispatch(qclhip_ctx* ctx, const GAArgs& a) {
    const int N = a.N;
    if (N <= 8) return launch_green_adjoint<1>(ctx, a);
    if (N <= 16) return launch_green_adjoint<2>(ctx, a);
    if (N <= 24) return launch_green_adjoint<3>(ctx, a);
    if (N <= 32) return launch_green_adjoint<4>(ctx, a);
    if (N <= 40) return launch_green_adjoint<5>(ctx, a);
    if (N <= 48) return launch_green_adjoint<6>(ctx, a);
    return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");
}

}  // namespace qclhip

using namespace qclhip;

extern "C" {

int qclhip_current_green_adjoint(qclhip_ctx* ctx, const qclhip_grid* grid, int nfam, const double* const* famL,
                                 const double* const* famG, double target, const double* coef_host,
                                 const double* GR, const double* GL, const double* GG, const double* GRbar,
                                 const double* GLbar, const double* GGbar, double* SRbar, double* SLbar,
                                 double* SGbar, double* hbar) {
    if (!ctx || !grid || !famL || !famG || !coef_host || !GR || !GL || !GG || !GLbar || !GGbar || !SRbar ||
        !SLbar || !SGbar)
        return QCLHIP_ERR_ARGUMENT;
    if (nfam < 1 || nfam > kMaxFamilies)
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "the adjoint takes 1..8 self-energy families");
    const double number = coef_host[1], holes = coef_host[2], a = coef_host[3], c = coef_host[4];
    if (!(std::fabs(a - c) > 1e-300))
        return fail(ctx, QCLHIP_ERR_DOMAIN, "paired map is singular (a = c); raw Keldysh fields are unrecoverable");
    if (number < target && !(holes > 0))
        return fail(ctx, QCLHIP_ERR_DOMAIN, "hole functional must be positive on the c > 0 branch");
    const long long nblk = (long long)grid->NE * grid->Nk;
    const long long n = nblk * grid->Nb * grid->Nb;
    char* ws = (char*)workspace(ctx, (size_t)2 * kPairBlocks * sizeof(double) + 256);
    if (!ws) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    double* partial = (double*)ws;
    PairAdjoint* coef = (PairAdjoint*)(ws + (size_t)2 * kPairBlocks * sizeof(double));
    const double k = (double)grid->gs / (2.0 * M_PI);
    hipLaunchKernelGGL(k_pair_adjoint_partial, dim3(kPairBlocks), dim3(kPairThreads), 0, ctx->stream, n,
                       (const c64*)GL, (const c64*)GG, (const c64*)GLbar, (const c64*)GGbar, a, c, partial);
    hipLaunchKernelGGL(k_pair_adjoint_coef, dim3(1), dim3(kPairBlocks), 0, ctx->stream, partial, a, c, number,
                       holes, target, k, coef);
    int rc = launch_check(ctx, "pair adjoint");
    if (rc) return rc;
    GAArgs g{};
    g.NE = grid->NE;
    g.Nk = grid->Nk;
    g.N = grid->Nb;
    g.GR = (const c64*)GR;
    for (int f = 0; f < nfam; ++f) {
        g.famL[f] = (const c64*)famL[f];
        g.famG[f] = (const c64*)famG[f];
    }
    g.nfam = nfam;
    g.GRbar = (const c64*)GRbar;
    g.GLbar = (const c64*)GLbar;
    g.GGbar = (const c64*)GGbar;
    g.wE = grid->wE;
    g.wk = grid->wk;
    g.coef = coef;
    g.SRbar = (c64*)SRbar;
    g.SLbar = (c64*)SLbar;
    g.SGbar = (c64*)SGbar;
    rc = green_adjoint_dispatch(ctx, g);
    if (rc || !hbar) return rc;
    const long long nout = (long long)grid->Nk * grid->Nb * grid->Nb;
    hipLaunchKernelGGL(k_energy_sum, dim3((unsigned)((nout + 3) / 4)), dim3(256), 0, ctx->stream, grid->NE, nout,
                       (const c64*)SRbar, (c64*)hbar);
    return launch_check(ctx, "k_energy_sum");
}

int qclhip_boundary_flux_adjoint(qclhip_ctx* ctx, const qclhip_grid* grid, double Jbar, const double* plusL,
                                 const double* plusG, const double* GL, const double* GG, double* plusLbar,
                                 double* plusGbar, double* GLbar, double* GGbar) {
    if (!ctx || !grid || !plusL || !plusG || !GL || !GG) return QCLHIP_ERR_ARGUMENT;
    const long long n = (long long)grid->NE * grid->Nk * grid->Nb * grid->Nb;
    const double sk = Jbar * (double)grid->gs / (2.0 * M_PI);
    hipLaunchKernelGGL(k_flux_adjoint, dim3(grid_blocks(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream,
                       grid->NE, grid->Nk, grid->Nb, grid->wE, grid->wk, sk, (const c64*)plusL, (const c64*)plusG,
                       (const c64*)GL, (const c64*)GG, (c64*)plusLbar, (c64*)plusGbar, (c64*)GLbar, (c64*)GGbar);
    return launch_check(ctx, "k_flux_adjoint");
}

int qclhip_shift_transpose(qclhip_ctx* ctx, const qclhip_grid* grid, const qclhip_shift* shift, const double* Y,
                           double alpha, int accumulate, double* X) {
    if (!ctx || !grid || !shift || !Y || !X) return QCLHIP_ERR_ARGUMENT;
    if (shift->NE != grid->NE) return fail(ctx, QCLHIP_ERR_DIMENSION, "shift plan has the wrong energy size");
    if (X == Y) return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    return shift_transpose(ctx, shift, (long long)grid->Nk * grid->Nb * grid->Nb, (const c64*)Y, alpha,
                           accumulate ? 1 : 0, (c64*)X);
}

}  // extern "C"
