// Per-iteration residual suite on the device (production_residual_suite,
// residuals.jl:537-772, and production_selfenergy_residual, :802-865): the
// convergence gates of the SCBA loop without downloading fields.
//
// One wavefront per (e,m) block. Block products reuse the lane-grid product
// of the Keldysh kernels; the positivity and causality checks need Hermitian
// extreme eigenvalues, computed by a Householder reduction to tridiagonal
// form in registers and Sturm multisection (eig_core, wave_eig.hpp), after a
// Cholesky screen that settles positive blocks.
// Per-block values are reduced afterwards in a fixed order.
#include <algorithm>
#include <vector>

#include "wave_eig.hpp"

namespace qclhip {

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
    // 0: the whole suite; 1: every term but r_K (no candidate; the state
    // pass, before the candidates exist)
    int mode;
    double* out;  // [7][NE*Nk]: rD, specN, specD, kelN, kelD, psd, caus
    // Blocks whose Cholesky screens did not settle: (blk << 8) | screen bits,
    // appended through the counter (zeroed before the first pass).
    unsigned long long* flagged;
    unsigned* nflagged;
    // Prescreened mode: the Keldysh kernel's screen word per block for the
    // raw products (NULL: the suite forms and screens them itself).
    const unsigned char* raw_screen;
};

// A = i (G - G^dagger) of the wave's own G^R tile into X, the arithmetic of
// k_spectral (bitwise the stored spectral function).
template <int LD>
__device__ inline void tile_spectral(const c64* G, c64* X, int N, int lane) {
    for (int q = lane; q < N * N; q += 64) {
        const int i = q % N, j = q / N;
        const c64 g = G[i * LD + j], gt = G[j * LD + i];
        const double dx = g.x - gt.x, dy = g.y + gt.y;
        X[i * LD + j] = cmk(-dy, dx);
    }
}

// Gamma = i (S> - S<) of the summed families, cooperatively into the W
// per-wave tiles (the thread mapping of tile_load).
template <int W, int LD, int TH = 64 * W>
__device__ inline void tile_load_gamma(c64* lds, int wstride, const c64* const* famG, const c64* const* famL,
                                       int nfam, long long blk0, int nvalid, int N, size_t S) {
    const int total = N * N * W;
    int f = threadIdx.x;
    const int w = f % W;
    int ij = f / W;
    int i = ij % N, j = ij / N;
    for (; f < total; f += TH) {
        if (w < nvalid) {
            const size_t idx = blk0 + w + S * (i + (size_t)N * j);
            const c64 g = fam_load(famG, nfam, idx), l = fam_load(famL, nfam, idx);
            lds[w * wstride + i * LD + j] = cmk(-(g.y - l.y), g.x - l.x);
        }
        i += TH / W;
        while (i >= N) {
            i -= N;
            ++j;
        }
    }
}

// LDS of one wave of the residual suite: two tiles and the eigen work vectors.
template <int B>
constexpr size_t res_wave_bytes() {
    return 2 * (size_t)TileStride<B>::value * sizeof(c64) + sizeof(EigWork<B>);
}

// Waves (consecutive blocks) per workgroup: as many as keep two workgroups
// within the 160 KiB of LDS of a CU, at most 4 (64-byte load segments);
// four in one workgroup per CU where two would not fit.
template <int B>
constexpr int res_waves() {
    return 4 * res_wave_bytes<B>() <= 81920 ? 4
           : (4 * res_wave_bytes<B>() <= 163840 ? 4 : (2 * res_wave_bytes<B>() <= 81920 ? 2 : 1));
}

// Register prefetch of Gamma = i (S> - S<) of the summed families with the
// thread mapping of TileMap (the value of tile_load_gamma).
template <int B, int W>
__device__ inline void fetch_gamma(const TileMap<B, W>& tm, c64 (&r)[TileMap<B, W>::EPT], const c64* const* famG,
                                   const c64* const* famL, int nfam, long long blk0, int nvalid, size_t S) {
    const bool on = tm.w < nvalid;
#pragma unroll
    for (int k = 0; k < TileMap<B, W>::EPT; ++k)
        if (on && k < tm.n) {
            const size_t idx = blk0 + tm.w + S * (size_t)tm.gidx[k];
            const c64 g = fam_load(famG, nfam, idx), l = fam_load(famL, nfam, idx);
            r[k] = cmk(-(g.y - l.y), g.x - l.x);
        }
}

// T = G X G^dagger of one wave's tiles (X is overwritten by G X); with
// STORE the result also replaces X.
template <int B, bool STORE>
__device__ inline void res_sandwich(const c64* G, c64* X, int N, int lane, c64 (&T)[B][B]) {
    rprod<B>(G, X, N, lane, T, false);
    wave_sync();
    rstore<B>(X, N, lane, T);
    wave_sync();
    rprod<B>(X, G, N, lane, T, true);
    wave_sync();
    if (STORE) {
        rstore<B>(X, N, lane, T);
        wave_sync();
    }
}

// W waves per workgroup, one (e,m) block per wave, blocks blk0 .. blk0+W-1.
// Each wave owns two LDS tiles: G^R, and a working tile X that successively
// holds S^R, D, Gamma, G Gamma, A, the candidate S<, G S<, G<, G>, and the
// raw products; the right factor of each product stays in registers. Fields
// reach X through a register prefetch: the cooperative loads of the next
// field (consecutive threads = consecutive blocks, 16W-byte segments) are
// issued before the current phase's products and screens run, and written
// to LDS between two workgroup barriers once X is free. The positivity and
// causality checks run their Cholesky screens here; a block whose screen
// does not settle is appended to a list with one bit per unsettled check,
// and k_residual_fallback computes its eigenvalues afterwards. Waves past the
// last block only join the barriers.
template <int B, int W>
__global__ __launch_bounds__(64 * W) void k_residual_suite(ResArgs p) {
    constexpr int LD = 8 * B + 1;
    using TM = TileMap<B, W>;
    const int wstride = TileStride<B>::value;
    extern __shared__ __align__(16) unsigned char smem[];
    c64* Gall = reinterpret_cast<c64*>(smem);
    c64* Xall = Gall + W * wstride;
    EigWork<B>* ewall = reinterpret_cast<EigWork<B>*>(Xall + W * wstride);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 7, lc = lane >> 3;
    const long long nblk = (long long)p.NE * p.Nk;
    // XCD-aware: consecutive tiles (sharing cache lines) on one XCD's L2
    const long long blk0 = (long long)W * xcd_tile(blockIdx.x, (nblk + W - 1) / W);
    const int nvalid = (int)min((long long)W, nblk - blk0);
    const bool valid = W == 1 || wave < nvalid;  // a one-wave tile is never empty
    const long long blk = blk0 + (valid ? wave : 0);  // e + NE m
    const int e = (int)(blk % p.NE), m = (int)(blk / p.NE);
    const int N = p.N;
    const size_t S = (size_t)nblk;
    c64* G = Gall + wave * wstride;
    c64* X = Xall + wave * wstride;
    EigWork<B>* ew = ewall + wave;
    c64 T[B][B];
    // Register prefetch up to N = 32; larger tiles would spill, and their
    // fields are loaded straight into LDS once X is free.
    constexpr bool kPrefetch = B <= 4;
    TM tm;
    tm.init(N);
    c64 P[kPrefetch ? TM::EPT : 1];  // the next field of X, in flight
    const c64* one[1] = {p.GR};
    const c64* fA[1] = {p.A};
    const c64* fL[1] = {p.GL};
    const c64* fGg[1] = {p.GG};
    tile_load<W, LD>(Gall, wstride, one, 1, blk0, nvalid, N, S);
    tile_load<W, LD>(Xall, wstride, p.famR, p.nfam, blk0, nvalid, N, S);
    __syncthreads();
    // The next field of X: fetched into registers now, or remembered and
    // loaded by put().
    const c64* const* pend = nullptr;
    int pend_n = 0;
    bool pend_gamma = true;
    auto fetch = [&](const c64* const* fam, int nfam) {
        pend = fam;
        pend_n = nfam;
        pend_gamma = false;
        if constexpr (kPrefetch) tm.fetch(P, fam, nfam, blk0, nvalid, S);
    };
    if constexpr (kPrefetch) fetch_gamma<B, W>(tm, P, p.famG, p.famL, p.nfam, blk0, nvalid, S);
    // Moves the next field into the X tiles once every wave is done with them.
    auto put = [&]() {
        __syncthreads();
        if constexpr (kPrefetch) {
            tm.put(Xall, wstride, P, nvalid);
        } else {
            if (pend_gamma)
                tile_load_gamma<W, LD>(Xall, wstride, p.famG, p.famL, p.nfam, blk0, nvalid, N, S);
            else
                tile_load<W, LD>(Xall, wstride, pend, pend_n, blk0, nvalid, N, S);
        }
        __syncthreads();
    };
    unsigned unsettled = 0;
    auto screen = [&](unsigned bit, double fr, double fi) {
        bool u = false;
        const double v = psd_screen_tile<B, LD>(X, N, lane, ew, fr, fi, u);
        if (u) unsettled |= bit;
        wave_sync();
        return v;
    };
    double caus = 0.0, rD = 0.0;
    if (valid) {
        bool u = false;
        caus = causality_screen_tile<B, LD>(X, N, lane, ew, u);
        if (u) unsettled |= kScrCaus;
        wave_sync();
        // r_D: D = eps - h - S^R in X, ||D G - I|| / (||D|| ||G|| + sqrt(N));
        // each lane rewrites its register-block positions (all h loads in flight).
        const double ee = p.eps[e];
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const int i = lr * B + r, j = lc * B + c;
                if (i < N && j < N) {
                    const c64 hv = p.h[(size_t)m + (size_t)p.Nk * (i + (size_t)N * j)];
                    const c64 sv = X[i * LD + j];
                    X[i * LD + j] = cmk((i == j ? ee : 0.0) - hv.x - sv.x, -hv.y - sv.y);
                }
            }
        wave_sync();
        rprod<B>(X, G, N, lane, T, false);
        double dn = 0.0;
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const int i = lr * B + r, j = lc * B + c;
                if (i < N && j < N) {
                    const c64 v = cmk(T[r][c].x - (i == j ? 1.0 : 0.0), T[r][c].y);
                    dn += v.x * v.x + v.y * v.y;
                }
            }
        rD = sqrt(wave_sum(dn)) /
             (sqrt(tile_frob2<LD>(X, N, lane, 0.0)) * sqrt(tile_frob2<LD>(G, N, lane, 0.0)) + sqrt((double)N));
    }
    // Gamma = i (S> - S<): its screen, then T = G Gamma G^dagger (r_A).
    put();
    if (p.A) fetch(fA, 1);
    double psd = 0.0, specN = 0.0, specD = 0.0, kelN = 0.0, kelD = 0.0;
    if (valid) {
        psd = fmax(psd, screen(kScrGamma, 1.0, 0.0));
        res_sandwich<B, false>(G, X, N, lane, T);
    }
    // A (stored, or formed from the G^R tile): the r_A terms and its screen.
    if (p.A) {
        put();
    } else {
        __syncthreads();
        if (valid) tile_spectral<LD>(G, X, N, lane);
        __syncthreads();
    }
    const bool cand = p.mode != 1;  // the state pass has no candidate
    if (cand)
        fetch(p.cand, p.ncand);
    else
        fetch(fL, 1);
    if (valid) {
        specN = tile_reg_diff2<B, LD>(X, T, N, lane);
        specD = tile_frob2<LD>(X, N, lane, 0.0);
        psd = fmax(psd, screen(kScrA, 1.0, 0.0));
    }
    // The candidate S<: T = G S<_cand G^dagger (r_K).
    if (cand) {
        put();
        fetch(fL, 1);
        if (valid) res_sandwich<B, false>(G, X, N, lane, T);
    }
    // G<: the r_K terms and its screen.
    put();
    fetch(fGg, 1);
    if (valid) {
        if (cand) {
            kelN = tile_reg_diff2<B, LD>(X, T, N, lane);
            kelD = tile_frob2<LD>(X, N, lane, 0.0);
        }
        psd = fmax(psd, screen(kScrLess, 0.0, -1.0));
    }
    // G>.
    put();
    if (!p.raw_screen) fetch(p.famL, p.nfam);
    if (valid) psd = fmax(psd, screen(kScrGrt, 0.0, 1.0));
    // The raw products G S< G^dagger and G S> G^dagger: screened here, or
    // (prescreened mode) by the Keldysh kernel that formed them from the same
    // G^R and total self-energy with this arithmetic, whose screen word is
    // taken instead (a raw product with a non-finite entry gives Inf).
    if (p.raw_screen) {
        if (valid) {
            const unsigned rb = p.raw_screen[blk];
            if (rb & kScrNonFinite) psd = INFINITY;
            unsettled |= rb & (kScrRawL | kScrRawG);
        }
    } else {
        put();
        fetch(p.famG, p.nfam);
        if (valid) {
            res_sandwich<B, true>(G, X, N, lane, T);
            psd = fmax(psd, screen(kScrRawL, 0.0, -1.0));
        }
        put();
        if (valid) {
            res_sandwich<B, true>(G, X, N, lane, T);
            psd = fmax(psd, screen(kScrRawG, 0.0, 1.0));
        }
    }
    if (valid && lane == 0) {
        p.out[0 * nblk + blk] = rD;
        p.out[1 * nblk + blk] = specN;
        p.out[2 * nblk + blk] = specD;
        p.out[3 * nblk + blk] = kelN;
        p.out[4 * nblk + blk] = kelD;
        p.out[5 * nblk + blk] = psd;
        p.out[6 * nblk + blk] = caus;
        if (unsettled) {
            const unsigned slot = atomicAdd(p.nflagged, 1u);
            p.flagged[slot] = ((unsigned long long)blk << 8) | unsettled;
        }
    }
}

// The suite in prescreened mode with A formed from G^R (the SCBA loop's
// call) at N = 25..32 (B = 4), B waves per block: wave s of a block computes
// row 8 s + lr of every block product (block_product_rows, so the entries are
// bitwise those of the one-wave products) and forms D on those rows; the
// block's first wave runs the screens and the norms, reading the stored
// products from LDS in the one-wave lane order, so that every per-block value
// is bitwise that of k_residual_suite. Sharing the two tiles of a block over
// B waves raises the waves a CU holds from 4 (B = 4: two 17 KB tiles per
// wave) to 16. The sandwich products are stored into X; G< goes into the G
// tile once G^R is no longer read.
template <int B, int W>
__global__ __launch_bounds__(64 * B * W) void k_residual_suite_split(ResArgs p) {
    constexpr int SPLIT = B;
    constexpr int LD = 8 * B + 1;
    constexpr int TH = 64 * SPLIT * W;
    const int wstride = TileStride<B>::value;
    extern __shared__ __align__(16) unsigned char smem[];
    c64* Gall = reinterpret_cast<c64*>(smem);
    c64* Xall = Gall + W * wstride;
    EigWork<B>* ewall = reinterpret_cast<EigWork<B>*>(Xall + W * wstride);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = wave / SPLIT, s = wave % SPLIT;
    const int lr = lane & 7, lc = lane >> 3;
    const long long nblk = (long long)p.NE * p.Nk;
    const long long blk0 = (long long)W * xcd_tile(blockIdx.x, (nblk + W - 1) / W);
    const int nvalid = (int)min((long long)W, nblk - blk0);
    const bool valid = b < nvalid;
    const bool lead = valid && s == 0;
    const long long blk = blk0 + (valid ? b : 0);  // e + NE m
    const int e = (int)(blk % p.NE), m = (int)(blk / p.NE);
    const int N = p.N;
    const size_t S = (size_t)nblk;
    c64* G = Gall + b * wstride;
    c64* X = Xall + b * wstride;
    EigWork<B>* ew = ewall + b;
    const int i0 = 8 * s + lr;  // this lane's row of the block products
    c64 T[1][B];
    // the rows of T into X
    auto store_rows = [&]() {
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int j = lc * B + c;
            if (i0 < N && j < N) X[i0 * LD + j] = T[0][c];
        }
    };
    // T = G X G^dagger on this wave's rows, stored into X (X holds G X in
    // between); ends with every row stored
    auto sandwich = [&]() {
        if (valid) block_product_rows<B, SPLIT, false>(G, X, N, s, lr, lc, T);
        __syncthreads();  // every wave of the block has read X
        if (valid) store_rows();
        wave_sync();  // T G^dagger reads only this wave's rows of G X
        if (valid) block_product_rows<B, SPLIT, true>(X, G, N, s, lr, lc, T);
        wave_sync();
        if (valid) store_rows();
        __syncthreads();
    };
    const c64* one[1] = {p.GR};
    tile_load<W, LD, TH>(Gall, wstride, one, 1, blk0, nvalid, N, S);
    tile_load<W, LD, TH>(Xall, wstride, p.famR, p.nfam, blk0, nvalid, N, S);
    __syncthreads();
    unsigned unsettled = 0;
    double caus = 0.0, rD = 0.0, psd = 0.0, specN = 0.0, specD = 0.0, kelN = 0.0, kelD = 0.0;
    if (lead) {
        bool u = false;
        caus = causality_screen_tile<B, LD>(X, N, lane, ew, u);
        if (u) unsettled |= kScrCaus;
    }
    __syncthreads();  // S^R read before D replaces it
    // r_D: D = eps - h - S^R on this wave's rows, ||D G - I|| / (||D|| ||G|| + sqrt(N))
    if (valid) {
        const double ee = p.eps[e];
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int i = i0, j = lc * B + c;
            if (i < N && j < N) {
                const c64 hv = p.h[(size_t)m + (size_t)p.Nk * (i + (size_t)N * j)];
                const c64 sv = X[i * LD + j];
                X[i * LD + j] = cmk((i == j ? ee : 0.0) - hv.x - sv.x, -hv.y - sv.y);
            }
        }
    }
    __syncthreads();
    double fD = 0.0, fG = 0.0;
    if (lead) {
        fD = tile_frob2<LD>(X, N, lane, 0.0);
        fG = tile_frob2<LD>(G, N, lane, 0.0);
    }
    if (valid) block_product_rows<B, SPLIT, false>(X, G, N, s, lr, lc, T);
    __syncthreads();  // D read by every wave before D G replaces it
    if (valid) store_rows();
    __syncthreads();
    if (lead) {
        double dn = 0.0;
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const int i = lr * B + r, j = lc * B + c;
                if (i < N && j < N) {
                    const c64 t = X[i * LD + j];
                    const c64 v = cmk(t.x - (i == j ? 1.0 : 0.0), t.y);
                    dn += v.x * v.x + v.y * v.y;
                }
            }
        rD = sqrt(wave_sum(dn)) / (sqrt(fD) * sqrt(fG) + sqrt((double)N));
    }
    __syncthreads();  // before Gamma replaces X
    // Gamma = i (S> - S<): its screen, then G Gamma G^dagger (r_A)
    tile_load_gamma<W, LD, TH>(Xall, wstride, p.famG, p.famL, p.nfam, blk0, nvalid, N, S);
    __syncthreads();
    if (lead) {
        bool u = false;
        psd = fmax(psd, psd_screen_tile<B, LD>(X, N, lane, ew, 1.0, 0.0, u));
        if (u) unsettled |= kScrGamma;
        wave_sync();
    }
    sandwich();
    // A = i (G - G^dagger) from the G^R tile: the r_A terms and its screen
    if (lead) {
        double sn = 0.0, sd = 0.0;
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const int i = lr * B + r, j = lc * B + c;
                if (i < N && j < N) {
                    const c64 av = spectral_at<LD>(G, i, j);
                    const c64 d = csub(av, X[i * LD + j]);
                    sn += d.x * d.x + d.y * d.y;
                    c64 v = av;
                    if (i == j) v.x -= 0.0;
                    sd += v.x * v.x + v.y * v.y;
                }
            }
        specN = wave_sum(sn);
        specD = wave_sum(sd);
        bool u = false;
        psd = fmax(psd, psd_screen_at<B>([&](int i, int j) { return spectral_at<LD>(G, i, j); }, N, lane, ew, 1.0,
                                         0.0, u));
        if (u) unsettled |= kScrA;
        wave_sync();
    }
    const bool cand = p.mode != 1;  // the state pass has no candidate
    if (cand) {
        // the candidate S<: G S<_cand G^dagger (r_K)
        __syncthreads();  // X read by the first wave before the candidate replaces it
        tile_load<W, LD, TH>(Xall, wstride, p.cand, p.ncand, blk0, nvalid, N, S);
        __syncthreads();
        sandwich();
    }
    // G< into the G tile (G^R is no longer read): the r_K terms and its screen
    __syncthreads();
    const c64* fL[1] = {p.GL};
    tile_load<W, LD, TH>(Gall, wstride, fL, 1, blk0, nvalid, N, S);
    __syncthreads();
    if (lead) {
        if (cand) {
            double kn = 0.0;
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int c = 0; c < B; ++c) {
                    const int i = lr * B + r, j = lc * B + c;
                    if (i < N && j < N) {
                        const c64 d = csub(G[i * LD + j], X[i * LD + j]);
                        kn += d.x * d.x + d.y * d.y;
                    }
                }
            kelN = wave_sum(kn);
            kelD = tile_frob2<LD>(G, N, lane, 0.0);
        }
        bool u = false;
        psd = fmax(psd, psd_screen_tile<B, LD>(G, N, lane, ew, 0.0, -1.0, u));
        if (u) unsettled |= kScrLess;
        wave_sync();
    }
    __syncthreads();  // X read by the first wave before G> replaces it
    const c64* fGg[1] = {p.GG};
    tile_load<W, LD, TH>(Xall, wstride, fGg, 1, blk0, nvalid, N, S);
    __syncthreads();
    if (lead) {
        bool u = false;
        psd = fmax(psd, psd_screen_tile<B, LD>(X, N, lane, ew, 0.0, 1.0, u));
        if (u) unsettled |= kScrGrt;
        const unsigned rb = p.raw_screen[blk];
        if (rb & kScrNonFinite) psd = INFINITY;
        unsettled |= rb & (kScrRawL | kScrRawG);
        if (lane == 0) {
            p.out[0 * nblk + blk] = rD;
            p.out[1 * nblk + blk] = specN;
            p.out[2 * nblk + blk] = specD;
            p.out[3 * nblk + blk] = kelN;
            p.out[4 * nblk + blk] = kelD;
            p.out[5 * nblk + blk] = psd;
            p.out[6 * nblk + blk] = caus;
            if (unsettled) {
                const unsigned slot = atomicAdd(p.nflagged, 1u);
                p.flagged[slot] = ((unsigned long long)blk << 8) | unsettled;
            }
        }
    }
}

// Second pass of the suite: the listed blocks whose Cholesky screens did not
// settle. One wave per block (grid-stride over the list); each unsettled
// check rebuilds its matrix as the first pass did (same loads, same product
// arithmetic) and runs the full check (psd_from_tile, causality_from_tile:
// Householder + Sturm eigenvalues), raising the block's r_PSD entry and
// setting its r_caus entry. The list holds each block at most once.
template <int B>
__global__ __launch_bounds__(64) void k_residual_fallback(ResArgs p) {
    constexpr int LD = 8 * B + 1;
    const int wstride = TileStride<B>::value;
    extern __shared__ __align__(16) unsigned char smem[];
    c64* G = reinterpret_cast<c64*>(smem);
    c64* X = G + wstride;
    EigWork<B>* ew = reinterpret_cast<EigWork<B>*>(X + wstride);
    const int lane = threadIdx.x;
    const long long nblk = (long long)p.NE * p.Nk;
    const int N = p.N;
    const size_t S = (size_t)nblk;
    const unsigned count = *p.nflagged;
    c64 T[B][B];
    const c64* one[1];
    for (unsigned i = blockIdx.x; i < count; i += gridDim.x) {
        const unsigned long long entry = p.flagged[i];
        const long long blk = (long long)(entry >> 8);
        const unsigned bits = (unsigned)(entry & 0xffu);
        double psd = 0.0;
        if ((bits & (kScrRawL | kScrRawG)) || ((bits & kScrA) && !p.A)) {
            one[0] = p.GR;
            tile_load<1, LD>(G, wstride, one, 1, blk, 1, N, S);
        }
        if (bits & kScrCaus) {
            tile_load<1, LD>(X, wstride, p.famR, p.nfam, blk, 1, N, S);
            wave_sync();
            const double c = causality_from_tile<B, LD>(X, N, lane, ew, T);
            if (lane == 0) p.out[6 * nblk + blk] = c;
        }
        wave_sync();
        if (bits & kScrGamma) {
            tile_load_gamma<1, LD>(X, wstride, p.famG, p.famL, p.nfam, blk, 1, N, S);
            wave_sync();
            psd = fmax(psd, psd_from_tile<B, LD>(X, N, lane, ew, 1.0, 0.0));
            wave_sync();
        }
        if (bits & kScrA) {
            if (p.A) {
                one[0] = p.A;
                tile_load<1, LD>(X, wstride, one, 1, blk, 1, N, S);
            } else {
                wave_sync();
                tile_spectral<LD>(G, X, N, lane);
            }
            wave_sync();
            psd = fmax(psd, psd_from_tile<B, LD>(X, N, lane, ew, 1.0, 0.0));
            wave_sync();
        }
        if (bits & (kScrLess | kScrGrt)) {
            for (int k = 0; k < 2; ++k) {
                if (!(bits & (k ? kScrGrt : kScrLess))) continue;
                one[0] = k ? p.GG : p.GL;
                tile_load<1, LD>(X, wstride, one, 1, blk, 1, N, S);
                wave_sync();
                psd = fmax(psd, psd_from_tile<B, LD>(X, N, lane, ew, 0.0, k ? 1.0 : -1.0));
                wave_sync();
            }
        }
        for (int k = 0; k < 2; ++k) {
            if (!(bits & (k ? kScrRawG : kScrRawL))) continue;
            tile_load<1, LD>(X, wstride, k ? p.famG : p.famL, p.nfam, blk, 1, N, S);
            wave_sync();
            res_sandwich<B, true>(G, X, N, lane, T);
            psd = fmax(psd, psd_from_tile<B, LD>(X, N, lane, ew, 0.0, k ? 1.0 : -1.0));
            wave_sync();
        }
        if (lane == 0) p.out[5 * nblk + blk] = fmax(p.out[5 * nblk + blk], psd);
        wave_sync();
    }
}

// r_K alone (residuals.jl:643-664): ||G< - G^R S<_cand G^R^dagger||^2 and
// ||G<||^2 per block into rows 3 and 4, the candidate families summed on
// load. Two LDS tiles per wave (G^R, then the candidate and G<).
template <int B, int W>
__global__ __launch_bounds__(64 * W) void k_residual_keldysh(ResArgs p) {
    constexpr int LD = 8 * B + 1;
    const int wstride = TileStride<B>::value;
    extern __shared__ __align__(16) unsigned char smem[];
    c64* Gall = reinterpret_cast<c64*>(smem);
    c64* Xall = Gall + W * wstride;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long nblk = (long long)p.NE * p.Nk;
    const long long blk0 = (long long)W * xcd_tile(blockIdx.x, (nblk + W - 1) / W);
    const int nvalid = (int)min((long long)W, nblk - blk0);
    const bool valid = W == 1 || wave < nvalid;
    const long long blk = blk0 + (valid ? wave : 0);
    const int N = p.N;
    const size_t S = (size_t)nblk;
    c64* G = Gall + wave * wstride;
    c64* X = Xall + wave * wstride;
    c64 T[B][B];
    const c64* one[1];
    one[0] = p.GR;
    tile_load<W, LD>(Gall, wstride, one, 1, blk0, nvalid, N, S);
    tile_load<W, LD>(Xall, wstride, p.cand, p.ncand, blk0, nvalid, N, S);
    __syncthreads();
    if (valid) {
        rprod<B>(G, X, N, lane, T, false);
        wave_sync();
        rstore<B>(X, N, lane, T);
        wave_sync();
        rprod<B>(X, G, N, lane, T, true);
    }
    __syncthreads();  // every wave is done with its X
    one[0] = p.GL;
    tile_load<W, LD>(Xall, wstride, one, 1, blk0, nvalid, N, S);
    __syncthreads();
    if (valid) {
        const double dn2 = tile_reg_diff2<B, LD>(X, T, N, lane);
        const double xn2 = tile_frob2<LD>(X, N, lane, 0.0);
        if (lane == 0) {
            p.out[3 * nblk + blk] = dn2;
            p.out[4 * nblk + blk] = xn2;
        }
    }
}

// Fixed-order reduction of the per-block values: sums for rows 1-4, maxima
// for rows 0, 5, 6; a non-finite value makes its residual Inf. Rows outside
// the mask `rows` are not read (their results are 0).
__global__ __launch_bounds__(1024) void k_residual_reduce(long long nblk, const double* v, double* out,
                                                          unsigned rows) {
    __shared__ double red[7][1024];
    const int t = threadIdx.x;
    const long long per = (nblk + 1023) / 1024;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long long i = t * per; i < min(nblk, (t + 1) * per); ++i)
        for (int k = 0; k < 7; ++k) {
            if (!((rows >> k) & 1u)) continue;
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
        out[1] = ((rows >> 1) & 1u) ? sqrt(red[1][0]) / (sqrt(red[2][0]) + 1e-14) : 0.0;
        out[2] = ((rows >> 3) & 1u) ? sqrt(red[3][0]) / (sqrt(red[4][0]) + 1e-14) : 0.0;
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

// Linear mixing fused with the self-energy residual of the unmixed pairs: a
// group is one pair, or two pairs whose sum is also a residual term (plus
// and minus, whose sum is the embedding). Grid (kMixChunks, groups); each
// block strides over a fixed element set, so the partial sums and their
// reduction order are fixed.
constexpr int kMixChunks = 512;
constexpr int kMaxResPairs = 32;
struct MixResArgs {
    c64* old_[kMaxResPairs];
    const c64* cand[kMaxResPairs];
    int ga[kMaxResPairs], gb[kMaxResPairs], gt[kMaxResPairs];  // pairs and sum term of each group
    int ngroups;
};

__global__ __launch_bounds__(256) void k_mix_residual(long long n, MixResArgs a, double alpha, double* part) {
    __shared__ double red[6][256];
    const int g = blockIdx.y;
    const int pa = a.ga[g], pb = a.gb[g];
    c64* __restrict__ oa = a.old_[pa];
    const c64* __restrict__ ca = a.cand[pa];
    c64* __restrict__ ob = pb >= 0 ? a.old_[pb] : nullptr;
    const c64* __restrict__ cb = pb >= 0 ? a.cand[pb] : nullptr;
    const double beta = 1.0 - alpha;
    double acc[6] = {0, 0, 0, 0, 0, 0};  // |c-o|^2, |c|^2 of a, b and of the sum
    const long long stride = (long long)kMixChunks * 256;
    // Two elements per iteration, all loads first (more bytes in flight);
    // the sums still run over i, i + stride, ... in order.
    auto step = [&](long long i, c64 o1, c64 c1, c64 o2, c64 c2) {
        const double dx = c1.x - o1.x, dy = c1.y - o1.y;
        acc[0] += dx * dx + dy * dy;
        acc[1] += c1.x * c1.x + c1.y * c1.y;
        if (alpha != 0.0) oa[i] = mix1(o1, c1, alpha, beta);
        if (pb >= 0) {
            const double ex = c2.x - o2.x, ey = c2.y - o2.y;
            acc[2] += ex * ex + ey * ey;
            acc[3] += c2.x * c2.x + c2.y * c2.y;
            // the sum term: (c1 + c2) - (o1 + o2), as the stored sums would give it
            const double sx = (c1.x + c2.x) - (o1.x + o2.x), sy = (c1.y + c2.y) - (o1.y + o2.y);
            const double tx = c1.x + c2.x, ty = c1.y + c2.y;
            acc[4] += sx * sx + sy * sy;
            acc[5] += tx * tx + ty * ty;
            if (alpha != 0.0) ob[i] = mix1(o2, c2, alpha, beta);
        }
    };
    const c64 z = cmk(0.0, 0.0);
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += 2 * stride) {
        const long long j = i + stride;
        const bool hj = j < n;
        const c64 o1 = oa[i], c1 = ca[i];
        const c64 o2 = pb >= 0 ? ob[i] : z, c2 = pb >= 0 ? cb[i] : z;
        const c64 p1 = hj ? oa[j] : z, q1 = hj ? ca[j] : z;
        const c64 p2 = (hj && pb >= 0) ? ob[j] : z, q2 = (hj && pb >= 0) ? cb[j] : z;
        step(i, o1, c1, o2, c2);
        if (hj) step(j, p1, q1, p2, q2);
    }
    const int nacc = pb >= 0 ? 6 : 2;
    for (int k = 0; k < nacc; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < nacc; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // part[(2 t + {0,1}) * kMixChunks + chunk], t the residual term
        part[(2 * pa + 0) * kMixChunks + blockIdx.x] = red[0][0];
        part[(2 * pa + 1) * kMixChunks + blockIdx.x] = red[1][0];
        if (pb >= 0) {
            part[(2 * pb + 0) * kMixChunks + blockIdx.x] = red[2][0];
            part[(2 * pb + 1) * kMixChunks + blockIdx.x] = red[3][0];
            part[(2 * a.gt[g] + 0) * kMixChunks + blockIdx.x] = red[4][0];
            part[(2 * a.gt[g] + 1) * kMixChunks + blockIdx.x] = red[5][0];
        }
    }
}

// One block per term: ratio ||c - o|| / (||c|| + 1e-14), Inf when non-finite;
// with raw != NULL also the sums themselves (raw[2t], raw[2t+1]).
// The same mixing pass grouped by total: group t mixes its member pairs (in
// summation order) and writes totals[t] = the left-fold sum of their mixed
// values, i.e. what the family sum of the next GF solve would load
// (fam[0] + fam[1] + ...), so that solve reads one family. A couple (the
// plus/minus embedding pair) lies inside one group. Residual partials as in
// k_mix_residual, accumulated over the same elements in the same order.
constexpr int kMaxTotals = 4;
struct MixTotArgs {
    c64* old_[kMaxResPairs];
    const c64* cand[kMaxResPairs];
    int mem[kMaxTotals][kMaxFamilies];  // member pairs of each total
    int nmem[kMaxTotals];
    int ca[kMaxTotals], cb[kMaxTotals], ct[kMaxTotals];  // couple: member positions and term, -1 none
    c64* tot[kMaxTotals];
    // kTotAccumulate: the fold starts from the total's current value;
    // kTotCandidate: the total sums the members' candidate values and
    // neither mixes them nor measures them (another total owns the pairs)
    int flags[kMaxTotals];
    // Bit p: the candidate of pair p holds one momentum's columns (NE x Nb^2),
    // read for every momentum of the NE x Nk x Nb^2 field (n < 2^31).
    unsigned cmask;
    int NE, Nk;
};
constexpr int kTotAccumulate = 1;
constexpr int kTotCandidate = 2;

__global__ __launch_bounds__(256) void k_mix_totals(long long n, MixTotArgs a, double alpha, double* part) {
    constexpr int kAcc = 2 * kMaxFamilies + 2;
    __shared__ double red[kAcc][256];
    const int t = blockIdx.y;
    const int nm = a.nmem[t], pa = a.ca[t], pb = a.cb[t];
    const bool accumulate = a.flags[t] & kTotAccumulate, cand_only = a.flags[t] & kTotCandidate;
    c64* __restrict__ tot = a.tot[t];
    const double beta = 1.0 - alpha;
    double acc[kAcc];
#pragma unroll
    for (int k = 0; k < kAcc; ++k) acc[k] = 0.0;
    const long long stride = (long long)kMixChunks * 256;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        c64 o[kMaxFamilies], c[kMaxFamilies];
        long long ci = i;  // the element of a one-momentum candidate
        if (a.cmask) {
            const unsigned ii = (unsigned)i, col = ii / (unsigned)a.NE;
            ci = (long long)(ii - col * (unsigned)a.NE) + (long long)a.NE * (col / (unsigned)a.Nk);
        }
#pragma unroll
        for (int f = 0; f < kMaxFamilies; ++f)  // all loads first
            if (f < nm) {
                const int pm = a.mem[t][f];
                if (!cand_only) o[f] = a.old_[pm][i];
                c[f] = a.cand[pm][((a.cmask >> pm) & 1u) ? ci : i];
            }
        const c64 prior = accumulate ? tot[i] : cmk(0.0, 0.0);
        c64 sum = cmk(0.0, 0.0), o1 = sum, c1 = sum, o2 = sum, c2 = sum;
        if (cand_only) {
#pragma unroll
            for (int f = 0; f < kMaxFamilies; ++f)
                if (f < nm) sum = (f == 0 && !accumulate) ? c[f] : cadd(f == 0 ? prior : sum, c[f]);
            tot[i] = sum;
            continue;
        }
#pragma unroll
        for (int f = 0; f < kMaxFamilies; ++f)
            if (f < nm) {
                const double dx = c[f].x - o[f].x, dy = c[f].y - o[f].y;
                acc[2 * f] += dx * dx + dy * dy;
                acc[2 * f + 1] += c[f].x * c[f].x + c[f].y * c[f].y;
                const c64 v = alpha != 0.0 ? mix1(o[f], c[f], alpha, beta) : o[f];
                if (alpha != 0.0) a.old_[a.mem[t][f]][i] = v;
                sum = (f == 0 && !accumulate) ? v : cadd(f == 0 ? prior : sum, v);
                if (f == pa) {
                    o1 = o[f];
                    c1 = c[f];
                }
                if (f == pb) {
                    o2 = o[f];
                    c2 = c[f];
                }
            }
        if (pb >= 0) {
            const double sx = (c1.x + c2.x) - (o1.x + o2.x), sy = (c1.y + c2.y) - (o1.y + o2.y);
            const double tx = c1.x + c2.x, ty = c1.y + c2.y;
            acc[2 * kMaxFamilies] += sx * sx + sy * sy;
            acc[2 * kMaxFamilies + 1] += tx * tx + ty * ty;
        }
        tot[i] = sum;
    }
#pragma unroll
    for (int k = 0; k < kAcc; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int k = 0; k < kAcc; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0 && !cand_only) {
        for (int f = 0; f < nm; ++f) {
            const int p = a.mem[t][f];
            part[(2 * p + 0) * kMixChunks + blockIdx.x] = red[2 * f][0];
            part[(2 * p + 1) * kMixChunks + blockIdx.x] = red[2 * f + 1][0];
        }
        if (pb >= 0) {
            part[(2 * a.ct[t] + 0) * kMixChunks + blockIdx.x] = red[2 * kMaxFamilies][0];
            part[(2 * a.ct[t] + 1) * kMixChunks + blockIdx.x] = red[2 * kMaxFamilies + 1][0];
        }
    }
}

__global__ __launch_bounds__(kMixChunks) void k_mix_residual_reduce(const double* part, double* ratios, double* raw) {
    __shared__ double red[2][kMixChunks];
    const int t = blockIdx.x;
    red[0][threadIdx.x] = part[(2 * t + 0) * kMixChunks + threadIdx.x];
    red[1][threadIdx.x] = part[(2 * t + 1) * kMixChunks + threadIdx.x];
    __syncthreads();
    for (int w = kMixChunks / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double d = red[0][0], c = red[1][0];
        const double v = sqrt(d) / (sqrt(c) + 1e-14);
        ratios[t] = (isfinite(d) && isfinite(c) && isfinite(v)) ? v : INFINITY;
        if (raw) {
            raw[2 * t] = d;
            raw[2 * t + 1] = c;
        }
    }
}

__global__ void k_fill_zero(long long n, c64* __restrict__ x) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] = cmk(0.0, 0.0);
}

__global__ void k_zero_counter(unsigned* c) {
    if (threadIdx.x == 0) *c = 0u;
}

template <int B, int W>
int launch_residual_split(qclhip_ctx* ctx, const ResArgs& a) {
    const size_t lds = (size_t)W * (2 * (size_t)TileStride<B>::value * sizeof(c64) + sizeof(EigWork<B>));
    const size_t lds_fb = res_wave_bytes<B>();
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_residual_suite_split<B, W>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (lds_fb > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_residual_fallback<B>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_fb));
    const long long nblk = (long long)a.NE * a.Nk;
    const long long ntile = (nblk + W - 1) / W;
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, ctx->stream, a.nflagged);
    hipLaunchKernelGGL((k_residual_suite_split<B, W>), dim3((unsigned)ntile), dim3(64 * B * W), lds, ctx->stream, a);
    const unsigned fb_grid = (unsigned)std::min<long long>(nblk, 1024);
    hipLaunchKernelGGL((k_residual_fallback<B>), dim3(fb_grid), dim3(64), lds_fb, ctx->stream, a);
    return launch_check(ctx, "k_residual_suite_split");
}

// The screening pass over every block, then the eigenvalue pass over the
// blocks it listed (a fixed grid that reads the list length on the device).
template <int B>
int launch_residual(qclhip_ctx* ctx, const ResArgs& a) {
    // N in 25..32, prescreened, A formed from G^R: four waves per block, two
    // blocks per workgroup (measured at N = 32, NE = 4096, Nk = 128: 74.8 ms
    // for the state pass with the observables against 82.7 ms; four blocks
    // per workgroup 94.9 ms; at N <= 24 the split ran slower than one wave
    // per block: 11.4 against 8.3 ms for the full suite at N = 24)
    if constexpr (B == 4) {
        if (a.raw_screen && !a.A) return launch_residual_split<B, 2>(ctx, a);
    }
    constexpr int W = res_waves<B>();
    const size_t lds = W * res_wave_bytes<B>();
    const size_t lds_fb = res_wave_bytes<B>();
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_residual_suite<B, W>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (lds_fb > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_residual_fallback<B>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_fb));
    const long long nblk = (long long)a.NE * a.Nk;
    const long long ntile = (nblk + W - 1) / W;
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, ctx->stream, a.nflagged);
    hipLaunchKernelGGL((k_residual_suite<B, W>), dim3((unsigned)ntile), dim3(64 * W), lds, ctx->stream, a);
    const unsigned fb_grid = (unsigned)std::min<long long>(nblk, 1024);
    hipLaunchKernelGGL((k_residual_fallback<B>), dim3(fb_grid), dim3(64), lds_fb, ctx->stream, a);
    return launch_check(ctx, "k_residual_suite");
}

// The r_K kernel: two tiles per wave, four waves per workgroup (64-byte
// segments) while two workgroups fit a CU's LDS.
template <int B>
int launch_residual_keldysh(qclhip_ctx* ctx, const ResArgs& a) {
    constexpr size_t wave_bytes = 2 * (size_t)TileStride<B>::value * sizeof(c64);
    constexpr int W = 4 * wave_bytes <= 81920 ? 4 : (2 * wave_bytes <= 81920 ? 2 : 1);
    const size_t lds = W * wave_bytes;
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_residual_keldysh<B, W>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long ntile = ((long long)a.NE * a.Nk + W - 1) / W;
    hipLaunchKernelGGL((k_residual_keldysh<B, W>), dim3((unsigned)ntile), dim3(64 * W), lds, ctx->stream, a);
    return launch_check(ctx, "k_residual_keldysh");
}

// Runs the suite kernel (keldysh_only = false) or the r_K kernel for the
// basis class of N, reduces the rows of `rows` and returns the 5 results
// (r_D, r_A, r_K, r_PSD, r_caus) to out_host or the deferred buffer.
int residual_run(qclhip_ctx* ctx, ResArgs& a, bool keldysh_only, unsigned rows, double* out_host) {
    const long long nblk = (long long)a.NE * a.Nk;
    // per-block values, the 5 results (8 slots), the list counter (8
    // slots) and the list of unsettled blocks
    double* v = (double*)workspace(ctx, (size_t)(7 * nblk + 16) * sizeof(double) +
                                            (keldysh_only ? 0 : (size_t)nblk * sizeof(unsigned long long)));
    if (!v) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    a.out = v;
    a.nflagged = reinterpret_cast<unsigned*>(v + 7 * nblk + 8);
    a.flagged = reinterpret_cast<unsigned long long*>(v + 7 * nblk + 16);
    int rc;
    const int N = a.N;
    const int B = N <= 8 ? 1 : (N <= 16 ? 2 : (N <= 24 ? 3 : (N <= 32 ? 4 : (N <= 40 ? 5 : (N <= 48 ? 6 : 0)))));
    switch (B) {
        case 1: rc = keldysh_only ? launch_residual_keldysh<1>(ctx, a) : launch_residual<1>(ctx, a); break;
        case 2: rc = keldysh_only ? launch_residual_keldysh<2>(ctx, a) : launch_residual<2>(ctx, a); break;
        case 3: rc = keldysh_only ? launch_residual_keldysh<3>(ctx, a) : launch_residual<3>(ctx, a); break;
        case 4: rc = keldysh_only ? launch_residual_keldysh<4>(ctx, a) : launch_residual<4>(ctx, a); break;
        case 5: rc = keldysh_only ? launch_residual_keldysh<5>(ctx, a) : launch_residual<5>(ctx, a); break;
        case 6: rc = keldysh_only ? launch_residual_keldysh<6>(ctx, a) : launch_residual<6>(ctx, a); break;
        default: return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");
    }
    if (rc) return rc;
    double* res = v + 7 * nblk;
    hipLaunchKernelGGL(k_residual_reduce, dim3(1), dim3(1024), 0, ctx->stream, nblk, v, res, rows);
    rc = launch_check(ctx, "k_residual_reduce");
    if (rc) return rc;
    bool deferred = false;
    rc = defer_result(ctx, res, 5, &deferred);
    if (rc || deferred) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(out_host, res, 5 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return qclhip_ctx_synchronize(ctx);
}

}  // namespace qclhip

using namespace qclhip;

extern "C" {

static int residual_args(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, const double* GR,
                         const double* GL, const double* GG, const double* A, int nfam, const double* const* famR,
                         const double* const* famL, const double* const* famG, int ncand,
                         const double* const* candL, ResArgs* a) {
    if (nfam < 0 || nfam > kMaxFamilies || ncand < 0 || ncand > kMaxFamilies)
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "1..8 families and candidate families");
    *a = ResArgs{};
    a->NE = grid->NE;
    a->Nk = grid->Nk;
    a->N = grid->Nb;
    a->eps = grid->eps;
    a->h = (const c64*)h;
    a->GR = (const c64*)GR;
    a->GL = (const c64*)GL;
    a->GG = (const c64*)GG;
    a->A = (const c64*)A;
    for (int f = 0; f < nfam; ++f) {
        if (!famR[f] || !famL[f] || !famG[f]) return fail(ctx, QCLHIP_ERR_ARGUMENT, "null family field");
        a->famR[f] = (const c64*)famR[f];
        a->famL[f] = (const c64*)famL[f];
        a->famG[f] = (const c64*)famG[f];
    }
    for (int f = 0; f < ncand; ++f) {
        if (!candL[f]) return fail(ctx, QCLHIP_ERR_ARGUMENT, "null candidate field");
        a->cand[f] = (const c64*)candL[f];
    }
    a->nfam = nfam;
    a->ncand = ncand;
    return QCLHIP_OK;
}

int qclhip_residual_suite(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, const double* GR,
                          const double* GL, const double* GG, const double* A, int nfam,
                          const double* const* famR, const double* const* famL, const double* const* famG,
                          int ncand, const double* const* candL, double* out_host) {
    if (!ctx || !grid || !h || !GR || !GL || !GG || !famR || !famL || !famG || !candL || !out_host)
        return QCLHIP_ERR_ARGUMENT;
    if (nfam < 1 || ncand < 1) return fail(ctx, QCLHIP_ERR_ARGUMENT, "1..8 families and candidate families");
    ResArgs a;
    int rc = residual_args(ctx, grid, h, GR, GL, GG, A, nfam, famR, famL, famG, ncand, candL, &a);
    if (rc) return rc;
    a.mode = 0;
    return residual_run(ctx, a, false, 0x7fu, out_host);
}

int qclhip_residual_suite_prescreened(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, const double* GR,
                                      const double* GL, const double* GG, const double* A, int nfam,
                                      const double* const* famR, const double* const* famL,
                                      const double* const* famG, int ncand, const double* const* candL,
                                      const unsigned char* raw_screen, double* out_host) {
    if (!ctx || !grid || !h || !GR || !GL || !GG || !famR || !famL || !famG || !candL || !raw_screen || !out_host)
        return QCLHIP_ERR_ARGUMENT;
    if (nfam < 1 || ncand < 1) return fail(ctx, QCLHIP_ERR_ARGUMENT, "1..8 families and candidate families");
    ResArgs a;
    int rc = residual_args(ctx, grid, h, GR, GL, GG, A, nfam, famR, famL, famG, ncand, candL, &a);
    if (rc) return rc;
    a.mode = 0;
    a.raw_screen = raw_screen;
    return residual_run(ctx, a, false, 0x7fu, out_host);
}

int qclhip_residual_state_prescreened(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, const double* GR,
                                      const double* GL, const double* GG, const double* A, int nfam,
                                      const double* const* famR, const double* const* famL,
                                      const double* const* famG, const unsigned char* raw_screen,
                                      double* out_host) {
    if (!ctx || !grid || !h || !GR || !GL || !GG || !famR || !famL || !famG || !raw_screen || !out_host)
        return QCLHIP_ERR_ARGUMENT;
    if (nfam < 1) return fail(ctx, QCLHIP_ERR_ARGUMENT, "1..8 families");
    ResArgs a;
    int rc = residual_args(ctx, grid, h, GR, GL, GG, A, nfam, famR, famL, famG, 0, nullptr, &a);
    if (rc) return rc;
    a.mode = 1;
    a.raw_screen = raw_screen;
    return residual_run(ctx, a, false, 0x67u, out_host);  // rows 0, 1, 2, 5, 6
}

int qclhip_residual_state(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, const double* GR,
                          const double* GL, const double* GG, const double* A, int nfam,
                          const double* const* famR, const double* const* famL, const double* const* famG,
                          double* out_host) {
    if (!ctx || !grid || !h || !GR || !GL || !GG || !famR || !famL || !famG || !out_host)
        return QCLHIP_ERR_ARGUMENT;
    if (nfam < 1) return fail(ctx, QCLHIP_ERR_ARGUMENT, "1..8 families");
    ResArgs a;
    int rc = residual_args(ctx, grid, h, GR, GL, GG, A, nfam, famR, famL, famG, 0, nullptr, &a);
    if (rc) return rc;
    a.mode = 1;
    return residual_run(ctx, a, false, 0x67u, out_host);  // rows 0, 1, 2, 5, 6
}

int qclhip_residual_keldysh(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GR, const double* GL,
                            int ncand, const double* const* candL, double* out_host) {
    if (!ctx || !grid || !GR || !GL || !candL || !out_host) return QCLHIP_ERR_ARGUMENT;
    if (ncand < 1) return fail(ctx, QCLHIP_ERR_ARGUMENT, "1..8 candidate families");
    ResArgs a;
    int rc = residual_args(ctx, grid, nullptr, GR, GL, nullptr, nullptr, 0, nullptr, nullptr, nullptr, ncand,
                           candL, &a);
    if (rc) return rc;
    a.mode = 2;
    return residual_run(ctx, a, true, 0x18u, out_host);  // rows 3, 4
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

static int mix_fields_residual_impl(qclhip_ctx* ctx, int64_t n, int npairs, double* const* old_,
                                    const double* const* cand, double alpha, const int* partner, double* ratios_host,
                                    double* raw_host) {
    if (!ctx || !old_ || !cand || !partner || !(ratios_host || raw_host) || n < 0) return QCLHIP_ERR_ARGUMENT;
    if (npairs < 1 || npairs > kMaxResPairs) return fail(ctx, QCLHIP_ERR_ARGUMENT, "1..32 mixing pairs");
    if (!(alpha >= 0 && alpha <= 1)) return fail(ctx, QCLHIP_ERR_ARGUMENT, "mixing factor must lie in [0,1]");
    MixResArgs a{};
    int nterms = npairs;
    a.ngroups = 0;
    for (int p = 0; p < npairs; ++p) {
        if (!old_[p] || !cand[p]) return QCLHIP_ERR_ARGUMENT;
        const int q = partner[p];
        if (q >= npairs || q == p || (q >= 0 && partner[q] != p))
            return fail(ctx, QCLHIP_ERR_ARGUMENT, "partner must be a symmetric pairing of distinct pairs");
        a.old_[p] = (c64*)old_[p];
        a.cand[p] = (const c64*)cand[p];
        if (q >= 0 && q < p) continue;  // grouped with its lower partner
        a.ga[a.ngroups] = p;
        a.gb[a.ngroups] = q;
        a.gt[a.ngroups] = q >= 0 ? nterms++ : -1;
        ++a.ngroups;
    }
    double* part = (double*)workspace(ctx, ((size_t)2 * nterms * kMixChunks + 3 * (size_t)nterms) * sizeof(double));
    if (!part) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    double* ratios = part + (size_t)2 * nterms * kMixChunks;
    double* raw = ratios + nterms;
    hipLaunchKernelGGL(k_mix_residual, dim3(kMixChunks, a.ngroups), dim3(256), 0, ctx->stream, (long long)n, a,
                       alpha, part);
    hipLaunchKernelGGL(k_mix_residual_reduce, dim3(nterms), dim3(kMixChunks), 0, ctx->stream, part, ratios,
                       raw_host ? raw : nullptr);
    int rc = launch_check(ctx, "mix with self-energy residual");
    if (rc) return rc;
    if (raw_host) {
        QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(raw_host, raw, 2 * nterms * sizeof(double), hipMemcpyDeviceToHost,
                                           ctx->stream));
        return qclhip_ctx_synchronize(ctx);
    }
    bool deferred = false;
    rc = defer_result(ctx, ratios, nterms, &deferred);
    if (rc || deferred) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(ratios_host, ratios, nterms * sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->stream));
    return qclhip_ctx_synchronize(ctx);
}

int qclhip_mix_fields_residual(qclhip_ctx* ctx, int64_t n, int npairs, double* const* old_,
                               const double* const* cand, double alpha, const int* partner, double* ratios_host) {
    if (!ratios_host) return QCLHIP_ERR_ARGUMENT;
    return mix_fields_residual_impl(ctx, n, npairs, old_, cand, alpha, partner, ratios_host, nullptr);
}

static int mix_totals_impl(qclhip_ctx* ctx, int64_t n, int npairs, double* const* old_, const double* const* cand,
                           double alpha, const int* partner, int ntot, const int* tot_count, const int* tot_members,
                           const int* tot_flags, double* const* totals, double* ratios_host,
                           const qclhip_grid* grid = nullptr, const int* compact = nullptr) {
    if (!ctx || !old_ || !cand || !partner || !tot_count || !tot_members || !totals || !ratios_host || n < 0)
        return QCLHIP_ERR_ARGUMENT;
    if (npairs < 1 || npairs > kMaxResPairs) return fail(ctx, QCLHIP_ERR_ARGUMENT, "1..32 mixing pairs");
    if (ntot < 1 || ntot > kMaxTotals) return fail(ctx, QCLHIP_ERR_ARGUMENT, "1..4 totals");
    if (!(alpha >= 0 && alpha <= 1)) return fail(ctx, QCLHIP_ERR_ARGUMENT, "mixing factor must lie in [0,1]");
    MixTotArgs a{};
    if (compact) {
        if (!grid) return QCLHIP_ERR_ARGUMENT;
        if (n != (int64_t)grid->NE * grid->Nk * grid->Nb * grid->Nb)
            return fail(ctx, QCLHIP_ERR_DIMENSION, "the fields do not match the grid");
        if (n >= ((int64_t)1 << 31)) return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "one-momentum candidates need n < 2^31");
        for (int p = 0; p < npairs; ++p)
            if (compact[p]) a.cmask |= 1u << p;
        a.NE = grid->NE;
        a.Nk = grid->Nk;
    }
    std::vector<int> owner(npairs, -1), pos(npairs, -1);
    int k = 0;
    for (int t = 0; t < ntot; ++t) {
        if (!totals[t] || tot_count[t] < 1 || tot_count[t] > kMaxFamilies)
            return fail(ctx, QCLHIP_ERR_ARGUMENT, "each total sums 1..8 pairs");
        a.tot[t] = (c64*)totals[t];
        a.nmem[t] = tot_count[t];
        a.ca[t] = a.cb[t] = a.ct[t] = -1;
        a.flags[t] = tot_flags ? tot_flags[t] : 0;
        if (a.flags[t] & ~(kTotAccumulate | kTotCandidate))
            return fail(ctx, QCLHIP_ERR_ARGUMENT, "unknown total flag");
        for (int u = 0; u < t; ++u)
            if (a.tot[u] == a.tot[t]) return fail(ctx, QCLHIP_ERR_ARGUMENT, "totals must be distinct fields");
        for (int f = 0; f < tot_count[t]; ++f, ++k) {
            const int p = tot_members[k];
            if (p < 0 || p >= npairs) return fail(ctx, QCLHIP_ERR_ARGUMENT, "total member out of range");
            a.mem[t][f] = p;
            if (a.flags[t] & kTotCandidate) continue;  // reads candidates only; owns no pair
            if (owner[p] >= 0)
                return fail(ctx, QCLHIP_ERR_ARGUMENT, "every pair belongs to exactly one mixing total");
            owner[p] = t;
            pos[p] = f;
        }
    }
    int nterms = npairs;
    for (int p = 0; p < npairs; ++p) {
        if (!old_[p] || !cand[p]) return QCLHIP_ERR_ARGUMENT;
        if (owner[p] < 0) return fail(ctx, QCLHIP_ERR_ARGUMENT, "every pair belongs to exactly one total");
        const int q = partner[p];
        if (q >= npairs || q == p || (q >= 0 && partner[q] != p))
            return fail(ctx, QCLHIP_ERR_ARGUMENT, "partner must be a symmetric pairing of distinct pairs");
        a.old_[p] = (c64*)old_[p];
        a.cand[p] = (const c64*)cand[p];
        if (q > p) {  // couple terms numbered in ascending p, after the pairs
            const int t = owner[p];
            if (owner[q] != t || a.cb[t] >= 0)
                return fail(ctx, QCLHIP_ERR_ARGUMENT, "a couple lies inside one total, at most one per total");
            a.ca[t] = pos[p];
            a.cb[t] = pos[q];
            a.ct[t] = nterms++;
        }
    }
    double* part = (double*)workspace(ctx, ((size_t)2 * nterms * kMixChunks + (size_t)nterms) * sizeof(double));
    if (!part) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    double* ratios = part + (size_t)2 * nterms * kMixChunks;
    for (int t = 0; t < ntot; ++t)
        for (int p = 0; p < npairs; ++p)
            if ((const void*)a.tot[t] == (const void*)old_[p] || (const void*)a.tot[t] == (const void*)cand[p])
                return fail(ctx, QCLHIP_ERR_ARGUMENT, "a total must not alias a mixing pair");
    hipLaunchKernelGGL(k_mix_totals, dim3(kMixChunks, ntot), dim3(256), 0, ctx->stream, (long long)n, a, alpha, part);
    hipLaunchKernelGGL(k_mix_residual_reduce, dim3(nterms), dim3(kMixChunks), 0, ctx->stream, part, ratios, nullptr);
    int rc = launch_check(ctx, "mix with totals and self-energy residual");
    if (rc) return rc;
    bool deferred = false;
    rc = defer_result(ctx, ratios, nterms, &deferred);
    if (rc || deferred) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(ratios_host, ratios, nterms * sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->stream));
    return qclhip_ctx_synchronize(ctx);
}

int qclhip_mix_fields_residual_totals(qclhip_ctx* ctx, int64_t n, int npairs, double* const* old_,
                                      const double* const* cand, double alpha, const int* partner, int ntot,
                                      const int* tot_count, const int* tot_members, double* const* totals,
                                      double* ratios_host) {
    return mix_totals_impl(ctx, n, npairs, old_, cand, alpha, partner, ntot, tot_count, tot_members, nullptr,
                           totals, ratios_host);
}

int qclhip_mix_fields_residual_accumulate(qclhip_ctx* ctx, int64_t n, int npairs, double* const* old_,
                                          const double* const* cand, double alpha, const int* partner, int ntot,
                                          const int* tot_count, const int* tot_members, const int* tot_flags,
                                          double* const* totals, double* ratios_host) {
    if (!tot_flags) return QCLHIP_ERR_ARGUMENT;
    return mix_totals_impl(ctx, n, npairs, old_, cand, alpha, partner, ntot, tot_count, tot_members, tot_flags,
                           totals, ratios_host);
}

int qclhip_mix_fields_residual_compact(qclhip_ctx* ctx, const qclhip_grid* grid, int64_t n, int npairs,
                                       double* const* old_, const double* const* cand, double alpha,
                                       const int* partner, int ntot, const int* tot_count, const int* tot_members,
                                       const int* tot_flags, double* const* totals, const int* compact,
                                       double* ratios_host) {
    if (!grid || !tot_flags || !compact) return QCLHIP_ERR_ARGUMENT;
    return mix_totals_impl(ctx, n, npairs, old_, cand, alpha, partner, ntot, tot_count, tot_members, tot_flags,
                           totals, ratios_host, grid, compact);
}

int qclhip_mix_fields_change(qclhip_ctx* ctx, int64_t n, int npairs, double* const* old_, const double* const* cand,
                             double alpha, double* sums_host) {
    if (!ctx || !sums_host || npairs < 1 || npairs > kMaxResPairs) return QCLHIP_ERR_ARGUMENT;
    std::vector<int> partner(npairs, -1);
    std::vector<double> raw(2 * (size_t)npairs);
    int rc = mix_fields_residual_impl(ctx, n, npairs, old_, cand, alpha, partner.data(), nullptr, raw.data());
    if (rc) return rc;
    double d = 0.0, c = 0.0;
    for (int p = 0; p < npairs; ++p) {  // pair order
        d += raw[2 * p];
        c += raw[2 * p + 1];
    }
    sums_host[0] = d;
    sums_host[1] = c;
    return QCLHIP_OK;
}

int qclhip_fill_zero(qclhip_ctx* ctx, int64_t n, double* X) {
    if (!ctx || !X || n < 0) return QCLHIP_ERR_ARGUMENT;
    long long blocks = std::min<long long>(((long long)n + 255) / 256, (long long)std::max(ctx->sm_count, 1) * 32);
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::max<long long>(1, blocks)), dim3(256), 0, ctx->stream,
                       (long long)n, (c64*)X);
    return launch_check(ctx, "k_fill_zero");
}

}  // extern "C"
