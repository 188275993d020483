// Per-block (e,m) tile machinery shared by the Dyson/Keldysh kernels:
// argument block, LDS shapes, XCD-aware tile order, coalesced tile transfers
// and the 8-lane DPP argmax of the pivot search.
#pragma once

#include "common.hpp"

namespace qclhip {

// LDS stride between the per-wave tiles of one workgroup, in elements: the
// (8B) x (8B+1) tile rounded up so that the stride is 2 (mod 8). Cooperative
// transfers put the same (i,j) of consecutive energies in consecutive lanes;
// with this stride those addresses start 8 dwords (complex double) or 4
// dwords (complex float) apart instead of falling into the same LDS banks.
template <int B>
struct TileStride {
    static constexpr int value = ((8 * B) * (8 * B + 1) + 7) / 8 * 8 + 2;
};

template <int B>
struct DKShape {
    static constexpr int NP = 8 * B;      // padded matrix size
    static constexpr int LD = NP + 1;     // LDS row stride (complex elements)
    static constexpr size_t bytes =
        (size_t)2 * TileStride<B>::value * sizeof(c64) + 2 * NP * sizeof(c64) + 2 * NP * sizeof(int);
};

constexpr int kMaxFamilies = 8;
// Energies per workgroup of the FP64 Dyson kernel.
constexpr int kDysonTileWaves = 4;

struct DKArgs {
    int NE, Nk, N;
    const double* eps;
    const c64* h;
    const c64* famR[kMaxFamilies];  // summed in order: the total self-energy
    const c64* famL[kMaxFamilies];
    const c64* famG[kMaxFamilies];
    int nfamR, nfamL, nfamG;
    double eta;
    c64* GR;
    c64* GL;
    c64* GG;
    double* scale;
    c64* traces;  // [2][NE*Nk] raw traces of G< and G> blocks
    DeviceStatus* status;
    int keldysh;
    // Mixed precision (FP32 factorization + FP64 refinement): refinement
    // steps, the bound on the predicted final residual norm, and one flag
    // per block (1 = re-solve in FP64). When block_flags is set for the FP64
    // Dyson kernel it only redoes tiles that hold a flagged block.
    int refine;
    double refine_tol;
    unsigned* block_flags;
    // Optional by-products of the FP64 vector path: the summed families
    // (total self-energy) as loaded, for later passes of the iteration.
    c64* totR;
    c64* totL;
    c64* totG;
    // Elements between the (a,b) planes of every field (element (e,m,a,b) at
    // e + NE m + plane (a + N b)); 0: NE*Nk, the contiguous layout. Honoured
    // by the FP64 vector Dyson and Keldysh kernels.
    long long plane;
    // Optional: A = i (G^R - G^R^dagger), written by the FP64 vector Keldysh
    // kernel from its G^R tile.
    c64* A;
    // Energies per workgroup of the FP64 vector Dyson and Keldysh kernels
    // (0: the built-in rule; else one of the compiled shapes, chosen by the
    // autotuner), and waves per block of the vector Keldysh kernel (0: the
    // rule; S > 1 splits a block's rows over S waves, for N > 24).
    int wd, wk, ks;
    // Optional (FP64 vector Keldysh kernels): one screen word per block, the
    // Cholesky screens of the raw products -i G<raw and i G>raw as the
    // residual suite runs them (kScrRawL, kScrRawG: not settled;
    // kScrNonFinite), for the suite's prescreened mode.
    unsigned char* raw_screen;
};

// out[r][c] (lane block) = Gs[rows] * Xs[:, cols]  or  Xs[rows] * Gs[cols]^dagger
template <int B, bool ADJ>
__device__ inline void block_product(const c64* __restrict__ L, const c64* __restrict__ R, int N,
                                     int lr, int lc, c64 (&out)[B][B]) {
    constexpr int LD = DKShape<B>::LD;
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) out[r][c] = cmk(0.0, 0.0);
#pragma unroll 2
    for (int q = 0; q < N; ++q) {
        c64 lv[B], rv[B];
#pragma unroll
        for (int r = 0; r < B; ++r) lv[r] = L[(lr * B + r) * LD + q];
#pragma unroll
        for (int c = 0; c < B; ++c) rv[c] = ADJ ? R[(lc * B + c) * LD + q] : R[q * LD + lc * B + c];
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c)
                out[r][c] = ADJ ? cfmac(out[r][c], lv[r], rv[c]) : cfma(out[r][c], lv[r], rv[c]);
    }
}

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2):
// give every XCD a contiguous range of tiles so that the energies sharing a
// 128-byte line are served by one L2. Bijective for any block count.
__device__ inline long long xcd_tile(long long b, long long nb) {
    const long long q = nb >> 3, rem = nb & 7, x = b & 7, s = b >> 3;
    return x * q + (x < rem ? x : rem) + s;
}

// Element (i,j) of the summed families of one block (the reference's
// _total_family_production order: fam[0] + fam[1] + ...).
__device__ inline c64 fam_load(const c64* const* fam, int nfam, size_t idx) {
    c64 v = fam[0][idx];
    // Statically unrolled so that a family table copied into a kernel's
    // registers is never indexed dynamically (which would move it to scratch).
#pragma unroll
    for (int f = 1; f < kMaxFamilies; ++f)
        if (f < nfam) v = cadd(v, fam[f][idx]);
    return v;
}

// Cooperative tile transfer between HBM and the per-wave LDS matrices of a
// workgroup whose W waves own the W consecutive energies blk0 .. blk0+W-1 of
// one momentum. Thread f handles element (i,j) of energy w with
// f = w + W*(i + N*j): consecutive threads touch consecutive energies, so one
// instruction moves contiguous 16*W-byte segments. Wave w's matrix sits at
// lds + w*wstride, element (i,j) at i*LD + j. T: threads of the workgroup
// (64 W, or 64 S W when S waves share a block).
template <int W, int LD, int T = 64 * W>
__device__ inline void tile_load(c64* lds, int wstride, const c64* const* fam, int nfam, long long blk0,
                                 int nvalid, int N, size_t S) {
    const int total = N * N * W;
    int f = threadIdx.x;
    const int w = f % W;
    int ij = f / W;
    int i = ij % N, j = ij / N;
    for (; f < total; f += T) {
        if (w < nvalid) lds[w * wstride + i * LD + j] = fam_load(fam, nfam, blk0 + w + S * (i + (size_t)N * j));
        i += T / W;
        while (i >= N) {
            i -= N;
            ++j;
        }
    }
}

template <int W, int LD, bool ACC = false, int T = 64 * W>
__device__ inline void tile_store(const c64* lds, int wstride, c64* dst, long long blk0, unsigned wave_mask, int N,
                                  size_t S) {
    const int total = N * N * W;
    int f = threadIdx.x;
    const int w = f % W;
    int ij = f / W;
    int i = ij % N, j = ij / N;
    const bool on = (wave_mask >> w) & 1u;
    for (; f < total; f += T) {
        if (on) {
            const size_t idx = blk0 + w + S * (i + (size_t)N * j);
            const c64 v = lds[w * wstride + i * LD + j];
            dst[idx] = ACC ? cadd(dst[idx], v) : v;
        }
        i += T / W;
        while (i >= N) {
            i -= N;
            ++j;
        }
    }
}

// A = i (G - G^dagger) of the W per-wave G tiles, stored with the thread
// mapping of tile_store (the arithmetic of k_spectral).
template <int W, int LD, int T = 64 * W>
__device__ inline void tile_store_spectral(const c64* lds, int wstride, c64* dst, long long blk0, unsigned wave_mask,
                                           int N, size_t S) {
    const int total = N * N * W;
    int f = threadIdx.x;
    const int w = f % W;
    int ij = f / W;
    int i = ij % N, j = ij / N;
    const bool on = (wave_mask >> w) & 1u;
    for (; f < total; f += T) {
        if (on) {
            const c64 g = lds[w * wstride + i * LD + j], gt = lds[w * wstride + j * LD + i];
            const double dx = g.x - gt.x, dy = g.y + gt.y;  // G - conj(G^T)
            dst[blk0 + w + S * (i + (size_t)N * j)] = cmk(-dy, dx);
        }
        i += T / W;
        while (i >= N) {
            i -= N;
            ++j;
        }
    }
}

// 8-lane argmax (largest |Re|+|Im|, then smallest row) over the lanes of one
// column group, by DPP lane permutes: xor 1, xor 2 (quad_perm) and the
// half-row mirror (lane i <-> 7-i of each 8 lanes).
template <int CTRL>
__device__ inline void dpp_argmax_step(double& mag, int& idx) {
    long long bits = __double_as_longlong(mag);
    // Every lane reads a lane of its own row (quad_perm, row_half_mirror), so
    // the old value of the destination is never used.
    int lo = __builtin_amdgcn_mov_dpp((int)(bits & 0xffffffffll), CTRL, 0xF, 0xF, true);
    int hi = __builtin_amdgcn_mov_dpp((int)(bits >> 32), CTRL, 0xF, 0xF, true);
    int oi = __builtin_amdgcn_mov_dpp(idx, CTRL, 0xF, 0xF, true);
    double om = __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    // Larger magnitude, ties to the lower row; selects, no divergent branch.
    const bool take = (om > mag) | ((om == mag) & (oi < idx));
    mag = take ? om : mag;
    idx = take ? oi : idx;
}

__device__ inline void argmax8(double& mag, int& idx) {
    dpp_argmax_step<0xB1>(mag, idx);   // quad_perm [1,0,3,2]
    dpp_argmax_step<0x4E>(mag, idx);   // quad_perm [2,3,0,1]
    dpp_argmax_step<0x141>(mag, idx);  // row_half_mirror
}

// One DPP lane permute of a double (every lane reads a lane of its own row).
template <int CTRL>
__device__ inline double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The pivot of the column group whose lanes are 8g .. 8g+7 (g wave-uniform):
// each lane offers its first-largest candidate (mag, idx); the group's
// largest magnitude comes from a DPP max (exact: the max is one of the
// values), and the lowest lane that holds it (a ballot) gives the row: the
// first largest |Re|+|Im| in row order, as argmax8 selects it. Returns the
// magnitude; piv is wave-uniform. Magnitudes are never NaN (-1 and -inf
// mark a lane without a candidate).
__device__ inline double pivot8(double mag, int idx, int g, int& piv) {
    double mx = fmax(mag, dpp_d<0xB1>(mag));  // quad_perm [1,0,3,2]
    mx = fmax(mx, dpp_d<0x4E>(mx));           // quad_perm [2,3,0,1]
    mx = fmax(mx, dpp_d<0x141>(mx));          // row_half_mirror
    const double pmag = readlane_d(mx, 8 * g);
    const unsigned long long hit = __ballot(mag == pmag) & (0xffull << (8 * g));
    const int src = hit ? __ffsll((long long)hit) - 1 : 8 * g;
    piv = __builtin_amdgcn_readlane(idx, src);
    return pmag;
}

// Converting variants of the cooperative tile transfer: the LDS tile holds
// elements of type V (float2 or double2); HBM fields stay complex double.
template <typename V>
__device__ inline V to_lds(c64 z) {
    V v;
    v.x = z.x;
    v.y = z.y;
    return v;
}

template <int W, int LD, typename V>
__device__ inline void tile_load_as(V* lds, int wstride, const c64* const* fam, int nfam, long long blk0, int nvalid,
                                    int N, size_t S) {
    const int total = N * N * W;
    int f = threadIdx.x;
    const int w = f % W;
    int ij = f / W;
    int i = ij % N, j = ij / N;
    for (; f < total; f += 64 * W) {
        if (w < nvalid) lds[w * wstride + i * LD + j] = to_lds<V>(fam_load(fam, nfam, blk0 + w + S * (i + (size_t)N * j)));
        i += 64;
        while (i >= N) {
            i -= N;
            ++j;
        }
    }
}

template <int W, int LD, typename V>
__device__ inline void tile_store_from(const V* lds, int wstride, c64* dst, long long blk0, unsigned wave_mask, int N,
                                       size_t S) {
    const int total = N * N * W;
    int f = threadIdx.x;
    const int w = f % W;
    int ij = f / W;
    int i = ij % N, j = ij / N;
    const bool on = (wave_mask >> w) & 1u;
    for (; f < total; f += 64 * W) {
        if (on) {
            const V v = lds[w * wstride + i * LD + j];
            dst[blk0 + w + S * (i + (size_t)N * j)] = cmk((double)v.x, (double)v.y);
        }
        i += 64;
        while (i >= N) {
            i -= N;
            ++j;
        }
    }
}

// Batched: the workgroups of one launch cover the tiles of every instance of
// one tile class (instances listed in `cls`, tile prefix sums in `begin`).
struct BatchClass {
    const DKArgs* insts;      // device table, one entry per instance
    const int* inst;          // instances of this class
    const long long* begin;   // [n+1] tile prefix over them
    int n;
};

// Batched number functionals and paired coefficients: one workgroup per
// instance.
struct BatchScalars {
    const c64* traces;
    int NE, Nk, gs;
    const double* wE;
    const double* wk;
    double target, limit;
    PairCoefficients* coef;
    c64* GL;
    c64* GG;
    long long n;
};

// Batched launches (dyson.hip): one launch per tile class and kernel.
int batch_tile_waves(int B, bool keldysh);
int launch_dk_batch(qclhip_ctx* ctx, int B, bool keldysh, const BatchClass& c, long long total);
int launch_pair_batch(qclhip_ctx* ctx, const BatchScalars* d, int ninst, long long maxn);

// Workgroups of a persistent launch: the number resident on the device at
// once (occupancy x CUs), a multiple of 8 so that each workgroup keeps one
// XCD, and at most the number of tiles (runtime.hip).
int persistent_grid(qclhip_ctx* ctx, const void* kernel, int threads, size_t lds, long long ntiles, unsigned* grid);
// Launchers (dyson.hip, mixed.hip).
int launch_dyson_fp64(qclhip_ctx* ctx, const DKArgs& a);
int launch_keldysh_valu(qclhip_ctx* ctx, const DKArgs& a);
int launch_fused_fp64(qclhip_ctx* ctx, const DKArgs& a);  // k_dyson_keldysh
int launch_keldysh_mfma(qclhip_ctx* ctx, const DKArgs& a, bool fp32);
int launch_mixed_dyson(qclhip_ctx* ctx, const DKArgs& a);
// Tile shapes the FP64 vector kernels can take for basis size N (dyson.hip).
bool dyson_waves_possible(int N, int W);
bool keldysh_waves_possible(int N, int W);
bool keldysh_split_possible(int N, int S);  // waves per block of the vector Keldysh kernel
// out[e, col] (+)= alpha * (S^T Y)[e, col] over ncols columns (adjoint.hip).
int shift_transpose(qclhip_ctx* ctx, const qclhip_shift* s, long long ncols, const c64* Y, double alpha,
                    int accumulate, c64* out);

// Register-staged tile transfer for persistent kernels. Thread f of a
// W-wave workgroup owns, in every tile, the elements ij = f/W + 64k
// (k < EPT) of energy w = f % W: the mapping of tile_load, so loads stay
// coalesced (consecutive threads, consecutive energies). The next tile's
// inputs are fetched into registers while the current tile computes and
// then written to LDS in one step.
template <int B, int W>
struct TileMap {
    static constexpr int EPT = ((8 * B) * (8 * B) + 63) / 64;
    static constexpr int LD = 8 * B + 1;
    int w, n;
    int gidx[EPT];  // i + N j
    int lidx[EPT];  // i LD + j
    __device__ inline void init(int N) {
        w = threadIdx.x % W;
        const int first = threadIdx.x / W;
        n = 0;
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
            const int ij = first + 64 * k;
            const int i = ij % N, j = ij / N;
            const bool on = ij < N * N;
            gidx[k] = on ? i + N * j : 0;
            lidx[k] = on ? i * LD + j : 0;
            n += on ? 1 : 0;
        }
    }
    __device__ inline void fetch(c64 (&r)[EPT], const c64* const* fam, int nfam, long long blk0, int nvalid,
                                 size_t S) const {
        const bool on = w < nvalid;
#pragma unroll
        for (int k = 0; k < EPT; ++k)
            if (on && k < n) r[k] = fam_load(fam, nfam, blk0 + w + S * (size_t)gidx[k]);
    }
    __device__ inline void put(c64* lds, int wstride, const c64 (&r)[EPT], int nvalid) const {
        const bool on = w < nvalid;
#pragma unroll
        for (int k = 0; k < EPT; ++k)
            if (on && k < n) lds[w * wstride + lidx[k]] = r[k];
    }
};

// Tile d of a persistent loop: XCD-aware tile index, first energy, momentum,
// valid energies and first block.
struct TileCoord {
    int e0, m, nvalid;
    long long blk0;
};

template <int W>
__device__ inline TileCoord tile_coord(long long d, long long ntiles, int NE) {
    const int nte = (NE + W - 1) / W;
    const long long t = xcd_tile(d, ntiles);
    TileCoord c;
    c.e0 = (int)(t % nte) * W;
    c.m = (int)(t / nte);
    c.nvalid = min(W, NE - c.e0);
    c.blk0 = c.e0 + (long long)NE * c.m;
    return c;
}

// Rows [s 8 B/SPLIT, (s+1) 8 B/SPLIT) of the lane-grid product L R (ADJ:
// L R^dagger) of LDS tiles, the share of wave s of SPLIT waves working on one
// block: each lane a (B/SPLIT) x B register block, with block_product's
// per-element arithmetic (FMAs in q order), so the entries are bitwise those
// of the one-wave product.
template <int B, int SPLIT, bool ADJ>
__device__ inline void block_product_rows(const c64* __restrict__ L, const c64* __restrict__ R, int N, int s, int lr,
                                          int lc, c64 (&out)[B / SPLIT][B]) {
    constexpr int LD = DKShape<B>::LD;
    constexpr int BR = B / SPLIT;
    const int row0 = s * 8 * BR + lr * BR;
#pragma unroll
    for (int r = 0; r < BR; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) out[r][c] = cmk(0.0, 0.0);
#pragma unroll 2
    for (int q = 0; q < N; ++q) {
        c64 lv[BR], rv[B];
#pragma unroll
        for (int r = 0; r < BR; ++r) lv[r] = L[(row0 + r) * LD + q];
#pragma unroll
        for (int c = 0; c < B; ++c) rv[c] = ADJ ? R[(lc * B + c) * LD + q] : R[q * LD + lc * B + c];
#pragma unroll
        for (int r = 0; r < BR; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c)
                out[r][c] = ADJ ? cfmac(out[r][c], lv[r], rv[c]) : cfma(out[r][c], lv[r], rv[c]);
    }
}

}  // namespace qclhip
