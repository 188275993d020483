// Dyson / Keldysh kernels (FP64, vector ALUs) and the number-functional
// reductions.
//
// One (e,m) block per wavefront. The 64 lanes form an 8x8
// grid (lr = lane&7 row group, lc = lane>>3 column group); lane (lr,lc) keeps
// the B x B sub-block rows lr*B.., columns lc*B.. of the matrix in registers
// (N <= 8B, padding rows/columns are zero and never pivot).
//
// Inversion: in-place Gauss-Jordan with partial pivoting. The pivot of
// column k is the first not-yet-used row of maximal |Re|+|Im| (the LAPACK
// izamax choice of zgetrf, production_operators.jl:20-41, and of the owned
// small inverse, :84-139). Rows are never swapped: the pivot row is
// remembered and the permutation is undone when the inverse is written to
// LDS. Per step the pivot column and the pivot row travel through LDS;
// every lane scales the row by the pivot's reciprocal and does B*B complex
// FMAs on its own registers.
//
// Keldysh: with G^R in LDS, T = G^R S and G = T G^R^dagger are formed by the
// same 8x8 lane grid, each lane accumulating its B x B block.
#include <climits>
#include <cstddef>
#include <mutex>

#include "wave_eig.hpp"

namespace qclhip {

// Dyson inversion only (A1): one block per wavefront, 8x8 lane grid, B x B
// register sub-blocks, Gauss-Jordan with partial pivoting. The workgroup
// stages the W self-energy blocks through LDS with coalesced loads; during
// the elimination only the pivot column and pivot row travel through LDS.
// The inverse stays in the pivoted order in registers and is un-permuted
// on its way back to LDS, G[steprow[i]][pivrow[j]] = M[i][j] / s, then
// stored with coalesced writes.
template <int B, int W>
struct DyShape {
    static constexpr int NP = 8 * B;
    static constexpr int LD = NP + 1;
    // The pivot column and row and the two permutation vectors live in the
    // wave's own tile during the elimination (the tile is dead between the
    // load of D and the store of G), so a wave needs only its tile:
    // 4 workgroups of 4 waves per CU at N <= 24.
    static constexpr size_t per_wave = (size_t)TileStride<B>::value * sizeof(c64);
    static constexpr size_t bytes = per_wave * W;
    static_assert(2 * NP * sizeof(c64) + 2 * NP * sizeof(int) <= per_wave, "elimination buffers exceed the tile");
};

// One tile (W consecutive energies of one momentum) of the FP64 Dyson step;
// called by the single-instance, batched and fused kernels. The self-energy
// tiles are loaded into in_tiles (whose wave tiles then hold the elimination
// buffers) and the inverse is written to out_tiles (the same LDS for k_dyson;
// the G^R tiles that the fused kernel's products read), then stored.
template <int B, int W>
__device__ inline void dyson_tile(const DKArgs& p, const TileCoord& tc, c64* in_tiles, c64* out_tiles,
                                  unsigned& s_ok) {
    constexpr int NP = DyShape<B, W>::NP;
    constexpr int LD = DyShape<B, W>::LD;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long long nblk = (long long)p.NE * p.Nk;
    const int wstride = TileStride<B>::value;
    c64* Ms = in_tiles + (size_t)wave * wstride;
    c64* Os = out_tiles + (size_t)wave * wstride;
    c64* colbuf = Ms;
    c64* rowbuf = colbuf + NP;
    int* pivrow = reinterpret_cast<int*>(rowbuf + NP);
    int* steprow = pivrow + NP;
    const int N = p.N;
    const size_t S = p.plane ? (size_t)p.plane : (size_t)nblk;
    const int lr = lane & 7, lc = lane >> 3;
    const int e0 = tc.e0, m = tc.m, nvalid = tc.nvalid;
    const long long blk0 = tc.blk0;
    const int e = e0 + wave;
    const long long blk = blk0 + wave;
    if (threadIdx.x == 0) s_ok = 0;
    if (p.nfamR) tile_load<W, LD>(in_tiles, wstride, p.famR, p.nfamR, blk0, nvalid, N, S);
    __syncthreads();
    if (p.totR && p.nfamR) {
        const unsigned vmask = nvalid >= 32 ? 0xffffffffu : ((1u << nvalid) - 1u);
        tile_store<W, LD>(in_tiles, wstride, p.totR, blk0, vmask, N, S);
    }
    bool ok = wave < nvalid;
    double s = 0.0, rs = 0.0;
    c64 A[B][B];
    if (ok) {
        // D = (eps_e + i eta) I - h_m - SigmaR_em (production_operators.jl:1360-1363)
        const double ee = p.eps[e];
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const int i = lr * B + r, j = lc * B + c;
                c64 v = cmk(0.0, 0.0);
                if (i < N && j < N) {
                    const c64 hv = p.h[(size_t)m + (size_t)p.Nk * (i + (size_t)N * j)];
                    const c64 sv = p.nfamR ? Ms[i * LD + j] : cmk(0.0, 0.0);
                    const double dre = (i == j) ? ee : 0.0, dim = (i == j) ? p.eta : 0.0;
                    v = cmk(dre - hv.x - sv.x, dim - hv.y - sv.y);
                }
                A[r][c] = v;
            }
        // s = max_a sum_b |D_ab| (production_operators.jl:1364-1372)
#pragma unroll
        for (int r = 0; r < B; ++r) {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < B; ++c) acc += hypot(A[r][c].x, A[r][c].y);
            acc += shfl_xor_d(acc, 8);
            acc += shfl_xor_d(acc, 16);
            acc += shfl_xor_d(acc, 32);
            s = fmax(s, acc);
        }
        s = fmax(s, shfl_xor_d(s, 1));
        s = fmax(s, shfl_xor_d(s, 2));
        s = fmax(s, shfl_xor_d(s, 4));
        s = readlane_d(s, 0);
        ok = s > 0.0;
    }
    // Every wave has read its tile (and the totals store has read all of
    // them) before the elimination buffers overwrite it.
    __syncthreads();
    if (ok) {
        rs = 1.0 / s;
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) A[r][c] = cmk(A[r][c].x * rs, A[r][c].y * rs);
        // Padding rows and columns (index >= N) hold zeros and stay zero: a
        // padding row never wins a pivot search against a nonzero entry and
        // ties go to the lower (real) row, so the steps need no bounds
        // checks. piv, pr and ck are wave-uniform and the lane conditions are
        // applied with selects: every step is straight-line code.
        // pen[r]: 0 for an own row not yet pivoted, -inf once it has been, so
        // that its candidate magnitude |Re|+|Im| + pen never wins.
        double pen[B];
#pragma unroll
        for (int r = 0; r < B; ++r) pen[r] = 0.0;
        const int nck = (N + B - 1) / B;
        for (int ck = 0; ck < nck && ok; ++ck) {
            const bool is_c = lc == ck;
#pragma unroll
            for (int kk = 0; kk < B; ++kk) {
                const int k = ck * B + kk;
                if (k >= N) break;
                double mag = -1.0;
                int idx = lr * B;
#pragma unroll
                for (int r = 0; r < B; ++r) {
                    const double mg = cabs1(A[r][kk]) + pen[r];
                    const bool gt = mg > mag;
                    mag = gt ? mg : mag;
                    idx = gt ? lr * B + r : idx;
                }
                int piv;
                const double pmag = pivot8(mag, idx, ck, piv);
                if (!(pmag > 0.0) || piv >= N) {
                    ok = false;
                    break;
                }
                const int pg = piv / B, pr = piv - pg * B;
                const bool in_pg = lr == pg;
                if (is_c) {
#pragma unroll
                    for (int r = 0; r < B; ++r) colbuf[lr * B + r] = A[r][kk];
                }
                static_for<0, B>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    if (pr == r && in_pg) {
#pragma unroll
                        for (int c = 0; c < B; ++c) rowbuf[lc * B + c] = A[r][c];
                    }
                });
                if (lane == 0) {
                    pivrow[k] = piv;
                    steprow[piv] = k;
                }
                wave_sync();
                const c64 pv = colbuf[piv];
                c64 f[B], R[B];
#pragma unroll
                for (int r = 0; r < B; ++r) f[r] = colbuf[lr * B + r];
#pragma unroll
                for (int c = 0; c < B; ++c) R[c] = rowbuf[lc * B + c];
                const c64 inv = crecip_sel(pv);
#pragma unroll
                for (int c = 0; c < B; ++c) R[c] = cmul(R[c], inv);
                // Column k of the pivot column group becomes -f/pivot: zero
                // it before the update and let R[kk] = 1/pivot there, which
                // is also the new pivot entry.
                // (Selections are written as products with 0/1 weights, exact
                // for finite entries, so that no select of two array loads is
                // formed: such a select becomes a dynamic register index.)
                const double zc = is_c ? 0.0 : 1.0, oc = 1.0 - zc;
                R[kk] = cmk(fma(inv.x, oc, R[kk].x * zc), fma(inv.y, oc, R[kk].y * zc));
#pragma unroll
                for (int r = 0; r < B; ++r) A[r][kk] = cmk(A[r][kk].x * zc, A[r][kk].y * zc);
#pragma unroll
                for (int r = 0; r < B; ++r) pen[r] = (in_pg && pr == r) ? -INFINITY : pen[r];
#pragma unroll
                for (int r = 0; r < B; ++r)
#pragma unroll
                    for (int c = 0; c < B; ++c) A[r][c] = cfms(A[r][c], f[r], R[c]);
                // The pivot row becomes the scaled pivot row (pr is
                // wave-uniform: one register row, on the lanes of group pg).
                static_for<0, B>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    if (pr == r && in_pg) {
#pragma unroll
                        for (int c = 0; c < B; ++c) A[r][c] = R[c];
                    }
                });
                wave_sync();
            }
        }
        if (!ok && lane == 0) report_status(p.status, 1, (unsigned long long)blk + 1);
    } else if (wave < nvalid && lane == 0) {
        report_status(p.status, 1, (unsigned long long)blk + 1);
    }
    if (ok) {
        // Un-permute into LDS: G[k(i)][p(j)] = M[i][j] / s.
        int si[B], pj[B];
#pragma unroll
        for (int r = 0; r < B; ++r) si[r] = (lr * B + r < N) ? steprow[lr * B + r] : 0;
#pragma unroll
        for (int c = 0; c < B; ++c) pj[c] = (lc * B + c < N) ? pivrow[lc * B + c] : 0;
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const int i = lr * B + r, j = lc * B + c;
                if (i < N && j < N) Os[si[r] * LD + pj[c]] = cmk(A[r][c].x * rs, A[r][c].y * rs);
            }
        if (lane == 0) {
            atomicOr(&s_ok, 1u << wave);
            if (p.scale) p.scale[blk] = s;
        }
    }
    __syncthreads();
    tile_store<W, LD>(out_tiles, wstride, p.GR, blk0, s_ok, N, S);
}

// One tile per workgroup. With block_flags set (the FP64 re-solve of the
// mixed-precision path) only tiles holding a flagged block are solved.
template <int B, int W>
__global__ __launch_bounds__(64 * W) void k_dyson(DKArgs p, long long ntiles) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ unsigned s_ok;
    const TileCoord tc = tile_coord<W>(blockIdx.x, ntiles, p.NE);
    if (p.block_flags) {
        unsigned any = 0;
        for (int w = 0; w < tc.nvalid; ++w) any |= p.block_flags[tc.blk0 + w];
        if (!any) return;
    }
    c64* tiles = reinterpret_cast<c64*>(smem);
    dyson_tile<B, W>(p, tc, tiles, tiles, s_ok);
}

__device__ inline int batch_find(const BatchClass& c, long long g) {
    int lo = 0, hi = c.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (c.begin[mid] <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <int B, int W>
__global__ __launch_bounds__(64 * W) void k_dyson_batch(BatchClass c) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ unsigned s_ok;
    const long long total = c.begin[c.n];
    const long long g = xcd_tile(blockIdx.x, total);
    const int j = batch_find(c, g);
    const DKArgs p = c.insts[c.inst[j]];
    const long long t = g - c.begin[j];
    const int nte = (p.NE + W - 1) / W;
    TileCoord tc;
    tc.e0 = (int)(t % nte) * W;
    tc.m = (int)(t / nte);
    tc.nvalid = min(W, p.NE - tc.e0);
    tc.blk0 = tc.e0 + (long long)p.NE * tc.m;
    c64* tiles = reinterpret_cast<c64*>(smem);
    dyson_tile<B, W>(p, tc, tiles, tiles, s_ok);
}

// Keldysh products (A2): G<raw = G^R S< G^R^dagger and G>raw likewise, plus
// the block traces. The workgroup stages G^R and S tiles through LDS with
// coalesced loads; each lane accumulates its B x B block of T = G S and of
// T G^dagger; results return through LDS with coalesced stores.
// One Keldysh component of one block (one wavefront): Xs <- G Xs G^dagger
// with G in Gs (both LDS tiles), and the block trace to *trace when given.
template <int B>
__device__ inline void keldysh_products(c64* Gs, c64* Xs, int N, int lane, c64* trace) {
    constexpr int LD = DKShape<B>::LD;
    const int lr = lane & 7, lc = lane >> 3;
    c64 T[B][B];
    block_product<B, false>(Gs, Xs, N, lr, lc, T);
    wave_sync();
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) Xs[(lr * B + r) * LD + lc * B + c] = T[r][c];
    wave_sync();
    block_product<B, true>(Xs, Gs, N, lr, lc, T);
    wave_sync();
    c64 tr = cmk(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
            const int i = lr * B + r, j = lc * B + c;
            Xs[i * LD + j] = T[r][c];
            if (i == j && i < N) tr = cadd(tr, T[r][c]);
        }
    if (trace) {
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            tr.x += shfl_xor_d(tr.x, off);
            tr.y += shfl_xor_d(tr.y, off);
        }
        if (lane == 0) *trace = tr;
    }
}

// Keldysh products of one block shared by SPLIT waves, for the large blocks
// whose two LDS tiles per wave would hold a CU to one wave per SIMD. Wave s
// of the block computes rows [s NP/SPLIT, (s+1) NP/SPLIT) of T = G X and of
// T G^dagger, each lane a (B/SPLIT) x B register block, with block_product's
// per-element arithmetic (FMAs in q order), so G< and G> are bitwise those
// of keldysh_products. The block trace gathers the diagonal through diag and
// one lane sums it as keldysh_products does: per diagonal group of B entries
// in order, then the 64-leaf xor-butterfly tree over the lanes of the 8x8
// grid (group g sits at lane 9g, every other leaf is +0), so the traces are
// bitwise equal too. Every thread of the workgroup calls it (workgroup
// barriers); active marks the waves with a valid block.
template <int B, int SPLIT>
__device__ inline void keldysh_products_split(c64* Gs, c64* Xs, c64* diag, int N, int lane, int s, bool active,
                                              c64* trace) {
    static_assert(B % SPLIT == 0, "the rows of a block split evenly over its waves");
    constexpr int LD = DKShape<B>::LD;
    constexpr int BR = B / SPLIT;
    const int lr = lane & 7, lc = lane >> 3;
    const int row0 = s * 8 * BR + lr * BR;
    c64 T[BR][B];
    if (active) block_product_rows<B, SPLIT, false>(Gs, Xs, N, s, lr, lc, T);
    __syncthreads();  // every wave of the block has read X before T replaces it
    if (active) {
#pragma unroll
        for (int r = 0; r < BR; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) Xs[(row0 + r) * LD + lc * B + c] = T[r][c];
    }
    wave_sync();  // T G^dagger reads only this wave's rows of T
    if (active) block_product_rows<B, SPLIT, true>(Xs, Gs, N, s, lr, lc, T);
    wave_sync();
    if (active) {
#pragma unroll
        for (int r = 0; r < BR; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const int i = row0 + r, j = lc * B + c;
                Xs[i * LD + j] = T[r][c];
                if (diag && i == j) diag[i] = T[r][c];
            }
    }
    __syncthreads();
    if (trace && active && s == 0 && lane == 0) {
        c64 leaf[64];
#pragma unroll
        for (int l = 0; l < 64; ++l) leaf[l] = cmk(0.0, 0.0);
#pragma unroll
        for (int g = 0; g < 8; ++g) {
            c64 t = cmk(0.0, 0.0);
#pragma unroll
            for (int r = 0; r < B; ++r)
                if (g * B + r < N) t = cadd(t, diag[g * B + r]);
            leaf[9 * g] = t;
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1)
#pragma unroll
            for (int l = 0; l < 64; l += 2 * off) leaf[l] = cadd(leaf[l], leaf[l + off]);
        *trace = leaf[0];
    }
}

// The Keldysh products with SPLIT waves per block (W blocks per workgroup),
// no register prefetch; otherwise the k_keldysh loop.
// With SCR the first wave of each block also runs the raw-product screens
// (the residual suite's Cholesky screens on -i G<raw and i G>raw) and stores
// the block's screen word in p.raw_screen.
template <int B, int W, int SPLIT, bool SCR>
__device__ __forceinline__ void keldysh_split_body(const DKArgs& p, long long ntiles) {
    constexpr int LD = DKShape<B>::LD;
    constexpr int T = 64 * SPLIT * W;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ c64 diag[W][8 * B];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int b = wave / SPLIT, s = wave % SPLIT;
    const long long nblk = (long long)p.NE * p.Nk;
    const int wstride = TileStride<B>::value;
    c64* G0 = reinterpret_cast<c64*>(smem);
    c64* X0 = G0 + (size_t)W * wstride;
    c64* Gs = G0 + (size_t)b * wstride;
    c64* Xs = X0 + (size_t)b * wstride;
    EigWork<B>* ew = reinterpret_cast<EigWork<B>*>(X0 + (size_t)W * wstride) + b;  // SCR only
    const int N = p.N;
    const size_t S = p.plane ? (size_t)p.plane : (size_t)nblk;
    const c64* gr_fam[1] = {p.GR};
    for (long long d = blockIdx.x; d < ntiles; d += gridDim.x) {
        const TileCoord tc = tile_coord<W>(d, ntiles, p.NE);
        const int nvalid = tc.nvalid;
        const long long blk0 = tc.blk0;
        const long long blk = blk0 + b;
        const bool active = b < nvalid;
        unsigned sbits = 0;
        const unsigned mask = (nvalid >= 32) ? 0xffffffffu : ((1u << nvalid) - 1u);
        __syncthreads();  // the previous tile's LDS is free
        tile_load<W, LD, T>(G0, wstride, gr_fam, 1, blk0, nvalid, N, S);
        tile_load<W, LD, T>(X0, wstride, p.famL, p.nfamL, blk0, nvalid, N, S);
        __syncthreads();
        if (p.A) tile_store_spectral<W, LD, T>(G0, wstride, p.A, blk0, mask, N, S);  // G0 stays read-only
        for (int comp = 0; comp < 2; ++comp) {
            c64* dst = comp ? p.GG : p.GL;
            if (comp == 1) {
                tile_load<W, LD, T>(X0, wstride, p.famG, p.nfamG, blk0, nvalid, N, S);
                __syncthreads();
            }
            c64* tot = comp ? p.totG : p.totL;
            if (tot) {
                tile_store<W, LD, false, T>(X0, wstride, tot, blk0, mask, N, S);
                __syncthreads();  // the products overwrite the tiles
            }
            keldysh_products_split<B, SPLIT>(Gs, Xs, diag[b], N, lane, s, active,
                                             p.traces ? p.traces + (size_t)comp * nblk + blk : nullptr);
            if constexpr (SCR) {
                if (active && s == 0) {
                    bool u = false;
                    const double v = psd_screen_tile<B, LD>(Xs, N, lane, ew, 0.0, comp ? 1.0 : -1.0, u);
                    sbits |= (u ? (comp ? kScrRawG : kScrRawL) : 0u) | (isinf(v) ? kScrNonFinite : 0u);
                    if (comp == 1 && lane == 0) p.raw_screen[blk] = (unsigned char)sbits;
                }
            }
            tile_store<W, LD, false, T>(X0, wstride, dst, blk0, mask, N, S);
            if (comp == 0) __syncthreads();  // before S> replaces the tile
        }
    }
}

template <int B, int W, int SPLIT>
__global__ __launch_bounds__(64 * SPLIT * W) void k_keldysh_split(DKArgs p, long long ntiles) {
    keldysh_split_body<B, W, SPLIT, false>(p, ntiles);
}

// The screening variant, held to four waves per SIMD (the screens would
// otherwise raise the register count above 128 and cost a wave per SIMD).
template <int B, int W, int SPLIT>
__global__ __launch_bounds__(64 * SPLIT * W) __attribute__((amdgpu_waves_per_eu(4))) void k_keldysh_split_scr(
    DKArgs p, long long ntiles) {
    keldysh_split_body<B, W, SPLIT, true>(p, ntiles);
}

// With SCR each wave also runs the raw-product screens of its block (the
// residual suite's Cholesky screens on -i G<raw and i G>raw, on the product
// tile before it is stored) and stores the block's screen word in
// p.raw_screen.
template <int B, int W, bool SCR = false>
__global__ __launch_bounds__(64 * W) void k_keldysh(DKArgs p, long long ntiles) {
    constexpr int NP = DKShape<B>::NP;
    constexpr int LD = DKShape<B>::LD;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long long nblk = (long long)p.NE * p.Nk;
    const int wstride = TileStride<B>::value;
    c64* G0 = reinterpret_cast<c64*>(smem);
    c64* X0 = G0 + (size_t)W * wstride;
    c64* Gs = G0 + (size_t)wave * wstride;
    c64* Xs = X0 + (size_t)wave * wstride;
    EigWork<B>* ew = reinterpret_cast<EigWork<B>*>(X0 + (size_t)W * wstride) + wave;  // SCR only
    const int N = p.N;
    const size_t S = p.plane ? (size_t)p.plane : (size_t)nblk;
    const int lr = lane & 7, lc = lane >> 3;
    (void)NP;
    // Persistent loop: G^R and S< of the next tile, and S> of the current
    // one, are fetched into registers while the products run.
    // Register prefetch for B <= 3 (N <= 24); larger tiles would spill and
    // load straight into LDS.
    constexpr bool kPrefetch = B <= 3;
    TileMap<B, W> tm;
    tm.init(N);
    constexpr int EPT = kPrefetch ? TileMap<B, W>::EPT : 1;
    c64 rG[EPT], rS[EPT];
    const c64* gr_fam[1] = {p.GR};
    long long d = blockIdx.x;
    if constexpr (kPrefetch) {
        if (d < ntiles) {
            const TileCoord c = tile_coord<W>(d, ntiles, p.NE);
            tm.fetch(rG, gr_fam, 1, c.blk0, c.nvalid, S);
            tm.fetch(rS, p.famL, p.nfamL, c.blk0, c.nvalid, S);
        }
    }
    for (; d < ntiles; d += gridDim.x) {
        const TileCoord tc = tile_coord<W>(d, ntiles, p.NE);
        const int nvalid = tc.nvalid;
        const long long blk0 = tc.blk0;
        const long long blk = blk0 + wave;
        const bool active = wave < nvalid;
        const unsigned mask = (nvalid >= 32) ? 0xffffffffu : ((1u << nvalid) - 1u);
        unsigned sbits = 0;
        __syncthreads();  // the previous tile's LDS is free
        if constexpr (kPrefetch) {
            tm.put(G0, wstride, rG, nvalid);
            tm.put(X0, wstride, rS, nvalid);
        } else {
            tile_load<W, LD>(G0, wstride, gr_fam, 1, blk0, nvalid, N, S);
            tile_load<W, LD>(X0, wstride, p.famL, p.nfamL, blk0, nvalid, N, S);
        }
        __syncthreads();
        if constexpr (kPrefetch) tm.fetch(rS, p.famG, p.nfamG, blk0, nvalid, S);
        if (p.A) tile_store_spectral<W, LD>(G0, wstride, p.A, blk0, mask, N, S);  // G0 stays read-only
        for (int comp = 0; comp < 2; ++comp) {
            c64* dst = comp ? p.GG : p.GL;
            if (comp == 1) {
                if constexpr (kPrefetch) {
                    tm.put(X0, wstride, rS, nvalid);
                    if (d + gridDim.x < ntiles) {
                        const TileCoord c = tile_coord<W>(d + gridDim.x, ntiles, p.NE);
                        tm.fetch(rG, gr_fam, 1, c.blk0, c.nvalid, S);
                        tm.fetch(rS, p.famL, p.nfamL, c.blk0, c.nvalid, S);
                    }
                } else {
                    tile_load<W, LD>(X0, wstride, p.famG, p.nfamG, blk0, nvalid, N, S);
                }
                __syncthreads();
            }
            c64* tot = comp ? p.totG : p.totL;
            if (tot) {
                tile_store<W, LD>(X0, wstride, tot, blk0, mask, N, S);
                __syncthreads();  // the products overwrite the tiles
            }
            if (active) keldysh_products<B>(Gs, Xs, N, lane, p.traces ? p.traces + (size_t)comp * nblk + blk : nullptr);
            if constexpr (SCR) {
                if (active) {
                    bool u = false;
                    const double v = psd_screen_tile<B, LD>(Xs, N, lane, ew, 0.0, comp ? 1.0 : -1.0, u);
                    sbits |= (u ? (comp ? kScrRawG : kScrRawL) : 0u) | (isinf(v) ? kScrNonFinite : 0u);
                    if (comp == 1 && lane == 0) p.raw_screen[blk] = (unsigned char)sbits;
                }
            }
            __syncthreads();
            tile_store<W, LD>(X0, wstride, dst, blk0, mask, N, S);
            // before S> replaces the tile; after comp 1 the next tile's
            // barrier at the top of the loop does this
            if (comp == 0) __syncthreads();
        }
    }
}

// Dyson and Keldysh in one pass (QCLHIP_ENGINE_FUSED), one tile per
// workgroup: the self-energy tile goes into the X tiles, the elimination
// leaves G^R in the G tiles (stored once, as k_dyson stores it), and the
// Keldysh products of k_keldysh read it from there, so G^R is never read
// back from HBM. Same arithmetic as k_dyson + k_keldysh: the results are
// bitwise those of the two launches. The two LDS tiles per wave hold the
// elimination to two waves per SIMD at N <= 24, where k_dyson runs four;
// the autotuner weighs that against the saved G^R traffic per shape.
template <int B, int W>
__global__ __launch_bounds__(64 * W) void k_dyson_keldysh(DKArgs p, long long ntiles) {
    constexpr int LD = DKShape<B>::LD;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ unsigned s_ok;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long long nblk = (long long)p.NE * p.Nk;
    const int wstride = TileStride<B>::value;
    c64* G0 = reinterpret_cast<c64*>(smem);
    c64* X0 = G0 + (size_t)W * wstride;
    c64* Gs = G0 + (size_t)wave * wstride;
    c64* Xs = X0 + (size_t)wave * wstride;
    const int N = p.N;
    const size_t S = p.plane ? (size_t)p.plane : (size_t)nblk;
    const TileCoord tc = tile_coord<W>(blockIdx.x, ntiles, p.NE);
    dyson_tile<B, W>(p, tc, X0, G0, s_ok);  // ends with a barrier and the G^R store from G0: X0 is free
    const int nvalid = tc.nvalid;
    const long long blk0 = tc.blk0;
    const long long blk = blk0 + wave;
    const bool active = wave < nvalid;
    const unsigned mask = (nvalid >= 32) ? 0xffffffffu : ((1u << nvalid) - 1u);
    tile_load<W, LD>(X0, wstride, p.famL, p.nfamL, blk0, nvalid, N, S);
    __syncthreads();
    if (p.A) tile_store_spectral<W, LD>(G0, wstride, p.A, blk0, mask, N, S);  // G0 stays read-only
    for (int comp = 0; comp < 2; ++comp) {
        c64* dst = comp ? p.GG : p.GL;
        if (comp == 1) {
            tile_load<W, LD>(X0, wstride, p.famG, p.nfamG, blk0, nvalid, N, S);
            __syncthreads();
        }
        c64* tot = comp ? p.totG : p.totL;
        if (tot) {
            tile_store<W, LD>(X0, wstride, tot, blk0, mask, N, S);
            __syncthreads();  // the products overwrite the tiles
        }
        if (active) keldysh_products<B>(Gs, Xs, N, lane, p.traces ? p.traces + (size_t)comp * nblk + blk : nullptr);
        __syncthreads();
        tile_store<W, LD>(X0, wstride, dst, blk0, mask, N, S);
        if (comp == 0) __syncthreads();  // before S> replaces the tile
    }
}

// Batched Keldysh products: a persistent grid over the tiles of every
// instance of one tile class (each tile finds its instance; no prefetch);
// SPLIT waves per block for the large blocks, as k_keldysh_split.
template <int B, int W, int SPLIT = 1>
__global__ __launch_bounds__(64 * SPLIT * W) void k_keldysh_batch(BatchClass cl) {
    constexpr int LD = DKShape<B>::LD;
    constexpr int TH = 64 * SPLIT * W;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ c64 diag[W][8 * B];
    const long long total = cl.begin[cl.n];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int bw = wave / SPLIT, part = wave % SPLIT;
    const int wstride = TileStride<B>::value;
    c64* G0 = reinterpret_cast<c64*>(smem);
    c64* X0 = G0 + (size_t)W * wstride;
    c64* Gs = G0 + (size_t)bw * wstride;
    c64* Xs = X0 + (size_t)bw * wstride;
    for (long long d = blockIdx.x; d < total; d += gridDim.x) {
        const long long g = xcd_tile(d, total);
        const int jj = batch_find(cl, g);
        const DKArgs p = cl.insts[cl.inst[jj]];
        const long long t = g - cl.begin[jj];
        const int nte = (p.NE + W - 1) / W;
        const int e0 = (int)(t % nte) * W;
        const int m = (int)(t / nte);
        const int nvalid = min(W, p.NE - e0);
        const long long nblk = (long long)p.NE * p.Nk;
        const long long blk0 = e0 + (long long)p.NE * m;
        const long long blk = blk0 + bw;
        const int N = p.N;
        const size_t S = (size_t)nblk;
        const bool active = bw < nvalid;
        const unsigned mask = (nvalid >= 32) ? 0xffffffffu : ((1u << nvalid) - 1u);
        const c64* gr_fam[1] = {p.GR};
        __syncthreads();  // the previous tile's LDS is free
        tile_load<W, LD, TH>(G0, wstride, gr_fam, 1, blk0, nvalid, N, S);
        for (int comp = 0; comp < 2; ++comp) {
            tile_load<W, LD, TH>(X0, wstride, comp ? p.famG : p.famL, comp ? p.nfamG : p.nfamL, blk0, nvalid, N,
                                 S);
            __syncthreads();
            c64* tr = p.traces ? p.traces + (size_t)comp * nblk + blk : nullptr;
            if constexpr (SPLIT > 1) {
                keldysh_products_split<B, SPLIT>(Gs, Xs, diag[bw], N, lane, part, active, tr);
            } else {
                if (active) keldysh_products<B>(Gs, Xs, N, lane, tr);
                __syncthreads();
            }
            tile_store<W, LD, false, TH>(X0, wstride, comp ? p.GG : p.GL, blk0, mask, N, S);
            __syncthreads();
        }
    }
}

// Batched sandwich out_em = T (S X)_em T^dagger with a constant T (the
// embedding components, production_operators.jl:886-922). Same lane grid and
// LDS products as the Keldysh step; zero outside the shift window.
struct SWArgs {
    int NE, Nk, N;
    const c64* T;
    const c64* X;
    const int* left;
    const int* right;
    const double* wr;
    const unsigned char* valid;
    c64* out;
    int tadj;        // 1: use T^dagger in place of T (T^dagger X T)
    int accumulate;  // 1: out += result
};

// SPLIT waves per block share its tile (keldysh_products_split) for the
// large blocks; the results are bitwise those of one wave per block.
template <int B, int W, int SPLIT = 1>
__global__ __launch_bounds__(64 * SPLIT * W) void k_sandwich(SWArgs p) {
    constexpr int NP = DKShape<B>::NP;
    constexpr int LD = DKShape<B>::LD;
    constexpr int TH = 64 * SPLIT * W;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int blkw = wave / SPLIT, part = wave % SPLIT;
    const int nte = (p.NE + W - 1) / W;
    const long long t = xcd_tile(blockIdx.x, gridDim.x);
    const int e0 = (int)(t % nte) * W;
    const int m = (int)(t / nte);
    const int nvalid = min(W, p.NE - e0);
    const long long nblk = (long long)p.NE * p.Nk;
    const long long mbase = (long long)p.NE * m;
    const int wstride = TileStride<B>::value;
    c64* Ts = reinterpret_cast<c64*>(smem);          // the constant T, shared
    c64* X0 = Ts + wstride;
    c64* Xs = X0 + (size_t)blkw * wstride;
    const int N = p.N;
    const size_t S = (size_t)nblk;
    const int lr = lane & 7, lc = lane >> 3;
    for (int f = threadIdx.x; f < N * N; f += TH) {
        if (p.tadj)
            Ts[(f / N) * LD + f % N] = cconj(p.T[f]);
        else
            Ts[(f % N) * LD + f / N] = p.T[f];
    }
    // Shifted input X(e + delta) for the W energies: two coalesced gathers.
    {
        const int total = N * N * W;
        int f = threadIdx.x;
        const int w = f % W;
        int ij = f / W;
        int i = ij % N, j = ij / N;
        const int e = e0 + w;
        const bool valid = w < nvalid && (!p.valid || p.valid[e]);
        int l = e, r = e;
        double th = 0.0;
        if (valid && p.valid) {
            l = p.left[e];
            r = p.right[e];
            th = p.wr[e];
        }
        for (; f < total; f += TH) {
            if (w < nvalid) {
                c64 v = cmk(0.0, 0.0);
                if (valid) {
                    const size_t off = mbase + S * (i + (size_t)N * j);
                    const c64 xl = p.X[l + off];
                    v = xl;
                    if (!(l == r || th == 0.0)) {
                        const c64 xr = p.X[r + off];
                        v = cmk((1.0 - th) * xl.x + th * xr.x, (1.0 - th) * xl.y + th * xr.y);
                    }
                }
                X0[w * wstride + i * LD + j] = v;
            }
            i += TH / W;
            while (i >= N) {
                i -= N;
                ++j;
            }
        }
    }
    __syncthreads();
    if constexpr (SPLIT > 1) {
        keldysh_products_split<B, SPLIT>(Ts, Xs, nullptr, N, lane, part, blkw < nvalid, nullptr);
    } else if (wave < nvalid) {
        c64 Tm[B][B];
        block_product<B, false>(Ts, Xs, N, lr, lc, Tm);
        wave_sync();
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) Xs[(lr * B + r) * LD + lc * B + c] = Tm[r][c];
        wave_sync();
        block_product<B, true>(Xs, Ts, N, lr, lc, Tm);
        wave_sync();
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) Xs[(lr * B + r) * LD + lc * B + c] = Tm[r][c];
    }
    __syncthreads();
    const unsigned mask = (nvalid >= 32) ? 0xffffffffu : ((1u << nvalid) - 1u);
    if (p.accumulate)
        tile_store<W, LD, true, TH>(X0, wstride, p.out, e0 + mbase, mask, N, S);
    else
        tile_store<W, LD, false, TH>(X0, wstride, p.out, e0 + mbase, mask, N, S);
}

// Deterministic reduction of the weighted block traces in the reference's
// term order linear = e*Nk + m (production_operators.jl:1426-1440), then the
// realness gates, lambda and the paired coefficients (production_state.jl:
// 49-70, reference/greens.jl:155-185). One workgroup; fixed partition.
// The realness limit is the reference's 1e-10 (numerical_invariants.jl:5)
// for FP64 fields and realness_limit() of the grid otherwise.
constexpr int kReduceThreads = 1024;

__device__ inline void number_pair_body(const c64* traces, int ncomp, int NE, int Nk, const double* wE,
                                        const double* wk, int gs, double target, int pair, double limit,
                                        PairCoefficients* coef, DeviceStatus* st) {
    __shared__ double red[2][2][kReduceThreads];
    const long long nblk = (long long)NE * Nk;
    const int t = threadIdx.x;
    const long long per = (nblk + kReduceThreads - 1) / kReduceThreads;
    const long long lo = t * per, hi = min(nblk, lo + per);
    for (int comp = 0; comp < ncomp; ++comp) {
        double sx = 0.0, sy = 0.0;
        for (long long linear = lo; linear < hi; ++linear) {
            int e = (int)(linear / Nk), mm = (int)(linear % Nk);
            c64 tr = traces[(size_t)comp * nblk + e + (size_t)NE * mm];
            double w = wE[e] * wk[mm];
            sx += w * tr.x;
            sy += w * tr.y;
        }
        red[comp][0][t] = sx;
        red[comp][1][t] = sy;
    }
    __syncthreads();
    for (int width = kReduceThreads / 2; width > 0; width >>= 1) {
        if (t < width)
            for (int comp = 0; comp < ncomp; ++comp) {
                red[comp][0][t] += red[comp][0][t + width];
                red[comp][1][t] += red[comp][1][t + width];
            }
        __syncthreads();
    }
    if (t != 0) return;
    const double k = -(double)gs / (2.0 * M_PI);  // value *= -i g_s/(2 pi)
    double val[2], imv[2];
    for (int comp = 0; comp < ncomp; ++comp) {
        double x = red[comp][0][0], y = red[comp][1][0];
        val[comp] = -y * k;  // Re[(x + iy)(0 + ik)]
        imv[comp] = x * k;
        if (!(fabs(imv[comp]) <= limit * fmax(fabs(val[comp]), 1.0))) {
            report_status(st, QCLHIP_ERR_DOMAIN, 0);
        }
    }
    PairCoefficients out;
    out.number = val[0];
    out.holes = ncomp > 1 ? -val[1] : 0.0;
    out.lambda = out.number / target;
    out.a = 1.0;
    out.c = 0.0;
    if (pair == 2) {
        // :scalar_lesser (production_state.jl:54-61): only lambda > 0 is required.
        if (!(isfinite(out.lambda) && out.lambda > 0)) report_status(st, QCLHIP_ERR_DOMAIN, 0);
    } else if (pair) {
        if (!(isfinite(out.lambda) && out.lambda > 0)) report_status(st, QCLHIP_ERR_DOMAIN, 0);
        double number = out.number, holes = out.holes;
        if (!(isfinite(number) && number > 0) || !(isfinite(holes) && holes >= 0) ||
            !(target <= number + holes)) {
            report_status(st, QCLHIP_ERR_DOMAIN, 0);
        } else if (number >= target) {
            out.a = target / number;
            out.c = 0.0;
        } else {
            out.a = 1.0;
            out.c = (target - number) / holes;
        }
    }
    *coef = out;
}

__global__ __launch_bounds__(kReduceThreads) void k_number_pair(const c64* traces, int ncomp,
                                                                  int NE, int Nk, const double* wE,
                                                                  const double* wk, int gs,
                                                                  double target, int pair, double limit,
                                                                  PairCoefficients* coef,
                                                                  DeviceStatus* st) {
    number_pair_body(traces, ncomp, NE, Nk, wE, wk, gs, target, pair, limit, coef, st);
}


__global__ __launch_bounds__(kReduceThreads) void k_number_pair_batch(const BatchScalars* b, DeviceStatus* st) {
    const BatchScalars& d = b[blockIdx.x];
    number_pair_body(d.traces, 2, d.NE, d.Nk, d.wE, d.wk, d.gs, d.target, 1, d.limit, d.coef, st);
}

// Batched paired map: blockIdx.y is the instance.
__global__ void k_normalize_pair_batch(const BatchScalars* b) {
    const BatchScalars& d = b[blockIdx.y];
    const double a = d.coef->a, c = d.coef->c;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < d.n;
         i += (long long)gridDim.x * blockDim.x) {
        const c64 l = d.GL[i], g = d.GG[i];
        d.GL[i] = cmk(a * l.x - c * g.x, a * l.y - c * g.y);
        d.GG[i] = cmk((1.0 - c) * g.x - (1.0 - a) * l.x, (1.0 - c) * g.y - (1.0 - a) * l.y);
    }
}

// Block traces of one field (for the standalone number functional).
__global__ void k_block_traces(const c64* G, long long nblk, int N, c64* traces) {
    long long blk = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= nblk) return;
    c64 tr = cmk(0.0, 0.0);
    for (int a = 0; a < N; ++a) tr = cadd(tr, G[blk + (size_t)nblk * (a + (size_t)N * a)]);
    traces[blk] = tr;
}

__global__ void k_normalize_pair(long long n, c64* __restrict__ GL, c64* __restrict__ GG,
                                 const PairCoefficients* coef, double a_host, double c_host) {
    const double a = coef ? coef->a : a_host;
    const double c = coef ? coef->c : c_host;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        c64 l = GL[i], g = GG[i];
        GL[i] = cmk(a * l.x - c * g.x, a * l.y - c * g.y);
        GG[i] = cmk((1.0 - c) * g.x - (1.0 - a) * l.x, (1.0 - c) * g.y - (1.0 - a) * l.y);
    }
}

// G< <- G< / lambda, the legacy scalar normalization (production_state.jl:54-61;
// a complex divided by a real divides both parts).
__global__ void k_scalar_lesser(long long n, c64* __restrict__ GL, const PairCoefficients* coef) {
    const double lam = coef->lambda;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const c64 l = GL[i];
        GL[i] = cmk(l.x / lam, l.y / lam);
    }
}

// A_em(a,b) = i (G(a,b) - conj(G(b,a))); consecutive threads = consecutive energies.
__global__ void k_spectral(long long nblk, int N, const c64* __restrict__ GR, c64* __restrict__ A) {
    long long n = nblk * N * N;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        long long blk = idx % nblk, ab = idx / nblk;
        int a = (int)(ab % N), b = (int)(ab / N);
        c64 g = GR[idx], gt = GR[blk + nblk * (b + (long long)N * a)];
        double dx = g.x - gt.x, dy = g.y + gt.y;  // G - conj(G^T)
        A[idx] = cmk(-dy, dx);                    // i * (dx + i dy)
    }
}

// G> = (G - G^dagger) + G< and A = i (G - G^dagger), either optional.
__global__ void k_green_state(long long nblk, int N, const c64* __restrict__ GR, const c64* __restrict__ GL,
                              c64* __restrict__ GG, c64* __restrict__ A) {
    long long n = nblk * N * N;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        long long blk = idx % nblk, ab = idx / nblk;
        int a = (int)(ab % N), b = (int)(ab / N);
        c64 g = GR[idx], gt = GR[blk + nblk * (b + (long long)N * a)];
        double dx = g.x - gt.x, dy = g.y + gt.y;
        if (GG) {
            c64 l = GL[idx];
            GG[idx] = cmk(dx + l.x, dy + l.y);
        }
        if (A) A[idx] = cmk(-dy, dx);
    }
}

// G< = -f (G(a,b) - conj(G(b,a))) (seed.jl:113-126).
__global__ void k_seed_lesser(int NE, long long nblk, int N, const c64* __restrict__ GR,
                              const double* eps, double mu, double kBT, c64* __restrict__ GL) {
    long long n = nblk * N * N;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        long long blk = idx % nblk, ab = idx / nblk;
        int a = (int)(ab % N), b = (int)(ab / N);
        int e = (int)(blk % NE);
        double x = (eps[e] - mu) / kBT;
        double f = x > 40.0 ? 0.0 : (x < -40.0 ? 1.0 : 1.0 / (exp(x) + 1.0));
        c64 g = GR[idx], gt = GR[blk + nblk * (b + (long long)N * a)];
        GL[idx] = cmk(-f * (g.x - gt.x), -f * (g.y + gt.y));
    }
}

// q_e = g_s/(2 pi) wE_e sum_m wk_m sum_a Re[i (z - conj z)] (seed.jl:52-64).
__global__ void k_seed_weights(int NE, int Nk, int N, const c64* __restrict__ GR, const double* wE,
                               const double* wk, int gs, double* q) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NE) return;
    const long long nblk = (long long)NE * Nk;
    double msum = 0.0;
    for (int mm = 0; mm < Nk; ++mm) {
        double trA = 0.0;
        for (int a = 0; a < N; ++a) {
            c64 z = GR[e + (long long)NE * mm + nblk * (a + (long long)N * a)];
            trA += -(z.y - (-z.y));  // Re[i (z - conj z)] = -(2 Im z)
        }
        msum += wk[mm] * trA;
    }
    q[e] = ((double)gs / (2.0 * M_PI)) * wE[e] * msum;
}

// out = x0 + x1 + ... (up to 16 addends per launch, in order).
struct SumArgs {
    const c64* x[16];
    int nx;
};

__global__ void k_sum_fields(long long n, SumArgs a, int accumulate, c64* __restrict__ out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        c64 v = accumulate ? out[i] : a.x[0][i];
        for (int f = accumulate ? 0 : 1; f < a.nx; ++f) v = cadd(v, a.x[f][i]);
        out[i] = v;
    }
}

__global__ void k_mix(long long n, c64* __restrict__ old_, const c64* __restrict__ cand, double alpha) {
    const double beta = 1.0 - alpha;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        c64 o = old_[i], c = cand[i];
        old_[i] = mix1(o, c, alpha, beta);
    }
}

// All (old, candidate) pairs of one mixing phase in a single pass.
constexpr int kMaxMixPairs = 32;
struct MixArgs {
    c64* old_[kMaxMixPairs];
    const c64* cand[kMaxMixPairs];
    int npairs;
};

__global__ void k_mix_many(long long n, MixArgs a, double alpha) {
    const double beta = 1.0 - alpha;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        for (int p = 0; p < a.npairs; ++p) {
            const c64 o = a.old_[p][i], c = a.cand[p][i];
            a.old_[p][i] = mix1(o, c, alpha, beta);
        }
    }
}

inline unsigned grid_for(long long n, int threads, int sm_count) {
    long long blocks = (n + threads - 1) / threads;
    long long cap = (long long)sm_count * 32;
    if (cap < 1) cap = 1024;
    return (unsigned)std::max<long long>(1, std::min(blocks, cap));
}

template <int B>
constexpr int tile_waves(size_t per_wave) {
    // Energies per workgroup: 4 (64-byte segments) when it fits in half the LDS
    // of a CU, else fewer.
    return per_wave * 4 <= 80 * 1024 ? 4 : (per_wave * 2 <= 80 * 1024 ? 2 : 1);
}

template <int B, int W>
int launch_dyson_w(qclhip_ctx* ctx, const DKArgs& a) {
    const size_t lds = DyShape<B, W>::bytes;
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_dyson<B, W>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long tiles = (long long)((a.NE + W - 1) / W) * a.Nk;
    // One tile per workgroup (the loop runs once): measured faster than a
    // persistent grid for the elimination, which hides latency by occupancy.
    hipLaunchKernelGGL((k_dyson<B, W>), dim3((unsigned)tiles), dim3(64 * W), lds, ctx->stream, a, tiles);
    return launch_check(ctx, "k_dyson");
}

// Largest Dyson tile: 8 energies per workgroup (128-byte segments: the tile
// access pattern moves 4.2 TB/s there against 2.7 TB/s with 64-byte
// segments) while two workgroups still fit in the LDS of a CU.
template <int B>
constexpr bool dyson_w8_fits() {
    return DyShape<B, 8>::per_wave * 8 <= 80 * 1024;
}

template <int B>
int launch_dyson(qclhip_ctx* ctx, const DKArgs& a) {
    // a.wd: the autotuner's choice (4 or 8), else 8 where it fits
    if (a.wd == 4) return launch_dyson_w<B, kDysonTileWaves>(ctx, a);
    if constexpr (dyson_w8_fits<B>()) return launch_dyson_w<B, 8>(ctx, a);
    return launch_dyson_w<B, kDysonTileWaves>(ctx, a);
}

template <int B, int W, bool SCR>
int launch_keldysh_ws(qclhip_ctx* ctx, const DKArgs& a) {
    constexpr size_t per_wave = 2 * (size_t)TileStride<B>::value * sizeof(c64) + (SCR ? sizeof(EigWork<B>) : 0);
    const size_t lds = per_wave * W;
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_keldysh<B, W, SCR>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long tiles = (long long)((a.NE + W - 1) / W) * a.Nk;
    unsigned grid = 0;
    int rc = persistent_grid(ctx, (const void*)k_keldysh<B, W, SCR>, 64 * W, lds, tiles, &grid);
    if (rc) return rc;
    hipLaunchKernelGGL((k_keldysh<B, W, SCR>), dim3(grid), dim3(64 * W), lds, ctx->stream, a, tiles);
    return launch_check(ctx, "k_keldysh");
}

template <int B, int W>
int launch_keldysh_w(qclhip_ctx* ctx, const DKArgs& a) {
    return a.raw_screen ? launch_keldysh_ws<B, W, true>(ctx, a) : launch_keldysh_ws<B, W, false>(ctx, a);
}

// Waves per block when the autotuner has not chosen (index B): one wave
// per block up to N = 24, where two LDS tiles per wave still leave two waves
// per SIMD; above, one row of the 8x8 lane grid per wave (measured at
// NE = 2048 or 1024, Nk = 64: 9.4 -> 6.2 ms at N = 32, 9.1 -> 7.2 ms at
// N = 40, 19.7 -> 9.3 ms at N = 48).
constexpr int kDefaultKeldyshSplit[7] = {1, 1, 1, 1, 4, 5, 6};

template <int B, int W, int SPLIT, bool SCR>
int launch_keldysh_split_ws(qclhip_ctx* ctx, const DKArgs& a) {
    constexpr size_t per_block = 2 * (size_t)TileStride<B>::value * sizeof(c64) + (SCR ? sizeof(EigWork<B>) : 0);
    const size_t lds = per_block * W;
    const void* kern;
    if constexpr (SCR)
        kern = (const void*)k_keldysh_split_scr<B, W, SPLIT>;
    else
        kern = (const void*)k_keldysh_split<B, W, SPLIT>;
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long tiles = (long long)((a.NE + W - 1) / W) * a.Nk;
    unsigned grid = 0;
    int rc = persistent_grid(ctx, kern, 64 * SPLIT * W, lds, tiles, &grid);
    if (rc) return rc;
    if constexpr (SCR)
        hipLaunchKernelGGL((k_keldysh_split_scr<B, W, SPLIT>), dim3(grid), dim3(64 * SPLIT * W), lds, ctx->stream, a,
                           tiles);
    else
        hipLaunchKernelGGL((k_keldysh_split<B, W, SPLIT>), dim3(grid), dim3(64 * SPLIT * W), lds, ctx->stream, a,
                           tiles);
    return launch_check(ctx, "k_keldysh_split");
}

// Raw-product screens in the split kernels up to N = 32 (B = 4); above, the
// screening variant would spill, and the caller leaves the raw products to
// the residual suite (qclhip_current_green_screened refuses N > 32).
template <int B, int W, int SPLIT>
int launch_keldysh_split_w(qclhip_ctx* ctx, const DKArgs& a) {
    if constexpr (B <= 4) {
        if (a.raw_screen) return launch_keldysh_split_ws<B, W, SPLIT, true>(ctx, a);
    }
    return launch_keldysh_split_ws<B, W, SPLIT, false>(ctx, a);
}

// Waves per block of the vector Keldysh kernel compiled for each tile size:
// B = 4 (N <= 32): 2 or 4; B = 5: 5; B = 6: 2, 3 or 6 (the rows of a block
// split evenly).
constexpr bool split_compiled(int B, int S) {
    return S == 1 || (B == 4 && (S == 2 || S == 4)) || (B == 5 && S == 5) || (B == 6 && (S == 2 || S == 3 || S == 6));
}

template <int B, int W>
int launch_keldysh_split(qclhip_ctx* ctx, const DKArgs& a, int ks) {
    if constexpr (B == 4) {
        if (ks == 2) return launch_keldysh_split_w<B, W, 2>(ctx, a);
        if (ks == 4) return launch_keldysh_split_w<B, W, 4>(ctx, a);
    } else if constexpr (B == 5) {
        if (ks == 5) return launch_keldysh_split_w<B, W, 5>(ctx, a);
    } else if constexpr (B == 6) {
        if (ks == 2) return launch_keldysh_split_w<B, W, 2>(ctx, a);
        if (ks == 3) return launch_keldysh_split_w<B, W, 3>(ctx, a);
        if (ks == 6) return launch_keldysh_split_w<B, W, 6>(ctx, a);
    }
    return fail(ctx, QCLHIP_ERR_ARGUMENT, "waves per block not available for this basis size");
}

// Twice the energies per workgroup for the split kernels (N > 24), where one
// workgroup per CU still holds several waves per SIMD: 64-byte segments at
// N <= 32 (4 energies), 32-byte ones above. Fits when the two tiles of every
// block stay within the LDS of a CU.
template <int B>
constexpr bool split_wide_fits() {
    constexpr int W = tile_waves<B>(2 * (size_t)TileStride<B>::value * sizeof(c64));
    return B >= 4 && 2 * (size_t)TileStride<B>::value * sizeof(c64) * (2 * W) <= 160 * 1024;
}

template <int B>
int launch_keldysh(qclhip_ctx* ctx, const DKArgs& a) {
    // a.wk: the autotuner's choice (2 or 4 energies per workgroup), else 4
    // (64-byte segments) when two workgroups fit in a CU's LDS
    constexpr int W = tile_waves<B>(2 * (size_t)TileStride<B>::value * sizeof(c64));
    // a.ks: waves per block (the autotuner's choice, else the rule below)
    const int ks = a.ks ? a.ks : kDefaultKeldyshSplit[B];
    if (ks > 1) {
        if constexpr (split_wide_fits<B>()) {
            if (a.wk == 2 * W) return launch_keldysh_split<B, 2 * W>(ctx, a, ks);
        }
        return launch_keldysh_split<B, W>(ctx, a, ks);
    }
    if constexpr (W >= 2) {
        if (a.wk == 2) return launch_keldysh_w<B, 2>(ctx, a);
    }
    return launch_keldysh_w<B, W>(ctx, a);
}

// The fused kernel: the Keldysh kernel's two tiles per wave and its
// energies per workgroup (4 while two workgroups fit a CU's LDS), one tile
// per workgroup as k_dyson.
template <int B>
int launch_fused(qclhip_ctx* ctx, const DKArgs& a) {
    constexpr size_t per_wave = 2 * (size_t)TileStride<B>::value * sizeof(c64);
    constexpr int W = tile_waves<B>(per_wave);
    static_assert(DyShape<B, W>::per_wave <= per_wave / 2, "elimination buffers exceed the X tile");
    const size_t lds = per_wave * W;
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_dyson_keldysh<B, W>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long tiles = (long long)((a.NE + W - 1) / W) * a.Nk;
    hipLaunchKernelGGL((k_dyson_keldysh<B, W>), dim3((unsigned)tiles), dim3(64 * W), lds, ctx->stream, a, tiles);
    return launch_check(ctx, "k_dyson_keldysh");
}

int launch_fused_fp64(qclhip_ctx* ctx, const DKArgs& a) {
    const int N = a.N;
    if (N <= 8) return launch_fused<1>(ctx, a);
    if (N <= 16) return launch_fused<2>(ctx, a);
    if (N <= 24) return launch_fused<3>(ctx, a);
    if (N <= 32) return launch_fused<4>(ctx, a);
    if (N <= 40) return launch_fused<5>(ctx, a);
    if (N <= 48) return launch_fused<6>(ctx, a);
    return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");
}

// The sandwich takes the Keldysh kernel's waves per block (the same
// products on the same tiles).
template <int B>
int launch_sw(qclhip_ctx* ctx, const SWArgs& a) {
    constexpr size_t per_wave = (size_t)TileStride<B>::value * sizeof(c64);
    // (twice the energies per workgroup measured slower at N = 32: 12.6 ms
    // against 9.9 ms at NE = 4096, Nk = 128)
    constexpr int W = tile_waves<B>(2 * per_wave);
    constexpr int SPLIT = kDefaultKeldyshSplit[B];
    const size_t lds = per_wave * (W + 1);
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_sandwich<B, W, SPLIT>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long tiles = (long long)((a.NE + W - 1) / W) * a.Nk;
    hipLaunchKernelGGL((k_sandwich<B, W, SPLIT>), dim3((unsigned)tiles), dim3(64 * SPLIT * W), lds, ctx->stream, a);
    return launch_check(ctx, "k_sandwich");
}

int sandwich_dispatch(qclhip_ctx* ctx, const SWArgs& a) {
    const int N = a.N;
    if (N <= 8) return launch_sw<1>(ctx, a);
    if (N <= 16) return launch_sw<2>(ctx, a);
    if (N <= 24) return launch_sw<3>(ctx, a);
    if (N <= 32) return launch_sw<4>(ctx, a);
    if (N <= 40) return launch_sw<5>(ctx, a);
    if (N <= 48) return launch_sw<6>(ctx, a);
    return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");
}

int launch_dyson_fp64(qclhip_ctx* ctx, const DKArgs& a) {
    const int N = a.N;
    if (N <= 8) return launch_dyson<1>(ctx, a);
    if (N <= 16) return launch_dyson<2>(ctx, a);
    if (N <= 24) return launch_dyson<3>(ctx, a);
    if (N <= 32) return launch_dyson<4>(ctx, a);
    if (N <= 40) return launch_dyson<5>(ctx, a);
    if (N <= 48) return launch_dyson<6>(ctx, a);
    return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");
}

// Whether a tile shape exists (is compiled and fits the LDS) for basis size N.
bool dyson_waves_possible(int N, int W) {
    if (W == kDysonTileWaves) return true;
    if (W != 8) return false;
    switch ((N + 7) / 8) {
        case 1: return dyson_w8_fits<1>();
        case 2: return dyson_w8_fits<2>();
        case 3: return dyson_w8_fits<3>();
        case 4: return dyson_w8_fits<4>();
        case 5: return dyson_w8_fits<5>();
        case 6: return dyson_w8_fits<6>();
        default: return false;
    }
}

bool keldysh_waves_possible(int N, int W) {
    auto widest = [](int B) {
        switch (B) {
            case 1: return tile_waves<1>(2 * (size_t)TileStride<1>::value * sizeof(c64));
            case 2: return tile_waves<2>(2 * (size_t)TileStride<2>::value * sizeof(c64));
            case 3: return tile_waves<3>(2 * (size_t)TileStride<3>::value * sizeof(c64));
            case 4: return tile_waves<4>(2 * (size_t)TileStride<4>::value * sizeof(c64));
            case 5: return tile_waves<5>(2 * (size_t)TileStride<5>::value * sizeof(c64));
            default: return tile_waves<6>(2 * (size_t)TileStride<6>::value * sizeof(c64));
        }
    };
    const int w = widest((N + 7) / 8);
    bool wide = false;  // the split kernels' twice-as-wide tile
    switch ((N + 7) / 8) {
        case 4: wide = split_wide_fits<4>(); break;
        case 5: wide = split_wide_fits<5>(); break;
        case 6: wide = split_wide_fits<6>(); break;
        default: break;
    }
    return W == w || (W == 2 && w >= 2) || (wide && W == 2 * w);
}

bool keldysh_split_possible(int N, int S) {
    const int B = (N + 7) / 8;
    return B >= 1 && B <= 6 && split_compiled(B, S);
}

int launch_keldysh_valu(qclhip_ctx* ctx, const DKArgs& a) {
    const int N = a.N;
    if (N <= 8) return launch_keldysh<1>(ctx, a);
    if (N <= 16) return launch_keldysh<2>(ctx, a);
    if (N <= 24) return launch_keldysh<3>(ctx, a);
    if (N <= 32) return launch_keldysh<4>(ctx, a);
    if (N <= 40) return launch_keldysh<5>(ctx, a);
    if (N <= 48) return launch_keldysh<6>(ctx, a);
    return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");
}

// Dyson (+ Keldysh) at the grid's precision. FP64: k_dyson. FP32 +
// refinement: k_dyson_mixed, then the FP64 re-solve of flagged tiles. The
// Keldysh products run in FP64 on the engine the grid selects, or in FP32 on
// the matrix cores (QCLHIP_PRECISION_FP32_PRODUCTS).
// Realness gate of the number functional: the reference's limit for FP64;
// FP32 Keldysh products leave an imaginary part of order 1e-7 relative.
double realness_limit(const qclhip_grid* grid) {
    return grid->precision == QCLHIP_PRECISION_FP32_PRODUCTS ? 1.0e-5 : 1.0e-10;
}

int dyson_keldysh_dispatch(qclhip_ctx* ctx, const qclhip_grid* grid, DKArgs a, bool keldysh) {
    if (a.N > 48) return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");
    a.wd = grid->dyson_waves;
    a.wk = grid->keldysh_waves;
    a.ks = grid->keldysh_split;
    int rc;
    if (grid->precision != QCLHIP_PRECISION_FP64) {
        a.block_flags = grid->block_flags;  // one flag per block, owned by the grid
        if (!a.block_flags) return fail(ctx, QCLHIP_ERR_RUNTIME, "grid has no refinement flags");
        a.refine = grid->refine_steps;
        a.refine_tol = grid->refine_tol;
        rc = launch_mixed_dyson(ctx, a);
        if (rc || !keldysh) return rc;
        if (grid->precision == QCLHIP_PRECISION_FP32_PRODUCTS) return launch_keldysh_mfma(ctx, a, true);
    } else {
        a.block_flags = nullptr;
        if (keldysh && grid->keldysh_engine == QCLHIP_ENGINE_FUSED) return launch_fused_fp64(ctx, a);
        rc = launch_dyson_fp64(ctx, a);
        if (rc || !keldysh) return rc;
    }
    if (grid->keldysh_engine == QCLHIP_ENGINE_MATRIX) return launch_keldysh_mfma(ctx, a, false);
    return launch_keldysh_valu(ctx, a);
}

template <int B>
int launch_dyson_batch_b(qclhip_ctx* ctx, const BatchClass& c, long long total) {
    constexpr int W = kDysonTileWaves;
    const size_t lds = DyShape<B, W>::bytes;
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_dyson_batch<B, W>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_dyson_batch<B, W>), dim3((unsigned)total), dim3(64 * W), lds, ctx->stream, c);
    return launch_check(ctx, "k_dyson_batch");
}

template <int B>
constexpr int keldysh_waves() {
    return tile_waves<B>(2 * (size_t)TileStride<B>::value * sizeof(c64));
}

template <int B>
int launch_keldysh_batch_b(qclhip_ctx* ctx, const BatchClass& c, long long total) {
    constexpr int W = keldysh_waves<B>();
    constexpr int SPLIT = kDefaultKeldyshSplit[B];
    const size_t lds = 2 * (size_t)TileStride<B>::value * sizeof(c64) * W;
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_keldysh_batch<B, W, SPLIT>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    unsigned grid = 0;
    int rc = persistent_grid(ctx, (const void*)k_keldysh_batch<B, W, SPLIT>, 64 * SPLIT * W, lds, total, &grid);
    if (rc) return rc;
    hipLaunchKernelGGL((k_keldysh_batch<B, W, SPLIT>), dim3(grid), dim3(64 * SPLIT * W), lds, ctx->stream, c);
    return launch_check(ctx, "k_keldysh_batch");
}

int batch_tile_waves(int B, bool keldysh) {
    if (!keldysh) return kDysonTileWaves;
    switch (B) {
        case 1: return keldysh_waves<1>();
        case 2: return keldysh_waves<2>();
        case 3: return keldysh_waves<3>();
        case 4: return keldysh_waves<4>();
        case 5: return keldysh_waves<5>();
        default: return keldysh_waves<6>();
    }
}

int launch_dk_batch(qclhip_ctx* ctx, int B, bool keldysh, const BatchClass& c, long long total) {
    switch (B) {
        case 1: return keldysh ? launch_keldysh_batch_b<1>(ctx, c, total) : launch_dyson_batch_b<1>(ctx, c, total);
        case 2: return keldysh ? launch_keldysh_batch_b<2>(ctx, c, total) : launch_dyson_batch_b<2>(ctx, c, total);
        case 3: return keldysh ? launch_keldysh_batch_b<3>(ctx, c, total) : launch_dyson_batch_b<3>(ctx, c, total);
        case 4: return keldysh ? launch_keldysh_batch_b<4>(ctx, c, total) : launch_dyson_batch_b<4>(ctx, c, total);
        case 5: return keldysh ? launch_keldysh_batch_b<5>(ctx, c, total) : launch_dyson_batch_b<5>(ctx, c, total);
        case 6: return keldysh ? launch_keldysh_batch_b<6>(ctx, c, total) : launch_dyson_batch_b<6>(ctx, c, total);
        default: return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");
    }
}

int launch_pair_batch(qclhip_ctx* ctx, const BatchScalars* d, int ninst, long long maxn) {
    hipLaunchKernelGGL(k_number_pair_batch, dim3(ninst), dim3(kReduceThreads), 0, ctx->stream, d, ctx->status);
    const unsigned blocks = std::max(1u, std::min(grid_for(maxn, 256, ctx->sm_count), 65535u));
    hipLaunchKernelGGL(k_normalize_pair_batch, dim3(blocks, ninst), dim3(256), 0, ctx->stream, d);
    return launch_check(ctx, "batched number functionals");
}

}  // namespace qclhip

// ---- fixed scratch of the allocation-free current_green ------------------
// current_green's only scratch (with at most kMaxFamilies families, summed in
// the kernels' loads) is 2 NE Nk complex block traces, every one written by
// the Keldysh kernel before the number functionals read it, so nothing here
// needs clearing between calls. Each device holds its own copy of this array
// for the life of the process; in host code its name is not a device address,
// so the address is looked up per device (hipGetSymbolAddress) and only that
// address is handed to kernels.
static __device__ __align__(256) unsigned char scratch[qclhip::kStaticScratchBytes];

namespace qclhip {
namespace {
constexpr int kMaxStaticDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxStaticDevices] = {};
}  // namespace

int static_scratch_setup(int device) {
    if (device < 0 || device >= kMaxStaticDevices) return QCLHIP_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return QCLHIP_OK;
    void* ptr = nullptr;
    // the calling context has made `device` current
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return QCLHIP_ERR_RUNTIME;
    g_scratch_ptr[device] = ptr;
    return QCLHIP_OK;
}

void* static_scratch(int device) {
    if (device < 0 || device >= kMaxStaticDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}
}  // namespace qclhip

using namespace qclhip;

extern "C" {

int qclhip_max_basis(void) { return 48; }

static int dyson_keldysh_impl(qclhip_ctx* ctx, const qclhip_grid* grid, long long plane, const double* h,
                              const double* sigR, const double* sigL, const double* sigG, double eta, double* GR,
                              double* GL, double* GG, double* scale, double* traces) {
    if (!ctx || !grid || !h || !GR) return QCLHIP_ERR_ARGUMENT;
    const bool keldysh = sigL || sigG || GL || GG;
    if (keldysh && !(sigL && sigG && GL && GG))
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "Keldysh products need sigL, sigG, GL and GG");
    const long long nblk = (long long)grid->NE * grid->Nk;
    if (keldysh && !traces) {
        traces = (double*)workspace(ctx, (size_t)2 * nblk * sizeof(c64));
        if (!traces) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    }
    DKArgs a{};
    a.NE = grid->NE;
    a.Nk = grid->Nk;
    a.N = grid->Nb;
    a.eps = grid->eps;
    a.h = (const c64*)h;
    a.famR[0] = (const c64*)sigR;
    a.famL[0] = (const c64*)sigL;
    a.famG[0] = (const c64*)sigG;
    a.nfamR = sigR ? 1 : 0;
    a.nfamL = a.nfamG = 1;
    a.eta = eta;
    a.GR = (c64*)GR;
    a.GL = (c64*)GL;
    a.GG = (c64*)GG;
    a.scale = scale;
    a.traces = (c64*)traces;
    a.status = ctx->status;
    a.keldysh = keldysh ? 1 : 0;
    if (plane == 0) return dyson_keldysh_dispatch(ctx, grid, a, keldysh);
    if (grid->precision != QCLHIP_PRECISION_FP64 || grid->keldysh_engine == QCLHIP_ENGINE_MATRIX)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "the plane-strided layout is implemented by the FP64 vector engine");
    if (a.N > 48) return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");
    a.plane = plane;
    a.wd = grid->dyson_waves;
    a.wk = grid->keldysh_waves;
    a.ks = grid->keldysh_split;
    if (keldysh && grid->keldysh_engine == QCLHIP_ENGINE_FUSED) return launch_fused_fp64(ctx, a);
    int rc = launch_dyson_fp64(ctx, a);
    if (rc || !keldysh) return rc;
    return launch_keldysh_valu(ctx, a);
}

int qclhip_dyson_keldysh(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, const double* sigR,
                         const double* sigL, const double* sigG, double eta, double* GR, double* GL,
                         double* GG, double* scale, double* traces) {
    return dyson_keldysh_impl(ctx, grid, 0, h, sigR, sigL, sigG, eta, GR, GL, GG, scale, traces);
}

int qclhip_dyson_keldysh_strided(qclhip_ctx* ctx, const qclhip_grid* grid, int64_t plane, const double* h,
                                 const double* sigR, const double* sigL, const double* sigG, double eta,
                                 double* GR, double* GL, double* GG, double* scale, double* traces) {
    if (!ctx || !grid) return QCLHIP_ERR_ARGUMENT;
    if (plane < (int64_t)grid->NE * grid->Nk)
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "the plane stride must be at least NE*Nk");
    return dyson_keldysh_impl(ctx, grid, (long long)plane, h, sigR, sigL, sigG, eta, GR, GL, GG, scale, traces);
}

int qclhip_retarded_green(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, const double* sigR,
                          double eta, double* GR, double* scale) {
    return qclhip_dyson_keldysh(ctx, grid, h, sigR, nullptr, nullptr, eta, GR, nullptr, nullptr, scale,
                                nullptr);
}

static int current_green_impl(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, int nfam,
                              const double* const* famR, const double* const* famL, const double* const* famG,
                              double target, double* GR, double* GL, double* GG, double* A, double* scale,
                              double* out_host, double* const* totals, bool scalar_lesser = false,
                              c64* traces_ext = nullptr, unsigned char* raw_screen = nullptr) {
    if (!ctx || !grid || !h || !famR || !famL || !famG || !GR || !GL || !GG) return QCLHIP_ERR_ARGUMENT;
    if (raw_screen && (grid->keldysh_engine != QCLHIP_ENGINE_VECTOR ||
                       grid->precision == QCLHIP_PRECISION_FP32_PRODUCTS || grid->Nb > 32))
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED,
                    "raw-product screens run in the FP64 vector Keldysh kernels at Nb <= 32 only");
    if (nfam < 1) return fail(ctx, QCLHIP_ERR_ARGUMENT, "at least the embedding family is required");
    if (!(std::isfinite(target) && target > 0))
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "target charge must be finite and positive");
    const long long nblk = (long long)grid->NE * grid->Nk;
    const long long n = nblk * grid->Nb * grid->Nb;
    // Scratch: traces, then (more families than one launch takes) the totals.
    const bool fused = nfam <= kMaxFamilies;
    size_t tr_bytes = (size_t)2 * nblk * sizeof(c64);
    size_t tot_bytes = fused ? 0 : (size_t)3 * n * sizeof(c64);
    char* ws = nullptr;
    if (traces_ext) {
        // caller-supplied trace storage: nothing is allocated here
        if (!fused) return fail(ctx, QCLHIP_ERR_ARGUMENT, "the allocation-free call sums at most 8 families");
    } else {
        ws = (char*)workspace(ctx, tr_bytes + tot_bytes);
        if (!ws) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    }
    c64* traces = traces_ext ? traces_ext : (c64*)ws;
    DKArgs a{};
    a.NE = grid->NE;
    a.Nk = grid->Nk;
    a.N = grid->Nb;
    a.eps = grid->eps;
    a.h = (const c64*)h;
    if (fused) {
        // The family sum happens in the kernels' loads, in the reference order.
        for (int f = 0; f < nfam; ++f) {
            a.famR[f] = (const c64*)famR[f];
            a.famL[f] = (const c64*)famL[f];
            a.famG[f] = (const c64*)famG[f];
        }
        a.nfamR = a.nfamL = a.nfamG = nfam;
    } else {
        c64* tot = (c64*)(ws + tr_bytes);
        const double* const* fams[3] = {famR, famL, famG};
        for (int k = 0; k < 3; ++k) {
            int done = 0;
            while (done < nfam) {
                SumArgs sa;
                sa.nx = std::min(16, nfam - done);
                for (int f = 0; f < sa.nx; ++f) sa.x[f] = (const c64*)fams[k][done + f];
                hipLaunchKernelGGL(k_sum_fields, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream,
                                   n, sa, done > 0 ? 1 : 0, tot + (size_t)k * n);
                done += sa.nx;
            }
        }
        int rc = launch_check(ctx, "k_sum_fields");
        if (rc) return rc;
        a.famR[0] = tot;
        a.famL[0] = tot + n;
        a.famG[0] = tot + 2 * n;
        a.nfamR = a.nfamL = a.nfamG = 1;
    }
    a.eta = 0.0;
    a.GR = (c64*)GR;
    a.GL = (c64*)GL;
    a.GG = (c64*)GG;
    a.scale = scale;
    a.traces = traces;
    a.status = ctx->status;
    a.keldysh = 1;
    a.raw_screen = raw_screen;
    if (totals) {
        const bool in_kernels = fused && grid->precision == QCLHIP_PRECISION_FP64 &&
                                grid->keldysh_engine != QCLHIP_ENGINE_MATRIX && a.N <= 48;
        if (in_kernels) {
            a.totR = (c64*)totals[0];
            a.totL = (c64*)totals[1];
            a.totG = (c64*)totals[2];
        } else {
            const c64* const* src[3] = {a.famR, a.famL, a.famG};
            for (int k = 0; k < 3; ++k) {
                SumArgs sa;
                sa.nx = 1;
                sa.x[0] = src[k][0];
                if (fused) {
                    sa.nx = nfam;
                    for (int f = 0; f < nfam; ++f) sa.x[f] = src[k][f];
                }
                hipLaunchKernelGGL(k_sum_fields, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream,
                                   n, sa, 0, (c64*)totals[k]);
            }
            int rc0 = launch_check(ctx, "k_sum_fields");
            if (rc0) return rc0;
        }
    }
    // A from the Keldysh kernel's G^R tiles on the FP64 vector path, else k_spectral.
    const bool a_in_kernel = A && !scalar_lesser && grid->precision == QCLHIP_PRECISION_FP64 &&
                             grid->keldysh_engine != QCLHIP_ENGINE_MATRIX && a.N <= 48;
    if (a_in_kernel) a.A = (c64*)A;
    int rc = dyson_keldysh_dispatch(ctx, grid, a, true);
    if (rc) return rc;
    if (scalar_lesser) {
        // number of G<raw only, lambda = number/target > 0, G< /= lambda, then
        // _green_state_production without a given G>: G> = G^R - G^R^dagger + G<.
        hipLaunchKernelGGL(k_number_pair, dim3(1), dim3(kReduceThreads), 0, ctx->stream, traces, 1, grid->NE,
                           grid->Nk, grid->wE, grid->wk, grid->gs, target, 2, realness_limit(grid), ctx->coef,
                           ctx->status);
        hipLaunchKernelGGL(k_scalar_lesser, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream, n,
                           (c64*)GL, (const PairCoefficients*)ctx->coef);
        hipLaunchKernelGGL(k_green_state, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream, nblk,
                           grid->Nb, (const c64*)GR, (const c64*)GL, (c64*)GG, (c64*)A);
    } else {
        hipLaunchKernelGGL(k_number_pair, dim3(1), dim3(kReduceThreads), 0, ctx->stream, traces, 2, grid->NE,
                           grid->Nk, grid->wE, grid->wk, grid->gs, target, 1, realness_limit(grid), ctx->coef,
                           ctx->status);
        hipLaunchKernelGGL(k_normalize_pair, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream, n,
                           (c64*)GL, (c64*)GG, (const PairCoefficients*)ctx->coef, 0.0, 0.0);
    }
    if (A && !a_in_kernel && !scalar_lesser)
        hipLaunchKernelGGL(k_spectral, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream, nblk,
                           grid->Nb, (const c64*)GR, (c64*)A);
    rc = launch_check(ctx, "current_green");
    if (rc) return rc;
    if (out_host) {
        bool deferred = false;
        rc = defer_coefficients(ctx, ctx->coef, &deferred);
        if (rc || deferred) return rc;
        QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, ctx->coef, sizeof(PairCoefficients),
                                           hipMemcpyDeviceToHost, ctx->stream));
        rc = qclhip_ctx_synchronize(ctx);
        if (rc) return rc;
        const PairCoefficients* pc = (const PairCoefficients*)ctx->host_scalars;
        out_host[0] = pc->lambda;
        out_host[1] = pc->number;
        out_host[2] = pc->holes;
        out_host[3] = pc->a;
        out_host[4] = pc->c;
    }
    return QCLHIP_OK;
}

int qclhip_current_green(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, int nfam,
                         const double* const* famR, const double* const* famL, const double* const* famG,
                         double target, double* GR, double* GL, double* GG, double* A, double* scale,
                         double* out_host) {
    return current_green_impl(ctx, grid, h, nfam, famR, famL, famG, target, GR, GL, GG, A, scale, out_host,
                              nullptr);
}

int qclhip_current_green_totals(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, int nfam,
                                const double* const* famR, const double* const* famL, const double* const* famG,
                                double target, double* GR, double* GL, double* GG, double* A, double* scale,
                                double* out_host, double* totR, double* totL, double* totG) {
    if (!totR || !totL || !totG) return QCLHIP_ERR_ARGUMENT;
    double* const tot[3] = {totR, totL, totG};
    return current_green_impl(ctx, grid, h, nfam, famR, famL, famG, target, GR, GL, GG, A, scale, out_host, tot);
}

int qclhip_current_green_screened(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, int nfam,
                                  const double* const* famR, const double* const* famL,
                                  const double* const* famG, double target, double* GR, double* GL, double* GG,
                                  double* A, double* scale, double* out_host, double* totR, double* totL,
                                  double* totG, unsigned char* raw_screen) {
    if (!raw_screen) return QCLHIP_ERR_ARGUMENT;
    if ((totR || totL || totG) && !(totR && totL && totG)) return QCLHIP_ERR_ARGUMENT;
    double* const tot[3] = {totR, totL, totG};
    return current_green_impl(ctx, grid, h, nfam, famR, famL, famG, target, GR, GL, GG, A, scale, out_host,
                              totR ? tot : nullptr, false, nullptr, raw_screen);
}

int qclhip_current_green_noalloc(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, int nfam,
                                 const double* const* famR, const double* const* famL, const double* const* famG,
                                 double target, double* GR, double* GL, double* GG, double* A, double* scale,
                                 double* out_host) {
    if (!ctx || !grid) return QCLHIP_ERR_ARGUMENT;
    if (nfam > kMaxFamilies) return fail(ctx, QCLHIP_ERR_ARGUMENT, "the allocation-free call sums at most 8 families");
    const size_t need = (size_t)2 * grid->NE * grid->Nk * sizeof(c64);
    if (need > kStaticScratchBytes)
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "the allocation-free call takes at most NE*Nk = " +
                                                  std::to_string(kStaticScratchBytes / (2 * sizeof(c64))) +
                                                  " blocks (" + std::to_string(grid->NE) + " x " +
                                                  std::to_string(grid->Nk) + " given)");
    c64* traces = (c64*)static_scratch(ctx->device);
    if (!traces) return fail(ctx, QCLHIP_ERR_RUNTIME, "the fixed scratch of this device was not set up");
    return current_green_impl(ctx, grid, h, nfam, famR, famL, famG, target, GR, GL, GG, A, scale, out_host,
                              nullptr, false, traces);
}

int qclhip_current_green_scalar_lesser(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, int nfam,
                                       const double* const* famR, const double* const* famL,
                                       const double* const* famG, double target, double* GR, double* GL,
                                       double* GG, double* A, double* scale, double* out_host) {
    return current_green_impl(ctx, grid, h, nfam, famR, famL, famG, target, GR, GL, GG, A, scale, out_host,
                              nullptr, true);
}

int qclhip_number_functional(qclhip_ctx* ctx, const qclhip_grid* grid, const double* G, double* value_host) {
    if (!ctx || !grid || !G || !value_host) return QCLHIP_ERR_ARGUMENT;
    const long long nblk = (long long)grid->NE * grid->Nk;
    c64* traces = (c64*)workspace(ctx, (size_t)nblk * sizeof(c64));
    if (!traces) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    hipLaunchKernelGGL(k_block_traces, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const c64*)G, nblk, grid->Nb, traces);
    hipLaunchKernelGGL(k_number_pair, dim3(1), dim3(kReduceThreads), 0, ctx->stream, traces, 1, grid->NE,
                       grid->Nk, grid->wE, grid->wk, grid->gs, 1.0, 0, realness_limit(grid), ctx->coef,
                       ctx->status);
    int rc = launch_check(ctx, "number_functional");
    if (rc) return rc;
    bool deferred = false;
    rc = defer_result(ctx, (const double*)((const char*)ctx->coef + offsetof(PairCoefficients, number)), 1,
                      &deferred);
    if (rc || deferred) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, ctx->coef, sizeof(PairCoefficients),
                                       hipMemcpyDeviceToHost, ctx->stream));
    rc = qclhip_ctx_synchronize(ctx);
    if (rc) return rc;
    *value_host = ((const PairCoefficients*)ctx->host_scalars)->number;
    return QCLHIP_OK;
}

int qclhip_spectral(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GR, double* A) {
    if (!ctx || !grid || !GR || !A) return QCLHIP_ERR_ARGUMENT;
    if (GR == A) return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    const long long nblk = (long long)grid->NE * grid->Nk;
    const long long n = nblk * grid->Nb * grid->Nb;
    hipLaunchKernelGGL(k_spectral, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream, nblk,
                       grid->Nb, (const c64*)GR, (c64*)A);
    return launch_check(ctx, "k_spectral");
}

int qclhip_green_state(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GR, const double* GL, double* GG,
                       double* A) {
    if (!ctx || !grid || !GR || (GG && !GL)) return QCLHIP_ERR_ARGUMENT;
    if (GG == GR || A == GR || (GG && GG == A)) return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    const long long nblk = (long long)grid->NE * grid->Nk;
    const long long n = nblk * grid->Nb * grid->Nb;
    hipLaunchKernelGGL(k_green_state, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream, nblk,
                       grid->Nb, (const c64*)GR, (const c64*)GL, (c64*)GG, (c64*)A);
    return launch_check(ctx, "k_green_state");
}

int qclhip_normalize_pair(qclhip_ctx* ctx, int64_t n, double* GL, double* GG, double a, double c) {
    if (!ctx || !GL || !GG || n < 0) return QCLHIP_ERR_ARGUMENT;
    if (n == 0) return QCLHIP_OK;
    hipLaunchKernelGGL(k_normalize_pair, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream,
                       (long long)n, (c64*)GL, (c64*)GG, (const PairCoefficients*)nullptr, a, c);
    return launch_check(ctx, "k_normalize_pair");
}

int qclhip_seed_weights(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GR, double* q_host) {
    if (!ctx || !grid || !GR || !q_host) return QCLHIP_ERR_ARGUMENT;
    double* q = (double*)workspace(ctx, (size_t)grid->NE * sizeof(double));
    if (!q) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    hipLaunchKernelGGL(k_seed_weights, dim3((grid->NE + 255) / 256), dim3(256), 0, ctx->stream, grid->NE,
                       grid->Nk, grid->Nb, (const c64*)GR, grid->wE, grid->wk, grid->gs, q);
    int rc = launch_check(ctx, "k_seed_weights");
    if (rc) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(q_host, q, (size_t)grid->NE * sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->stream));
    rc = qclhip_ctx_synchronize(ctx);
    if (rc) return rc;
    for (int e = 0; e < grid->NE; ++e)
        if (!std::isfinite(q_host[e])) return fail(ctx, QCLHIP_ERR_DOMAIN, "seed spectral number weights must be finite");
    return QCLHIP_OK;
}

int qclhip_seed_lesser(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GR, double mu, double kBT,
                       double* GL) {
    if (!ctx || !grid || !GR || !GL) return QCLHIP_ERR_ARGUMENT;
    if (!(kBT > 0)) return fail(ctx, QCLHIP_ERR_ARGUMENT, "kBT must be positive");
    const long long nblk = (long long)grid->NE * grid->Nk;
    const long long n = nblk * grid->Nb * grid->Nb;
    hipLaunchKernelGGL(k_seed_lesser, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream,
                       grid->NE, nblk, grid->Nb, (const c64*)GR, grid->eps, mu, kBT, (c64*)GL);
    return launch_check(ctx, "k_seed_lesser");
}

int qclhip_sum_fields(qclhip_ctx* ctx, int64_t n, int nx, const double* const* x, double* out) {
    if (!ctx || !x || !out || nx < 1 || n < 0) return QCLHIP_ERR_ARGUMENT;
    int done = 0;
    while (done < nx) {
        SumArgs sa;
        sa.nx = std::min(16, nx - done);
        for (int f = 0; f < sa.nx; ++f) sa.x[f] = (const c64*)x[done + f];
        hipLaunchKernelGGL(k_sum_fields, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream,
                           (long long)n, sa, done > 0 ? 1 : 0, (c64*)out);
        done += sa.nx;
    }
    return launch_check(ctx, "k_sum_fields");
}

int qclhip_mix_fields(qclhip_ctx* ctx, int64_t n, int npairs, double* const* old_, const double* const* cand,
                      double alpha) {
    if (!ctx || !old_ || !cand || n < 0 || npairs < 0) return QCLHIP_ERR_ARGUMENT;
    if (!(alpha > 0 && alpha <= 1)) return fail(ctx, QCLHIP_ERR_ARGUMENT, "mixing factor must lie in (0,1]");
    for (int done = 0; done < npairs;) {
        MixArgs a;
        a.npairs = std::min(kMaxMixPairs, npairs - done);
        for (int p = 0; p < a.npairs; ++p) {
            a.old_[p] = (c64*)old_[done + p];
            a.cand[p] = (const c64*)cand[done + p];
        }
        hipLaunchKernelGGL(k_mix_many, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream,
                           (long long)n, a, alpha);
        done += a.npairs;
    }
    return launch_check(ctx, "k_mix_many");
}

int qclhip_mix(qclhip_ctx* ctx, int64_t n, double* old_, const double* cand, double alpha) {
    if (!ctx || !old_ || !cand || n < 0) return QCLHIP_ERR_ARGUMENT;
    if (!(alpha > 0 && alpha <= 1)) return fail(ctx, QCLHIP_ERR_ARGUMENT, "mixing factor must lie in (0,1]");
    hipLaunchKernelGGL(k_mix, dim3(grid_for(n, 256, ctx->sm_count)), dim3(256), 0, ctx->stream, (long long)n,
                       (c64*)old_, (const c64*)cand, alpha);
    return launch_check(ctx, "k_mix");
}

int qclhip_embedding_component(qclhip_ctx* ctx, const qclhip_grid* grid, const double* T, const double* X,
                               const qclhip_shift* shift, double* S) {
    if (!ctx || !grid || !T || !X || !shift || !S) return QCLHIP_ERR_ARGUMENT;
    if (shift->NE != grid->NE) return fail(ctx, QCLHIP_ERR_DIMENSION, "shift plan has the wrong energy size");
    if (X == S) return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    SWArgs a{};
    a.NE = grid->NE;
    a.Nk = grid->Nk;
    a.N = grid->Nb;
    a.T = (const c64*)T;
    a.X = (const c64*)X;
    a.left = shift->left;
    a.right = shift->right;
    a.wr = shift->wr;
    a.valid = shift->valid;
    a.out = (c64*)S;
    return sandwich_dispatch(ctx, a);
}

// Reverse of the embedding component: Xbar += S^T (T^dagger Sbar T); the
// shift acts on energy and T on the block indices, so the transpose of the
// shift is applied first, then the sandwich with T^dagger.
int qclhip_embedding_adjoint(qclhip_ctx* ctx, const qclhip_grid* grid, const double* T, const qclhip_shift* shift,
                             const double* Sbar, double* Xbar) {
    if (!ctx || !grid || !T || !shift || !Sbar || !Xbar) return QCLHIP_ERR_ARGUMENT;
    if (shift->NE != grid->NE) return fail(ctx, QCLHIP_ERR_DIMENSION, "shift plan has the wrong energy size");
    if (Sbar == Xbar) return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    const long long ncols = (long long)grid->Nk * grid->Nb * grid->Nb;
    c64* Z = (c64*)workspace(ctx, (size_t)grid->NE * ncols * sizeof(c64));
    if (!Z) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    int rc = shift_transpose(ctx, shift, ncols, (const c64*)Sbar, 1.0, 0, Z);
    if (rc) return rc;
    SWArgs a{};
    a.NE = grid->NE;
    a.Nk = grid->Nk;
    a.N = grid->Nb;
    a.T = (const c64*)T;
    a.X = Z;
    a.out = (c64*)Xbar;
    a.tadj = 1;
    a.accumulate = 1;
    return sandwich_dispatch(ctx, a);
}

}  // extern "C"
