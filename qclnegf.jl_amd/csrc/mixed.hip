// Matrix-core GF solve: the Keldysh products on MFMA (FP32 or FP64) and the
// mixed-precision Dyson step (FP32 Gauss-Jordan + FP64 residual refinement).
//
// Complex products go to the 16x16x4 MFMA tiles through the real embedding
//   [Re C; Im C] = [[Re A, -Im A], [Im A, Re A]] [Re B; Im B],
// a (2N x 2N)(2N x N) real product: for N = 24 that is 3 x 2 output tiles
// and 12 k-steps, with no padding in the row and k directions.
//
// Mixed precision, per block (A1 of the production operators,
// production_operators.jl:1337-1385, with the same D = (eps + i eta) I - h -
// Sigma^R and row-sum scale s):
//   X0 = inv(D/s) by Gauss-Jordan in FP32 (same pivot rule as the FP64 path),
//   then `refine` Newton steps  R = I - (D/s) X  (FP64 residual, matrix cores,
//   X0 and the FP32 correction sum Delta enter exactly),  Delta += X0 R (FP32),
//   and G^R = (X0 + Delta) / s in FP64.
// A guard bounds the final residual by ||R_1||^(2^refine) (infinity norm,
// over-estimated through |Re| + |Im|): a block whose bound exceeds
// refine_tol, or whose FP32 elimination meets a zero pivot, is flagged and
// re-solved (with its tile) by the FP64 Dyson kernel, which also reports
// truly singular blocks.
#include "dk_common.hpp"

namespace qclhip {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

// 16x16x4 tiles. Operands: lane l holds A[l%16][l/16] and B[l/16][l%16].
// Result element v of lane l: FP32 row 4*(l/16)+v, FP64 row (l/16)+4*v;
// column l%16 in both (checked on gfx950).
template <>
struct Mfma<float> {
    using acc = f32x4;
    using cplx = float2;
    __device__ static inline acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lk, int v) { return 4 * lk + v; }
};

template <>
struct Mfma<double> {
    using acc = f64x4;
    using cplx = double2;
    __device__ static inline acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lk, int v) { return lk + 4 * v; }
};

template <typename T, int RT, int CT>
__device__ inline void acc_zero(typename Mfma<T>::acc (&acc)[RT][CT]) {
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int u = 0; u < CT; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[t][u][v] = (T)0;
}

// acc += A op(B) for complex N x N matrices in LDS (row-major, strides lda,
// ldb), op(B) = B or B^dagger (ADJ). RT >= ceil(2N/16), CT >= ceil(N/16).
// Must be called by all 64 lanes of the wave.
template <typename T, int RT, int CT, bool ADJ, typename VA, typename VB>
__device__ inline void wave_cgemm(const VA* __restrict__ A, int lda, const VB* __restrict__ Bm, int ldb, int N,
                                  int lane, typename Mfma<T>::acc (&acc)[RT][CT]) {
    const int li = lane & 15, lk = lane >> 4;
    const int n2 = 2 * N;
    const int nsteps = (n2 + 3) >> 2;
    for (int s = 0; s < nsteps; ++s) {
        const int q = 4 * s + lk;
        const bool qv = q < n2;
        const bool qim = q >= N;
        const int qc = qim ? q - N : q;
        T a[RT], b[CT];
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            const int r = 16 * t + li;
            T v = (T)0;
            if (qv && r < n2) {
                const bool rim = r >= N;
                const VA z = A[(rim ? r - N : r) * lda + qc];
                v = (rim == qim) ? (T)z.x : (rim ? (T)z.y : -(T)z.y);
            }
            a[t] = v;
        }
#pragma unroll
        for (int u = 0; u < CT; ++u) {
            const int j = 16 * u + li;
            T v = (T)0;
            if (qv && j < N) {
                if (ADJ) {
                    const VB z = Bm[j * ldb + qc];  // conj(B[j][qc])
                    v = qim ? -(T)z.y : (T)z.x;
                } else {
                    const VB z = Bm[qc * ldb + j];
                    v = qim ? (T)z.y : (T)z.x;
                }
            }
            b[u] = v;
        }
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int u = 0; u < CT; ++u)
                if (16 * t < n2 && 16 * u < N) acc[t][u] = Mfma<T>::mma(a[t], b[u], acc[t][u]);
    }
}

// Visit this lane's elements of C: f(row, col, imag, value, t, u, v).
template <typename T, int RT, int CT, typename F>
__device__ inline void acc_for_each(const typename Mfma<T>::acc (&acc)[RT][CT], int N, int lane, F f) {
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int u = 0; u < CT; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * t + Mfma<T>::row(lk, v);
                const int j = 16 * u + li;
                if (r < 2 * N && j < N) {
                    const bool im = r >= N;
                    f(im ? r - N : r, j, im, acc[t][u][v], t, u, v);
                }
            }
}

template <typename V, typename S>
__device__ inline void set_part(V& z, bool im, S value) {
    if (im)
        z.y = value;
    else
        z.x = value;
}

// ---------------------------------------------------------------------------
// Keldysh products on the matrix cores: G<raw = G^R S< G^R^dagger and
// G>raw likewise (production_state.jl:24-48), plus the block traces. Tiles
// move as in k_keldysh; the LDS copies hold type T.
template <typename T, int B, int W>
__global__ __launch_bounds__(64 * W) void k_keldysh_mfma(DKArgs p) {
    using V = typename Mfma<T>::cplx;
    using Acc = typename Mfma<T>::acc;
    constexpr int NP = 8 * B;
    constexpr int LD = NP + 1;
    constexpr int RT = B, CT = (B + 1) / 2;
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
    const long long blk = blk0 + wave;
    const int wstride = TileStride<B>::value;
    V* G0 = reinterpret_cast<V*>(smem);
    V* X0 = G0 + (size_t)W * wstride;
    V* Gs = G0 + (size_t)wave * wstride;
    V* Xs = X0 + (size_t)wave * wstride;
    const int N = p.N;
    const size_t S = (size_t)nblk;
    const bool active = wave < nvalid;
    const unsigned mask = (nvalid >= 32) ? 0xffffffffu : ((1u << nvalid) - 1u);
    const c64* gr_fam[1] = {p.GR};
    tile_load_as<W, LD, V>(G0, wstride, gr_fam, 1, blk0, nvalid, N, S);
    for (int comp = 0; comp < 2; ++comp) {
        const c64* const* fam = comp ? p.famG : p.famL;
        const int nf = comp ? p.nfamG : p.nfamL;
        c64* dst = comp ? p.GG : p.GL;
        tile_load_as<W, LD, V>(X0, wstride, fam, nf, blk0, nvalid, N, S);
        __syncthreads();
        if (active) {
            Acc acc[RT][CT];
            acc_zero<T, RT, CT>(acc);
            wave_cgemm<T, RT, CT, false>(Gs, LD, Xs, LD, N, lane, acc);
            wave_sync();
            acc_for_each<T, RT, CT>(acc, N, lane, [&](int r, int j, bool im, T v, int, int, int) {
                set_part(Xs[r * LD + j], im, v);
            });
            wave_sync();
            acc_zero<T, RT, CT>(acc);
            wave_cgemm<T, RT, CT, true>(Xs, LD, Gs, LD, N, lane, acc);
            wave_sync();
            double trx = 0.0, trY = 0.0;
            acc_for_each<T, RT, CT>(acc, N, lane, [&](int r, int j, bool im, T v, int, int, int) {
                set_part(Xs[r * LD + j], im, v);
                if (r == j) {
                    if (im)
                        trY += (double)v;
                    else
                        trx += (double)v;
                }
            });
            if (p.traces) {
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    trx += shfl_xor_d(trx, off);
                    trY += shfl_xor_d(trY, off);
                }
                if (lane == 0) p.traces[(size_t)comp * nblk + blk] = cmk(trx, trY);
            }
        }
        __syncthreads();
        tile_store_from<W, LD, V>(X0, wstride, dst, blk0, mask, N, S);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// FP32 Gauss-Jordan pieces (same scheme as the FP64 Dyson kernel: 8x8 lane
// grid, B x B register blocks, implicit partial pivoting).
__device__ inline float2 fmk(float re, float im) {
    float2 z;
    z.x = re;
    z.y = im;
    return z;
}
__device__ inline float2 fmul(float2 a, float2 b) { return fmk(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x)); }
__device__ inline float2 ffms(float2 c, float2 a, float2 b) {
    return fmk(fmaf(-a.x, b.x, fmaf(a.y, b.y, c.x)), fmaf(-a.x, b.y, fmaf(-a.y, b.x, c.y)));
}
__device__ inline float2 frecip(float2 z) {
    if (fabsf(z.x) >= fabsf(z.y)) {
        const float r = z.y / z.x, d = z.x + z.y * r;
        return fmk(1.0f / d, -r / d);
    }
    const float r = z.x / z.y, d = z.y + z.x * r;
    return fmk(r / d, -1.0f / d);
}

template <int CTRL>
__device__ inline void dpp_argmax_step_f(float& mag, int& idx) {
    const float om = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(mag), CTRL, 0xF, 0xF, false));
    const int oi = __builtin_amdgcn_update_dpp(0, idx, CTRL, 0xF, 0xF, false);
    if (om > mag || (om == mag && oi < idx)) {
        mag = om;
        idx = oi;
    }
}

__device__ inline void argmax8_f(float& mag, int& idx) {
    dpp_argmax_step_f<0xB1>(mag, idx);
    dpp_argmax_step_f<0x4E>(mag, idx);
    dpp_argmax_step_f<0x141>(mag, idx);
}

// In-place inverse of the register-held matrix; the result stays in the
// pivoted order (pivrow[k] = row chosen at step k, steprow[row] = k).
// Returns false on a zero pivot. Wave-uniform result.
template <int B>
__device__ inline bool gj_invert_f32(float2 (&A)[B][B], int N, int lane, float2* colbuf, float2* rowbuf, int* pivrow,
                                     int* steprow) {
    const int lr = lane & 7, lc = lane >> 3;
    unsigned used = 0;
    for (int ck = 0; ck < 8; ++ck) {
        if (ck * B >= N) break;
#pragma unroll
        for (int kk = 0; kk < B; ++kk) {
            const int k = ck * B + kk;
            if (k >= N) break;
            float mag = -1.0f;
            int idx = 0x7fffffff;
            if (lc == ck) {
#pragma unroll
                for (int r = 0; r < B; ++r) {
                    const int i = lr * B + r;
                    if (i < N) {
                        colbuf[i] = A[r][kk];
                        if (!((used >> r) & 1u)) {
                            const float mg = fabsf(A[r][kk].x) + fabsf(A[r][kk].y);
                            if (mg > mag) {
                                mag = mg;
                                idx = i;
                            }
                        }
                    }
                }
            }
            argmax8_f(mag, idx);
            const int piv = __builtin_amdgcn_readlane(idx, 8 * ck);
            const float pmag = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mag), 8 * ck));
            if (!(pmag > 0.0f) || piv >= N) return false;
            const int pg = piv / B, pr = piv - pg * B;
            if (lr == pg) {
#pragma unroll
                for (int r = 0; r < B; ++r)
                    if (r == pr) {
#pragma unroll
                        for (int c = 0; c < B; ++c) {
                            const int j = lc * B + c;
                            if (j < N) rowbuf[j] = A[r][c];
                        }
                    }
            }
            if (lane == 0) {
                pivrow[k] = piv;
                steprow[piv] = k;
            }
            wave_sync();
            const float2 pv = colbuf[piv];
            float2 f[B], R[B];
#pragma unroll
            for (int r = 0; r < B; ++r) {
                const int i = lr * B + r;
                f[r] = (i < N) ? colbuf[i] : fmk(0.0f, 0.0f);
            }
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const int j = lc * B + c;
                R[c] = (j < N) ? rowbuf[j] : fmk(0.0f, 0.0f);
            }
            const float2 inv = frecip(pv);
#pragma unroll
            for (int c = 0; c < B; ++c) R[c] = fmul(R[c], inv);
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int c = 0; c < B; ++c) A[r][c] = ffms(A[r][c], f[r], R[c]);
            if (lr == pg) {
#pragma unroll
                for (int r = 0; r < B; ++r)
                    if (r == pr) {
#pragma unroll
                        for (int c = 0; c < B; ++c) A[r][c] = R[c];
                        used |= 1u << r;
                    }
            }
            if (lc == ck) {
#pragma unroll
                for (int r = 0; r < B; ++r) {
                    const float2 tt = fmul(f[r], inv);
                    A[r][kk] = (lr == pg && r == pr) ? inv : fmk(-tt.x, -tt.y);
                }
            }
            wave_sync();
        }
    }
    return true;
}

// LDS per wave: D (FP64, the Sigma^R tile in, D/s for the residual, G^R
// out), X0 (FP32), R (FP32; also the Gauss-Jordan scratch and Delta).
template <int B, int W>
struct MxShape {
    static constexpr int NP = 8 * B;
    static constexpr int LD = NP + 1;
    static constexpr size_t per_wave = (size_t)TileStride<B>::value * (sizeof(c64) + 2 * sizeof(float2));
    static constexpr size_t bytes = per_wave * W;
};

template <int B, int W>
__global__ __launch_bounds__(64 * W) void k_dyson_mixed(DKArgs p) {
    using Acc32 = Mfma<float>::acc;
    using Acc64 = Mfma<double>::acc;
    constexpr int NP = 8 * B;
    constexpr int LD = NP + 1;
    constexpr int RT = B, CT = (B + 1) / 2;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ unsigned s_ok;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nte = (p.NE + W - 1) / W;
    const long long t = xcd_tile(blockIdx.x, gridDim.x);
    const int e0 = (int)(t % nte) * W;
    const int m = (int)(t / nte);
    const int nvalid = min(W, p.NE - e0);
    const long long nblk = (long long)p.NE * p.Nk;
    const long long blk0 = e0 + (long long)p.NE * m;
    const int e = e0 + wave;
    const long long blk = blk0 + wave;
    const int wstride = TileStride<B>::value;
    c64* D0 = reinterpret_cast<c64*>(smem);
    float2* X0all = reinterpret_cast<float2*>(D0 + (size_t)W * wstride);
    float2* R0all = X0all + (size_t)W * wstride;
    c64* Ds = D0 + (size_t)wave * wstride;
    float2* Xs = X0all + (size_t)wave * wstride;
    float2* Rs = R0all + (size_t)wave * wstride;
    float2* colbuf = Rs;
    float2* rowbuf = Rs + NP;
    int* pivrow = reinterpret_cast<int*>(Rs + 2 * NP);
    int* steprow = pivrow + NP;
    const int N = p.N;
    const size_t S = (size_t)nblk;
    const int lr = lane & 7, lc = lane >> 3;
    if (threadIdx.x == 0) s_ok = 0;
    if (p.nfamR) tile_load<W, LD>(D0, wstride, p.famR, p.nfamR, blk0, nvalid, N, S);
    __syncthreads();
    const bool valid = wave < nvalid;
    bool ok = valid;
    double s = 0.0;
    float2 A[B][B];
    if (ok) {
        const double ee = p.eps[e];
        c64 Dv[B][B];
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const int i = lr * B + r, j = lc * B + c;
                c64 v = cmk(0.0, 0.0);
                if (i < N && j < N) {
                    const c64 hv = p.h[(size_t)m + (size_t)p.Nk * (i + (size_t)N * j)];
                    const c64 sv = p.nfamR ? Ds[i * LD + j] : cmk(0.0, 0.0);
                    const double dre = (i == j) ? ee : 0.0, dim = (i == j) ? p.eta : 0.0;
                    v = cmk(dre - hv.x - sv.x, dim - hv.y - sv.y);
                }
                Dv[r][c] = v;
            }
#pragma unroll
        for (int r = 0; r < B; ++r) {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < B; ++c) acc += hypot(Dv[r][c].x, Dv[r][c].y);
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
        if (ok) {
            const double rs = 1.0 / s;
#pragma unroll
            for (int r = 0; r < B; ++r)
#pragma unroll
                for (int c = 0; c < B; ++c) {
                    const int i = lr * B + r, j = lc * B + c;
                    const c64 a = cmk(Dv[r][c].x * rs, Dv[r][c].y * rs);
                    if (i < N && j < N) Ds[i * LD + j] = a;
                    A[r][c] = fmk((float)a.x, (float)a.y);
                }
        }
    }
    if (ok) ok = gj_invert_f32<B>(A, N, lane, colbuf, rowbuf, pivrow, steprow);
    if (ok) {
        // X0 = inv(D/s), un-permuted into LDS.
        int si[B], pj[B];
#pragma unroll
        for (int r = 0; r < B; ++r) si[r] = (lr * B + r < N) ? steprow[lr * B + r] : 0;
#pragma unroll
        for (int c = 0; c < B; ++c) pj[c] = (lc * B + c < N) ? pivrow[lc * B + c] : 0;
#pragma unroll
        for (int r = 0; r < B; ++r)
#pragma unroll
            for (int c = 0; c < B; ++c)
                if (lr * B + r < N && lc * B + c < N) Xs[si[r] * LD + pj[c]] = A[r][c];
        wave_sync();
        Acc32 dl[RT][CT];
        acc_zero<float, RT, CT>(dl);
        for (int it = 0; it < p.refine && ok; ++it) {
            Acc64 r64[RT][CT];
            acc_zero<double, RT, CT>(r64);
            wave_cgemm<double, RT, CT, false>(Ds, LD, Xs, LD, N, lane, r64);
            if (it > 0) {
                // Delta -> LDS (FP32, exact), then r64 += (D/s) Delta.
                acc_for_each<float, RT, CT>(dl, N, lane, [&](int r, int j, bool im, float v, int, int, int) {
                    set_part(Rs[r * LD + j], im, v);
                });
                wave_sync();
                wave_cgemm<double, RT, CT, false>(Ds, LD, Rs, LD, N, lane, r64);
            }
            wave_sync();
            // R = I - (D/s) X, rounded to FP32; row sums of |R| for the guard.
            double rowsum[RT][4];
#pragma unroll
            for (int a = 0; a < RT; ++a)
#pragma unroll
                for (int v = 0; v < 4; ++v) rowsum[a][v] = 0.0;
            acc_for_each<double, RT, CT>(r64, N, lane, [&](int r, int j, bool im, double v, int ti, int, int vi) {
                const double rv = ((r == j && !im) ? 1.0 : 0.0) - v;
                set_part(Rs[r * LD + j], im, (float)rv);
                rowsum[ti][vi] += fabs(rv);
            });
            if (it == 0) {
                // Sum over the 16 columns held by lanes of equal l/16, then max.
                double nr = 0.0;
#pragma unroll
                for (int a = 0; a < RT; ++a)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        double x = rowsum[a][v];
                        x += shfl_xor_d(x, 1);
                        x += shfl_xor_d(x, 2);
                        x += shfl_xor_d(x, 4);
                        x += shfl_xor_d(x, 8);
                        nr = fmax(nr, x);
                    }
                nr = fmax(nr, shfl_xor_d(nr, 16));
                nr = fmax(nr, shfl_xor_d(nr, 32));
                // ||R||_inf <= 2 max over embedded rows (Re and Im rows of one row).
                double bound = 2.0 * readlane_d(nr, 0);
                for (int k = 0; k < p.refine; ++k) bound *= bound;
                if (!(bound <= p.refine_tol)) ok = false;
            }
            wave_sync();
            if (!ok) break;
            Acc32 c32[RT][CT];
            acc_zero<float, RT, CT>(c32);
            wave_cgemm<float, RT, CT, false>(Xs, LD, Rs, LD, N, lane, c32);
#pragma unroll
            for (int a = 0; a < RT; ++a)
#pragma unroll
                for (int u = 0; u < CT; ++u) dl[a][u] += c32[a][u];
        }
        if (ok) {
            // G^R = (X0 + Delta) / s in FP64, staged in the D tile.
            const double rs = 1.0 / s;
            acc_for_each<float, RT, CT>(dl, N, lane, [&](int r, int j, bool im, float v, int, int, int) {
                const float2 x0 = Xs[r * LD + j];
                const double x = ((double)(im ? x0.y : x0.x) + (double)v) * rs;
                set_part(Ds[r * LD + j], im, x);
            });
            if (lane == 0) {
                atomicOr(&s_ok, 1u << wave);
                if (p.scale) p.scale[blk] = s;
            }
        }
    }
    if (valid && lane == 0) p.block_flags[blk] = ok ? 0u : 1u;
    __syncthreads();
    tile_store<W, LD>(D0, wstride, p.GR, blk0, s_ok, N, S);
}

template <typename T, int B>
int launch_keldysh_mfma_impl(qclhip_ctx* ctx, const DKArgs& a) {
    using V = typename Mfma<T>::cplx;
    constexpr size_t per_wave = 2 * (size_t)TileStride<B>::value * sizeof(V);
    constexpr int W = per_wave * 4 <= 80 * 1024 ? 4 : (per_wave * 2 <= 80 * 1024 ? 2 : 1);
    const size_t lds = per_wave * W;
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_keldysh_mfma<T, B, W>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const long long tiles = (long long)((a.NE + W - 1) / W) * a.Nk;
    hipLaunchKernelGGL((k_keldysh_mfma<T, B, W>), dim3((unsigned)tiles), dim3(64 * W), lds, ctx->stream, a);
    return launch_check(ctx, "k_keldysh_mfma");
}

template <int B>
int launch_mixed_impl(qclhip_ctx* ctx, const DKArgs& a) {
    // Energies per workgroup: 4 when the tiles fit in half the LDS of a CU.
    constexpr int W = MxShape<B, 4>::bytes <= 80 * 1024 ? 4 : (MxShape<B, 2>::bytes <= 80 * 1024 ? 2 : 1);
    {
        const size_t lds = MxShape<B, W>::bytes;
        if (lds > ctx->max_lds_per_block)
            return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "basis size needs more LDS than the device offers");
        if (lds > 65536)
            QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_dyson_mixed<B, W>,
                                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const long long tiles = (long long)((a.NE + W - 1) / W) * a.Nk;
        hipLaunchKernelGGL((k_dyson_mixed<B, W>), dim3((unsigned)tiles), dim3(64 * W), lds, ctx->stream, a);
        int rc = launch_check(ctx, "k_dyson_mixed");
        if (rc) return rc;
    }
    // FP64 re-solve of the tiles that hold a flagged block.
    return launch_dyson_fp64(ctx, a);
}

#define QCLHIP_FOR_B(N, CALL)                                                                      \
    do {                                                                                           \
        if ((N) <= 8) return CALL(1);                                                              \
        if ((N) <= 16) return CALL(2);                                                             \
        if ((N) <= 24) return CALL(3);                                                             \
        if ((N) <= 32) return CALL(4);                                                             \
        if ((N) <= 40) return CALL(5);                                                             \
        if ((N) <= 48) return CALL(6);                                                             \
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "Nb > 48 is not supported by this build");      \
    } while (0)

int launch_mixed_dyson(qclhip_ctx* ctx, const DKArgs& a) {
#define QCLHIP_MX(b) launch_mixed_impl<b>(ctx, a)
    QCLHIP_FOR_B(a.N, QCLHIP_MX);
#undef QCLHIP_MX
}

int launch_keldysh_mfma(qclhip_ctx* ctx, const DKArgs& a, bool fp32) {
#define QCLHIP_K32(b) launch_keldysh_mfma_impl<float, b>(ctx, a)
#define QCLHIP_K64(b) launch_keldysh_mfma_impl<double, b>(ctx, a)
    if (fp32) QCLHIP_FOR_B(a.N, QCLHIP_K32);
    QCLHIP_FOR_B(a.N, QCLHIP_K64);
#undef QCLHIP_K32
#undef QCLHIP_K64
}

// Number of flagged blocks of the last mixed-precision Dyson call
// (deterministic: per-thread partial counts, then a fixed-order tree).
__global__ __launch_bounds__(1024) void k_count_flags(const unsigned* flags, long long n,
                                                       unsigned long long* out) {
    __shared__ unsigned long long red[1024];
    unsigned long long c = 0;
    for (long long i = threadIdx.x; i < n; i += 1024) c += flags[i] ? 1u : 0u;
    red[threadIdx.x] = c;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace qclhip

using namespace qclhip;

extern "C" int qclhip_grid_refine_fallbacks(qclhip_ctx* ctx, const qclhip_grid* grid, int64_t* count_host) {
    if (!ctx || !grid || !count_host) return QCLHIP_ERR_ARGUMENT;
    if (!grid->block_flags) {
        *count_host = 0;
        return QCLHIP_OK;
    }
    auto* d = (unsigned long long*)workspace(ctx, sizeof(unsigned long long));
    if (!d) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    hipLaunchKernelGGL(k_count_flags, dim3(1), dim3(1024), 0, ctx->stream, (const unsigned*)grid->block_flags,
                       (long long)grid->NE * grid->Nk, d);
    int rc = launch_check(ctx, "k_count_flags");
    if (rc) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, d, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                       ctx->stream));
    rc = qclhip_ctx_synchronize(ctx);
    if (rc) return rc;
    *count_host = (int64_t)(*(const unsigned long long*)ctx->host_scalars);
    return QCLHIP_OK;
}
