// FFT Hilbert transform of every column alpha = (m,a,b) of a field
// (fft_parallel.jl:27-291; production_operators.jl:735-884).
//
// Lambda_e = sum_{e' != e} wE_e' Gamma_e' / (2 pi (e - e') de) is evaluated as
// the zero-padded linear convolution of wE*Gamma with the kernel
// 1/(2 pi lag de), lag = -(NE-1)..NE-1, through length-L FFTs
// (L = nextpow2(2NE-1)), extracting y[e + NE - 1]. The round-off witness
// evaluates the transform at 2L as well and reports
// ||L2 - L|| / (||L|| + 1e-14) (see k_hilbert).
//
// One workgroup per column. Each thread keeps 8 complex values in registers
// (T = L/8 threads, 16 values above L = 8192); every Stockham pass does
// radix-R butterflies in registers and exchanges data through LDS one real
// component at a time, so the LDS footprint is L doubles. FFT lengths are
// template parameters: all index arithmetic is shifts and masks.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "common.hpp"

struct qclhip_hilbert {
    const qclhip_grid* grid = nullptr;
    int NE = 0;
    long long ncols = 0;
    int L = 0;
    int verify = 0;
    qclhip::c64* twiddle = nullptr;  // exp(-2 pi i idx / L), idx < L
    // kernel spectrum for L, and (with the witness) the odd-frequency half
    // of the 2L spectrum: DFT_L(k_n exp(-i pi n / L))
    qclhip::c64* kf[2] = {nullptr, nullptr};
    qclhip::c64* phase = nullptr;  // exp(-i pi n / L), n < L (witness)
    int Lmax = 0;
    // Four-step path through HBM (L = L1 * L2, see k_fs_columns): the two
    // spectra above in row order, kp[k1 * L2 + k2] = kf[k1 + L1 * k2].
    qclhip::c64* kp[2] = {nullptr, nullptr};
    int L1 = 0, L2 = 0;
    int path = 0;  // QCLHIP_HILBERT_PATH_*; AUTO resolves by L
    // Identical columns each column stands for in the witness's norms (a
    // momentum-independent self-energy transformed once for every momentum)
    double col_mult = 1.0;
};

namespace qclhip {

namespace {


__constant__ double kC16[16] = {
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
__constant__ double kS16[16] = {
    0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
    1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
    0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
    -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};

template <int R>
__device__ constexpr int bitrev(int x) {
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

// In-register DFT of R points (radix-2 decimation in time), forward sign
// exp(-2 pi i jk/R), or its conjugate when INV.
template <int R, bool INV>
__device__ inline void dft(double* re, double* im) {
    double tr[R], ti[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        tr[i] = re[bitrev<R>(i)];
        ti[i] = im[bitrev<R>(i)];
    }
#pragma unroll
    for (int m = 2; m <= R; m <<= 1) {
#pragma unroll
        for (int s = 0; s < R; s += m) {
#pragma unroll
            for (int j = 0; j < m / 2; ++j) {
                const int idx = j * (16 / m);
                const double wc = kC16[idx];
                const double ws = INV ? kS16[idx] : -kS16[idx];
                const int a = s + j, b = s + j + m / 2;
                double xr, xi;
                if (j == 0) {
                    xr = tr[b];
                    xi = ti[b];
                } else {
                    xr = wc * tr[b] - ws * ti[b];
                    xi = wc * ti[b] + ws * tr[b];
                }
                const double ur = tr[a], ui = ti[a];
                tr[a] = ur + xr;
                ti[a] = ui + xi;
                tr[b] = ur - xr;
                ti[b] = ui - xi;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i) {
        re[i] = tr[i];
        im[i] = ti[i];
    }
}

// ---- compile-time Stockham plan for a power-of-two length L ---------------
// Passes use the base radix except the last, which takes the remaining factor; the
// inverse runs the radices in reverse order so that the forward outputs are
// already in the inverse's input order.
constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x >> 1); }
// 16 values per thread (radix-16 passes) up to L = 2048, 8 (radix 8; 512 and
// 1024 threads) at 4096 and 8192, where 16 would hold occupancy at two waves
// per SIMD, and 16 above.
constexpr int vals_per_thread(int L) { return L < 16 ? L : ((L == 4096 || L == 8192) ? 8 : 16); }
constexpr int threads_for(int L) { return L / vals_per_thread(L); }
constexpr int base_log(int L) { return ilog2c(vals_per_thread(L)); }
constexpr int npasses(int L) { return (ilog2c(L) + base_log(L) - 1) / base_log(L); }
constexpr int fwd_radix(int L, int p) {
    return p < npasses(L) - 1 ? vals_per_thread(L) : (1 << (ilog2c(L) - base_log(L) * (npasses(L) - 1)));
}
constexpr int inv_radix(int L, int q) { return fwd_radix(L, npasses(L) - 1 - q); }
constexpr int fwd_span(int L, int p) { return p == 0 ? 1 : fwd_span(L, p - 1) * fwd_radix(L, p - 1); }
constexpr int inv_span(int L, int q) { return q == 0 ? 1 : inv_span(L, q - 1) * inv_radix(L, q - 1); }

// One Stockham pass: butterfly j (inputs j + r L/R) with twiddles
// exp(-+2 pi i r (j mod NS) / (NS R)) and an in-register DFT of R points.
template <int L, int R, int NS, bool INV>
__device__ inline void stockham_pass(double* re, double* im, int t, const c64* __restrict__ tw, int twstride) {
    constexpr int V = vals_per_thread(L), T = threads_for(L);
#pragma unroll
    for (int i = 0; i < V / R; ++i) {
        const int j = t + i * T;
        if (NS > 1) {
            const int k = j & (NS - 1);
            const int step = (L / (NS * R)) * twstride;
            if constexpr (R == 8 || R == 16) {
                // w^1, w^2, w^4 (and w^8) from the table; each other power is
                // formed where it is applied, as the product of the base powers
                // of its binary digits (at most three products, a few ulps),
                // against 15 table loads per butterfly
                c64 b[4];
                b[0] = tw[k * step];
                b[1] = tw[2 * k * step];
                b[2] = tw[4 * k * step];
                b[3] = R == 16 ? tw[8 * k * step] : cmk(1.0, 0.0);
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    c64 w = cmk(1.0, 0.0);
                    bool first = true;
#pragma unroll
                    for (int bit = 0; bit < 4; ++bit)
                        if (r & (1 << bit)) {
                            w = first ? b[bit] : cmul(w, b[bit]);
                            first = false;
                        }
                    const double wi = INV ? -w.y : w.y;
                    const double xr = re[i * R + r], xi = im[i * R + r];
                    re[i * R + r] = w.x * xr - wi * xi;
                    im[i * R + r] = w.x * xi + wi * xr;
                }
            } else {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const c64 w = tw[r * k * step];
                    const double wi = INV ? -w.y : w.y;
                    const double xr = re[i * R + r], xi = im[i * R + r];
                    re[i * R + r] = w.x * xr - wi * xi;
                    im[i * R + r] = w.x * xi + wi * xr;
                }
            }
        }
        dft<R, INV>(re + i * R, im + i * R);
    }
}

// LDS slot of element x: one pad double per 32 keeps the stride-R writes of
// the first pass off a single bank.
__device__ constexpr int lds_slot(int x) { return x + (x >> 5); }

// At L = 8192 the registers of a 1024-thread workgroup (88-90 VGPRs) allow
// one workgroup per CU whatever its LDS, so the exchange moves whole complex
// values (16-byte slots, 128 KiB): one barrier pair per pass instead of two
// and half the LDS instructions (5% of the transform; same values). The
// slot of element x has its low three bits XORed with bits 3-5, so that the
// stride-8 writes of the first pass and the unit-stride reads of 8
// consecutive lanes fall on different banks.
constexpr bool complex_exchange(int L) { return L == 8192; }
__device__ constexpr int cslot(int x) { return x ^ ((x >> 3) & 7); }

// Redistribute through LDS from the outputs of a pass (radix R, span NS) to
// the inputs of the next pass (radix RN): one real component at a time, or
// complex values where complex_exchange(L).
template <int L, int R, int NS, int RN>
__device__ inline void stockham_exchange(double* re, double* im, double* lds, int t) {
    constexpr int V = vals_per_thread(L), T = threads_for(L);
    if constexpr (complex_exchange(L)) {
        c64* cl = reinterpret_cast<c64*>(lds);
        __syncthreads();
#pragma unroll
        for (int f = 0; f < V; ++f) {
            const int i = f / R, r = f % R;
            const int j = t + i * T;
            cl[cslot((j / NS) * NS * R + (j & (NS - 1)) + r * NS)] = cmk(re[f], im[f]);
        }
        __syncthreads();
#pragma unroll
        for (int f = 0; f < V; ++f) {
            const int i = f / RN, r = f % RN;
            const int j = t + i * T;
            const c64 v = cl[cslot(j + r * (L / RN))];
            re[f] = v.x;
            im[f] = v.y;
        }
        return;
    }
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
        double* v = comp ? im : re;
        __syncthreads();
#pragma unroll
        for (int f = 0; f < V; ++f) {
            const int i = f / R, r = f % R;
            const int j = t + i * T;
            lds[lds_slot((j / NS) * NS * R + (j & (NS - 1)) + r * NS)] = v[f];
        }
        __syncthreads();
#pragma unroll
        for (int f = 0; f < V; ++f) {
            const int i = f / RN, r = f % RN;
            const int j = t + i * T;
            v[f] = lds[lds_slot(j + r * (L / RN))];
        }
    }
}

template <int L, int P, bool INV>
struct StockhamPasses {
    static __device__ inline void run(double* re, double* im, double* lds, int t, const c64* tw, int twstride) {
        constexpr int NP = npasses(L);
        constexpr int R = INV ? inv_radix(L, P) : fwd_radix(L, P);
        constexpr int NS = INV ? inv_span(L, P) : fwd_span(L, P);
        stockham_pass<L, R, NS, INV>(re, im, t, tw, twstride);
        if constexpr (P + 1 < NP) {
            constexpr int RN = INV ? inv_radix(L, P + 1) : fwd_radix(L, P + 1);
            stockham_exchange<L, R, NS, RN>(re, im, lds, t);
            StockhamPasses<L, P + 1, INV>::run(re, im, lds, t, tw, twstride);
        }
    }
};

struct HilbertArgs {
    int NE;
    long long ncols;
    const double* wE;
    const c64* gamma;   // Gamma input, or NULL when (SL, SG) are given
    const c64* SL;
    const c64* SG;
    c64* out;           // Lambda (mode 0) or S^R = Lambda - i Gamma/2 (mode 1)
    double* partial;    // witness: [2][ncols] sums |L2-L|^2 and |L|^2
    const c64* tw;
    int twstride;
    const c64* kf;
    int mode;           // 0: Lambda, 1: S^R, 3: adjoint of S^R
    c64* out2;          // adjoint mode: S>_bar (accumulated); out holds S<_bar
    const c64* kodd;    // witness: odd half of the 2L kernel spectrum (NULL: no witness)
    const c64* phase;   // witness: exp(-i pi n / L)
    int stash;          // mode 1: Gamma kept in LDS behind the FFT buffer for the output
};

// LDS of the FFT exchange (and of the witness reduction), in doubles,
// rounded up to a 16-byte multiple for the Gamma stash behind it.
template <int L>
constexpr int fft_lds_len() {
    if (complex_exchange(L)) return 2 * L;
    return (((L + L / 32) > 2 * threads_for(L) ? (L + L / 32) : 2 * threads_for(L)) + 1) / 2 * 2;
}

__device__ inline c64 load_gamma(const HilbertArgs& a, size_t idx) {
    if (a.gamma) return a.gamma[idx];
    c64 l = a.SL[idx], g = a.SG[idx];
    // Gamma = i (S> - S<)
    return cmk(-(g.y - l.y), g.x - l.x);
}

// Odd-frequency half of the 2L transform of one column (see k_hilbert):
// x_n exp(-i pi n/L) through the L transform with the odd kernel spectrum;
// writes exp(i pi n/L) y'_n / L to the output positions of the column.
template <int L>
__global__ __launch_bounds__(threads_for(L)) void k_hilbert_odd(HilbertArgs a) {
    constexpr int V = vals_per_thread(L), T = threads_for(L);
    constexpr int NPASS = npasses(L);
    constexpr int R0 = fwd_radix(L, 0), RL = fwd_radix(L, NPASS - 1);
    extern __shared__ __align__(16) double lds[];
    const int t = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * a.NE;
    const int NE = a.NE;
    const double invL = 1.0 / (double)L;
    double re[V], im[V];
#pragma unroll
    for (int f = 0; f < V; ++f) {
        const int i = f / R0, r = f % R0;
        const int n = t + i * T + r * (L / R0);
        double xr = 0.0, xi = 0.0;
        if (n < NE) {
            const c64 gm = load_gamma(a, base + n);
            const double we = a.wE[n];
            const c64 ph = a.phase[n];
            const double gr = we * gm.x, gi = we * gm.y;
            xr = gr * ph.x - gi * ph.y;
            xi = gr * ph.y + gi * ph.x;
        }
        re[f] = xr;
        im[f] = xi;
    }
    StockhamPasses<L, 0, false>::run(re, im, lds, t, a.tw, a.twstride);
#pragma unroll
    for (int f = 0; f < V; ++f) {
        const int i = f / RL, r = f % RL;
        const c64 k = a.kodd[t + i * T + r * (L / RL)];
        const double xr = re[f], xi = im[f];
        re[f] = k.x * xr - k.y * xi;
        im[f] = k.x * xi + k.y * xr;
    }
    StockhamPasses<L, 0, true>::run(re, im, lds, t, a.tw, a.twstride);
#pragma unroll
    for (int f = 0; f < V; ++f) {
        const int i = f / R0, r = f % R0;
        const int n = t + i * T + r * (L / R0);
        const int e = n - (NE - 1);
        if (e < 0 || e >= NE) continue;
        const c64 ph = a.phase[n];  // exp(+i pi n/L) = conj(ph)
        const double yr = re[f] * invL, yi = im[f] * invL;
        a.out[base + e] = cmk(yr * ph.x + yi * ph.y, yi * ph.x - yr * ph.y);
    }
}

// The 2L witness: the 2L DFT of the zero-padded column splits into even
// outputs, the L transform itself, and odd outputs, the L transform of
// x_n exp(-i pi n/L) (k_hilbert_odd, run first, or this kernel itself when
// Gamma is stashed in LDS; either parks exp(i pi n/L) y'_n in the output
// field). With the odd half of the 2L
// kernel spectrum, the 2L result at n < L is (y_n + exp(i pi n/L) y'_n)/2, so
// Lambda2 - Lambda = (exp(i pi n/L) y'_n - y_n)/2 compares two independently
// rounded evaluations of the same convolution, as the reference's 2L
// transform does, at the cost of one more length-L transform instead of a
// length-2L one.
template <int L, bool FUSE>
__global__ __launch_bounds__(threads_for(L)) void k_hilbert(HilbertArgs a) {
    constexpr int V = vals_per_thread(L), T = threads_for(L);
    constexpr int NPASS = npasses(L);
    constexpr int R0 = fwd_radix(L, 0), RL = fwd_radix(L, NPASS - 1);
    extern __shared__ __align__(16) double lds[];
    const int t = threadIdx.x;
    const long long col = blockIdx.x;
    const int NE = a.NE;
    const size_t base = (size_t)col * NE;
    const double invL = 1.0 / (double)L;
    const bool witness = a.kodd != nullptr;
    c64* gst = reinterpret_cast<c64*>(lds + fft_lds_len<L>());  // Gamma stash (a.stash)
    double dn = 0.0, nn = 0.0;
    double re[V], im[V];
    if constexpr (FUSE) {
        // Fused witness (stash and witness both on): Gamma is read once, into the stash, and the odd half
        // runs first (the computation of k_hilbert_odd), parking its outputs
        // in the output field; each thread reads back only positions it wrote.
#pragma unroll
        for (int f = 0; f < V; ++f) {
            const int i = f / R0, r = f % R0;
            const int n = t + i * T + r * (L / R0);
            double xr = 0.0, xi = 0.0;
            if (n < NE) {
                const c64 gm = load_gamma(a, base + n);
                gst[n] = gm;
                const double we = a.wE[n];
                const c64 ph = a.phase[n];
                const double gr = we * gm.x, gi = we * gm.y;
                xr = gr * ph.x - gi * ph.y;
                xi = gr * ph.y + gi * ph.x;
            }
            re[f] = xr;
            im[f] = xi;
        }
        StockhamPasses<L, 0, false>::run(re, im, lds, t, a.tw, a.twstride);
#pragma unroll
        for (int f = 0; f < V; ++f) {
            const int i = f / RL, r = f % RL;
            const c64 k = a.kodd[t + i * T + r * (L / RL)];
            const double xr = re[f], xi = im[f];
            re[f] = k.x * xr - k.y * xi;
            im[f] = k.x * xi + k.y * xr;
        }
        StockhamPasses<L, 0, true>::run(re, im, lds, t, a.tw, a.twstride);
#pragma unroll
        for (int f = 0; f < V; ++f) {
            const int i = f / R0, r = f % R0;
            const int n = t + i * T + r * (L / R0);
            const int e = n - (NE - 1);
            if (e < 0 || e >= NE) continue;
            const c64 ph = a.phase[n];
            const double yr = re[f] * invL, yi = im[f] * invL;
            a.out[base + e] = cmk(yr * ph.x + yi * ph.y, yi * ph.x - yr * ph.y);
        }
        // The main transform's input from the stash (written before the
        // passes' barriers).
#pragma unroll
        for (int f = 0; f < V; ++f) {
            const int i = f / R0, r = f % R0;
            const int e = t + i * T + r * (L / R0);
            double xr = 0.0, xi = 0.0;
            if (e < NE) {
                const c64 gm = gst[e];
                const double we = a.wE[e];
                xr = we * gm.x;
                xi = we * gm.y;
            }
            re[f] = xr;
            im[f] = xi;
        }
    } else {
        // Load wE*Gamma, zero padded, in the first pass's input order.
#pragma unroll
        for (int f = 0; f < V; ++f) {
            const int i = f / R0, r = f % R0;
            const int e = t + i * T + r * (L / R0);
            double xr = 0.0, xi = 0.0;
            if (e < NE) {
                const c64 gm = load_gamma(a, base + e);
                if (a.stash) gst[e] = gm;  // read back at the output, after the passes' barriers
                const double we = a.mode == 3 ? 1.0 : a.wE[e];  // the adjoint weights the output
                xr = we * gm.x;
                xi = we * gm.y;
            }
            re[f] = xr;
            im[f] = xi;
        }
    }
    StockhamPasses<L, 0, false>::run(re, im, lds, t, a.tw, a.twstride);
    // Kernel spectrum, at the forward outputs j + r L/RL.
#pragma unroll
    for (int f = 0; f < V; ++f) {
        const int i = f / RL, r = f % RL;
        const c64 k = a.kf[t + i * T + r * (L / RL)];
        const double xr = re[f], xi = im[f];
        re[f] = k.x * xr - k.y * xi;
        im[f] = k.x * xi + k.y * xr;
    }
    StockhamPasses<L, 0, true>::run(re, im, lds, t, a.tw, a.twstride);
    // Outputs of the inverse sit at j + r L/R0; Lambda_e = y[e + NE - 1] / L.
#pragma unroll
    for (int f = 0; f < V; ++f) {
        const int i = f / R0, r = f % R0;
        const int n = t + i * T + r * (L / R0);
        const int e = n - (NE - 1);
        if (e < 0 || e >= NE) continue;
        const c64 lam = cmk(re[f] * invL, im[f] * invL);
        const size_t gi = base + e;
        if (witness) {
            const c64 od = a.out[gi];  // parked by k_hilbert_odd or above
            const double dx = 0.5 * (lam.x + od.x) - lam.x, dy = 0.5 * (lam.y + od.y) - lam.y;
            dn += dx * dx + dy * dy;
            nn += lam.x * lam.x + lam.y * lam.y;
        }
        if (a.mode == 3) {
            // Gamma_bar = -wE * M SR_bar + (i/2) SR_bar (M antisymmetric);
            // S<_bar += i Gamma_bar, S>_bar -= i Gamma_bar.
            const c64 rb = a.gamma[gi];
            const double we = a.wE[e];
            const c64 ig = cmk(we * lam.y - 0.5 * rb.x, -we * lam.x - 0.5 * rb.y);
            a.out[gi] = cadd(a.out[gi], ig);
            a.out2[gi] = csub(a.out2[gi], ig);
            continue;
        }
        if (a.mode == 1) {
            const c64 gm = a.stash ? gst[e] : load_gamma(a, gi);
            a.out[gi] = cmk(lam.x + 0.5 * gm.y, lam.y - 0.5 * gm.x);  // Lambda - 0.5 i Gamma
        } else {
            a.out[gi] = lam;
        }
    }
    if (!witness) return;
    // Deterministic block reduction of the witness sums (fixed tree in LDS).
    double* rd = lds;
    double* rn = lds + T;
    __syncthreads();
    rd[t] = dn;
    rn[t] = nn;
    __syncthreads();
#pragma unroll
    for (int width = T / 2; width > 0; width >>= 1) {
        if (t < width) {
            rd[t] += rd[t + width];
            rn[t] += rn[t + width];
        }
        __syncthreads();
    }
    if (t == 0) {
        a.partial[col] = rd[0];
        a.partial[a.ncols + col] = rn[0];
    }
}

__global__ __launch_bounds__(1024) void k_witness_reduce(const double* partial, long long ncols, double mult,
                                                        double* out) {
    __shared__ double red[2][1024];
    const int t = threadIdx.x;
    const long long per = (ncols + 1023) / 1024;
    double sd = 0.0, sn = 0.0;
    for (long long c = t * per; c < min(ncols, (t + 1) * per); ++c) {
        sd += partial[c];
        sn += partial[ncols + c];
    }
    red[0][t] = sd;
    red[1][t] = sn;
    __syncthreads();
    for (int width = 512; width > 0; width >>= 1) {
        if (t < width) {
            red[0][t] += red[0][t + width];
            red[1][t] += red[1][t + width];
        }
        __syncthreads();
    }
    if (t == 0) out[0] = sqrt(mult * red[0][0]) / (sqrt(mult * red[1][0]) + 1e-14);
}

// ---- energy axes beyond one workgroup's LDS: a four-step FFT through HBM ---
// L = L1 * L2 (L1 <= 16, L2 a one-workgroup length), n = L2 n1 + n2 and
// k = k1 + L1 k2, so that
//   X[k1 + L1 k2] = sum_n2 w_L2^(n2 k2) w_L^(n2 k1) sum_n1 w_L1^(n1 k1) x[L2 n1 + n2]
// (w_M = exp(-2 pi i / M)). The convolution runs in three passes over a
// column-chunk scratch of L1 x L2 rows:
//  k_fs_columns: one thread per (column, n2) loads x[L2 n1 + n2], n1 < L1
//    (consecutive threads read consecutive n2, so the loads coalesce), runs
//    the L1-point DFT in registers, applies w_L^(n2 k1) and writes row k1;
//  k_fs_rows: one workgroup per row runs the L2-point transform with the
//    Stockham plan of k_hilbert, multiplies by the kernel spectrum (stored
//    in the same row order, so no transpose is ever made), runs the inverse
//    L2-point transform and applies w_L^(-n2 k1);
//  k_fs_output: the inverse L1-point DFT in registers gives y[L2 n1 + n2];
//    the outputs y[e + NE - 1] / L take the modes of k_hilbert.
// With the witness every pass also carries the odd half of the 2L transform
// (input x_n exp(-i pi n/L), the odd kernel spectrum), as k_hilbert_odd.
struct FourStepArgs {
    HilbertArgs h;       // NE, wE, inputs, outputs, mode, witness tables and partials
    c64* work;           // [nc][L1][L2] rows, then the odd half's (witness)
    const c64* kp;       // kernel spectrum in row order
    const c64* kpodd;    // odd half of the 2L spectrum in row order (NULL: no witness)
    long long col0, nc;  // the chunk's first column and its column count
    long long part0;     // the chunk's first witness partial
    long long nparts;    // witness partials of the whole call (one per output workgroup)
    int L;
};

constexpr int kFsThreads = 256;
constexpr int kMaxHilbertLength = 262144;  // L1 = 16 rows of 16384
constexpr int kMinFourStepLength = 4096;

// w_L^(n2 k1), k1 < L1, from w^(n2), w^(2 n2), w^(4 n2), w^(8 n2) (the loads
// of consecutive n2 coalesce, where tw[n2 k1] would stride by k1) with at
// most three products, as stockham_pass forms its powers.
template <int L1>
__device__ inline void fs_twiddles(const c64* __restrict__ tw, int n2, c64* w) {
    c64 b[4];
#pragma unroll
    for (int bit = 0; bit < 4; ++bit) b[bit] = (1 << bit) < L1 ? tw[(1 << bit) * n2] : cmk(1.0, 0.0);
    w[0] = cmk(1.0, 0.0);
#pragma unroll
    for (int k1 = 1; k1 < L1; ++k1) {
        c64 x = cmk(1.0, 0.0);
        bool first = true;
#pragma unroll
        for (int bit = 0; bit < 4; ++bit)
            if (k1 & (1 << bit)) {
                x = first ? b[bit] : cmul(x, b[bit]);
                first = false;
            }
        w[k1] = x;
    }
}

template <int L1>
__global__ __launch_bounds__(kFsThreads) void k_fs_columns(FourStepArgs f) {
    const int L = f.L, L2 = L / L1, NE = f.h.NE;
    const long long gid = (long long)blockIdx.x * kFsThreads + threadIdx.x;
    if (gid >= f.nc * L2) return;
    const long long c = gid / L2;
    const int n2 = (int)(gid - c * L2);
    const size_t base = (size_t)(f.col0 + c) * NE;
    const bool witness = f.kpodd != nullptr;
    double re[L1], im[L1], ro[L1], io[L1];
#pragma unroll
    for (int n1 = 0; n1 < L1; ++n1) {
        const int n = n2 + L2 * n1;
        double xr = 0.0, xi = 0.0;
        if (n < NE) {
            const c64 gm = load_gamma(f.h, base + n);
            const double we = f.h.mode == 3 ? 1.0 : f.h.wE[n];  // the adjoint weights the output
            xr = we * gm.x;
            xi = we * gm.y;
        }
        re[n1] = xr;
        im[n1] = xi;
        if (witness) {
            double yr = 0.0, yi = 0.0;
            if (n < NE) {
                const c64 ph = f.h.phase[n];
                yr = xr * ph.x - xi * ph.y;
                yi = xr * ph.y + xi * ph.x;
            }
            ro[n1] = yr;
            io[n1] = yi;
        }
    }
    dft<L1, false>(re, im);
    c64 w[L1];
    fs_twiddles<L1>(f.h.tw, n2, w);  // 8 n2 < L / 2
    c64* row = f.work + (size_t)c * L + n2;
#pragma unroll
    for (int k1 = 0; k1 < L1; ++k1)
        row[(size_t)k1 * L2] = cmk(w[k1].x * re[k1] - w[k1].y * im[k1], w[k1].x * im[k1] + w[k1].y * re[k1]);
    if (!witness) return;
    dft<L1, false>(ro, io);
    c64* rowo = row + (size_t)f.nc * L;
#pragma unroll
    for (int k1 = 0; k1 < L1; ++k1)
        rowo[(size_t)k1 * L2] = cmk(w[k1].x * ro[k1] - w[k1].y * io[k1], w[k1].x * io[k1] + w[k1].y * ro[k1]);
}

// At most 128 VGPRs: two 512-thread workgroups per CU at L2 = 4096, so that
// one row's loads overlap another's transforms.
template <int L2>
__global__ __launch_bounds__(threads_for(L2)) __attribute__((amdgpu_waves_per_eu(L2 == 4096 ? 4 : 1))) void k_fs_rows(
    FourStepArgs f) {
    constexpr int V = vals_per_thread(L2), T = threads_for(L2);
    constexpr int NPASS = npasses(L2);
    constexpr int R0 = fwd_radix(L2, 0), RL = fwd_radix(L2, NPASS - 1);
    extern __shared__ __align__(16) double lds[];
    const int t = threadIdx.x;
    const int L = f.L, L1 = L / L2;
    const long long rowi = blockIdx.x;  // c * L1 + k1
    const int k1 = (int)(rowi % L1);
    const bool odd = blockIdx.y == 1;
    c64* row = f.work + (odd ? (size_t)f.nc * L : 0) + (size_t)rowi * L2;
    const c64* kp = (odd ? f.kpodd : f.kp) + (size_t)k1 * L2;
    double re[V], im[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int i = v / R0, r = v % R0;
        const c64 x = row[t + i * T + r * (L2 / R0)];
        re[v] = x.x;
        im[v] = x.y;
    }
    // the table holds w_L; stride L1 gives w_L2
    StockhamPasses<L2, 0, false>::run(re, im, lds, t, f.h.tw, L1);
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int i = v / RL, r = v % RL;
        const c64 k = kp[t + i * T + r * (L2 / RL)];
        const double xr = re[v], xi = im[v];
        re[v] = k.x * xr - k.y * xi;
        im[v] = k.x * xi + k.y * xr;
    }
    StockhamPasses<L2, 0, true>::run(re, im, lds, t, f.h.tw, L1);
    // each thread writes back the positions it read, times w_L^(-n2 k1):
    // w_L^(k1 (t + i T)) (one gather per i) times the workgroup-uniform
    // w_L^(k1 r L2/R0)
#pragma unroll
    for (int i = 0; i < V / R0; ++i) {
        const c64 a = f.h.tw[k1 * (t + i * T)];
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int v = i * R0 + r;
            const int n2 = t + i * T + r * (L2 / R0);
            const c64 w = r == 0 ? a : cmul(a, f.h.tw[k1 * r * (L2 / R0)]);  // conjugated below
            row[n2] = cmk(w.x * re[v] + w.y * im[v], w.x * im[v] - w.y * re[v]);
        }
    }
}

template <int L1>
__global__ __launch_bounds__(kFsThreads) void k_fs_output(FourStepArgs f) {
    __shared__ double rd[kFsThreads], rn[kFsThreads];
    const int L = f.L, L2 = L / L1, NE = f.h.NE;
    const int t = threadIdx.x;
    const long long gid = (long long)blockIdx.x * kFsThreads + t;
    const bool witness = f.kpodd != nullptr;
    const double invL = 1.0 / (double)L;
    double dn = 0.0, nn = 0.0;
    if (gid < f.nc * L2) {
        const long long c = gid / L2;
        const int n2 = (int)(gid - c * L2);
        const size_t base = (size_t)(f.col0 + c) * NE;
        const c64* row = f.work + (size_t)c * L + n2;
        double re[L1], im[L1], ro[L1], io[L1];
#pragma unroll
        for (int k1 = 0; k1 < L1; ++k1) {
            const c64 x = row[(size_t)k1 * L2];
            re[k1] = x.x;
            im[k1] = x.y;
        }
        dft<L1, true>(re, im);
        if (witness) {
            const c64* rowo = row + (size_t)f.nc * L;
#pragma unroll
            for (int k1 = 0; k1 < L1; ++k1) {
                const c64 x = rowo[(size_t)k1 * L2];
                ro[k1] = x.x;
                io[k1] = x.y;
            }
            dft<L1, true>(ro, io);
        }
#pragma unroll
        for (int n1 = 0; n1 < L1; ++n1) {
            const int n = n2 + L2 * n1;
            const int e = n - (NE - 1);
            if (e < 0 || e >= NE) continue;
            const c64 lam = cmk(re[n1] * invL, im[n1] * invL);
            const size_t gi = base + e;
            if (witness) {
                const c64 ph = f.h.phase[n];  // exp(+i pi n/L) = conj(ph)
                const double yr = ro[n1] * invL, yi = io[n1] * invL;
                const c64 od = cmk(yr * ph.x + yi * ph.y, yi * ph.x - yr * ph.y);
                const double dx = 0.5 * (lam.x + od.x) - lam.x, dy = 0.5 * (lam.y + od.y) - lam.y;
                dn += dx * dx + dy * dy;
                nn += lam.x * lam.x + lam.y * lam.y;
            }
            if (f.h.mode == 3) {
                const c64 rb = f.h.gamma[gi];
                const double we = f.h.wE[e];
                const c64 ig = cmk(we * lam.y - 0.5 * rb.x, -we * lam.x - 0.5 * rb.y);
                f.h.out[gi] = cadd(f.h.out[gi], ig);
                f.h.out2[gi] = csub(f.h.out2[gi], ig);
            } else if (f.h.mode == 1) {
                const c64 gm = load_gamma(f.h, gi);
                f.h.out[gi] = cmk(lam.x + 0.5 * gm.y, lam.y - 0.5 * gm.x);  // Lambda - 0.5 i Gamma
            } else {
                f.h.out[gi] = lam;
            }
        }
    }
    if (!witness) return;  // uniform over the workgroup
    rd[t] = dn;
    rn[t] = nn;
    __syncthreads();
#pragma unroll
    for (int width = kFsThreads / 2; width > 0; width >>= 1) {
        if (t < width) {
            rd[t] += rd[t + width];
            rn[t] += rn[t + width];
        }
        __syncthreads();
    }
    if (t == 0) {
        f.h.partial[f.part0 + blockIdx.x] = rd[0];
        f.h.partial[f.nparts + f.part0 + blockIdx.x] = rn[0];
    }
}

// Row length of the four-step split: L / L2 <= 16 and L2 a length the
// one-workgroup transform is tuned for (4096 while L1 <= 16 allows it).
int fourstep_rows(int L) { return L <= 8192 ? L / 2 : (L <= 65536 ? 4096 : L / 16); }

// Host FFT in long double for the plan tables.
void host_fft(std::vector<std::complex<long double>>& x) {
    const size_t L = x.size();
    for (size_t i = 1, j = 0; i < L; ++i) {
        size_t bit = L >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(x[i], x[j]);
    }
    const long double pi = 3.141592653589793238462643383279502884L;
    for (size_t len = 2; len <= L; len <<= 1) {
        for (size_t k = 0; k < len / 2; ++k) {
            long double ang = -2.0L * pi * (long double)k / (long double)len;
            std::complex<long double> w(cosl(ang), sinl(ang));
            for (size_t s = 0; s < L; s += len) {
                auto u = x[s + k], v = x[s + k + len / 2] * w;
                x[s + k] = u + v;
                x[s + k + len / 2] = u - v;
            }
        }
    }
}

// Spectrum of the kernel sequence k[lag + NE - 1] = 1/(2 pi lag de) on L
// points; odd: the odd-frequency half of its 2L spectrum,
// DFT_L(k_n exp(-i pi n/L)) (the sequence is shorter than L).
std::vector<c64> kernel_spectrum(int NE, int L, double de, bool odd) {
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<std::complex<long double>> k(L, 0.0L);
    for (int lag = -(NE - 1); lag <= NE - 1; ++lag) {
        if (lag == 0) continue;
        // The reference builds the kernel in Float64: inv(2 pi lag de).
        double v = 1.0 / (2.0 * M_PI * (double)lag * de);
        k[lag + NE - 1] = (long double)v;
    }
    if (odd)
        for (int n = 0; n < L; ++n)
            k[n] *= std::complex<long double>(cosl(-pi * n / (long double)L), sinl(-pi * n / (long double)L));
    host_fft(k);
    std::vector<c64> out(L);
    for (int i = 0; i < L; ++i) out[i] = cmk((double)k[i].real(), (double)k[i].imag());
    (void)pi;
    return out;
}

}  // namespace

template <int L>
int launch_hilbert_len(qclhip_ctx* ctx, long long ncols, const HilbertArgs& a) {
    constexpr int T = threads_for(L);
    const size_t lds = (size_t)fft_lds_len<L>() * sizeof(double);
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "FFT length exceeds the LDS of one workgroup");
    // S^R output: keep Gamma in LDS instead of reading S< and S> again,
    // while two workgroups still fit on a CU (L <= 4096).
    HilbertArgs b = a;
    const size_t lds_main = lds + (size_t)a.NE * sizeof(c64);
    b.stash = a.mode == 1 && lds_main <= 80 * 1024 && lds_main <= ctx->max_lds_per_block;
    const size_t lds_k = b.stash ? lds_main : lds;
    // At L = 4096 (8 values per thread) the fused kernel needs no more
    // registers than the plain one; at 16 values per thread it would cost occupancy.
    if constexpr (L == 4096) {
        if (a.kodd && b.stash) {  // one kernel runs the odd half and the transform
            if (lds_k > 65536)
                QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_hilbert<L, true>,
                                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_k));
            hipLaunchKernelGGL((k_hilbert<L, true>), dim3((unsigned)ncols), dim3(T), lds_k, ctx->stream, b);
            return launch_check(ctx, "k_hilbert");
        }
    }
    if (lds_k > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_hilbert<L, false>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_k));
    if (a.kodd) {
        if (lds > 65536)
            QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_hilbert_odd<L>,
                                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_hilbert_odd<L>, dim3((unsigned)ncols), dim3(T), lds, ctx->stream, a);
    }
    hipLaunchKernelGGL((k_hilbert<L, false>), dim3((unsigned)ncols), dim3(T), lds_k, ctx->stream, b);
    return launch_check(ctx, "k_hilbert");
}

template <int L1>
int launch_fs_columns_output(qclhip_ctx* ctx, const FourStepArgs& f, bool output, unsigned blocks) {
    if (output)
        hipLaunchKernelGGL(k_fs_output<L1>, dim3(blocks), dim3(kFsThreads), 0, ctx->stream, f);
    else
        hipLaunchKernelGGL(k_fs_columns<L1>, dim3(blocks), dim3(kFsThreads), 0, ctx->stream, f);
    return launch_check(ctx, output ? "k_fs_output" : "k_fs_columns");
}

int fs_columns_output(qclhip_ctx* ctx, const FourStepArgs& f, int L1, bool output, unsigned blocks) {
    switch (L1) {
        case 2: return launch_fs_columns_output<2>(ctx, f, output, blocks);
        case 4: return launch_fs_columns_output<4>(ctx, f, output, blocks);
        case 8: return launch_fs_columns_output<8>(ctx, f, output, blocks);
        case 16: return launch_fs_columns_output<16>(ctx, f, output, blocks);
        default: return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "four-step split outside L1 = 2..16");
    }
}

template <int L2>
int launch_fs_rows(qclhip_ctx* ctx, const FourStepArgs& f, unsigned rows) {
    const size_t lds = (size_t)fft_lds_len<L2>() * sizeof(double);
    if (lds > ctx->max_lds_per_block)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "FFT row length exceeds the LDS of one workgroup");
    if (lds > 65536)
        QCLHIP_TRY_HIP(ctx, hipFuncSetAttribute((const void*)k_fs_rows<L2>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_fs_rows<L2>, dim3(rows, f.kpodd ? 2 : 1), dim3(threads_for(L2)), lds, ctx->stream, f);
    return launch_check(ctx, "k_fs_rows");
}

int fs_rows(qclhip_ctx* ctx, const FourStepArgs& f, int L2, unsigned rows) {
    switch (L2) {
        case 2048: return launch_fs_rows<2048>(ctx, f, rows);
        case 4096: return launch_fs_rows<4096>(ctx, f, rows);
        case 8192: return launch_fs_rows<8192>(ctx, f, rows);
        case 16384: return launch_fs_rows<16384>(ctx, f, rows);
        default: return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "four-step row length outside 2048..16384");
    }
}

// Columns per four-step chunk: the chunk's rows (and the odd half's) within
// kFourStepChunkBytes of scratch.
constexpr size_t kFourStepChunkBytes = size_t(128) << 20;

long long fourstep_chunk(const qclhip_hilbert* p, bool witness) {
    const size_t per_col = (size_t)p->L * sizeof(c64) * (witness ? 2 : 1);
    return std::max<long long>(1, std::min<long long>(p->ncols, (long long)(kFourStepChunkBytes / per_col)));
}

bool fourstep_path(const qclhip_hilbert* p) {
    return p->path == QCLHIP_HILBERT_PATH_HBM || (p->path == QCLHIP_HILBERT_PATH_AUTO && p->L > 16384);
}

// The four-step transform of every column, chunk by chunk, in the scratch
// `work` (fourstep_chunk columns of rows, times two with the witness).
int hilbert_fourstep(qclhip_ctx* ctx, const qclhip_hilbert* p, const HilbertArgs& a, c64* work) {
    FourStepArgs f{};
    f.h = a;
    f.h.tw = p->twiddle;
    f.work = work;
    f.kp = p->kp[0];
    f.kpodd = a.kodd ? p->kp[1] : nullptr;
    f.L = p->L;
    const int L1 = p->L1, L2 = p->L2;
    const long long chunk = fourstep_chunk(p, a.kodd != nullptr);
    f.nparts = p->ncols * L2 / kFsThreads;
    for (long long c0 = 0; c0 < p->ncols; c0 += chunk) {
        f.col0 = c0;
        f.nc = std::min(chunk, p->ncols - c0);
        f.part0 = c0 * L2 / kFsThreads;
        const unsigned blocks = (unsigned)(f.nc * L2 / kFsThreads);  // L2 >= 2048
        int rc = fs_columns_output(ctx, f, L1, false, blocks);
        if (rc) return rc;
        rc = fs_rows(ctx, f, L2, (unsigned)(f.nc * L1));
        if (rc) return rc;
        rc = fs_columns_output(ctx, f, L1, true, blocks);
        if (rc) return rc;
    }
    return QCLHIP_OK;
}

int hilbert_launch(qclhip_ctx* ctx, const qclhip_hilbert* p, const HilbertArgs& a_in) {
    HilbertArgs a = a_in;
    const int L = p->L;
    a.kf = p->kf[0];
    a.tw = p->twiddle;
    a.twstride = 1;
    switch (L) {
        case 4: return launch_hilbert_len<4>(ctx, p->ncols, a);
        case 8: return launch_hilbert_len<8>(ctx, p->ncols, a);
        case 16: return launch_hilbert_len<16>(ctx, p->ncols, a);
        case 32: return launch_hilbert_len<32>(ctx, p->ncols, a);
        case 64: return launch_hilbert_len<64>(ctx, p->ncols, a);
        case 128: return launch_hilbert_len<128>(ctx, p->ncols, a);
        case 256: return launch_hilbert_len<256>(ctx, p->ncols, a);
        case 512: return launch_hilbert_len<512>(ctx, p->ncols, a);
        case 1024: return launch_hilbert_len<1024>(ctx, p->ncols, a);
        case 2048: return launch_hilbert_len<2048>(ctx, p->ncols, a);
        case 4096: return launch_hilbert_len<4096>(ctx, p->ncols, a);
        case 8192: return launch_hilbert_len<8192>(ctx, p->ncols, a);
        case 16384: return launch_hilbert_len<16384>(ctx, p->ncols, a);
        default: return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "FFT length outside 4..16384");
    }
}

int hilbert_run(qclhip_ctx* ctx, const qclhip_hilbert* p, const double* Gamma, const double* SL,
                const double* SG, double* out, int mode, double* roundoff_host, double* out2 = nullptr) {
    HilbertArgs a{};
    a.NE = p->NE;
    a.ncols = p->ncols;
    a.wE = p->grid->wE;
    a.gamma = (const c64*)Gamma;
    a.SL = (const c64*)SL;
    a.SG = (const c64*)SG;
    a.out = (c64*)out;
    a.out2 = (c64*)out2;
    a.mode = mode;
    if (roundoff_host && !p->verify)
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "the FFT plan was built without roundoff verification");
    const bool fs = fourstep_path(p);
    const bool witness = roundoff_host != nullptr;
    // One scratch request: the witness partials (one per column, or one per
    // output workgroup of the four-step path) and the four-step rows.
    const long long nparts = fs ? p->ncols * p->L2 / kFsThreads : p->ncols;
    const size_t part_bytes = witness ? ((size_t)(2 * nparts + 8) * sizeof(double) + 255) / 256 * 256 : 0;
    const size_t work_bytes = fs ? (size_t)fourstep_chunk(p, witness) * p->L * sizeof(c64) * (witness ? 2 : 1) : 0;
    char* scratch = nullptr;
    if (part_bytes + work_bytes) {
        scratch = (char*)workspace(ctx, part_bytes + work_bytes);
        if (!scratch) return fail(ctx, QCLHIP_ERR_RUNTIME, "scratch allocation failed");
    }
    double* partial = witness ? (double*)scratch : nullptr;
    if (witness) {
        a.kodd = p->kf[1];
        a.phase = p->phase;
        a.partial = partial;
    }
    int rc = fs ? hilbert_fourstep(ctx, p, a, (c64*)(scratch + part_bytes)) : hilbert_launch(ctx, p, a);
    if (rc || !witness) return rc;
    double* res = partial + 2 * nparts;
    hipLaunchKernelGGL(k_witness_reduce, dim3(1), dim3(1024), 0, ctx->stream, partial, nparts, p->col_mult, res);
    rc = launch_check(ctx, "k_witness_reduce");
    if (rc) return rc;
    bool deferred = false;
    rc = defer_result(ctx, res, 1, &deferred);
    if (rc || deferred) return rc;
    QCLHIP_TRY_HIP(ctx, hipMemcpyAsync(ctx->host_scalars, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    rc = qclhip_ctx_synchronize(ctx);
    if (rc) return rc;
    *roundoff_host = ctx->host_scalars[0];
    return QCLHIP_OK;
}

}  // namespace qclhip

using namespace qclhip;

extern "C" {

int qclhip_hilbert_create(qclhip_ctx* ctx, const qclhip_grid* grid, int64_t ncols, int verify,
                          qclhip_hilbert** out) {
    if (!ctx || !grid || !out) return QCLHIP_ERR_ARGUMENT;
    *out = nullptr;
    if (ncols <= 0) return fail(ctx, QCLHIP_ERR_ARGUMENT, "Ncolumns must be positive");
    if (grid->NE < 2) return fail(ctx, QCLHIP_ERR_ARGUMENT, "at least two energy nodes are required");
    if (!(grid->de > 0)) return fail(ctx, QCLHIP_ERR_ARGUMENT, "FFT Hilbert transform requires a uniform energy grid");
    int L = 1;
    while (L < 2 * grid->NE - 1) L <<= 1;
    const int Lmax = L;  // the witness runs at length L (odd-frequency half of the 2L transform)
    if (Lmax > kMaxHilbertLength)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "FFT length above 262144 (NE above 131072)");
    auto* p = new qclhip_hilbert();
    p->grid = grid;
    p->NE = grid->NE;
    p->ncols = ncols;
    p->L = L;
    p->Lmax = Lmax;
    p->verify = verify ? 1 : 0;
    // Twiddles exp(-2 pi i idx/Lmax) in long double, rounded once.
    std::vector<c64> tw(Lmax);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < Lmax; ++i) {
        long double ang = -2.0L * pi * (long double)i / (long double)Lmax;
        tw[i] = cmk((double)cosl(ang), (double)sinl(ang));
    }
    bool ok = hipMalloc(&p->twiddle, (size_t)Lmax * sizeof(c64)) == hipSuccess &&
              hipMemcpy(p->twiddle, tw.data(), (size_t)Lmax * sizeof(c64), hipMemcpyHostToDevice) == hipSuccess;
    if (L >= kMinFourStepLength) {
        p->L2 = fourstep_rows(L);
        p->L1 = L / p->L2;
    }
    for (int w = 0; w < (verify ? 2 : 1) && ok; ++w) {
        std::vector<c64> kf = kernel_spectrum(grid->NE, L, grid->de, w == 1);
        ok = hipMalloc(&p->kf[w], (size_t)L * sizeof(c64)) == hipSuccess &&
             hipMemcpy(p->kf[w], kf.data(), (size_t)L * sizeof(c64), hipMemcpyHostToDevice) == hipSuccess;
        if (ok && p->L2) {
            // row order of the four-step path: kp[k1 * L2 + k2] = kf[k1 + L1 * k2]
            std::vector<c64> kp((size_t)L);
            for (int k1 = 0; k1 < p->L1; ++k1)
                for (int k2 = 0; k2 < p->L2; ++k2) kp[(size_t)k1 * p->L2 + k2] = kf[(size_t)k1 + (size_t)p->L1 * k2];
            ok = hipMalloc(&p->kp[w], (size_t)L * sizeof(c64)) == hipSuccess &&
                 hipMemcpy(p->kp[w], kp.data(), (size_t)L * sizeof(c64), hipMemcpyHostToDevice) == hipSuccess;
        }
    }
    if (verify && ok) {
        std::vector<c64> ph(L);
        for (int i = 0; i < L; ++i) {
            long double ang = -pi * (long double)i / (long double)L;
            ph[i] = cmk((double)cosl(ang), (double)sinl(ang));
        }
        ok = hipMalloc(&p->phase, (size_t)L * sizeof(c64)) == hipSuccess &&
             hipMemcpy(p->phase, ph.data(), (size_t)L * sizeof(c64), hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
        qclhip_hilbert_destroy(p);
        return fail(ctx, QCLHIP_ERR_RUNTIME, "Hilbert plan allocation failed");
    }
    *out = p;
    return QCLHIP_OK;
}

int qclhip_hilbert_destroy(qclhip_hilbert* p) {
    if (!p) return QCLHIP_OK;
    hipDeviceSynchronize();
    if (p->twiddle) hipFree(p->twiddle);
    if (p->kf[0]) hipFree(p->kf[0]);
    if (p->kf[1]) hipFree(p->kf[1]);
    if (p->phase) hipFree(p->phase);
    if (p->kp[0]) hipFree(p->kp[0]);
    if (p->kp[1]) hipFree(p->kp[1]);
    delete p;
    return QCLHIP_OK;
}

int qclhip_hilbert_set_multiplicity(qclhip_ctx* ctx, qclhip_hilbert* plan, int64_t mult) {
    if (!ctx || !plan) return QCLHIP_ERR_ARGUMENT;
    if (mult < 1) return fail(ctx, QCLHIP_ERR_ARGUMENT, "the column multiplicity must be positive");
    plan->col_mult = (double)mult;
    return QCLHIP_OK;
}

int qclhip_hilbert_set_path(qclhip_ctx* ctx, qclhip_hilbert* plan, int path) {
    if (!ctx || !plan) return QCLHIP_ERR_ARGUMENT;
    if (path == QCLHIP_HILBERT_PATH_LDS && plan->L > 16384)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "the one-workgroup transform takes L <= 16384");
    if (path == QCLHIP_HILBERT_PATH_HBM && !plan->L2)
        return fail(ctx, QCLHIP_ERR_UNSUPPORTED, "the four-step transform takes L >= 4096");
    if (path < QCLHIP_HILBERT_PATH_AUTO || path > QCLHIP_HILBERT_PATH_HBM)
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "unknown Hilbert transform path");
    plan->path = path;
    return QCLHIP_OK;
}

int qclhip_hilbert_get_path(qclhip_ctx* ctx, const qclhip_hilbert* plan, int* path_host) {
    if (!ctx || !plan || !path_host) return QCLHIP_ERR_ARGUMENT;
    *path_host = fourstep_path(plan) ? QCLHIP_HILBERT_PATH_HBM : QCLHIP_HILBERT_PATH_LDS;
    return QCLHIP_OK;
}

int qclhip_fft_hilbert(qclhip_ctx* ctx, const qclhip_hilbert* plan, const double* Gamma, double* Lambda,
                       double* roundoff_host) {
    if (!ctx || !plan || !Gamma || !Lambda) return QCLHIP_ERR_ARGUMENT;
    if (Gamma == Lambda) return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    return hilbert_run(ctx, plan, Gamma, nullptr, nullptr, Lambda, 0, roundoff_host);
}

int qclhip_retarded_self_energy_adjoint(qclhip_ctx* ctx, const qclhip_hilbert* plan, const double* SRbar,
                                        double* SLbar, double* SGbar) {
    if (!ctx || !plan || !SRbar || !SLbar || !SGbar) return QCLHIP_ERR_ARGUMENT;
    if (SLbar == SGbar || SRbar == SLbar || SRbar == SGbar)
        return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    return hilbert_run(ctx, plan, SRbar, nullptr, nullptr, SLbar, 3, nullptr, SGbar);
}

int qclhip_retarded_self_energy(qclhip_ctx* ctx, const qclhip_hilbert* plan, const double* SL, const double* SG,
                                double* SR, double* Gamma, double* roundoff_host) {
    if (!ctx || !plan || !SL || !SG || !SR) return QCLHIP_ERR_ARGUMENT;
    if (SR == SL || SR == SG) return fail(ctx, QCLHIP_ERR_ARGUMENT, "in-place aliasing is not supported");
    int rc = hilbert_run(ctx, plan, nullptr, SL, SG, SR, 1, roundoff_host);
    if (rc || !Gamma) return rc;
    // Gamma output requested: i (S> - S<), elementwise.
    return qclhip::gamma_field(ctx, (long long)plan->NE * plan->ncols, SL, SG, Gamma);
}

}  // extern "C"
